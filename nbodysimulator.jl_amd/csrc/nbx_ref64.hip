// nbx_ref64.hip -- FP64 kernels that reproduce the reference's arithmetic bit for bit.
//
// Built with -ffp-contract=off: every a*b+c rounds twice, as Julia does without
// @fastmath/@muladd.  Every per-target sum walks the source atoms in ascending index
// order (the reference's `for j in indxs`): gravity keeps one thread per target with
// sources staged through LDS tiles, Lennard-Jones and Coulomb give each target a
// wavefront whose lanes compute the pair terms of 64 sources at once before the terms
// are added in j order.  Neither changes the order or the rounding of the sums.  Device
// state is SoA: x[k * n + i].
#include <mutex>

#include "nbx_internal.h"

#define REF_TILE 256

// ---------------------------------------------------------------------------
// Layout / fill helpers
// ---------------------------------------------------------------------------

__global__ void k_fill(double* __restrict__ p, int64_t count, double value) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < count) p[t] = value;
}

__global__ void k_aos_to_soa(const double* __restrict__ aos, double* __restrict__ soa,
                             int64_t n) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < 3 * n) {
        int64_t i = t / 3, k = t - 3 * i;
        soa[k * n + i] = aos[t];
    }
}

__global__ void k_soa_to_aos(const double* __restrict__ soa, double* __restrict__ aos,
                             int64_t n) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < 3 * n) {
        int64_t i = t / 3, k = t - 3 * i;
        aos[t] = soa[k * n + i];
    }
}

void ref64_fill(hipStream_t s, double* p, int64_t count, double value) {
    if (count <= 0) return;
    hipLaunchKernelGGL(k_fill, dim3(nbx_blocks(count, 256)), dim3(256), 0, s, p, count, value);
}
void ref64_aos_to_soa(hipStream_t s, const double* aos, double* soa, int64_t n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_aos_to_soa, dim3(nbx_blocks(3 * n, 256)), dim3(256), 0, s, aos, soa, n);
}
void ref64_soa_to_aos(hipStream_t s, const double* soa, double* aos, int64_t n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_soa_to_aos, dim3(nbx_blocks(3 * n, 256)), dim3(256), 0, s, soa, aos, n);
}

// ---------------------------------------------------------------------------
// Lennard-Jones: src/basic_potentials.jl:240-272
// ---------------------------------------------------------------------------

// Wavefront per target.  The one-thread-per-target loop leaves the chip idle at small n
// (n = 864 is 14 wavefronts on 1,024 SIMDs), and its cost is the per-pair work: the wrap
// loops and two FP64 divides.  Here the 64 lanes of a wavefront evaluate the pair terms
// factor * r_ij of 64 consecutive sources j at once, and the x, y, z sums then take those
// terms one j after another (each term read from its lane with v_readlane, the adds in
// ascending j), so every sum sees the reference's terms in the reference's order: the
// result is bit for bit the one-thread loop's.  A skipped pair contributes +0.0, which
// leaves a sum unchanged (the sums start at +0.0 and never become -0.0 under
// round-to-nearest).
#define REF_WAVES 4   // targets (wavefronts) per block

__device__ __forceinline__ double ref_lane(double v, int q) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), q);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), q);
    return __hiloint2double(hi, lo);
}

// f += t_q for q = 0..cnt-1 (lane q holds t_q), in q order
__device__ __forceinline__ void ref_add_lanes(double& f1, double& f2, double& f3, double t1,
                                              double t2, double t3, int cnt) {
    if (cnt == 64) {
#pragma unroll
        for (int q = 0; q < 64; ++q) {
            f1 += ref_lane(t1, q);
            f2 += ref_lane(t2, q);
            f3 += ref_lane(t3, q);
        }
    } else {
        for (int q = 0; q < cnt; ++q) {
            f1 += ref_lane(t1, q);
            f2 += ref_lane(t2, q);
            f3 += ref_lane(t3, q);
        }
    }
}

__global__ void __launch_bounds__(64 * REF_WAVES)
k_ref_lj(const double* __restrict__ x, double* __restrict__ dv, const double* __restrict__ ms,
         int64_t n, int64_t ib, int64_t cnt, int istride, int jstride, int64_t nj, double s2,
         double R2, double eps, nbx_bcv bc) {
    const int lane = threadIdx.x & 63;
    const int64_t t = (int64_t)blockIdx.x * REF_WAVES + (threadIdx.x >> 6);
    if (t >= cnt) return;   // the whole wavefront
    const int64_t i = ib + t * istride;
    const double xi = x[i], yi = x[n + i], zi = x[2 * n + i];
    double f1 = 0.0, f2 = 0.0, f3 = 0.0;
    for (int64_t base = 0; base < nj; base += 64) {
        const int64_t jj = base + lane;
        const int64_t j = jj * jstride;
        double t1 = 0.0, t2 = 0.0, t3 = 0.0;
        if (jj < nj && j != i) {
            double rx, ry, rz, rij_2;
            nbx_displacement(xi, yi, zi, x[j], x[n + j], x[2 * n + j], bc, rx, ry, rz, rij_2);
            if (rij_2 < R2) {
                double qq = s2 / rij_2;
                double s6 = qq * qq * qq;
                double s12 = s6 * s6;
                double factor = (2.0 * s12 - s6) / rij_2;
                t1 = factor * rx;
                t2 = factor * ry;
                t3 = factor * rz;
            }
        }
        ref_add_lanes(f1, f2, f3, t1, t2, t3, (int)((nj - base) < 64 ? (nj - base) : 64));
    }
    if (lane == 0) {
        double coeff = 24.0 * eps / ms[i];
        dv[i] += coeff * f1;
        dv[n + i] += coeff * f2;
        dv[2 * n + i] += coeff * f3;
    }
}

void ref64_lj(hipStream_t s, const double* x, double* dv, const double* ms, int64_t n,
              int64_t ib, int64_t ie, int istride, int jstride, int64_t nj, double eps,
              double sigma, double R, const nbx_bcv& bc) {
    if (ie <= ib) return;
    const int64_t cnt = (ie - ib + istride - 1) / istride;
    hipLaunchKernelGGL(k_ref_lj, dim3(nbx_blocks(cnt, REF_WAVES)), dim3(64 * REF_WAVES), 0, s, x,
                       dv, ms, n, ib, cnt, istride, jstride, nj, sigma * sigma, R * R, eps, bc);
}

// ---------------------------------------------------------------------------
// Electrostatics (truncated Coulomb, src/basic_potentials.jl:274-304) and reaction field
// ---------------------------------------------------------------------------

__device__ __forceinline__ bool ref_excluded(int64_t i, int64_t j, int water) {
    return water ? (i / 3) == (j / 3) : i == j;   // src/nbody_to_ode.jl:326, :346-348
}

template <bool RF>
__global__ void __launch_bounds__(64 * REF_WAVES)
k_ref_coulomb(const double* __restrict__ x, double* __restrict__ dv,
              const double* __restrict__ ms, const double* __restrict__ qs, int64_t n,
              int64_t ib, int64_t ie, int water, double k, double R2, double krf2,
              nbx_bcv bc) {
    const int lane = threadIdx.x & 63;
    const int64_t i = ib + (int64_t)blockIdx.x * REF_WAVES + (threadIdx.x >> 6);
    if (i >= ie) return;   // the whole wavefront
    const double xi = x[i], yi = x[n + i], zi = x[2 * n + i];
    double f1 = 0.0, f2 = 0.0, f3 = 0.0;
    for (int64_t base = 0; base < n; base += 64) {
        const int64_t jj = base + lane;
        double t1 = 0.0, t2 = 0.0, t3 = 0.0;
        if (jj < n && !ref_excluded(i, jj, water)) {
            double rx, ry, rz, r2;
            nbx_displacement(xi, yi, zi, x[jj], x[n + jj], x[2 * n + jj], bc, rx, ry, rz, r2);
            if (r2 < R2) {
                const double qj = qs[jj];
                double r = sqrt(r2);
                double factor;
                if (RF) factor = qj * (1.0 / (r * r2) - krf2);
                else factor = qj / (r * r2);
                t1 = factor * rx;
                t2 = factor * ry;
                t3 = factor * rz;
            }
        }
        ref_add_lanes(f1, f2, f3, t1, t2, t3, (int)((n - base) < 64 ? (n - base) : 64));
    }
    if (lane == 0) {
        double coeff = k * qs[i] / ms[i];
        dv[i] += coeff * f1;
        dv[n + i] += coeff * f2;
        dv[2 * n + i] += coeff * f3;
    }
}

static double rf_krf(double R, double eps_rf) {
    if (__builtin_isinf(eps_rf)) return 1.0 / (2.0 * R * R * R);
    return (eps_rf - 1.0) / ((2.0 * eps_rf + 1.0) * (R * R * R));
}

void ref64_elec(hipStream_t s, const double* x, double* dv, const double* ms,
                const double* qs, int64_t n, int64_t ib, int64_t ie, int water, double k,
                double R, const nbx_bcv& bc) {
    if (ie <= ib) return;
    hipLaunchKernelGGL(k_ref_coulomb<false>, dim3(nbx_blocks(ie - ib, REF_WAVES)),
                       dim3(64 * REF_WAVES), 0, s, x, dv, ms, qs, n, ib, ie, water, k, R * R,
                       0.0, bc);
}

void ref64_rf(hipStream_t s, const double* x, double* dv, const double* ms, const double* qs,
              int64_t n, int64_t ib, int64_t ie, int water, double k, double R,
              double eps_rf, const nbx_bcv& bc) {
    if (ie <= ib) return;
    hipLaunchKernelGGL(k_ref_coulomb<true>, dim3(nbx_blocks(ie - ib, REF_WAVES)),
                       dim3(64 * REF_WAVES), 0, s, x, dv, ms, qs, n, ib, ie, water, k, R * R,
                       2.0 * rf_krf(R, eps_rf), bc);
}

// ---------------------------------------------------------------------------
// Gravity: src/basic_potentials.jl:306-331 (no boundary conditions, no cutoff)
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(REF_TILE)
k_ref_grav(const double* __restrict__ x, double* __restrict__ dv, const double* __restrict__ ms,
           int64_t n, int64_t ib, int64_t ie, double G) {
    __shared__ double sx[REF_TILE], sy[REF_TILE], sz[REF_TILE], sm[REF_TILE];
    const int64_t i = ib + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = i < ie;
    double xi = 0.0, yi = 0.0, zi = 0.0;
    if (active) { xi = x[i]; yi = x[n + i]; zi = x[2 * n + i]; }
    double a1 = 0.0, a2 = 0.0, a3 = 0.0;
    for (int64_t base = 0; base < n; base += REF_TILE) {
        const int64_t j = base + threadIdx.x;
        if (j < n) {
            sx[threadIdx.x] = x[j]; sy[threadIdx.x] = x[n + j];
            sz[threadIdx.x] = x[2 * n + j]; sm[threadIdx.x] = ms[j];
        }
        __syncthreads();
        const int lim = (int)((n - base) < REF_TILE ? (n - base) : REF_TILE);
        if (active) {
            for (int q = 0; q < lim; ++q) {
                if (base + q == i) continue;
                double rx = xi - sx[q], ry = yi - sy[q], rz = zi - sz[q];
                double ss = rx * rx + ry * ry;
                double nr = sqrt(ss + rz * rz);
                double factor = -G * sm[q] / (nr * nr * nr);
                a1 += factor * rx;
                a2 += factor * ry;
                a3 += factor * rz;
            }
        }
        __syncthreads();
    }
    if (active) {
        dv[i] += a1;
        dv[n + i] += a2;
        dv[2 * n + i] += a3;
    }
}

void ref64_grav(hipStream_t s, const double* x, double* dv, const double* ms, int64_t n,
                int64_t ib, int64_t ie, double G) {
    if (ie <= ib) return;
    hipLaunchKernelGGL(k_ref_grav, dim3(nbx_blocks(ie - ib, REF_TILE)), dim3(REF_TILE), 0, s,
                       x, dv, ms, n, ib, ie, G);
}

// sqrt((x^2 + y^2) + z^2)
__device__ __forceinline__ double ref_norm3(double x, double y, double z) {
    double s = x * x + y * y;
    return sqrt(s + z * z);
}

// ---------------------------------------------------------------------------
// Magnetic dipoles: src/basic_potentials.jl:333-365 (no boundary conditions, as the
// reference).  Per source j, in the reference's expression order:
//   rij4 = dot(rij, rij)^2, r = rij / norm(rij), mir = mi.r, mij = mj.r,
//   c_k = (((mi_k mij + mj_k mir) + r_k (mi.mj)) - ((5 r_k) mir) mij) / rij4
// then dv += (3 mu_4pi / m_i) * sum_j c.  mm: moments SoA (component k at k*n + j).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(REF_TILE)
k_ref_mag(const double* __restrict__ x, double* __restrict__ dv, const double* __restrict__ ms,
          const double* __restrict__ mm, int64_t n, int64_t ib, int64_t ie, double mu_4pi) {
    __shared__ double sx[REF_TILE], sy[REF_TILE], sz[REF_TILE];
    __shared__ double smx[REF_TILE], smy[REF_TILE], smz[REF_TILE];
    const int64_t i = ib + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = i < ie;
    double xi = 0.0, yi = 0.0, zi = 0.0, mix = 0.0, miy = 0.0, miz = 0.0;
    if (active) {
        xi = x[i]; yi = x[n + i]; zi = x[2 * n + i];
        mix = mm[i]; miy = mm[n + i]; miz = mm[2 * n + i];
    }
    double f1 = 0.0, f2 = 0.0, f3 = 0.0;
    for (int64_t base = 0; base < n; base += REF_TILE) {
        const int64_t j = base + threadIdx.x;
        if (j < n) {
            sx[threadIdx.x] = x[j]; sy[threadIdx.x] = x[n + j]; sz[threadIdx.x] = x[2 * n + j];
            smx[threadIdx.x] = mm[j]; smy[threadIdx.x] = mm[n + j]; smz[threadIdx.x] = mm[2 * n + j];
        }
        __syncthreads();
        const int lim = (int)((n - base) < REF_TILE ? (n - base) : REF_TILE);
        if (active) {
            for (int q = 0; q < lim; ++q) {
                if (base + q == i) continue;
                const double rx = xi - sx[q], ry = yi - sy[q], rz = zi - sz[q];
                const double d = (rx * rx + ry * ry) + rz * rz;
                const double rij4 = d * d;
                const double nr = ref_norm3(rx, ry, rz);
                const double r1 = rx / nr, r2 = ry / nr, r3 = rz / nr;
                const double mjx = smx[q], mjy = smy[q], mjz = smz[q];
                const double mir = (mix * r1 + miy * r2) + miz * r3;
                const double mij = (mjx * r1 + mjy * r2) + mjz * r3;
                const double dmm = (mix * mjx + miy * mjy) + miz * mjz;
                f1 += (((mix * mij + mjx * mir) + r1 * dmm) - ((5.0 * r1) * mir) * mij) / rij4;
                f2 += (((miy * mij + mjy * mir) + r2 * dmm) - ((5.0 * r2) * mir) * mij) / rij4;
                f3 += (((miz * mij + mjz * mir) + r3 * dmm) - ((5.0 * r3) * mir) * mij) / rij4;
            }
        }
        __syncthreads();
    }
    if (active) {
        const double coeff = 3.0 * mu_4pi / ms[i];
        dv[i] += coeff * f1;
        dv[n + i] += coeff * f2;
        dv[2 * n + i] += coeff * f3;
    }
}

void ref64_mag(hipStream_t s, const double* x, double* dv, const double* ms, const double* mm,
               int64_t n, int64_t ib, int64_t ie, double mu_4pi) {
    if (ie <= ib) return;
    hipLaunchKernelGGL(k_ref_mag, dim3(nbx_blocks(ie - ib, REF_TILE)), dim3(REF_TILE), 0, s, x,
                       dv, ms, mm, n, ib, ie, mu_4pi);
}

// ---------------------------------------------------------------------------
// SPC/Fw bonded terms: src/basic_potentials.jl:367-433
// ---------------------------------------------------------------------------

__global__ void k_ref_bond(const double* __restrict__ x, double* __restrict__ dv,
                           const double* __restrict__ ms, int64_t n, int64_t ib, int64_t ie,
                           double rOH, double kbond) {
    const int64_t i = ib + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ie) return;
    const int64_t o = 3 * (i / 3);
    int64_t nb[2];
    int nn;
    if (i == o) { nb[0] = o + 1; nb[1] = o + 2; nn = 2; }   // src/nbody_to_ode.jl:274-284
    else { nb[0] = o; nb[1] = o; nn = 1; }
    double f1 = 0.0, f2 = 0.0, f3 = 0.0;
    const double xi = x[i], yi = x[n + i], zi = x[2 * n + i];
    for (int t = 0; t < nn; ++t) {
        const int64_t j = nb[t];
        double rx = xi - x[j], ry = yi - x[n + j], rz = zi - x[2 * n + j];
        double r = ref_norm3(rx, ry, rz);
        double d = r - rOH;
        double factor = -d * kbond / r;
        f1 += factor * rx;
        f2 += factor * ry;
        f3 += factor * rz;
    }
    double coeff = 1.0 / ms[i];
    dv[i] += coeff * f1;
    dv[n + i] += coeff * f2;
    dv[2 * n + i] += coeff * f3;
}

void ref64_bond(hipStream_t s, const double* x, double* dv, const double* ms, int64_t n,
                int64_t ib, int64_t ie, double rOH, double kbond) {
    if (ie <= ib) return;
    hipLaunchKernelGGL(k_ref_bond, dim3(nbx_blocks(ie - ib, 256)), dim3(256), 0, s, x, dv, ms,
                       n, ib, ie, rOH, kbond);
}

__device__ __forceinline__ void ref_cross(double a0, double a1, double a2, double b0, double b1,
                                          double b2, double& c0, double& c1, double& c2) {
    c0 = a1 * b2 - a2 * b1;
    c1 = a2 * b0 - a0 * b2;
    c2 = a0 * b1 - a1 * b0;
}

// valence_angle_potential_acceleration!(dv, rs, a=H1, b=O, c=H2, ...) per molecule
__global__ void k_ref_angle(const double* __restrict__ x, double* __restrict__ dv,
                            const double* __restrict__ ms, int64_t n, int64_t mb, int64_t me,
                            double aHOH, double ka) {
    const int64_t m = mb + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= me) return;
    const int64_t a = 3 * m + 1, b = 3 * m, c = 3 * m + 2;
    double ra[3] = {x[a], x[n + a], x[2 * n + a]};
    double rb[3] = {x[b], x[n + b], x[2 * n + b]};
    double rc[3] = {x[c], x[n + c], x[2 * n + c]};
    double rba[3], rbc[3], rcb[3];
    for (int k = 0; k < 3; ++k) {
        rba[k] = ra[k] - rb[k];
        rbc[k] = rc[k] - rb[k];
        rcb[k] = rb[k] - rc[k];
    }
    double X[3], ta[3], tc[3];
    ref_cross(rba[0], rba[1], rba[2], rbc[0], rbc[1], rbc[2], X[0], X[1], X[2]);
    ref_cross(rba[0], rba[1], rba[2], X[0], X[1], X[2], ta[0], ta[1], ta[2]);
    ref_cross(rcb[0], rcb[1], rcb[2], X[0], X[1], X[2], tc[0], tc[1], tc[2]);
    double ia = 1.0 / ref_norm3(ta[0], ta[1], ta[2]);
    double ic = 1.0 / ref_norm3(tc[0], tc[1], tc[2]);
    double nba = ref_norm3(rba[0], rba[1], rba[2]);
    double nbc = ref_norm3(rbc[0], rbc[1], rbc[2]);
    double dt = rba[0] * rbc[0] + rba[1] * rbc[1];
    dt = dt + rba[2] * rbc[2];
    double cosine = dt / (nba * nbc);
    if (cosine > 1.0) cosine = 1.0;
    else if (cosine < -1.0) cosine = -1.0;
    double ang = acos(cosine);
    double force = -ka * (ang - aHOH);
    for (int k = 0; k < 3; ++k) {
        double pa = ia * ta[k], pc = ic * tc[k];
        double fa = pa * force / nba;
        double fc = pc * force / nbc;
        double fb = -(fa + fc);
        dv[k * n + a] += fa / ms[a];
        dv[k * n + b] += fb / ms[b];
        dv[k * n + c] += fc / ms[c];
    }
}

void ref64_angle(hipStream_t s, const double* x, double* dv, const double* ms, int64_t n,
                 int64_t mb, int64_t me, double aHOH, double ka) {
    if (me <= mb) return;
    hipLaunchKernelGGL(k_ref_angle, dim3(nbx_blocks(me - mb, 256)), dim3(256), 0, s, x, dv, ms,
                       n, mb, me, aHOH, ka);
}

// ---------------------------------------------------------------------------
// Temperature and Berendsen: src/thermostats.jl:76-91
// ---------------------------------------------------------------------------

// per-atom terms in parallel, then one thread sums them in index order
__global__ void k_ref_m_v2(const double* __restrict__ v, const double* __restrict__ ms,
                           int64_t n, double half_mass, double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a = v[i], b = v[n + i], c = v[2 * n + i];
    double s = a * a + b * b;
    s = s + c * c;
    if (half_mass != 0.0) out[i] = s * (ms[i] / 2.0);   // kinetic_energy term
    else out[i] = ms[i] * s;                             // md_temperature term
}

// mode 0: out = e; 1: out = e / c; 2: out = c * e
__global__ void k_seq_sum(const double* __restrict__ terms, int64_t count, int mode, double c,
                          double* __restrict__ out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double e = 0.0;
    for (int64_t i = 0; i < count; ++i) e += terms[i];
    out[0] = mode == 1 ? e / c : (mode == 2 ? c * e : e);
}

void ref64_temperature(hipStream_t s, const double* v, const double* ms, int64_t n,
                       double kb, int64_t Nc, double* d_tmp, double* d_T) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_ref_m_v2, dim3(nbx_blocks(n, 256)), dim3(256), 0, s, v, ms, n, 0.0,
                       d_tmp);
    hipLaunchKernelGGL(k_seq_sum, dim3(1), dim3(64), 0, s, d_tmp, n, 1,
                       kb * (double)(3 * n - Nc), d_T);
}

__global__ void k_ref_berendsen(double* __restrict__ dv, const double* __restrict__ v,
                                int64_t n3, const double* __restrict__ d_T, double T0,
                                double gamma) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n3) return;
    const double T = d_T[0];
    if (1.0 / T == __builtin_inf()) dv[t] += gamma * v[t];
    else {
        double c = gamma * (T0 / T - 1.0);
        dv[t] += c * v[t];
    }
}

void ref64_berendsen(hipStream_t s, double* dv, const double* v, int64_t n, const double* d_T,
                     double T0, double gamma) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_ref_berendsen, dim3(nbx_blocks(3 * n, 256)), dim3(256), 0, s, dv, v,
                       3 * n, d_T, T0, gamma);
}

// ---------------------------------------------------------------------------
// Velocity Verlet and stochastic thermostats
// ---------------------------------------------------------------------------

__global__ void k_ref_drift(double* __restrict__ x, const double* __restrict__ v,
                            const double* __restrict__ a, double* __restrict__ vh, int64_t n3,
                            double dt) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n3) return;
    const double half = 0.5, dtsq = dt * dt;
    x[t] = x[t] + dt * v[t] + dtsq * (half * a[t]);
    vh[t] = v[t] + dt * (half * a[t]);
}

__global__ void k_ref_kick(double* __restrict__ v, const double* __restrict__ vh,
                           const double* __restrict__ a, int64_t n3, double dt) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n3) return;
    v[t] = vh[t] + dt * (0.5 * a[t]);
}

void ref64_drift(hipStream_t s, double* x, const double* v, const double* a, double* vh,
                 int64_t n, double dt) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_ref_drift, dim3(nbx_blocks(3 * n, 256)), dim3(256), 0, s, x, v, a, vh,
                       3 * n, dt);
}

void ref64_kick(hipStream_t s, double* v, const double* vh, const double* a, int64_t n,
                double dt) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_ref_kick, dim3(nbx_blocks(3 * n, 256)), dim3(256), 0, s, v, vh, a,
                       3 * n, dt);
}

__global__ void k_ref_langevin_half(double* __restrict__ v, const double* __restrict__ ms,
                                    int64_t n, double kb, double T, double c1, double c2,
                                    uint64_t seed, int64_t step, uint32_t stream) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double sig = c2 * sqrt(kb * T / ms[i]);
    double g[3];
    nbx_rng_normals3(seed, i, step, stream, g);
    for (int k = 0; k < 3; ++k) v[k * n + i] = c1 * v[k * n + i] + sig * g[k];
}

void ref64_langevin_half(hipStream_t s, double* v, const double* ms, int64_t n, double kb,
                         double T, double gamma, double dt, uint64_t seed, int64_t step,
                         uint32_t stream) {
    if (n <= 0) return;
    const double c1 = exp(-gamma * (0.5 * dt));
    const double c2 = sqrt(1.0 - c1 * c1);
    hipLaunchKernelGGL(k_ref_langevin_half, dim3(nbx_blocks(n, 256)), dim3(256), 0, s, v, ms, n,
                       kb, T, c1, c2, seed, step, stream);
}

// Euler-Maruyama step of the Langevin SDE (src/nbody_to_ode.jl:567-598), one thread per
// atom, from the state (x, v) and a = f(x): x' = x + dt v and
// v' = v + dt (a - gamma v) + sigma (sqrt(dt) xi), sigma = sqrt(2 gamma kb T / m_1)
__global__ void k_ref_em(double* __restrict__ x, double* __restrict__ v,
                         const double* __restrict__ a, int64_t n, double dt, double gamma,
                         double sigma, double sqrt_dt, uint64_t seed, int64_t step) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double g[3];
    nbx_rng_normals3(seed, i, step, NBX_RNG_EM, g);
    for (int k = 0; k < 3; ++k) {
        const double vk = v[k * n + i];
        x[k * n + i] = x[k * n + i] + dt * vk;
        v[k * n + i] = vk + dt * (a[k * n + i] - gamma * vk) + sigma * (sqrt_dt * g[k]);
    }
}

void ref64_em(hipStream_t s, double* x, double* v, const double* a, int64_t n, double dt,
              double gamma, double sigma, uint64_t seed, int64_t step) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_ref_em, dim3(nbx_blocks(n, 256)), dim3(256), 0, s, x, v, a, n, dt, gamma,
                       sigma, sqrt(dt), seed, step);
}

// Andersen collisions (src/nbody_simulation_result.jl:511-540), one thread per body
__global__ void k_ref_andersen(double* __restrict__ v, const double* __restrict__ ms, int64_t n,
                               int water, double kb, double T, double prob, uint64_t seed,
                               int64_t step) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nb = water ? n / 3 : n;
    if (b >= nb) return;
    const int64_t first = water ? 3 * b : b;
    if (!(nbx_rng_uniform(seed, first, step, NBX_RNG_ANDERSEN) < prob)) return;
    const int natoms = water ? 3 : 1;
    for (int t = 0; t < natoms; ++t) {
        const int64_t i = first + t;
        const double m = water ? ms[i] : ms[0];
        const double vdev = sqrt(kb * T / m);
        double g[3];
        nbx_rng_normals3(seed, i, step, NBX_RNG_ANDERSEN, g);
        v[i] = vdev * g[0];
        v[n + i] = vdev * g[1];
        v[2 * n + i] = vdev * g[2];
    }
}

void ref64_andersen(hipStream_t s, double* v, const double* ms, int64_t n, int water,
                    double kb, double T, double prob, uint64_t seed, int64_t step) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_ref_andersen, dim3(nbx_blocks(n, 256)), dim3(256), 0, s, v, ms, n,
                       water, kb, T, prob, seed, step);
}

// ---------------------------------------------------------------------------
// Energies: src/nbody_simulation_result.jl:209-397
// ---------------------------------------------------------------------------

void ref64_kinetic(hipStream_t s, const double* v, const double* ms, int64_t n, double* d_tmp,
                   double* d_out) {
    if (n <= 0) { ref64_fill(s, d_out, 1, 0.0); return; }
    hipLaunchKernelGGL(k_ref_m_v2, dim3(nbx_blocks(n, 256)), dim3(256), 0, s, v, ms, n, 1.0,
                       d_tmp);
    hipLaunchKernelGGL(k_seq_sum, dim3(1), dim3(64), 0, s, d_tmp, n, 0, 0.0, d_out);
}

// LJ energy: per-i partial sums over the later atoms of the index set
__global__ void k_ref_lj_pe(const double* __restrict__ x, int64_t n, int jstride, int64_t nj,
                            double s2, double R2, nbx_bcv bc, double* __restrict__ out) {
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nj) return;
    const int64_t i = a * jstride;
    const double xi = x[i], yi = x[n + i], zi = x[2 * n + i];
    double e = 0.0;
    for (int64_t b = a + 1; b < nj; ++b) {
        const int64_t j = b * jstride;
        double rx, ry, rz, r2;
        nbx_displacement(xi, yi, zi, x[j], x[n + j], x[2 * n + j], bc, rx, ry, rz, r2);
        if (!(r2 < R2)) r2 = R2;                    // :308-310
        double q = s2 / r2;
        double s6 = q * q * q;
        double s12 = s6 * s6;
        e += (s12 - s6);
    }
    out[a] = e;
}

void ref64_lj_pe(hipStream_t s, const double* x, int64_t n, int jstride, int64_t nj, double eps,
                 double sigma, double R, const nbx_bcv& bc, double* d_tmp, double* d_out) {
    if (nj <= 0) { ref64_fill(s, d_out, 1, 0.0); return; }
    hipLaunchKernelGGL(k_ref_lj_pe, dim3(nbx_blocks(nj, 128)), dim3(128), 0, s, x, n, jstride,
                       nj, sigma * sigma, R * R, bc, d_tmp);
    hipLaunchKernelGGL(k_seq_sum, dim3(1), dim3(64), 0, s, d_tmp, nj, 0, 0.0, d_out);
}

// Coulomb energy (:321-351): e += (sum_j>i q_j / r) * q_i, then * k
__global__ void k_ref_elec_pe(const double* __restrict__ x, const double* __restrict__ qs,
                              int64_t n, int water, double R, double R2, int rf, double krf,
                              double crf, nbx_bcv bc, double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double xi = x[i], yi = x[n + i], zi = x[2 * n + i];
    double ei = 0.0;
    for (int64_t j = i + 1; j < n; ++j) {
        if (ref_excluded(i, j, water)) continue;
        double rx, ry, rz, r2;
        nbx_displacement(xi, yi, zi, x[j], x[n + j], x[2 * n + j], bc, rx, ry, rz, r2);
        if (rf) {
            if (r2 < R2) ei += qs[j] * (1.0 / sqrt(r2) + krf * r2 - crf);
        } else {
            if (r2 < R2) ei += qs[j] / sqrt(r2);
            else ei += qs[j] / R;
        }
    }
    out[i] = ei * qs[i];
}

void ref64_elec_pe(hipStream_t s, const double* x, const double* qs, int64_t n, int water,
                   double k, double R, int rf, double eps_rf, const nbx_bcv& bc,
                   double* d_tmp, double* d_out) {
    if (n <= 0) { ref64_fill(s, d_out, 1, 0.0); return; }
    const double krf = rf ? rf_krf(R, eps_rf) : 0.0;
    const double crf = rf ? 1.0 / R + krf * (R * R) : 0.0;
    hipLaunchKernelGGL(k_ref_elec_pe, dim3(nbx_blocks(n, 128)), dim3(128), 0, s, x, qs, n,
                       water, R, R * R, rf, krf, crf, bc, d_tmp);
    hipLaunchKernelGGL(k_seq_sum, dim3(1), dim3(64), 0, s, d_tmp, n, 0, 0.0, d_out);
}

__global__ void k_ref_grav_pe(const double* __restrict__ x, const double* __restrict__ ms,
                              int64_t n, double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double xi = x[i], yi = x[n + i], zi = x[2 * n + i];
    double ei = 0.0;
    for (int64_t j = i + 1; j < n; ++j)
        ei += ms[j] / ref_norm3(xi - x[j], yi - x[n + j], zi - x[2 * n + j]);
    out[i] = ei * ms[i];
}

void ref64_grav_pe(hipStream_t s, const double* x, const double* ms, int64_t n, double G,
                   double* d_tmp, double* d_out) {
    if (n <= 0) { ref64_fill(s, d_out, 1, 0.0); return; }
    hipLaunchKernelGGL(k_ref_grav_pe, dim3(nbx_blocks(n, 128)), dim3(128), 0, s, x, ms, n,
                       d_tmp);
    hipLaunchKernelGGL(k_seq_sum, dim3(1), dim3(64), 0, s, d_tmp, n, 2, -G, d_out);
}

// Bonds (:353-372, each bond seen from both ends, / 4) and angles (:374-397, / 2)
__global__ void k_ref_bonded_terms(const double* __restrict__ x, int64_t n, double rOH,
                                   double kbond, double aHOH, double ka,
                                   double* __restrict__ out) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nmol = n / 3;
    if (m >= nmol) return;
    const int64_t o = 3 * m;
    // bond terms in visit order: O->H1, O->H2, H1->O, H2->O
    const int64_t pi[4] = {o, o, o + 1, o + 2};
    const int64_t pj[4] = {o + 1, o + 2, o, o};
    for (int t = 0; t < 4; ++t) {
        double d = ref_norm3(x[pi[t]] - x[pj[t]], x[n + pi[t]] - x[n + pj[t]],
                             x[2 * n + pi[t]] - x[2 * n + pj[t]]) - rOH;
        out[4 * m + t] = d * d * kbond;
    }
    const int64_t a = o + 1, b = o, c = o + 2;
    double rba0 = x[a] - x[b], rba1 = x[n + a] - x[n + b], rba2 = x[2 * n + a] - x[2 * n + b];
    double rbc0 = x[c] - x[b], rbc1 = x[n + c] - x[n + b], rbc2 = x[2 * n + c] - x[2 * n + b];
    double dt = rba0 * rbc0 + rba1 * rbc1;
    dt = dt + rba2 * rbc2;
    double ang = acos(dt / (ref_norm3(rba0, rba1, rba2) * ref_norm3(rbc0, rbc1, rbc2)));
    double d = ang - aHOH;
    out[4 * nmol + m] = ka * (d * d);
}

__global__ void k_ref_bonded_sum(const double* __restrict__ terms, int64_t nmol,
                                 double* __restrict__ out2) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double eb = 0.0, ea = 0.0;
    for (int64_t t = 0; t < 4 * nmol; ++t) eb += terms[t];
    for (int64_t m = 0; m < nmol; ++m) ea += terms[4 * nmol + m];
    out2[0] = eb / 4.0;
    out2[1] = ea / 2.0;
}

void ref64_bonded_pe(hipStream_t s, const double* x, int64_t n, double rOH, double kbond,
                     double aHOH, double ka, double* d_tmp, double* d_out2) {
    const int64_t nmol = n / 3;
    if (nmol <= 0) { ref64_fill(s, d_out2, 2, 0.0); return; }
    hipLaunchKernelGGL(k_ref_bonded_terms, dim3(nbx_blocks(nmol, 256)), dim3(256), 0, s, x, n,
                       rOH, kbond, aHOH, ka, d_tmp);
    hipLaunchKernelGGL(k_ref_bonded_sum, dim3(1), dim3(64), 0, s, d_tmp, nmol, d_out2);
}

// ---------------------------------------------------------------------------
// RDF histogram (src/nbody_simulation_result.jl:664-693): frame in AoS (3 x n)
// ---------------------------------------------------------------------------

__global__ void k_ref_rdf(const double* __restrict__ f, int jstride, int64_t nj, double L,
                          double dr, int64_t maxbin, unsigned long long* __restrict__ hist) {
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nj) return;
    nbx_bcv bc;
    bc.kind = 1;
    bc.pad = 0;
    bc.b[0] = L;
    for (int k = 1; k < 6; ++k) bc.b[k] = 0.0;
    const double lim = (0.5 * L) * (0.5 * L);
    const int64_t i = a * jstride;
    const double xi = f[3 * i], yi = f[3 * i + 1], zi = f[3 * i + 2];
    for (int64_t b = a + 1; b < nj; ++b) {
        const int64_t j = b * jstride;
        double rx, ry, rz, r2;
        nbx_displacement(xi, yi, zi, f[3 * j], f[3 * j + 1], f[3 * j + 2], bc, rx, ry, rz, r2);
        if (r2 < lim) {
            double fb = ceil(sqrt(r2) / dr);
            if (fb > 1.0 && fb <= (double)maxbin) atomicAdd(&hist[(int64_t)fb], 2ull);
        }
    }
}

void ref64_rdf_hist(hipStream_t s, const double* frame_aos, int64_t n, int jstride, int64_t nj,
                    double L, int64_t maxbin, unsigned long long* d_hist) {
    (void)n;
    if (nj <= 0) return;
    hipLaunchKernelGGL(k_ref_rdf, dim3(nbx_blocks(nj, 128)), dim3(128), 0, s, frame_aos, jstride,
                       nj, L, L / (double)maxbin, maxbin, d_hist);
}

// Histogram of nbx_rdf_noalloc: library-owned device memory (bins 0..NBX_RDF_NOALLOC_MAXBIN,
// 128 KiB), one copy per GPU, held until the process ends.  Host code never passes the
// array's name to a kernel: it uses the address that hipGetSymbolAddress returned for the
// device (ref64_rdf_scratch).
static __device__ __align__(256) unsigned long long scratch[NBX_RDF_NOALLOC_MAXBIN + 1];

int ref64_rdf_scratch(int device, unsigned long long** out) {
    static std::mutex mu;
    static unsigned long long* ptr[NBX_MAX_DEVICES] = {};
    if (device < 0 || device >= NBX_MAX_DEVICES) return -1;
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[device]) {
        int cur = -1;
        if (hipGetDevice(&cur) != hipSuccess) return -1;
        if (cur != device && hipSetDevice(device) != hipSuccess) return -1;
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p) return -1;
        ptr[device] = static_cast<unsigned long long*>(p);
    }
    *out = ptr[device];
    return 0;
}

__global__ void k_fill_u64(unsigned long long* __restrict__ p, int64_t count) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < count) p[t] = 0ull;
}

void ref64_zero_u64(hipStream_t s, unsigned long long* p, int64_t count) {
    if (count <= 0) return;
    hipLaunchKernelGGL(k_fill_u64, dim3(nbx_blocks(count, 256)), dim3(256), 0, s, p, count);
}

__global__ void k_check_finite(const double* __restrict__ p, int64_t count, int* flag) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < count && !isfinite(p[t])) flag[0] = 1;
}

int64_t ref64_check_finite(hipStream_t s, const double* p, int64_t count, int* d_flag) {
    if (count <= 0) return 0;
    hipLaunchKernelGGL(k_check_finite, dim3(nbx_blocks(count, 256)), dim3(256), 0, s, p, count,
                       d_flag);
    return 0;
}

// ---------------------------------------------------------------------------
// Nose-Hoover thermostat: src/thermostats.jl:116-122.  The reference keeps zeta in the
// position part of an extra coordinate column and writes its rate into the velocity part
// on every right-hand side; here both live in device scalars.
// ---------------------------------------------------------------------------
__global__ void k_ref_nosehoover(double* __restrict__ dv, const double* __restrict__ v, int64_t n,
                                 const double* __restrict__ scal, double* __restrict__ xi_out,
                                 double T0, double tau, int64_t ndf) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const double zeta = scal[NBX_SCAL_ZETA];
    if (i < 3 * n) dv[i] -= zeta * v[i];               // @. dv -= u[zind] * v
    if (i == 0) {
        const double it = 1.0 / tau;                     // inv(p.tau)^2
        const double T = scal[NBX_SCAL_T];
        xi_out[0] = (it * it) * (T / T0 - (double)(ndf + 1) / (double)ndf);
    }
}

void ref64_nosehoover(hipStream_t s, double* dv, const double* v, int64_t n, const double* scal,
                      double* xi_out, double T0, double tau, int64_t ndf) {
    const int64_t cnt = 3 * n > 0 ? 3 * n : 1;
    hipLaunchKernelGGL(k_ref_nosehoover, dim3(nbx_blocks(cnt, 256)), dim3(256), 0, s, dv, v, n,
                       scal, xi_out, T0, tau, ndf);
}

__global__ void k_ref_nh_drift(double* scal, double dt) {
    if (threadIdx.x == 0 && blockIdx.x == 0)
        scal[NBX_SCAL_ZETA] = scal[NBX_SCAL_ZETA] + dt * scal[NBX_SCAL_XI];
}

void ref64_nh_drift(hipStream_t s, double* scal, double dt) {
    hipLaunchKernelGGL(k_ref_nh_drift, dim3(1), dim3(64), 0, s, scal, dt);
}
