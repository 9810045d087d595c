// nbx_nonbonded.hip -- mixed-precision non-bonded path for Lennard-Jones, Coulomb and
// reaction-field systems (potential systems and SPC/Fw water) on gfx950.
//
// Data layout in HBM
//   x, v, vh, a : FP64 SoA (component k of atom i at k*n + i), unwrapped positions
//   pos         : double4 per atom = position wrapped into [0, L)^3 and its charge;
//                 the pair displacement is taken in FP64 from these, minimum-imaged,
//                 then cast to FP32 for the pair law
//   nbr         : Verlet list, interleaved for G lanes per atom: slot k of atom i is
//                 at (k / G) * (n * G) + i * G + (k % G), so lane l of atom i reads
//                 consecutive words with the other lanes of the wavefront
//   cells       : fixed-capacity bins (count + ccap slots per cell), rebuilt with the
//                 list whenever an atom moved more than skin/2 since the last build
//                 (flag set by the drift kernel; every rebuild kernel exits at once
//                 when the flag is clear, so the step is a fixed kernel sequence)
//
// Pair law (per ordered pair, FP32): r2 = |r|^2; LJ (2 s^12 - s^6)/r2 * r with
// s^2 = sigma^2/r2; Coulomb q_j / r^3 * r; RF q_j (1/r^3 - 2 k_rf) * r.  Per-atom
// sums use Kahan compensation per lane; lanes are combined in FP64.
#include "nbx_mixed.h"

#define NB_THREADS 256

// ---------------------------------------------------------------------------
// Pair law
// ---------------------------------------------------------------------------

template <bool LJ, int EL>
__device__ __forceinline__ void nb_pair(float rx, float ry, float rz, float qj, bool lj_pair,
                                        const mx_pair_params& pp, nbx_kahan* acc) {
    const float r2 = rx * rx + ry * ry + rz * rz;
    if (LJ) {
        if (lj_pair && r2 < pp.lj_R2) {
            const float inv = __builtin_amdgcn_rcpf(r2);
            const float q = pp.lj_s2 * inv;
            const float s6 = q * q * q;
            const float f = (2.0f * s6 * s6 - s6) * inv;
            acc[0].add(f * rx);
            acc[1].add(f * ry);
            acc[2].add(f * rz);
        }
    }
    if (EL != 0) {
        if (r2 < pp.el_R2) {
            const float rinv = __builtin_amdgcn_rsqf(r2);
            const float r3 = rinv * rinv * rinv;
            const float f = (EL == 2) ? qj * (r3 - pp.krf2) : qj * r3;
            acc[3].add(f * rx);
            acc[4].add(f * ry);
            acc[5].add(f * rz);
        }
    }
}

// Combine per-lane sums of G lanes (G divides the wavefront) in FP64.
template <int G>
__device__ __forceinline__ double lane_sum(double x) {
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) x += __shfl_xor(x, off, G);
    return x;
}

// Final per-atom accumulation + fused end-of-step update (lane 0 of each atom)
template <bool LJ, int EL>
__device__ __forceinline__ void nb_finish(int64_t i, int64_t n, const double* F,
                                          const mx_pair_params& pp,
                                          const double* __restrict__ ms,
                                          const double* __restrict__ qs, double* __restrict__ a,
                                          double* __restrict__ v, const double* __restrict__ vh,
                                          const mx_kick_params& kp, bool water) {
    const double mi = ms[i];
    double ax = 0.0, ay = 0.0, az = 0.0;
    if (LJ) {
        const double c = pp.lj_c / mi;
        ax += c * F[0]; ay += c * F[1]; az += c * F[2];
    }
    if (EL != 0) {
        const double c = pp.el_k * qs[i] / mi;
        ax += c * F[3]; ay += c * F[4]; az += c * F[5];
    }
    if (water) {  // bonded terms and the step's end are done per molecule
        a[i] = ax; a[n + i] = ay; a[2 * n + i] = az;
        return;
    }
    if (kp.therm == NBX_THERM_BERENDSEN) {
        const double cb = mx_berendsen_coeff(kp);
        ax += cb * vh[i]; ay += cb * vh[n + i]; az += cb * vh[2 * n + i];
    }
    a[i] = ax; a[n + i] = ay; a[2 * n + i] = az;
    if (kp.do_kick) {
        mx_finish_atom(i, n, ax, ay, az, vh, v, ms, kp);
        if (kp.therm == NBX_THERM_ANDERSEN &&
            nbx_rng_uniform(kp.seed, i, kp.step, NBX_RNG_ANDERSEN) < kp.prob)
            mx_andersen_atom(i, n, kp.m0, v, kp);
    }
}

// ---------------------------------------------------------------------------
// Verlet-list force kernel: G lanes per atom
// ---------------------------------------------------------------------------

template <int G, bool LJ, int EL, bool WATER>
__global__ void __launch_bounds__(NB_THREADS)
k_nb_list(const double4* __restrict__ pos, const int* __restrict__ nbr,
          const int* __restrict__ nnb, int n, mx_pair_params pp,
          const double* __restrict__ ms, const double* __restrict__ qs,
          double* __restrict__ a, double* __restrict__ v, const double* __restrict__ vh,
          mx_kick_params kp, int* __restrict__ clear_flag) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (clear_flag != nullptr && gid == 0) clear_flag[0] = 0;
    const int i = gid / G;
    const int lane = gid & (G - 1);
    if (i >= n) return;
    const double4 pi = pos[i];
    const int cnt = nnb[i];
    const bool io = !WATER || (i % 3) == 0;
    nbx_kahan acc[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) acc[q].init();
    const size_t stride = (size_t)n * G;
    int k = lane;
    const int* p = nbr + gid;
    for (; k < cnt; k += G, p += stride) {
        const int j = *p;
        const double4 pj = pos[j];
        double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        dx -= pp.L * rint(dx * pp.invL);
        dy -= pp.L * rint(dy * pp.invL);
        dz -= pp.L * rint(dz * pp.invL);
        const bool ljp = io && (!WATER || (j % 3) == 0);
        nb_pair<LJ, EL>((float)dx, (float)dy, (float)dz, (float)pj.w, ljp, pp, acc);
    }
    double F[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) F[q] = lane_sum<G>(acc[q].value());
    if (lane == 0) nb_finish<LJ, EL>(i, n, F, pp, ms, qs, a, v, vh, kp, WATER);
}

// ---------------------------------------------------------------------------
// All-pairs force kernel (any boundary condition, any cutoff): one thread per atom,
// sources staged through LDS, displacement by the reference's boundary rule in FP64.
// ---------------------------------------------------------------------------

#define AP_TILE 256

template <bool LJ, int EL, bool WATER>
__global__ void __launch_bounds__(AP_TILE)
k_nb_allpairs(const double* __restrict__ x, int n, mx_pair_params pp, nbx_bcv bc,
              const double* __restrict__ ms, const double* __restrict__ qs,
              double* __restrict__ a, double* __restrict__ v, const double* __restrict__ vh,
              mx_kick_params kp) {
    __shared__ double sx[AP_TILE], sy[AP_TILE], sz[AP_TILE];
    __shared__ float sq[AP_TILE];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = i < n;
    double xi = 0.0, yi = 0.0, zi = 0.0;
    if (active) { xi = x[i]; yi = x[n + i]; zi = x[2 * n + i]; }
    const bool io = !WATER || (i % 3) == 0;
    nbx_kahan acc[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) acc[q].init();
    for (int base = 0; base < n; base += AP_TILE) {
        const int j = base + threadIdx.x;
        if (j < n) {
            sx[threadIdx.x] = x[j]; sy[threadIdx.x] = x[n + j]; sz[threadIdx.x] = x[2 * n + j];
            sq[threadIdx.x] = (EL != 0) ? (float)qs[j] : 0.0f;
        }
        __syncthreads();
        const int lim = (n - base) < AP_TILE ? (n - base) : AP_TILE;
        if (active) {
            for (int q = 0; q < lim; ++q) {
                const int jj = base + q;
                if (WATER ? (jj / 3 == i / 3) : (jj == i)) continue;
                double rx, ry, rz, r2;
                nbx_displacement(xi, yi, zi, sx[q], sy[q], sz[q], bc, rx, ry, rz, r2);
                const bool ljp = io && (!WATER || (jj % 3) == 0);
                nb_pair<LJ, EL>((float)rx, (float)ry, (float)rz, sq[q], ljp, pp, acc);
            }
        }
        __syncthreads();
    }
    if (!active) return;
    double F[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) F[q] = acc[q].value();
    nb_finish<LJ, EL>(i, n, F, pp, ms, qs, a, v, vh, kp, WATER);
}

// ---------------------------------------------------------------------------
// Dispatch over the template parameters
// ---------------------------------------------------------------------------

template <int G, bool LJ, int EL, bool WATER>
static void launch_list(hipStream_t s, const nbx_ctx* c, const mx_pair_params& pp,
                        const mx_kick_params& kp, int* clear_flag) {
    const int64_t threads = c->n * G;
    hipLaunchKernelGGL((k_nb_list<G, LJ, EL, WATER>), dim3(nbx_blocks(threads, NB_THREADS)),
                       dim3(NB_THREADS), 0, s, (const double4*)c->d_pos, c->nl.d_nbr,
                       c->nl.d_nnb, (int)c->n, pp, c->d_ms, c->d_qs, c->d_a, c->d_v, c->d_vh,
                       kp, clear_flag);
}

template <bool LJ, int EL, bool WATER>
static void launch_list_g(hipStream_t s, int G, const nbx_ctx* c, const mx_pair_params& pp,
                          const mx_kick_params& kp, int* clear_flag) {
    switch (G) {
    case 1: launch_list<1, LJ, EL, WATER>(s, c, pp, kp, clear_flag); break;
    case 2: launch_list<2, LJ, EL, WATER>(s, c, pp, kp, clear_flag); break;
    case 4: launch_list<4, LJ, EL, WATER>(s, c, pp, kp, clear_flag); break;
    case 16: launch_list<16, LJ, EL, WATER>(s, c, pp, kp, clear_flag); break;
    default: launch_list<8, LJ, EL, WATER>(s, c, pp, kp, clear_flag); break;
    }
}

template <bool LJ, int EL, bool WATER>
static void launch_nb(hipStream_t s, const mx_nb_launch& L, const nbx_ctx* c,
                      const mx_pair_params& pp, const mx_kick_params& kp, int* clear_flag) {
    if (L.mode_list) {
        launch_list_g<LJ, EL, WATER>(s, L.lanes, c, pp, kp, clear_flag);
    } else {
        hipLaunchKernelGGL((k_nb_allpairs<LJ, EL, WATER>), dim3(nbx_blocks(c->n, AP_TILE)),
                           dim3(AP_TILE), 0, s, c->d_x, (int)c->n, pp, c->bc, c->d_ms,
                           c->d_qs, c->d_a, c->d_v, c->d_vh, kp);
    }
}

void mx_launch_nonbonded(hipStream_t s, const mx_nb_launch& L, const nbx_ctx* c,
                         const mx_pair_params& pp, const mx_kick_params& kp,
                         int* clear_flag) {
    if (c->n <= 0) return;
    const int key = (L.has_lj ? 1 : 0) | (L.el_kind << 1) | (L.water ? 8 : 0);
    switch (key) {
    case 1: launch_nb<true, 0, false>(s, L, c, pp, kp, clear_flag); break;
    case 2: launch_nb<false, 1, false>(s, L, c, pp, kp, clear_flag); break;
    case 3: launch_nb<true, 1, false>(s, L, c, pp, kp, clear_flag); break;
    case 4: launch_nb<false, 2, false>(s, L, c, pp, kp, clear_flag); break;
    case 5: launch_nb<true, 2, false>(s, L, c, pp, kp, clear_flag); break;
    case 1 | 2 | 8: launch_nb<true, 1, true>(s, L, c, pp, kp, clear_flag); break;
    case 1 | 4 | 8: launch_nb<true, 2, true>(s, L, c, pp, kp, clear_flag); break;
    case 2 | 8: launch_nb<false, 1, true>(s, L, c, pp, kp, clear_flag); break;
    case 4 | 8: launch_nb<false, 2, true>(s, L, c, pp, kp, clear_flag); break;
    case 1 | 8: launch_nb<true, 0, true>(s, L, c, pp, kp, clear_flag); break;
    default: break;  // no non-bonded term (e.g. bonded-only water)
    }
}

// ---------------------------------------------------------------------------
// Drift: x' = x + dt v + dt^2 (a/2), vh = v + dt (a/2), wrapped copy, skin check
// ---------------------------------------------------------------------------

// write_pos_kind: 0 none, 1 double4 wrapped (+charge), 2 float4 (+mass) for gravity
__global__ void k_mx_drift(double* __restrict__ x, double* __restrict__ v,
                           const double* __restrict__ a, double* __restrict__ vh, int n,
                           double dt, int therm, double c1, double c2, double T, double kb,
                           const double* __restrict__ ms, const double* __restrict__ qs,
                           uint64_t seed, int64_t step, int write_pos_kind, void* pos,
                           double L, const double* __restrict__ xref, double half_skin2,
                           int* __restrict__ flags, int* __restrict__ cell_count, int ncells) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell_count != nullptr && i < ncells) cell_count[i] = 0;
    if (i >= n) return;
    double vx = v[i], vy = v[n + i], vz = v[2 * n + i];
    if (therm == NBX_THERM_LANGEVIN) {
        const double sig = c2 * sqrt(kb * T / ms[i]);
        double g[3];
        nbx_rng_normals3(seed, i, step, NBX_RNG_LANGEVIN_A, g);
        vx = c1 * vx + sig * g[0];
        vy = c1 * vy + sig * g[1];
        vz = c1 * vz + sig * g[2];
        v[i] = vx; v[n + i] = vy; v[2 * n + i] = vz;
    }
    const double dtsq = dt * dt;
    const double ax = a[i], ay = a[n + i], az = a[2 * n + i];
    const double px = x[i] + dt * vx + dtsq * (0.5 * ax);
    const double py = x[n + i] + dt * vy + dtsq * (0.5 * ay);
    const double pz = x[2 * n + i] + dt * vz + dtsq * (0.5 * az);
    x[i] = px; x[n + i] = py; x[2 * n + i] = pz;
    vh[i] = vx + dt * (0.5 * ax);
    vh[n + i] = vy + dt * (0.5 * ay);
    vh[2 * n + i] = vz + dt * (0.5 * az);
    if (write_pos_kind == 1) {
        double wx = px - L * floor(px / L), wy = py - L * floor(py / L),
               wz = pz - L * floor(pz / L);
        if (wx >= L) wx -= L;
        if (wy >= L) wy -= L;
        if (wz >= L) wz -= L;
        ((double4*)pos)[i] = make_double4(wx, wy, wz, qs != nullptr ? qs[i] : 0.0);
    } else if (write_pos_kind == 2) {
        ((float4*)pos)[i] = make_float4((float)px, (float)py, (float)pz, (float)ms[i]);
    }
    if (xref != nullptr) {
        const double ex = px - xref[i], ey = py - xref[n + i], ez = pz - xref[2 * n + i];
        if (ex * ex + ey * ey + ez * ez > half_skin2) flags[0] = 1;
    }
}

void mx_launch_drift(hipStream_t s, const nbx_ctx* c, double dt, int therm, double c1,
                     double c2, double T, double kb, uint64_t seed, int64_t step,
                     int write_pos_kind, double half_skin2) {
    const bool list = c->nl.ready && write_pos_kind == 1;
    const int ncells = list ? c->nl.ncell * c->nl.ncell * c->nl.ncell : 0;
    const int64_t work = c->n > ncells ? c->n : ncells;
    hipLaunchKernelGGL(k_mx_drift, dim3(nbx_blocks(work, 256)), dim3(256), 0, s, c->d_x,
                       c->d_v, c->d_a, c->d_vh, (int)c->n, dt, therm, c1, c2, T, kb, c->d_ms,
                       c->has_q ? c->d_qs : nullptr, seed, step, write_pos_kind, c->d_pos,
                       c->bc.b[0], list ? c->nl.d_xref : nullptr, half_skin2, c->nl.d_flags,
                       list ? c->nl.d_cell_count : nullptr, ncells);
}

// positions -> wrapped copy (after set_state) and a forced rebuild
__global__ void k_mx_positions(const double* __restrict__ x, int n, int kind, void* pos,
                               double L, const double* __restrict__ ms,
                               const double* __restrict__ qs, int* flags,
                               int* __restrict__ cell_count, int ncells) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0 && flags != nullptr) flags[0] = 1;
    if (cell_count != nullptr && i < ncells) cell_count[i] = 0;
    if (i >= n) return;
    const double px = x[i], py = x[n + i], pz = x[2 * n + i];
    if (kind == 1) {
        double wx = px - L * floor(px / L), wy = py - L * floor(py / L),
               wz = pz - L * floor(pz / L);
        if (wx >= L) wx -= L;
        if (wy >= L) wy -= L;
        if (wz >= L) wz -= L;
        ((double4*)pos)[i] = make_double4(wx, wy, wz, qs != nullptr ? qs[i] : 0.0);
    } else if (kind == 2) {
        ((float4*)pos)[i] = make_float4((float)px, (float)py, (float)pz, (float)ms[i]);
    }
}

void mx_launch_positions(hipStream_t s, const nbx_ctx* c, int kind) {
    const bool list = c->nl.ready && kind == 1;
    const int ncells = list ? c->nl.ncell * c->nl.ncell * c->nl.ncell : 0;
    const int64_t work = c->n > ncells ? c->n : ncells;
    if (work <= 0) return;
    hipLaunchKernelGGL(k_mx_positions, dim3(nbx_blocks(work, 256)), dim3(256), 0, s, c->d_x,
                       (int)c->n, kind, c->d_pos, c->bc.b[0], c->d_ms,
                       c->has_q ? c->d_qs : nullptr, list ? c->nl.d_flags : nullptr,
                       list ? c->nl.d_cell_count : nullptr, ncells);
}

// ---------------------------------------------------------------------------
// Cell list and Verlet list (gated by flags[0])
// ---------------------------------------------------------------------------

__device__ __forceinline__ int cell_coord(double w, double inv_cs, int nc) {
    int cx = (int)(w * inv_cs);
    return cx < 0 ? 0 : (cx >= nc ? nc - 1 : cx);
}

__global__ void k_cells_bin(const double4* __restrict__ pos, int n, int nc, double inv_cs,
                            int* __restrict__ count, int* __restrict__ atoms, int ccap,
                            int* __restrict__ flags) {
    if (flags[0] == 0) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 p = pos[i];
    const int cid = (cell_coord(p.z, inv_cs, nc) * nc + cell_coord(p.y, inv_cs, nc)) * nc +
                    cell_coord(p.x, inv_cs, nc);
    const int slot = atomicAdd(&count[cid], 1);
    if (slot < ccap) atoms[(size_t)cid * ccap + slot] = i;
    else flags[1] = 1;
}

// sort each bin by atom index so that list order (and summation order) is reproducible
__global__ void k_cells_sort(int* __restrict__ count, int* __restrict__ atoms, int ncells,
                             int ccap, const int* __restrict__ flags) {
    if (flags[0] == 0) return;
    const int cid = blockIdx.x * blockDim.x + threadIdx.x;
    if (cid >= ncells) return;
    int m = count[cid];
    if (m > ccap) m = ccap;
    int* a = atoms + (size_t)cid * ccap;
    for (int p = 1; p < m; ++p) {
        const int key = a[p];
        int q = p - 1;
        while (q >= 0 && a[q] > key) { a[q + 1] = a[q]; --q; }
        a[q + 1] = key;
    }
}

template <bool WATER>
__global__ void k_nl_build(const double4* __restrict__ pos, const double* __restrict__ x,
                           int n, int nc, double inv_cs, double L, double invL,
                           const int* __restrict__ count, const int* __restrict__ atoms,
                           int ccap, double rlist2, int G, int cap, int* __restrict__ nbr,
                           int* __restrict__ nnb, double* __restrict__ xref,
                           int* __restrict__ flags) {
    if (flags[0] == 0) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (i == 0) atomicAdd(&flags[2], 1);
    const double4 pi = pos[i];
    const size_t stride = (size_t)n * G;
    int c = 0;
    auto consider = [&](int j) {
        if (j == i) return;
        if (WATER && (j / 3) == (i / 3)) return;
        const double4 pj = pos[j];
        double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        dx -= L * rint(dx * invL);
        dy -= L * rint(dy * invL);
        dz -= L * rint(dz * invL);
        if (dx * dx + dy * dy + dz * dz < rlist2) {
            if (c < cap) nbr[(size_t)(c / G) * stride + (size_t)i * G + (c % G)] = j;
            ++c;
        }
    };
    if (nc >= 3) {
        const int cx = cell_coord(pi.x, inv_cs, nc), cy = cell_coord(pi.y, inv_cs, nc),
                  cz = cell_coord(pi.z, inv_cs, nc);
        for (int oz = -1; oz <= 1; ++oz) {
            const int zz = (cz + oz + nc) % nc;
            for (int oy = -1; oy <= 1; ++oy) {
                const int yy = (cy + oy + nc) % nc;
                for (int ox = -1; ox <= 1; ++ox) {
                    const int xx = (cx + ox + nc) % nc;
                    const int cid = (zz * nc + yy) * nc + xx;
                    int m = count[cid];
                    if (m > ccap) m = ccap;
                    const int* a = atoms + (size_t)cid * ccap;
                    for (int s = 0; s < m; ++s) consider(a[s]);
                }
            }
        }
    } else {
        for (int j = 0; j < n; ++j) consider(j);
    }
    if (c > cap) { flags[1] = 1; c = cap; }
    nnb[i] = c;
    xref[i] = x[i];
    xref[n + i] = x[n + i];
    xref[2 * n + i] = x[2 * n + i];
}

void mx_launch_rebuild(hipStream_t s, const nbx_ctx* c, int water) {
    const nbx_nlist& nl = c->nl;
    const int n = (int)c->n;
    const double L = c->bc.b[0];
    const int nc = nl.ncell;
    const double inv_cs = nc > 0 ? (double)nc / L : 0.0;
    if (nc >= 3) {
        hipLaunchKernelGGL(k_cells_bin, dim3(nbx_blocks(n, 256)), dim3(256), 0, s,
                           (const double4*)c->d_pos, n, nc, inv_cs, nl.d_cell_count,
                           nl.d_cell_atoms, nl.ccap, nl.d_flags);
        const int ncells = nc * nc * nc;
        hipLaunchKernelGGL(k_cells_sort, dim3(nbx_blocks(ncells, 64)), dim3(64), 0, s,
                           nl.d_cell_count, nl.d_cell_atoms, ncells, nl.ccap, nl.d_flags);
    }
    const double r2 = nl.rlist * nl.rlist;
    if (water)
        hipLaunchKernelGGL(k_nl_build<true>, dim3(nbx_blocks(n, 128)), dim3(128), 0, s,
                           (const double4*)c->d_pos, c->d_x, n, nc, inv_cs, L, 1.0 / L,
                           nl.d_cell_count, nl.d_cell_atoms, nl.ccap, r2, nl.lanes, nl.cap,
                           nl.d_nbr, nl.d_nnb, nl.d_xref, nl.d_flags);
    else
        hipLaunchKernelGGL(k_nl_build<false>, dim3(nbx_blocks(n, 128)), dim3(128), 0, s,
                           (const double4*)c->d_pos, c->d_x, n, nc, inv_cs, L, 1.0 / L,
                           nl.d_cell_count, nl.d_cell_atoms, nl.ccap, r2, nl.lanes, nl.cap,
                           nl.d_nbr, nl.d_nnb, nl.d_xref, nl.d_flags);
}

// ---------------------------------------------------------------------------
// SPC/Fw molecule kernel: bonds + angle (FP64), Berendsen, end of step.
// One thread per molecule (O, H1, H2 = 3m, 3m+1, 3m+2); the non-bonded accelerations
// are already in a.  Formulas of src/basic_potentials.jl:367-433.
// ---------------------------------------------------------------------------

__device__ __forceinline__ void mol_bond(const double* ri, const double* rj, double rOH,
                                         double kbond, double* f) {
    const double rx = ri[0] - rj[0], ry = ri[1] - rj[1], rz = ri[2] - rj[2];
    const double r = sqrt(rx * rx + ry * ry + rz * rz);
    const double fac = -(r - rOH) * kbond / r;
    f[0] += fac * rx;
    f[1] += fac * ry;
    f[2] += fac * rz;
}

__global__ void k_water_molecules(const double* __restrict__ x, double* __restrict__ a,
                                  double* __restrict__ v, const double* __restrict__ vh,
                                  const double* __restrict__ ms, int n, double rOH,
                                  double kbond, double aHOH, double ka, mx_kick_params kp) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n / 3) return;
    const int o = 3 * m;
    double r[3][3];
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        r[t][0] = x[o + t];
        r[t][1] = x[n + o + t];
        r[t][2] = x[2 * n + o + t];
    }
    double f[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    mol_bond(r[0], r[1], rOH, kbond, f[0]);
    mol_bond(r[0], r[2], rOH, kbond, f[0]);
    mol_bond(r[1], r[0], rOH, kbond, f[1]);
    mol_bond(r[2], r[0], rOH, kbond, f[2]);
    double acc[3][3];
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        const double im = 1.0 / ms[o + t];
        for (int k = 0; k < 3; ++k) acc[t][k] = a[k * n + o + t] + im * f[t][k];
    }
    // angle H1-O-H2
    double rba[3], rbc[3], rcb[3];
    for (int k = 0; k < 3; ++k) {
        rba[k] = r[1][k] - r[0][k];
        rbc[k] = r[2][k] - r[0][k];
        rcb[k] = r[0][k] - r[2][k];
    }
    double X[3] = {rba[1] * rbc[2] - rba[2] * rbc[1], rba[2] * rbc[0] - rba[0] * rbc[2],
                   rba[0] * rbc[1] - rba[1] * rbc[0]};
    double ta[3] = {rba[1] * X[2] - rba[2] * X[1], rba[2] * X[0] - rba[0] * X[2],
                    rba[0] * X[1] - rba[1] * X[0]};
    double tc[3] = {rcb[1] * X[2] - rcb[2] * X[1], rcb[2] * X[0] - rcb[0] * X[2],
                    rcb[0] * X[1] - rcb[1] * X[0]};
    const double ia = 1.0 / sqrt(ta[0] * ta[0] + ta[1] * ta[1] + ta[2] * ta[2]);
    const double ic = 1.0 / sqrt(tc[0] * tc[0] + tc[1] * tc[1] + tc[2] * tc[2]);
    const double nba = sqrt(rba[0] * rba[0] + rba[1] * rba[1] + rba[2] * rba[2]);
    const double nbc = sqrt(rbc[0] * rbc[0] + rbc[1] * rbc[1] + rbc[2] * rbc[2]);
    double cosine = (rba[0] * rbc[0] + rba[1] * rbc[1] + rba[2] * rbc[2]) / (nba * nbc);
    cosine = cosine > 1.0 ? 1.0 : (cosine < -1.0 ? -1.0 : cosine);
    const double force = -ka * (acos(cosine) - aHOH);
    for (int k = 0; k < 3; ++k) {
        const double fa = ia * ta[k] * force / nba;
        const double fc = ic * tc[k] * force / nbc;
        acc[1][k] += fa / ms[o + 1];
        acc[0][k] += -(fa + fc) / ms[o];
        acc[2][k] += fc / ms[o + 2];
    }
    if (kp.therm == NBX_THERM_BERENDSEN) {
        const double cb = mx_berendsen_coeff(kp);
        for (int t = 0; t < 3; ++t)
            for (int k = 0; k < 3; ++k) acc[t][k] += cb * vh[k * n + o + t];
    }
    for (int t = 0; t < 3; ++t)
        for (int k = 0; k < 3; ++k) a[k * n + o + t] = acc[t][k];
    if (!kp.do_kick) return;
    for (int t = 0; t < 3; ++t) mx_finish_atom(o + t, n, acc[t][0], acc[t][1], acc[t][2], vh, v,
                                               ms, kp);
    if (kp.therm == NBX_THERM_ANDERSEN &&
        nbx_rng_uniform(kp.seed, o, kp.step, NBX_RNG_ANDERSEN) < kp.prob) {
        for (int t = 0; t < 3; ++t) mx_andersen_atom(o + t, n, ms[o + t], v, kp);
    }
}

void mx_launch_water_molecules(hipStream_t s, const nbx_ctx* c, const mx_kick_params& kp) {
    const int64_t nmol = c->n / 3;
    if (nmol <= 0) return;
    hipLaunchKernelGGL(k_water_molecules, dim3(nbx_blocks(nmol, 128)), dim3(128), 0, s, c->d_x,
                       c->d_a, c->d_v, c->d_vh, c->d_ms, (int)c->n, c->rOH, c->kbond, c->aHOH,
                       c->kangle, kp);
}

// ---------------------------------------------------------------------------
// Temperature of a velocity array (Berendsen input): per-block partial sums of
// m v^2, then one block adds them in block order.
// ---------------------------------------------------------------------------

__global__ void k_mv2_partial(const double* __restrict__ v, const double* __restrict__ ms,
                              int n, double* __restrict__ part) {
    __shared__ double red[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0;
    if (i < n) {
        const double a = v[i], b = v[n + i], c = v[2 * n + i];
        e = ms[i] * (a * a + b * b + c * c);
    }
    red[threadIdx.x] = e;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_mv2_final(const double* __restrict__ part, int nparts, double denom,
                            double* __restrict__ out) {
    __shared__ double red[256];
    double e = 0.0;
    for (int p = threadIdx.x; p < nparts; p += 256) e += part[p];
    red[threadIdx.x] = e;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0] / denom;
}

void mx_launch_temperature(hipStream_t s, const nbx_ctx* c, const double* v, int64_t Nc,
                           double* d_T) {
    const int n = (int)c->n;
    const int nb = (int)nbx_blocks(n, 256);
    hipLaunchKernelGGL(k_mv2_partial, dim3(nb), dim3(256), 0, s, v, c->d_ms, n, c->d_red);
    hipLaunchKernelGGL(k_mv2_final, dim3(1), dim3(256), 0, s, c->d_red, nb,
                       c->kb * (double)(3 * c->n - Nc), d_T);
}

// ---------------------------------------------------------------------------
// In-cutoff ordered pair count (the pair-interaction metric)
// ---------------------------------------------------------------------------

template <bool WATER>
__global__ void k_count_pairs_list(const double4* __restrict__ pos, const int* __restrict__ nbr,
                                   const int* __restrict__ nnb, int n, int G, double L,
                                   double invL, double R2lj, double R2el, int has_lj,
                                   int has_el, unsigned long long* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 pi = pos[i];
    const size_t stride = (size_t)n * G;
    unsigned long long cnt = 0;
    for (int k = 0; k < nnb[i]; ++k) {
        const int j = nbr[(size_t)(k / G) * stride + (size_t)i * G + (k % G)];
        const double4 pj = pos[j];
        double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        dx -= L * rint(dx * invL);
        dy -= L * rint(dy * invL);
        dz -= L * rint(dz * invL);
        const float rx = (float)dx, ry = (float)dy, rz = (float)dz;
        const float r2 = rx * rx + ry * ry + rz * rz;
        const bool ljp = has_lj && (!WATER || ((i % 3) == 0 && (j % 3) == 0)) &&
                         r2 < (float)R2lj;
        const bool elp = has_el && r2 < (float)R2el;
        if (ljp || elp) ++cnt;
    }
    atomicAdd(out, cnt);
}

void mx_launch_count_pairs(hipStream_t s, const nbx_ctx* c, const mx_nb_launch& L,
                           const mx_pair_params& pp, unsigned long long* d_count) {
    const int n = (int)c->n;
    if (L.water)
        hipLaunchKernelGGL(k_count_pairs_list<true>, dim3(nbx_blocks(n, 256)), dim3(256), 0, s,
                           (const double4*)c->d_pos, c->nl.d_nbr, c->nl.d_nnb, n, c->nl.lanes,
                           pp.L, pp.invL, (double)pp.lj_R2, (double)pp.el_R2, L.has_lj,
                           L.el_kind != 0, d_count);
    else
        hipLaunchKernelGGL(k_count_pairs_list<false>, dim3(nbx_blocks(n, 256)), dim3(256), 0, s,
                           (const double4*)c->d_pos, c->nl.d_nbr, c->nl.d_nnb, n, c->nl.lanes,
                           pp.L, pp.invL, (double)pp.lj_R2, (double)pp.el_R2, L.has_lj,
                           L.el_kind != 0, d_count);
}
