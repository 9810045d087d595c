// nbx_batch.hip -- batched Lennard-Jones replicas in one persistent launch (gfx950).
//
// One workgroup of BATCH_THREADS threads owns one replica for every step of a launch:
//   - the replica's positions, wrapped into [0, L) and rounded to FP32, live in LDS
//     (BATCH_NMAX float4 = 64 KB), so the pair loop never reads HBM for positions;
//   - x, v, a, vh and the list reference positions stay in FP64 in HBM (SoA over the
//     whole batch: component k of global atom g at k * ntot + g); a thread owns atoms
//     tid, tid + BATCH_THREADS, ... in every phase, so it reads back its own writes;
//   - LDS positions are kept in cell order ("slots"); the Verlet list (entry k of local
//     atom i at nbr[list_off + k * n + i]) holds slots.  It is rebuilt inside the launch
//     whenever an atom of the replica moved more than skin/2 since the last build: atoms
//     are binned into an LDS cell list (cells >= list radius), placed, rank-sorted by
//     index within each cell (reproducible lists), and every atom scans the 27
//     neighbouring cells; replicas with fewer than 3 cells per side scan all atoms;
//   - per step: [Langevin A] drift -> barrier -> [rebuild -> barrier] -> forces (FP32,
//     Kahan), kick, [Langevin B | Andersen] -> barrier.  Only workgroup barriers: the
//     replicas never synchronise with each other, so nothing crosses an XCD.
// Blocks are ordered by decreasing replica size (longest first).
#include "nbx_internal.h"

#define BATCH_THREADS 512
#define BATCH_NMAX 4096
#define BATCH_MAXCELLS 512

struct nbx_batch_rep {
    int64_t off;       // first global atom
    int64_t list_off;  // first list slot
    int32_t n;
    int32_t therm;
    int32_t nc;        // cells per side of the in-kernel cell list (< 3: all pairs)
    int32_t pad;
    double inv_cs;     // nc / L
    double L;
    double T, param;
    double c1, c2, prob;  // per-dt thermostat coefficients (set by the host per run)
    double m0;
    uint64_t key;         // Philox key of the replica
};

struct nbx_batch_params {
    float s2, R2, rlist2f;
    double c24eps;
    double half_skin2;
    double dt, kb;
    int cap;
    int64_t ntot;
};

__device__ __forceinline__ float4 batch_wrap(double x, double y, double z, double L) {
    double wx = x - L * floor(x / L), wy = y - L * floor(y / L), wz = z - L * floor(z / L);
    if (wx >= L) wx -= L;
    if (wy >= L) wy -= L;
    if (wz >= L) wz -= L;
    return make_float4((float)wx, (float)wy, (float)wz, 0.0f);
}

__global__ void __launch_bounds__(BATCH_THREADS)
k_batch_steps(const nbx_batch_rep* __restrict__ reps, const int* __restrict__ order,
              double* __restrict__ x, double* __restrict__ v, double* __restrict__ a,
              double* __restrict__ vh, double* __restrict__ xref, int* __restrict__ nbr,
              int* __restrict__ nnb, const double* __restrict__ ms, int* __restrict__ slot_of,
              int* __restrict__ cell_of, int* __restrict__ tmp_idx, nbx_batch_params P,
              const int64_t* __restrict__ step_base, int nsteps, int prologue,
              int* __restrict__ flags) {
    __shared__ float4 spos[BATCH_NMAX];           // positions by slot (cell order)
    __shared__ int ccount[BATCH_MAXCELLS + 1];    // atoms per cell, then fill counters
    __shared__ int cstart[BATCH_MAXCELLS + 1];    // first slot of each cell
    __shared__ int rebuild;
    const nbx_batch_rep rep = reps[order[blockIdx.x]];
    const int n = rep.n;
    const int tid = threadIdx.x;
    const int64_t nt = P.ntot, off = rep.off;
    double* X = x + off;
    double* V = v + off;
    double* A = a + off;
    double* VH = vh + off;
    double* XR = xref + off;
    const double* M = ms + off;
    int* NB = nbr + rep.list_off;
    int* NN = nnb + off;
    int* SL = slot_of + off;
    int* CI = cell_of + off;
    int* TI = tmp_idx + off;
    const float Lf = (float)rep.L, invLf = (float)(1.0 / rep.L);
    const int nc = rep.nc;
    const int ncells = nc * nc * nc;
    const double dt = P.dt, dtsq = dt * dt;
    const int64_t step0 = *step_base;

    // Verlet list of every atom, as slots of the cell-ordered LDS positions.  With
    // nc >= 3: bin (LDS counters), scan, place, then rank-sort each cell's slots by atom
    // index (so lists are reproducible) and scan the 27 neighbour cells; otherwise slots
    // are atom indices and every atom scans all atoms.
    auto build = [&]() {
        if (nc >= 3) {
            for (int c = tid; c < ncells; c += BATCH_THREADS) ccount[c] = 0;
            __syncthreads();
            for (int i = tid; i < n; i += BATCH_THREADS) {
                const float4 w = batch_wrap(X[i], X[nt + i], X[2 * nt + i], rep.L);
                int cx = (int)(w.x * rep.inv_cs), cy = (int)(w.y * rep.inv_cs),
                    cz = (int)(w.z * rep.inv_cs);
                cx = cx < nc ? cx : nc - 1; cy = cy < nc ? cy : nc - 1; cz = cz < nc ? cz : nc - 1;
                const int c = (cz * nc + cy) * nc + cx;
                CI[i] = c;
                atomicAdd(&ccount[c], 1);
            }
            __syncthreads();
            if (tid == 0) {
                int s0 = 0;
                for (int c = 0; c < ncells; ++c) { cstart[c] = s0; s0 += ccount[c]; ccount[c] = 0; }
                cstart[ncells] = s0;
            }
            __syncthreads();
            for (int i = tid; i < n; i += BATCH_THREADS) {
                const int c = CI[i];
                TI[cstart[c] + atomicAdd(&ccount[c], 1)] = i;
            }
            __syncthreads();
            for (int s = tid; s < n; s += BATCH_THREADS) {
                const int i = TI[s];
                const int c = CI[i];
                int r = 0;
                for (int u = cstart[c]; u < cstart[c + 1]; ++u) r += TI[u] < i;
                const int slot = cstart[c] + r;
                SL[i] = slot;
                spos[slot] = batch_wrap(X[i], X[nt + i], X[2 * nt + i], rep.L);
            }
            __syncthreads();
            for (int i = tid; i < n; i += BATCH_THREADS) {
                const int si = SL[i];
                const float4 pi = spos[si];
                const int c0 = CI[i];
                const int cz = c0 / (nc * nc), cy = (c0 / nc) % nc, cx = c0 % nc;
                int cnt = 0;
                for (int q = 0; q < 27; ++q) {
                    const int zz = (cz + q / 9 - 1 + nc) % nc, yy = (cy + (q / 3) % 3 - 1 + nc) % nc,
                              xx = (cx + q % 3 - 1 + nc) % nc;
                    const int c = (zz * nc + yy) * nc + xx;
                    for (int s = cstart[c]; s < cstart[c + 1]; ++s) {
                        const float4 pj = spos[s];
                        float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
                        dx -= Lf * rintf(dx * invLf);
                        dy -= Lf * rintf(dy * invLf);
                        dz -= Lf * rintf(dz * invLf);
                        if (s != si && dx * dx + dy * dy + dz * dz < P.rlist2f) {
                            if (cnt < P.cap) NB[(int64_t)cnt * n + i] = s;
                            ++cnt;
                        }
                    }
                }
                if (cnt > P.cap) { flags[0] = 1; cnt = P.cap; }
                NN[i] = cnt;
                XR[i] = X[i];
                XR[nt + i] = X[nt + i];
                XR[2 * nt + i] = X[2 * nt + i];
            }
        } else {
            for (int i = tid; i < n; i += BATCH_THREADS) {
                SL[i] = i;
                spos[i] = batch_wrap(X[i], X[nt + i], X[2 * nt + i], rep.L);
            }
            __syncthreads();
            for (int i = tid; i < n; i += BATCH_THREADS) {
                const float4 pi = spos[i];
                int c = 0;
                for (int j = 0; j < n; ++j) {
                    const float4 pj = spos[j];
                    float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
                    dx -= Lf * rintf(dx * invLf);
                    dy -= Lf * rintf(dy * invLf);
                    dz -= Lf * rintf(dz * invLf);
                    if (j != i && dx * dx + dy * dy + dz * dz < P.rlist2f) {
                        if (c < P.cap) NB[(int64_t)c * n + i] = j;
                        ++c;
                    }
                }
                if (c > P.cap) { flags[0] = 1; c = P.cap; }
                NN[i] = c;
                XR[i] = X[i];
                XR[nt + i] = X[nt + i];
                XR[2 * nt + i] = X[2 * nt + i];
            }
        }
        if (tid == 0) atomicAdd(&flags[1], 1);
    };
    // forces of the current positions, then (do_kick) the end of step `step`
    auto force = [&](int do_kick, int64_t step) {
        for (int i = tid; i < n; i += BATCH_THREADS) {
            const float4 pi = spos[SL[i]];
            nbx_kahan fx, fy, fz;
            fx.init(); fy.init(); fz.init();
            const int cnt = NN[i];
            for (int k = 0; k < cnt; ++k) {
                const float4 pj = spos[NB[(int64_t)k * n + i]];
                float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
                dx -= Lf * rintf(dx * invLf);
                dy -= Lf * rintf(dy * invLf);
                dz -= Lf * rintf(dz * invLf);
                const float r2 = dx * dx + dy * dy + dz * dz;
                if (r2 < P.R2) {
                    const float inv = __builtin_amdgcn_rcpf(r2);
                    const float q = P.s2 * inv;
                    const float s6 = q * q * q;
                    const float f = (2.0f * s6 * s6 - s6) * inv;
                    fx.add(f * dx);
                    fy.add(f * dy);
                    fz.add(f * dz);
                }
            }
            const double c = P.c24eps / M[i];
            const double ax = c * fx.value(), ay = c * fy.value(), az = c * fz.value();
            A[i] = ax;
            A[nt + i] = ay;
            A[2 * nt + i] = az;
            if (!do_kick) continue;
            double v3[3] = {VH[i] + dt * (0.5 * ax), VH[nt + i] + dt * (0.5 * ay),
                            VH[2 * nt + i] + dt * (0.5 * az)};
            if (rep.therm == NBX_THERM_LANGEVIN) {
                const double sig = rep.c2 * sqrt(P.kb * rep.T / M[i]);
                double g[3];
                nbx_rng_normals3(rep.key, i, step, NBX_RNG_LANGEVIN_B, g);
                for (int k = 0; k < 3; ++k) v3[k] = rep.c1 * v3[k] + sig * g[k];
            } else if (rep.therm == NBX_THERM_ANDERSEN &&
                       nbx_rng_uniform(rep.key, i, step, NBX_RNG_ANDERSEN) < rep.prob) {
                const double vdev = sqrt(P.kb * rep.T / rep.m0);
                double g[3];
                nbx_rng_normals3(rep.key, i, step, NBX_RNG_ANDERSEN, g);
                for (int k = 0; k < 3; ++k) v3[k] = vdev * g[k];
            }
            V[i] = v3[0];
            V[nt + i] = v3[1];
            V[2 * nt + i] = v3[2];
        }
    };

    if (prologue) {
        build();
        __syncthreads();
        force(0, 0);
        __syncthreads();
    }
    for (int s = 1; s <= nsteps; ++s) {
        const int64_t step = step0 + s;
        if (tid == 0) rebuild = 0;
        __syncthreads();
        for (int i = tid; i < n; i += BATCH_THREADS) {
            double vx = V[i], vy = V[nt + i], vz = V[2 * nt + i];
            if (rep.therm == NBX_THERM_LANGEVIN) {
                const double sig = rep.c2 * sqrt(P.kb * rep.T / M[i]);
                double g[3];
                nbx_rng_normals3(rep.key, i, step, NBX_RNG_LANGEVIN_A, g);
                vx = rep.c1 * vx + sig * g[0];
                vy = rep.c1 * vy + sig * g[1];
                vz = rep.c1 * vz + sig * g[2];
            }
            const double ax = A[i], ay = A[nt + i], az = A[2 * nt + i];
            const double px = X[i] + dt * vx + dtsq * (0.5 * ax);
            const double py = X[nt + i] + dt * vy + dtsq * (0.5 * ay);
            const double pz = X[2 * nt + i] + dt * vz + dtsq * (0.5 * az);
            X[i] = px; X[nt + i] = py; X[2 * nt + i] = pz;
            VH[i] = vx + dt * (0.5 * ax);
            VH[nt + i] = vy + dt * (0.5 * ay);
            VH[2 * nt + i] = vz + dt * (0.5 * az);
            spos[SL[i]] = batch_wrap(px, py, pz, rep.L);
            const double ex = px - XR[i], ey = py - XR[nt + i], ez = pz - XR[2 * nt + i];
            if (ex * ex + ey * ey + ez * ez > P.half_skin2) rebuild = 1;
        }
        __syncthreads();
        if (rebuild) {
            build();
            __syncthreads();
        }
        force(1, step);
        __syncthreads();
    }
}

// per-replica kinetic energy, LJ potential energy with the reference's tail convention
// (src/nbody_simulation_result.jl:209-218, :293-319) and in-cutoff ordered pair count;
// one block per replica, FP64
__global__ void __launch_bounds__(256)
k_batch_energies(const nbx_batch_rep* __restrict__ reps, const double* __restrict__ x,
                 const double* __restrict__ v, const double* __restrict__ ms, int64_t nt,
                 double s2, double R2, double eps, double* __restrict__ ke,
                 double* __restrict__ pe, long long* __restrict__ pairs) {
    __shared__ double rk[256], rp[256];
    __shared__ long long rc[256];
    const nbx_batch_rep rep = reps[blockIdx.x];
    const int n = rep.n;
    const int64_t off = rep.off;
    const double L = rep.L;
    double k = 0.0, p = 0.0;
    long long c = 0;
    for (int i = threadIdx.x; i < n; i += 256) {
        const int64_t g = off + i;
        const double a = v[g], b = v[nt + g], d = v[2 * nt + g];
        k += (a * a + b * b + d * d) * (ms[g] / 2.0);
        const double xi = x[g], yi = x[nt + g], zi = x[2 * nt + g];
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            const int64_t h = off + j;
            double dx = xi - x[h], dy = yi - x[nt + h], dz = zi - x[2 * nt + h];
            dx -= L * rint(dx / L);
            dy -= L * rint(dy / L);
            dz -= L * rint(dz / L);
            double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 < R2) ++c;
            if (j < i) continue;
            if (!(r2 < R2)) r2 = R2;
            const double q = s2 / r2;
            const double s6 = q * q * q;
            p += s6 * s6 - s6;
        }
    }
    rk[threadIdx.x] = k;
    rp[threadIdx.x] = p;
    rc[threadIdx.x] = c;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            rk[threadIdx.x] += rk[threadIdx.x + o];
            rp[threadIdx.x] += rp[threadIdx.x + o];
            rc[threadIdx.x] += rc[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        ke[blockIdx.x] = rk[0];
        pe[blockIdx.x] = 4.0 * eps * rp[0];
        pairs[blockIdx.x] = rc[0];
    }
}

void batch_launch_steps(hipStream_t s, int nblocks, const void* reps, const int* order,
                        double* x, double* v, double* a, double* vh, double* xref, int* nbr,
                        int* nnb, const double* ms, int* scratch3, int64_t ntot,
                        const nbx_batch_params& P, const int64_t* step_base, int nsteps,
                        int prologue, int* flags) {
    hipLaunchKernelGGL(k_batch_steps, dim3(nblocks), dim3(BATCH_THREADS), 0, s,
                       (const nbx_batch_rep*)reps, order, x, v, a, vh, xref, nbr, nnb, ms,
                       scratch3, scratch3 + ntot, scratch3 + 2 * ntot, P, step_base, nsteps,
                       prologue, flags);
}

void batch_launch_energies(hipStream_t s, int nrep, const void* reps, const double* x,
                           const double* v, const double* ms, int64_t nt, double s2, double R2,
                           double eps, double* ke, double* pe, long long* pairs) {
    hipLaunchKernelGGL(k_batch_energies, dim3(nrep), dim3(256), 0, s, (const nbx_batch_rep*)reps,
                       x, v, ms, nt, s2, R2, eps, ke, pe, pairs);
}

size_t batch_rep_size() { return sizeof(nbx_batch_rep); }
int batch_nmax() { return BATCH_NMAX; }

// host-side fill of the replica records (layout private to this file)
void batch_fill_rep(void* dst, int64_t off, int64_t list_off, int32_t n, int32_t therm, double L,
                    double T, double param, double dt, double m0, uint64_t key, double rlist) {
    nbx_batch_rep r{};
    int nc = (int)floor(L / rlist);
    while (nc * nc * nc > BATCH_MAXCELLS) --nc;
    r.nc = nc >= 3 ? nc : 0;
    r.inv_cs = r.nc ? (double)r.nc / L : 0.0;
    r.off = off;
    r.list_off = list_off;
    r.n = n;
    r.therm = therm;
    r.L = L;
    r.T = T;
    r.param = param;
    r.m0 = m0;
    r.key = key;
    if (therm == NBX_THERM_LANGEVIN) {
        r.c1 = exp(-param * (0.5 * dt));
        r.c2 = sqrt(1.0 - r.c1 * r.c1);
    } else if (therm == NBX_THERM_ANDERSEN) {
        r.prob = param * dt;
    }
    *(nbx_batch_rep*)dst = r;
}

void batch_fill_params(void* dst, double eps, double sigma, double R, double rlist, double skin,
                       double dt, double kb, int cap, int64_t ntot) {
    nbx_batch_params P{};
    P.s2 = (float)(sigma * sigma);
    P.R2 = (float)(R * R);
    P.rlist2f = (float)(rlist * rlist) * 1.0001f;
    P.c24eps = 24.0 * eps;
    P.half_skin2 = 0.25 * skin * skin;
    P.dt = dt;
    P.kb = kb;
    P.cap = cap;
    P.ntot = ntot;
    *(nbx_batch_params*)dst = P;
}

size_t batch_params_size() { return sizeof(nbx_batch_params); }

void batch_launch_steps_raw(hipStream_t s, int nblocks, const void* reps, const int* order,
                            double* x, double* v, double* a, double* vh, double* xref, int* nbr,
                            int* nnb, const double* ms, int* scratch3, const void* params,
                            const int64_t* step_base, int nsteps, int prologue, int* flags) {
    const nbx_batch_params& P = *(const nbx_batch_params*)params;
    batch_launch_steps(s, nblocks, reps, order, x, v, a, vh, xref, nbr, nnb, ms, scratch3, P.ntot,
                       P, step_base, nsteps, prologue, flags);
}
