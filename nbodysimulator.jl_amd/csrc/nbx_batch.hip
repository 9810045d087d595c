// nbx_batch.hip -- batched Lennard-Jones replicas in one persistent launch (gfx950).
//
// One workgroup of BATCH_THREADS threads owns one replica for every step of a launch:
//   - the replica's positions, wrapped into [0, L), centred on the box, in units of the
//     box side (fractional, [-1/2, 1/2)) and rounded to FP32, live in LDS (BATCH_NMAX
//     float4 = 64 KB), so the pair loop never reads HBM for positions; the minimum image
//     of a fractional difference d is d - rint(d), and the cutoffs and the LJ
//     coefficients are scaled per replica to box units; the reference positions of the
//     last list build live in LDS too (fractional FP32, by atom), so the per-step
//     displacement check reads no HBM;
//   - x, v, a, vh stay in FP64 in HBM (SoA over the whole batch: component k of global
//     atom g at k * ntot + g); a thread owns atoms tid, tid + BATCH_THREADS, ... in every
//     phase, so it reads back its own writes.  The FP64 build positions are kept in HBM
//     only to carry the reference from one launch to the next;
//   - LDS positions are kept in cell order ("slots"); the Verlet list holds 16-bit slots
//     in groups of 8 (entry k of local atom i at nbr[list_off + ((k / 8) * n + i) * 8 +
//     k % 8]): a thread fetches 8 neighbours with one 16-byte load and prefetches the next
//     group (and the next atom's first group) while it computes, so the pair loop is not
//     a chain of dependent HBM round trips.  The list is rebuilt inside the launch
//     whenever an atom of the replica moved more than skin/2 since the last build: atoms
//     are binned into an LDS cell list (cells >= list radius), placed, rank-sorted by
//     index within each cell (reproducible lists), and every atom scans the 27
//     neighbouring cells; replicas with fewer than 3 cells per side scan all atoms;
//   - pair forces: FP32, branch-free (out-of-range pairs select a zero force factor), the
//     8 pairs of a list group summed in FP32 and the group sums added to Kahan sums;
//   - thermostat noise: the Philox words of the single-system path, turned into normals
//     by Box-Muller on the FP32 hardware logarithm, square root, sine and cosine;
//   - per step: [Langevin A] drift -> barrier -> [rebuild -> barrier] -> forces, kick,
//     [Langevin B | Andersen] -> barrier.  Only workgroup barriers: the replicas never
//     synchronise with each other, so nothing crosses an XCD.
// Blocks are ordered by decreasing replica size (longest first).
#include "nbx_internal.h"
#include <type_traits>

// 16 waves per replica (one workgroup per CU at the 128-VGPR budget): 4 waves per SIMD
// hide the LDS and HBM latency of one replica's pair loop (8 waves measured 26 % slower
// on the 512-replica batch, whose step time is that of its 4,000-atom replicas)
#define BATCH_THREADS 1024
#define BATCH_NMAX 4096
#define BATCH_MAXCELLS 512
#define BATCH_PER_THREAD (BATCH_NMAX / BATCH_THREADS)

struct nbx_batch_rep {
    int64_t off;       // first global atom
    int64_t list_off;  // first list slot
    int32_t n;
    int32_t therm;
    int32_t nc;        // cells per side of the in-kernel cell list (< 3: all pairs)
    int32_t pad;
    double L, invL;
    // in units of the box side (positions in LDS are fractional): (sigma/L)^6,
    // 2 (sigma/L)^12, (R/L)^2, list radius^2 (with a 1e-4 margin), (skin/2)^2
    float c6, c12, R2, rlist2, hs2;
    double T, param;
    double c1, c2, prob;  // per-dt thermostat coefficients (set by the host per run)
    double m0;
    uint64_t key;
    float c2kT;           // c2 * sqrt(kb T): the Langevin noise amplitude is c2kT / sqrt(m)         // Philox key of the replica
};

struct nbx_batch_params {
    double c24eps;
    double dt, kb;
    int cap;
    int64_t ntot;
};

// position wrapped into [0, L) in FP64, then centred on the box and rounded to FP32:
// coordinates in [-L/2, L/2) carry half the rounding error of [0, L)
__device__ __forceinline__ float4 batch_wrap(double x, double y, double z, double L,
                                             double invL) {
    double wx = x - L * floor(x * invL), wy = y - L * floor(y * invL),
           wz = z - L * floor(z * invL);
    if (wx >= L) wx -= L;
    if (wy >= L) wy -= L;
    if (wz >= L) wz -= L;
    if (wx < 0.0) wx += L;
    if (wy < 0.0) wy += L;
    if (wz < 0.0) wz += L;
    return make_float4((float)(wx * invL - 0.5), (float)(wy * invL - 0.5),
                       (float)(wz * invL - 0.5), 0.0f);
}

// three standard normals from the Philox words of nbx_rng_normals3 (same counters and
// Box-Muller pairing), the logarithm, square root, sine and cosine taken by the FP32
// hardware instructions (v_log_f32 is log2; v_sin_f32 / v_cos_f32 take revolutions, which
// is what the second uniform of each pair is)
__device__ __forceinline__ void batch_normals3(uint64_t seed, int64_t atom, int64_t step,
                                               uint32_t stream, float* g) {
    uint32_t w[4];
    nbx_philox((uint32_t)atom, (uint32_t)step, stream, 1u, (uint32_t)seed,
               (uint32_t)(seed >> 32), w);
    const float m1 = (float)(1.0 - nbx_u01(w[0], w[1]));
    const float t1 = (float)nbx_u01(w[2], w[3]);
    const float r1 = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(m1));
    g[0] = r1 * __builtin_amdgcn_cosf(t1);
    g[1] = r1 * __builtin_amdgcn_sinf(t1);
    nbx_philox((uint32_t)atom, (uint32_t)step, stream, 2u, (uint32_t)seed,
               (uint32_t)(seed >> 32), w);
    const float m2 = (float)(1.0 - nbx_u01(w[0], w[1]));
    const float t2 = (float)nbx_u01(w[2], w[3]);
    const float r2 = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(m2));
    g[2] = r2 * __builtin_amdgcn_cosf(t2);
}

// cell of a centred fractional position (nc cells per side)
__device__ __forceinline__ int batch_cell(float4 w, int nc) {
    const float fn = (float)nc;
    int cx = (int)((w.x + 0.5f) * fn), cy = (int)((w.y + 0.5f) * fn),
        cz = (int)((w.z + 0.5f) * fn);
    cx = cx < 0 ? 0 : (cx < nc ? cx : nc - 1);
    cy = cy < 0 ? 0 : (cy < nc ? cy : nc - 1);
    cz = cz < 0 ? 0 : (cz < nc ? cz : nc - 1);
    return (cz * nc + cy) * nc + cx;
}

// the lane's thread index through an empty asm: per-lane addresses derived from it are
// formed where a phase uses them instead of being hoisted out of the step loop, where
// one 64-bit address per array would stay live (and spill) for the whole launch
__device__ __forceinline__ int batch_lane() {
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

__global__ void __launch_bounds__(BATCH_THREADS) __attribute__((amdgpu_waves_per_eu(4)))
k_batch_steps(const nbx_batch_rep* __restrict__ reps, const int* __restrict__ order,
              double* __restrict__ x, double* __restrict__ v, double* __restrict__ a,
              double* __restrict__ vh, double* __restrict__ xref, uint16_t* __restrict__ nbr,
              int* __restrict__ nnb, const double* __restrict__ ms, int* __restrict__ slot_of,
              nbx_batch_params P,
              const int64_t* __restrict__ step_base, int nsteps, int prologue,
              int* __restrict__ flags) {
    __shared__ float4 spos[BATCH_NMAX];           // positions by slot (cell order)
    __shared__ int ccount[BATCH_MAXCELLS + 1];    // atoms per cell, then fill counters
    __shared__ int cstart[BATCH_MAXCELLS + 1];    // first slot of each cell
    __shared__ uint16_t tidx[BATCH_NMAX];         // atom indices by slot during a build
    __shared__ int wsum[BATCH_THREADS / 64];      // wave totals of the cell-count scan
    __shared__ int rebuild;
    __shared__ uint4 ebuf[BATCH_THREADS];         // list entries being packed, per thread
    // reference positions of the last build (FP32, wrapped and centred), by atom
    __shared__ float xr0[BATCH_NMAX], xr1[BATCH_NMAX], xr2[BATCH_NMAX];
    const nbx_batch_rep rep = reps[order[blockIdx.x]];
    const int n = rep.n;
    const int tid = threadIdx.x;
    const int64_t nt = P.ntot, off = rep.off;
    // one uniform base pointer per component (X0, X1, X2 = x, y, z of the replica), so
    // per-lane addressing is a single 32-bit offset shared by every array
    double* const X0 = x + off, * const X1 = X0 + nt, * const X2 = X1 + nt;
    double* const V0 = v + off, * const V1 = V0 + nt, * const V2 = V1 + nt;
    double* const A0 = a + off, * const A1 = A0 + nt, * const A2 = A1 + nt;
    double* const VH0 = vh + off, * const VH1 = VH0 + nt, * const VH2 = VH1 + nt;
    double* const XR0 = xref + off, * const XR1 = XR0 + nt, * const XR2 = XR1 + nt;
    const double* M = ms + off;
    uint16_t* NB = nbr + rep.list_off;
    const uint4* NBG = reinterpret_cast<const uint4*>(NB);  // group g of atom i: NBG[g * n + i]
    int* NN = nnb + off;
    int* SL = slot_of + off;
    const int nc = rep.nc;
    const int ncells = nc * nc * nc;
    const double dt = P.dt, dtsq = dt * dt;
    const int64_t step0 = *step_base;

    // Verlet list of every atom, as slots of the cell-ordered LDS positions.  With
    // nc >= 3: bin into LDS counters (the arrival rank and cell of a thread's atoms stay
    // in registers), scan the counts in parallel, place the atom indices in LDS, rank
    // each atom among its cell's atoms by index (so slots and lists are reproducible),
    // then every slot scans the 27 neighbour cells with one periodic shift per cell.
    // Otherwise slots are atom indices and every atom scans all atoms.  List entries are
    // packed 8 to a 16-byte group in an LDS staging slot of the thread and written a group
    // at a time.
    auto build = [&]() {
        const int tid = batch_lane();
        uint4* NBW = reinterpret_cast<uint4*>(NB);
        // entry `cnt` of atom i is slot u; a full group (or the final partial one) is stored
        // the thread's group of 8 entries is staged in LDS (entries of a partial final group
        // past the count keep older values: the pair loop never reads them)
        uint16_t* const eb = reinterpret_cast<uint16_t*>(&ebuf[tid]);
        auto emit = [&](int i, int& cnt, unsigned u) {
            const int e = cnt & 7;
            eb[e] = (uint16_t)u;
            if (e == 7 && cnt < P.cap) NBW[(int64_t)(cnt >> 3) * n + i] = ebuf[tid];
            ++cnt;
        };
        auto finish = [&](int i, int cnt) {
            if (cnt > P.cap) { flags[0] = 1; cnt = P.cap; }
            else if (cnt & 7) NBW[(int64_t)(cnt >> 3) * n + i] = ebuf[tid];
            NN[i] = cnt;
        };
        if (nc >= 3) {
            for (int c = tid; c < ncells; c += BATCH_THREADS) ccount[c] = 0;
            __syncthreads();
            int code[BATCH_PER_THREAD];  // cell * BATCH_NMAX + arrival rank in the cell
#pragma unroll
            for (int m = 0; m < BATCH_PER_THREAD; ++m) {
                const int i = tid + m * BATCH_THREADS;
                code[m] = 0;
                if (i < n) {
                    const double x0 = X0[i], x1 = X1[i], x2 = X2[i];
                    XR0[i] = x0; XR1[i] = x1; XR2[i] = x2;
                    const float4 w = batch_wrap(x0, x1, x2, rep.L, rep.invL);
                    xr0[i] = w.x; xr1[i] = w.y; xr2[i] = w.z;
                    const int c = batch_cell(w, nc);
                    code[m] = c * BATCH_NMAX + atomicAdd(&ccount[c], 1);
                }
            }
            __syncthreads();
            {   // exclusive scan of the ncells <= BATCH_THREADS counts: wave scans + wave totals
                const int lane = tid & 63, w = tid >> 6;
                const int own = tid < ncells ? ccount[tid] : 0;
                int v = own;
                for (int d = 1; d < 64; d <<= 1) {
                    const int u = __shfl_up(v, d, 64);
                    if (lane >= d) v += u;
                }
                if (lane == 63) wsum[w] = v;
                __syncthreads();
                int pre = 0;
                for (int q = 0; q < w; ++q) pre += wsum[q];
                if (tid < ncells) cstart[tid] = pre + v - own;
                if (tid == ncells - 1) cstart[ncells] = pre + v;
            }
            __syncthreads();
#pragma unroll
            for (int m = 0; m < BATCH_PER_THREAD; ++m) {
                const int i = tid + m * BATCH_THREADS;
                if (i < n) tidx[cstart[code[m] / BATCH_NMAX] + code[m] % BATCH_NMAX] = (uint16_t)i;
            }
            __syncthreads();
            int slot[BATCH_PER_THREAD];
#pragma unroll
            for (int m = 0; m < BATCH_PER_THREAD; ++m) {
                const int i = tid + m * BATCH_THREADS;
                slot[m] = 0;
                if (i < n) {
                    const int c = code[m] / BATCH_NMAX;
                    const int u0 = cstart[c], u1 = cstart[c + 1];
                    int r = 0;
                    for (int u = u0; u < u1; ++u) r += tidx[u] < i;
                    slot[m] = u0 + r;
                    SL[i] = u0 + r;
                    spos[u0 + r] = batch_wrap(X0[i], X1[i], X2[i], rep.L, rep.invL);
                }
            }
            __syncthreads();
            // tidx becomes the atom at each slot (the ranks above sort every cell by index)
#pragma unroll
            for (int m = 0; m < BATCH_PER_THREAD; ++m) {
                const int i = tid + m * BATCH_THREADS;
                if (i < n) tidx[slot[m]] = (uint16_t)i;
            }
            __syncthreads();
            for (int s = tid; s < n; s += BATCH_THREADS) {
                const int i = tidx[s];
                const float4 pi = spos[s];
                const int c0 = batch_cell(pi, nc);
                const int cz = c0 / (nc * nc), cy = (c0 / nc) % nc, cx = c0 % nc;
                int cnt = 0;
                for (int dz = -1; dz <= 1; ++dz) {
                    int zz = cz + dz;
                    float oz = 0.0f;
                    if (zz < 0) { zz += nc; oz = -1.0f; } else if (zz >= nc) { zz -= nc; oz = 1.0f; }
                    for (int dy = -1; dy <= 1; ++dy) {
                        int yy = cy + dy;
                        float oy = 0.0f;
                        if (yy < 0) { yy += nc; oy = -1.0f; } else if (yy >= nc) { yy -= nc; oy = 1.0f; }
                        for (int dx = -1; dx <= 1; ++dx) {
                            int xx = cx + dx;
                            float ox = 0.0f;
                            if (xx < 0) { xx += nc; ox = -1.0f; } else if (xx >= nc) { xx -= nc; ox = 1.0f; }
                            // the neighbour cell's atoms at their image next to this cell
                            const float qx = pi.x - ox, qy = pi.y - oy, qz = pi.z - oz;
                            const int c = (zz * nc + yy) * nc + xx;
                            const int u1 = cstart[c + 1];
                            // hits of up to 32 candidates as a bit mask (positions read 4 at a
                            // time ahead of the tests, then the last 0..3), the atom's own slot
                            // cleared from the mask, then the set bits in order
                            for (int ub = cstart[c]; ub < u1; ub += 32) {
                                const int nb = u1 - ub < 32 ? u1 - ub : 32;
                                unsigned hits = 0;
                                int k = 0;
                                for (; k + 4 <= nb; k += 4) {
                                    float4 pj[4];
#pragma unroll
                                    for (int t = 0; t < 4; ++t) pj[t] = spos[ub + k + t];
#pragma unroll
                                    for (int t = 0; t < 4; ++t) {
                                        const float ex = qx - pj[t].x, ey = qy - pj[t].y,
                                                    ez = qz - pj[t].z;
                                        hits |= (unsigned)(ex * ex + ey * ey + ez * ez < rep.rlist2)
                                                << (k + t);
                                    }
                                }
                                for (; k < nb; ++k) {
                                    const float4 pj = spos[ub + k];
                                    const float ex = qx - pj.x, ey = qy - pj.y, ez = qz - pj.z;
                                    hits |= (unsigned)(ex * ex + ey * ey + ez * ez < rep.rlist2) << k;
                                }
                                if ((unsigned)(s - ub) < 32u) hits &= ~(1u << (s - ub));
                                while (hits) {
                                    const int b = __builtin_ctz(hits);
                                    hits &= hits - 1;
                                    emit(i, cnt, (unsigned)(ub + b));
                                }
                            }
                        }
                    }
                }
                finish(i, cnt);
            }
        } else {
            for (int i = tid; i < n; i += BATCH_THREADS) {
                SL[i] = i;
                const double x0 = X0[i], x1 = X1[i], x2 = X2[i];
                XR0[i] = x0; XR1[i] = x1; XR2[i] = x2;
                const float4 w = batch_wrap(x0, x1, x2, rep.L, rep.invL);
                spos[i] = w;
                xr0[i] = w.x; xr1[i] = w.y; xr2[i] = w.z;
            }
            __syncthreads();
            for (int i = tid; i < n; i += BATCH_THREADS) {
                const float4 pi = spos[i];
                int cnt = 0;
                for (int j = 0; j < n; ++j) {
                    const float4 pj = spos[j];
                    float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
                    dx -= rintf(dx);
                    dy -= rintf(dy);
                    dz -= rintf(dz);
                    if (j != i && dx * dx + dy * dy + dz * dz < rep.rlist2) emit(i, cnt, (unsigned)j);
                }
                finish(i, cnt);
            }
        }
        if (tid == 0) atomicAdd(&flags[1], 1);
    };
    // forces of the current positions, then (do_kick) the end of step `step`.  The next
    // group of 8 slots and the next atom's count, slot and first group are loaded one
    // group ahead; entries past the count are masked to the atom's own slot (in range)
    // and skipped.
    auto force = [&](int do_kick, int64_t step) {
        const int tid = batch_lane();
        int i = tid;
        int cnt_n = 0, sl_n = 0;
        uint4 g_n = make_uint4(0u, 0u, 0u, 0u);
        if (i < n) { cnt_n = NN[i]; sl_n = SL[i]; g_n = NBG[i]; }
        for (; i < n; i += BATCH_THREADS) {
            const int cnt = cnt_n, si = sl_n;
            uint4 cur = g_n;
            const int inx = i + BATCH_THREADS;
            if (inx < n) { cnt_n = NN[inx]; sl_n = SL[inx]; g_n = NBG[inx]; }
            const double mi = M[i];
            double vh3[3] = {0.0, 0.0, 0.0};
            if (do_kick) { vh3[0] = VH0[i]; vh3[1] = VH1[i]; vh3[2] = VH2[i]; }
            const float4 pi = spos[si];
            nbx_kahan fx, fy, fz;
            fx.init(); fy.init(); fz.init();
            // one pair, branch-free: the force factor of an out-of-range pair (or of a
            // masked entry) is selected to zero and the products enter the group sums
            auto pair = [&](const float4 pj, bool ok, float& gx, float& gy, float& gz) {
                float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
                dx -= rintf(dx);
                dy -= rintf(dy);
                dz -= rintf(dz);
                // fractional units: with t = (L/r)^6, 2 (sigma/r)^12 - (sigma/r)^6 =
                // t (c12 t - c6); the factor 24 eps / L is applied to the sums
                const float r2 = dx * dx + dy * dy + dz * dz;
                const float inv = __builtin_amdgcn_rcpf(r2);
                const float t = inv * inv * inv;
                float f = inv * (t * fmaf(rep.c12, t, -rep.c6));
                f = (ok && r2 < rep.R2) ? f : 0.0f;
                gx = fmaf(f, dx, gx);
                gy = fmaf(f, dy, gy);
                gz = fmaf(f, dz, gz);
            };
            // a group of 8 slots: positions read 4 at a time, the pair forces summed in FP32,
            // the group sum added to the Kahan sums.  TAIL masks entries at or past `rem`.
            auto group = [&](uint4 cur, int rem, auto tail) {
                float gx = 0.0f, gy = 0.0f, gz = 0.0f;
                const unsigned w[4] = {cur.x, cur.y, cur.z, cur.w};
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const unsigned sl[4] = {w[2 * h] & 0xffffu, w[2 * h] >> 16,
                                            w[2 * h + 1] & 0xffffu, w[2 * h + 1] >> 16};
                    float4 pj[4];
#pragma unroll
                    for (int t = 0; t < 4; ++t)
                        pj[t] = spos[(decltype(tail)::value && 4 * h + t >= rem) ? (unsigned)si : sl[t]];
#pragma unroll
                    for (int t = 0; t < 4; ++t)
                        pair(pj[t], decltype(tail)::value ? (4 * h + t < rem) : true, gx, gy, gz);
                }
                fx.add(gx);
                fy.add(gy);
                fz.add(gz);
            };
            const int ng = (cnt + 7) >> 3;
            const int nfull = cnt >> 3;
            for (int g = 0; g < nfull; ++g) {
                uint4 nxt = cur;
                if (g + 1 < ng) nxt = NBG[(int64_t)(g + 1) * n + i];
                group(cur, 8, std::false_type{});
                cur = nxt;
            }
            if (nfull < ng) group(cur, cnt - 8 * nfull, std::true_type{});
            const double c = P.c24eps * rep.invL / mi;
            const double ax = c * fx.value(), ay = c * fy.value(), az = c * fz.value();
            A0[i] = ax;
            A1[i] = ay;
            A2[i] = az;
            if (!do_kick) continue;
            double v3[3] = {vh3[0] + dt * (0.5 * ax), vh3[1] + dt * (0.5 * ay),
                            vh3[2] + dt * (0.5 * az)};
            if (rep.therm == NBX_THERM_LANGEVIN) {
                const float sig = rep.c2kT * __builtin_amdgcn_rsqf((float)mi);
                float g[3];
                batch_normals3(rep.key, i, step, NBX_RNG_LANGEVIN_B, g);
                for (int k = 0; k < 3; ++k) v3[k] = rep.c1 * v3[k] + (double)(sig * g[k]);
            } else if (rep.therm == NBX_THERM_ANDERSEN &&
                       nbx_rng_uniform(rep.key, i, step, NBX_RNG_ANDERSEN) < rep.prob) {
                const float vdev = (float)sqrt(P.kb * rep.T / rep.m0);
                float g[3];
                batch_normals3(rep.key, i, step, NBX_RNG_ANDERSEN, g);
                for (int k = 0; k < 3; ++k) v3[k] = (double)(vdev * g[k]);
            }
            V0[i] = v3[0];
            V1[i] = v3[1];
            V2[i] = v3[2];
        }
    };

    if (!prologue) {
        for (int i = tid; i < n; i += BATCH_THREADS) {
            const float4 w = batch_wrap(XR0[i], XR1[i], XR2[i], rep.L, rep.invL);
            xr0[i] = w.x; xr1[i] = w.y; xr2[i] = w.z;
        }
    }
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
        // drift (and the first Langevin half step).  Without Langevin the next atom's
        // state is loaded while the current one is computed; with it, the current atom's
        // loads are in flight during the random numbers (prefetching as well would
        // exceed the register budget of four waves per SIMD)
        struct dstate { double v[3], a[3], x[3], m; int sl; };
        auto dload = [&](int j, dstate& d) {
            d.v[0] = V0[j]; d.v[1] = V1[j]; d.v[2] = V2[j];
            d.a[0] = A0[j]; d.a[1] = A1[j]; d.a[2] = A2[j];
            d.x[0] = X0[j]; d.x[1] = X1[j]; d.x[2] = X2[j];
            d.m = M[j];
            d.sl = SL[j];
        };
        auto dstore = [&](int i, const dstate& d, double vx, double vy, double vz) {
            const double ax = d.a[0], ay = d.a[1], az = d.a[2];
            const double px = d.x[0] + dt * vx + dtsq * (0.5 * ax);
            const double py = d.x[1] + dt * vy + dtsq * (0.5 * ay);
            const double pz = d.x[2] + dt * vz + dtsq * (0.5 * az);
            X0[i] = px; X1[i] = py; X2[i] = pz;
            VH0[i] = vx + dt * (0.5 * ax);
            VH1[i] = vy + dt * (0.5 * ay);
            VH2[i] = vz + dt * (0.5 * az);
            const float4 w = batch_wrap(px, py, pz, rep.L, rep.invL);
            spos[d.sl] = w;
            float ex = w.x - xr0[i], ey = w.y - xr1[i], ez = w.z - xr2[i];
            ex -= rintf(ex);
            ey -= rintf(ey);
            ez -= rintf(ez);
            if (ex * ex + ey * ey + ez * ez > rep.hs2) rebuild = 1;
        };
        const int t0 = batch_lane();
        if (rep.therm == NBX_THERM_LANGEVIN) {
            for (int i = t0; i < n; i += BATCH_THREADS) {
                dstate d;
                dload(i, d);
                float g[3];
                batch_normals3(rep.key, i, step, NBX_RNG_LANGEVIN_A, g);
                const float sig = rep.c2kT * __builtin_amdgcn_rsqf((float)d.m);
                dstore(i, d, rep.c1 * d.v[0] + (double)(sig * g[0]),
                       rep.c1 * d.v[1] + (double)(sig * g[1]),
                       rep.c1 * d.v[2] + (double)(sig * g[2]));
            }
        } else {
            dstate dn;
            if (t0 < n) dload(t0, dn);
            for (int i = t0; i < n; i += BATCH_THREADS) {
                const dstate d = dn;
                if (i + BATCH_THREADS < n) dload(i + BATCH_THREADS, dn);
                dstore(i, d, d.v[0], d.v[1], d.v[2]);
            }
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
                        double* x, double* v, double* a, double* vh, double* xref, uint16_t* nbr,
                        int* nnb, const double* ms, int* slot_of,
                        const nbx_batch_params& P, const int64_t* step_base, int nsteps,
                        int prologue, int* flags) {
    hipLaunchKernelGGL(k_batch_steps, dim3(nblocks), dim3(BATCH_THREADS), 0, s,
                       (const nbx_batch_rep*)reps, order, x, v, a, vh, xref, nbr, nnb, ms,
                       slot_of, P, step_base, nsteps, prologue, flags);
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
                    double T, double param, double dt, double m0, uint64_t key, double sigma,
                    double R, double skin, double kb) {
    const double rlist = R + skin;
    nbx_batch_rep r{};
    int nc = (int)floor(L / rlist);
    while (nc * nc * nc > BATCH_MAXCELLS) --nc;
    r.nc = nc >= 3 ? nc : 0;
    r.off = off;
    r.list_off = list_off;
    r.n = n;
    r.therm = therm;
    r.L = L;
    r.invL = 1.0 / L;
    const double c6 = pow(sigma / L, 6.0);
    r.c6 = (float)c6;
    r.c12 = (float)(2.0 * c6 * c6);
    r.R2 = (float)((R / L) * (R / L));
    r.rlist2 = (float)((rlist / L) * (rlist / L)) * 1.0001f;
    r.hs2 = (float)(0.25 * (skin / L) * (skin / L));
    r.T = T;
    r.param = param;
    r.m0 = m0;
    r.key = key;
    if (therm == NBX_THERM_LANGEVIN) {
        r.c1 = exp(-param * (0.5 * dt));
        r.c2 = sqrt(1.0 - r.c1 * r.c1);
        r.c2kT = (float)(r.c2 * sqrt(kb * T));
    } else if (therm == NBX_THERM_ANDERSEN) {
        r.prob = param * dt;
    }
    *(nbx_batch_rep*)dst = r;
}

void batch_fill_params(void* dst, double eps, double sigma, double R, double rlist, double skin,
                       double dt, double kb, int cap, int64_t ntot) {
    nbx_batch_params P{};
    (void)sigma; (void)R; (void)rlist; (void)skin;
    P.c24eps = 24.0 * eps;
    P.dt = dt;
    P.kb = kb;
    P.cap = cap;
    P.ntot = ntot;
    *(nbx_batch_params*)dst = P;
}

size_t batch_params_size() { return sizeof(nbx_batch_params); }

void batch_launch_steps_raw(hipStream_t s, int nblocks, const void* reps, const int* order,
                            double* x, double* v, double* a, double* vh, double* xref, uint16_t* nbr,
                            int* nnb, const double* ms, int* slot_of, const void* params,
                            const int64_t* step_base, int nsteps, int prologue, int* flags) {
    const nbx_batch_params& P = *(const nbx_batch_params*)params;
    batch_launch_steps(s, nblocks, reps, order, x, v, a, vh, xref, nbr, nnb, ms, slot_of, P,
                       step_base, nsteps, prologue, flags);
}
