// nbx_nonbonded.hip -- mixed-precision non-bonded path for Lennard-Jones, Coulomb and
// reaction-field systems (potential systems and SPC/Fw water) on gfx950.
//
// Data layout in HBM
//   X[2], v, vh, a : FP64 SoA (component k of atom i at k*n + i), unwrapped positions;
//                    X[k & 1] holds x_k (double-buffered so one kernel can read x_k for
//                    every atom while writing x_{k+1})
//   P[2]           : double4 per atom = x wrapped into [0, L)^3 and the charge; pair
//                    displacements are taken in FP64 from these, minimum-imaged, then
//                    cast to FP32 for the pair law (all-pairs kernels).  After the n
//                    double4 follows a fixed-point copy (mx_posq: the box fraction of each
//                    coordinate in 32 bits, the charge in FP32) that the list force kernel
//                    gathers: the integer difference is the minimum image
//   nbr            : Verlet list, interleaved for G lanes per atom: slot k of atom i is
//                    at (k / G) * (n * G) + i * G + (k % G), so lane l of atom i reads
//                    consecutive entries with the other lanes of the wavefront; entries
//                    are 16-bit atom indices up to 65,536 atoms (half the bytes of the
//                    list stream, no decoding), 32-bit above
//   B[3]           : fixed-capacity cell bins (count + ccap slots per cell); a list is
//                    rebuilt when an atom moved more than skin/2 since the last build:
//                    the drift sets a request flag, and the build launch of the next
//                    step (and, with per-step bins, the binning launch after the drift)
//                    read the flag and exit at once when it is clear.  Bin modes: per-
//                    step bins (two sets by step parity); fixed bins (3 cells per side:
//                    the 27 neighbour cells are the whole box, so the set-up bins serve
//                    every build); bins made by each build for the next one (half-size
//                    cells, three rotating sets, 5 x 5 x 5 stencil)
//
// Pair law (per ordered pair, FP32): r2 = |r|^2; LJ (2 s^12 - s^6)/r2 * r with
// s^2 = sigma^2/r2; Coulomb q_j / r^3 * r; RF q_j (1/r^3 - 2 k_rf) * r.  Per-atom
// sums use Kahan compensation per lane (list kernel of charged plans: the plain FP32 sum
// of each pass of NB_UNROLL pairs is the compensated term); lanes are combined in FP64.
#include "nbx_mixed.h"

#define NB_THREADS 256
#define NB_UNROLL 4

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

// The same pair law added into plain FP32 sums g[0..5] (the list kernel of charged
// plans sums a pass of NB_UNROLL pairs this way and adds the pass sum to its Kahan sums:
// the compensation once per pass instead of once per pair)
template <bool LJ, int EL>
__device__ __forceinline__ void nb_pair_sum(float rx, float ry, float rz, float qj,
                                            bool lj_pair, const mx_pair_params& pp,
                                            float* g) {
    const float r2 = rx * rx + ry * ry + rz * rz;
    if (LJ) {
        if (lj_pair && r2 < pp.lj_R2) {
            const float inv = __builtin_amdgcn_rcpf(r2);
            const float q = pp.lj_s2 * inv;
            const float s6 = q * q * q;
            const float f = (2.0f * s6 * s6 - s6) * inv;
            g[0] = fmaf(f, rx, g[0]);
            g[1] = fmaf(f, ry, g[1]);
            g[2] = fmaf(f, rz, g[2]);
        }
    }
    if (EL != 0) {
        if (r2 < pp.el_R2) {
            const float rinv = __builtin_amdgcn_rsqf(r2);
            const float r3 = rinv * rinv * rinv;
            const float f = (EL == 2) ? qj * (r3 - pp.krf2) : qj * r3;
            g[3] = fmaf(f, rx, g[3]);
            g[4] = fmaf(f, ry, g[4]);
            g[5] = fmaf(f, rz, g[5]);
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

// Per-atom end of the step (one thread per atom): accelerations, kick, thermostats,
// drift.  Returns m vh^2 of the next step (0 when not drifting).
template <bool LJ, int EL, bool WATER>
__device__ __forceinline__ double nb_finish(int64_t i, int64_t n, const double* F,
                                            const mx_atom& A, const mx_pair_params& pp,
                                            const mx_kick_params& kp, const mx_io& io) {
    const double mi = A.m;
    double ax = 0.0, ay = 0.0, az = 0.0;
    if (LJ) {
        const double c = pp.lj_c / mi;
        ax += c * F[0]; ay += c * F[1]; az += c * F[2];
    }
    if (EL != 0) {
        const double c = pp.el_k * A.q / mi;
        ax += c * F[3]; ay += c * F[4]; az += c * F[5];
    }
    if (WATER) {  // bonded terms and the end of the step are done per molecule
        io.a[i] = ax; io.a[n + i] = ay; io.a[2 * n + i] = az;
        return 0.0;
    }
    if (mx_therm_force(kp)) {
        const double cb = mx_berendsen_coeff(kp);
        ax += cb * A.vh[0]; ay += cb * A.vh[1]; az += cb * A.vh[2];
    }
    io.a[i] = ax; io.a[n + i] = ay; io.a[2 * n + i] = az;
    if (!kp.do_kick) return 0.0;
    double v3[3];
    mx_kick_atom(i, ax, ay, az, A, kp, v3);
    if (kp.therm == NBX_THERM_ANDERSEN &&
        nbx_rng_uniform(kp.seed, mx_atom_id(kp, i), kp.step, NBX_RNG_ANDERSEN) < kp.prob)
        mx_andersen_atom(i, kp.m0, kp, v3);
    io.v[i] = v3[0]; io.v[n + i] = v3[1]; io.v[2 * n + i] = v3[2];
    if (!io.do_drift) return 0.0;
    return mx_drift_atom(i, n, v3, ax, ay, az, A, kp, io);
}

// ---------------------------------------------------------------------------
// Verlet-list force kernel: G lanes per atom
// ---------------------------------------------------------------------------

template <int G, bool LJ, int EL, bool WATER, typename IDX>
__global__ void __launch_bounds__(NB_THREADS)
k_nb_list(const double4* __restrict__ pos, const IDX* __restrict__ nbr,
          const int* __restrict__ nnb, int n, mx_pair_params pp, mx_kick_params kp, mx_io io) {
    const int gid = mx_xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    mx_resolve_step(kp);
    const int i = gid / G;
    const int lane = gid & (G - 1);
    double tval = 0.0;
    if (i < n) {
        const int cnt = nnb[i];
        const bool io_ = !WATER || (i % 3) == 0;
        // the atom's inputs to the end of the step, loaded while the pairs are summed
        mx_atom A;
        if (lane == 0) {
            if (WATER) {
                A.m = io.ms[i];
                A.q = EL != 0 ? io.qs[i] : 0.0;
            } else {
                mx_load_atom(i, n, io, A);
            }
        }
        nbx_kahan acc[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) acc[q].init();
        const size_t stride = (size_t)n * G;
        const IDX* p = nbr + gid;
        // NB_UNROLL slots per lane and pass: the position gathers of a pass are issued
        // together, and the slots of the next pass are loaded while this pass computes,
        // so a lane waits for about one gather per pass rather than two dependent loads
        // per pair (slots past the count read the atom itself and are skipped).  The
        // gathers read the 16-byte fixed-point copy: integer minimum image, FP32 charge
        const uint4* qpos = mx_posq(pos, n);
        const uint4 qi = qpos[i];
        // displacements stay in units of L / 2^32 (one rounding, integer -> FP32); the pair
        // law's constants carry the unit instead: sigma^2 and the cutoffs^2 divided by
        // u^2, 2 k_rf times u^3 (u = L / 2^32).  The LJ sums are then 1/u, the
        // electrostatic ones 1/u^2 times the forces, rescaled in FP64 below.  Every scaled
        // quantity is a normal FP32 number: a displacement is below 2^31 units.
        const double un = pp.L * 0x1p-32, iu = 1.0 / un;
        mx_pair_params ps = pp;
        ps.lj_s2 = (float)((double)pp.lj_s2 * iu * iu);
        ps.lj_R2 = (float)((double)pp.lj_R2 * iu * iu);
        ps.el_R2 = (float)((double)pp.el_R2 * iu * iu);
        ps.krf2 = (float)((double)pp.krf2 * un * un * un);
        int jn[NB_UNROLL];
#pragma unroll
        for (int u = 0; u < NB_UNROLL; ++u) jn[u] = lane + u * G < cnt ? (int)p[u * stride] : i;
        for (int k = lane; k < cnt; k += NB_UNROLL * G) {
            int jj[NB_UNROLL];
            uint4 pq[NB_UNROLL];
#pragma unroll
            for (int u = 0; u < NB_UNROLL; ++u) jj[u] = jn[u];
#pragma unroll
            for (int u = 0; u < NB_UNROLL; ++u) pq[u] = qpos[jj[u]];
            p += NB_UNROLL * stride;
            const int kn = k + NB_UNROLL * G;
#pragma unroll
            for (int u = 0; u < NB_UNROLL; ++u) jn[u] = kn + u * G < cnt ? (int)p[u * stride] : i;
            // charged plans (VALU bound: about 430 entries per water atom) add each pass of
            // pairs in plain FP32 and compensate the pass sum; Lennard-Jones-only plans
            // compensate every pair (their close-range terms are large: the pass sums
            // measured 1.5e-6 against 8.5e-7 of the largest force at N = 32,000)
            constexpr bool PASS_SUMS = EL != 0;
            float g[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int u = 0; u < NB_UNROLL; ++u) {
                if (k + u * G >= cnt) break;
                const float rx = (float)(int)(qi.x - pq[u].x);
                const float ry = (float)(int)(qi.y - pq[u].y);
                const float rz = (float)(int)(qi.z - pq[u].z);
                const bool ljp = io_ && (!WATER || (jj[u] % 3) == 0);
                if constexpr (PASS_SUMS)
                    nb_pair_sum<LJ, EL>(rx, ry, rz, __uint_as_float(pq[u].w), ljp, ps, g);
                else
                    nb_pair<LJ, EL>(rx, ry, rz, __uint_as_float(pq[u].w), ljp, ps, acc);
            }
            if constexpr (PASS_SUMS) {
                if (LJ) {
#pragma unroll
                    for (int q = 0; q < 3; ++q) acc[q].add(g[q]);
                }
#pragma unroll
                for (int q = 3; q < 6; ++q) acc[q].add(g[q]);
            }
        }
        double F[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) F[q] = lane_sum<G>(acc[q].value());
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            F[q] *= iu;
            F[3 + q] *= iu * iu;
        }
        if (lane == 0) tval = nb_finish<LJ, EL, WATER>(i, n, F, A, pp, kp, io);
    }
    if (!WATER) mx_epilogue(tval, io, n);
}

// ---------------------------------------------------------------------------
// All-pairs force kernel (any boundary condition, any cutoff): one thread per atom,
// sources staged through LDS, displacement by the reference's boundary rule in FP64.
// ---------------------------------------------------------------------------

#define AP_TILE 256

template <bool LJ, int EL, bool WATER>
__global__ void __launch_bounds__(AP_TILE)
k_nb_allpairs(int n, mx_pair_params pp, nbx_bcv bc, mx_kick_params kp, mx_io io) {
    __shared__ double sx[AP_TILE], sy[AP_TILE], sz[AP_TILE];
    __shared__ float sq[AP_TILE];
    mx_resolve_step(kp);
    const double* __restrict__ x = io.x_in;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = i < n;
    double xi = 0.0, yi = 0.0, zi = 0.0;
    if (active) { xi = x[i]; yi = x[n + i]; zi = x[2 * n + i]; }
    const bool io_ = !WATER || (i % 3) == 0;
    nbx_kahan acc[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) acc[q].init();
    for (int base = 0; base < n; base += AP_TILE) {
        const int j = base + threadIdx.x;
        if (j < n) {
            sx[threadIdx.x] = x[j]; sy[threadIdx.x] = x[n + j]; sz[threadIdx.x] = x[2 * n + j];
            sq[threadIdx.x] = (EL != 0) ? (float)io.qs[j] : 0.0f;
        }
        __syncthreads();
        const int lim = (n - base) < AP_TILE ? (n - base) : AP_TILE;
        if (active) {
            for (int q = 0; q < lim; ++q) {
                const int jj = base + q;
                if (WATER ? (jj / 3 == i / 3) : (jj == i)) continue;
                double rx, ry, rz, r2;
                nbx_displacement(xi, yi, zi, sx[q], sy[q], sz[q], bc, rx, ry, rz, r2);
                const bool ljp = io_ && (!WATER || (jj % 3) == 0);
                nb_pair<LJ, EL>((float)rx, (float)ry, (float)rz, sq[q], ljp, pp, acc);
            }
        }
        __syncthreads();
    }
    double tval = 0.0;
    if (active) {
        double F[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) F[q] = acc[q].value();
        mx_atom A;
        if (WATER) {
            A.m = io.ms[i];
            A.q = EL != 0 ? io.qs[i] : 0.0;
        } else {
            mx_load_atom(i, n, io, A);
        }
        tval = nb_finish<LJ, EL, WATER>(i, n, F, A, pp, kp, io);
    }
    if (!WATER) mx_epilogue(tval, io, n);
}

// ---------------------------------------------------------------------------
// Dispatch over the template parameters
// ---------------------------------------------------------------------------

template <int G, bool LJ, int EL, bool WATER>
static void launch_list(hipStream_t s, const nbx_ctx* c, const void* pos,
                        const mx_pair_params& pp, const mx_kick_params& kp, const mx_io& io) {
    const int64_t threads = c->n * G;
    if (c->nl.nbr_wide)
        hipLaunchKernelGGL((k_nb_list<G, LJ, EL, WATER, int>), dim3(nbx_blocks(threads, NB_THREADS)),
                           dim3(NB_THREADS), 0, s, (const double4*)pos,
                           static_cast<const int*>(c->nl.d_nbr), c->nl.d_nnb, (int)c->n, pp, kp, io);
    else
        hipLaunchKernelGGL((k_nb_list<G, LJ, EL, WATER, uint16_t>),
                           dim3(nbx_blocks(threads, NB_THREADS)), dim3(NB_THREADS), 0, s,
                           (const double4*)pos, static_cast<const uint16_t*>(c->nl.d_nbr),
                           c->nl.d_nnb, (int)c->n, pp, kp, io);
}

template <bool LJ, int EL, bool WATER>
static void launch_nb(hipStream_t s, const mx_nb_launch& L, const nbx_ctx* c, const void* pos,
                      const mx_pair_params& pp, const mx_kick_params& kp, const mx_io& io) {
    if (L.mode_list) {
        switch (L.lanes) {
        case 1: launch_list<1, LJ, EL, WATER>(s, c, pos, pp, kp, io); break;
        case 2: launch_list<2, LJ, EL, WATER>(s, c, pos, pp, kp, io); break;
        case 4: launch_list<4, LJ, EL, WATER>(s, c, pos, pp, kp, io); break;
        case 16: launch_list<16, LJ, EL, WATER>(s, c, pos, pp, kp, io); break;
        case 32: launch_list<32, LJ, EL, WATER>(s, c, pos, pp, kp, io); break;
        default: launch_list<8, LJ, EL, WATER>(s, c, pos, pp, kp, io); break;
        }
    } else {
        hipLaunchKernelGGL((k_nb_allpairs<LJ, EL, WATER>), dim3(nbx_blocks(c->n, AP_TILE)),
                           dim3(AP_TILE), 0, s, (int)c->n, pp, c->bc, kp, io);
    }
}

void mx_launch_nonbonded(hipStream_t s, const mx_nb_launch& L, const nbx_ctx* c,
                         const void* pos, const mx_pair_params& pp, const mx_kick_params& kp,
                         const mx_io& io) {
    if (c->n <= 0) return;
    const int key = (L.has_lj ? 1 : 0) | (L.el_kind << 1) | (L.water ? 8 : 0);
    switch (key) {
    case 1: launch_nb<true, 0, false>(s, L, c, pos, pp, kp, io); break;
    case 2: launch_nb<false, 1, false>(s, L, c, pos, pp, kp, io); break;
    case 3: launch_nb<true, 1, false>(s, L, c, pos, pp, kp, io); break;
    case 4: launch_nb<false, 2, false>(s, L, c, pos, pp, kp, io); break;
    case 5: launch_nb<true, 2, false>(s, L, c, pos, pp, kp, io); break;
    case 1 | 2 | 8: launch_nb<true, 1, true>(s, L, c, pos, pp, kp, io); break;
    case 1 | 4 | 8: launch_nb<true, 2, true>(s, L, c, pos, pp, kp, io); break;
    case 2 | 8: launch_nb<false, 1, true>(s, L, c, pos, pp, kp, io); break;
    case 4 | 8: launch_nb<false, 2, true>(s, L, c, pos, pp, kp, io); break;
    case 1 | 8: launch_nb<true, 0, true>(s, L, c, pos, pp, kp, io); break;
    default: break;
    }
}

// ---------------------------------------------------------------------------
// Drift of every atom (start of a run / chunk): x_{k+1}, vh_{k+1} from x_k, v_k, a_k
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(256) k_mx_drift(int n, mx_kick_params kp, mx_io io) {
    mx_resolve_step(kp);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double tval = 0.0;
    if (i < n) {
        mx_atom A;
        mx_load_atom(i, n, io, A);
        const double v3[3] = {io.v[i], io.v[n + i], io.v[2 * n + i]};
        tval = mx_drift_atom(i, n, v3, io.a[i], io.a[n + i], io.a[2 * n + i], A, kp, io);
    }
    mx_epilogue_waves(tval, io, n);
}

void mx_launch_drift(hipStream_t s, const nbx_ctx* c, const mx_kick_params& kp,
                     const mx_io& io) {
    if (c->n <= 0) return;
    hipLaunchKernelGGL(k_mx_drift, dim3(nbx_blocks(c->n, 256)), dim3(256), 0, s, (int)c->n, kp,
                       io);
}

// step base bookkeeping for replayed graphs: set (set = 1) or advance by delta
__global__ void k_step_advance(int64_t* p, int64_t delta, int set) {
    if (blockIdx.x == 0 && threadIdx.x == 0) p[0] = set ? delta : p[0] + delta;
}

void mx_launch_step_advance(hipStream_t s, int64_t* step_base, int64_t delta, int set) {
    hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(64), 0, s, step_base, delta, set);
}

__global__ void k_fill_int(int* __restrict__ p, int64_t count, int value) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < count) p[t] = value;
}

void mx_launch_fill_int(hipStream_t s, int* p, int64_t count, int value) {
    if (count <= 0) return;
    hipLaunchKernelGGL(k_fill_int, dim3(nbx_blocks(count, 256)), dim3(256), 0, s, p, count, value);
}

// wrapped copy of x (after set_state), binned; requests a list build
__global__ void k_mx_positions(const double* __restrict__ x, int n, int kind, void* pos,
                               double L, double wscale, const double* __restrict__ ms,
                               const double* __restrict__ qs, int* __restrict__ bin_count,
                               int* __restrict__ bin_atoms, int ccap, int nc, double inv_cs,
                               int* flags, int* req) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0 && req != nullptr) req[0] = 1;
    if (i >= n) return;
    const double px = x[i], py = x[n + i], pz = x[2 * n + i];
    if (kind == 1) {
        const double wx = mx_wrap(px, L), wy = mx_wrap(py, L), wz = mx_wrap(pz, L);
        const double q = qs != nullptr ? qs[i] : 0.0;
        ((double4*)pos)[i] = make_double4(wx, wy, wz, q);
        mx_posq(pos, n)[i] = mx_fixed_position(px, py, pz, 1.0 / L, q);
        if (bin_count != nullptr) {
            const int cid = (mx_cell_coord(wz, inv_cs, nc) * nc + mx_cell_coord(wy, inv_cs, nc)) *
                                nc + mx_cell_coord(wx, inv_cs, nc);
            const int slot = atomicAdd(&bin_count[cid], 1);
            if (slot < ccap) bin_atoms[(size_t)cid * ccap + slot] = i;
            else flags[1] = 1;
        }
    } else if (kind == 2) {
        ((float4*)pos)[i] = make_float4((float)px, (float)py, (float)pz, (float)(ms[i] * wscale));
    }
}

void mx_launch_positions(hipStream_t s, const nbx_ctx* c, const double* x, void* pos,
                         int pos_kind, double wscale, int* bin_count, int* bin_atoms, int* req) {
    if (c->n <= 0) return;
    const int nc = c->nl.ncell;
    hipLaunchKernelGGL(k_mx_positions, dim3(nbx_blocks(c->n, 256)), dim3(256), 0, s, x,
                       (int)c->n, pos_kind, pos, c->bc.b[0], wscale, c->d_ms,
                       c->has_q ? c->d_qs : nullptr, nc >= 3 ? bin_count : nullptr, bin_atoms,
                       c->nl.ccap, nc, nc > 0 ? (double)nc / c->bc.b[0] : 0.0, c->nl.d_flags, req);
}

// bins of the wrapped positions (only when a list build is requested for them)
// Consecutive atoms of a workgroup that fall into the same cell (runs: the state is kept
// in cell order during long runs) take their slots with one device-scope atomic per run
// instead of one per atom; the order of atoms within a cell does not matter (the builds
// rank the candidates by index).
#define BIN_THREADS 256
__global__ void __launch_bounds__(BIN_THREADS)
k_bin(const double4* __restrict__ pos, int n, int nc, double inv_cs,
      int* __restrict__ bin_count, int* __restrict__ bin_atoms, int ccap,
      int* __restrict__ flags, const int* __restrict__ req) {
    if (req[0] == 0) return;
    __shared__ int scid[BIN_THREADS], sbase[BIN_THREADS];
    const int t = threadIdx.x;
    const int i = blockIdx.x * BIN_THREADS + t;
    int cid = -1;
    if (i < n) {
        const double4 p = pos[i];
        cid = (mx_cell_coord(p.z, inv_cs, nc) * nc + mx_cell_coord(p.y, inv_cs, nc)) * nc +
              mx_cell_coord(p.x, inv_cs, nc);
    }
    scid[t] = cid;
    __syncthreads();
    // the run of equal cells that this thread belongs to: [head, end)
    int head = t;
    while (head > 0 && scid[head - 1] == cid) --head;
    if (cid >= 0 && head == t) {
        int end = t + 1;
        while (end < BIN_THREADS && scid[end] == cid) ++end;
        sbase[t] = atomicAdd(&bin_count[cid], end - t);
    }
    __syncthreads();
    if (cid < 0) return;
    const int slot = sbase[head] + (t - head);
    if (slot < ccap) bin_atoms[(size_t)cid * ccap + slot] = i;
    else flags[1] = 1;
}

void mx_launch_bin(hipStream_t s, const nbx_ctx* c, const void* pos, const int* req,
                   int* bin_count, int* bin_atoms) {
    const int nc = c->nl.ncell;
    if (c->n <= 0 || nc < 3) return;
    hipLaunchKernelGGL(k_bin, dim3(nbx_blocks(c->n, BIN_THREADS)), dim3(BIN_THREADS), 0, s,
                       (const double4*)pos,
                       (int)c->n, nc, (double)nc / c->bc.b[0], bin_count, bin_atoms, c->nl.ccap,
                       c->nl.d_flags, req);
}

// xi = (1/tau)^2 (T/T0 - (ndf+1)/ndf) from scal[0] (the initial right-hand side)
__global__ void k_nh_rate(double* scal, mx_nh_params nh) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double it = 1.0 / nh.tau;
    scal[NBX_SCAL_XI] = (it * it) * (scal[0] / nh.T0 - (double)(nh.ndf + 1) / (double)nh.ndf);
}

void mx_launch_nh_rate(hipStream_t s, double* scal, const mx_nh_params& nh) {
    hipLaunchKernelGGL(k_nh_rate, dim3(1), dim3(64), 0, s, scal, nh);
}

// the temperature of vh_{k+1}: the block partials of the step kernel before it on the
// stream, added in block order (each thread a strided run of blocks, then a fixed tree),
// and the Nose-Hoover update (src/thermostats.jl:116-128)
__global__ void __launch_bounds__(256) k_temperature_finish(mx_io io) {
    __shared__ double red[256];
    const int t = threadIdx.x;
    const int nb = *io.red_count;
    double s = 0.0;
    for (int b = t; b < nb; b += 256) s += io.red_part[b];
    red[t] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (t < off) red[t] += red[t + off];
        __syncthreads();
    }
    if (t == 0) {
        const double T = red[0] / io.T_denom;
        io.T_out[0] = T;
        if (io.nh.nh) {
            if (io.nh.advance) {
                const double z0 = io.T_out[NBX_SCAL_ZETA], x0 = io.T_out[NBX_SCAL_XI];
                io.T_out[NBX_SCAL_ZETA_PREV] = z0;
                io.T_out[NBX_SCAL_XI_PREV] = x0;
                io.T_out[NBX_SCAL_ZETA] = z0 + io.nh.dt * x0;
            }
            const double it = 1.0 / io.nh.tau;
            io.T_out[NBX_SCAL_XI] =
                (it * it) * (T / io.nh.T0 - (double)(io.nh.ndf + 1) / (double)io.nh.ndf);
        }
    }
}

void mx_launch_temperature_finish(hipStream_t s, const mx_io& io) {
    hipLaunchKernelGGL(k_temperature_finish, dim3(1), dim3(256), 0, s, io);
}

// one component-major (SoA) array reordered by a permutation
__global__ void __launch_bounds__(256) k_gather(const double* __restrict__ src,
                                                double* __restrict__ dst,
                                                const int* __restrict__ map, int64_t n,
                                                int comps) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int64_t j = map[s];
    for (int k = 0; k < comps; ++k) dst[k * n + s] = src[k * n + j];
}

void mx_launch_gather(hipStream_t s, const double* src, double* dst, const int* map,
                      int64_t n, int comps) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_gather, dim3(nbx_blocks(n, 256)), dim3(256), 0, s, src, dst, map, n,
                       comps);
}

// ---------------------------------------------------------------------------
// Verlet list build (gated by *req): one block of 4 wavefronts per cell.
//   1. the cell's own atoms (<= 256 per pass) are staged in LDS (rotate: and binned by
//      their current positions into the next bin set);
//   2. the atoms of the stencil cells -- the 27 neighbour cells, or with rotate the 125
//      cells within two half-size cells (fixed order; in batches of <= 1024) -- are
//      staged in LDS -- indices, then FP32 positions (with per-step bins: at the periodic
//      image of their cell next to the home cell), every load independent -- and sorted
//      by index within each cell, so every list comes out in a reproducible order (cell
//      order, then index order);
//   3. wave w takes own atoms 8 at a time; its 64 lanes read 64 candidates at a time and
//      test each against the 8 (per-step bins: a plain FP32 difference, the images being
//      the near ones; other bins: the difference of the fixed-point box fractions, whose
//      integer wrap-around is the minimum image), appending every atom's hits in
//      candidate order (ballot + prefix count).
//   The squared list radius carries a 1e-4 relative margin so that the FP32 test never
//   drops a pair that the FP64 distance puts inside the list radius.
// Per-step bins: every block also clears its cell in the other bin set (used by the next
// step).  Rotate (bins made by the builds): the build reads set b % 3 (the bins of the
// previous build; every atom has since moved at most 3/4 skin, which the drift enforces,
// and two half cells reach list radius + 1.5 skin), fills set (b + 1) % 3 (cleared by the
// previous build) and clears set (b + 2) % 3; the last block to finish (device-scope
// counter) advances b, which every block read at its start.
// ---------------------------------------------------------------------------

#define BUILD_THREADS 256
#define BUILD_SLOTS 1024
#define BUILD_OWN 256
#define BUILD_OWN_BATCH 8

__device__ __forceinline__ int build_segment(const int* seg, int nseg, int t) {
    int lo = 0, hi = nseg - 1;   // last s with seg[s] <= t
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (seg[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

#define BUILD_STENCIL 125

template <bool WATER, bool SHIFT>
__global__ void __launch_bounds__(BUILD_THREADS)
k_build_cells(const double4* __restrict__ pos, const double* __restrict__ x, int n, int nc,
              double Ld, int* __restrict__ bin_count, int* __restrict__ bin_atoms,
              int ccap, double inv_cs, float rlist2f, int G, int cap,
              void* __restrict__ nbr, int nbr_wide, int* __restrict__ nnb,
              double* __restrict__ xref, int* __restrict__ flags, int* __restrict__ req,
              int clear_bins, int rotate, int parts) {
    __shared__ __attribute__((aligned(16))) int skey[BUILD_SLOTS];
    __shared__ float4 spos[BUILD_SLOTS];
    __shared__ int seg[BUILD_STENCIL + 1], segcell[BUILD_STENCIL];
    __shared__ int ncnt[BUILD_STENCIL], ncell[BUILD_STENCIL];
    __shared__ float4 nshift[BUILD_STENCIL], segshift[BUILD_STENCIL];
    __shared__ int npre[BUILD_STENCIL + 1];   // exclusive prefix sums of ncnt
    __shared__ int okey[BUILD_OWN], ocnt[BUILD_OWN];
    __shared__ float4 opos[BUILD_OWN];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    // the squared list radius in the units of the test: length, or L / 2^32 for the
    // fixed-point positions
    const float r2test = SHIFT ? rlist2f
                               : (float)((double)rlist2f * (4294967296.0 / Ld) * (4294967296.0 / Ld));
    // `parts` blocks share a home cell, each taking a contiguous slice of its atoms (more
    // blocks in flight when the cells are few and large, as for water)
    const int cid = blockIdx.x / parts, part = blockIdx.x % parts;
    if (req[0] == 0) return;   // uniform: only this kernel's last block clears the request
    const int ncells = nc * nc * nc;
    // rotate: the bins of the previous build (set `bset`) are read, the bins of the
    // current positions go to the next set, and the set after it is cleared
    int* state = flags + NBX_FLAG_BINSET;
    const int bset = rotate ? state[0] : 0;
    const int* count = bin_count + (size_t)bset * ncells;
    const int* atoms = bin_atoms + (size_t)bset * ncells * ccap;
    int* next_count = bin_count + (size_t)((bset + 1) % 3) * ncells;
    int* next_atoms = bin_atoms + (size_t)((bset + 1) % 3) * ncells * ccap;
    if (rotate && tid == 0 && part == 0) bin_count[(size_t)((bset + 2) % 3) * ncells + cid] = 0;
    if (blockIdx.x == 0 && tid == 0) atomicAdd(&flags[2], 1);
    const int cz = cid / (nc * nc), cy = (cid / nc) % nc, cx = cid % nc;
    const int nst = rotate ? 125 : 27;
    if (tid < nst) {  // the stencil cells (fixed order) and their counts, in parallel
        const int w = rotate ? 5 : 3, h = rotate ? 2 : 1;
        const int oz = tid / (w * w) - h, oy = (tid / w) % w - h, ox = tid % w - h;
        const int nid = (((cz + oz + nc) % nc) * nc + (cy + oy + nc) % nc) * nc + (cx + ox + nc) % nc;
        const int m2 = count[nid];
        ncell[tid] = nid;
        ncnt[tid] = m2 < ccap ? m2 : ccap;
        // the periodic image of the stencil cell next to the home cell: -1, 0 or +1 box
        // sides per axis
        const auto img = [nc](int c) { return c < 0 ? -1.0f : (c >= nc ? 1.0f : 0.0f); };
        nshift[tid] = make_float4(img(cx + ox), img(cy + oy), img(cz + oz), 0.0f);
    }
    int m = count[cid];
    m = m < ccap ? m : ccap;
    const int per = (m + parts - 1) / parts;
    const int hb = part * per, he = (hb + per) < m ? (hb + per) : m;
    if (hb < he) {   // uniform over the block
    __syncthreads();
    if (wave == 0) {   // prefix sums of the stencil counts (nst <= 128: two per lane)
        const int a0 = 2 * lane < nst ? ncnt[2 * lane] : 0;
        const int a1 = 2 * lane + 1 < nst ? ncnt[2 * lane + 1] : 0;
        int v = a0 + a1;
        for (int d = 1; d < 64; d <<= 1) {
            const int u = __shfl_up(v, d, 64);
            if (lane >= d) v += u;
        }
        const int ex = v - a0 - a1;
        if (2 * lane <= nst) npre[2 * lane] = ex;
        if (2 * lane + 1 <= nst) npre[2 * lane + 1] = ex + a0;
    }
    __syncthreads();
    // G is a power of two: slot r of atom i is at (r >> lgG) n G + i G + r % G
    const uint32_t stride32 = (uint32_t)n * (uint32_t)G;
    const int lgG = __builtin_ctz((unsigned)G);
    for (int own0 = hb; own0 < he; own0 += BUILD_OWN) {
        const int mo = (he - own0) < BUILD_OWN ? (he - own0) : BUILD_OWN;
        for (int t = tid; t < mo; t += BUILD_THREADS) {
            const int i = atoms[(size_t)cid * ccap + own0 + t];
            const double4 p = pos[i];
            okey[t] = i;
            if (SHIFT) {
                opos[t] = make_float4((float)p.x, (float)p.y, (float)p.z, 0.0f);
            } else {   // the fixed-point box fractions (their bits), minimum image by integers
                const uint4 qv = mx_posq(pos, n)[i];
                opos[t] = make_float4(__uint_as_float(qv.x), __uint_as_float(qv.y),
                                      __uint_as_float(qv.z), 0.0f);
            }
            ocnt[t] = 0;
            if (rotate) {   // the bin of the atom's current position, for the next build
                const int c2 = (mx_cell_coord(p.z, inv_cs, nc) * nc +
                                mx_cell_coord(p.y, inv_cs, nc)) * nc + mx_cell_coord(p.x, inv_cs, nc);
                const int slot = atomicAdd(&next_count[c2], 1);
                if (slot < ccap) next_atoms[(size_t)c2 * ccap + slot] = i;
                else flags[1] = 1;
            }
        }
        int nb = 0;
        while (nb < nst) {
            // a batch of consecutive stencil cells that fits in LDS: the most cells from nb
            // on whose counts sum to at most BUILD_SLOTS (a cell holds at most ccap <=
            // BUILD_SLOTS atoms), found by bisection on the prefix sums
            int lo = nb + 1, hi = nst;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (npre[mid] - npre[nb] <= BUILD_SLOTS) lo = mid;
                else hi = mid - 1;
            }
            const int nseg = lo - nb, M = npre[lo] - npre[nb];
            if (tid < nseg) {
                seg[tid] = npre[nb + tid] - npre[nb];
                segcell[tid] = ncell[nb + tid];
                segshift[tid] = nshift[nb + tid];
            }
            if (tid == 0) seg[nseg] = M;
            nb = lo;
            __syncthreads();
            // candidate t = tid + q BUILD_THREADS: its stencil cell (segment), index and
            // position; the position gathers are issued here and stay in flight through
            // the barrier and the rank counts below
            constexpr int QN = BUILD_SLOTS / BUILD_THREADS;
            int ck[QN], cs[QN], cr[QN];
            // per-step bins: FP64 positions shifted to the image next to the home cell;
            // other bins: the 16-byte fixed-point positions
            using gather_t = typename std::conditional<SHIFT, double4, uint4>::type;
            const gather_t* gsrc;
            if constexpr (SHIFT) gsrc = pos;
            else gsrc = mx_posq(pos, n);
            gather_t pg[QN];
#pragma unroll
            for (int q = 0; q < QN; ++q) {
                const int t = tid + q * BUILD_THREADS;
                ck[q] = -1;
                cs[q] = 0;
                if (t < M) {
                    const int s = build_segment(seg, nseg, t);
                    cs[q] = s;
                    ck[q] = atoms[(size_t)segcell[s] * ccap + (t - seg[s])];
                    skey[t] = ck[q];
                }
            }
#pragma unroll
            for (int q = 0; q < QN; ++q)
                if (ck[q] >= 0) pg[q] = gsrc[ck[q]];
            __syncthreads();
            // rank by index within the candidate's cell (the cell's indices read four at
            // a time where aligned), so that candidates are stored in list order
#pragma unroll
            for (int q = 0; q < QN; ++q) {
                if (ck[q] >= 0) {
                    const int key = ck[q];
                    const int b = seg[cs[q]], e1 = seg[cs[q] + 1];
                    int r = 0, u = b;
                    for (; u < e1 && (u & 3) != 0; ++u) r += skey[u] < key;
                    for (; u + 4 <= e1; u += 4) {
                        const int4 k4 = *reinterpret_cast<const int4*>(&skey[u]);
                        r += (k4.x < key) + (k4.y < key) + (k4.z < key) + (k4.w < key);
                    }
                    for (; u < e1; ++u) r += skey[u] < key;
                    cr[q] = b + r;
                }
            }
            // positions at the candidate's image next to the home cell (FP64 sum, one FP32
            // rounding), or the fixed-point fractions' bits
            float4 cp[QN];
#pragma unroll
            for (int q = 0; q < QN; ++q) {
                if (ck[q] >= 0) {
                    if constexpr (SHIFT) {
                        const float4 sh = segshift[cs[q]];
                        cp[q] = make_float4((float)(pg[q].x + (double)sh.x * Ld),
                                            (float)(pg[q].y + (double)sh.y * Ld),
                                            (float)(pg[q].z + (double)sh.z * Ld), 0.0f);
                    } else {
                        cp[q] = make_float4(__uint_as_float(pg[q].x), __uint_as_float(pg[q].y),
                                            __uint_as_float(pg[q].z), 0.0f);
                    }
                }
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < QN; ++q) {
                if (ck[q] >= 0) {
                    skey[cr[q]] = ck[q];
                    spos[cr[q]] = cp[q];
                }
            }
            __syncthreads();
            // wave w tests the candidates against `per` of its own atoms at a time (own
            // atoms per*w .. per*w+per-1, then per*(w+4) ..; per = ceil(own / 4) up to 8, so
            // every wave has work): each candidate is read from LDS once per `per` tests,
            // and those tests are independent
            const int per = min(BUILD_OWN_BATCH, (mo + BUILD_THREADS / 64 - 1) / (BUILD_THREADS / 64));
            for (int o0 = wave * per; o0 < mo; o0 += BUILD_THREADS / 64 * per) {
                // the batch's own atoms are the same for the whole wave: their index and list
                // count are held in scalar registers
                const int no = __builtin_amdgcn_readfirstlane(min(per, mo - o0));
                float4 pi[BUILD_OWN_BATCH];
                int ii[BUILD_OWN_BATCH], im[BUILD_OWN_BATCH], cc[BUILD_OWN_BATCH];
#pragma unroll
                for (int k = 0; k < BUILD_OWN_BATCH; ++k) {
                    const int o = k < no ? o0 + k : o0;
                    pi[k] = opos[o];
                    ii[k] = __builtin_amdgcn_readfirstlane(okey[o]);
                    im[k] = WATER ? ii[k] / 3 : 0;
                    cc[k] = __builtin_amdgcn_readfirstlane(ocnt[o]);
                }
                for (int tb = 0; tb < M; tb += 64) {
                    const int t = tb + lane;
                    const bool v = t < M;
                    const int j = skey[v ? t : 0];
                    const float4 pj = spos[v ? t : 0];
                    const int jm = WATER ? j / 3 : 0;
#pragma unroll
                    for (int k = 0; k < BUILD_OWN_BATCH; ++k) {
                        if (k >= no) break;   // scalar
                        float dx, dy, dz;
                        if constexpr (SHIFT) {
                            dx = pi[k].x - pj.x;
                            dy = pi[k].y - pj.y;
                            dz = pi[k].z - pj.z;
                        } else {   // minimum image in units of L / 2^32
                            dx = (float)(int)(__float_as_uint(pi[k].x) - __float_as_uint(pj.x));
                            dy = (float)(int)(__float_as_uint(pi[k].y) - __float_as_uint(pj.y));
                            dz = (float)(int)(__float_as_uint(pi[k].z) - __float_as_uint(pj.z));
                        }
                        // every term evaluated (no short-circuit branches)
                        const bool hit = v & (dx * dx + dy * dy + dz * dz < r2test) &
                                         (j != ii[k]) & !(WATER && jm == im[k]);
                        const unsigned long long mask = __ballot(hit);
                        if (mask == 0ull) continue;   // scalar
                        // slot r of a hit: the hits before this lane in the wave, after
                        // the atom's earlier hits (slots past the capacity are counted, not
                        // stored: the count flags the overflow)
                        const unsigned lo = (unsigned)mask, hi = (unsigned)(mask >> 32);
                        const int r = cc[k] + (int)__builtin_amdgcn_mbcnt_hi(
                                                  hi, __builtin_amdgcn_mbcnt_lo(lo, 0u));
                        if (hit & (r < cap))
                            mx_nbr_store(nbr, nbr_wide,
                                         (uint64_t)(uint32_t)(r >> lgG) * stride32 +
                                             (uint32_t)(ii[k] * G + (r & (G - 1))), j);
                        cc[k] += __popcll(mask);
                    }
                }
                if (lane == 0) {
#pragma unroll
                    for (int k = 0; k < BUILD_OWN_BATCH; ++k)
                        if (k < no) ocnt[o0 + k] = cc[k];
                }
            }
            __syncthreads();
        }
        for (int t = tid; t < mo; t += BUILD_THREADS) {
            const int i = okey[t];
            int c = ocnt[t];
            if (c > cap) { flags[1] = 1; c = cap; }
            nnb[i] = c;
            xref[i] = x[i];
            xref[n + i] = x[n + i];
            xref[2 * n + i] = x[2 * n + i];
        }
        __syncthreads();
    }
    }
    // the last block to finish (every block has read the request, the set index and its
    // bins) clears the request and, with per-step bins, the counts of the bins it used;
    // rotate: advances the set.  No fence: the last block reads nothing that the other
    // blocks wrote (the device-scope counter aside), and every block's reads of what it
    // clears returned before its barrier (their values are in LDS); a device-scope fence
    // here would write back the L2 of the block's XCD once per cell
    __shared__ int last;
    __syncthreads();
    if (tid == 0) last = atomicAdd(reinterpret_cast<unsigned*>(&state[1]), 1u) == gridDim.x - 1;
    __syncthreads();
    if (!last) return;
    if (clear_bins)
        for (int c = tid; c < ncells; c += BUILD_THREADS) bin_count[c] = 0;
    if (tid == 0) {
        state[1] = 0;
        if (rotate) state[0] = (bset + 1) % 3;
        req[0] = 0;
    }
}

// small boxes (fewer than 3 cells per side): every atom scans all atoms
template <bool WATER>
__device__ __forceinline__ void build_allpairs_atom(const double4* __restrict__ pos,
                                                    const double* __restrict__ x, int i, int n,
                                                    double L, double invL, double rlist2, int G,
                                                    int cap, void* __restrict__ nbr,
                                                    int nbr_wide, int* __restrict__ nnb,
                                                    double* __restrict__ xref,
                                                    int* __restrict__ flags) {
    if (i == 0) atomicAdd(&flags[2], 1);
    const double4 pi = pos[i];
    const size_t stride = (size_t)n * G;
    const int lgG = __builtin_ctz((unsigned)G);
    int c = 0;
    for (int j = 0; j < n; ++j) {
        if (j == i) continue;
        if (WATER && (j / 3) == (i / 3)) continue;
        const double4 pj = pos[j];
        double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        dx -= L * rint(dx * invL);
        dy -= L * rint(dy * invL);
        dz -= L * rint(dz * invL);
        if (dx * dx + dy * dy + dz * dz < rlist2) {
            if (c < cap)
                mx_nbr_store(nbr, nbr_wide, (size_t)(c >> lgG) * stride + (size_t)i * G + (c & (G - 1)), j);
            ++c;
        }
    }
    if (c > cap) { flags[1] = 1; c = cap; }
    nnb[i] = c;
    xref[i] = x[i];
    xref[n + i] = x[n + i];
    xref[2 * n + i] = x[2 * n + i];
}

template <bool WATER>
__global__ void k_build_allpairs(const double4* __restrict__ pos, const double* __restrict__ x,
                                 int n, double L, double invL, double rlist2, int G, int cap,
                                 void* __restrict__ nbr, int nbr_wide, int* __restrict__ nnb,
                                 double* __restrict__ xref, int* __restrict__ flags,
                                 int* __restrict__ req) {
    if (req[0] == 0) return;   // uniform: only this kernel's last block clears the request
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        build_allpairs_atom<WATER>(pos, x, i, n, L, invL, rlist2, G, cap, nbr, nbr_wide, nnb,
                                   xref, flags);
    __shared__ int last;
    __syncthreads();
    if (threadIdx.x == 0)   // no fence: the last block only clears what every block read
        last = atomicAdd(reinterpret_cast<unsigned*>(&flags[NBX_FLAG_BINSET + 1]), 1u) ==
               gridDim.x - 1;
    __syncthreads();
    if (last && threadIdx.x == 0) {
        flags[NBX_FLAG_BINSET + 1] = 0;
        req[0] = 0;
    }
}

template <bool WATER>
static void launch_build_t(hipStream_t s, const nbx_ctx* c, const void* pos, const double* x,
                           int* req, int* bin_count, int* bin_atoms, int clear_bins) {
    const nbx_nlist& nl = c->nl;
    const int n = (int)c->n;
    const double L = c->bc.b[0];
    const int nc = nl.ncell;
    const double r2 = nl.rlist * nl.rlist;
    void* nbr = nl.d_nbr;
    if (nc >= 3) {
        const int ncells = nc * nc * nc;
        const float r2f = (float)r2 * 1.0001f;
        const int rotate = nl.bin_mode == NBX_BINS_BUILD ? 1 : 0;
        // at least ~512 blocks (two per CU) when the cells are few: up to 8 per cell
        int parts = (512 + ncells - 1) / ncells;
        parts = parts < 1 ? 1 : (parts > 8 ? 8 : parts);
        const dim3 grid(ncells * parts);
        // per-step bins hold every atom in the cell of its current position, so a stencil
        // cell's atoms are shifted once to their image next to the home cell; bins made
        // earlier (set-up bins of 3 cells per side, bins of the previous build) need the
        // minimum image per candidate
        if (nl.bin_mode == NBX_BINS_STEP)
            hipLaunchKernelGGL((k_build_cells<WATER, true>), grid, dim3(BUILD_THREADS), 0, s,
                               (const double4*)pos, x, n, nc, L, bin_count,
                               bin_atoms, nl.ccap, (double)nc / L, r2f, nl.lanes, nl.cap, nbr,
                               nl.nbr_wide, nl.d_nnb, nl.d_xref, nl.d_flags, req, clear_bins,
                               rotate, parts);
        else
            hipLaunchKernelGGL((k_build_cells<WATER, false>), grid, dim3(BUILD_THREADS), 0, s,
                               (const double4*)pos, x, n, nc, L, bin_count,
                               bin_atoms, nl.ccap, (double)nc / L, r2f, nl.lanes, nl.cap, nbr,
                               nl.nbr_wide, nl.d_nnb, nl.d_xref, nl.d_flags, req, clear_bins,
                               rotate, parts);
    } else {
        hipLaunchKernelGGL((k_build_allpairs<WATER>), dim3(nbx_blocks(n, 128)), dim3(128), 0,
                           s, (const double4*)pos, x, n, L, 1.0 / L, r2, nl.lanes, nl.cap, nbr,
                           nl.nbr_wide, nl.d_nnb, nl.d_xref, nl.d_flags, req);
    }
}

void mx_launch_build(hipStream_t s, const nbx_ctx* c, int water, const void* pos,
                     const double* x, int* req, int* bin_count, int* bin_atoms,
                     int clear_bins) {
    if (c->n <= 0) return;
    if (water) launch_build_t<true>(s, c, pos, x, req, bin_count, bin_atoms, clear_bins);
    else launch_build_t<false>(s, c, pos, x, req, bin_count, bin_atoms, clear_bins);
}

// ---------------------------------------------------------------------------
// SPC/Fw molecule kernel: bonds + angle (FP64), Berendsen, end of step, drift.
// One thread per molecule (O, H1, H2 = 3m, 3m+1, 3m+2); the non-bonded accelerations
// are already in io.a.  Formulas of src/basic_potentials.jl:367-433.
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

// sum over the block's 3 * WATER_MOLS threads, the same value in every thread and the same
// order in every block: a butterfly within each wavefront (a + b == b + a, so every lane
// ends with the same sum), then the three wavefront sums in order
#define WATER_MOLS 64
__device__ __forceinline__ double water_block_sum(double s) {
    __shared__ double ws[3];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    return (ws[0] + ws[1]) + ws[2];
}

// One thread per atom, 64 molecules (192 atoms) per block: the block's positions are
// shared through LDS, every thread evaluates its molecule's bonds and angle (the same
// expressions in the same order for all three) and keeps its own atom's terms, then ends
// its atom's step.  Andersen draws one uniform per molecule (counter of its O atom).
// With kp.T_parts (Berendsen) every block adds the partials of m vh^2 that the drift
// before this step left (in block order of that drift), so no kernel between the steps
// finishes the temperature.
__global__ void __launch_bounds__(3 * WATER_MOLS)
k_water_atoms(int n, double rOH, double kbond, double aHOH, double ka, mx_kick_params kp,
              mx_io io) {
    mx_resolve_step(kp);
    __shared__ double sx[3][3 * WATER_MOLS];
    const int t = threadIdx.x;
    const int i = 3 * WATER_MOLS * blockIdx.x + t;
    const bool act = i < n;
    // every input of the atom is requested before the barrier (one round trip to memory
    // for the whole kernel: it has 12,288 threads for water's 4,096 molecules, too few
    // waves to hide several)
    mx_atom A;
    double anb[3] = {0.0, 0.0, 0.0};
    if (act) {
        mx_load_atom(i, n, io, A);
#pragma unroll
        for (int k = 0; k < 3; ++k) anb[k] = io.a[(size_t)k * n + i];
    }
    // the partial count and one partial per thread are requested with the atom's inputs
    // (the buffer holds at least 4,095 partials, so partial t is in bounds)
    const bool fold = kp.T_parts != nullptr;
    int tnb = 0;
    double tpart = 0.0;
    if (fold) {
        tnb = *reinterpret_cast<const int*>(kp.T_parts);
        tpart = kp.T_parts[1 + t];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) sx[k][t] = act ? io.x_in[(size_t)k * n + i] : 0.0;
    __syncthreads();
    double Tb = 0.0;
    if (fold) {
        double ps = t < tnb ? tpart : 0.0;
        for (int b = t + 3 * WATER_MOLS; b < tnb; b += 3 * WATER_MOLS) ps += kp.T_parts[1 + b];
        Tb = water_block_sum(ps) / kp.T_denom;
    }
    double tval = 0.0;
    if (act) {
        const int w = t % 3;               // 0: O, 1: H1, 2: H2
        const int b = t - w;               // the molecule's O in this block
        const int o = i - w;
        double r[3][3];
#pragma unroll
        for (int q = 0; q < 3; ++q)
#pragma unroll
            for (int k = 0; k < 3; ++k) r[q][k] = sx[k][b + q];
        double f[3] = {0, 0, 0};
        if (w == 0) {
            mol_bond(r[0], r[1], rOH, kbond, f);
            mol_bond(r[0], r[2], rOH, kbond, f);
        } else {
            mol_bond(r[w], r[0], rOH, kbond, f);
        }
        const double mi = A.m;
        const double im = 1.0 / mi;
        double acc[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) acc[k] = anb[k] + im * f[k];
        // angle H1-O-H2 (src/basic_potentials.jl:395-433)
        double rba[3], rbc[3], rcb[3];
        for (int k = 0; k < 3; ++k) {
            rba[k] = r[1][k] - r[0][k];
            rbc[k] = r[2][k] - r[0][k];
            rcb[k] = r[0][k] - r[2][k];
        }
        const double X[3] = {rba[1] * rbc[2] - rba[2] * rbc[1], rba[2] * rbc[0] - rba[0] * rbc[2],
                             rba[0] * rbc[1] - rba[1] * rbc[0]};
        const double ta[3] = {rba[1] * X[2] - rba[2] * X[1], rba[2] * X[0] - rba[0] * X[2],
                              rba[0] * X[1] - rba[1] * X[0]};
        const double tc[3] = {rcb[1] * X[2] - rcb[2] * X[1], rcb[2] * X[0] - rcb[0] * X[2],
                              rcb[0] * X[1] - rcb[1] * X[0]};
        const double ia = 1.0 / sqrt(ta[0] * ta[0] + ta[1] * ta[1] + ta[2] * ta[2]);
        const double ic = 1.0 / sqrt(tc[0] * tc[0] + tc[1] * tc[1] + tc[2] * tc[2]);
        const double nba = sqrt(rba[0] * rba[0] + rba[1] * rba[1] + rba[2] * rba[2]);
        const double nbc = sqrt(rbc[0] * rbc[0] + rbc[1] * rbc[1] + rbc[2] * rbc[2]);
        double cosine = (rba[0] * rbc[0] + rba[1] * rbc[1] + rba[2] * rbc[2]) / (nba * nbc);
        cosine = cosine > 1.0 ? 1.0 : (cosine < -1.0 ? -1.0 : cosine);
        const double force = -ka * (acos(cosine) - aHOH);
        // the common factors divided once (the kernel is one dependent FP64 chain per
        // thread: each IEEE division is about 30 instructions on it)
        const double sa = ia * force / nba, sc = ic * force / nbc;
        for (int k = 0; k < 3; ++k) {
            const double fa = sa * ta[k];
            const double fc = sc * tc[k];
            if (w == 1) acc[k] += fa * im;
            else if (w == 0) acc[k] -= (fa + fc) * im;
            else acc[k] += fc * im;
        }
        if (mx_therm_force(kp)) {
            const double cb = fold ? mx_berendsen_from_T(kp, Tb) : mx_berendsen_coeff(kp);
            for (int k = 0; k < 3; ++k) acc[k] += cb * A.vh[k];
        }
        for (int k = 0; k < 3; ++k) io.a[(size_t)k * n + i] = acc[k];
        if (kp.do_kick) {
            double v3[3];
            mx_kick_atom(i, acc[0], acc[1], acc[2], A, kp, v3);
            if (kp.therm == NBX_THERM_ANDERSEN &&
                nbx_rng_uniform(kp.seed, mx_atom_id(kp, o), kp.step, NBX_RNG_ANDERSEN) < kp.prob)
                mx_andersen_atom(i, mi, kp, v3);
            io.v[i] = v3[0];
            io.v[n + i] = v3[1];
            io.v[2 * n + i] = v3[2];
            if (io.do_drift) tval = mx_drift_atom(i, n, v3, acc[0], acc[1], acc[2], A, kp, io);
        }
    }
    mx_epilogue_waves(tval, io, n);
}

void mx_launch_water_molecules(hipStream_t s, const nbx_ctx* c, const mx_kick_params& kp,
                               const mx_io& io) {
    const int64_t nmol = c->n / 3;
    if (nmol <= 0) return;
    hipLaunchKernelGGL(k_water_atoms, dim3(nbx_blocks(nmol, WATER_MOLS)), dim3(3 * WATER_MOLS),
                       0, s, (int)c->n, c->rOH, c->kbond, c->aHOH, c->kangle, kp, io);
}

// ---------------------------------------------------------------------------
// Temperature of a velocity array (prologue Berendsen input): per-block partial sums
// of m v^2, then one block adds them in block order.
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
__global__ void k_count_pairs_list(const double4* __restrict__ pos, const void* __restrict__ nbr,
                                   int nbr_wide,
                                   const int* __restrict__ nnb, int n, int G, double L,
                                   double invL, double R2lj, double R2el, int has_lj,
                                   int has_el, unsigned long long* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 pi = pos[i];
    const size_t stride = (size_t)n * G;
    const int lgG = __builtin_ctz((unsigned)G);
    unsigned long long cnt = 0;
    for (int k = 0; k < nnb[i]; ++k) {
        const int j = mx_nbr_load(nbr, nbr_wide, (size_t)(k >> lgG) * stride + (size_t)i * G + (k & (G - 1)));
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
                           const void* pos, const mx_pair_params& pp,
                           unsigned long long* d_count) {
    const int n = (int)c->n;
    if (L.water)
        hipLaunchKernelGGL(k_count_pairs_list<true>, dim3(nbx_blocks(n, 256)), dim3(256), 0, s,
                           (const double4*)pos, c->nl.d_nbr, c->nl.nbr_wide, c->nl.d_nnb, n, c->nl.lanes,
                           pp.L, pp.invL, (double)pp.lj_R2, (double)pp.el_R2, L.has_lj,
                           L.el_kind != 0, d_count);
    else
        hipLaunchKernelGGL(k_count_pairs_list<false>, dim3(nbx_blocks(n, 256)), dim3(256), 0, s,
                           (const double4*)pos, c->nl.d_nbr, c->nl.nbr_wide, c->nl.d_nnb, n, c->nl.lanes,
                           pp.L, pp.invL, (double)pp.lj_R2, (double)pp.el_R2, L.has_lj,
                           L.el_kind != 0, d_count);
}
