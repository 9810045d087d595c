// nbx_gravity.hip -- all-pairs gravity (src/basic_potentials.jl:306-331) in mixed
// precision on gfx950.
//
// Layout: pos = float4 (x, y, z, m) per body, written by the drift kernel from the FP64
// state.  Grid = (target blocks) x (source chunks): a block owns TPT*256 targets and
// one contiguous chunk of sources; it stages 256 sources at a time in LDS and every
// lane reads the same LDS entry (a broadcast read) per inner iteration.  Within a tile
// the per-target sums are plain FP32; tile sums are Kahan-added into the per-target
// accumulator, and chunk partials are stored in FP64 and added by the kick kernel in
// chunk order (deterministic).
//   a_i = G * sum_j m_j (r_j - r_i) / |r_j - r_i|^3      (= -G m_j r_ij / |r_ij|^3)
// The self term is removed by adding 1e-24 to r^2 inside the FMA chain (no extra
// instruction): for j == i the displacement is exactly zero and 1/r^3 = 1e36 is finite,
// so the contribution is exactly zero.  The mass in .w is m / m_max <= 1 (k_grav_finish
// multiplies by G m_max), so m_j / r^3 stays finite for any mass scale; for distinct
// bodies the 1e-24 rounds away unless r^2 < 2^24 * 1e-24 (r < 4e-9 in the caller's
// length unit).
// When every body has the same mass m (EQM), the tiles sum (r_j - r_i) / |r_j - r_i|^3
// and k_grav_finish multiplies by G m in FP64: one multiply fewer per pair (12 VALU
// instructions and one v_rsq_f32 instead of 13 and one).
#include "nbx_mixed.h"

#define GR_THREADS 256
#define GR_TILE 256

template <int TPT, bool EQM>
__global__ void __launch_bounds__(GR_THREADS)
k_grav_tiled(const float4* __restrict__ pos, int n, int jchunk, double* __restrict__ part) {
    __shared__ float4 tile[GR_TILE];
    const int tid = threadIdx.x;
    const int chunk = blockIdx.y;
    const int ibase = blockIdx.x * (GR_THREADS * TPT) + tid;
    float xi[TPT], yi[TPT], zi[TPT];
    nbx_kahan ax[TPT], ay[TPT], az[TPT];
#pragma unroll
    for (int t = 0; t < TPT; ++t) {
        const int i = ibase + t * GR_THREADS;
        const float4 p = i < n ? pos[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        xi[t] = p.x; yi[t] = p.y; zi[t] = p.z;
        ax[t].init(); ay[t].init(); az[t].init();
    }
    const int j0 = chunk * jchunk;
    const int j1 = (j0 + jchunk) < n ? (j0 + jchunk) : n;
    for (int base = j0; base < j1; base += GR_TILE) {
        const int j = base + tid;
        tile[tid] = j < j1 ? pos[j] : make_float4(0.f, 0.f, 0.f, 0.f);
        __syncthreads();
        float sx[TPT], sy[TPT], sz[TPT];
#pragma unroll
        for (int t = 0; t < TPT; ++t) { sx[t] = 0.f; sy[t] = 0.f; sz[t] = 0.f; }
        // one target per thread: 16 sources per unrolled iteration, so the scheduler
        // interleaves more independent r^-3 chains around each v_rsq_f32 (measured 6 %
        // faster than 8 at N = 65,536, at 33 VGPRs)
        constexpr int unroll = TPT == 1 ? 16 : 8;
#pragma unroll unroll
        for (int q = 0; q < GR_TILE; ++q) {
            const float4 pj = tile[q];
#pragma unroll
            for (int t = 0; t < TPT; ++t) {
                const float dx = pj.x - xi[t], dy = pj.y - yi[t], dz = pj.z - zi[t];
                // r^2 + 1e-24 inside the FMA chain: the self term (dx = dy = dz = 0) gets a
                // finite 1/r^3 (1e36, times a mass factor <= 1) times a zero displacement
                const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, 1e-24f)));
                const float rinv = __builtin_amdgcn_rsqf(r2);
                const float s = EQM ? rinv * rinv * rinv : pj.w * (rinv * rinv * rinv);
                sx[t] += dx * s;
                sy[t] += dy * s;
                sz[t] += dz * s;
            }
        }
#pragma unroll
        for (int t = 0; t < TPT; ++t) { ax[t].add(sx[t]); ay[t].add(sy[t]); az[t].add(sz[t]); }
        __syncthreads();
    }
#pragma unroll
    for (int t = 0; t < TPT; ++t) {
        const int i = ibase + t * GR_THREADS;
        if (i < n) {
            double* pc = part + (size_t)chunk * 3 * n;
            pc[i] = ax[t].value();
            pc[n + i] = ay[t].value();
            pc[2 * n + i] = az[t].value();
        }
    }
}

// a_i = gscale * sum over chunks (chunk order), then the end of the step and the drift;
// gscale = G m_max (the tiles summed m_j / m_max), or G m when the tiles left out the
// common mass m
__global__ void __launch_bounds__(256)
k_grav_finish(const double* __restrict__ part, int chunks, int n, double gscale,
              mx_kick_params kp, mx_io io) {
    mx_resolve_step(kp);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double tval = 0.0;
    if (i < n) {
        mx_atom A;
        mx_load_atom(i, n, io, A);
        double s[3] = {0.0, 0.0, 0.0};
        for (int c = 0; c < chunks; ++c) {
            const double* pc = part + (size_t)c * 3 * n;
            s[0] += pc[i];
            s[1] += pc[n + i];
            s[2] += pc[2 * n + i];
        }
        double ax = gscale * s[0], ay = gscale * s[1], az = gscale * s[2];
        if (mx_therm_force(kp)) {
            const double cb = mx_berendsen_coeff(kp);
            ax += cb * A.vh[0]; ay += cb * A.vh[1]; az += cb * A.vh[2];
        }
        io.a[i] = ax; io.a[n + i] = ay; io.a[2 * n + i] = az;
        if (kp.do_kick) {
            double v3[3];
            mx_kick_atom(i, ax, ay, az, A, kp, v3);
            if (kp.therm == NBX_THERM_ANDERSEN &&
                nbx_rng_uniform(kp.seed, mx_atom_id(kp, i), kp.step, NBX_RNG_ANDERSEN) < kp.prob)
                mx_andersen_atom(i, kp.m0, kp, v3);
            io.v[i] = v3[0]; io.v[n + i] = v3[1]; io.v[2 * n + i] = v3[2];
            if (io.do_drift) tval = mx_drift_atom(i, n, v3, ax, ay, az, A, kp, io);
        }
    }
    mx_epilogue(tval, io, n);
}

static int grav_tpt(int64_t n) { return n >= 16384 ? 2 : 1; }
static int grav_tpt_of(const nbx_ctx* c) { return c->grav_tpt_set ? c->grav_tpt_set : grav_tpt(c->n); }

// ev0/ev1 (optional): HIP events recorded around the k_grav_tiled launch
void mx_launch_gravity(hipStream_t s, const nbx_ctx* c, const void* pos, double G,
                       const mx_kick_params& kp, const mx_io& io, int force_only_kernel,
                       hipEvent_t ev0, hipEvent_t ev1) {
    const int n = (int)c->n;
    if (n <= 0) return;
    const int tpt = grav_tpt_of(c);
    const int chunks = c->grav_chunks;
    const int jchunk = (int)((n + chunks - 1) / chunks);
    const dim3 grid(nbx_blocks(n, GR_THREADS * tpt), chunks);
    const float4* p4 = (const float4*)pos;
    const bool eqm = c->equal_masses;
    if (ev0) (void)hipEventRecord(ev0, s);
    if (tpt == 4) {
        if (eqm) hipLaunchKernelGGL((k_grav_tiled<4, true>), grid, dim3(GR_THREADS), 0, s, p4, n, jchunk, c->d_part);
        else hipLaunchKernelGGL((k_grav_tiled<4, false>), grid, dim3(GR_THREADS), 0, s, p4, n, jchunk, c->d_part);
    } else if (tpt == 2) {
        if (eqm) hipLaunchKernelGGL((k_grav_tiled<2, true>), grid, dim3(GR_THREADS), 0, s, p4, n, jchunk, c->d_part);
        else hipLaunchKernelGGL((k_grav_tiled<2, false>), grid, dim3(GR_THREADS), 0, s, p4, n, jchunk, c->d_part);
    } else {
        if (eqm) hipLaunchKernelGGL((k_grav_tiled<1, true>), grid, dim3(GR_THREADS), 0, s, p4, n, jchunk, c->d_part);
        else hipLaunchKernelGGL((k_grav_tiled<1, false>), grid, dim3(GR_THREADS), 0, s, p4, n, jchunk, c->d_part);
    }
    if (ev1) (void)hipEventRecord(ev1, s);
    if (force_only_kernel) return;
    const double scale = eqm ? G * c->h_ms[0] : G / c->mass_wscale;
    hipLaunchKernelGGL(k_grav_finish, dim3(nbx_blocks(n, 256)), dim3(256), 0, s, c->d_part,
                       chunks, n, scale, kp, io);
}

// number of source chunks: enough blocks to fill 256 CUs several times over, with
// chunks of at least 2048 sources
int mx_gravity_chunks(const nbx_ctx* c) {
    const int64_t n = c->n;
    if (c->grav_chunks_set) return c->grav_chunks_set;
    const int tpt = grav_tpt_of(c);
    const int64_t tblocks = (n + GR_THREADS * tpt - 1) / (GR_THREADS * tpt);
    int chunks = 1;
    while (tblocks * chunks < 2048 && (n / (chunks * 2)) >= 2048) chunks *= 2;
    return chunks;
}

// ---------------------------------------------------------------------------
// All-pairs magnetic dipoles (src/basic_potentials.jl:333-365) in mixed precision: FP32
// pair law on float4 positions and float4 moments staged 256 at a time in LDS, plain FP32
// sums per tile, Kahan across tiles; the per-target sum divided by m_i goes to `part`
// (FP64), and k_grav_finish applies 3 mu/4pi, the kick and the drift.  No boundary
// conditions (as the reference); j == i skipped.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(GR_THREADS)
k_mag_tiled(const float4* __restrict__ pos, const float4* __restrict__ mm, int n,
            double* __restrict__ part) {
    __shared__ float4 tp[GR_TILE];
    __shared__ float4 tm[GR_TILE];
    const int tid = threadIdx.x;
    const int i = blockIdx.x * GR_THREADS + tid;
    const float4 pi = i < n ? pos[i] : make_float4(0.f, 0.f, 0.f, 1.f);
    const float4 mi = i < n ? mm[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    nbx_kahan ax, ay, az;
    ax.init(); ay.init(); az.init();
    for (int base = 0; base < n; base += GR_TILE) {
        const int j = base + tid;
        tp[tid] = j < n ? pos[j] : make_float4(0.f, 0.f, 0.f, 0.f);
        tm[tid] = j < n ? mm[j] : make_float4(0.f, 0.f, 0.f, 0.f);
        __syncthreads();
        const int lim = (n - base) < GR_TILE ? (n - base) : GR_TILE;
        float sx = 0.f, sy = 0.f, sz = 0.f;
        for (int q = 0; q < lim; ++q) {
            if (base + q == i) continue;
            const float4 pj = tp[q];
            const float4 mj = tm[q];
            const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
            const float d = rx * rx + ry * ry + rz * rz;
            const float inv = __builtin_amdgcn_rsqf(d);
            const float inv2 = inv * inv;
            const float inv4 = inv2 * inv2;
            const float r1 = rx * inv, r2 = ry * inv, r3 = rz * inv;
            const float mir = mi.x * r1 + mi.y * r2 + mi.z * r3;
            const float mij = mj.x * r1 + mj.y * r2 + mj.z * r3;
            const float dmm = mi.x * mj.x + mi.y * mj.y + mi.z * mj.z;
            const float f5 = 5.0f * mir * mij;
            sx += (mi.x * mij + mj.x * mir + r1 * (dmm - f5)) * inv4;
            sy += (mi.y * mij + mj.y * mir + r2 * (dmm - f5)) * inv4;
            sz += (mi.z * mij + mj.z * mir + r3 * (dmm - f5)) * inv4;
        }
        ax.add(sx); ay.add(sy); az.add(sz);
        __syncthreads();
    }
    if (i < n) {
        const double im = 1.0 / (double)pi.w;
        part[i] = ax.value() * im;
        part[n + i] = ay.value() * im;
        part[2 * n + i] = az.value() * im;
    }
}

void mx_launch_magnetic(hipStream_t s, const nbx_ctx* c, const void* pos, double mu_4pi,
                        const mx_kick_params& kp, const mx_io& io, int force_only_kernel) {
    const int n = (int)c->n;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_mag_tiled, dim3(nbx_blocks(n, GR_THREADS)), dim3(GR_THREADS), 0, s,
                       (const float4*)pos, (const float4*)c->d_mm4, n, c->d_part);
    if (force_only_kernel) return;
    hipLaunchKernelGGL(k_grav_finish, dim3(nbx_blocks(n, 256)), dim3(256), 0, s, c->d_part, 1, n,
                       3.0 * mu_4pi, kp, io);
}

// ---------------------------------------------------------------------------
// Plain device-to-device copy: the calibration kernel for HBM byte counters and the
// achievable-bandwidth probe.  Reads `n16` 16-byte words from src, writes them to dst.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_copy_probe(const float4* __restrict__ src,
                                                    float4* __restrict__ dst, int64_t n16) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
        dst[i] = src[i];
}

// VALU issue-rate calibration: 8 independent chains per lane of one instruction, its
// encoding pinned by inline asm (0: v_fmac_f32, 1: v_mul_f32, 2: v_rsq_f32)
#define VP_ITERS 2048
template <int KIND>
__global__ void __launch_bounds__(256) k_valu_probe(float* __restrict__ out, float b0, float c0) {
    float a[8];
    const float b = b0 + threadIdx.x * 1e-9f, cc = c0;
#pragma unroll
    for (int k = 0; k < 8; ++k) a[k] = 1.0f + threadIdx.x * 1e-3f + k;
    for (int it = 0; it < VP_ITERS; ++it) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (KIND == 0) asm volatile("v_fmac_f32 %0, %1, %2" : "+v"(a[k]) : "v"(b), "v"(cc));
            if (KIND == 1) asm volatile("v_mul_f32 %0, %0, %1" : "+v"(a[k]) : "v"(b));
            if (KIND == 2) asm volatile("v_rsq_f32 %0, %0" : "+v"(a[k]));
        }
    }
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < 8; ++k) s += a[k];
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

int mx_valu_probe(nbx_ctx* c, int kind, int reps, double* rate) {
    const int blocks = 8192, threads = 256;
    float* out = nullptr;
    if (hipMalloc((void**)&out, sizeof(float) * blocks * threads) != hipSuccess)
        return nbx_fail(c, NBX_ERR_DEVICE, "valu probe: out of memory");
    int cus = 0;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device);
    hipStream_t s = c->stream;
    auto launch = [&]() {
        if (kind == 0) hipLaunchKernelGGL(k_valu_probe<0>, dim3(blocks), dim3(threads), 0, s, out, 0.999f, 1e-3f);
        else if (kind == 1) hipLaunchKernelGGL(k_valu_probe<1>, dim3(blocks), dim3(threads), 0, s, out, 0.999f, 1e-3f);
        else hipLaunchKernelGGL(k_valu_probe<2>, dim3(blocks), dim3(threads), 0, s, out, 0.999f, 1e-3f);
    };
    launch();
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    float best = 0.f;
    hipError_t err = hipSuccess;
    for (int r = 0; r < reps && err == hipSuccess; ++r) {
        (void)hipEventRecord(e0, s);
        launch();
        (void)hipEventRecord(e1, s);
        err = hipEventSynchronize(e1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (r == 0 || ms < best) best = ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(out);
    if (err != hipSuccess) return nbx_fail(c, NBX_ERR_DEVICE, hipGetErrorString(err));
    const double wave_instr = (double)blocks * threads / 64 * VP_ITERS * 8;
    *rate = wave_instr / (4.0 * (cus > 0 ? cus : 256)) / ((double)best * 1e6);
    return NBX_OK;
}

int mx_copy_probe(nbx_ctx* c, int64_t bytes, int iters, double* avg_ms) {
    const int64_t n16 = bytes / 16;
    void* a = nullptr;
    void* b = nullptr;
    if (hipMalloc(&a, n16 * 16) != hipSuccess) return nbx_fail(c, NBX_ERR_DEVICE, "copy probe: out of memory");
    if (hipMalloc(&b, n16 * 16) != hipSuccess) {
        (void)hipFree(a);
        return nbx_fail(c, NBX_ERR_DEVICE, "copy probe: out of memory");
    }
    hipStream_t s = c->stream;
    (void)hipMemsetAsync(a, 0, n16 * 16, s);
    const int blocks = 8192;
    hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, s, (const float4*)a, (float4*)b, n16);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, s);
    for (int k = 0; k < iters; ++k)
        hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, s, (const float4*)a, (float4*)b,
                           n16);
    (void)hipEventRecord(e1, s);
    hipError_t err = hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (err != hipSuccess) return nbx_fail(c, NBX_ERR_DEVICE, hipGetErrorString(err));
    *avg_ms = (double)ms / iters;
    return NBX_OK;
}
