// nbx_learn.hip -- kernels of the learned stride (no counterpart in the reference):
//   * k_lr_chain   : MLP forward (and backward) for a batch of samples, one wavefront per
//                    32 samples, activations kept in MFMA registers between layers
//   * k_lr_dw      : weight gradients dW = dZ^T H over sample chunks (BF16 MFMA)
//   * k_lr_adam    : partial-sum reduction + Adam on the FP32 master weights + BF16
//                    fragment refresh
//   * k_lr_features: radial descriptor + force-frame velocity from the Verlet list
//   * k_lr_mlp_infer: persistent inference with the forward fragments staged in LDS,
//                    corrections rotated back to the lab frame
//   * k_lr_drift   : momentum-neutral corrected drift of the 8-dt step
//   * k_lr_targets, k_lr_energy: training targets and the energy-drift monitor
#include <hip/hip_runtime.h>

#include "nbx_learn.h"
#include "nbx_learn_ctx.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint16_t lr_bits(__bf16 b) { return __builtin_bit_cast(uint16_t, b); }
__device__ __forceinline__ float lr_bf2f(__bf16 b) {
    return __uint_as_float((uint32_t)lr_bits(b) << 16);
}

__device__ __forceinline__ f32x16 lr_mfma(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// logistic function from the hardware 2^x and reciprocal (relative error ~1e-6, far
// below the BF16 rounding that follows it)
__device__ __forceinline__ float lr_sigmoid(float z) {
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.44269504088896341f * z));
}

// One 128-output layer with its epilogue.  The four M tiles accumulate side by side (kt
// outer, mt inner: four independent MFMA chains in flight, 64 accumulator registers), so
// no MFMA waits for the previous one's result; then, per M tile, acc = sum_kt W(mt, kt)
// Bin(kt) and per register r (unit u = 32 mt + drow(h, r)):
//   forward  (BWD = false): z = acc + bias[u], Bout = bf16(SiLU(z)); optionally z (D
//            order) and the activation unit-major for the weight gradients
//   backward (BWD = true) : dz = acc * SiLU'(z) with z from zsave; Bout = bf16(dz) and the
//            unit-major copy
// Bout(2 mt + r / 8) slot r % 8 is then the next layer's B operand.
template <int KT, bool BWD>
__device__ __forceinline__ void lr_layer(const bf16x8* __restrict__ W, int fbase, const bf16x8* Bin,
                                         bf16x8* Bout, const float* __restrict__ bias,
                                         float* zsave, uint16_t* tU, int B, int bs, int h, int lane) {
    f32x16 acc[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) acc[mt] = f32x16{};
#pragma unroll
    for (int kt = 0; kt < KT; ++kt)
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
            acc[mt] = lr_mfma(W[(fbase + mt * KT + kt) * 64 + lane], Bin[kt], acc[mt]);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        const f32x16 a = acc[mt];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            bf16x8 t;
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                const int r = q * 8 + s;
                const int u = mt * 32 + lr_drow(h, r);
                float val;
                if (BWD) {
                    const float z = zsave[(mt * 16 + r) * 64 + lane];
                    const float sg = lr_sigmoid(z);
                    val = a[r] * (sg * (1.0f + z * (1.0f - sg)));
                } else {
                    const float z = a[r] + bias[u];
                    if (zsave) zsave[(mt * 16 + r) * 64 + lane] = z;
                    val = z * lr_sigmoid(z);
                }
                const __bf16 vb = (__bf16)val;
                t[s] = vb;
                if (tU) tU[(size_t)u * B + bs] = lr_bits(vb);
            }
            if (Bout) Bout[2 * mt + q] = t;
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

// ---------------------------------------------------------------------------------
// Training chain: one wavefront per 32 samples, 4 wavefronts per block, no block
// synchronisation (a wavefront past the last tile returns at once).
// backward = 0: forward only, outputs to b.out[B][6].
// ---------------------------------------------------------------------------------
template <bool backward>
__global__ void __launch_bounds__(256) k_lr_chain(lr_train_bufs b, float gscale) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ntiles = b.B >> 5;
    const int tile = blockIdx.x * 4 + wave;
    if (tile >= ntiles) return;
    const int h = lane >> 5, j = lane & 31;
    const int B = b.B;
    const int bs = tile * 32 + j;
    const int64_t d = b.idx ? (int64_t)b.idx[bs] : (int64_t)bs;
    const bf16x8* __restrict__ F = reinterpret_cast<const bf16x8*>(b.frag);
    const float* __restrict__ P = b.params;
    float* z1 = backward ? b.z + ((size_t)(0 * ntiles + tile) * 64 * 64) : nullptr;
    float* z2 = backward ? b.z + ((size_t)(1 * ntiles + tile) * 64 * 64) : nullptr;
    float* z3 = backward ? b.z + ((size_t)(2 * ntiles + tile) * 64 * 64) : nullptr;
    uint16_t* tH1 = backward ? b.tH : nullptr;
    uint16_t* tH2 = backward ? b.tH + (size_t)128 * B : nullptr;
    uint16_t* tH3 = backward ? b.tH + (size_t)256 * B : nullptr;

    // layer-1 input: normalized features in the k order of the W1 fragments
    bf16x8 Bx[3];
#pragma unroll
    for (int kt = 0; kt < 3; ++kt) {
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const int f = kt * 16 + lr_drow(h, s);
            float xv = 0.0f;
            if (f < LR_NIN) xv = (b.X[d * LR_NIN + f] - b.norm[f]) * b.norm[LR_NIN + f];
            const __bf16 xb = (__bf16)xv;
            Bx[kt][s] = xb;
            if (backward) b.tH0[(size_t)f * B + bs] = lr_bits(xb);
        }
    }
    bf16x8 Bh[8], Bg[8];
    lr_layer<3, false>(F, LR_F1, Bx, Bh, P + LR_B1, z1, tH1, B, bs, h, lane);
    lr_layer<8, false>(F, LR_F2, Bh, Bg, P + LR_B2, z2, tH2, B, bs, h, lane);
    lr_layer<8, false>(F, LR_F3, Bg, Bh, P + LR_B3, z3, tH3, B, bs, h, lane);
    f32x16 y = {};
#pragma unroll
    for (int kt = 0; kt < 8; ++kt) y = lr_mfma(F[(LR_F4 + kt) * 64 + lane], Bh[kt], y);

    if constexpr (!backward) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = lr_drow(h, r);
            if (row < LR_NOUT) b.out[(size_t)bs * LR_NOUT + row] = y[r] + P[LR_B4 + row];
        }
        return;
    }
    // output gradient of mean((y - t/scale)^2) over B x 6
    bf16x8 Bd[1];
    float l2 = 0.0f;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int row = lr_drow(h, r);
        float g = 0.0f;
        if (row < LR_NOUT) {
            const float yv = y[r] + P[LR_B4 + row];
            const float tv = b.T[d * LR_NOUT + row] / b.norm[2 * LR_NIN + row];
            const float e = yv - tv;
            l2 += e * e;
            g = gscale * e;
        }
        const __bf16 gb = (__bf16)g;
        Bd[0][r] = gb;
        if (row < LR_NOUT) b.tdY[(size_t)row * B + bs] = lr_bits(gb);
    }
    // dH3 = W4^T dY (one k tile), dZ3; dH2 = W3^T dZ3, dZ2; dH1 = W2^T dZ2, dZ1
    uint16_t* tdZ1 = b.tdZ;
    uint16_t* tdZ2 = b.tdZ + (size_t)128 * B;
    uint16_t* tdZ3 = b.tdZ + (size_t)256 * B;
    lr_layer<1, true>(F, LR_T4, Bd, Bh, nullptr, z3, tdZ3, B, bs, h, lane);
    lr_layer<8, true>(F, LR_T3, Bh, Bg, nullptr, z2, tdZ2, B, bs, h, lane);
    lr_layer<8, true>(F, LR_T2, Bg, nullptr, nullptr, z1, tdZ1, B, bs, h, lane);
    // per-tile loss
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) l2 += __shfl_xor(l2, off);
    if (lane == 0) b.loss[tile] = l2;
}

// ---------------------------------------------------------------------------------
// Weight gradients.  Block (t, chunk), one wavefront: a 32 x 32 tile of dW_l summed over
// LR_CHUNK samples; bias gradients from the A operand of the nt == 0 tiles.  Tiles:
// t < 8 layer 1 (4 x 2), < 24 layer 2 (4 x 4), < 40 layer 3 (4 x 4), < 44 layer 4 (1 x 4).
// ---------------------------------------------------------------------------------
__global__ void __launch_bounds__(64) k_lr_dw(lr_train_bufs b) {
    const int t = blockIdx.x, chunk = blockIdx.y;
    const int lane = threadIdx.x, h = lane >> 5, j = lane & 31;
    const int B = b.B;
    int mt, nt, Mout, Kin, Woff, Boff;
    const uint16_t* A;
    const uint16_t* Bm;
    if (t < 8) {
        mt = t >> 1; nt = t & 1; Mout = LR_NH; Kin = LR_NIN; Woff = LR_W1; Boff = LR_B1;
        A = b.tdZ; Bm = b.tH0;
    } else if (t < 24) {
        const int u = t - 8;
        mt = u >> 2; nt = u & 3; Mout = LR_NH; Kin = LR_NH; Woff = LR_W2; Boff = LR_B2;
        A = b.tdZ + (size_t)128 * B; Bm = b.tH;
    } else if (t < 40) {
        const int u = t - 24;
        mt = u >> 2; nt = u & 3; Mout = LR_NH; Kin = LR_NH; Woff = LR_W3; Boff = LR_B3;
        A = b.tdZ + (size_t)256 * B; Bm = b.tH + (size_t)128 * B;
    } else {
        mt = 0; nt = t - 40; Mout = LR_NOUT; Kin = LR_NH; Woff = LR_W4; Boff = LR_B4;
        A = b.tdY; Bm = b.tH + (size_t)256 * B;
    }
    const int s0 = chunk * LR_CHUNK;
    const bf16x8* Arow = reinterpret_cast<const bf16x8*>(A + (size_t)(mt * 32 + j) * B + s0 + h * 8);
    const bf16x8* Brow = reinterpret_cast<const bf16x8*>(Bm + (size_t)(nt * 32 + j) * B + s0 + h * 8);
    f32x16 acc = {};
    float dbs = 0.0f;
    for (int k = 0; k < LR_CHUNK / 16; ++k) {
        const bf16x8 a = Arow[2 * k];
        const bf16x8 bb = Brow[2 * k];
        acc = lr_mfma(a, bb, acc);
        if (nt == 0) {
#pragma unroll
            for (int s = 0; s < 8; ++s) dbs += lr_bf2f(a[s]);
        }
    }
    float* out = b.part + (size_t)chunk * LR_NPARAM;
    const int ci = nt * 32 + j;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int o = mt * 32 + lr_drow(h, r);
        if (o < Mout && ci < Kin) out[Woff + o * Kin + ci] = acc[r];
    }
    if (nt == 0) {
        dbs += __shfl_down(dbs, 32);
        const int o = mt * 32 + j;
        if (h == 0 && o < Mout) out[Boff + o] = dbs;
    }
}

// BF16 fragment entries of parameter p with value w
__device__ __forceinline__ void lr_write_frag(int p, float w, uint16_t* frag) {
    const uint16_t bits = lr_bits((__bf16)w);
    if (p < LR_B1) {
        frag[lr_frag_fwd(1, p / LR_NIN, p % LR_NIN)] = bits;
    } else if (p >= LR_W2 && p < LR_B2) {
        const int q = p - LR_W2, o = q / LR_NH, i = q % LR_NH;
        frag[lr_frag_fwd(2, o, i)] = bits;
        frag[lr_frag_bwd(2, o, i)] = bits;
    } else if (p >= LR_W3 && p < LR_B3) {
        const int q = p - LR_W3, o = q / LR_NH, i = q % LR_NH;
        frag[lr_frag_fwd(3, o, i)] = bits;
        frag[lr_frag_bwd(3, o, i)] = bits;
    } else if (p >= LR_W4 && p < LR_B4) {
        const int q = p - LR_W4, o = q / LR_NH, i = q % LR_NH;
        frag[lr_frag_fwd(4, o, i)] = bits;
        frag[lr_frag_bwd(4, o, i)] = bits;
    }
}

__global__ void k_lr_adam(int nchunks, const float* __restrict__ part, float* __restrict__ P,
                          float* __restrict__ M, float* __restrict__ V, uint16_t* __restrict__ frag,
                          float lr, float b1, float b2, float eps, float c1, float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= LR_NPARAM) return;
    float g = 0.0f;
    for (int c = 0; c < nchunks; ++c) g += part[(size_t)c * LR_NPARAM + p];
    const float m = b1 * M[p] + (1.0f - b1) * g;
    const float v = b2 * V[p] + (1.0f - b2) * g * g;
    M[p] = m;
    V[p] = v;
    const float w = P[p] - lr * (m * c1) / (sqrtf(v * c2) + eps);
    P[p] = w;
    lr_write_frag(p, w, frag);
}

__global__ void k_lr_reduce(int nchunks, const float* __restrict__ part, float* __restrict__ grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= LR_NPARAM) return;
    float g = 0.0f;
    for (int c = 0; c < nchunks; ++c) g += part[(size_t)c * LR_NPARAM + p];
    grad[p] = g;
}

__global__ void k_lr_write_frags(const float* __restrict__ P, uint16_t* __restrict__ frag) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= LR_NPARAM) return;
    lr_write_frag(p, P[p], frag);
}

__global__ void k_lr_gather_index(int* idx, int B, int64_t count, uint32_t s0, uint32_t s1,
                                  int64_t iter) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= B) return;
    uint32_t r[4];
    nbx_philox((uint32_t)q, (uint32_t)iter, (uint32_t)(iter >> 32), 0x6C6561u, s0, s1, r);
    const uint64_t u = ((uint64_t)r[0] << 32) | r[1];
    idx[q] = (int)(u % (uint64_t)count);
}

__global__ void k_lr_loss_sum(const float* __restrict__ loss, int ntiles, float scale,
                              float* __restrict__ out) {
    __shared__ float red[256];
    float acc = 0.0f;
    for (int q = threadIdx.x; q < ntiles; q += 256) acc += loss[q];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0] * scale;
}

void lr_launch_loss_sum(hipStream_t s, const float* loss, int ntiles, float scale, float* out) {
    hipLaunchKernelGGL(k_lr_loss_sum, dim3(1), dim3(256), 0, s, loss, ntiles, scale, out);
}

void lr_launch_train_chain(hipStream_t s, const lr_train_bufs& b, int backward, float gscale) {
    const int ntiles = b.B >> 5;
    if (backward)
        hipLaunchKernelGGL(k_lr_chain<true>, dim3((ntiles + 3) / 4), dim3(256), 0, s, b, gscale);
    else
        hipLaunchKernelGGL(k_lr_chain<false>, dim3((ntiles + 3) / 4), dim3(256), 0, s, b, gscale);
}

void lr_launch_dw(hipStream_t s, const lr_train_bufs& b) {
    hipLaunchKernelGGL(k_lr_dw, dim3(44, b.B / LR_CHUNK), dim3(64), 0, s, b);
}

void lr_launch_adam(hipStream_t s, int nchunks, const float* part, float* params, float* m,
                    float* v, uint16_t* frag, float lr, float b1, float b2, float eps, int step) {
    const float c1 = 1.0f / (1.0f - powf(b1, (float)step));
    const float c2 = 1.0f / (1.0f - powf(b2, (float)step));
    hipLaunchKernelGGL(k_lr_adam, dim3((LR_NPARAM + 255) / 256), dim3(256), 0, s, nchunks, part,
                       params, m, v, frag, lr, b1, b2, eps, c1, c2);
}

void lr_launch_reduce(hipStream_t s, int nchunks, const float* part, float* grad) {
    hipLaunchKernelGGL(k_lr_reduce, dim3((LR_NPARAM + 255) / 256), dim3(256), 0, s, nchunks, part,
                       grad);
}

void lr_launch_write_frags(hipStream_t s, const float* params, uint16_t* frag) {
    hipLaunchKernelGGL(k_lr_write_frags, dim3((LR_NPARAM + 255) / 256), dim3(256), 0, s, params,
                       frag);
}

void lr_launch_gather_index(hipStream_t s, int* idx, int B, int64_t count, uint64_t seed,
                            int64_t iter) {
    hipLaunchKernelGGL(k_lr_gather_index, dim3((B + 255) / 256), dim3(256), 0, s, idx, B, count,
                       (uint32_t)seed, (uint32_t)(seed >> 32), iter);
}

// =================================================================================
// Context side: descriptor, inference, corrected drift, targets, energy monitor
// =================================================================================

// Gaussian centre of bin b (0.8 .. 2.5 sigma in 31 equal steps)
__device__ __forceinline__ float lr_centre(int b) { return (float)(0.8 + 1.7 * (double)b / 31.0); }

// force frame and frame features of one atom (FP64)
__device__ __forceinline__ void lr_frame(double ax, double ay, double az, double vx, double vy,
                                         double vz, double* e1, double* e2, double* ff) {
    const double an = sqrt(ax * ax + ay * ay + az * az);
    if (an > 0.0) { e1[0] = ax / an; e1[1] = ay / an; e1[2] = az / an; }
    else { e1[0] = 1.0; e1[1] = 0.0; e1[2] = 0.0; }
    const double vp = vx * e1[0] + vy * e1[1] + vz * e1[2];
    const double wx = vx - vp * e1[0], wy = vy - vp * e1[1], wz = vz - vp * e1[2];
    const double wn = sqrt(wx * wx + wy * wy + wz * wz);
    if (wn > 0.0) { e2[0] = wx / wn; e2[1] = wy / wn; e2[2] = wz / wn; }
    else {
        // the axis least aligned with e1, made normal to it
        const double ux = fabs(e1[0]), uy = fabs(e1[1]), uz = fabs(e1[2]);
        double c[3] = {0.0, 0.0, 0.0};
        if (ux <= uy && ux <= uz) c[0] = 1.0; else if (uy <= uz) c[1] = 1.0; else c[2] = 1.0;
        const double cp = c[0] * e1[0] + c[1] * e1[1] + c[2] * e1[2];
        double qx = c[0] - cp * e1[0], qy = c[1] - cp * e1[1], qz = c[2] - cp * e1[2];
        const double qn = sqrt(qx * qx + qy * qy + qz * qz);
        e2[0] = qx / qn; e2[1] = qy / qn; e2[2] = qz / qn;
    }
    ff[0] = vp;
    ff[1] = wn;
    ff[2] = an;
}

template <int G>
__global__ void __launch_bounds__(256) k_lr_features(lr_feat_params fp) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int q = gid / G;
    const int lane = gid & (G - 1);
    const int count = fp.sel ? fp.nsel : fp.n;
    if (q >= count) return;  // the G lanes of one atom leave together
    const int n = fp.n;
    const int i = fp.sel ? fp.sel[q] : q;
    const double4 pi = fp.pos[i];
    const int cnt = fp.nnb[i];
    float g[LR_NBINS];
#pragma unroll
    for (int b = 0; b < LR_NBINS; ++b) g[b] = 0.0f;
    const size_t stride = (size_t)n * G;
    size_t e = (size_t)i * G + lane;
    for (int k = lane; k < cnt; k += G, e += stride) {
        const double4 pj = fp.pos[mx_nbr_load(fp.nbr, fp.nbr_wide, e)];
        double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        dx -= fp.L * rint(dx * fp.invL);
        dy -= fp.L * rint(dy * fp.invL);
        dz -= fp.L * rint(dz * fp.invL);
        const double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 < 6.25) {
            const float r = sqrtf((float)r2);
            const float sw = r < 2.0f ? 1.0f : 0.5f * (1.0f + cosf(3.14159265358979f * (r - 2.0f) * 2.0f));
            // exp(-50 e^2) = 2^(-50 log2(e) e^2)
#pragma unroll
            for (int b = 0; b < LR_NBINS; ++b) {
                const float e = r - lr_centre(b);
                g[b] += sw * __builtin_amdgcn_exp2f(-72.1347520444482f * e * e);
            }
        }
    }
#pragma unroll
    for (int b = 0; b < LR_NBINS; ++b) {
#pragma unroll
        for (int off = G >> 1; off > 0; off >>= 1) g[b] += __shfl_xor(g[b], off, G);
    }
    if (lane != 0) return;
    const double ax = fp.a[i], ay = fp.a[n + i], az = fp.a[2 * n + i];
    const double vx = fp.v[i], vy = fp.v[n + i], vz = fp.v[2 * n + i];
    double e1[3], e2[3], ff[3];
    lr_frame(ax, ay, az, vx, vy, vz, e1, e2, ff);
    float* fr = fp.frames + (size_t)q * 6;
    fr[0] = (float)e1[0]; fr[1] = (float)e1[1]; fr[2] = (float)e1[2];
    fr[3] = (float)e2[0]; fr[4] = (float)e2[1]; fr[5] = (float)e2[2];
    if (fp.fraw) {
        float* o = fp.fraw + (size_t)q * LR_NIN;
#pragma unroll
        for (int b = 0; b < LR_NBINS; ++b) o[b] = g[b];
        o[32] = (float)ff[0]; o[33] = (float)ff[1]; o[34] = (float)ff[2];
    }
    if (fp.snap_x) {
        double* sx = fp.snap_x + (size_t)q * 9;
        sx[0] = fp.x[i]; sx[1] = fp.x[n + i]; sx[2] = fp.x[2 * n + i];
        sx[3] = vx; sx[4] = vy; sx[5] = vz;
        sx[6] = ax; sx[7] = ay; sx[8] = az;
    }
    if (fp.ffrag) {
        const float* mean = fp.norm;
        const float* inv = fp.norm + LR_NIN;
#pragma unroll
        for (int f = 0; f < 48; ++f) {
            float xv = 0.0f;
            if (f < LR_NBINS) xv = g[f];
            else if (f < LR_NIN) xv = (float)ff[f - LR_NBINS];
            if (f < LR_NIN) xv = (xv - mean[f]) * inv[f];
            fp.ffrag[lr_feat_frag(i, f)] = lr_bits((__bf16)xv);
        }
    }
}

void lr_launch_features(hipStream_t s, int lanes, const lr_feat_params& fp) {
    const int count = fp.sel ? fp.nsel : fp.n;
    if (count <= 0) return;
    const int64_t threads = (int64_t)count * lanes;
    const dim3 grid((unsigned)((threads + 255) / 256));
    switch (lanes) {
    case 1: hipLaunchKernelGGL(k_lr_features<1>, grid, dim3(256), 0, s, fp); break;
    case 2: hipLaunchKernelGGL(k_lr_features<2>, grid, dim3(256), 0, s, fp); break;
    case 4: hipLaunchKernelGGL(k_lr_features<4>, grid, dim3(256), 0, s, fp); break;
    case 16: hipLaunchKernelGGL(k_lr_features<16>, grid, dim3(256), 0, s, fp); break;
    case 32: hipLaunchKernelGGL(k_lr_features<32>, grid, dim3(256), 0, s, fp); break;
    default: hipLaunchKernelGGL(k_lr_features<8>, grid, dim3(256), 0, s, fp); break;
    }
}

// Persistent inference: forward fragments staged once per block in LDS (84 KiB), one
// wavefront per 32 atoms.  Outputs the lab-frame corrections dlab[i][6] (dx, dv) and per
// block sums of them (FP64) for the momentum-neutral drift.
__global__ void __launch_bounds__(256) k_lr_mlp_infer(lr_infer_params ip) {
    __shared__ bf16x8 sW[LR_NFWD * 64];
    __shared__ double sred[4][6];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int h = lane >> 5, j = lane & 31;
    const bf16x8* gW = reinterpret_cast<const bf16x8*>(ip.frag);
    for (int q = threadIdx.x; q < LR_NFWD * 64; q += 256) sW[q] = gW[q];
    __syncthreads();
    const float* __restrict__ P = ip.params;
    const float* osc = ip.norm + 2 * LR_NIN;
    const bf16x8* X = reinterpret_cast<const bf16x8*>(ip.ffrag);
    const int ntiles = (ip.n + 31) >> 5;
    double ps[3] = {0.0, 0.0, 0.0};
    for (int tile = blockIdx.x * 4 + wave; tile < ntiles; tile += gridDim.x * 4) {
        bf16x8 Bx[3];
#pragma unroll
        for (int kt = 0; kt < 3; ++kt) Bx[kt] = X[(tile * 3 + kt) * 64 + lane];
        bf16x8 Bh[8], Bg[8];
        lr_layer<3, false>(sW, LR_F1, Bx, Bh, P + LR_B1, nullptr, nullptr, 0, 0, h, lane);
        lr_layer<8, false>(sW, LR_F2, Bh, Bg, P + LR_B2, nullptr, nullptr, 0, 0, h, lane);
        lr_layer<8, false>(sW, LR_F3, Bg, Bh, P + LR_B3, nullptr, nullptr, 0, 0, h, lane);
        f32x16 y = {};
#pragma unroll
        for (int kt = 0; kt < 8; ++kt) y = lr_mfma(sW[(LR_F4 + kt) * 64 + lane], Bh[kt], y);
        // h = 0 lanes hold outputs 0..3, h = 1 lanes outputs 4, 5 (registers 0, 1)
        float o[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = lr_drow(h, r);
            o[r] = row < LR_NOUT ? (y[r] + P[LR_B4 + row]) * osc[row] : 0.0f;
        }
        const float dv0 = __shfl(o[3], j);  // output 3 of the same atom, from its h = 0 lane
        const int i = tile * 32 + j;
        if (i < ip.n) {
            const float* fr = ip.frames + (size_t)i * 6;
            const float e1x = fr[0], e1y = fr[1], e1z = fr[2];
            const float e2x = fr[3], e2y = fr[4], e2z = fr[5];
            const float e3x = e1y * e2z - e1z * e2y;
            const float e3y = e1z * e2x - e1x * e2z;
            const float e3z = e1x * e2y - e1y * e2x;
            float c0, c1, c2;
            if (h == 0) { c0 = o[0]; c1 = o[1]; c2 = o[2]; }
            else { c0 = dv0; c1 = o[0]; c2 = o[1]; }
            const float lx = c0 * e1x + c1 * e2x + c2 * e3x;
            const float ly = c0 * e1y + c1 * e2y + c2 * e3y;
            const float lz = c0 * e1z + c1 * e2z + c2 * e3z;
            float* out = ip.dlab + (size_t)i * 6 + 3 * h;
            out[0] = lx; out[1] = ly; out[2] = lz;
            ps[0] += lx; ps[1] += ly; ps[2] += lz;
        }
    }
    // block sums: components 0..2 from h = 0 lanes, 3..5 from h = 1 lanes
    double t6[6];
#pragma unroll
    for (int c = 0; c < 3; ++c) { t6[c] = h == 0 ? ps[c] : 0.0; t6[3 + c] = h == 1 ? ps[c] : 0.0; }
#pragma unroll
    for (int c = 0; c < 6; ++c) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) t6[c] += __shfl_xor(t6[c], off);
    }
    if (lane == 0) {
#pragma unroll
        for (int c = 0; c < 6; ++c) sred[wave][c] = t6[c];
    }
    __syncthreads();
    if (threadIdx.x < 6)
        ip.part[(size_t)blockIdx.x * 6 + threadIdx.x] =
            sred[0][threadIdx.x] + sred[1][threadIdx.x] + sred[2][threadIdx.x] + sred[3][threadIdx.x];
}

int lr_infer_blocks(int64_t n) {
    const int64_t tiles = (n + 31) / 32;
    int64_t blocks = (tiles + 3) / 4;
    if (blocks > 256) blocks = 256;
    return (int)(blocks < 1 ? 1 : blocks);
}

void lr_launch_mlp_infer(hipStream_t s, const lr_infer_params& ip) {
    if (ip.n <= 0) return;
    hipLaunchKernelGGL(k_lr_mlp_infer, dim3(lr_infer_blocks(ip.n)), dim3(256), 0, s, ip);
}

// Corrected drift of the strided step: x' = x + D v + D^2 a/2 + dx, vh = v + D a/2 + dv,
// with the mean correction removed (total momentum and centre of mass unchanged by the
// network); wrapped copy, skin test.
__global__ void __launch_bounds__(256) k_lr_drift(const double* __restrict__ part, int nparts,
                                                  const float* __restrict__ dlab, double D,
                                                  mx_io io, int n) {
    __shared__ double smean[6];
    if (threadIdx.x < 64) {
        double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (int q = threadIdx.x; q < nparts; q += 64) {
#pragma unroll
            for (int c = 0; c < 6; ++c) acc[c] += part[(size_t)q * 6 + c];
        }
#pragma unroll
        for (int c = 0; c < 6; ++c) {
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) acc[c] += __shfl_xor(acc[c], off);
        }
        if (threadIdx.x < 6) {
            double m = acc[0];
#pragma unroll
            for (int c = 1; c < 6; ++c) if ((int)threadIdx.x == c) m = acc[c];
            smean[threadIdx.x] = m / (double)n;
        }
    }
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* dl = dlab + (size_t)i * 6;
    double v3[3], a3[3], p3[3], h3[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        v3[k] = io.v[(size_t)k * n + i];
        a3[k] = io.a[(size_t)k * n + i];
        p3[k] = io.x_in[(size_t)k * n + i] + D * v3[k] + (D * D) * (0.5 * a3[k]) +
                ((double)dl[k] - smean[k]);
        h3[k] = v3[k] + D * (0.5 * a3[k]) + ((double)dl[3 + k] - smean[3 + k]);
        io.x_out[(size_t)k * n + i] = p3[k];
        io.vh[(size_t)k * n + i] = h3[k];
    }
    const double qi = io.qs != nullptr ? io.qs[i] : 0.0;
    ((double4*)io.pos_out)[i] =
        make_double4(mx_wrap(p3[0], io.L), mx_wrap(p3[1], io.L), mx_wrap(p3[2], io.L), qi);
    mx_posq(io.pos_out, n)[i] = mx_fixed_position(p3[0], p3[1], p3[2], io.invL, qi);
    const double ex = p3[0] - io.xref[i], ey = p3[1] - io.xref[n + i], ez = p3[2] - io.xref[2 * n + i];
    const double e2 = ex * ex + ey * ey + ez * ez;
    if (e2 > io.half_skin2) io.req_out[0] = 1;
    if (e2 > io.guard2) io.flags[1] = 1;   // past what the bins of the last build cover
}

void lr_launch_drift(hipStream_t s, const double* part, int nparts, const float* dlab, double D,
                     const mx_io& io, int64_t n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_lr_drift, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, part, nparts,
                       dlab, D, io, (int)n);
}

// Training targets of the sampled atoms from their snapshot at t (x, v, a) and the state
// at t + D (SoA), in the snapshot's force frame; the features are copied alongside.
__global__ void k_lr_targets(lr_target_params tp) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= tp.nsel) return;
    const int n = tp.n;
    const int i = tp.sel[q];
    const double* s9 = tp.snap + (size_t)q * 9;
    const double D = tp.D;
    double dx[3], dv[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double x0 = s9[k], v0 = s9[3 + k], a0 = s9[6 + k];
        dx[k] = tp.x1[(size_t)k * n + i] - (x0 + D * v0 + (D * D) * (0.5 * a0));
        dv[k] = tp.v1[(size_t)k * n + i] - (v0 + D * (0.5 * (a0 + tp.a1[(size_t)k * n + i])));
    }
    const float* fr = tp.frames + (size_t)q * 6;
    const double e1[3] = {fr[0], fr[1], fr[2]};
    const double e2[3] = {fr[3], fr[4], fr[5]};
    const double e3[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                          e1[0] * e2[1] - e1[1] * e2[0]};
    float* T = tp.T + (size_t)(tp.row0 + q) * LR_NOUT;
    T[0] = (float)(dx[0] * e1[0] + dx[1] * e1[1] + dx[2] * e1[2]);
    T[1] = (float)(dx[0] * e2[0] + dx[1] * e2[1] + dx[2] * e2[2]);
    T[2] = (float)(dx[0] * e3[0] + dx[1] * e3[1] + dx[2] * e3[2]);
    T[3] = (float)(dv[0] * e1[0] + dv[1] * e1[1] + dv[2] * e1[2]);
    T[4] = (float)(dv[0] * e2[0] + dv[1] * e2[1] + dv[2] * e2[2]);
    T[5] = (float)(dv[0] * e3[0] + dv[1] * e3[1] + dv[2] * e3[2]);
    float* X = tp.X + (size_t)(tp.row0 + q) * LR_NIN;
    const float* f = tp.fstage + (size_t)q * LR_NIN;
#pragma unroll
    for (int k = 0; k < LR_NIN; ++k) X[k] = f[k];
}

void lr_launch_targets(hipStream_t s, const lr_target_params& tp) {
    if (tp.nsel <= 0) return;
    hipLaunchKernelGGL(k_lr_targets, dim3((tp.nsel + 255) / 256), dim3(256), 0, s, tp);
}

__global__ void k_lr_select(int* sel, int nsel, int n, uint32_t s0, uint32_t s1, int64_t frame) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nsel) return;
    uint32_t r[4];
    nbx_philox((uint32_t)q, (uint32_t)frame, (uint32_t)(frame >> 32), 0x73656Cu, s0, s1, r);
    sel[q] = (int)(r[0] % (uint32_t)n);
}

void lr_launch_select(hipStream_t s, int* sel, int nsel, int64_t n, uint64_t seed, int64_t frame) {
    if (nsel <= 0) return;
    hipLaunchKernelGGL(k_lr_select, dim3((nsel + 255) / 256), dim3(256), 0, s, sel, nsel, (int)n,
                       (uint32_t)seed, (uint32_t)(seed >> 32), frame);
}

// Energy monitor: per-block FP64 sums of kinetic energy and the shifted Lennard-Jones
// energy 4 eps sum_{r < R} [(s/r)^12 - (s/r)^6 - (s/R)^12 + (s/R)^6] over the list (each
// ordered pair contributes half).
template <int G>
__global__ void __launch_bounds__(256) k_lr_energy(lr_energy_params ep) {
    __shared__ double sk[256], sp[256];
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = gid / G, lane = gid & (G - 1);
    const int n = ep.n;
    double ke = 0.0, pe = 0.0;
    if (i < n) {
        const double4 pi = ep.pos[i];
        const int cnt = ep.nnb[i];
        const size_t stride = (size_t)n * G;
        size_t e = (size_t)gid;
        for (int k = lane; k < cnt; k += G, e += stride) {
            const double4 pj = ep.pos[mx_nbr_load(ep.nbr, ep.nbr_wide, e)];
            double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            dx -= ep.L * rint(dx * ep.invL);
            dy -= ep.L * rint(dy * ep.invL);
            dz -= ep.L * rint(dz * ep.invL);
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 < ep.R2) {
                const double s6 = ep.s2 * ep.s2 * ep.s2 / (r2 * r2 * r2);
                pe += s6 * s6 - s6 - ep.uc;
            }
        }
        if (lane == 0) {
            const double vx = ep.v[i], vy = ep.v[n + i], vz = ep.v[2 * n + i];
            ke = 0.5 * ep.ms[i] * (vx * vx + vy * vy + vz * vz);
        }
        pe *= 0.5 * ep.c4eps;
    }
    sk[threadIdx.x] = ke;
    sp[threadIdx.x] = pe;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            sk[threadIdx.x] += sk[threadIdx.x + off];
            sp[threadIdx.x] += sp[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        ep.part[2 * blockIdx.x] = sk[0];
        ep.part[2 * blockIdx.x + 1] = sp[0];
    }
}

__global__ void __launch_bounds__(256) k_lr_energy_final(const double* part, int nparts,
                                                         double* out) {
    __shared__ double sk[256], sp[256];
    double k = 0.0, p = 0.0;
    for (int q = threadIdx.x; q < nparts; q += 256) { k += part[2 * q]; p += part[2 * q + 1]; }
    sk[threadIdx.x] = k;
    sp[threadIdx.x] = p;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            sk[threadIdx.x] += sk[threadIdx.x + off];
            sp[threadIdx.x] += sp[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = sk[0];
        out[1] = sp[0];
    }
}

int lr_energy_blocks(int64_t n, int lanes) { return (int)((n * lanes + 255) / 256); }

void lr_launch_energy(hipStream_t s, int lanes, const lr_energy_params& ep, double* out) {
    if (ep.n <= 0) return;
    const int blocks = lr_energy_blocks(ep.n, lanes);
    switch (lanes) {
    case 1: hipLaunchKernelGGL(k_lr_energy<1>, dim3(blocks), dim3(256), 0, s, ep); break;
    case 2: hipLaunchKernelGGL(k_lr_energy<2>, dim3(blocks), dim3(256), 0, s, ep); break;
    case 4: hipLaunchKernelGGL(k_lr_energy<4>, dim3(blocks), dim3(256), 0, s, ep); break;
    case 16: hipLaunchKernelGGL(k_lr_energy<16>, dim3(blocks), dim3(256), 0, s, ep); break;
    case 32: hipLaunchKernelGGL(k_lr_energy<32>, dim3(blocks), dim3(256), 0, s, ep); break;
    default: hipLaunchKernelGGL(k_lr_energy<8>, dim3(blocks), dim3(256), 0, s, ep); break;
    }
    hipLaunchKernelGGL(k_lr_energy_final, dim3(1), dim3(256), 0, s, ep.part, blocks, out);
}
