// nbx_learn.h -- the learned stride: per-atom descriptor, a 35-128-128-128-6 SiLU MLP on
// BF16 MFMA (FP32 accumulation), its training kernels, and the corrected 8-dt drift.
//
// Per-atom input (35): 32 Gaussian radial density bins (centres 0.8..2.5 sigma, standard
// deviation 0.1 sigma, cosine switch from 2.0 to 2.5 sigma over the neighbours of the
// Verlet list) followed by the atom's velocity expressed in its force frame
// (v . e1, |v - (v . e1) e1|) and |a|.  The frame is e1 = a/|a|, e2 = the unit
// component of v normal to e1, e3 = e1 x e2.
// Per-atom output (6): position and velocity corrections along (e1, e2, e3), in units
// of the per-component scale held with the model.
//
// MFMA layout (v_mfma_f32_32x32x16_bf16, 64 lanes, lane = 32 h + j):
//   A operand: lane (h, i) holds row i, k-slots s = 0..7
//   B operand: lane (h, j) holds column j, the same k-slots
//   D        : lane (h, j) holds column j, register r is row drow(h, r)
// The k index of slot s in a 16-wide k tile is kperm(h, s) = drow(h, s): a layer's D
// registers 0..7 / 8..15 of an M tile are then directly the next layer's B operand for
// the two k tiles of that M tile, so activations stay in registers between layers.
// Weights are stored as A fragments (64 lanes x 8 bf16 = 1 KiB each) in that order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define LR_NIN 35
#define LR_NH 128
#define LR_NOUT 6
#define LR_NPARAM 38406
#define LR_NNORM 76
#define LR_NBINS 32

// offsets of the flat FP32 parameter vector (W row-major [out][in])
#define LR_W1 0
#define LR_B1 4480
#define LR_W2 4608
#define LR_B2 20992
#define LR_W3 21120
#define LR_B3 37504
#define LR_W4 37632
#define LR_B4 38400

// fragment numbering
#define LR_F1 0     // W1: mt * 3 + kt      (12)
#define LR_F2 12    // W2: mt * 8 + kt      (32)
#define LR_F3 44    // W3: mt * 8 + kt      (32)
#define LR_F4 76    // W4: kt               (8)
#define LR_NFWD 84
#define LR_T4 84    // W4^T: mt             (4)
#define LR_T3 88    // W3^T: mt * 8 + kt    (32)
#define LR_T2 120   // W2^T: mt * 8 + kt    (32)
#define LR_NFRAG 152
#define LR_FRAG_ELEMS 512   // bf16 per fragment

// samples per weight-gradient chunk (a multiple of 16)
#define LR_CHUNK 256

__host__ __device__ inline int lr_drow(int h, int r) { return ((r >> 2) << 3) | (h << 2) | (r & 3); }

// element offset (in bf16) of W[o][i] of layer l (1..4) inside the forward fragments
__host__ __device__ inline int lr_frag_fwd(int l, int o, int i) {
    const int kt = i >> 4, w = i & 15;
    const int h = (w >> 2) & 1, s = ((w >> 3) << 2) | (w & 3);
    const int lane = h * 32 + (o & 31), mt = o >> 5;
    int f;
    if (l == 1) f = LR_F1 + mt * 3 + kt;
    else if (l == 2) f = LR_F2 + mt * 8 + kt;
    else if (l == 3) f = LR_F3 + mt * 8 + kt;
    else f = LR_F4 + kt;
    return (f * 64 + lane) * 8 + s;
}

// element offset of W[o][i] of layer l (2..4) inside the transposed (backward) fragments:
// rows = the layer's inputs i, k = its outputs o
__host__ __device__ inline int lr_frag_bwd(int l, int o, int i) {
    const int kt = o >> 4, w = o & 15;
    const int h = (w >> 2) & 1, s = ((w >> 3) << 2) | (w & 3);
    const int lane = h * 32 + (i & 31), mt = i >> 5;
    int f;
    if (l == 4) f = LR_T4 + mt;
    else if (l == 3) f = LR_T3 + mt * 8 + kt;
    else f = LR_T2 + mt * 8 + kt;
    return (f * 64 + lane) * 8 + s;
}

// element offset of feature f of atom i in the fragment-ordered bf16 input of layer 1
__host__ __device__ inline int64_t lr_feat_frag(int64_t i, int f) {
    const int64_t tile = i >> 5;
    const int j = (int)(i & 31), kt = f >> 4, w = f & 15;
    const int h = (w >> 2) & 1, s = ((w >> 3) << 2) | (w & 3);
    return ((tile * 3 + kt) * 64 + h * 32 + j) * 8 + s;
}

// Training-step buffers for a batch of B samples (B a multiple of LR_CHUNK)
struct lr_train_bufs {
    int B;
    const int* idx;          // [B] dataset rows
    const float* X;          // dataset features [count][35] (raw)
    const float* T;          // dataset targets [count][6] (raw)
    const float* norm;       // [76] mean[35], inv_scale[35], out_scale[6]
    const float* params;     // master FP32 parameters
    const uint16_t* frag;    // bf16 fragments (LR_NFRAG)
    uint16_t* tH0;           // [64][B]  normalized inputs (rows 35..63 zero)
    uint16_t* tH;            // [3][128][B] hidden activations
    uint16_t* tdZ;           // [3][128][B] pre-activation gradients
    uint16_t* tdY;           // [32][B] output gradients (rows 6..31 zero)
    float* z;                // [3][B/32][64 regs][64 lanes] pre-activations (D order)
    float* loss;             // [B/32] per-tile sums of squared errors
    float* part;             // [B/LR_CHUNK][LR_NPARAM] gradient partials
    float* out;              // [B][6] network outputs (forward-only mode) or null
};

// launchers (nbx_learn.hip)
void lr_launch_train_chain(hipStream_t s, const lr_train_bufs& b, int backward, float inv_count);
void lr_launch_dw(hipStream_t s, const lr_train_bufs& b);
void lr_launch_adam(hipStream_t s, int nchunks, const float* part, float* params, float* m,
                    float* v, uint16_t* frag, float lr, float b1, float b2, float eps, int step);
void lr_launch_reduce(hipStream_t s, int nchunks, const float* part, float* grad);
void lr_launch_write_frags(hipStream_t s, const float* params, uint16_t* frag);
void lr_launch_loss_sum(hipStream_t s, const float* loss, int ntiles, float scale, float* out);
void lr_launch_gather_index(hipStream_t s, int* idx, int B, int64_t count, uint64_t seed,
                            int64_t iter);
