// nbx_learn_ctx.h -- parameter blocks of the learned-stride kernels that work on a
// context's state (descriptor, inference, corrected drift, targets, energy monitor).
#pragma once
#include "nbx_learn.h"
#include "nbx_mixed.h"

struct lr_feat_params {
    const double4* pos;   // wrapped positions (list mode)
    const void* nbr;      // Verlet list (lane-interleaved; 16- or 32-bit entries)
    int nbr_wide;         // 1: 32-bit entries
    const int* nnb;
    int n;
    double L, invL;
    const double* a;      // SoA accelerations and velocities
    const double* v;
    const double* x;      // SoA positions (snapshot mode)
    const int* sel;       // atoms to describe (null: all n)
    int nsel;
    float* frames;        // [count][6] e1, e2
    float* fraw;          // [count][35] raw features, or null
    double* snap_x;       // [count][9] x, v, a of the described atoms, or null
    uint16_t* ffrag;      // bf16 normalized features in layer-1 fragment order, or null
    const float* norm;    // normalization (ffrag)
};

struct lr_infer_params {
    const uint16_t* ffrag;
    const float* frames;
    const float* params;  // FP32 master parameters (biases)
    const float* norm;    // output scales at [70..75]
    const uint16_t* frag; // forward fragments
    int n;
    float* dlab;          // [n][6] lab-frame dx, dv
    double* part;         // [blocks][6]
};

struct lr_target_params {
    const int* sel;
    int nsel;
    int n;
    const double* snap;   // [nsel][9]
    const double* x1;     // SoA state after the stride
    const double* v1;
    const double* a1;
    double D;             // stride length in time
    const float* frames;  // [nsel][6] frames of the snapshot
    const float* fstage;  // [nsel][35] features of the snapshot
    float* X;             // dataset features
    float* T;             // dataset targets
    int64_t row0;
};

struct lr_energy_params {
    const double4* pos;
    const void* nbr;
    int nbr_wide;
    const int* nnb;
    int n;
    double L, invL;
    double s2, R2;        // sigma^2, cutoff^2
    double uc;            // (s/R)^12 - (s/R)^6
    double c4eps;         // 4 eps
    const double* v;
    const double* ms;
    double* part;         // [blocks][2]
};

void lr_launch_features(hipStream_t s, int lanes, const lr_feat_params& fp);
int lr_infer_blocks(int64_t n);
void lr_launch_mlp_infer(hipStream_t s, const lr_infer_params& ip);
void lr_launch_drift(hipStream_t s, const double* part, int nparts, const float* dlab, double D,
                     const mx_io& io, int64_t n);
void lr_launch_targets(hipStream_t s, const lr_target_params& tp);
void lr_launch_select(hipStream_t s, int* sel, int nsel, int64_t n, uint64_t seed, int64_t frame);
int lr_energy_blocks(int64_t n, int lanes);
void lr_launch_energy(hipStream_t s, int lanes, const lr_energy_params& ep, double* out);
