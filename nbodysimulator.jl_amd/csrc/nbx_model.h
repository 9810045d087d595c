// nbx_model.h -- the learned-stride model object: FP32 master parameters with Adam state,
// BF16 MFMA fragments, normalization, the training set in HBM and training scratch.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "nbx.h"
#include "nbx_learn.h"

struct nbx_model {
    int device = 0;
    hipStream_t stream = nullptr;
    float* d_params = nullptr;     // [LR_NPARAM]
    float* d_m = nullptr;          // Adam first moment
    float* d_v = nullptr;          // Adam second moment
    uint16_t* d_frag = nullptr;    // [LR_NFRAG][512] bf16
    float* d_norm = nullptr;       // [LR_NNORM]
    double h_norm[LR_NNORM] = {};
    bool norm_set = false;
    int adam_step = 0;
    // training set
    float* d_X = nullptr;          // [cap][35]
    float* d_T = nullptr;          // [cap][6]
    int64_t count = 0, cap = 0;
    int64_t frames_collected = 0;  // selection counter of the collector
    // training scratch for batches of tb_B samples
    int tb_B = 0;
    int* d_idx = nullptr;
    uint16_t *d_tH0 = nullptr, *d_tH = nullptr, *d_tdZ = nullptr, *d_tdY = nullptr;
    float *d_z = nullptr, *d_loss = nullptr, *d_part = nullptr, *d_out = nullptr;
    float* d_hist = nullptr;       // per-iteration losses of one train call
    int64_t hist_cap = 0;
    std::string err;
};

int nbx_model_fail(nbx_model* m, int code, const std::string& msg);
// grow the training set to hold `need` rows
int lr_model_reserve(nbx_model* m, int64_t need);
