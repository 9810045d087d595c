// nbx_learn.cpp -- C ABI of the learned-stride model (no counterpart in the reference):
// creation and initialisation, parameter I/O, forward and gradient on host-given samples,
// the training set, normalization, and Adam training with the BF16 MFMA kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "nbx_model.h"

namespace {
std::string g_model_create_error;

#define LR_HIP(m, call)                                                                 \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            return nbx_model_fail((m), NBX_ERR_DEVICE,                                  \
                                  std::string(#call ": ") + hipGetErrorString(e_));     \
    } while (0)

// splitmix64 (Steele, Lea, Flood 2014) for the initial weights
uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void free_train(nbx_model* m) {
    void* ps[] = {m->d_idx, m->d_tH0, m->d_tH, m->d_tdZ, m->d_tdY, m->d_z, m->d_loss, m->d_part,
                  m->d_out};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    m->d_idx = nullptr;
    m->d_tH0 = m->d_tH = m->d_tdZ = m->d_tdY = nullptr;
    m->d_z = m->d_loss = m->d_part = m->d_out = nullptr;
    m->tb_B = 0;
}

int alloc_train(nbx_model* m, int B) {
    if (m->tb_B == B) return NBX_OK;
    free_train(m);
    const size_t b = (size_t)B;
    hipStream_t s = m->stream;
    LR_HIP(m, hipMalloc((void**)&m->d_idx, sizeof(int) * b));
    LR_HIP(m, hipMalloc((void**)&m->d_tH0, 2 * 64 * b));
    LR_HIP(m, hipMalloc((void**)&m->d_tH, 2 * 3 * 128 * b));
    LR_HIP(m, hipMalloc((void**)&m->d_tdZ, 2 * 3 * 128 * b));
    LR_HIP(m, hipMalloc((void**)&m->d_tdY, 2 * 32 * b));
    LR_HIP(m, hipMalloc((void**)&m->d_z, sizeof(float) * 3 * 128 * b));
    LR_HIP(m, hipMalloc((void**)&m->d_loss, sizeof(float) * (b / 32)));
    LR_HIP(m, hipMalloc((void**)&m->d_part, sizeof(float) * (b / LR_CHUNK) * LR_NPARAM));
    LR_HIP(m, hipMalloc((void**)&m->d_out, sizeof(float) * 6 * b));
    // rows that no kernel writes (inputs 48..63, output gradients 6..31) stay zero
    LR_HIP(m, hipMemsetAsync(m->d_tH0, 0, 2 * 64 * b, s));
    LR_HIP(m, hipMemsetAsync(m->d_tdY, 0, 2 * 32 * b, s));
    m->tb_B = B;
    return NBX_OK;
}

lr_train_bufs bufs(const nbx_model* m, const float* X, const float* T, const int* idx) {
    lr_train_bufs b{};
    b.B = m->tb_B;
    b.idx = idx;
    b.X = X;
    b.T = T;
    b.norm = m->d_norm;
    b.params = m->d_params;
    b.frag = m->d_frag;
    b.tH0 = m->d_tH0;
    b.tH = m->d_tH;
    b.tdZ = m->d_tdZ;
    b.tdY = m->d_tdY;
    b.z = m->d_z;
    b.loss = m->d_loss;
    b.part = m->d_part;
    b.out = m->d_out;
    return b;
}

int upload_norm(nbx_model* m) {
    float f[LR_NNORM];
    for (int k = 0; k < LR_NNORM; ++k) f[k] = (float)m->h_norm[k];
    LR_HIP(m, hipMemcpyAsync(m->d_norm, f, sizeof(f), hipMemcpyHostToDevice, m->stream));
    LR_HIP(m, hipStreamSynchronize(m->stream));
    return NBX_OK;
}

int reset_adam(nbx_model* m) {
    LR_HIP(m, hipMemsetAsync(m->d_m, 0, sizeof(float) * LR_NPARAM, m->stream));
    LR_HIP(m, hipMemsetAsync(m->d_v, 0, sizeof(float) * LR_NPARAM, m->stream));
    m->adam_step = 0;
    return NBX_OK;
}

// features / targets from the host into temporary device rows (padded with zeros)
struct dev_rows {
    float* X = nullptr;
    float* T = nullptr;
    ~dev_rows() {
        if (X) (void)hipFree(X);
        if (T) (void)hipFree(T);
    }
};

int upload_rows(nbx_model* m, const double* feats, const double* targets, int64_t n, int64_t npad,
                dev_rows& r) {
    std::vector<float> hx((size_t)npad * LR_NIN, 0.0f), ht;
    for (int64_t k = 0; k < n * LR_NIN; ++k) hx[k] = (float)feats[k];
    LR_HIP(m, hipMalloc((void**)&r.X, sizeof(float) * hx.size()));
    LR_HIP(m, hipMemcpyAsync(r.X, hx.data(), sizeof(float) * hx.size(), hipMemcpyHostToDevice,
                             m->stream));
    if (targets) {
        ht.assign((size_t)npad * LR_NOUT, 0.0f);
        for (int64_t k = 0; k < n * LR_NOUT; ++k) ht[k] = (float)targets[k];
        LR_HIP(m, hipMalloc((void**)&r.T, sizeof(float) * ht.size()));
        LR_HIP(m, hipMemcpyAsync(r.T, ht.data(), sizeof(float) * ht.size(), hipMemcpyHostToDevice,
                                 m->stream));
    }
    LR_HIP(m, hipStreamSynchronize(m->stream));
    return NBX_OK;
}

bool finite_all(const double* p, int64_t n) {
    for (int64_t k = 0; k < n; ++k)
        if (!std::isfinite(p[k])) return false;
    return true;
}
}  // namespace

int nbx_model_fail(nbx_model* m, int code, const std::string& msg) {
    if (m) m->err = msg;
    else g_model_create_error = msg;
    return code;
}

int lr_model_reserve(nbx_model* m, int64_t need) {
    if (need <= m->cap) return NBX_OK;
    int64_t cap = std::max<int64_t>(need, std::max<int64_t>(2 * m->cap, 1 << 16));
    float* X = nullptr;
    float* T = nullptr;
    LR_HIP(m, hipMalloc((void**)&X, sizeof(float) * LR_NIN * (size_t)cap));
    if (hipMalloc((void**)&T, sizeof(float) * LR_NOUT * (size_t)cap) != hipSuccess) {
        (void)hipFree(X);
        return nbx_model_fail(m, NBX_ERR_DEVICE, "training set: out of device memory");
    }
    if (m->count > 0) {
        LR_HIP(m, hipMemcpyAsync(X, m->d_X, sizeof(float) * LR_NIN * (size_t)m->count,
                                 hipMemcpyDeviceToDevice, m->stream));
        LR_HIP(m, hipMemcpyAsync(T, m->d_T, sizeof(float) * LR_NOUT * (size_t)m->count,
                                 hipMemcpyDeviceToDevice, m->stream));
    }
    LR_HIP(m, hipStreamSynchronize(m->stream));
    if (m->d_X) (void)hipFree(m->d_X);
    if (m->d_T) (void)hipFree(m->d_T);
    m->d_X = X;
    m->d_T = T;
    m->cap = cap;
    return NBX_OK;
}

extern "C" int nbx_model_create(nbx_model** out, uint64_t seed) {
    if (!out) return NBX_ERR_INVALID;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return nbx_model_fail(nullptr, NBX_ERR_DEVICE, "no HIP device visible");
    nbx_model* m = new nbx_model();
    int rc = NBX_OK;
    auto fail = [&](int code) {
        g_model_create_error = m->err;
        nbx_model_destroy(m);
        return code;
    };
    if (hipGetDevice(&m->device) != hipSuccess || hipStreamCreate(&m->stream) != hipSuccess) {
        m->err = "stream creation failed";
        return fail(NBX_ERR_DEVICE);
    }
    if (hipMalloc((void**)&m->d_params, sizeof(float) * LR_NPARAM) != hipSuccess ||
        hipMalloc((void**)&m->d_m, sizeof(float) * LR_NPARAM) != hipSuccess ||
        hipMalloc((void**)&m->d_v, sizeof(float) * LR_NPARAM) != hipSuccess ||
        hipMalloc((void**)&m->d_frag, 2 * LR_NFRAG * LR_FRAG_ELEMS) != hipSuccess ||
        hipMalloc((void**)&m->d_norm, sizeof(float) * LR_NNORM) != hipSuccess) {
        m->err = "out of device memory";
        return fail(NBX_ERR_DEVICE);
    }
    // PyTorch-style initialisation: U(-1/sqrt(fan_in), 1/sqrt(fan_in)) for W and b
    std::vector<double> p(LR_NPARAM);
    uint64_t st = seed ^ 0x6E62786D6F64656Cull;
    const int offs[5] = {LR_W1, LR_W2, LR_W3, LR_W4, LR_NPARAM};
    const int fan[4] = {LR_NIN, LR_NH, LR_NH, LR_NH};
    for (int l = 0; l < 4; ++l) {
        const double bound = 1.0 / std::sqrt((double)fan[l]);
        for (int k = offs[l]; k < offs[l + 1]; ++k)
            p[k] = bound * (2.0 * ((double)(splitmix(st) >> 11) * 0x1.0p-53) - 1.0);
    }
    double norm[LR_NNORM];
    for (int k = 0; k < LR_NIN; ++k) { norm[k] = 0.0; norm[LR_NIN + k] = 1.0; }
    for (int k = 0; k < LR_NOUT; ++k) norm[2 * LR_NIN + k] = 1.0;
    if (hipMemsetAsync(m->d_frag, 0, 2 * LR_NFRAG * LR_FRAG_ELEMS, m->stream) != hipSuccess) {
        m->err = "hipMemsetAsync failed";
        return fail(NBX_ERR_DEVICE);
    }
    if ((rc = nbx_model_set(m, p.data(), norm))) return fail(rc);
    m->norm_set = false;
    *out = m;
    return NBX_OK;
}

extern "C" int nbx_model_destroy(nbx_model* m) {
    if (!m) return NBX_OK;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    free_train(m);
    void* ps[] = {m->d_params, m->d_m, m->d_v, m->d_frag, m->d_norm, m->d_X, m->d_T, m->d_hist};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return NBX_OK;
}

extern "C" const char* nbx_model_last_error(const nbx_model* m) {
    return m ? m->err.c_str() : g_model_create_error.c_str();
}

extern "C" int nbx_model_set(nbx_model* m, const double* params, const double* norm) {
    if (!m) return NBX_ERR_INVALID;
    if (params) {
        if (!finite_all(params, LR_NPARAM))
            return nbx_model_fail(m, NBX_ERR_NONFINITE, "non-finite parameter");
        std::vector<float> f(LR_NPARAM);
        for (int k = 0; k < LR_NPARAM; ++k) f[k] = (float)params[k];
        LR_HIP(m, hipMemcpyAsync(m->d_params, f.data(), sizeof(float) * LR_NPARAM,
                                 hipMemcpyHostToDevice, m->stream));
        lr_launch_write_frags(m->stream, m->d_params, m->d_frag);
        LR_HIP(m, hipGetLastError());
        int rc = reset_adam(m);
        if (rc) return rc;
    }
    if (norm) {
        if (!finite_all(norm, LR_NNORM))
            return nbx_model_fail(m, NBX_ERR_NONFINITE, "non-finite normalization");
        for (int k = 0; k < LR_NOUT; ++k)
            if (!(norm[2 * LR_NIN + k] > 0.0))
                return nbx_model_fail(m, NBX_ERR_INVALID, "output scales must be > 0");
        std::memcpy(m->h_norm, norm, sizeof(m->h_norm));
        m->norm_set = true;
        int rc = upload_norm(m);
        if (rc) return rc;
    }
    LR_HIP(m, hipStreamSynchronize(m->stream));
    return NBX_OK;
}

extern "C" int nbx_model_get(nbx_model* m, double* params, double* norm) {
    if (!m) return NBX_ERR_INVALID;
    if (params) {
        std::vector<float> f(LR_NPARAM);
        LR_HIP(m, hipMemcpyAsync(f.data(), m->d_params, sizeof(float) * LR_NPARAM,
                                 hipMemcpyDeviceToHost, m->stream));
        LR_HIP(m, hipStreamSynchronize(m->stream));
        for (int k = 0; k < LR_NPARAM; ++k) params[k] = f[k];
    }
    if (norm) std::memcpy(norm, m->h_norm, sizeof(m->h_norm));
    return NBX_OK;
}

extern "C" int nbx_model_forward(nbx_model* m, const double* feats, int64_t n, double* out) {
    if (!m || (!feats && n > 0) || (!out && n > 0) || n < 0) return NBX_ERR_INVALID;
    if (n == 0) return NBX_OK;
    if (n > (int64_t)1 << 24) return nbx_model_fail(m, NBX_ERR_INVALID, "at most 2^24 samples");
    const int64_t npad = (n + LR_CHUNK - 1) / LR_CHUNK * LR_CHUNK;
    int rc;
    if ((rc = alloc_train(m, (int)npad))) return rc;
    dev_rows r;
    if ((rc = upload_rows(m, feats, nullptr, n, npad, r))) return rc;
    lr_launch_train_chain(m->stream, bufs(m, r.X, nullptr, nullptr), 0, 0.0f);
    LR_HIP(m, hipGetLastError());
    std::vector<float> o((size_t)npad * LR_NOUT);
    LR_HIP(m, hipMemcpyAsync(o.data(), m->d_out, sizeof(float) * o.size(), hipMemcpyDeviceToHost,
                             m->stream));
    LR_HIP(m, hipStreamSynchronize(m->stream));
    for (int64_t k = 0; k < n * LR_NOUT; ++k) out[k] = o[k];
    return NBX_OK;
}

extern "C" int nbx_model_gradient(nbx_model* m, const double* feats, const double* targets,
                                  int64_t n, double* grad, double* loss) {
    if (!m || !feats || !targets || !grad || n <= 0) return NBX_ERR_INVALID;
    if (n % LR_CHUNK != 0 || n > (int64_t)1 << 24)
        return nbx_model_fail(m, NBX_ERR_INVALID, "sample count must be a multiple of 256, <= 2^24");
    int rc;
    if ((rc = alloc_train(m, (int)n))) return rc;
    dev_rows r;
    if ((rc = upload_rows(m, feats, targets, n, n, r))) return rc;
    const lr_train_bufs b = bufs(m, r.X, r.T, nullptr);
    lr_launch_train_chain(m->stream, b, 1, (float)(2.0 / (6.0 * (double)n)));
    lr_launch_dw(m->stream, b);
    float* d_grad = nullptr;
    LR_HIP(m, hipMalloc((void**)&d_grad, sizeof(float) * (LR_NPARAM + 1)));
    lr_launch_reduce(m->stream, (int)(n / LR_CHUNK), m->d_part, d_grad);
    lr_launch_loss_sum(m->stream, m->d_loss, (int)(n / 32), (float)(1.0 / (6.0 * (double)n)),
                       d_grad + LR_NPARAM);
    std::vector<float> g(LR_NPARAM + 1);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(g.data(), d_grad, sizeof(float) * g.size(), hipMemcpyDeviceToHost,
                           m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    (void)hipFree(d_grad);
    if (e != hipSuccess) return nbx_model_fail(m, NBX_ERR_DEVICE, hipGetErrorString(e));
    for (int k = 0; k < LR_NPARAM; ++k) grad[k] = g[k];
    if (loss) *loss = g[LR_NPARAM];
    return NBX_OK;
}

extern "C" int nbx_model_add_samples(nbx_model* m, const double* feats, const double* targets,
                                     int64_t n) {
    if (!m || n < 0 || (n > 0 && (!feats || !targets))) return NBX_ERR_INVALID;
    if (n == 0) return NBX_OK;
    if (!finite_all(feats, n * LR_NIN) || !finite_all(targets, n * LR_NOUT))
        return nbx_model_fail(m, NBX_ERR_NONFINITE, "non-finite sample");
    int rc;
    if ((rc = lr_model_reserve(m, m->count + n))) return rc;
    std::vector<float> hx((size_t)n * LR_NIN), ht((size_t)n * LR_NOUT);
    for (size_t k = 0; k < hx.size(); ++k) hx[k] = (float)feats[k];
    for (size_t k = 0; k < ht.size(); ++k) ht[k] = (float)targets[k];
    LR_HIP(m, hipMemcpyAsync(m->d_X + (size_t)m->count * LR_NIN, hx.data(), sizeof(float) * hx.size(),
                             hipMemcpyHostToDevice, m->stream));
    LR_HIP(m, hipMemcpyAsync(m->d_T + (size_t)m->count * LR_NOUT, ht.data(),
                             sizeof(float) * ht.size(), hipMemcpyHostToDevice, m->stream));
    LR_HIP(m, hipStreamSynchronize(m->stream));
    m->count += n;
    return NBX_OK;
}

extern "C" int nbx_model_get_samples(nbx_model* m, int64_t first, int64_t n, double* feats,
                                     double* targets, int64_t* count) {
    if (!m) return NBX_ERR_INVALID;
    if (count) *count = m->count;
    if (n == 0) return NBX_OK;
    if (first < 0 || n < 0 || first + n > m->count)
        return nbx_model_fail(m, NBX_ERR_BOUNDS, "sample range outside the training set");
    std::vector<float> hx((size_t)n * LR_NIN), ht((size_t)n * LR_NOUT);
    LR_HIP(m, hipMemcpyAsync(hx.data(), m->d_X + (size_t)first * LR_NIN, sizeof(float) * hx.size(),
                             hipMemcpyDeviceToHost, m->stream));
    LR_HIP(m, hipMemcpyAsync(ht.data(), m->d_T + (size_t)first * LR_NOUT,
                             sizeof(float) * ht.size(), hipMemcpyDeviceToHost, m->stream));
    LR_HIP(m, hipStreamSynchronize(m->stream));
    if (feats)
        for (size_t k = 0; k < hx.size(); ++k) feats[k] = hx[k];
    if (targets)
        for (size_t k = 0; k < ht.size(); ++k) targets[k] = ht[k];
    return NBX_OK;
}

extern "C" int nbx_model_clear_samples(nbx_model* m) {
    if (!m) return NBX_ERR_INVALID;
    m->count = 0;
    return NBX_OK;
}

// feature mean / inverse standard deviation and per-component RMS of the targets over
// (at most 2^20 evenly spaced rows of) the training set
extern "C" int nbx_model_fit_normalization(nbx_model* m) {
    if (!m) return NBX_ERR_INVALID;
    if (m->count < 2) return nbx_model_fail(m, NBX_ERR_INVALID, "training set has < 2 samples");
    const int64_t take = std::min<int64_t>(m->count, (int64_t)1 << 20);
    const int64_t step = m->count / take;
    std::vector<double> s1(LR_NIN, 0.0), s2(LR_NIN, 0.0), t2(LR_NOUT, 0.0);
    // strided rows copied in blocks
    const int64_t blk = 4096;
    std::vector<float> bx((size_t)blk * LR_NIN), bt((size_t)blk * LR_NOUT);
    int64_t used = 0;
    for (int64_t r0 = 0; r0 < take; r0 += blk) {
        const int64_t nb = std::min(blk, take - r0);
        if (step == 1) {
            LR_HIP(m, hipMemcpyAsync(bx.data(), m->d_X + (size_t)r0 * LR_NIN,
                                     sizeof(float) * nb * LR_NIN, hipMemcpyDeviceToHost, m->stream));
            LR_HIP(m, hipMemcpyAsync(bt.data(), m->d_T + (size_t)r0 * LR_NOUT,
                                     sizeof(float) * nb * LR_NOUT, hipMemcpyDeviceToHost, m->stream));
        } else {
            LR_HIP(m, hipMemcpy2DAsync(bx.data(), sizeof(float) * LR_NIN,
                                       m->d_X + (size_t)r0 * step * LR_NIN,
                                       sizeof(float) * LR_NIN * step, sizeof(float) * LR_NIN, nb,
                                       hipMemcpyDeviceToHost, m->stream));
            LR_HIP(m, hipMemcpy2DAsync(bt.data(), sizeof(float) * LR_NOUT,
                                       m->d_T + (size_t)r0 * step * LR_NOUT,
                                       sizeof(float) * LR_NOUT * step, sizeof(float) * LR_NOUT, nb,
                                       hipMemcpyDeviceToHost, m->stream));
        }
        LR_HIP(m, hipStreamSynchronize(m->stream));
        for (int64_t q = 0; q < nb; ++q) {
            for (int k = 0; k < LR_NIN; ++k) {
                const double x = bx[q * LR_NIN + k];
                s1[k] += x;
                s2[k] += x * x;
            }
            for (int k = 0; k < LR_NOUT; ++k) {
                const double t = bt[q * LR_NOUT + k];
                t2[k] += t * t;
            }
        }
        used += nb;
    }
    double norm[LR_NNORM];
    for (int k = 0; k < LR_NIN; ++k) {
        const double mean = s1[k] / used;
        const double var = std::max(0.0, s2[k] / used - mean * mean);
        norm[k] = mean;
        norm[LR_NIN + k] = 1.0 / std::max(std::sqrt(var), 1e-6);
    }
    for (int k = 0; k < LR_NOUT; ++k) norm[2 * LR_NIN + k] = std::max(std::sqrt(t2[k] / used), 1e-12);
    return nbx_model_set(m, nullptr, norm);
}

extern "C" int nbx_model_train(nbx_model* m, int64_t iters, int32_t batch, double lr,
                               uint64_t seed, double* losses) {
    if (!m || iters < 0) return NBX_ERR_INVALID;
    if (batch <= 0 || batch % LR_CHUNK != 0 || batch > (1 << 20))
        return nbx_model_fail(m, NBX_ERR_INVALID, "batch must be a positive multiple of 256, <= 2^20");
    if (!(lr > 0.0) || !std::isfinite(lr)) return nbx_model_fail(m, NBX_ERR_INVALID, "lr must be > 0");
    if (m->count < 1) return nbx_model_fail(m, NBX_ERR_INVALID, "empty training set");
    if (iters == 0) return NBX_OK;
    int rc;
    if ((rc = alloc_train(m, batch))) return rc;
    if (m->hist_cap < iters) {
        if (m->d_hist) (void)hipFree(m->d_hist);
        m->d_hist = nullptr;
        m->hist_cap = 0;
        LR_HIP(m, hipMalloc((void**)&m->d_hist, sizeof(float) * iters));
        m->hist_cap = iters;
    }
    hipStream_t s = m->stream;
    const lr_train_bufs b = bufs(m, m->d_X, m->d_T, m->d_idx);
    const float gscale = (float)(2.0 / (6.0 * (double)batch));
    const float lscale = (float)(1.0 / (6.0 * (double)batch));
    for (int64_t it = 0; it < iters; ++it) {
        lr_launch_gather_index(s, m->d_idx, batch, m->count, seed, (int64_t)m->adam_step);
        lr_launch_train_chain(s, b, 1, gscale);
        lr_launch_dw(s, b);
        lr_launch_loss_sum(s, m->d_loss, batch / 32, lscale, m->d_hist + it);
        ++m->adam_step;
        lr_launch_adam(s, batch / LR_CHUNK, m->d_part, m->d_params, m->d_m, m->d_v, m->d_frag,
                       (float)lr, 0.9f, 0.999f, 1e-8f, m->adam_step);
    }
    LR_HIP(m, hipGetLastError());
    std::vector<float> h((size_t)iters);
    LR_HIP(m, hipMemcpyAsync(h.data(), m->d_hist, sizeof(float) * iters, hipMemcpyDeviceToHost, s));
    LR_HIP(m, hipStreamSynchronize(s));
    if (losses)
        for (int64_t k = 0; k < iters; ++k) losses[k] = h[k];
    return NBX_OK;
}

extern "C" int nbx_model_evaluate(nbx_model* m, int64_t first, int64_t n, double* mse6,
                                  double* base6) {
    if (!m || !mse6 || n <= 0) return NBX_ERR_INVALID;
    if (first < 0 || first + n > m->count)
        return nbx_model_fail(m, NBX_ERR_BOUNDS, "sample range outside the training set");
    const int B = 1 << 16;
    int rc;
    if ((rc = alloc_train(m, B))) return rc;
    std::vector<double> acc(LR_NOUT, 0.0), base(LR_NOUT, 0.0);
    std::vector<float> o((size_t)B * LR_NOUT), t((size_t)B * LR_NOUT);
    for (int64_t r0 = 0; r0 < n; r0 += B) {
        const int64_t nb = std::min<int64_t>(B, n - r0);
        const int64_t row = first + r0;
        // rows beyond the set are never read: pad the batch by repeating row `row`
        std::vector<int> idx((size_t)B);
        for (int64_t q = 0; q < B; ++q) idx[q] = (int)(row + (q < nb ? q : 0));
        LR_HIP(m, hipMemcpyAsync(m->d_idx, idx.data(), sizeof(int) * B, hipMemcpyHostToDevice, m->stream));
        lr_launch_train_chain(m->stream, bufs(m, m->d_X, m->d_T, m->d_idx), 0, 0.0f);
        LR_HIP(m, hipGetLastError());
        LR_HIP(m, hipMemcpyAsync(o.data(), m->d_out, sizeof(float) * nb * LR_NOUT,
                                 hipMemcpyDeviceToHost, m->stream));
        LR_HIP(m, hipMemcpyAsync(t.data(), m->d_T + (size_t)row * LR_NOUT, sizeof(float) * nb * LR_NOUT,
                                 hipMemcpyDeviceToHost, m->stream));
        LR_HIP(m, hipStreamSynchronize(m->stream));
        for (int64_t q = 0; q < nb; ++q)
            for (int k = 0; k < LR_NOUT; ++k) {
                const double tn = t[q * LR_NOUT + k] / m->h_norm[2 * LR_NIN + k];
                const double e = o[q * LR_NOUT + k] - tn;
                acc[k] += e * e;
                base[k] += tn * tn;
            }
    }
    for (int k = 0; k < LR_NOUT; ++k) {
        mse6[k] = acc[k] / n;
        if (base6) base6[k] = base[k] / n;
    }
    return NBX_OK;
}
