// nbx_batch.cpp -- C ABI of the batched-replica engine (include/nbx.h, nbx_batch_*):
// replica table, device buffers, and the launch sequence: an optional prologue launch
// (list build + forces of the initial state), then blocks of steps_per_launch steps, each
// one persistent launch plus a step-base advance, captured once into a hipGraph and
// replayed.  A run that overflows a neighbour list is replayed from a device backup
// with larger lists (the counter-based RNG makes the replay identical).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <numeric>
#include <string>
#include <vector>

#include "nbx_internal.h"

// nbx_batch.hip
size_t batch_rep_size();
size_t batch_params_size();
int batch_nmax();
void batch_fill_rep(void* dst, int64_t off, int64_t list_off, int32_t n, int32_t therm, double L,
                    double T, double param, double dt, double m0, uint64_t key, double sigma,
                    double R, double skin, double kb);
void batch_fill_params(void* dst, double eps, double sigma, double R, double rlist, double skin,
                       double dt, double kb, int cap, int64_t ntot);
void batch_launch_steps_raw(hipStream_t s, int nblocks, const void* reps, const int* order,
                            double* x, double* v, double* a, double* vh, double* xref, uint16_t* nbr,
                            int* nnb, const double* ms, int* slot_of, const void* params,
                            const int64_t* step_base, int nsteps, int prologue, int* flags);
void batch_launch_energies(hipStream_t s, int nrep, const void* reps, const double* x,
                           const double* v, const double* ms, int64_t nt, double s2, double R2,
                           double eps, double* ke, double* pe, long long* pairs);
// nbx_nonbonded.hip
void mx_launch_step_advance(hipStream_t s, int64_t* step_base, int64_t delta, int set);

struct nbx_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    int nrep = 0;
    int64_t ntot = 0;
    std::vector<int64_t> n, off, list_off;
    std::vector<double> L, T, param, m0;
    std::vector<int32_t> therm;
    std::vector<int> order;
    double eps = 1, sigma = 1, R = 2.5, kb = 1, skin = 0.3;
    int cap = 0;
    double* d_x = nullptr;
    double* d_v = nullptr;
    double* d_a = nullptr;
    double* d_vh = nullptr;
    double* d_xref = nullptr;
    double* d_ms = nullptr;
    double* d_tmp = nullptr;
    uint16_t* d_nbr = nullptr;  // 16-bit slots in groups of 8 per atom
    int* d_nnb = nullptr;
    int* d_slot = nullptr;    // LDS slot of each atom (cell order) since the last list build
    int* d_flags = nullptr;   // [0] overflow, [1] list builds
    int* d_order = nullptr;
    int64_t* d_step = nullptr;
    void* d_reps = nullptr;
    void* h_params = nullptr;
    bool state_set = false, accel_valid = false;
    int64_t steps_taken = 0;
    int64_t version = 0;
    hipGraphExec_t gexec = nullptr;
    double g_dt = 0;
    uint64_t g_seed = 0;
    int64_t g_steps = 0, g_version = -1;
    std::string err;
};

static int bfail(nbx_batch* b, int code, const std::string& msg) {
    if (b) b->err = msg;
    else nbx_fail(nullptr, code, msg);
    return code;
}

#define BHIP(b, call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            return bfail((b), NBX_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

static void bfree(void* p) { if (p) (void)hipFree(p); }

static void batch_release(nbx_batch* b) {
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
    bfree(b->d_x); bfree(b->d_v); bfree(b->d_a); bfree(b->d_vh); bfree(b->d_xref);
    bfree(b->d_ms); bfree(b->d_tmp); bfree(b->d_nbr); bfree(b->d_nnb); bfree(b->d_flags);
    bfree(b->d_slot);
    bfree(b->d_order); bfree(b->d_step); bfree(b->d_reps);
    if (b->h_params) free(b->h_params);
    b->h_params = nullptr;
    if (b->stream) (void)hipStreamDestroy(b->stream);
    b->stream = nullptr;
}

static double rlist_of(const nbx_batch* b) { return b->R + b->skin; }

static int alloc_lists(nbx_batch* b, double grow) {
    // a captured step graph holds the list pointer: it must not outlive the list
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
    ++b->version;
    bfree(b->d_nbr);
    b->d_nbr = nullptr;
    double rho = 0.0;
    int64_t nmax = 1;
    for (int r = 0; r < b->nrep; ++r) {
        rho = std::max(rho, (double)b->n[r] / (b->L[r] * b->L[r] * b->L[r]));
        nmax = std::max(nmax, b->n[r]);
    }
    const double rl = rlist_of(b);
    const double est = 4.0 / 3.0 * M_PI * rl * rl * rl * rho;
    // a multiple of 8 entries (whole 16-byte groups), so every replica's list starts on
    // a 16-byte boundary
    const int cap = (int)std::min<double>((double)nmax, std::ceil(grow * (1.6 * est + 24.0)));
    b->cap = (cap + 7) / 8 * 8;
    b->list_off.assign(b->nrep, 0);
    int64_t total = 0;
    for (int r = 0; r < b->nrep; ++r) {
        b->list_off[r] = total;
        total += (int64_t)b->cap * b->n[r];
    }
    BHIP(b, hipMalloc((void**)&b->d_nbr, sizeof(uint16_t) * (size_t)std::max<int64_t>(total, 8)));
    return NBX_OK;
}

extern "C" int nbx_batch_create(nbx_batch** out, int32_t nrep, const int64_t* natoms,
                                const double* box_L, const double* masses, double eps,
                                double sigma, double R, double kb) {
    if (!out) return nbx_fail(nullptr, NBX_ERR_INVALID, "nbx_batch_create: out is NULL");
    *out = nullptr;
    if (nrep <= 0 || !natoms || !box_L || !masses)
        return nbx_fail(nullptr, NBX_ERR_INVALID, "nbx_batch_create: need nrep > 0 and arrays");
    if (!(sigma > 0) || !(R > 0) || !std::isfinite(R) || !std::isfinite(eps) || !(kb > 0))
        return nbx_fail(nullptr, NBX_ERR_INVALID, "nbx_batch_create: bad LJ parameters or kb");
    int64_t ntot = 0;
    for (int r = 0; r < nrep; ++r) {
        if (natoms[r] < 1 || natoms[r] > batch_nmax())
            return nbx_fail(nullptr, NBX_ERR_INVALID,
                            "nbx_batch_create: replica sizes must be 1.." + std::to_string(batch_nmax()));
        if (!(box_L[r] > 0) || !std::isfinite(box_L[r]))
            return nbx_fail(nullptr, NBX_ERR_INVALID, "nbx_batch_create: box side must be > 0");
        ntot += natoms[r];
    }
    for (int64_t g = 0; g < ntot; ++g)
        if (!(masses[g] > 0) || !std::isfinite(masses[g]))
            return nbx_fail(nullptr, NBX_ERR_INVALID, "nbx_batch_create: masses must be > 0");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return nbx_fail(nullptr, NBX_ERR_DEVICE, "nbx_batch_create: no HIP device visible");
    nbx_batch* b = new nbx_batch();
    b->nrep = nrep;
    b->ntot = ntot;
    b->n.assign(natoms, natoms + nrep);
    b->L.assign(box_L, box_L + nrep);
    b->off.resize(nrep);
    b->m0.resize(nrep);
    int64_t o = 0;
    for (int r = 0; r < nrep; ++r) {
        b->off[r] = o;
        b->m0[r] = masses[o];
        o += natoms[r];
    }
    b->therm.assign(nrep, NBX_THERM_NONE);
    b->T.assign(nrep, 0.0);
    b->param.assign(nrep, 1.0);
    b->eps = eps; b->sigma = sigma; b->R = R; b->kb = kb;
    b->order.resize(nrep);
    std::iota(b->order.begin(), b->order.end(), 0);
    std::stable_sort(b->order.begin(), b->order.end(),
                     [&](int p, int q) { return b->n[p] > b->n[q]; });
    const size_t b3 = sizeof(double) * 3 * (size_t)ntot;
    hipError_t e = hipGetDevice(&b->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    double** bufs[6] = {&b->d_x, &b->d_v, &b->d_a, &b->d_vh, &b->d_xref, &b->d_tmp};
    for (double** p : bufs)
        if (e == hipSuccess) e = hipMalloc((void**)p, b3);
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_ms, sizeof(double) * (size_t)ntot);
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_nnb, sizeof(int) * (size_t)ntot);
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_slot, sizeof(int) * (size_t)ntot);
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_flags, sizeof(int) * 8);
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_order, sizeof(int) * (size_t)nrep);
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_step, sizeof(int64_t));
    if (e == hipSuccess) e = hipMalloc(&b->d_reps, batch_rep_size() * (size_t)nrep);
    if (e == hipSuccess)
        e = hipMemcpy(b->d_ms, masses, sizeof(double) * (size_t)ntot, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(b->d_order, b->order.data(), sizeof(int) * (size_t)nrep, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(b->d_flags, 0, sizeof(int) * 8);
    if (e == hipSuccess) e = hipMemset(b->d_nnb, 0, sizeof(int) * (size_t)ntot);
    b->h_params = calloc(1, batch_params_size());
    if (e != hipSuccess) {
        std::string msg = std::string("nbx_batch_create: ") + hipGetErrorString(e);
        batch_release(b);
        delete b;
        return nbx_fail(nullptr, NBX_ERR_DEVICE, msg);
    }
    if (int rc = alloc_lists(b, 1.0)) {
        std::string msg = b->err;
        batch_release(b);
        delete b;
        return nbx_fail(nullptr, rc, msg);
    }
    *out = b;
    return NBX_OK;
}

extern "C" int nbx_batch_destroy(nbx_batch* b) {
    if (!b) return NBX_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_release(b);
    delete b;
    return NBX_OK;
}

extern "C" const char* nbx_batch_last_error(const nbx_batch* b) {
    return b ? b->err.c_str() : nbx_last_error(nullptr);
}

extern "C" int nbx_batch_set_thermostat(nbx_batch* b, int32_t r, int32_t kind, double T,
                                        double param) {
    if (!b) return NBX_ERR_INVALID;
    if (r < 0 || r >= b->nrep) return bfail(b, NBX_ERR_BOUNDS, "replica index out of range");
    if (kind != NBX_THERM_NONE && kind != NBX_THERM_ANDERSEN && kind != NBX_THERM_LANGEVIN)
        return bfail(b, NBX_ERR_UNSUPPORTED, "batched replicas take Andersen or Langevin");
    if (kind != NBX_THERM_NONE && (!(T >= 0) || !(param > 0) || !std::isfinite(T) ||
                                   !std::isfinite(param)))
        return bfail(b, NBX_ERR_INVALID, "thermostat needs T >= 0 and a parameter > 0");
    b->therm[r] = kind;
    b->T[r] = T;
    b->param[r] = param;
    ++b->version;
    return NBX_OK;
}

extern "C" int nbx_batch_set_skin(nbx_batch* b, double skin) {
    if (!b) return NBX_ERR_INVALID;
    if (!(skin >= 0) || !std::isfinite(skin)) return bfail(b, NBX_ERR_INVALID, "skin must be >= 0");
    b->skin = skin;
    b->accel_valid = false;
    ++b->version;
    return alloc_lists(b, 1.0);
}

static int up3(nbx_batch* b, const double* host, double* d) {
    const size_t bytes = sizeof(double) * 3 * (size_t)b->ntot;
    BHIP(b, hipMemcpyAsync(b->d_tmp, host, bytes, hipMemcpyHostToDevice, b->stream));
    ref64_aos_to_soa(b->stream, b->d_tmp, d, b->ntot);
    return NBX_OK;
}

static int down3(const nbx_batch* bc, const double* d, double* host) {
    nbx_batch* b = const_cast<nbx_batch*>(bc);
    const size_t bytes = sizeof(double) * 3 * (size_t)b->ntot;
    ref64_soa_to_aos(b->stream, d, b->d_tmp, b->ntot);
    BHIP(b, hipMemcpyAsync(host, b->d_tmp, bytes, hipMemcpyDeviceToHost, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    return NBX_OK;
}

extern "C" int nbx_batch_set_state(nbx_batch* b, const double* u, const double* v) {
    if (!b) return NBX_ERR_INVALID;
    if (!u) return bfail(b, NBX_ERR_INVALID, "positions are NULL");
    for (int64_t k = 0; k < 3 * b->ntot; ++k)
        if (!std::isfinite(u[k]) || (v && !std::isfinite(v[k])))
            return bfail(b, NBX_ERR_NONFINITE, "state holds NaN or Inf");
    int rc = up3(b, u, b->d_x);
    if (rc) return rc;
    if (v) { if ((rc = up3(b, v, b->d_v))) return rc; }
    else ref64_fill(b->stream, b->d_v, 3 * b->ntot, 0.0);
    BHIP(b, hipStreamSynchronize(b->stream));
    b->state_set = true;
    b->accel_valid = false;
    b->steps_taken = 0;
    return NBX_OK;
}

extern "C" int nbx_batch_get_state(const nbx_batch* b, double* u, double* v) {
    if (!b) return NBX_ERR_INVALID;
    if (!b->state_set) return bfail(const_cast<nbx_batch*>(b), NBX_ERR_INVALID, "no state set");
    if (u) { int rc = down3(b, b->d_x, u); if (rc) return rc; }
    if (v) { int rc = down3(b, b->d_v, v); if (rc) return rc; }
    return NBX_OK;
}

static int upload_tables(nbx_batch* b, double dt, uint64_t seed) {
    std::vector<char> reps(batch_rep_size() * (size_t)b->nrep);
    for (int r = 0; r < b->nrep; ++r) {
        const uint64_t key = seed + (uint64_t)r * 0x9E3779B97F4A7C15ull;
        batch_fill_rep(reps.data() + batch_rep_size() * r, b->off[r], b->list_off[r],
                       (int32_t)b->n[r], b->therm[r], b->L[r], b->T[r], b->param[r], dt,
                       b->m0[r], key, b->sigma, b->R, b->skin, b->kb);
    }
    BHIP(b, hipMemcpyAsync(b->d_reps, reps.data(), reps.size(), hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    batch_fill_params(b->h_params, b->eps, b->sigma, b->R, rlist_of(b), b->skin, dt, b->kb, b->cap,
                      b->ntot);
    return NBX_OK;
}

static void launch(nbx_batch* b, int nsteps, int prologue) {
    batch_launch_steps_raw(b->stream, b->nrep, b->d_reps, b->d_order, b->d_x, b->d_v, b->d_a,
                           b->d_vh, b->d_xref, b->d_nbr, b->d_nnb, b->d_ms, b->d_slot,
                           b->h_params, b->d_step, nsteps, prologue, b->d_flags);
}

extern "C" int nbx_batch_run(nbx_batch* b, double dt, int64_t nsteps, uint64_t seed,
                             int64_t steps_per_launch) {
    if (!b) return NBX_ERR_INVALID;
    if (!b->state_set) return bfail(b, NBX_ERR_INVALID, "no state set");
    if (!(dt > 0) || !std::isfinite(dt)) return bfail(b, NBX_ERR_INVALID, "dt must be > 0");
    if (nsteps < 0) return bfail(b, NBX_ERR_INVALID, "nsteps < 0");
    const int64_t K = steps_per_launch > 0 ? steps_per_launch : 1000;
    if (K > (1 << 30)) return bfail(b, NBX_ERR_INVALID, "steps_per_launch too large");
    hipStream_t s = b->stream;
    const size_t b3 = sizeof(double) * 3 * (size_t)b->ntot;
    double* bk = nullptr;
    BHIP(b, hipMalloc((void**)&bk, 3 * b3));
    struct guard { double* p; ~guard() { if (p) (void)hipFree(p); } } g{bk};
    double grow = 1.0;
    const int64_t step0 = b->steps_taken;
    const bool had_accel = b->accel_valid;
    BHIP(b, hipMemcpyAsync(bk, b->d_x, b3, hipMemcpyDeviceToDevice, s));
    BHIP(b, hipMemcpyAsync((char*)bk + b3, b->d_v, b3, hipMemcpyDeviceToDevice, s));
    BHIP(b, hipMemcpyAsync((char*)bk + 2 * b3, b->d_a, b3, hipMemcpyDeviceToDevice, s));
    for (int attempt = 0;; ++attempt) {
        int rc = upload_tables(b, dt, seed);
        if (rc) return rc;
        mx_launch_step_advance(s, b->d_step, step0, 1);
        if (!b->accel_valid) launch(b, 0, 1);
        const int64_t full = nsteps / K, rem = nsteps % K;
        if (full > 0) {
            const bool same = b->gexec && b->g_dt == dt && b->g_seed == seed && b->g_steps == K &&
                              b->g_version == b->version;
            if (!same) {
                if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
                b->gexec = nullptr;
                hipGraph_t gr = nullptr;
                BHIP(b, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                launch(b, (int)K, 0);
                mx_launch_step_advance(s, b->d_step, K, 0);
                BHIP(b, hipStreamEndCapture(s, &gr));
                hipError_t e = hipGraphInstantiate(&b->gexec, gr, nullptr, nullptr, 0);
                (void)hipGraphDestroy(gr);
                if (e != hipSuccess) {
                    b->gexec = nullptr;
                    return bfail(b, NBX_ERR_DEVICE, std::string("hipGraphInstantiate: ") +
                                                        hipGetErrorString(e));
                }
                b->g_dt = dt; b->g_seed = seed; b->g_steps = K; b->g_version = b->version;
            }
            for (int64_t r = 0; r < full; ++r) BHIP(b, hipGraphLaunch(b->gexec, s));
        }
        if (rem > 0) {
            launch(b, (int)rem, 0);
            mx_launch_step_advance(s, b->d_step, rem, 0);
        }
        BHIP(b, hipGetLastError());
        int h[8];
        BHIP(b, hipMemcpyAsync(h, b->d_flags, sizeof(h), hipMemcpyDeviceToHost, s));
        BHIP(b, hipStreamSynchronize(s));
        if (!h[0]) break;
        // a list overflowed: restore, grow the lists and replay the whole call
        if (attempt >= 6 || grow > 40)
            return bfail(b, NBX_ERR_CAPACITY, "batched neighbour list overflow persists");
        grow *= 1.6;
        BHIP(b, hipMemcpyAsync(b->d_x, bk, b3, hipMemcpyDeviceToDevice, s));
        BHIP(b, hipMemcpyAsync(b->d_v, (char*)bk + b3, b3, hipMemcpyDeviceToDevice, s));
        BHIP(b, hipMemcpyAsync(b->d_a, (char*)bk + 2 * b3, b3, hipMemcpyDeviceToDevice, s));
        BHIP(b, hipMemsetAsync(b->d_flags, 0, sizeof(int) * 8, s));
        if ((rc = alloc_lists(b, grow))) return rc;
        ++b->version;
        b->accel_valid = false;  // rebuild the (new) lists before the first step
        (void)had_accel;
    }
    b->accel_valid = true;
    b->steps_taken = step0 + nsteps;
    return NBX_OK;
}

// Verlet skin by timing: for every candidate (in units of sigma) `trial_steps` steps of
// length dt from the current state, which is restored after each trial (positions,
// velocities, accelerations, step count).  Keeps the fastest skin.
extern "C" int nbx_batch_autotune(nbx_batch* b, double dt, int64_t trial_steps, double* out) {
    if (!b) return NBX_ERR_INVALID;
    if (!b->state_set) return bfail(b, NBX_ERR_INVALID, "no state set");
    if (!(dt > 0) || !std::isfinite(dt) || trial_steps <= 0)
        return bfail(b, NBX_ERR_INVALID, "dt > 0 and trial_steps > 0 required");
    static const double cand[] = {0.2, 0.3, 0.4, 0.5, 0.6, 0.8};
    hipStream_t s = b->stream;
    const size_t b3 = sizeof(double) * 3 * (size_t)b->ntot;
    double* bk = nullptr;
    BHIP(b, hipMalloc((void**)&bk, 3 * b3));
    struct guard { double* p; ~guard() { if (p) (void)hipFree(p); } } g{bk};
    BHIP(b, hipMemcpyAsync(bk, b->d_x, b3, hipMemcpyDeviceToDevice, s));
    BHIP(b, hipMemcpyAsync((char*)bk + b3, b->d_v, b3, hipMemcpyDeviceToDevice, s));
    BHIP(b, hipMemcpyAsync((char*)bk + 2 * b3, b->d_a, b3, hipMemcpyDeviceToDevice, s));
    const double skin0 = b->skin;
    const int64_t steps0 = b->steps_taken;
    const bool accel0 = b->accel_valid;
    double best_ms = 0, best_skin = skin0, entry_ms = -1;
    auto restore = [&]() -> int {
        BHIP(b, hipMemcpyAsync(b->d_x, bk, b3, hipMemcpyDeviceToDevice, s));
        BHIP(b, hipMemcpyAsync(b->d_v, (char*)bk + b3, b3, hipMemcpyDeviceToDevice, s));
        BHIP(b, hipMemcpyAsync(b->d_a, (char*)bk + 2 * b3, b3, hipMemcpyDeviceToDevice, s));
        BHIP(b, hipStreamSynchronize(s));
        b->steps_taken = steps0;
        return NBX_OK;
    };
    for (int k = -1; k < (int)(sizeof(cand) / sizeof(cand[0])); ++k) {
        const double skin = k < 0 ? skin0 : cand[k] * b->sigma;
        int rc = nbx_batch_set_skin(b, skin);
        if (rc) return rc;
        if ((rc = nbx_batch_run(b, dt, 20, 0, 0))) return rc;           // warm: lists, graph
        auto t0 = std::chrono::steady_clock::now();
        if ((rc = nbx_batch_run(b, dt, trial_steps, 0, 0))) return rc;  // synchronous
        const double ms = std::chrono::duration<double, std::milli>(
                              std::chrono::steady_clock::now() - t0).count() / trial_steps;
        if ((rc = restore())) return rc;
        if (k < 0) { entry_ms = ms; best_ms = ms; best_skin = skin0; }
        else if (ms < best_ms) { best_ms = ms; best_skin = skin; }
    }
    int rc = nbx_batch_set_skin(b, best_skin);
    if (rc) return rc;
    b->accel_valid = false;   // new lists: the first run rebuilds them (a is still a(x))
    (void)accel0;
    if (out) { out[0] = best_skin; out[1] = best_ms; out[2] = entry_ms; }
    return NBX_OK;
}

extern "C" int nbx_batch_energies(nbx_batch* b, double* ke, double* pe, int64_t* pairs) {
    if (!b) return NBX_ERR_INVALID;
    if (!b->state_set) return bfail(b, NBX_ERR_INVALID, "no state set");
    std::vector<char> reps(batch_rep_size() * (size_t)b->nrep);
    for (int r = 0; r < b->nrep; ++r)
        batch_fill_rep(reps.data() + batch_rep_size() * r, b->off[r], b->list_off[r],
                       (int32_t)b->n[r], b->therm[r], b->L[r], b->T[r], b->param[r], 1.0,
                       b->m0[r], 0, b->sigma, b->R, b->skin, b->kb);
    void* d_reps = nullptr;
    double* d_out = nullptr;
    BHIP(b, hipMalloc(&d_reps, reps.size()));
    hipError_t e = hipMalloc((void**)&d_out, sizeof(double) * 3 * (size_t)b->nrep);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_reps, reps.data(), reps.size(), hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) {
        batch_launch_energies(b->stream, b->nrep, d_reps, b->d_x, b->d_v, b->d_ms, b->ntot,
                              b->sigma * b->sigma, b->R * b->R, b->eps, d_out, d_out + b->nrep,
                              reinterpret_cast<long long*>(d_out + 2 * b->nrep));
        e = hipGetLastError();
    }
    std::vector<double> h(3 * (size_t)b->nrep);
    if (e == hipSuccess)
        e = hipMemcpyAsync(h.data(), d_out, sizeof(double) * h.size(), hipMemcpyDeviceToHost,
                           b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    bfree(d_reps);
    bfree(d_out);
    if (e != hipSuccess) return bfail(b, NBX_ERR_DEVICE, hipGetErrorString(e));
    for (int r = 0; r < b->nrep; ++r) {
        if (ke) ke[r] = h[r];
        if (pe) pe[r] = h[b->nrep + r];
        if (pairs) pairs[r] = reinterpret_cast<const long long*>(h.data() + 2 * b->nrep)[r];
    }
    return NBX_OK;
}

extern "C" int nbx_batch_stats(nbx_batch* b, int64_t* out, int32_t nout) {
    if (!b || !out || nout < 0 || nout > 3) return NBX_ERR_INVALID;
    int h[8] = {0};
    BHIP(b, hipMemcpyAsync(h, b->d_flags, sizeof(h), hipMemcpyDeviceToHost, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    const int64_t v[3] = {h[1], b->steps_taken, b->cap};
    for (int k = 0; k < nout; ++k) out[k] = v[k];
    return NBX_OK;
}
