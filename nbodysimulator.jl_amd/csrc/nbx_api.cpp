// nbx_api.cpp -- C ABI of libnbx (declared in include/nbx.h): context management,
// argument validation (the reference's BoundsError / MethodError become error codes),
// host <-> device transfers (Julia column-major 3 x n <-> device SoA) and the FP64
// reference-order orchestration of soode_system! (src/nbody_to_ode.jl:474-532) and of
// the velocity-Verlet step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "nbx_internal.h"

static thread_local std::string g_create_error;

int nbx_fail(nbx_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    else g_create_error = msg;
    return code;
}

extern "C" const char* nbx_version(void) { return "nbx 0.1 (gfx950)"; }

extern "C" int nbx_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

extern "C" const char* nbx_last_error(const nbx_ctx* c) {
    return c ? c->err.c_str() : g_create_error.c_str();
}

static void free_ptr(void* p) {
    if (p) (void)hipFree(p);
}

static void release(nbx_ctx* c) {
    mx_release(c);
    free_ptr(c->d_x); free_ptr(c->d_v); free_ptr(c->d_a); free_ptr(c->d_vh);
    free_ptr(c->d_tmp); free_ptr(c->d_ms); free_ptr(c->d_qs); free_ptr(c->d_red);
    free_ptr(c->d_scal);
    free_ptr(c->d_mm);
    free_ptr(c->d_mm4);
    if (c->d_perm) (void)hipFree(c->d_perm);
    if (c->d_iperm) (void)hipFree(c->d_iperm);
    if (c->d_pmap) (void)hipFree(c->d_pmap);
    if (c->d_sortbuf) (void)hipFree(c->d_sortbuf);
    if (c->d_permbuf) (void)hipFree(c->d_permbuf);
    c->d_perm = c->d_iperm = c->d_pmap = nullptr;
    c->d_sortbuf = nullptr;
    c->d_permbuf = nullptr;
    c->sortbuf_bytes = 0;
    c->d_mm = nullptr;
    c->d_mm4 = nullptr;
    c->d_x = c->d_v = c->d_a = c->d_vh = c->d_tmp = c->d_ms = c->d_qs = nullptr;
    c->d_red = c->d_scal = nullptr;
    if (c->stream) (void)hipStreamDestroy(c->stream);
    c->stream = nullptr;
}

extern "C" int nbx_create(nbx_ctx** out, int64_t n, int32_t system_kind, const double* masses,
                          const double* charges) {
    if (!out) return nbx_fail(nullptr, NBX_ERR_INVALID, "nbx_create: out is NULL");
    *out = nullptr;
    if (n < 0) return nbx_fail(nullptr, NBX_ERR_INVALID, "nbx_create: n < 0");
    if (system_kind != NBX_SYS_POTENTIAL && system_kind != NBX_SYS_WATER)
        return nbx_fail(nullptr, NBX_ERR_INVALID, "nbx_create: unknown system kind");
    if (system_kind == NBX_SYS_WATER && n % 3 != 0)
        return nbx_fail(nullptr, NBX_ERR_INVALID,
                        "nbx_create: a water system has 3 atoms (O, H1, H2) per molecule");
    if (n > 0 && !masses) return nbx_fail(nullptr, NBX_ERR_INVALID, "nbx_create: masses NULL");
    if (n > (int64_t)1 << 30)
        return nbx_fail(nullptr, NBX_ERR_INVALID, "nbx_create: n above 2^30 atoms");
    for (int64_t i = 0; i < n; ++i)
        if (!(masses[i] > 0.0) || !std::isfinite(masses[i]))
            return nbx_fail(nullptr, NBX_ERR_INVALID,
                            "nbx_create: mass " + std::to_string(i) + " not positive/finite");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return nbx_fail(nullptr, NBX_ERR_DEVICE, "nbx_create: no HIP device visible");
    nbx_ctx* c = new nbx_ctx();
    c->n = n;
    c->sys = system_kind;
    c->h_ms.assign(masses, masses + n);
    c->equal_masses = n > 0 && std::all_of(masses, masses + n,
                                           [&](double m) { return m == masses[0]; });
    if (n > 0) c->mass_wscale = 1.0 / *std::max_element(masses, masses + n);
    if (charges) {
        c->h_qs.assign(charges, charges + n);
        c->has_q = true;
    }
    c->bc.kind = NBX_BC_INFINITE;
    for (int k = 0; k < 6; ++k) c->bc.b[k] = (k % 2 == 0) ? -INFINITY : INFINITY;
    if (const char* g = std::getenv("NBX_GRAPH_STEPS")) c->graph_steps = std::max(0, std::atoi(g));
    if (const char* g = std::getenv("NBX_BUILD_EVERY")) {
        const int m = std::atoi(g);
        if (m == 1 || m == 2 || m == 5) c->build_every = m;
    }
    if (const char* g = std::getenv("NBX_LANES")) {
        const int l = std::atoi(g);
        if (l == 1 || l == 2 || l == 4 || l == 8 || l == 16 || l == 32) c->nl.lanes = l;
    }
    // spatial re-sort of the mixed path's state every 10,000 steps from 8,192 atoms on
    // (NBX_SORT_EVERY: another interval for any size, 0 to never sort)
    c->sort_every = n >= 8192 ? 10000 : 0;
    if (const char* g = std::getenv("NBX_SORT_EVERY")) c->sort_every = std::max(0, std::atoi(g));
    hipError_t e = hipGetDevice(&c->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    // set-up of nbx_rdf_noalloc: the device address of this GPU's histogram array
    if (e == hipSuccess && ref64_rdf_scratch(c->device, &c->d_rdf_hist) != 0)
        e = hipErrorInvalidSymbol;
    const size_t b3 = sizeof(double) * (size_t)(3 * (n > 0 ? n : 1));
    const size_t b1 = sizeof(double) * (size_t)(n > 0 ? n : 1);
    if (e == hipSuccess) e = hipMalloc((void**)&c->d_x, b3);
    if (e == hipSuccess) e = hipMalloc((void**)&c->d_v, b3);
    if (e == hipSuccess) e = hipMalloc((void**)&c->d_a, b3);
    if (e == hipSuccess) e = hipMalloc((void**)&c->d_vh, b3);
    if (e == hipSuccess) e = hipMalloc((void**)&c->d_tmp, b3);
    if (e == hipSuccess) e = hipMalloc((void**)&c->d_ms, b1);
    if (e == hipSuccess) e = hipMalloc((void**)&c->d_qs, b1);
    if (e == hipSuccess) e = hipMalloc((void**)&c->d_red, b1 + 8192 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&c->d_scal, NBX_SCAL_COUNT * sizeof(double));
    const size_t bi = sizeof(int) * (size_t)(n > 0 ? n : 1);
    if (e == hipSuccess) e = hipMalloc((void**)&c->d_perm, bi);
    if (e == hipSuccess) e = hipMalloc((void**)&c->d_iperm, bi);
    if (e == hipSuccess) e = hipMalloc((void**)&c->d_pmap, bi);
    if (e == hipSuccess && n > 0) {
        std::vector<int> iota((size_t)n);
        for (int64_t k = 0; k < n; ++k) iota[(size_t)k] = (int)k;
        e = hipMemcpy(c->d_perm, iota.data(), bi, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(c->d_iperm, iota.data(), bi, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess && n > 0)
        e = hipMemcpy(c->d_ms, masses, b1, hipMemcpyHostToDevice);
    if (e == hipSuccess && n > 0 && charges)
        e = hipMemcpy(c->d_qs, charges, b1, hipMemcpyHostToDevice);
    if (e == hipSuccess && n > 0 && !charges) e = hipMemset(c->d_qs, 0, b1);
    if (e != hipSuccess) {
        std::string msg = std::string("nbx_create: ") + hipGetErrorString(e);
        release(c);
        delete c;
        return nbx_fail(nullptr, NBX_ERR_DEVICE, msg);
    }
    *out = c;
    return NBX_OK;
}

extern "C" int nbx_destroy(nbx_ctx* c) {
    if (!c) return NBX_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    release(c);
    delete c;
    return NBX_OK;
}

static void invalidate(nbx_ctx* c) {
    c->accel_valid = false;
    c->nl.ready = false;
    c->list_sorted = false;
    ++c->version;
}

extern "C" int nbx_set_boundary(nbx_ctx* c, int32_t kind, const double* box) {
    if (!c) return NBX_ERR_INVALID;
    nbx_bcv bc{};
    bc.kind = kind;
    if (kind == NBX_BC_INFINITE) {
        for (int k = 0; k < 6; ++k) bc.b[k] = (k % 2 == 0) ? -INFINITY : INFINITY;
    } else if (kind == NBX_BC_CUBIC) {
        if (!box || !(box[0] > 0.0) || !std::isfinite(box[0]))
            return nbx_fail(c, NBX_ERR_INVALID, "cubic box side L must be positive and finite");
        bc.b[0] = box[0];
    } else if (kind == NBX_BC_PERIODIC) {
        if (!box) return nbx_fail(c, NBX_ERR_INVALID, "periodic bounds missing");
        for (int k = 0; k < 3; ++k)
            if (!(box[2 * k + 1] > box[2 * k]) || !std::isfinite(box[2 * k]) ||
                !std::isfinite(box[2 * k + 1]))
                return nbx_fail(c, NBX_ERR_INVALID, "periodic bounds need finite lo < hi");
        for (int k = 0; k < 6; ++k) bc.b[k] = box[k];
    } else {
        return nbx_fail(c, NBX_ERR_INVALID, "unknown boundary condition kind");
    }
    c->bc = bc;
    invalidate(c);
    return NBX_OK;
}

extern "C" int nbx_set_kb(nbx_ctx* c, double kb) {
    if (!c) return NBX_ERR_INVALID;
    if (!(kb > 0.0) || !std::isfinite(kb)) return nbx_fail(c, NBX_ERR_INVALID, "kb must be > 0");
    c->kb = kb;
    invalidate(c);
    return NBX_OK;
}

extern "C" int nbx_set_precision(nbx_ctx* c, int32_t mode) {
    if (!c) return NBX_ERR_INVALID;
    if (mode != NBX_PREC_FP64_REF && mode != NBX_PREC_MIXED)
        return nbx_fail(c, NBX_ERR_INVALID, "unknown precision mode");
    c->prec = mode;
    invalidate(c);
    return NBX_OK;
}

extern "C" int nbx_set_skin(nbx_ctx* c, double skin) {
    if (!c) return NBX_ERR_INVALID;
    if (!(skin >= 0.0) || !std::isfinite(skin))
        return nbx_fail(c, NBX_ERR_INVALID, "skin must be >= 0 and finite");
    c->skin = skin;
    c->skin_set = true;
    invalidate(c);
    return NBX_OK;
}

static int add_term(nbx_ctx* c, int kind, double p0, double p1, double p2) {
    nbx_term t{};
    t.kind = kind;
    t.p[0] = p0; t.p[1] = p1; t.p[2] = p2;
    if (c->sys == NBX_SYS_WATER) {
        if (kind == NBX_TERM_LJ) { c->water_lj = t; c->water_lj_set = true; }
        else if (kind == NBX_TERM_ELEC || kind == NBX_TERM_RF) {
            c->water_elec = t; c->water_elec_set = true;
        } else {
            return nbx_fail(c, NBX_ERR_UNSUPPORTED, "water systems take LJ and electrostatics only");
        }
    } else {
        if (c->terms.size() >= 8) return nbx_fail(c, NBX_ERR_INVALID, "at most 8 potentials");
        c->terms.push_back(t);
    }
    invalidate(c);
    return NBX_OK;
}

extern "C" int nbx_add_lennard_jones(nbx_ctx* c, double eps, double sigma, double R) {
    if (!c) return NBX_ERR_INVALID;
    if (!std::isfinite(eps) || !(sigma > 0.0) || !std::isfinite(sigma) || !(R > 0.0))
        return nbx_fail(c, NBX_ERR_INVALID, "LennardJonesParameters: need finite eps, sigma > 0, R > 0");
    return add_term(c, NBX_TERM_LJ, eps, sigma, R);
}

extern "C" int nbx_add_electrostatic(nbx_ctx* c, double k, double R) {
    if (!c) return NBX_ERR_INVALID;
    if (!c->has_q) return nbx_fail(c, NBX_ERR_INVALID, "electrostatics needs charges");
    if (!std::isfinite(k) || !(R > 0.0))
        return nbx_fail(c, NBX_ERR_INVALID, "ElectrostaticParameters: need finite k and R > 0");
    return add_term(c, NBX_TERM_ELEC, k, R, 0.0);
}

extern "C" int nbx_add_reaction_field(nbx_ctx* c, double k, double R, double eps_rf) {
    if (!c) return NBX_ERR_INVALID;
    if (!c->has_q) return nbx_fail(c, NBX_ERR_INVALID, "reaction field needs charges");
    if (!std::isfinite(k) || !(R > 0.0) || !std::isfinite(R) || !(eps_rf >= 1.0))
        return nbx_fail(c, NBX_ERR_INVALID, "reaction field: need finite k, finite R > 0, eps_rf >= 1");
    return add_term(c, NBX_TERM_RF, k, R, eps_rf);
}

extern "C" int nbx_add_gravitational(nbx_ctx* c, double G) {
    if (!c) return NBX_ERR_INVALID;
    if (!std::isfinite(G)) return nbx_fail(c, NBX_ERR_INVALID, "G must be finite");
    return add_term(c, NBX_TERM_GRAV, G, 0.0, 0.0);
}

static int upload_aos(nbx_ctx* c, const double* host, double* d_soa);
static bool all_finite(const double* p, int64_t count);

extern "C" int nbx_add_magnetostatic(nbx_ctx* c, double mu_4pi) {
    if (!c) return NBX_ERR_INVALID;
    if (c->sys != NBX_SYS_POTENTIAL)
        return nbx_fail(c, NBX_ERR_UNSUPPORTED, "magnetic dipoles apply to potential systems");
    if (!c->mm_set) return nbx_fail(c, NBX_ERR_INVALID, "magnetostatics needs moments (nbx_set_moments)");
    if (!std::isfinite(mu_4pi)) return nbx_fail(c, NBX_ERR_INVALID, "mu/4pi must be finite");
    return add_term(c, NBX_TERM_MAG, mu_4pi, 0.0, 0.0);
}

extern "C" int nbx_set_moments(nbx_ctx* c, const double* mm) {
    if (!c || !mm) return NBX_ERR_INVALID;
    const int64_t n = c->n;
    if (!all_finite(mm, 3 * n)) return nbx_fail(c, NBX_ERR_NONFINITE, "moments hold NaN or Inf");
    if (!c->d_mm) NBX_HIP(c, hipMalloc((void**)&c->d_mm, sizeof(double) * 3 * (size_t)std::max<int64_t>(n, 1)));
    if (!c->d_mm4) NBX_HIP(c, hipMalloc(&c->d_mm4, sizeof(float) * 4 * (size_t)std::max<int64_t>(n, 1)));
    if (n > 0) {
        int rc = upload_aos(c, mm, c->d_mm);
        if (rc) return rc;
        std::vector<float> f4((size_t)4 * n);
        for (int64_t i = 0; i < n; ++i) {
            f4[4 * i] = (float)mm[3 * i];
            f4[4 * i + 1] = (float)mm[3 * i + 1];
            f4[4 * i + 2] = (float)mm[3 * i + 2];
            f4[4 * i + 3] = 0.0f;
        }
        NBX_HIP(c, hipMemcpyAsync(c->d_mm4, f4.data(), sizeof(float) * f4.size(),
                                  hipMemcpyHostToDevice, c->stream));
        NBX_HIP(c, hipStreamSynchronize(c->stream));
    }
    c->mm_set = true;
    invalidate(c);
    return NBX_OK;
}

extern "C" int nbx_set_spcfw(nbx_ctx* c, double rOH, double aHOH, double kbond, double kangle) {
    if (!c) return NBX_ERR_INVALID;
    if (c->sys != NBX_SYS_WATER)
        return nbx_fail(c, NBX_ERR_UNSUPPORTED, "SPCFwParameters apply to water systems");
    if (!(rOH > 0.0) || !std::isfinite(aHOH) || !std::isfinite(kbond) || !std::isfinite(kangle))
        return nbx_fail(c, NBX_ERR_INVALID, "SPCFwParameters: need rOH > 0 and finite values");
    c->rOH = rOH; c->aHOH = aHOH; c->kbond = kbond; c->kangle = kangle;
    c->spcfw_set = true;
    invalidate(c);
    return NBX_OK;
}

// the extra coordinate column of the reference's Nose-Hoover state starts at zero
// (gather_bodies_initial_coordinates, src/nbody_to_ode.jl:1-20)
static int reset_nosehoover(nbx_ctx* c) {
    c->nh_advanced = false;
    NBX_HIP(c, hipMemsetAsync(c->d_scal + NBX_SCAL_ZETA, 0, 2 * sizeof(double), c->stream));
    NBX_HIP(c, hipStreamSynchronize(c->stream));
    return NBX_OK;
}

extern "C" int nbx_set_thermostat(nbx_ctx* c, int32_t kind, double T, double param) {
    if (!c) return NBX_ERR_INVALID;
    if (kind < NBX_THERM_NONE || kind > NBX_THERM_NOSEHOOVER)
        return nbx_fail(c, NBX_ERR_INVALID, "unknown thermostat kind");
    if (kind == NBX_THERM_NOSEHOOVER && T <= 0.0)
        return nbx_fail(c, NBX_ERR_INVALID, "Nose-Hoover needs a target temperature > 0");
    if (kind != NBX_THERM_NONE) {
        if (!(T >= 0.0) || !std::isfinite(T))
            return nbx_fail(c, NBX_ERR_INVALID, "thermostat temperature must be >= 0");
        if (!(param > 0.0) || !std::isfinite(param))
            return nbx_fail(c, NBX_ERR_INVALID, "thermostat parameter must be > 0");
    }
    c->therm = kind; c->therm_T = T; c->therm_p = param;
    invalidate(c);
    return reset_nosehoover(c);
}

// ---------------------------------------------------------------------------
// State transfer
// ---------------------------------------------------------------------------

static int upload_aos(nbx_ctx* c, const double* host, double* d_soa) {
    const size_t b = sizeof(double) * (size_t)(3 * c->n);
    NBX_HIP(c, hipMemcpyAsync(c->d_tmp, host, b, hipMemcpyHostToDevice, c->stream));
    ref64_aos_to_soa(c->stream, c->d_tmp, d_soa, c->n);
    NBX_HIP(c, hipGetLastError());
    return NBX_OK;
}

static int download_aos(const nbx_ctx* cc, const double* d_soa, double* host) {
    nbx_ctx* c = const_cast<nbx_ctx*>(cc);
    const size_t b = sizeof(double) * (size_t)(3 * c->n);
    ref64_soa_to_aos(c->stream, d_soa, c->d_tmp, c->n);
    NBX_HIP(c, hipMemcpyAsync(host, c->d_tmp, b, hipMemcpyDeviceToHost, c->stream));
    NBX_HIP(c, hipStreamSynchronize(c->stream));
    return NBX_OK;
}

static bool all_finite(const double* p, int64_t count) {
    for (int64_t i = 0; i < count; ++i)
        if (!std::isfinite(p[i])) return false;
    return true;
}

extern "C" int nbx_set_state(nbx_ctx* c, const double* u, const double* v) {
    if (!c) return NBX_ERR_INVALID;
    if (c->n == 0) { c->state_set = true; c->steps_taken = 0; return NBX_OK; }
    if (!u) return nbx_fail(c, NBX_ERR_INVALID, "positions are NULL");
    if (!all_finite(u, 3 * c->n) || (v && !all_finite(v, 3 * c->n)))
        return nbx_fail(c, NBX_ERR_NONFINITE, "state holds NaN or Inf");
    int rc = upload_aos(c, u, c->d_x);
    if (rc) return rc;
    if (v) {
        rc = upload_aos(c, v, c->d_v);
        if (rc) return rc;
    } else {
        ref64_fill(c->stream, c->d_v, 3 * c->n, 0.0);
    }
    NBX_HIP(c, hipStreamSynchronize(c->stream));
    c->state_set = true;
    c->accel_valid = false;
    c->nl.ready = false;
    c->steps_taken = 0;
    if ((rc = mx_reset_order(c))) return rc;
    return reset_nosehoover(c);
}

extern "C" int nbx_get_state(const nbx_ctx* c, double* u, double* v) {
    if (!c) return NBX_ERR_INVALID;
    if (!c->state_set) return nbx_fail(const_cast<nbx_ctx*>(c), NBX_ERR_INVALID, "no state set");
    if (c->n == 0) return NBX_OK;
    if (u) { int rc = download_aos(c, c->d_x, u); if (rc) return rc; }
    if (v) { int rc = download_aos(c, c->d_v, v); if (rc) return rc; }
    return NBX_OK;
}

extern "C" int64_t nbx_steps_taken(const nbx_ctx* c) { return c ? c->steps_taken : -1; }

// Nose-Hoover extended coordinate and its rate (the reference's u[3n+1] and v[3n+1]) at
// the state; after a mixed run the last drift has already advanced them for the next step,
// and the values before that advance are the state's
extern "C" int nbx_get_thermostat_state(const nbx_ctx* c, double* zeta, double* xi) {
    if (!c) return NBX_ERR_INVALID;
    nbx_ctx* m = const_cast<nbx_ctx*>(c);
    double h[2] = {0.0, 0.0};
    const bool ahead = c->nh_advanced && c->therm == NBX_THERM_NOSEHOOVER;
    const int slot = ahead ? NBX_SCAL_ZETA_PREV : NBX_SCAL_ZETA;
    NBX_HIP(m, hipMemcpyAsync(h, c->d_scal + slot, sizeof(h), hipMemcpyDeviceToHost,
                              c->stream));
    NBX_HIP(m, hipStreamSynchronize(c->stream));
    if (zeta) *zeta = h[0];
    if (xi) *xi = h[1];
    return NBX_OK;
}

// ---------------------------------------------------------------------------
// FP64 reference-order right-hand side
// ---------------------------------------------------------------------------

static int check_config(nbx_ctx* c) {
    if (c->sys == NBX_SYS_WATER) {
        if (!c->water_lj_set || !c->water_elec_set || !c->spcfw_set)
            return nbx_fail(c, NBX_ERR_INVALID,
                            "water needs Lennard-Jones, electrostatic and SPC/Fw parameters");
    }
    for (const nbx_term& t : c->terms)
        if ((t.kind == NBX_TERM_ELEC || t.kind == NBX_TERM_RF) && !c->has_q)
            return nbx_fail(c, NBX_ERR_INVALID, "electrostatics needs charges");
    return NBX_OK;
}

// one potential term's closure for atoms [ib, ie), adding into dv (SoA)
static void ref_term(nbx_ctx* c, const nbx_term& t, const double* x, double* dv, int64_t ib,
                     int64_t ie) {
    hipStream_t s = c->stream;
    const int water = c->sys == NBX_SYS_WATER;
    switch (t.kind) {
    case NBX_TERM_LJ:
        if (water) {
            // O atoms only, partners = O atoms (src/nbody_to_ode.jl:302-314)
            int64_t first = ((ib + 2) / 3) * 3;
            ref64_lj(s, x, dv, c->d_ms, c->n, first, ie, 3, 3, c->n / 3, t.p[0], t.p[1], t.p[2],
                     c->bc);
        } else {
            ref64_lj(s, x, dv, c->d_ms, c->n, ib, ie, 1, 1, c->n, t.p[0], t.p[1], t.p[2], c->bc);
        }
        break;
    case NBX_TERM_ELEC:
        ref64_elec(s, x, dv, c->d_ms, c->d_qs, c->n, ib, ie, water, t.p[0], t.p[1], c->bc);
        break;
    case NBX_TERM_RF:
        ref64_rf(s, x, dv, c->d_ms, c->d_qs, c->n, ib, ie, water, t.p[0], t.p[1], t.p[2], c->bc);
        break;
    case NBX_TERM_GRAV:
        ref64_grav(s, x, dv, c->d_ms, c->n, ib, ie, t.p[0]);
        break;
    case NBX_TERM_MAG:
        ref64_mag(s, x, dv, c->d_ms, c->d_mm, c->n, ib, ie, t.p[0]);
        break;
    default:
        break;
    }
}

// soode_system!(dv, v, u): dv overwritten (src/nbody_to_ode.jl:474-488, :502-532)
// xi_out: where the Nose-Hoover rate goes (the context's during runs, scratch otherwise)
static int ref_rhs(nbx_ctx* c, const double* x, const double* v, double* dv, double* xi_out) {
    hipStream_t s = c->stream;
    const int64_t n = c->n;
    ref64_fill(s, dv, 3 * n, 0.0);
    if (c->sys == NBX_SYS_WATER) {
        // per-atom closure order: O = (electrostatic, bond, LJ); H = (electrostatic, bond).
        // Each atom's dv only receives its own closures, so running the terms as
        // whole-array passes in that order reproduces the per-atom sums exactly.
        ref_term(c, c->water_elec, x, dv, 0, n);
        ref64_bond(s, x, dv, c->d_ms, n, 0, n, c->rOH, c->kbond);
        ref_term(c, c->water_lj, x, dv, 0, n);
        ref64_angle(s, x, dv, c->d_ms, n, 0, n / 3, c->aHOH, c->kangle);
    } else {
        for (const nbx_term& t : c->terms) ref_term(c, t, x, dv, 0, n);
    }
    const int64_t Nc = c->sys == NBX_SYS_WATER ? 2 * (n / 3) : 0;
    if (c->therm == NBX_THERM_BERENDSEN) {
        ref64_temperature(s, v, c->d_ms, n, c->kb, Nc, c->d_red, c->d_scal);
        ref64_berendsen(s, dv, v, n, c->d_scal, c->therm_T, 0.5 / c->therm_p);
    } else if (c->therm == NBX_THERM_NOSEHOOVER) {
        // src/thermostats.jl:116-122 (md_temperature over the n atoms, ndf = 3n - Nc)
        ref64_temperature(s, v, c->d_ms, n, c->kb, Nc, c->d_red, c->d_scal);
        ref64_nosehoover(s, dv, v, n, c->d_scal, xi_out, c->therm_T, c->therm_p, 3 * n - Nc);
    }
    NBX_HIP(c, hipGetLastError());
    return NBX_OK;
}

// Scratch triple for calls that must not touch the context state
struct scratch3 {
    double *x = nullptr, *v = nullptr, *dv = nullptr;
    ~scratch3() { free_ptr(x); free_ptr(v); free_ptr(dv); }
};

static int alloc_scratch(nbx_ctx* c, scratch3& s) {
    const size_t b = sizeof(double) * (size_t)(3 * c->n);
    NBX_HIP(c, hipMalloc((void**)&s.x, b));
    NBX_HIP(c, hipMalloc((void**)&s.v, b));
    NBX_HIP(c, hipMalloc((void**)&s.dv, b));
    return NBX_OK;
}

extern "C" int nbx_rhs(nbx_ctx* c, const double* u, const double* v, double* dv) {
    if (!c) return NBX_ERR_INVALID;
    if (c->n == 0) return NBX_OK;
    if (!u || !dv) return nbx_fail(c, NBX_ERR_INVALID, "u and dv are required");
    int rc = check_config(c);
    if (rc) return rc;
    if (!all_finite(u, 3 * c->n)) return nbx_fail(c, NBX_ERR_NONFINITE, "u holds NaN or Inf");
    if ((c->therm == NBX_THERM_BERENDSEN || c->therm == NBX_THERM_NOSEHOOVER) && !v)
        return nbx_fail(c, NBX_ERR_INVALID, "the Berendsen / Nose-Hoover term needs v");
    scratch3 s;
    if ((rc = alloc_scratch(c, s))) return rc;
    if ((rc = upload_aos(c, u, s.x))) return rc;
    if (v) { if ((rc = upload_aos(c, v, s.v))) return rc; }
    else ref64_fill(c->stream, s.v, 3 * c->n, 0.0);
    if ((rc = ref_rhs(c, s.x, s.v, s.dv, c->d_scal + NBX_SCAL_XI_SCRATCH))) return rc;
    return download_aos(c, s.dv, dv);
}

extern "C" int nbx_accelerations(nbx_ctx* c, int32_t term, const double* u, int64_t i_begin,
                                 int64_t i_end, double* dv) {
    if (!c) return NBX_ERR_INVALID;
    if (!u || !dv) return nbx_fail(c, NBX_ERR_INVALID, "u and dv are required");
    if (i_begin < 0 || i_end > c->n || i_begin > i_end)
        return nbx_fail(c, NBX_ERR_BOUNDS,
                        "atom range [" + std::to_string(i_begin) + ", " + std::to_string(i_end) +
                            ") outside 0.." + std::to_string(c->n));
    const int nterms = c->sys == NBX_SYS_WATER ? 4 : (int)c->terms.size();
    if (term < 0 || term >= nterms)
        return nbx_fail(c, NBX_ERR_BOUNDS, "potential term " + std::to_string(term) + " not defined");
    if (c->n == 0 || i_begin == i_end) return NBX_OK;
    if (c->sys == NBX_SYS_WATER) {
        if ((term == 0 && !c->water_elec_set) || (term == 1 && !c->water_lj_set) ||
            (term >= 2 && !c->spcfw_set))
            return nbx_fail(c, NBX_ERR_INVALID, "water term not configured");
    }
    if (!all_finite(u, 3 * c->n)) return nbx_fail(c, NBX_ERR_NONFINITE, "u holds NaN or Inf");
    scratch3 s;
    int rc;
    if ((rc = alloc_scratch(c, s))) return rc;
    if ((rc = upload_aos(c, u, s.x))) return rc;
    if ((rc = upload_aos(c, dv, s.dv))) return rc;
    if (c->sys == NBX_SYS_WATER) {
        if (term == 0) ref_term(c, c->water_elec, s.x, s.dv, i_begin, i_end);
        else if (term == 1) ref_term(c, c->water_lj, s.x, s.dv, i_begin, i_end);
        else if (term == 2) ref64_bond(c->stream, s.x, s.dv, c->d_ms, c->n, i_begin, i_end, c->rOH,
                                       c->kbond);
        else {
            // group closure over molecules whose O index lies in [i_begin, i_end)
            ref64_angle(c->stream, s.x, s.dv, c->d_ms, c->n, (i_begin + 2) / 3, (i_end + 2) / 3,
                        c->aHOH, c->kangle);
        }
    } else {
        ref_term(c, c->terms[term], s.x, s.dv, i_begin, i_end);
    }
    NBX_HIP(c, hipGetLastError());
    return download_aos(c, s.dv, dv);
}

// ---------------------------------------------------------------------------
// Velocity Verlet
// ---------------------------------------------------------------------------

static int save_frame(nbx_ctx* c, int64_t f, double* uf, double* vf) {
    const int64_t n3 = 3 * c->n;
    int rc = download_aos(c, c->d_x, uf + f * n3);
    if (rc) return rc;
    return download_aos(c, c->d_v, vf + f * n3);
}

static int ref_vv(nbx_ctx* c, double dt, int64_t nsteps, uint64_t seed, int64_t save_every,
                  double* uf, double* vf, int64_t* nframes) {
    hipStream_t s = c->stream;
    const int64_t n = c->n;
    int rc;
    if (c->nh_advanced && c->therm == NBX_THERM_NOSEHOOVER) {
        // a mixed run before this one left zeta and xi advanced for the next step: back to
        // the state's values, which this path advances at its first drift
        NBX_HIP(c, hipMemcpyAsync(c->d_scal + NBX_SCAL_ZETA, c->d_scal + NBX_SCAL_ZETA_PREV,
                                  2 * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    c->nh_advanced = false;
    if (!c->accel_valid) {
        if ((rc = ref_rhs(c, c->d_x, c->d_v, c->d_a, c->d_scal + NBX_SCAL_XI))) return rc;
        c->accel_valid = true;
    }
    const int water = c->sys == NBX_SYS_WATER;
    for (int64_t k = 1; k <= nsteps; ++k) {
        const int64_t step = c->steps_taken + 1;
        if (c->therm == NBX_THERM_LANGEVIN)
            ref64_langevin_half(s, c->d_v, c->d_ms, n, c->kb, c->therm_T, c->therm_p, dt, seed,
                                step, NBX_RNG_LANGEVIN_A);
        ref64_drift(s, c->d_x, c->d_v, c->d_a, c->d_vh, n, dt);
        if (c->therm == NBX_THERM_NOSEHOOVER) ref64_nh_drift(s, c->d_scal, dt);
        if ((rc = ref_rhs(c, c->d_x, c->d_vh, c->d_a, c->d_scal + NBX_SCAL_XI))) return rc;
        ref64_kick(s, c->d_v, c->d_vh, c->d_a, n, dt);
        if (c->therm == NBX_THERM_LANGEVIN)
            ref64_langevin_half(s, c->d_v, c->d_ms, n, c->kb, c->therm_T, c->therm_p, dt, seed,
                                step, NBX_RNG_LANGEVIN_B);
        if (c->therm == NBX_THERM_ANDERSEN)
            ref64_andersen(s, c->d_v, c->d_ms, n, water, c->kb, c->therm_T, c->therm_p * dt, seed,
                           step);
        NBX_HIP(c, hipGetLastError());
        c->steps_taken = step;
        if (save_every > 0 && uf && k % save_every == 0) {
            if ((rc = save_frame(c, *nframes, uf, vf))) return rc;
            ++*nframes;
        }
    }
    return NBX_OK;
}

int mx_vv(nbx_ctx* c, double dt, int64_t nsteps, uint64_t seed, int64_t save_every, double* uf,
          double* vf, int64_t* nframes);
int mx_supported(nbx_ctx* c, std::string* why);

static int check_finite_state(nbx_ctx* c) {
    int* d_flag = reinterpret_cast<int*>(c->d_scal + 8);
    NBX_HIP(c, hipMemsetAsync(d_flag, 0, sizeof(int), c->stream));
    ref64_check_finite(c->stream, c->d_x, 3 * c->n, d_flag);
    ref64_check_finite(c->stream, c->d_v, 3 * c->n, d_flag);
    int h = 0;
    NBX_HIP(c, hipMemcpyAsync(&h, d_flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    NBX_HIP(c, hipStreamSynchronize(c->stream));
    if (h) return nbx_fail(c, NBX_ERR_NONFINITE, "the state became NaN or Inf during the run");
    return NBX_OK;
}

extern "C" int nbx_vv_run(nbx_ctx* c, double dt, int64_t nsteps, uint64_t seed,
                          int64_t save_every, double* u_frames, double* v_frames) {
    if (!c) return NBX_ERR_INVALID;
    if (!c->state_set) return nbx_fail(c, NBX_ERR_INVALID, "no state set");
    if (!(dt > 0.0) || !std::isfinite(dt)) return nbx_fail(c, NBX_ERR_INVALID, "dt must be > 0");
    if (nsteps < 0) return nbx_fail(c, NBX_ERR_INVALID, "nsteps < 0");
    if (save_every > 0 && (!u_frames || !v_frames))
        return nbx_fail(c, NBX_ERR_INVALID, "frames requested but frame buffers are NULL");
    int rc = check_config(c);
    if (rc) return rc;
    if (c->n == 0) { c->steps_taken += nsteps; return NBX_OK; }
    int64_t nframes = 0;
    if (save_every > 0) {
        if ((rc = save_frame(c, 0, u_frames, v_frames))) return rc;
        nframes = 1;
    }
    if (c->prec == NBX_PREC_MIXED) {
        std::string why;
        if (!mx_supported(c, &why)) return nbx_fail(c, NBX_ERR_UNSUPPORTED, why);
        rc = mx_vv(c, dt, nsteps, seed, save_every, u_frames, v_frames, &nframes);
    } else {
        rc = ref_vv(c, dt, nsteps, seed, save_every, u_frames, v_frames, &nframes);
    }
    if (rc) return rc;
    NBX_HIP(c, hipStreamSynchronize(c->stream));
    return check_finite_state(c);
}

// ---------------------------------------------------------------------------
// Euler-Maruyama (the reference's Langevin SDE path)
// ---------------------------------------------------------------------------

static int ref_em(nbx_ctx* c, double dt, int64_t nsteps, uint64_t seed, int64_t save_every,
                  double* uf, double* vf, int64_t* nframes) {
    hipStream_t s = c->stream;
    int rc;
    if (!c->accel_valid) {
        if ((rc = ref_rhs(c, c->d_x, c->d_v, c->d_a, c->d_scal + NBX_SCAL_XI))) return rc;
        c->accel_valid = true;
    }
    const double sigma = std::sqrt(2 * c->therm_p * c->kb * c->therm_T / c->h_ms[0]);
    for (int64_t k = 1; k <= nsteps; ++k) {
        const int64_t step = c->steps_taken + 1;
        ref64_em(s, c->d_x, c->d_v, c->d_a, c->n, dt, c->therm_p, sigma, seed, step);
        if ((rc = ref_rhs(c, c->d_x, c->d_v, c->d_a, c->d_scal + NBX_SCAL_XI))) return rc;
        NBX_HIP(c, hipGetLastError());
        c->steps_taken = step;
        if (save_every > 0 && uf && k % save_every == 0) {
            if ((rc = save_frame(c, *nframes, uf, vf))) return rc;
            ++*nframes;
        }
    }
    return NBX_OK;
}

extern "C" int nbx_em_run(nbx_ctx* c, double dt, int64_t nsteps, uint64_t seed,
                          int64_t save_every, double* u_frames, double* v_frames) {
    if (!c) return NBX_ERR_INVALID;
    if (!c->state_set) return nbx_fail(c, NBX_ERR_INVALID, "no state set");
    if (!(dt > 0.0) || !std::isfinite(dt)) return nbx_fail(c, NBX_ERR_INVALID, "dt must be > 0");
    if (nsteps < 0) return nbx_fail(c, NBX_ERR_INVALID, "nsteps < 0");
    if (save_every > 0 && (!u_frames || !v_frames))
        return nbx_fail(c, NBX_ERR_INVALID, "frames requested but frame buffers are NULL");
    if (c->therm != NBX_THERM_LANGEVIN)
        return nbx_fail(c, NBX_ERR_INVALID,
                        "Euler-Maruyama integrates the Langevin SDE: set the Langevin thermostat");
    if (c->sys == NBX_SYS_WATER)
        return nbx_fail(c, NBX_ERR_UNSUPPORTED,
                        "the Langevin SDE is provided for potential systems only");
    int rc = check_config(c);
    if (rc) return rc;
    if (c->n == 0) { c->steps_taken += nsteps; return NBX_OK; }
    int64_t nframes = 0;
    if (save_every > 0) {
        if ((rc = save_frame(c, 0, u_frames, v_frames))) return rc;
        nframes = 1;
    }
    if (c->prec == NBX_PREC_MIXED) {
        std::string why;
        if (!mx_supported(c, &why)) return nbx_fail(c, NBX_ERR_UNSUPPORTED, why);
        c->em = 1;
        rc = mx_vv(c, dt, nsteps, seed, save_every, u_frames, v_frames, &nframes);
        c->em = 0;
    } else {
        rc = ref_em(c, dt, nsteps, seed, save_every, u_frames, v_frames, &nframes);
    }
    if (rc) return rc;
    NBX_HIP(c, hipStreamSynchronize(c->stream));
    return check_finite_state(c);
}

// ---------------------------------------------------------------------------
// Energies, temperature, RDF
// ---------------------------------------------------------------------------

extern "C" int nbx_energies(nbx_ctx* c, double* ke, double* pe, double* pe_grav) {
    if (!c) return NBX_ERR_INVALID;
    if (!c->state_set) return nbx_fail(c, NBX_ERR_INVALID, "no state set");
    int rc = check_config(c);
    if (rc) return rc;
    hipStream_t s = c->stream;
    const int64_t n = c->n;
    double* out = c->d_scal + 1;  // [1] ke, [2..7] terms
    ref64_fill(s, out, 7, 0.0);
    ref64_kinetic(s, c->d_v, c->d_ms, n, c->d_tmp, out);
    double grav = 0.0;
    auto fetch = [&](double* d) -> double {
        double h = 0.0;
        (void)hipMemcpyAsync(&h, d, sizeof(double), hipMemcpyDeviceToHost, s);
        (void)hipStreamSynchronize(s);
        return h;
    };
    const double h_ke = fetch(out);
    double h_pe = 0.0;
    if (c->sys == NBX_SYS_WATER) {
        // src/nbody_simulation_result.jl:266-291: LJ (O atoms), electrostatic, bonds, angles
        const nbx_term& lj = c->water_lj;
        ref64_lj_pe(s, c->d_x, n, 3, n / 3, lj.p[0], lj.p[1], lj.p[2], c->bc, c->d_tmp, out + 1);
        h_pe += 4.0 * lj.p[0] * fetch(out + 1);
        const nbx_term& el = c->water_elec;
        ref64_elec_pe(s, c->d_x, c->d_qs, n, 1, el.p[0], el.p[1], el.kind == NBX_TERM_RF, el.p[2],
                      c->bc, c->d_tmp, out + 2);
        h_pe += fetch(out + 2) * el.p[0];
        ref64_bonded_pe(s, c->d_x, n, c->rOH, c->kbond, c->aHOH, c->kangle, c->d_tmp, out + 3);
        h_pe += fetch(out + 3);
        h_pe += fetch(out + 4);
    } else {
        // src/nbody_simulation_result.jl:239-264: Lennard-Jones first, then electrostatics;
        // gravity has no potential energy there and is reported separately
        for (const nbx_term& t : c->terms) {
            if (t.kind != NBX_TERM_LJ) continue;
            ref64_lj_pe(s, c->d_x, n, 1, n, t.p[0], t.p[1], t.p[2], c->bc, c->d_tmp, out + 1);
            h_pe += 4.0 * t.p[0] * fetch(out + 1);
        }
        for (const nbx_term& t : c->terms) {
            if (t.kind != NBX_TERM_ELEC && t.kind != NBX_TERM_RF) continue;
            ref64_elec_pe(s, c->d_x, c->d_qs, n, 0, t.p[0], t.p[1], t.kind == NBX_TERM_RF,
                          t.p[2], c->bc, c->d_tmp, out + 2);
            h_pe += fetch(out + 2) * t.p[0];
        }
        for (const nbx_term& t : c->terms) {
            if (t.kind != NBX_TERM_GRAV) continue;
            ref64_grav_pe(s, c->d_x, c->d_ms, n, t.p[0], c->d_tmp, out + 3);
            grav += fetch(out + 3);
        }
    }
    NBX_HIP(c, hipGetLastError());
    if (ke) *ke = h_ke;
    if (pe) *pe = h_pe;
    if (pe_grav) *pe_grav = grav;
    return NBX_OK;
}

extern "C" int nbx_temperature(nbx_ctx* c, double* T) {
    if (!c || !T) return NBX_ERR_INVALID;
    if (!c->state_set) return nbx_fail(c, NBX_ERR_INVALID, "no state set");
    const int64_t Nc = c->sys == NBX_SYS_WATER ? 2 * (c->n / 3) : 0;
    if (c->n == 0) { *T = NAN; return NBX_OK; }
    ref64_temperature(c->stream, c->d_v, c->d_ms, c->n, c->kb, Nc, c->d_tmp, c->d_scal + 9);
    NBX_HIP(c, hipMemcpyAsync(T, c->d_scal + 9, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    NBX_HIP(c, hipStreamSynchronize(c->stream));
    return NBX_OK;
}

// g(r) and bin centres from the counts of bins 1..maxbin (src/nbody_simulation_result.jl:
// 695-706).  counts may alias gr: count k is read before g k is written.
static void rdf_normalize(const unsigned long long* counts, int64_t nframes, int64_t nj,
                          double L, int64_t maxbin, double* rs, double* gr) {
    const double dr = L / (double)maxbin;
    const double cc = 4.0 / 3.0 * 3.141592653589793 * (double)nj / (L * L * L);
    for (int64_t bin = 1; bin <= maxbin; ++bin) {
        unsigned long long cnt;
        std::memcpy(&cnt, counts + (bin - 1), sizeof(cnt));
        double rl = (double)(bin - 1) * dr;
        double ru = rl + dr;
        double nideal = cc * (ru * ru * ru - rl * rl * rl);
        gr[bin - 1] = ((double)cnt / (double)(nframes * nj)) / nideal;
        rs[bin - 1] = rl + dr / 2.0;
    }
}

extern "C" int nbx_rdf(nbx_ctx* c, const double* frames, int64_t nframes, int64_t maxbin,
                       double* rs, double* gr) {
    if (!c) return NBX_ERR_INVALID;
    if (c->bc.kind != NBX_BC_CUBIC)
        return nbx_fail(c, NBX_ERR_UNSUPPORTED, "rdf needs CubicPeriodicBoundaryConditions");
    if (!frames || nframes <= 0 || maxbin <= 0 || !rs || !gr)
        return nbx_fail(c, NBX_ERR_INVALID, "rdf: frames, maxbin, rs and gr are required");
    const int water = c->sys == NBX_SYS_WATER;
    const int jstride = water ? 3 : 1;
    const int64_t nj = water ? c->n / 3 : c->n;
    const double L = c->bc.b[0];
    unsigned long long* d_hist = nullptr;
    NBX_HIP(c, hipMalloc((void**)&d_hist, sizeof(unsigned long long) * (size_t)(maxbin + 1)));
    (void)hipMemsetAsync(d_hist, 0, sizeof(unsigned long long) * (size_t)(maxbin + 1), c->stream);
    const size_t b = sizeof(double) * (size_t)(3 * c->n);
    for (int64_t t = 0; t < nframes; ++t) {
        hipError_t e = hipMemcpyAsync(c->d_tmp, frames + t * 3 * c->n, b, hipMemcpyHostToDevice,
                                      c->stream);
        if (e != hipSuccess) { free_ptr(d_hist); return nbx_fail(c, NBX_ERR_DEVICE, "rdf copy"); }
        ref64_rdf_hist(c->stream, c->d_tmp, c->n, jstride, nj, L, maxbin, d_hist);
    }
    std::vector<unsigned long long> h((size_t)maxbin + 1);
    hipError_t e = hipMemcpyAsync(h.data(), d_hist, sizeof(unsigned long long) * h.size(),
                                  hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    free_ptr(d_hist);
    if (e != hipSuccess) return nbx_fail(c, NBX_ERR_DEVICE, "rdf histogram copy");
    rdf_normalize(h.data() + 1, nframes, nj, L, maxbin, rs, gr);
    return NBX_OK;
}

extern "C" int nbx_rdf_noalloc(nbx_ctx* c, const double* frames, int64_t nframes,
                               int64_t maxbin, double* rs, double* gr) {
    if (!c) return NBX_ERR_INVALID;
    if (c->bc.kind != NBX_BC_CUBIC)
        return nbx_fail(c, NBX_ERR_UNSUPPORTED, "rdf needs CubicPeriodicBoundaryConditions");
    if (!frames || nframes <= 0 || maxbin <= 0 || !rs || !gr)
        return nbx_fail(c, NBX_ERR_INVALID, "rdf: frames, maxbin, rs and gr are required");
    if (maxbin > NBX_RDF_NOALLOC_MAXBIN)
        return nbx_fail(c, NBX_ERR_CAPACITY,
                        "nbx_rdf_noalloc: maxbin " + std::to_string(maxbin) + " exceeds " +
                            std::to_string(NBX_RDF_NOALLOC_MAXBIN) +
                            ", the size of its fixed histogram (use nbx_rdf)");
    if (!c->d_rdf_hist) return nbx_fail(c, NBX_ERR_DEVICE, "rdf histogram address not set up");
    const int water = c->sys == NBX_SYS_WATER;
    const int jstride = water ? 3 : 1;
    const int64_t nj = water ? c->n / 3 : c->n;
    const double L = c->bc.b[0];
    unsigned long long* hist = c->d_rdf_hist;
    ref64_zero_u64(c->stream, hist, maxbin + 1);
    const size_t b = sizeof(double) * (size_t)(3 * c->n);
    for (int64_t t = 0; t < nframes; ++t) {
        NBX_HIP(c, hipMemcpyAsync(c->d_tmp, frames + t * 3 * c->n, b, hipMemcpyHostToDevice,
                                  c->stream));
        ref64_rdf_hist(c->stream, c->d_tmp, c->n, jstride, nj, L, maxbin, hist);
    }
    NBX_HIP(c, hipGetLastError());
    // the counts of bins 1..maxbin land in gr (8 bytes each) and are turned into g(r) there,
    // so no host buffer is allocated either
    static_assert(sizeof(double) == sizeof(unsigned long long), "bin counts are 8 bytes");
    NBX_HIP(c, hipMemcpyAsync(gr, hist + 1, sizeof(unsigned long long) * (size_t)maxbin,
                              hipMemcpyDeviceToHost, c->stream));
    NBX_HIP(c, hipStreamSynchronize(c->stream));
    rdf_normalize(reinterpret_cast<const unsigned long long*>(gr), nframes, nj, L, maxbin, rs, gr);
    return NBX_OK;
}

extern "C" int nbx_time_force_kernel(nbx_ctx* c, int32_t iters, double* avg_ms,
                                     const char** name) {
    if (!c || !avg_ms) return NBX_ERR_INVALID;
    if (!c->state_set) return nbx_fail(c, NBX_ERR_INVALID, "no state set");
    if (iters <= 0) return nbx_fail(c, NBX_ERR_INVALID, "iters must be > 0");
    if (c->prec != NBX_PREC_MIXED)
        return nbx_fail(c, NBX_ERR_UNSUPPORTED, "kernel timing is provided for the mixed path");
    return mx_time_force(c, iters, avg_ms, name);
}

extern "C" int nbx_set_kernel_timing(nbx_ctx* c, int32_t enable) {
    if (!c) return NBX_ERR_INVALID;
    c->kt_on = enable != 0;
    c->kt_sum_ms = 0.0;
    c->kt_launches = 0;
    c->kt_used = 0;
    return NBX_OK;
}

extern "C" int nbx_get_kernel_timing(nbx_ctx* c, double* avg_ms, int64_t* launches,
                                     const char** name) {
    if (!c || !avg_ms || !launches) return NBX_ERR_INVALID;
    *launches = c->kt_launches;
    *avg_ms = c->kt_launches > 0 ? c->kt_sum_ms / (double)c->kt_launches : 0.0;
    if (name) *name = mx_force_kernel_name(c);
    return NBX_OK;
}

extern "C" int nbx_copy_probe(nbx_ctx* c, int64_t bytes, int32_t iters, double* avg_ms) {
    if (!c || !avg_ms) return NBX_ERR_INVALID;
    if (bytes < 16 || iters <= 0) return nbx_fail(c, NBX_ERR_INVALID, "bytes >= 16 and iters > 0 required");
    return mx_copy_probe(c, bytes, iters, avg_ms);
}

extern "C" int nbx_valu_probe(nbx_ctx* c, int32_t kind, int32_t reps, double* rate) {
    if (!c || !rate) return NBX_ERR_INVALID;
    if (kind < 0 || kind > 2 || reps <= 0)
        return nbx_fail(c, NBX_ERR_INVALID, "kind 0..2 and reps > 0 required");
    return mx_valu_probe(c, kind, reps, rate);
}

static int learn_ready(nbx_ctx* c) {
    if (!c->state_set) return nbx_fail(c, NBX_ERR_INVALID, "no state set");
    if (c->prec != NBX_PREC_MIXED)
        return nbx_fail(c, NBX_ERR_UNSUPPORTED, "the learned stride runs on the mixed-precision path");
    return NBX_OK;
}

extern "C" int nbx_features(nbx_ctx* c, double* feats, double* frames) {
    if (!c) return NBX_ERR_INVALID;
    int rc;
    if ((rc = learn_ready(c))) return rc;
    return mx_features(c, feats, frames);
}

extern "C" int nbx_learn_energy(nbx_ctx* c, double* ke, double* pe) {
    if (!c || !ke || !pe) return NBX_ERR_INVALID;
    int rc;
    if ((rc = learn_ready(c))) return rc;
    return mx_learn_energy(c, ke, pe);
}

extern "C" int nbx_learn_collect(nbx_ctx* c, nbx_model* m, double dt, int32_t stride,
                                 int64_t nframes, int32_t atoms_per_frame, uint64_t seed) {
    if (!c || !m) return NBX_ERR_INVALID;
    int rc;
    if ((rc = learn_ready(c))) return rc;
    if (!(dt > 0.0) || !std::isfinite(dt)) return nbx_fail(c, NBX_ERR_INVALID, "dt must be > 0");
    if (stride < 1 || stride > 64) return nbx_fail(c, NBX_ERR_INVALID, "stride must be in 1..64");
    if (nframes < 0) return nbx_fail(c, NBX_ERR_INVALID, "nframes must be >= 0");
    if (atoms_per_frame < 1 || atoms_per_frame > c->n)
        return nbx_fail(c, NBX_ERR_INVALID, "atoms_per_frame must be in 1..n");
    if (nframes == 0) return NBX_OK;
    return mx_learn_collect(c, m, dt, stride, nframes, atoms_per_frame, seed);
}

extern "C" int nbx_learned_run(nbx_ctx* c, const nbx_model* m, double dt, int32_t stride,
                               int64_t nsteps, double drift_tol, int32_t check_every,
                               double* report) {
    if (!c || !m) return NBX_ERR_INVALID;
    int rc;
    if ((rc = learn_ready(c))) return rc;
    if (!(dt > 0.0) || !std::isfinite(dt)) return nbx_fail(c, NBX_ERR_INVALID, "dt must be > 0");
    if (stride < 1 || stride > 64) return nbx_fail(c, NBX_ERR_INVALID, "stride must be in 1..64");
    if (nsteps < 0) return nbx_fail(c, NBX_ERR_INVALID, "nsteps must be >= 0");
    if (!(drift_tol > 0.0)) return nbx_fail(c, NBX_ERR_INVALID, "drift_tol must be > 0");
    if (check_every < 1) return nbx_fail(c, NBX_ERR_INVALID, "check_every must be >= 1");
    return mx_learned_run(c, m, dt, stride, nsteps, drift_tol, check_every, report);
}

extern "C" int nbx_autotune(nbx_ctx* c, double dt, int32_t trial_steps, double* out,
                            int32_t nout) {
    if (!c || (nout > 0 && !out) || nout < 0 || nout > 6) return NBX_ERR_INVALID;
    if (!c->state_set) return nbx_fail(c, NBX_ERR_INVALID, "no state set");
    if (c->prec != NBX_PREC_MIXED)
        return nbx_fail(c, NBX_ERR_UNSUPPORTED, "autotuning is provided for the mixed path");
    if (!(dt > 0.0) || !std::isfinite(dt)) return nbx_fail(c, NBX_ERR_INVALID, "dt must be > 0");
    if (trial_steps < 1) return nbx_fail(c, NBX_ERR_INVALID, "trial_steps must be >= 1");
    return mx_autotune(c, dt, trial_steps, out, nout);
}

extern "C" int nbx_get_stats(nbx_ctx* c, int64_t* out, int32_t nout) {
    if (!c || !out || nout < 0 || nout > 8) return NBX_ERR_INVALID;
    int64_t v[8] = {0, c->nl.cap, c->nl.ncell, c->nl.lanes, c->graph_steps, c->build_every,
                    c->stale_replays, c->sorts_applied};
    if (c->nl.d_flags) {
        int h[8] = {0};
        NBX_HIP(c, hipMemcpyAsync(h, c->nl.d_flags, sizeof(h), hipMemcpyDeviceToHost, c->stream));
        NBX_HIP(c, hipStreamSynchronize(c->stream));
        v[0] = h[2];
    }
    for (int k = 0; k < nout; ++k) out[k] = v[k];
    return NBX_OK;
}

extern "C" int nbx_count_pairs(nbx_ctx* c, int64_t* pairs) {
    if (!c || !pairs) return NBX_ERR_INVALID;
    if (!c->state_set) return nbx_fail(c, NBX_ERR_INVALID, "no state set");
    return mx_count_pairs(c, pairs);
}
