// nbx_mixed.cpp -- host orchestration of the mixed-precision path: plan selection
// (Verlet list / all pairs / gravity tiles / water), list sizing, the step sequence
//   drift -> [temperature] -> list rebuild (device-gated) -> force + end of step
// and overflow recovery (a chunk of steps is replayed from a device-side backup with
// larger lists; the counter-based RNG makes the replay identical).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>

#include "nbx_mixed.h"

int mx_gravity_chunks(int64_t n);

namespace {

enum { PLAN_NONBONDED = 1, PLAN_WATER = 2, PLAN_GRAVITY = 3 };

struct mx_plan {
    int kind = 0;
    mx_nb_launch nb{};
    mx_pair_params pp{};
    double G = 0.0;
    double cutoff = 0.0;  // largest finite cutoff (list mode)
};

bool make_plan(const nbx_ctx* c, mx_plan& p, std::string* why) {
    p = mx_plan();
    const nbx_term* lj = nullptr;
    const nbx_term* el = nullptr;
    if (c->sys == NBX_SYS_WATER) {
        if (!c->water_lj_set || !c->water_elec_set || !c->spcfw_set) {
            if (why) *why = "water needs LJ, electrostatic and SPC/Fw parameters";
            return false;
        }
        lj = &c->water_lj;
        el = &c->water_elec;
        p.kind = PLAN_WATER;
        p.nb.water = 1;
    } else {
        int ngrav = 0;
        for (const nbx_term& t : c->terms) {
            if (t.kind == NBX_TERM_GRAV) { ++ngrav; p.G = t.p[0]; }
            else if (t.kind == NBX_TERM_LJ) {
                if (lj) { if (why) *why = "mixed path: one Lennard-Jones term at most"; return false; }
                lj = &t;
            } else {
                if (el) { if (why) *why = "mixed path: one electrostatic term at most"; return false; }
                el = &t;
            }
        }
        if (ngrav) {
            if (ngrav > 1 || lj || el) {
                if (why) *why = "mixed path: gravity cannot be combined with other potentials";
                return false;
            }
            p.kind = PLAN_GRAVITY;
            return true;
        }
        if (!lj && !el) {
            if (why) *why = "mixed path: no potential defined";
            return false;
        }
        p.kind = PLAN_NONBONDED;
    }
    p.nb.has_lj = lj != nullptr;
    p.nb.el_kind = el ? (el->kind == NBX_TERM_RF ? 2 : 1) : 0;
    if (lj) {
        p.pp.lj_s2 = (float)(lj->p[1] * lj->p[1]);
        p.pp.lj_R2 = (float)(lj->p[2] * lj->p[2]);
        p.pp.lj_c = 24.0 * lj->p[0];
    }
    if (el) {
        p.pp.el_k = el->p[0];
        p.pp.el_R2 = (float)(el->p[1] * el->p[1]);
        if (el->kind == NBX_TERM_RF) {
            const double R = el->p[1], e = el->p[2];
            const double krf = std::isinf(e) ? 1.0 / (2.0 * R * R * R)
                                             : (e - 1.0) / ((2.0 * e + 1.0) * (R * R * R));
            p.pp.krf2 = (float)(2.0 * krf);
        }
    }
    bool finite = true;
    if (lj) { finite = finite && std::isfinite(lj->p[2]); p.cutoff = std::max(p.cutoff, lj->p[2]); }
    if (el) { finite = finite && std::isfinite(el->p[1]); p.cutoff = std::max(p.cutoff, el->p[1]); }
    p.nb.mode_list = (c->bc.kind == NBX_BC_CUBIC && finite) ? 1 : 0;
    if (p.nb.mode_list) {
        p.pp.L = c->bc.b[0];
        p.pp.invL = 1.0 / c->bc.b[0];
    }
    p.nb.lanes = c->nl.lanes;
    return true;
}

void free_nl(nbx_nlist& nl) {
    if (nl.d_nbr) (void)hipFree(nl.d_nbr);
    if (nl.d_nnb) (void)hipFree(nl.d_nnb);
    if (nl.d_cell_count) (void)hipFree(nl.d_cell_count);
    if (nl.d_cell_atoms) (void)hipFree(nl.d_cell_atoms);
    if (nl.d_xref) (void)hipFree(nl.d_xref);
    if (nl.d_flags) (void)hipFree(nl.d_flags);
    nl.d_nbr = nl.d_nnb = nl.d_cell_count = nl.d_cell_atoms = nl.d_flags = nullptr;
    nl.d_xref = nullptr;
    nl.ready = false;
}

int round_up(int x, int m) { return ((x + m - 1) / m) * m; }

// size and allocate the lists for the current configuration (cap/ccap may be grown)
int alloc_lists(nbx_ctx* c, const mx_plan& p, double grow) {
    nbx_nlist& nl = c->nl;
    const int lanes = nl.lanes;
    free_nl(nl);
    nl.lanes = lanes;
    const int64_t n = c->n;
    const double L = c->bc.b[0];
    nl.cutoff = p.cutoff;
    nl.skin = c->skin_set ? c->skin : 0.12 * p.cutoff;
    nl.rlist = p.cutoff + nl.skin;
    const double rho = (double)n / (L * L * L);
    const double est = 4.0 / 3.0 * M_PI * nl.rlist * nl.rlist * nl.rlist * rho;
    int cap = (int)std::ceil(grow * (1.6 * est + 24.0));
    cap = std::min<int64_t>(cap, std::max<int64_t>(n, 1));
    nl.cap = round_up(std::max(cap, 1), lanes);
    int nc = (int)std::floor(L / nl.rlist);
    if (nc < 3) nc = 0;
    if (nc > 256) nc = 256;
    nl.ncell = nc;
    const double per_cell = nc ? (double)n / ((double)nc * nc * nc) : 0.0;
    nl.ccap = nc ? (int)std::ceil(grow * (2.0 * per_cell + 16.0)) : 0;
    const size_t ncells = (size_t)nc * nc * nc;
    NBX_HIP(c, hipMalloc((void**)&nl.d_nbr, sizeof(int) * (size_t)nl.cap * (size_t)n));
    NBX_HIP(c, hipMalloc((void**)&nl.d_nnb, sizeof(int) * (size_t)n));
    NBX_HIP(c, hipMalloc((void**)&nl.d_cell_count, sizeof(int) * std::max<size_t>(ncells, 1)));
    NBX_HIP(c, hipMalloc((void**)&nl.d_cell_atoms,
                         sizeof(int) * std::max<size_t>(ncells * (size_t)nl.ccap, 1)));
    NBX_HIP(c, hipMalloc((void**)&nl.d_xref, sizeof(double) * 3 * (size_t)n));
    NBX_HIP(c, hipMalloc((void**)&nl.d_flags, sizeof(int) * 8));
    NBX_HIP(c, hipMemsetAsync(nl.d_flags, 0, sizeof(int) * 8, c->stream));
    NBX_HIP(c, hipMemsetAsync(nl.d_nnb, 0, sizeof(int) * (size_t)n, c->stream));
    nl.cap_bytes = sizeof(int) * (size_t)nl.cap * (size_t)n;
    nl.ready = true;
    return NBX_OK;
}

int pos_kind(const mx_plan& p) {
    if (p.kind == PLAN_GRAVITY) return 2;
    return p.nb.mode_list ? 1 : 0;
}

mx_kick_params kick_params(const nbx_ctx* c, int do_kick, double dt, uint64_t seed,
                           int64_t step) {
    mx_kick_params kp{};
    kp.do_kick = do_kick;
    kp.therm = c->therm;
    kp.dt = dt;
    kp.kb = c->kb;
    kp.T = c->therm_T;
    kp.seed = seed;
    kp.step = step;
    kp.m0 = c->n > 0 ? c->h_ms[0] : 1.0;
    kp.d_T = c->d_scal;  // [0]
    if (c->therm == NBX_THERM_LANGEVIN) {
        kp.c1 = std::exp(-c->therm_p * (0.5 * dt));
        kp.c2 = std::sqrt(1.0 - kp.c1 * kp.c1);
    } else if (c->therm == NBX_THERM_ANDERSEN) {
        kp.prob = c->therm_p * dt;
    } else if (c->therm == NBX_THERM_BERENDSEN) {
        kp.ber_gamma = 0.5 / c->therm_p;
    }
    return kp;
}

int64_t constraints(const nbx_ctx* c) { return c->sys == NBX_SYS_WATER ? 2 * (c->n / 3) : 0; }

// force evaluation at the current positions, optionally finishing the step
void launch_force(nbx_ctx* c, const mx_plan& p, const mx_kick_params& kp) {
    hipStream_t s = c->stream;
    if (p.kind == PLAN_GRAVITY) {
        mx_launch_gravity(s, c, p.G, kp, 0);
        return;
    }
    if (p.nb.mode_list) mx_launch_rebuild(s, c, p.nb.water);
    mx_nb_launch nb = p.nb;
    nb.lanes = c->nl.lanes;
    mx_launch_nonbonded(s, nb, c, p.pp, kp, p.nb.mode_list ? c->nl.d_flags : nullptr);
    if (p.kind == PLAN_WATER) mx_launch_water_molecules(s, c, kp);
}

struct backup4 {
    double* p = nullptr;
    size_t n3 = 0;
    ~backup4() { if (p) (void)hipFree(p); }
};

}  // namespace

int mx_supported(nbx_ctx* c, std::string* why) {
    mx_plan p;
    return make_plan(c, p, why) ? 1 : 0;
}

static int prepare(nbx_ctx* c, mx_plan& p, double grow) {
    std::string why;
    if (!make_plan(c, p, &why)) return nbx_fail(c, NBX_ERR_UNSUPPORTED, why);
    const int64_t n = c->n;
    if (!c->d_pos) {
        NBX_HIP(c, hipMalloc(&c->d_pos, sizeof(double) * 4 * (size_t)n));
    }
    if (p.kind == PLAN_GRAVITY) {
        const int chunks = mx_gravity_chunks(n);
        if (!c->d_part || c->grav_chunks != chunks) {
            if (c->d_part) (void)hipFree(c->d_part);
            c->d_part = nullptr;
            NBX_HIP(c, hipMalloc((void**)&c->d_part, sizeof(double) * 3 * (size_t)n * chunks));
            c->grav_chunks = chunks;
        }
        free_nl(c->nl);
        NBX_HIP(c, hipMalloc((void**)&c->nl.d_flags, sizeof(int) * 8));
        NBX_HIP(c, hipMemsetAsync(c->nl.d_flags, 0, sizeof(int) * 8, c->stream));
        c->nl.ready = true;
        return NBX_OK;
    }
    if (!p.nb.mode_list) {
        free_nl(c->nl);
        NBX_HIP(c, hipMalloc((void**)&c->nl.d_flags, sizeof(int) * 8));
        NBX_HIP(c, hipMemsetAsync(c->nl.d_flags, 0, sizeof(int) * 8, c->stream));
        c->nl.ready = true;
        return NBX_OK;
    }
    if (c->nl.lanes <= 0) c->nl.lanes = 8;
    return alloc_lists(c, p, grow);
}

// d_x -> d_pos, list rebuild forced, a = f(x, v) (no step taken)
static int prologue(nbx_ctx* c, mx_plan& p) {
    hipStream_t s = c->stream;
    mx_launch_positions(s, c, pos_kind(p));
    if (c->therm == NBX_THERM_BERENDSEN)
        mx_launch_temperature(s, c, c->d_v, constraints(c), c->d_scal);
    // the Berendsen term of the first evaluation uses v itself (vh := v)
    NBX_HIP(c, hipMemcpyAsync(c->d_vh, c->d_v, sizeof(double) * 3 * (size_t)c->n,
                              hipMemcpyDeviceToDevice, s));
    launch_force(c, p, kick_params(c, 0, 0.0, 0, 0));
    NBX_HIP(c, hipGetLastError());
    c->accel_valid = true;
    return NBX_OK;
}

static int read_flags(nbx_ctx* c, int* h) {
    NBX_HIP(c, hipMemcpyAsync(h, c->nl.d_flags, sizeof(int) * 8, hipMemcpyDeviceToHost, c->stream));
    NBX_HIP(c, hipStreamSynchronize(c->stream));
    return NBX_OK;
}

static int grow_lists(nbx_ctx* c, mx_plan& p, double& grow) {
    grow *= 1.6;
    if (grow > 40.0)
        return nbx_fail(c, NBX_ERR_CAPACITY, "neighbour/cell list overflow persists after growth");
    int rc = prepare(c, p, grow);
    if (rc) return rc;
    mx_launch_positions(c->stream, c, pos_kind(p));  // also requests a rebuild
    NBX_HIP(c, hipGetLastError());
    return NBX_OK;
}

int mx_vv(nbx_ctx* c, double dt, int64_t nsteps, uint64_t seed, int64_t save_every, double* uf,
          double* vf, int64_t* nframes) {
    mx_plan p;
    double grow = 1.0;
    int rc;
    if (!c->nl.ready || !c->accel_valid) {
        if ((rc = prepare(c, p, grow))) return rc;
        if ((rc = prologue(c, p))) return rc;
        int h[8];
        if ((rc = read_flags(c, h))) return rc;
        while (h[1]) {
            if ((rc = grow_lists(c, p, grow))) return rc;
            if ((rc = prologue(c, p))) return rc;
            if ((rc = read_flags(c, h))) return rc;
        }
    } else {
        std::string why;
        if (!make_plan(c, p, &why)) return nbx_fail(c, NBX_ERR_UNSUPPORTED, why);
    }
    hipStream_t s = c->stream;
    const int64_t n = c->n;
    const size_t n3 = 3 * (size_t)n;
    backup4 bk;
    NBX_HIP(c, hipMalloc((void**)&bk.p, sizeof(double) * 4 * n3));
    const int pk = pos_kind(p);
    const double half_skin2 = 0.25 * c->nl.skin * c->nl.skin;
    const int64_t chunk_max = 500;
    int64_t done = 0;
    while (done < nsteps) {
        int64_t chunk = std::min<int64_t>(chunk_max, nsteps - done);
        if (save_every > 0) {
            const int64_t to_save = save_every - (done % save_every);
            chunk = std::min(chunk, to_save);
        }
        // device-side backup of the state for an overflow replay
        const double* src[4] = {c->d_x, c->d_v, c->d_a, c->d_vh};
        for (int k = 0; k < 4; ++k)
            NBX_HIP(c, hipMemcpyAsync(bk.p + k * n3, src[k], sizeof(double) * n3,
                                      hipMemcpyDeviceToDevice, s));
        const int64_t step0 = c->steps_taken;
        for (int attempt = 0;; ++attempt) {
            for (int64_t k = 1; k <= chunk; ++k) {
                const int64_t step = step0 + k;
                const mx_kick_params kp = kick_params(c, 1, dt, seed, step);
                mx_launch_drift(s, c, dt, c->therm, kp.c1, kp.c2, c->therm_T, c->kb, seed, step, pk,
                                half_skin2);
                if (c->therm == NBX_THERM_BERENDSEN)
                    mx_launch_temperature(s, c, c->d_vh, constraints(c), c->d_scal);
                launch_force(c, p, kp);
            }
            NBX_HIP(c, hipGetLastError());
            int h[8];
            if ((rc = read_flags(c, h))) return rc;
            if (!h[1]) break;
            // overflow: restore and replay with larger lists
            double* dst[4] = {c->d_x, c->d_v, c->d_a, c->d_vh};
            for (int k = 0; k < 4; ++k)
                NBX_HIP(c, hipMemcpyAsync(dst[k], bk.p + k * n3, sizeof(double) * n3,
                                          hipMemcpyDeviceToDevice, s));
            if ((rc = grow_lists(c, p, grow))) return rc;
            if (attempt > 8)
                return nbx_fail(c, NBX_ERR_CAPACITY, "list overflow replay did not converge");
        }
        done += chunk;
        c->steps_taken = step0 + chunk;
        if (save_every > 0 && uf && done % save_every == 0) {
            ref64_soa_to_aos(s, c->d_x, c->d_tmp, n);
            NBX_HIP(c, hipMemcpyAsync(uf + *nframes * n3, c->d_tmp, sizeof(double) * n3,
                                      hipMemcpyDeviceToHost, s));
            NBX_HIP(c, hipStreamSynchronize(s));
            ref64_soa_to_aos(s, c->d_v, c->d_tmp, n);
            NBX_HIP(c, hipMemcpyAsync(vf + *nframes * n3, c->d_tmp, sizeof(double) * n3,
                                      hipMemcpyDeviceToHost, s));
            NBX_HIP(c, hipStreamSynchronize(s));
            ++*nframes;
        }
    }
    return NBX_OK;
}

int mx_time_force(nbx_ctx* c, int iters, double* avg_ms, const char** name) {
    mx_plan p;
    int rc;
    if (!c->nl.ready || !c->accel_valid) {
        if ((rc = prepare(c, p, 1.0))) return rc;
        if ((rc = prologue(c, p))) return rc;
    } else {
        std::string why;
        if (!make_plan(c, p, &why)) return nbx_fail(c, NBX_ERR_UNSUPPORTED, why);
    }
    hipStream_t s = c->stream;
    const mx_kick_params kp = kick_params(c, 0, 0.0, 0, 0);
    mx_nb_launch nb = p.nb;
    nb.lanes = c->nl.lanes;
    auto launch = [&]() {
        if (p.kind == PLAN_GRAVITY) mx_launch_gravity(s, c, p.G, kp, 1);
        else mx_launch_nonbonded(s, nb, c, p.pp, kp, nullptr);
    };
    launch();  // warm
    hipEvent_t e0, e1;
    NBX_HIP(c, hipEventCreate(&e0));
    NBX_HIP(c, hipEventCreate(&e1));
    NBX_HIP(c, hipEventRecord(e0, s));
    for (int k = 0; k < iters; ++k) launch();
    NBX_HIP(c, hipEventRecord(e1, s));
    NBX_HIP(c, hipEventSynchronize(e1));
    float ms = 0.f;
    NBX_HIP(c, hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_ms = (double)ms / iters;
    if (name) {
        if (p.kind == PLAN_GRAVITY) *name = "k_grav_tiled";
        else if (p.nb.mode_list) *name = "k_nb_list";
        else *name = "k_nb_allpairs";
    }
    // the timed launches recomputed a = f(x) in place (force only, no kick): state unchanged
    return NBX_OK;
}

int mx_count_pairs(nbx_ctx* c, int64_t* pairs) {
    mx_plan p;
    std::string why;
    if (!make_plan(c, p, &why)) return nbx_fail(c, NBX_ERR_UNSUPPORTED, why);
    if (p.kind == PLAN_GRAVITY) {
        *pairs = c->n * (c->n - 1);
        return NBX_OK;
    }
    if (!p.nb.mode_list)
        return nbx_fail(c, NBX_ERR_UNSUPPORTED, "pair count needs a cubic box with finite cutoffs");
    const int prec = c->prec;
    int rc;
    if (!c->nl.ready || !c->accel_valid || prec != NBX_PREC_MIXED) {
        if ((rc = prepare(c, p, 1.0))) return rc;
        c->prec = NBX_PREC_MIXED;
        rc = prologue(c, p);
        c->prec = prec;
        if (prec != NBX_PREC_MIXED) c->accel_valid = false;  // the FP64 path recomputes a
        if (rc) return rc;
    }
    unsigned long long* d = reinterpret_cast<unsigned long long*>(c->d_scal + 12);
    NBX_HIP(c, hipMemsetAsync(d, 0, sizeof(unsigned long long), c->stream));
    mx_nb_launch nb = p.nb;
    nb.lanes = c->nl.lanes;
    mx_launch_count_pairs(c->stream, c, nb, p.pp, d);
    unsigned long long h = 0;
    NBX_HIP(c, hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    NBX_HIP(c, hipStreamSynchronize(c->stream));
    *pairs = (int64_t)h;
    return NBX_OK;
}

void mx_release(nbx_ctx* c) {
    free_nl(c->nl);
    if (c->d_pos) (void)hipFree(c->d_pos);
    if (c->d_part) (void)hipFree(c->d_part);
    c->d_pos = nullptr;
    c->d_part = nullptr;
}
