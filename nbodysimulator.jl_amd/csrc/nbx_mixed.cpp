// nbx_mixed.cpp -- host orchestration of the mixed-precision path.
//
// Plans: Verlet list (cubic box, finite cutoffs), all pairs (any other boundary or an
// infinite cutoff), LDS-tiled gravity, SPC/Fw water (non-bonded + per-molecule kernel).
// A run is a sequence of chunks (<= 500 steps, or up to the next saved frame); each
// chunk starts from x in X[0] with a drift kernel, then enqueues per step
//     build(k) [device-gated] ; step(k)      (water: step = non-bonded + molecule)
// either through a captured hipGraph of graph_steps steps (an even count, so buffer
// parities repeat) replayed with a device-side step base, or directly for the rest.
// A chunk that overflows a cell or neighbour list is replayed from a device backup with
// larger lists; the counter-based RNG makes the replay identical.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <utility>
#include <vector>
#include <cmath>
#include <cstdlib>
#include <string>

#include "nbx_mixed.h"

int mx_gravity_chunks(const nbx_ctx* c);

namespace {

enum { PLAN_NONBONDED = 1, PLAN_WATER = 2, PLAN_GRAVITY = 3, PLAN_MAGNETIC = 4 };

struct mx_plan {
    int kind = 0;
    mx_nb_launch nb{};
    mx_pair_params pp{};
    double G = 0.0;
    double mu = 0.0;      // mu / 4 pi (magnetic dipoles)
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
        int ngrav = 0, nmag = 0;
        for (const nbx_term& t : c->terms) {
            if (t.kind == NBX_TERM_GRAV) { ++ngrav; p.G = t.p[0]; }
            else if (t.kind == NBX_TERM_MAG) { ++nmag; p.mu = t.p[0]; }
            else if (t.kind == NBX_TERM_LJ) {
                if (lj) { if (why) *why = "mixed path: one Lennard-Jones term at most"; return false; }
                lj = &t;
            } else {
                if (el) { if (why) *why = "mixed path: one electrostatic term at most"; return false; }
                el = &t;
            }
        }
        if (nmag) {
            if (nmag > 1 || ngrav || lj || el) {
                if (why) *why = "mixed path: magnetic dipoles cannot be combined with other potentials";
                return false;
            }
            p.kind = PLAN_MAGNETIC;
            return true;
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
    nl.d_nbr = nullptr;
    nl.d_nnb = nl.d_cell_count = nl.d_cell_atoms = nl.d_flags = nullptr;
    nl.d_xref = nullptr;
    nl.ready = false;
}

int round_up(int x, int m) { return ((x + m - 1) / m) * m; }

int alloc_flags(nbx_ctx* c) {
    NBX_HIP(c, hipMalloc((void**)&c->nl.d_flags, sizeof(int) * 8));
    NBX_HIP(c, hipMemsetAsync(c->nl.d_flags, 0, sizeof(int) * 8, c->stream));
    return NBX_OK;
}

// size and allocate the lists for the current configuration (cap/ccap grown by `grow`)
int alloc_lists(nbx_ctx* c, const mx_plan& p, double grow) {
    nbx_nlist& nl = c->nl;
    const int lanes = nl.lanes;
    free_nl(nl);
    nl.lanes = lanes;
    const int64_t n = c->n;
    // the list kernels index one list row of n * lanes slots in 32 bits
    if (n * (int64_t)std::max(lanes, 1) >= (int64_t)1 << 32)
        return nbx_fail(c, NBX_ERR_CAPACITY, "n x lanes per atom must stay below 2^32");
    const double L = c->bc.b[0];
    nl.cutoff = p.cutoff;
    nl.skin = c->skin_set ? c->skin : 0.12 * p.cutoff;
    nl.rlist = p.cutoff + nl.skin;
    const double rho = (double)n / (L * L * L);
    const double est = 4.0 / 3.0 * M_PI * nl.rlist * nl.rlist * nl.rlist * rho;
    int cap = (int)std::ceil(grow * (1.6 * est + 24.0));
    cap = (int)std::min<int64_t>(cap, std::max<int64_t>(n, 1));
    nl.cap = round_up(std::max(cap, 1), lanes);
    int nc = (int)std::floor(L / nl.rlist);
    int mode = NBX_BINS_STEP;
    if (nc < 3) {
        nc = 0;
    } else if (nc == 3) {
        mode = NBX_BINS_FIXED;
    } else {
        // half-size cells binned by each build: with every atom moved at most 3/4 skin
        // since (the drift flags more), 2 cells reach list radius + 1.5 skin.  The build
        // gives every cell a workgroup, so this pays only with well-filled cells (water:
        // ~57 atoms per half cell; a Lennard-Jones liquid: ~4, which keeps per-step bins)
        const int nh = (int)std::floor(L / (0.5 * (nl.rlist + 1.5 * nl.skin)));
        if (nh >= 5 && (double)n / ((double)nh * nh * nh) >= 16.0) {
            nc = nh;
            mode = NBX_BINS_BUILD;
        }
    }
    if (nc > 128) nc = 128;
    nl.ncell = nc;
    nl.bin_mode = mode;
    const double per_cell = nc ? (double)n / ((double)nc * nc * nc) : 0.0;
    nl.ccap = nc ? (int)std::ceil(grow * (2.0 * per_cell + 16.0)) : 0;
    if (nl.ccap > 1024)
        return nbx_fail(c, NBX_ERR_CAPACITY, "cell occupancy above 1024 atoms (box too dense)");
    const size_t ncells = (size_t)nc * nc * nc;
    nl.nbr_wide = n > 65536 ? 1 : 0;
    NBX_HIP(c, hipMalloc(&nl.d_nbr, (nl.nbr_wide ? sizeof(int) : sizeof(uint16_t)) *
                                        (size_t)nl.cap * (size_t)n));
    NBX_HIP(c, hipMalloc((void**)&nl.d_nnb, sizeof(int) * (size_t)n));
    NBX_HIP(c, hipMalloc((void**)&nl.d_cell_count, sizeof(int) * 3 * std::max<size_t>(ncells, 1)));
    NBX_HIP(c, hipMalloc((void**)&nl.d_cell_atoms,
                         sizeof(int) * 3 * std::max<size_t>(ncells * (size_t)nl.ccap, 1)));
    NBX_HIP(c, hipMalloc((void**)&nl.d_xref, sizeof(double) * 3 * (size_t)n));
    int rc = alloc_flags(c);
    if (rc) return rc;
    NBX_HIP(c, hipMemsetAsync(nl.d_nnb, 0, sizeof(int) * (size_t)n, c->stream));
    NBX_HIP(c, hipMemsetAsync(nl.d_cell_count, 0, sizeof(int) * 3 * std::max<size_t>(ncells, 1),
                              c->stream));
    nl.cap_bytes = sizeof(int) * (size_t)nl.cap * (size_t)n;
    nl.ready = true;
    return NBX_OK;
}

bool all_pairs_tiled(const mx_plan& p) { return p.kind == PLAN_GRAVITY || p.kind == PLAN_MAGNETIC; }

// factor on the mass carried in the FP32 positions: gravity keeps m / m_max there (the
// magnetic kernel divides by the plain mass)
double mass_wscale(const nbx_ctx* c, const mx_plan& p) {
    return p.kind == PLAN_GRAVITY ? c->mass_wscale : 1.0;
}

int pos_kind(const mx_plan& p) {
    if (all_pairs_tiled(p)) return 2;
    return p.nb.mode_list ? 1 : 0;
}

mx_kick_params kick_params(const nbx_ctx* c, int do_kick, double dt, uint64_t seed) {
    mx_kick_params kp{};
    kp.do_kick = do_kick;
    kp.therm = c->therm;
    kp.dt = dt;
    kp.kb = c->kb;
    kp.T = c->therm_T;
    kp.seed = seed;
    kp.m0 = c->n > 0 ? c->h_ms[0] : 1.0;
    kp.d_T = c->d_scal;  // [0]
    if (c->em) {
        // sigma = sqrt(2 gamma kb T / m_1) (src/nbody_to_ode.jl:592)
        kp.em = 1;
        kp.em_gamma = c->therm_p;
        kp.em_sigma = std::sqrt(2 * c->therm_p * c->kb * c->therm_T / kp.m0);
        kp.em_sqrt_dt = std::sqrt(dt);
    } else if (c->therm == NBX_THERM_LANGEVIN) {
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

double* xbuf(nbx_ctx* c, int par) { return par ? c->d_x2 : c->d_x; }
void* pbuf(nbx_ctx* c, int par) { return par ? c->d_pos2 : c->d_pos; }
// the rebuild request: set by the drift, consumed (and cleared) by the next build
int* req_flag(nbx_ctx* c) { return c->nl.d_flags; }
// bins of the per-step and fixed modes (set 0; the rotating mode uses all three sets)
int* bin_count(nbx_ctx* c) { return c->nl.d_cell_count; }
int* bin_atoms(nbx_ctx* c) { return c->nl.d_cell_atoms; }

// steps between list-build opportunities
int build_every(const nbx_ctx* c) { return std::max(1, c->build_every); }

mx_nh_params nh_params(const nbx_ctx* c, double dt, int advance) {
    mx_nh_params nh{};
    nh.nh = c->therm == NBX_THERM_NOSEHOOVER ? 1 : 0;
    nh.advance = advance;
    nh.dt = dt;
    nh.T0 = c->therm_T;
    nh.tau = c->therm_p;
    nh.ndf = 3 * c->n - (c->sys == NBX_SYS_WATER ? 2 * (c->n / 3) : 0);
    return nh;
}

// the temperature partials of the drift that writes X[par]: one buffer per parity (an int
// count, then one partial per block; n/2 + 4,096 doubles each, more than any step kernel's
// blocks), so a step can add the partials of the drift before it while its own drift
// writes the other buffer
double* red_buf(const nbx_ctx* c, int par) {
    return c->d_red + (size_t)par * (size_t)(c->n / 2 + 4096);
}

// water with Berendsen: the molecule kernel adds the partials itself (kp.T_parts), and no
// temperature finish runs between the steps
bool folds_temperature(const nbx_ctx* c, const mx_plan& p) {
    return p.kind == PLAN_WATER && c->therm == NBX_THERM_BERENDSEN && !c->em;
}

// buffers of a launch that reads x in X[par] and (when drifting) writes X[par ^ 1];
// next_opp: the next step has a build opportunity
mx_io make_io(nbx_ctx* c, const mx_plan& p, int par, int do_drift, double dt = 0.0,
              int next_opp = 1) {
    mx_io io{};
    io.x_in = xbuf(c, par);
    io.x_out = xbuf(c, par ^ 1);
    io.pos_kind = pos_kind(p);
    io.wscale = mass_wscale(c, p);
    io.pos_out = pbuf(c, par ^ 1);
    io.a = c->d_a;
    io.v = c->d_v;
    io.vh = c->d_vh;
    io.ms = c->d_ms;
    io.qs = c->has_q ? c->d_qs : nullptr;
    io.do_drift = do_drift;
    io.flags = c->nl.d_flags;
    const bool list = p.kind != PLAN_GRAVITY && p.nb.mode_list;
    if (list) {
        // with a build opportunity every M steps a rebuild is requested at
        // (skin/2)(1 - (M-1)/(2M)), so that the M-1 steps to the next opportunity rarely take
        // an atom past skin/2; one that does marks the chunk stale (replayed with M = 1)
        const int M = build_every(c);
        const double trig = 0.5 * c->nl.skin * (1.0 - (double)(M - 1) / (2.0 * M));
        io.xref = c->nl.d_xref;
        io.half_skin2 = trig * trig;
        io.stale2 = 0.25 * c->nl.skin * c->nl.skin;
        io.next_opp = next_opp;
        io.guard2 = c->nl.bin_mode == NBX_BINS_BUILD ? 0.5625 * c->nl.skin * c->nl.skin
                                                      : INFINITY;
        io.req_out = req_flag(c);
        io.L = c->bc.b[0];
        io.invL = 1.0 / io.L;
    } else {
        io.next_opp = 1;
    }
    io.need_T = ((c->therm == NBX_THERM_BERENDSEN || c->therm == NBX_THERM_NOSEHOOVER) && do_drift)
                    ? 1 : 0;
    double* rb = red_buf(c, par ^ 1);
    io.red_count = reinterpret_cast<int*>(rb);
    io.red_part = rb + 1;
    io.T_out = c->d_scal;
    io.T_denom = c->kb * (double)(3 * c->n - constraints(c));
    io.nh = nh_params(c, dt, 1);   // with need_T: the Nose-Hoover update of this drift
    return io;
}


// after a kernel that wrote x into X[par_out]: bins for a requested list build (the
// thermostat temperature of vh, and the Nose-Hoover update, are finished by the last
// block of that kernel: mx_epilogue)
void after_drift(nbx_ctx* c, const mx_plan& p, int par_out) {
    hipStream_t s = c->stream;
    if (p.kind != PLAN_GRAVITY && p.nb.mode_list && c->nl.ncell >= 3 &&
        c->nl.bin_mode == NBX_BINS_STEP)
        mx_launch_bin(s, c, pbuf(c, par_out), req_flag(c), bin_count(c), bin_atoms(c));
}

// the next pair of force-kernel timing events, when timing is on and the pool has room
// (the pool holds two per step of the longest chunk)
void timing_events(nbx_ctx* c, hipEvent_t* e0, hipEvent_t* e1) {
    *e0 = *e1 = nullptr;
    if (!c->kt_on || c->kt_used + 2 > c->kt_ev.size()) return;
    *e0 = c->kt_ev[c->kt_used++];
    *e1 = c->kt_ev[c->kt_used++];
}

// force evaluation at x in X[par] (list: P[par]); with kp.do_kick the step's end, with
// do_drift the start of the next step (binned for a requested build when next_opp)
void launch_step(nbx_ctx* c, const mx_plan& p, const mx_kick_params& kp, int par, int do_drift,
                 bool timed = false, int next_opp = 1) {
    hipStream_t s = c->stream;
    const mx_io io = make_io(c, p, par, do_drift, kp.dt, next_opp);
    // the step's kick takes the temperature of the drift before it from its partials
    const bool fold = kp.do_kick && folds_temperature(c, p);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (timed) timing_events(c, &e0, &e1);
    if (p.kind == PLAN_GRAVITY) {
        mx_launch_gravity(s, c, pbuf(c, par), p.G, kp, io, 0, e0, e1);
    } else if (p.kind == PLAN_MAGNETIC) {
        mx_launch_magnetic(s, c, pbuf(c, par), p.mu, kp, io, 0);
    } else {
        mx_nb_launch nb = p.nb;
        nb.lanes = c->nl.lanes;
        if (e0) (void)hipEventRecord(e0, s);
        mx_launch_nonbonded(s, nb, c, pbuf(c, par), p.pp, kp, io);
        if (e1) (void)hipEventRecord(e1, s);
        if (p.kind == PLAN_WATER) {
            mx_kick_params kw = kp;
            if (fold) {
                kw.T_parts = red_buf(c, par);
                kw.T_denom = io.T_denom;
            }
            mx_launch_water_molecules(s, c, kw, io);
        }
    }
    if (io.need_T && !fold) mx_launch_temperature_finish(s, io);
    if (do_drift && next_opp) after_drift(c, p, par ^ 1);
}

void launch_build(nbx_ctx* c, const mx_plan& p, int par) {
    if (p.kind == PLAN_GRAVITY || !p.nb.mode_list) return;
    // 3 cells per side: the 27 neighbour cells are the whole box, so any partition of
    // the atoms into home cells finds every neighbour -- the bins made at set-up serve
    // every build and no step bins
    // per-step bins are cleared by the build that used them, ready for the next binning
    const int clear = c->nl.bin_mode == NBX_BINS_STEP && c->nl.ncell >= 3 ? 1 : 0;
    mx_launch_build(c->stream, c, p.nb.water, pbuf(c, par), xbuf(c, par), req_flag(c),
                    bin_count(c), bin_atoms(c), clear);
}

struct backup4 {
    double* p = nullptr;
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
    // double4 per atom, then the fixed-point copy (16 bytes per atom)
    const size_t pb = 48 * (size_t)std::max<int64_t>(n, 1);
    if (!c->d_pos) NBX_HIP(c, hipMalloc(&c->d_pos, pb));
    if (!c->d_pos2) NBX_HIP(c, hipMalloc(&c->d_pos2, pb));
    if (!c->d_x2)
        NBX_HIP(c, hipMalloc((void**)&c->d_x2, sizeof(double) * 3 * (size_t)std::max<int64_t>(n, 1)));
    if (!c->d_step) NBX_HIP(c, hipMalloc((void**)&c->d_step, sizeof(int64_t)));
    if (all_pairs_tiled(p)) {
        const int chunks = p.kind == PLAN_GRAVITY ? mx_gravity_chunks(c) : 1;
        if (!c->d_part || c->grav_chunks != chunks) {
            if (c->d_part) (void)hipFree(c->d_part);
            c->d_part = nullptr;
            NBX_HIP(c, hipMalloc((void**)&c->d_part, sizeof(double) * 3 * (size_t)n * chunks));
            c->grav_chunks = chunks;
        }
    }
    if (p.kind == PLAN_GRAVITY || !p.nb.mode_list) {
        free_nl(c->nl);
        int rc = alloc_flags(c);
        if (rc) return rc;
        c->nl.ready = true;
        return NBX_OK;
    }
    if (c->nl.lanes <= 0) c->nl.lanes = 8;
    return alloc_lists(c, p, grow);
}

// wrapped copy of X[0] into P[0], binned into B[0], list build requested
static void refresh_positions(nbx_ctx* c, const mx_plan& p, bool bins) {
    const int pk = pos_kind(p);
    if (pk == 0) return;
    const bool list = pk == 1 && c->nl.ncell >= 3 && bins;
    if (list) {   // fresh bins in set 0, the other sets empty, rotation at set 0
        const size_t ncells = (size_t)c->nl.ncell * c->nl.ncell * c->nl.ncell;
        mx_launch_fill_int(c->stream, c->nl.d_cell_count, (int64_t)(3 * ncells), 0);
        mx_launch_fill_int(c->stream, c->nl.d_flags + NBX_FLAG_BINSET, 2, 0);
    }
    mx_launch_positions(c->stream, c, c->d_x, c->d_pos, pk, mass_wscale(c, p),
                        list ? bin_count(c) : nullptr,
                        list ? bin_atoms(c) : nullptr,
                        (pk == 1 && bins) ? req_flag(c) : nullptr);
}

// a = f(x, v) at the current state (no step taken)
static int prologue(nbx_ctx* c, mx_plan& p) {
    hipStream_t s = c->stream;
    refresh_positions(c, p, true);
    if (c->therm == NBX_THERM_BERENDSEN || c->therm == NBX_THERM_NOSEHOOVER)
        mx_launch_temperature(s, c, c->d_v, constraints(c), c->d_scal);
    if (c->therm == NBX_THERM_NOSEHOOVER) mx_launch_nh_rate(s, c->d_scal, nh_params(c, 0.0, 0));
    // the Berendsen term of the first evaluation uses v itself (vh := v)
    NBX_HIP(c, hipMemcpyAsync(c->d_vh, c->d_v, sizeof(double) * 3 * (size_t)c->n,
                              hipMemcpyDeviceToDevice, s));
    launch_build(c, p, 0);
    launch_step(c, p, kick_params(c, 0, 0.0, 0), 0, 0);
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
    return prepare(c, p, grow);
}

static int rebuild_list(nbx_ctx* c, mx_plan& p);

// Lists and accelerations ready for the current state.  Outside a run (run = false) the
// state is in the user's order, so a list built for a run's sorted order is rebuilt for
// it (the accelerations were returned in the user's order with the state).
static int ensure_ready(nbx_ctx* c, mx_plan& p, bool run = false) {
    int rc;
    double grow = 1.0;
    if (!run && c->list_sorted) {
        c->list_sorted = false;
        if (c->nl.ready && c->accel_valid) {
            std::string why;
            if (!make_plan(c, p, &why)) return nbx_fail(c, NBX_ERR_UNSUPPORTED, why);
            return rebuild_list(c, p);
        }
    }
    if (!c->nl.ready || !c->accel_valid) {
        c->list_sorted = false;   // the list made here is for the state's current order
        if ((rc = prepare(c, p, grow))) return rc;
        if ((rc = prologue(c, p))) return rc;
        int h[8];
        if ((rc = read_flags(c, h))) return rc;
        while (h[1]) {
            if ((rc = grow_lists(c, p, grow))) return rc;
            if ((rc = prologue(c, p))) return rc;
            if ((rc = read_flags(c, h))) return rc;
        }
        return NBX_OK;
    }
    std::string why;
    if (!make_plan(c, p, &why)) return nbx_fail(c, NBX_ERR_UNSUPPORTED, why);
    return NBX_OK;
}

// ---- spatial order of the state during a run ----------------------------------------
// Runs of mixed Lennard-Jones-type Verlet lists with per-step bins (no water molecules, no
// Euler-Maruyama steps) keep the state sorted by cell when sort_every > 0: the list force
// kernel's gathers then stay near each other in memory as the liquid mixes.
static bool sort_enabled(const nbx_ctx* c, const mx_plan& p) {
    return c->sort_every > 0 && c->n > 1 && p.kind == PLAN_NONBONDED && p.nb.mode_list &&
           !p.nb.water && c->nl.ncell >= 3 && c->nl.bin_mode == NBX_BINS_STEP && !c->em;
}

int mx_reset_order(nbx_ctx* c) {
    mx_launch_iota(c->stream, c->d_perm, c->n);
    mx_launch_iota(c->stream, c->d_iperm, c->n);
    NBX_HIP(c, hipGetLastError());
    NBX_HIP(c, hipStreamSynchronize(c->stream));
    c->perm_identity = true;
    c->list_sorted = false;
    c->last_sort_step = -1;
    return NBX_OK;
}

// the per-atom state reordered in place: new[s] = old[map[s]] (positions, velocities,
// accelerations, half-kicked velocities, list reference positions, masses, charges), in
// two launches through a scratch buffer
static int permute_state(nbx_ctx* c, const int* d_map) {
    const int64_t n = c->n;
    if (!c->d_permbuf)
        NBX_HIP(c, hipMalloc((void**)&c->d_permbuf, sizeof(double) * 17 * (size_t)std::max<int64_t>(n, 1)));
    mx_perm_arrays A{};
    double* a3[5] = {c->d_x, c->d_v, c->d_a, c->d_vh, c->nl.d_xref};
    for (double* a : a3)
        if (a) A.a3[A.n3++] = a;
    double* a1[2] = {c->d_ms, c->has_q ? c->d_qs : nullptr};
    for (double* a : a1)
        if (a) A.a1[A.n1++] = a;
    mx_launch_permute(c->stream, d_map, n, A, c->d_permbuf);
    NBX_HIP(c, hipGetLastError());
    return NBX_OK;
}

// a Verlet list for the current positions (fresh bins, gated build with the request set,
// grown lists on overflow); the accelerations stay as they are
static int rebuild_list(nbx_ctx* c, mx_plan& p) {
    int rc;
    double grow = 1.0;
    for (int attempt = 0;; ++attempt) {
        refresh_positions(c, p, true);
        launch_build(c, p, 0);
        NBX_HIP(c, hipGetLastError());
        int h[8];
        if ((rc = read_flags(c, h))) return rc;
        if (!h[1]) return NBX_OK;
        if (attempt > 8) return nbx_fail(c, NBX_ERR_CAPACITY, "list overflow persists");
        if ((rc = grow_lists(c, p, grow))) return rc;
    }
}

// a new sorted order from the current positions: slots by cell of the list grid
// (z-major), the atoms of a cell by user index, found on the device (one radix sort); the
// state is reordered and the list rebuilt.  A state whose key sequence steps back at most
// n / 16 times is already nearly in cell order (an FCC lattice in generation order: about
// n / 80) and is kept.  The state must be in the current sorted order.
static int resort(nbx_ctx* c, mx_plan& p) {
    hipStream_t s = c->stream;
    const int64_t n = c->n;
    const int g = c->nl.ncell;
    const size_t temp = mx_sort_temp_bytes(n, g);
    if (temp == 0) return nbx_fail(c, NBX_ERR_DEVICE, "sort scratch size");
    // the next order's two maps (8 n bytes), the descent count, then the sort's scratch at
    // a 256-byte boundary
    const size_t maps = (8 * (size_t)n + 4 + 255) & ~(size_t)255;
    if (c->sortbuf_bytes < maps + temp) {
        if (c->d_sortbuf) (void)hipFree(c->d_sortbuf);
        c->d_sortbuf = nullptr;
        c->sortbuf_bytes = 0;
        NBX_HIP(c, hipMalloc(&c->d_sortbuf, maps + temp));
        c->sortbuf_bytes = maps + temp;
    }
    int* perm_new = static_cast<int*>(c->d_sortbuf);
    int* iperm_new = perm_new + n;
    // keys and values in the 24 n bytes of d_tmp: keys [0, 8n), sorted keys [8n, 16n),
    // slots [16n, 20n); the sorted slots (the map) go to d_pmap
    char* tb = reinterpret_cast<char*>(c->d_tmp);
    auto* keys_in = reinterpret_cast<unsigned long long*>(tb);
    auto* keys_out = reinterpret_cast<unsigned long long*>(tb + 8 * (size_t)n);
    int* vals_in = reinterpret_cast<int*>(tb + 16 * (size_t)n);
    unsigned* d_desc = reinterpret_cast<unsigned*>(iperm_new + n);
    if (mx_launch_sort_keys(s, c->d_x, n, c->bc.b[0], g, c->d_perm, keys_in, vals_in, d_desc))
        return nbx_fail(c, NBX_ERR_DEVICE, "device sort keys failed");
    unsigned desc = 0;
    NBX_HIP(c, hipMemcpyAsync(&desc, d_desc, sizeof(desc), hipMemcpyDeviceToHost, s));
    NBX_HIP(c, hipStreamSynchronize(s));
    if ((int64_t)desc * 16 <= n) return NBX_OK;
    if (mx_launch_sort_apply(s, n, g, keys_in, keys_out, vals_in, c->d_pmap,
                             static_cast<char*>(c->d_sortbuf) + maps, temp, c->d_perm, perm_new,
                             iperm_new))
        return nbx_fail(c, NBX_ERR_DEVICE, "device sort of the state failed");
    int rc;
    if ((rc = permute_state(c, c->d_pmap))) return rc;
    // the maps keep their addresses (captured step graphs read them)
    NBX_HIP(c, hipMemcpyAsync(c->d_perm, perm_new, sizeof(int) * n, hipMemcpyDeviceToDevice, s));
    NBX_HIP(c, hipMemcpyAsync(c->d_iperm, iperm_new, sizeof(int) * n, hipMemcpyDeviceToDevice, s));
    c->perm_identity = false;
    ++c->sorts_applied;
    return rebuild_list(c, p);
}

// allow_sort = false: internal callers that alternate short runs with work on the state
// in the user's order (training-set collection, the learned stride's plain steps)
static int mx_vv_run(nbx_ctx* c, double dt, int64_t nsteps, uint64_t seed, int64_t save_every,
                     double* uf, double* vf, int64_t* nframes, bool allow_sort) {
    mx_plan p;
    int rc;
    if ((rc = ensure_ready(c, p, true))) return rc;
    // the run keeps the state in the sorted order and returns it in the user's order
    const bool sorting = allow_sort && sort_enabled(c, p);
    struct order_guard {
        nbx_ctx* c = nullptr;
        ~order_guard() {
            if (!c || c->perm_identity) return;
            (void)permute_state(c, c->d_iperm);
            (void)hipStreamSynchronize(c->stream);
            c->list_sorted = true;
        }
    } og;
    if (sorting) {
        if (!c->perm_identity) {
            if ((rc = permute_state(c, c->d_perm))) return rc;
            og.c = c;
            if (!c->list_sorted && (rc = rebuild_list(c, p))) return rc;
        } else {
            og.c = c;
        }
    } else if (!c->perm_identity || c->list_sorted) {
        // not sorted from here on (an internal caller, or the plan no longer sorts): the
        // state is in the user's order, the list was made for the sorted one
        if ((rc = mx_reset_order(c))) return rc;
        if ((rc = rebuild_list(c, p))) return rc;
    }
    double grow = 1.0;
    hipStream_t s = c->stream;
    const int64_t n = c->n;
    const size_t n3 = 3 * (size_t)n;
    backup4 bk;
    NBX_HIP(c, hipMalloc((void**)&bk.p, sizeof(double) * (4 * n3 + NBX_SCAL_COUNT)));
    const int64_t chunk_max = 500;
    while (c->kt_on && c->kt_ev.size() < 2 * (size_t)chunk_max) {
        hipEvent_t e = nullptr;
        NBX_HIP(c, hipEventCreate(&e));
        c->kt_ev.push_back(e);
    }
    c->kt_used = 0;
    // steps k = k0+1..k0+count of the chunk, absolute step *d_step + k - k0 (the device
    // step base is advanced by count afterwards).  Absolute steps 1, M+1, 2M+1, ... have a
    // build opportunity (a gated list build, and the gated binning after the previous
    // step's drift); the others launch only the force kernel.  `base` is the absolute step
    // of k = 0 (0 when capturing a graph, whose replays start on opportunities); the
    // drift of the chunk's last step is redone at the start of the next chunk, so it is
    // treated as preceding an opportunity (no stale flag from a discarded state).  Step k
    // reads the buffers of parity (k + par_off) & 1 (par_off: the parity of the chunk step
    // a captured graph starts after).
    auto enqueue_steps = [&](int64_t k0, int64_t count, int64_t base, int64_t last,
                             int par_off) {
        const int64_t M = build_every(c);
        for (int64_t k = k0 + 1; k <= k0 + count; ++k) {
            const int par = (int)((k + par_off) & 1);
            mx_kick_params kp = kick_params(c, 1, dt, seed);
            kp.step_base = c->d_step;
            kp.step_off = k - k0;
            kp.ids = sorting ? c->d_perm : nullptr;
            if ((base + k - 1) % M == 0) launch_build(c, p, par);
            launch_step(c, p, kp, par, 1, true, ((base + k) % M == 0 || k == last) ? 1 : 0);
        }
        mx_launch_step_advance(s, c->d_step, count, 0);
    };
    auto graph_for = [&](int gsteps, int par_off, hipGraphExec_t* out) -> int {
        nbx_graph_key key;
        key.valid = 1;
        key.version = c->version;
        key.dt = dt;
        key.seed = seed;
        key.nbr = c->nl.d_nbr;
        key.nnb = c->nl.d_nnb;
        key.cell_count = c->nl.d_cell_count;
        key.cell_atoms = c->nl.d_cell_atoms;
        key.xref = c->nl.d_xref;
        key.flags = c->nl.d_flags;
        key.pos = c->d_pos;
        key.part = c->d_part;
        key.chunks = c->grav_chunks;
        key.tpt = c->grav_tpt_set;
        key.em = c->em;
        key.cap = c->nl.cap;
        key.ccap = c->nl.ccap;
        key.ncell = c->nl.ncell;
        key.lanes = c->nl.lanes;
        key.steps = gsteps;
        key.build_every = build_every(c);
        key.par_off = par_off;
        key.ids = sorting ? c->d_perm : nullptr;
        if (c->gexec && key == c->gkey) { *out = c->gexec; return NBX_OK; }
        if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
        hipGraph_t g = nullptr;
        NBX_HIP(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        enqueue_steps(0, gsteps, 0, gsteps, par_off);
        NBX_HIP(c, hipStreamEndCapture(s, &g));
        hipError_t e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            c->gexec = nullptr;
            return nbx_fail(c, NBX_ERR_DEVICE, std::string("hipGraphInstantiate: ") +
                                                   hipGetErrorString(e));
        }
        c->gkey = key;
        *out = c->gexec;
        return NBX_OK;
    };
    NBX_HIP(c, hipGetLastError());
    int64_t done = 0;
    while (done < nsteps) {
        int64_t chunk = std::min<int64_t>(chunk_max, nsteps - done);
        if (save_every > 0) chunk = std::min(chunk, save_every - (done % save_every));
        if (sorting) {
            // re-sorted at absolute steps that are multiples of sort_every (the chunks end
            // there), so that split runs keep the order of an unsplit one
            const int64_t S = c->sort_every, t0 = c->steps_taken;
            if (t0 % S == 0 && c->last_sort_step != t0) {
                if ((rc = resort(c, p))) return rc;
                c->last_sort_step = t0;
            }
            chunk = std::min(chunk, S - t0 % S);
        }
        // device-side backup of the state for an overflow replay
        const double* src[4] = {c->d_x, c->d_v, c->d_a, c->d_vh};
        for (int k = 0; k < 4; ++k)
            NBX_HIP(c, hipMemcpyAsync(bk.p + k * n3, src[k], sizeof(double) * n3,
                                      hipMemcpyDeviceToDevice, s));
        NBX_HIP(c, hipMemcpyAsync(bk.p + 4 * n3, c->d_scal, sizeof(double) * NBX_SCAL_COUNT,
                                  hipMemcpyDeviceToDevice, s));
        const int64_t step0 = c->steps_taken;
        for (int attempt = 0;; ++attempt) {
            mx_launch_step_advance(s, c->d_step, step0, 1);
            // start of step 1 of the chunk: X[0] -> X[1], binned for a requested build when
            // step 1 has a build opportunity.  The drift after the previous chunk's last
            // step is redone here with the same arithmetic, so its rebuild request stands
            // (as do earlier ones still waiting for an opportunity); its bins are cleared
            if (p.kind != PLAN_GRAVITY && p.nb.mode_list && c->nl.ncell >= 3 &&
                c->nl.bin_mode == NBX_BINS_STEP) {
                const size_t ncells = (size_t)c->nl.ncell * c->nl.ncell * c->nl.ncell;
                mx_launch_fill_int(s, bin_count(c), (int64_t)ncells, 0);
            }
            {
                mx_kick_params kp = kick_params(c, 1, dt, seed);
                kp.step_base = c->d_step;
                kp.step_off = 0;
                kp.ids = sorting ? c->d_perm : nullptr;
                const int first_opp = (step0 % build_every(c)) == 0 ? 1 : 0;
                mx_io io0 = make_io(c, p, 0, 1, dt, first_opp);
                // a redone drift recomputes xi but leaves zeta as its first pass left it
                if (c->nh_advanced) io0.nh.advance = 0;
                mx_launch_drift(s, c, kp, io0);
                if (io0.need_T && !folds_temperature(c, p)) mx_launch_temperature_finish(s, io0);
                if (first_opp) after_drift(c, p, 1);
            }
            const int gs = c->graph_steps;
            const int64_t M = build_every(c);
            // eager steps up to the first build opportunity, so that every graph replay
            // starts on one (a graph replays steps 1..gs of its own numbering; M divides gs)
            int64_t k0 = std::min<int64_t>(chunk, (M - step0 % M) % M);
            if (k0 > 0) enqueue_steps(0, k0, step0, chunk, 0);
            if (gs > 0 && (gs % 2) == 0 && (gs % M) == 0 && chunk - k0 >= gs && !c->kt_on) {
                hipGraphExec_t ge = nullptr;
                if ((rc = graph_for(gs, (int)(k0 & 1), &ge))) return rc;
                for (; k0 + gs <= chunk; k0 += gs) NBX_HIP(c, hipGraphLaunch(ge, s));
            }
            if (k0 < chunk) enqueue_steps(k0, chunk - k0, step0, chunk, 0);
            NBX_HIP(c, hipGetLastError());
            if (chunk & 1)  // the state is in X[1]
                NBX_HIP(c, hipMemcpyAsync(c->d_x, c->d_x2, sizeof(double) * n3,
                                          hipMemcpyDeviceToDevice, s));
            int h[8];
            if ((rc = read_flags(c, h))) return rc;
            if (!h[1] && !h[7]) {
                // the chunk stands: add its force-kernel times (the stream has drained)
                for (size_t k = 0; k + 1 < c->kt_used; k += 2) {
                    float ms = 0.f;
                    NBX_HIP(c, hipEventElapsedTime(&ms, c->kt_ev[k], c->kt_ev[k + 1]));
                    c->kt_sum_ms += ms;
                    ++c->kt_launches;
                }
                c->kt_used = 0;
                // the chunk's last drift advanced zeta for the step after the state
                c->nh_advanced = true;
                break;
            }
            c->kt_used = 0;   // replayed: its timings are dropped
            // overflow: restore and replay with larger lists; a list that went stale between
            // two build opportunities: restore and replay with an opportunity at every step
            double* dst[4] = {c->d_x, c->d_v, c->d_a, c->d_vh};
            for (int k = 0; k < 4; ++k)
                NBX_HIP(c, hipMemcpyAsync(dst[k], bk.p + k * n3, sizeof(double) * n3,
                                          hipMemcpyDeviceToDevice, s));
            NBX_HIP(c, hipMemcpyAsync(c->d_scal, bk.p + 4 * n3, sizeof(double) * NBX_SCAL_COUNT,
                                      hipMemcpyDeviceToDevice, s));
            if (h[1]) {
                if ((rc = grow_lists(c, p, grow))) return rc;
            } else {
                c->build_every = 1;
                ++c->stale_replays;
                mx_launch_fill_int(s, c->nl.d_flags + 7, 1, 0);
            }
            if ((rc = prologue(c, p))) return rc;
            if (attempt > 8)
                return nbx_fail(c, NBX_ERR_CAPACITY, "list overflow replay did not converge");
        }
        done += chunk;
        c->steps_taken = step0 + chunk;
        if (save_every > 0 && uf && done % save_every == 0) {
            // frames in the user's order (the backup buffer is free between chunks)
            const double* fx = c->d_x;
            const double* fv = c->d_v;
            if (sorting && !c->perm_identity) {
                mx_launch_gather(s, c->d_x, bk.p, c->d_iperm, n, 3);
                mx_launch_gather(s, c->d_v, bk.p + n3, c->d_iperm, n, 3);
                fx = bk.p;
                fv = bk.p + n3;
            }
            ref64_soa_to_aos(s, fx, c->d_tmp, n);
            NBX_HIP(c, hipMemcpyAsync(uf + *nframes * n3, c->d_tmp, sizeof(double) * n3,
                                      hipMemcpyDeviceToHost, s));
            NBX_HIP(c, hipStreamSynchronize(s));
            ref64_soa_to_aos(s, fv, c->d_tmp, n);
            NBX_HIP(c, hipMemcpyAsync(vf + *nframes * n3, c->d_tmp, sizeof(double) * n3,
                                      hipMemcpyDeviceToHost, s));
            NBX_HIP(c, hipStreamSynchronize(s));
            ++*nframes;
        }
    }
    return NBX_OK;
}

int mx_vv(nbx_ctx* c, double dt, int64_t nsteps, uint64_t seed, int64_t save_every, double* uf,
          double* vf, int64_t* nframes) {
    return mx_vv_run(c, dt, nsteps, seed, save_every, uf, vf, nframes, true);
}

// the step's dominant (force) kernel
const char* mx_force_kernel_name(nbx_ctx* c) {
    mx_plan p;
    if (!make_plan(c, p, nullptr)) return "";
    if (p.kind == PLAN_GRAVITY) return "k_grav_tiled";
    if (p.kind == PLAN_MAGNETIC) return "k_mag_tiled";
    return p.nb.mode_list ? "k_nb_list" : "k_nb_allpairs";
}

int mx_time_force(nbx_ctx* c, int iters, double* avg_ms, const char** name) {
    mx_plan p;
    int rc;
    if ((rc = ensure_ready(c, p))) return rc;
    hipStream_t s = c->stream;
    refresh_positions(c, p, false);  // P[0] <- X[0] (the list stays valid for it)
    const mx_kick_params kp = kick_params(c, 0, 0.0, 0);
    const mx_io io = make_io(c, p, 0, 0);
    mx_nb_launch nb = p.nb;
    nb.lanes = c->nl.lanes;
    auto launch = [&]() {
        if (p.kind == PLAN_GRAVITY) mx_launch_gravity(s, c, c->d_pos, p.G, kp, io, 1);
        else if (p.kind == PLAN_MAGNETIC) mx_launch_magnetic(s, c, c->d_pos, p.mu, kp, io, 1);
        else mx_launch_nonbonded(s, nb, c, c->d_pos, p.pp, kp, io);
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
    if (name) *name = mx_force_kernel_name(c);
    // the timed launches overwrote a with the force of the current positions alone
    // (water: the non-bonded part; Berendsen: with the look-ahead velocities), so the
    // next run re-evaluates a; x and v are unchanged
    c->accel_valid = false;
    return NBX_OK;
}

int mx_count_pairs(nbx_ctx* c, int64_t* pairs) {
    mx_plan p;
    std::string why;
    if (!make_plan(c, p, &why)) return nbx_fail(c, NBX_ERR_UNSUPPORTED, why);
    if (all_pairs_tiled(p)) {
        *pairs = c->n * (c->n - 1);
        return NBX_OK;
    }
    if (!p.nb.mode_list)
        return nbx_fail(c, NBX_ERR_UNSUPPORTED, "pair count needs a cubic box with finite cutoffs");
    const int prec = c->prec;
    int rc;
    // a list left by a sorted run is for the cell order, the state is in the user's
    if (!c->nl.ready || !c->accel_valid || c->list_sorted || prec != NBX_PREC_MIXED) {
        c->prec = NBX_PREC_MIXED;
        rc = ensure_ready(c, p);
        c->prec = prec;
        if (prec != NBX_PREC_MIXED) c->accel_valid = false;  // the FP64 path recomputes a
        if (rc) return rc;
    }
    refresh_positions(c, p, false);
    unsigned long long* d = reinterpret_cast<unsigned long long*>(c->d_scal + 12);
    NBX_HIP(c, hipMemsetAsync(d, 0, sizeof(unsigned long long), c->stream));
    mx_nb_launch nb = p.nb;
    nb.lanes = c->nl.lanes;
    mx_launch_count_pairs(c->stream, c, nb, c->d_pos, p.pp, d);
    unsigned long long h = 0;
    NBX_HIP(c, hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    NBX_HIP(c, hipStreamSynchronize(c->stream));
    *pairs = (int64_t)h;
    return NBX_OK;
}

void mx_release(nbx_ctx* c) {
    for (hipEvent_t e : c->kt_ev) (void)hipEventDestroy(e);
    c->kt_ev.clear();
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
    if (c->d_step) (void)hipFree(c->d_step);
    c->d_step = nullptr;
    free_nl(c->nl);
    if (c->d_pos) (void)hipFree(c->d_pos);
    if (c->d_pos2) (void)hipFree(c->d_pos2);
    if (c->d_x2) (void)hipFree(c->d_x2);
    if (c->d_part) (void)hipFree(c->d_part);
    c->d_pos = c->d_pos2 = nullptr;
    c->d_x2 = nullptr;
    c->d_part = nullptr;
}

// =================================================================================
// Learned stride (no counterpart in the reference): descriptor of the current state,
// training-set collection from plain velocity Verlet, the corrected strided integrator
// with its energy-drift monitor and fallback, and the list-based energy it monitors.
// =================================================================================
#include "nbx_learn_ctx.h"
#include "nbx_model.h"

namespace {

struct lj_consts {
    double eps = 0, sigma = 0, R = 0;
};

int learn_plan(nbx_ctx* c, mx_plan& p, lj_consts& lj) {
    std::string why;
    if (c->prec != NBX_PREC_MIXED)
        return nbx_fail(c, NBX_ERR_UNSUPPORTED, "the learned stride runs on the mixed-precision path");
    if (c->sys != NBX_SYS_POTENTIAL || c->therm != NBX_THERM_NONE)
        return nbx_fail(c, NBX_ERR_UNSUPPORTED, "the learned stride needs a potential system without thermostat");
    if (!make_plan(c, p, &why)) return nbx_fail(c, NBX_ERR_UNSUPPORTED, why);
    if (p.kind != PLAN_NONBONDED || !p.nb.mode_list || !p.nb.has_lj || p.nb.el_kind != 0)
        return nbx_fail(c, NBX_ERR_UNSUPPORTED,
                        "the learned stride needs one Lennard-Jones term in a cubic periodic box");
    for (const nbx_term& t : c->terms)
        if (t.kind == NBX_TERM_LJ) { lj.eps = t.p[0]; lj.sigma = t.p[1]; lj.R = t.p[2]; }
    return NBX_OK;
}

// P[0] <- wrap(X[0]) and a Verlet list built for it (the force a stays as it is)
int fresh_list(nbx_ctx* c, mx_plan& p) {
    int rc;
    double grow = 1.0;
    if ((rc = ensure_ready(c, p))) return rc;
    for (int attempt = 0;; ++attempt) {
        refresh_positions(c, p, true);
        launch_build(c, p, 0);
        NBX_HIP(c, hipGetLastError());
        int h[8];
        if ((rc = read_flags(c, h))) return rc;
        if (!h[1]) return NBX_OK;
        if (attempt > 8) return nbx_fail(c, NBX_ERR_CAPACITY, "list overflow persists");
        if ((rc = grow_lists(c, p, grow))) return rc;
        if ((rc = prologue(c, p))) return rc;
    }
}

lr_feat_params feat_params(nbx_ctx* c, int par) {
    lr_feat_params fp{};
    fp.pos = (const double4*)pbuf(c, par);
    fp.nbr = c->nl.d_nbr;
    fp.nbr_wide = c->nl.nbr_wide;
    fp.nnb = c->nl.d_nnb;
    fp.n = (int)c->n;
    fp.L = c->bc.b[0];
    fp.invL = 1.0 / c->bc.b[0];
    fp.a = c->d_a;
    fp.v = c->d_v;
    fp.x = c->d_x;
    return fp;
}

lr_energy_params energy_params(nbx_ctx* c, int par, const lj_consts& lj, double* part) {
    lr_energy_params ep{};
    ep.pos = (const double4*)pbuf(c, par);
    ep.nbr = c->nl.d_nbr;
    ep.nbr_wide = c->nl.nbr_wide;
    ep.nnb = c->nl.d_nnb;
    ep.n = (int)c->n;
    ep.L = c->bc.b[0];
    ep.invL = 1.0 / c->bc.b[0];
    ep.s2 = lj.sigma * lj.sigma;
    ep.R2 = lj.R * lj.R;
    const double sr6 = std::pow(ep.s2 / ep.R2, 3);
    ep.uc = sr6 * sr6 - sr6;
    ep.c4eps = 4.0 * lj.eps;
    ep.v = c->d_v;
    ep.ms = c->d_ms;
    ep.part = part;
    return ep;
}

struct dev_buf {
    void* p = nullptr;
    ~dev_buf() { if (p) (void)hipFree(p); }
    template <class T> T* as() { return static_cast<T*>(p); }
};

int dalloc(nbx_ctx* c, dev_buf& b, size_t bytes) {
    NBX_HIP(c, hipMalloc(&b.p, std::max<size_t>(bytes, 16)));
    return NBX_OK;
}

// kinetic and shifted Lennard-Jones energy of the state whose wrapped copy is P[par]
int list_energy(nbx_ctx* c, int par, const lj_consts& lj, dev_buf& part, double* ke, double* pe) {
    double* out = c->d_scal + 13;  // [13], [14]
    lr_launch_energy(c->stream, c->nl.lanes, energy_params(c, par, lj, part.as<double>()), out);
    double h[2];
    NBX_HIP(c, hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    NBX_HIP(c, hipStreamSynchronize(c->stream));
    *ke = h[0];
    *pe = h[1];
    return NBX_OK;
}

}  // namespace

int mx_learn_energy(nbx_ctx* c, double* ke, double* pe) {
    mx_plan p;
    lj_consts lj;
    int rc;
    if ((rc = learn_plan(c, p, lj))) return rc;
    if ((rc = fresh_list(c, p))) return rc;
    dev_buf part;
    if ((rc = dalloc(c, part, sizeof(double) * 2 * lr_energy_blocks(c->n, c->nl.lanes)))) return rc;
    return list_energy(c, 0, lj, part, ke, pe);
}

int mx_features(nbx_ctx* c, double* feats, double* frames) {
    mx_plan p;
    lj_consts lj;
    int rc;
    if ((rc = learn_plan(c, p, lj))) return rc;
    if ((rc = fresh_list(c, p))) return rc;
    const int64_t n = c->n;
    dev_buf f, fr;
    if ((rc = dalloc(c, f, sizeof(float) * LR_NIN * n))) return rc;
    if ((rc = dalloc(c, fr, sizeof(float) * 6 * n))) return rc;
    lr_feat_params fp = feat_params(c, 0);
    fp.fraw = f.as<float>();
    fp.frames = fr.as<float>();
    lr_launch_features(c->stream, c->nl.lanes, fp);
    NBX_HIP(c, hipGetLastError());
    std::vector<float> hf((size_t)LR_NIN * n), hr((size_t)6 * n);
    NBX_HIP(c, hipMemcpyAsync(hf.data(), f.p, sizeof(float) * hf.size(), hipMemcpyDeviceToHost, c->stream));
    NBX_HIP(c, hipMemcpyAsync(hr.data(), fr.p, sizeof(float) * hr.size(), hipMemcpyDeviceToHost, c->stream));
    NBX_HIP(c, hipStreamSynchronize(c->stream));
    if (feats)
        for (size_t k = 0; k < hf.size(); ++k) feats[k] = hf[k];
    if (frames)
        for (size_t k = 0; k < hr.size(); ++k) frames[k] = hr[k];
    return NBX_OK;
}

int mx_learn_collect(nbx_ctx* c, nbx_model* m, double dt, int stride, int64_t nframes, int nsel,
                     uint64_t seed) {
    mx_plan p;
    lj_consts lj;
    int rc;
    if ((rc = learn_plan(c, p, lj))) return rc;
    if (m->device != c->device)
        return nbx_fail(c, NBX_ERR_INVALID, "model and context live on different devices");
    if ((rc = lr_model_reserve(m, m->count + nframes * (int64_t)nsel)))
        return nbx_fail(c, rc, m->err);
    dev_buf sel, fst, frm, snap;
    if ((rc = dalloc(c, sel, sizeof(int) * nsel))) return rc;
    if ((rc = dalloc(c, fst, sizeof(float) * LR_NIN * nsel))) return rc;
    if ((rc = dalloc(c, frm, sizeof(float) * 6 * nsel))) return rc;
    if ((rc = dalloc(c, snap, sizeof(double) * 9 * nsel))) return rc;
    hipStream_t s = c->stream;
    const double D = dt * stride;
    for (int64_t f = 0; f < nframes; ++f) {
        if ((rc = fresh_list(c, p))) return rc;
        lr_launch_select(s, sel.as<int>(), nsel, c->n, seed, m->frames_collected + f);
        lr_feat_params fp = feat_params(c, 0);
        fp.sel = sel.as<int>();
        fp.nsel = nsel;
        fp.fraw = fst.as<float>();
        fp.frames = frm.as<float>();
        fp.snap_x = snap.as<double>();
        lr_launch_features(s, c->nl.lanes, fp);
        NBX_HIP(c, hipGetLastError());
        int64_t nf = 0;
        if ((rc = mx_vv_run(c, dt, stride, seed, 0, nullptr, nullptr, &nf, false))) return rc;
        lr_target_params tp{};
        tp.sel = sel.as<int>();
        tp.nsel = nsel;
        tp.n = (int)c->n;
        tp.snap = snap.as<double>();
        tp.x1 = c->d_x;
        tp.v1 = c->d_v;
        tp.a1 = c->d_a;
        tp.D = D;
        tp.frames = frm.as<float>();
        tp.fstage = fst.as<float>();
        tp.X = m->d_X;
        tp.T = m->d_T;
        tp.row0 = m->count;
        lr_launch_targets(s, tp);
        NBX_HIP(c, hipGetLastError());
        m->count += nsel;
    }
    NBX_HIP(c, hipStreamSynchronize(s));
    m->frames_collected += nframes;
    return NBX_OK;
}

// report: [0] strided steps taken, [1] plain steps taken after a trip, [2] tripped (0/1),
// [3] strided steps done before the trip, [4] E0, [5] final E, [6] largest relative
// deviation of the accepted blocks and the final state, [7] list builds in strided
// blocks, [8] the relative deviation that tripped the monitor (-1: list overflow or NaN)
int mx_learned_run(nbx_ctx* c, const nbx_model* m, double dt, int stride, int64_t nbig,
                   double tol, int check_every, double* report) {
    mx_plan p;
    lj_consts lj;
    int rc;
    if ((rc = learn_plan(c, p, lj))) return rc;
    if (m->device != c->device)
        return nbx_fail(c, NBX_ERR_INVALID, "model and context live on different devices");
    if (!m->norm_set) return nbx_fail(c, NBX_ERR_INVALID, "the model has no normalization");
    const int64_t n = c->n;
    const size_t n3 = 3 * (size_t)n;
    const int64_t ntiles = (n + 31) / 32;
    hipStream_t s = c->stream;
    dev_buf ffrag, frames, dlab, mpart, epart, bk;
    if ((rc = dalloc(c, ffrag, 2 * (size_t)ntiles * 3 * LR_FRAG_ELEMS))) return rc;
    if ((rc = dalloc(c, frames, sizeof(float) * 6 * n))) return rc;
    if ((rc = dalloc(c, dlab, sizeof(float) * 6 * n))) return rc;
    if ((rc = dalloc(c, mpart, sizeof(double) * 6 * 256))) return rc;
    if ((rc = dalloc(c, bk, sizeof(double) * 3 * n3))) return rc;
    NBX_HIP(c, hipMemsetAsync(ffrag.p, 0, 2 * (size_t)ntiles * 3 * LR_FRAG_ELEMS, s));
    if ((rc = fresh_list(c, p))) return rc;
    // energy-partials buffer sized for the largest lane count the lists may use (32)
    if ((rc = dalloc(c, epart, sizeof(double) * 2 * lr_energy_blocks(n, 32)))) return rc;
    double ke0, pe0;
    if ((rc = list_energy(c, 0, lj, epart, &ke0, &pe0))) return rc;
    const double E0 = ke0 + pe0;
    int hflags[8];
    if ((rc = read_flags(c, hflags))) return rc;
    const int builds0 = hflags[2];
    const double D = dt * stride;
    const int nblk = lr_infer_blocks(n);

    // the corrected start of a strided step from the state in X[par] / P[par]
    auto start_step = [&](int par) {
        lr_feat_params fp = feat_params(c, par);
        fp.frames = frames.as<float>();
        fp.ffrag = ffrag.as<uint16_t>();
        fp.norm = m->d_norm;
        lr_launch_features(s, c->nl.lanes, fp);
        lr_infer_params ip{};
        ip.ffrag = ffrag.as<uint16_t>();
        ip.frames = frames.as<float>();
        ip.params = m->d_params;
        ip.norm = m->d_norm;
        ip.frag = m->d_frag;
        ip.n = (int)n;
        ip.dlab = dlab.as<float>();
        ip.part = mpart.as<double>();
        lr_launch_mlp_infer(s, ip);
        lr_launch_drift(s, mpart.as<double>(), nblk, dlab.as<float>(), D, make_io(c, p, par, 1, D), n);
        after_drift(c, p, par ^ 1);
    };
    auto run_chunk = [&](int64_t msteps) -> int {
        int rc2;
        if ((rc2 = fresh_list(c, p))) return rc2;
        mx_launch_fill_int(s, req_flag(c), 1, 0);  // consumed by the fresh build
        start_step(0);
        const mx_kick_params kp = kick_params(c, 1, D, 0);
        for (int64_t k = 1; k <= msteps; ++k) {
            const int par = (int)(k & 1);
            launch_build(c, p, par);
            launch_step(c, p, kp, par, 0);
            start_step(par);
        }
        NBX_HIP(c, hipGetLastError());
        if (msteps & 1)
            NBX_HIP(c, hipMemcpyAsync(c->d_x, c->d_x2, sizeof(double) * n3, hipMemcpyDeviceToDevice, s));
        return NBX_OK;
    };

    int64_t done = 0, plain = 0, trip_at = -1;
    int tripped = 0;
    double maxdev = 0.0, E = E0, trip_dev = 0.0;
    int64_t builds = 0;
    int builds_prev = builds0;
    const int64_t every = std::max(1, check_every);
    while (done < nbig) {
        const int64_t msteps = std::min<int64_t>(every, nbig - done);
        const double* src[3] = {c->d_x, c->d_v, c->d_a};
        for (int k = 0; k < 3; ++k)
            NBX_HIP(c, hipMemcpyAsync(bk.as<double>() + k * n3, src[k], sizeof(double) * n3,
                                      hipMemcpyDeviceToDevice, s));
        auto restore = [&]() -> int {
            double* dst[3] = {c->d_x, c->d_v, c->d_a};
            for (int k = 0; k < 3; ++k)
                NBX_HIP(c, hipMemcpyAsync(dst[k], bk.as<double>() + k * n3, sizeof(double) * n3,
                                          hipMemcpyDeviceToDevice, s));
            return NBX_OK;
        };
        if ((rc = run_chunk(msteps))) return rc;
        if ((rc = read_flags(c, hflags))) return rc;
        builds += hflags[2] - builds_prev;
        builds_prev = hflags[2];
        // a list overflow inside a strided block means atoms piled up (the strided step
        // went unstable): the monitor trips as it does on an energy excursion
        bool bad = hflags[1] != 0;
        double ke, pe;
        trip_dev = -1.0;
        if (!bad) {
            if ((rc = list_energy(c, (int)(msteps & 1), lj, epart, &ke, &pe))) return rc;
            E = ke + pe;
            const double dev = std::fabs(E - E0) / std::fabs(E0);
            trip_dev = std::isfinite(dev) ? dev : -1.0;
            bad = !(dev <= tol);  // also catches NaN
            if (!bad) {
                maxdev = std::max(maxdev, dev);
                done += msteps;
                c->steps_taken += msteps * stride;
                continue;
            }
        }
        // the monitor tripped: back to the chunk's start, plain velocity Verlet from there
        tripped = 1;
        trip_at = done;
        if ((rc = restore())) return rc;
        // lists built during the failed block are not trusted: fresh lists and forces
        mx_launch_fill_int(s, c->nl.d_flags + 1, 1, 0);
        c->nl.ready = false;
        c->accel_valid = false;
        plain = (nbig - done) * stride;
        int64_t nf = 0;
        if ((rc = mx_vv_run(c, dt, plain, 0, 0, nullptr, nullptr, &nf, false))) return rc;
        if ((rc = fresh_list(c, p))) return rc;
        if ((rc = list_energy(c, 0, lj, epart, &ke, &pe))) return rc;
        E = ke + pe;
        maxdev = std::max(maxdev, std::fabs(E - E0) / std::fabs(E0));
        break;
    }
    if (report) {
        report[0] = (double)done;
        report[1] = (double)plain;
        report[2] = (double)tripped;
        report[3] = (double)trip_at;
        report[4] = E0;
        report[5] = E;
        report[6] = maxdev;
        report[7] = (double)builds;
        report[8] = tripped ? trip_dev : 0.0;
    }
    return NBX_OK;
}

// =================================================================================
// Autotuner (no counterpart in the reference): the mixed path's launch shape chosen by
// timing.  Each candidate runs `trial` velocity-Verlet steps from the current state after
// a warm-up run long enough to capture its step graph; the state, the step count and the
// integrator's inputs are restored after every trial, and the fastest candidate is kept.
//   list plans   : lanes per atom {2, 4, 8, 16, 32} x skin {0.06 .. 0.40} x cutoff
//   gravity      : targets per thread {1, 2, 4} x source chunks {1, 2, 4, ..., 64}
// out: [0] lanes | targets per thread, [1] skin | chunks, [2] best ms per step,
//      [3] ms per step of the configuration found on entry, [4] candidates timed.
// Results are kept per (plan kind, n, box side, cutoff, dt, trial steps, system kind,
// thermostat, device architecture): in the process, and in the tab-separated file named by
// NBX_TUNE_CACHE when that is set (one line per result, later lines win).  A result found
// there is applied without trials (out[4] = 0).
// =================================================================================
#include <chrono>
#include <cstdio>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>

namespace {

struct tune_result {
    double shape0, shape1, best_ms, base_ms, every;
};

std::mutex g_tune_mu;
std::map<std::string, tune_result> g_tune;
std::string g_tune_loaded;   // the cache file read into g_tune

const char* tune_file() {
    const char* f = std::getenv("NBX_TUNE_CACHE");
    return (f && *f) ? f : nullptr;
}

std::string tune_key(nbx_ctx* c, const mx_plan& p, bool grav, double dt, int trial) {
    hipDeviceProp_t prop{};
    std::string arch = "unknown";
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess) arch = prop.gcnArchName;
    arch = arch.substr(0, arch.find(':'));
    char buf[512];
    std::snprintf(buf, sizeof(buf),
                  "%s n=%lld L=%.17g rc=%.17g dt=%.17g trial=%d sys=%d therm=%d lj=%d el=%d arch=%s",
                  grav ? "gravity" : "list", (long long)c->n, grav ? 0.0 : c->bc.b[0],
                  grav ? 0.0 : p.cutoff, dt, trial, c->sys, c->therm, grav ? 0 : p.nb.has_lj,
                  grav ? 0 : p.nb.el_kind, arch.c_str());
    return buf;
}

// caller holds g_tune_mu
void tune_load_file() {
    const char* f = tune_file();
    if (!f || g_tune_loaded == f) return;
    g_tune_loaded = f;
    std::ifstream in(f);
    std::string line;
    while (std::getline(in, line)) {
        const size_t tab = line.find('\t');
        if (tab == std::string::npos) continue;
        std::istringstream vals(line.substr(tab + 1));
        tune_result r{};
        r.every = 1.0;
        if (vals >> r.shape0 >> r.shape1 >> r.best_ms >> r.base_ms) {
            double e;
            if (vals >> e) r.every = e;
            g_tune[line.substr(0, tab)] = r;
        }
    }
}

bool tune_lookup(const std::string& key, tune_result* r) {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    tune_load_file();
    auto it = g_tune.find(key);
    if (it == g_tune.end()) return false;
    *r = it->second;
    return true;
}

void tune_store(const std::string& key, const tune_result& r) {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tune[key] = r;
    if (const char* f = tune_file()) {
        std::ofstream o(f, std::ios::app);
        if (o) {
            char buf[200];
            std::snprintf(buf, sizeof(buf), "\t%.17g\t%.17g\t%.17g\t%.17g\t%.17g\n", r.shape0,
                          r.shape1, r.best_ms, r.base_ms, r.every);
            o << key << buf;
        }
    }
}

}  // namespace

int mx_autotune(nbx_ctx* c, double dt, int trial, double* out, int nout) {
    mx_plan p;
    std::string why;
    if (!make_plan(c, p, &why)) return nbx_fail(c, NBX_ERR_UNSUPPORTED, why);
    const bool grav = p.kind == PLAN_GRAVITY;
    if (!grav && !p.nb.mode_list)
        return nbx_fail(c, NBX_ERR_UNSUPPORTED, "autotuning covers list and gravity plans");
    const std::string key = tune_key(c, p, grav, dt, trial);
    {
        tune_result r{};
        // a kept result is applied only when its shape is one the trials could pick (the
        // file may have been edited): lanes a power of two up to 32, the build interval 1,
        // 2 or 5, a positive skin; gravity 1, 2 or 4 targets per thread and 1-64 chunks
        const auto valid = [grav](const tune_result& q) {
            const int a = (int)q.shape0, e = (int)q.every;
            if (grav) {
                const int ch = (int)q.shape1;
                return (double)a == q.shape0 && (a == 1 || a == 2 || a == 4) &&
                       (double)ch == q.shape1 && ch >= 1 && ch <= 64;
            }
            return (double)a == q.shape0 && a >= 1 && a <= 32 && (a & (a - 1)) == 0 &&
                   std::isfinite(q.shape1) && q.shape1 > 0.0 && (double)e == q.every &&
                   (e == 1 || e == 2 || e == 5);
        };
        if (tune_lookup(key, &r) && valid(r)) {
            if (grav) {
                c->grav_tpt_set = (int)r.shape0;
                c->grav_chunks_set = (int)r.shape1;
            } else {
                c->nl.lanes = (int)r.shape0;
                c->skin = r.shape1;
                c->skin_set = true;
                c->build_every = (int)r.every;
            }
            c->nl.ready = false;
            c->accel_valid = false;
            ++c->version;
            const double vals[6] = {r.shape0, r.shape1, r.best_ms, r.base_ms, 0.0, r.every};
            for (int k = 0; k < nout && k < 6; ++k) out[k] = vals[k];
            return NBX_OK;
        }
    }
    const int64_t n = c->n;
    const size_t n3 = 3 * (size_t)n;
    hipStream_t s = c->stream;
    dev_buf bx, bv, bs;
    int rc;
    if ((rc = dalloc(c, bx, sizeof(double) * n3))) return rc;
    if ((rc = dalloc(c, bv, sizeof(double) * n3))) return rc;
    if ((rc = dalloc(c, bs, sizeof(double) * NBX_SCAL_COUNT))) return rc;
    NBX_HIP(c, hipMemcpyAsync(bs.p, c->d_scal, sizeof(double) * NBX_SCAL_COUNT, hipMemcpyDeviceToDevice, s));
    NBX_HIP(c, hipMemcpyAsync(bx.p, c->d_x, sizeof(double) * n3, hipMemcpyDeviceToDevice, s));
    NBX_HIP(c, hipMemcpyAsync(bv.p, c->d_v, sizeof(double) * n3, hipMemcpyDeviceToDevice, s));
    NBX_HIP(c, hipStreamSynchronize(s));
    const int64_t steps0 = c->steps_taken;
    const bool nh0 = c->nh_advanced;
    auto restore = [&]() -> int {
        c->nh_advanced = nh0;
        NBX_HIP(c, hipMemcpyAsync(c->d_x, bx.p, sizeof(double) * n3, hipMemcpyDeviceToDevice, s));
        NBX_HIP(c, hipMemcpyAsync(c->d_v, bv.p, sizeof(double) * n3, hipMemcpyDeviceToDevice, s));
        NBX_HIP(c, hipMemcpyAsync(c->d_scal, bs.p, sizeof(double) * NBX_SCAL_COUNT, hipMemcpyDeviceToDevice, s));
        NBX_HIP(c, hipStreamSynchronize(s));
        c->steps_taken = steps0;
        c->accel_valid = false;
        return NBX_OK;
    };
    // ms per step of the current configuration
    auto time_config = [&](double* ms) -> int {
        int r;
        c->nl.ready = false;
        c->accel_valid = false;
        ++c->version;
        int64_t nf = 0;
        const int warm = std::max(c->graph_steps + 2, 8);
        if ((r = restore())) return r;
        if ((r = mx_vv(c, dt, warm, 0, 0, nullptr, nullptr, &nf))) return r;
        if ((r = restore())) return r;
        if ((r = mx_vv(c, dt, 2, 0, 0, nullptr, nullptr, &nf))) return r;  // prologue
        if ((r = restore())) return r;
        if ((r = ensure_ready(c, p))) return r;
        NBX_HIP(c, hipStreamSynchronize(s));
        const auto t0 = std::chrono::steady_clock::now();
        if ((r = mx_vv(c, dt, trial, 0, 0, nullptr, nullptr, &nf))) return r;
        const auto t1 = std::chrono::steady_clock::now();
        *ms = std::chrono::duration<double, std::milli>(t1 - t0).count() / trial;
        return NBX_OK;
    };
    double base_ms = 0.0, best_ms = 0.0;
    double best0 = 0.0, best1 = 0.0, best2 = 1.0;
    int count = 0;
    if (grav) {
        const int tpt0 = c->grav_tpt_set, ch0 = c->grav_chunks_set;
        if ((rc = time_config(&base_ms))) return rc;
        best_ms = base_ms;
        best0 = tpt0;
        best1 = ch0;
        // first pass: the best chunk count of each tile width (targets per thread)
        double wms[3] = {1e300, 1e300, 1e300};
        int wch[3] = {0, 0, 0};
        for (int k = 0, tpt = 1; tpt <= 4; tpt *= 2, ++k)
            for (int ch = 1; ch <= 64; ch *= 2) {
                if (n / ch < 256) break;
                c->grav_tpt_set = tpt;
                c->grav_chunks_set = ch;
                double ms;
                if ((rc = time_config(&ms))) return rc;
                ++count;
                if (ms < wms[k]) { wms[k] = ms; wch[k] = ch; }
            }
        // second pass: those shapes twice more each, the fastest repeat kept (one trial
        // varies by a few percent between runs, about the gap between the tile widths)
        for (int k = 0, tpt = 1; tpt <= 4; tpt *= 2, ++k) {
            if (!wch[k]) continue;
            c->grav_tpt_set = tpt;
            c->grav_chunks_set = wch[k];
            for (int rep = 0; rep < 2; ++rep) {
                double ms;
                if ((rc = time_config(&ms))) return rc;
                ++count;
                if (ms < best_ms) { best_ms = ms; best0 = tpt; best1 = wch[k]; }
            }
        }
        c->grav_tpt_set = (int)best0;
        c->grav_chunks_set = (int)best1;
        if (!c->grav_tpt_set) best0 = n >= 16384 ? 2 : 1;
        if (!c->grav_chunks_set) best1 = mx_gravity_chunks(c);
    } else {
        const int lanes0 = c->nl.lanes;
        const double skin0 = c->skin;
        const bool skin_set0 = c->skin_set;
        const int every0 = c->build_every;
        if ((rc = time_config(&base_ms))) return rc;
        best_ms = base_ms;
        best0 = lanes0;
        best1 = c->nl.skin;
        best2 = build_every(c);
        const int lanes_c[5] = {2, 4, 8, 16, 32};
        const double skin_c[8] = {0.06, 0.09, 0.12, 0.15, 0.18, 0.24, 0.32, 0.40};
        const int every_c[3] = {1, 2, 5};
        bool set_best = false;
        for (int ei = 0; ei < 3; ++ei)
            for (int li = 0; li < 5; ++li)
                for (int si = 0; si < 8; ++si) {
                    c->nl.lanes = lanes_c[li];
                    c->skin = skin_c[si] * p.cutoff;
                    c->skin_set = true;
                    c->build_every = every_c[ei];
                    c->stale_replays = 0;
                    double ms;
                    rc = time_config(&ms);
                    if (rc == NBX_ERR_CAPACITY) continue;  // too dense a cell for this skin
                    if (rc) return rc;
                    ++count;
                    // a candidate whose lists went stale ran with the fallback (every step)
                    if (c->stale_replays) continue;
                    if (ms < best_ms) {
                        best_ms = ms;
                        best0 = lanes_c[li];
                        best1 = c->skin;
                        best2 = every_c[ei];
                        set_best = true;
                    }
                }
        c->stale_replays = 0;
        if (set_best) {
            c->nl.lanes = (int)best0;
            c->skin = best1;
            c->skin_set = true;
            c->build_every = (int)best2;
        } else {
            c->nl.lanes = lanes0;
            c->skin = skin0;
            c->skin_set = skin_set0;
            c->build_every = every0;
        }
    }
    if ((rc = restore())) return rc;
    c->nl.ready = false;
    ++c->version;
    tune_store(key, tune_result{best0, best1, best_ms, base_ms, best2});
    const double vals[6] = {best0, best1, best_ms, base_ms, (double)count, best2};
    for (int k = 0; k < nout && k < 6; ++k) out[k] = vals[k];
    return NBX_OK;
}
