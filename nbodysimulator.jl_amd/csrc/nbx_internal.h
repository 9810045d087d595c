// nbx_internal.h -- context layout and kernel-launcher interfaces shared by the
// translation units of libnbx.so.  Device state is SoA FP64: component k of atom i
// lives at x[k * n + i].
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "nbx.h"
#include "nbx_device.h"

// device scalars (nbx_ctx::d_scal): [0] thermostat temperature, [1..7] energies, [8] finite
// flag, [9] temperature, [10] Nose-Hoover zeta, [11] its rate xi, [12] pair count,
// [13..14] list energies, [15] scratch rate for state-free RHS calls
// ZETA_PREV / XI_PREV: zeta and xi before the mixed path's last advance (that advance
// belongs to the drift of the step after the state when a run has ended)
enum { NBX_SCAL_T = 0, NBX_SCAL_ZETA = 10, NBX_SCAL_XI = 11, NBX_SCAL_XI_SCRATCH = 15,
       NBX_SCAL_ZETA_PREV = 16, NBX_SCAL_XI_PREV = 17, NBX_SCAL_COUNT = 32 };
enum { NBX_BINS_STEP = 0, NBX_BINS_FIXED = 1, NBX_BINS_BUILD = 2, NBX_FLAG_BINSET = 5 };

enum { NBX_TERM_LJ = 1, NBX_TERM_ELEC = 2, NBX_TERM_GRAV = 3, NBX_TERM_RF = 4, NBX_TERM_MAG = 5 };

struct nbx_term {
    int kind;
    double p[4];  // LJ: eps sigma R; ELEC: k R; GRAV: G; RF: k R eps_rf; MAG: mu/4pi
};

// Mixed-path neighbour machinery (cell list + Verlet list)
struct nbx_nlist {
    bool ready = false;
    int lanes = 8;            // lanes per atom in the force kernel (G)
    int cap = 0;              // neighbour slots per atom (multiple of lanes)
    int ncell = 0;            // cells per side (0: all-pairs neighbour search)
    int ccap = 0;             // atoms per cell bound
    // how bins follow the atoms: NBX_BINS_STEP -- every step bins (gated on a rebuild
    // request) and the build scans the 27 neighbour cells; NBX_BINS_FIXED -- 3 cells
    // per side, whose 27 neighbour cells are the whole box, so the set-up bins serve
    // every build; NBX_BINS_BUILD -- half-size cells (>= (list radius + 1.5 skin) / 2),
    // binned by each build for the next one, which scans 5 x 5 x 5 cells
    int bin_mode = 0;
    double rlist = 0.0;       // list radius = cutoff + skin
    double cutoff = 0.0;      // largest interaction cutoff
    double skin = 0.0;        // skin in use (rebuild when an atom moved skin/2)
    void* d_nbr = nullptr;    // neighbour slots, interleaved by lanes (see nbx_nonbonded.hip):
                              // 16-bit atom indices for n <= 65,536, else 32-bit
    int nbr_wide = 1;         // 1: 32-bit entries, 0: 16-bit
    int* d_nnb = nullptr;     // neighbour counts
    int* d_cell_count = nullptr;  // 3 x ncells (NBX_BINS_STEP: sets of even / odd steps;
                                  // NBX_BINS_BUILD: three sets rotating at each build)
    int* d_cell_atoms = nullptr;  // 3 x ncells x ccap
    double* d_xref = nullptr; // unwrapped positions at the last build (SoA)
    int* d_flags = nullptr;   // [0] rebuild request, [1] overflow, [2] builds done,
                              // [4] block count of the temperature partials, [5] current bin set, [6] build
                              // finish counter, [7] a list went stale between two build
                              // opportunities
    size_t cap_bytes = 0;
};

// What a captured step graph depends on; a mismatch forces a new capture.  Every device
// buffer the captured kernels reference is part of the key (lists are freed and
// allocated again when they grow or the configuration changes, and an allocation may or
// may not return a previous address).
struct nbx_graph_key {
    int valid = 0;
    int64_t version = 0;      // configuration version (bumped by every setter)
    double dt = 0.0;
    uint64_t seed = 0;
    const void* nbr = nullptr;
    const void* nnb = nullptr;
    const void* cell_count = nullptr;
    const void* cell_atoms = nullptr;
    const void* xref = nullptr;
    const void* flags = nullptr;
    const void* pos = nullptr;
    const void* part = nullptr;
    int cap = 0, ccap = 0, ncell = 0, lanes = 0, steps = 0, chunks = 0, tpt = 0;
    int build_every = 1;      // steps between build opportunities
    int par_off = 0;          // buffer parity of the chunk step a replay starts after
    int em = 0;               // Euler-Maruyama steps instead of velocity Verlet
    const void* ids = nullptr;  // slot -> user atom map of a sorted state (nullptr: none)
    bool operator==(const nbx_graph_key& o) const {
        return valid && o.valid && version == o.version && dt == o.dt && seed == o.seed &&
               nbr == o.nbr && nnb == o.nnb && cell_count == o.cell_count &&
               cell_atoms == o.cell_atoms && xref == o.xref && flags == o.flags && pos == o.pos &&
               part == o.part && cap == o.cap && ccap == o.ccap && ncell == o.ncell &&
               lanes == o.lanes && steps == o.steps && chunks == o.chunks && tpt == o.tpt &&
               em == o.em && build_every == o.build_every && par_off == o.par_off &&
               ids == o.ids;
    }
};

struct nbx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t n = 0;
    int sys = NBX_SYS_POTENTIAL;
    int prec = NBX_PREC_FP64_REF;
    std::vector<double> h_ms, h_qs;
    bool has_q = false;
    nbx_bcv bc{};
    double kb = 1.38e-23;  // kb_SI, src/nbody_simulation.jl:7
    std::vector<nbx_term> terms;                   // potential system, in order
    nbx_term water_lj{}, water_elec{};             // water system
    bool water_lj_set = false, water_elec_set = false;
    double rOH = 0, aHOH = 0, kbond = 0, kangle = 0;
    bool spcfw_set = false;
    int therm = NBX_THERM_NONE;
    double therm_T = 0, therm_p = 0;
    int em = 0;               // set while nbx_em_run steps the state (Euler-Maruyama)
    double skin = 0.0;        // user skin (when skin_set), else 0.12 x the largest cutoff
    bool skin_set = false;
    // Verlet-list build opportunities every build_every steps (1, 2 or 5; the autotuner's
    // choice): the steps in between launch no build or binning kernel.  An atom that moves
    // more than skin/2 between two opportunities invalidates the chunk, which is replayed
    // with a build opportunity at every step.
    int build_every = 1;
    int64_t stale_replays = 0;  // chunks replayed because a list went stale (build_every -> 1)
    // device state
    double* d_x = nullptr;    // positions (unwrapped), 3n
    double* d_v = nullptr;    // velocities, 3n
    double* d_a = nullptr;    // accelerations at d_x, 3n
    double* d_vh = nullptr;   // half-kicked velocities, 3n
    double* d_tmp = nullptr;  // scratch 3n (host transfers, rhs)
    unsigned long long* d_rdf_hist = nullptr;  // device address of the rdf_noalloc histogram
    double* d_ms = nullptr;
    double* d_qs = nullptr;
    double* d_mm = nullptr;   // magnetic moments (SoA 3n), when set
    void* d_mm4 = nullptr;    // the same as float4 (mx, my, mz, 0) for the mixed path
    bool mm_set = false;
    double* d_red = nullptr;  // reduction scratch (n + 64 doubles)
    double* d_scal = nullptr; // device scalars: [0] temperature for Berendsen
    void* d_pos = nullptr;    // mixed path: double4 wrapped positions (+charge) or float4 (+mass)
    void* d_pos2 = nullptr;   // second position buffer (odd steps)
    double* d_x2 = nullptr;   // second unwrapped-position buffer (odd steps)
    double* d_part = nullptr; // gravity: per-chunk partial accelerations
    int grav_chunks = 1;
    int grav_tpt_set = 0;     // targets per thread of the gravity kernel (0: by size)
    int grav_chunks_set = 0;  // source chunks of the gravity kernel (0: by size)
    bool equal_masses = false;  // every body has mass h_ms[0] (gravity takes m out of the sum)
    // 1 / (largest mass): the FP32 gravity positions carry m / m_max, so that the self
    // term's finite 1/r^3 guard times the mass stays finite for any mass scale
    double mass_wscale = 1.0;
    nbx_nlist nl;
    int64_t* d_step = nullptr;       // absolute step base read by replayed graphs
    hipGraphExec_t gexec = nullptr;  // captured block of mixed-path steps
    nbx_graph_key gkey;
    int64_t version = 0;             // bumped by every configuration change
    int graph_steps = 50;            // steps per captured graph (0: no graphs)
    // force-kernel timing over the steps of a run (measurement): a pair of HIP events
    // around every force-kernel launch of the mixed path, summed after each chunk
    bool kt_on = false;
    std::vector<hipEvent_t> kt_ev;
    size_t kt_used = 0;
    double kt_sum_ms = 0.0;
    int64_t kt_launches = 0;
    bool state_set = false;
    bool accel_valid = false; // d_a holds f(d_x, d_v) for the current state
    int64_t steps_taken = 0;
    // Spatial order of the mixed path (Lennard-Jones-type Verlet-list runs): inside
    // nbx_vv_run the per-atom state is kept in an order sorted by cell (slot s holds the
    // user's atom d_perm[s]; d_iperm is the inverse), re-sorted at absolute steps that are
    // multiples of sort_every, so that the neighbours of nearby atoms stay close in memory
    // as the liquid mixes.  Outside a run the state is in the user's order; the Verlet
    // list then belongs to the sorted order when list_sorted is set.  Thermostat random
    // numbers are drawn for the user's atom index.
    int sort_every = 0;       // 0: never sorted
    int* d_perm = nullptr;    // slot -> user atom (identity when perm_identity)
    int* d_iperm = nullptr;   // user atom -> slot
    int* d_pmap = nullptr;    // scratch: a permutation being applied
    void* d_sortbuf = nullptr;   // the next order's maps, then the radix sort's scratch
    size_t sortbuf_bytes = 0;
    double* d_permbuf = nullptr; // scratch of a reordering: 17 n doubles
    bool perm_identity = true;
    bool list_sorted = false;
    int64_t last_sort_step = -1;
    int64_t sorts_applied = 0;   // re-sorts that changed the order (diagnostics)
    // The mixed path's last drift of a run (or chunk) is redone by the next one; its
    // temperature finish advanced the Nose-Hoover zeta for the next step already, so the
    // redone drift must not advance it again
    bool nh_advanced = false;
    std::string err;
};

// error helpers ----------------------------------------------------------------
int nbx_fail(nbx_ctx* c, int code, const std::string& msg);
#define NBX_HIP(c, call)                                                           \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess)                                                      \
            return nbx_fail((c), NBX_ERR_DEVICE,                                   \
                            std::string(#call) + ": " + hipGetErrorString(e_));    \
    } while (0)

static inline unsigned nbx_blocks(int64_t work, int threads) {
    return (unsigned)((work + threads - 1) / threads);
}

// ---- FP64 reference-order kernels (nbx_ref64.hip, built with -ffp-contract=off) ----
// Each ADDS its contribution into dv (SoA) for atoms i in [ib, ie) (stride as noted).
void ref64_fill(hipStream_t s, double* p, int64_t count, double value);
void ref64_aos_to_soa(hipStream_t s, const double* aos, double* soa, int64_t n);
void ref64_soa_to_aos(hipStream_t s, const double* soa, double* aos, int64_t n);
// LJ: i = ib, ib+istride, ... < ie; j = 0, jstride, ..., (nj-1)*jstride
void ref64_lj(hipStream_t s, const double* x, double* dv, const double* ms, int64_t n,
              int64_t ib, int64_t ie, int istride, int jstride, int64_t nj, double eps,
              double sigma, double R, const nbx_bcv& bc);
void ref64_elec(hipStream_t s, const double* x, double* dv, const double* ms,
                const double* qs, int64_t n, int64_t ib, int64_t ie, int water, double k,
                double R, const nbx_bcv& bc);
void ref64_rf(hipStream_t s, const double* x, double* dv, const double* ms, const double* qs,
              int64_t n, int64_t ib, int64_t ie, int water, double k, double R,
              double eps_rf, const nbx_bcv& bc);
void ref64_mag(hipStream_t s, const double* x, double* dv, const double* ms, const double* mm,
               int64_t n, int64_t ib, int64_t ie, double mu_4pi);
// Nose-Hoover (src/thermostats.jl:116-122): dv -= zeta v; xi_out = (1/tau)^2 (T/T0 - (ndf+1)/ndf)
// with T = scal[0] (computed by ref64_temperature) and zeta = scal[NBX_SCAL_ZETA]
void ref64_nosehoover(hipStream_t s, double* dv, const double* v, int64_t n, const double* scal,
                      double* xi_out, double T0, double tau, int64_t ndf);
// zeta += dt xi (the extended coordinate's drift; its acceleration is zero)
void ref64_nh_drift(hipStream_t s, double* scal, double dt);
void ref64_grav(hipStream_t s, const double* x, double* dv, const double* ms, int64_t n,
                int64_t ib, int64_t ie, double G);
void ref64_bond(hipStream_t s, const double* x, double* dv, const double* ms, int64_t n,
                int64_t ib, int64_t ie, double rOH, double kbond);
// molecules m in [mb, me)
void ref64_angle(hipStream_t s, const double* x, double* dv, const double* ms, int64_t n,
                 int64_t mb, int64_t me, double aHOH, double ka);
// T = sum m v^2 / (kb (3N - Nc)) into *d_T (sequential order), then dv += ... (Berendsen)
void ref64_temperature(hipStream_t s, const double* v, const double* ms, int64_t n,
                       double kb, int64_t Nc, double* d_tmp, double* d_T);
void ref64_berendsen(hipStream_t s, double* dv, const double* v, int64_t n, const double* d_T,
                     double T0, double gamma);
// VV pieces
void ref64_drift(hipStream_t s, double* x, const double* v, const double* a, double* vh,
                 int64_t n, double dt);
void ref64_kick(hipStream_t s, double* v, const double* vh, const double* a, int64_t n,
                double dt);
void ref64_em(hipStream_t s, double* x, double* v, const double* a, int64_t n, double dt,
              double gamma, double sigma, uint64_t seed, int64_t step);
void ref64_langevin_half(hipStream_t s, double* v, const double* ms, int64_t n, double kb,
                         double T, double gamma, double dt, uint64_t seed, int64_t step,
                         uint32_t stream);
void ref64_andersen(hipStream_t s, double* v, const double* ms, int64_t n, int water,
                    double kb, double T, double prob, uint64_t seed, int64_t step);
// energies: per-term sums into *d_out (FP64, sequential final sum)
void ref64_kinetic(hipStream_t s, const double* v, const double* ms, int64_t n, double* d_tmp,
                   double* d_out);
void ref64_lj_pe(hipStream_t s, const double* x, int64_t n, int jstride, int64_t nj, double eps,
                 double sigma, double R, const nbx_bcv& bc, double* d_tmp, double* d_out);
void ref64_elec_pe(hipStream_t s, const double* x, const double* qs, int64_t n, int water,
                   double k, double R, int rf, double eps_rf, const nbx_bcv& bc,
                   double* d_tmp, double* d_out);
void ref64_grav_pe(hipStream_t s, const double* x, const double* ms, int64_t n, double G,
                   double* d_tmp, double* d_out);
void ref64_bonded_pe(hipStream_t s, const double* x, int64_t n, double rOH, double kbond,
                     double aHOH, double ka, double* d_tmp, double* d_out2);
void ref64_rdf_hist(hipStream_t s, const double* frame_aos, int64_t n, int jstride,
                    int64_t nj, double L, int64_t maxbin, unsigned long long* d_hist);
// address of the library-owned rdf histogram array on `device` (looked up once per device
// with hipGetSymbolAddress, with that device current); 0 or -1
#define NBX_MAX_DEVICES 64
int ref64_rdf_scratch(int device, unsigned long long** out);
void ref64_zero_u64(hipStream_t s, unsigned long long* p, int64_t count);
int64_t ref64_check_finite(hipStream_t s, const double* p, int64_t count, int* d_flag);

// ---- mixed path (nbx_mixed.cpp driving nbx_nonbonded.hip and nbx_gravity.hip) ----
int mx_time_force(nbx_ctx* c, int iters, double* avg_ms, const char** name);
const char* mx_force_kernel_name(nbx_ctx* c);
int mx_valu_probe(nbx_ctx* c, int kind, int reps, double* rate);
int mx_copy_probe(nbx_ctx* c, int64_t bytes, int iters, double* avg_ms);
int mx_autotune(nbx_ctx* c, double dt, int trial_steps, double* out, int nout);
// learned stride (nbx_mixed.cpp)
int mx_features(nbx_ctx* c, double* feats, double* frames);
int mx_learn_energy(nbx_ctx* c, double* ke, double* pe);
int mx_learn_collect(nbx_ctx* c, nbx_model* m, double dt, int stride, int64_t nframes, int nsel,
                     uint64_t seed);
int mx_learned_run(nbx_ctx* c, const nbx_model* m, double dt, int stride, int64_t nbig,
                   double tol, int check_every, double* report);
int mx_count_pairs(nbx_ctx* c, int64_t* pairs);
void mx_release(nbx_ctx* c);
// the mixed path's sorted order back to the identity (a new state)
int mx_reset_order(nbx_ctx* c);
