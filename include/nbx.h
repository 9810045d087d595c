/*
 * nbx.h -- C ABI of the MI355X (gfx950) force-evaluation and velocity-Verlet engine
 * for NBodySimulator.jl systems.
 *
 * Reference interfaces replaced (paths relative to the NBodySimulator.jl tree):
 *   - get_accelerating_function(p, sim) -> acceleration!(dv, u, v, t, i)
 *       src/nbody_to_ode.jl:156-242; force laws src/basic_potentials.jl:240-433
 *   - soode_system!(dv, v, u, p, t)        src/nbody_to_ode.jl:474-488, :502-532
 *   - the VelocityVerlet step driven by run_simulation / calculate_simulation
 *       src/nbody_simulation_result.jl:468-487 (+ OrdinaryDiffEqSymplecticRK)
 *   - thermostats: berendsen_acceleration! src/thermostats.jl:76-91,
 *       Andersen callback src/nbody_simulation_result.jl:504-540,
 *       Langevin src/nbody_to_ode.jl:567-598
 *   - energies / temperature / rdf          src/nbody_simulation_result.jl:182-446, :664-709
 *
 * Conventions
 *   - Coordinate arrays cross the boundary as Julia column-major 3 x n Float64
 *     matrices (x1 y1 z1 x2 ...), exactly the u/v/dv arrays the reference's closures
 *     see.  Atom indices are 0-based (Julia's i is i+1).
 *   - Every call is synchronous: host buffers may be reused on return.
 *   - Return value: NBX_OK (0) or a negative NBX_ERR_* code; nbx_last_error(ctx)
 *     gives the message (the reference throws BoundsError / KeyError / MethodError).
 *   - One context = one HIP stream on the current device; a context is not to be
 *     shared between threads without external locking (the reference is
 *     single-threaded and non-reentrant as well).
 */
#ifndef NBX_H
#define NBX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NBX_OK 0
#define NBX_ERR_INVALID (-1)     /* bad argument / unsupported combination        */
#define NBX_ERR_BOUNDS (-2)      /* index outside 0..n-1 (reference: BoundsError)   */
#define NBX_ERR_UNSUPPORTED (-3) /* feature not available for this system / mode    */
#define NBX_ERR_DEVICE (-4)      /* HIP runtime error (no device, launch failure)   */
#define NBX_ERR_CAPACITY (-5)    /* a cell or neighbour list overflowed its bound   */
#define NBX_ERR_NONFINITE (-6)   /* a state value became NaN/Inf during a run       */

/* Boundary conditions (src/boundary_conditions.jl:25-109) */
#define NBX_BC_INFINITE 0  /* InfiniteBox()                                        */
#define NBX_BC_CUBIC 1     /* CubicPeriodicBoundaryConditions(L): box[0] = L         */
#define NBX_BC_PERIODIC 2  /* PeriodicBoundaryConditions: box = xmin xmax ymin ...   */

/* System kinds (src/nbody_system.jl:72-75, :182-191) */
#define NBX_SYS_POTENTIAL 0 /* PotentialNBodySystem (GravitationalSystem, ChargedParticles) */
#define NBX_SYS_WATER 1     /* WaterSPCFw: atoms ordered O,H1,H2 per molecule (n = 3 nmol) */

/* Precision modes */
#define NBX_PREC_FP64_REF 0 /* FP64, all pairs in the reference's j order, no FMA
                               contraction: bit-for-bit the reference's arithmetic   */
#define NBX_PREC_MIXED 1    /* FP32 pair forces + Kahan accumulation, cell/Verlet
                               lists (LJ, Coulomb, RF) or LDS tiles (gravity); FP64
                               positions, velocities, integrator and reductions      */

/* Thermostats (src/thermostats.jl:7-156) */
#define NBX_THERM_NONE 0
#define NBX_THERM_ANDERSEN 1  /* param = collision frequency nu */
#define NBX_THERM_BERENDSEN 2 /* param = coupling time tau      */
#define NBX_THERM_LANGEVIN 3  /* param = friction gamma         */
#define NBX_THERM_NOSEHOOVER 4 /* param = relaxation time tau    */

typedef struct nbx_ctx nbx_ctx;

/* Create a context for n coordinate columns.  masses: n values (get_masses order,
 * src/nbody_simulation_result.jl:121-139); charges: n values or NULL.
 * system_kind: NBX_SYS_*.  For NBX_SYS_WATER, n must be a multiple of 3.
 * Replaces the host-side setup of src/nbody_to_ode.jl:263-375. */
int nbx_create(nbx_ctx** out, int64_t n, int32_t system_kind, const double* masses,
               const double* charges);
int nbx_destroy(nbx_ctx* ctx);
const char* nbx_last_error(const nbx_ctx* ctx);
/* Library version string, and the number of HIP devices visible (0 without a GPU). */
const char* nbx_version(void);
int nbx_device_count(void);

/* Boundary conditions: src/boundary_conditions.jl:25-109.  box: 6 doubles for
 * PERIODIC, box[0] = L for CUBIC, ignored for INFINITE. */
int nbx_set_boundary(nbx_ctx* ctx, int32_t kind, const double* box);
/* Boltzmann constant of the unit system (NBodySimulation.kb, src/nbody_simulation.jl:42-54) */
int nbx_set_kb(nbx_ctx* ctx, double kb);
int nbx_set_precision(nbx_ctx* ctx, int32_t mode);

/* Potential terms of a PotentialNBodySystem, applied per atom in the order added
 * (the reference iterates its Dict, src/nbody_to_ode.jl:87-99).  For WaterSPCFw they
 * set the system's lj_parameters / e_parameters (src/nbody_system.jl:182-191).  */
/* LennardJonesParameters(eps, sigma, R): src/basic_potentials.jl:59-69, :240-272 */
int nbx_add_lennard_jones(nbx_ctx* ctx, double eps, double sigma, double R);
/* ElectrostaticParameters(k, R) (R = +Inf: no cutoff): src/basic_potentials.jl:147-163, :274-304 */
int nbx_add_electrostatic(nbx_ctx* ctx, double k, double R);
/* Reaction-field electrostatics (no reference counterpart; Tironi et al. 1995):
 * eps_rf = dielectric constant beyond R (+Inf: conducting boundary). */
int nbx_add_reaction_field(nbx_ctx* ctx, double k, double R, double eps_rf);
/* GravitationalParameters(G): src/basic_potentials.jl:109-115, :306-331 */
int nbx_add_gravitational(nbx_ctx* ctx, double G);
/* Magnetic moments of MagneticParticle bodies (src/bodies.jl:102-107): mm is 3 x n
 * column-major (mm of body i in column i). */
int nbx_set_moments(nbx_ctx* ctx, const double* mm);
/* MagnetostaticParameters(mu_4pi): src/basic_potentials.jl:171-205, :333-365 (dipole-dipole,
 * all pairs, no boundary conditions, as the reference).  Needs nbx_set_moments first. */
int nbx_add_magnetostatic(nbx_ctx* ctx, double mu_4pi);
/* SPCFwParameters(rOH, aHOH, kb, ka): src/basic_potentials.jl:233-238, :367-433 */
int nbx_set_spcfw(nbx_ctx* ctx, double rOH, double aHOH, double kbond, double kangle);
/* Thermostat (NBX_THERM_*), target T and its parameter (nu, tau or gamma). */
int nbx_set_thermostat(nbx_ctx* ctx, int32_t kind, double T, double param);
/* Neighbour-list skin for the mixed path, in length units (default: 0.12 x the
 * largest cutoff, i.e. 0.3 sigma for R = 2.5 sigma).  The list is rebuilt on the
 * device whenever an atom has moved more than skin/2 since the last build. */
int nbx_set_skin(nbx_ctx* ctx, double skin);

/* State: u, v are 3 x n column-major (gather_bodies_initial_coordinates,
 * src/nbody_to_ode.jl:1-75).  get_state returns the current step's u and v. */
int nbx_set_state(nbx_ctx* ctx, const double* u, const double* v);
int nbx_get_state(const nbx_ctx* ctx, double* u, double* v);

/* One right-hand-side evaluation (soode_system!, src/nbody_to_ode.jl:474-488 and
 * :502-532): dv (3 x n) is OVERWRITTEN with the sum, per atom and in the reference's
 * order, of every potential closure, the SPC/Fw bond/angle group terms and the
 * Berendsen thermostat term.  u, v: host arrays (v is read only for Berendsen). */
int nbx_rhs(nbx_ctx* ctx, const double* u, const double* v, double* dv);

/* The per-atom closure protocol (acceleration!(dv, u, v, t, i) of one potential,
 * src/nbody_to_ode.jl:156-242): ADDS into dv (3 x n) the contribution of potential
 * term `term` (0-based, in the order added) for every atom i in [i_begin, i_end).
 * For water, term 0 = electrostatic, 1 = Lennard-Jones (O atoms only),
 * 2 = SPC/Fw bonds, 3 = SPC/Fw angles (group closure, src/nbody_to_ode.jl:244-261). */
int nbx_accelerations(nbx_ctx* ctx, int32_t term, const double* u, int64_t i_begin,
                      int64_t i_end, double* dv);

/* Velocity-Verlet run from the context state (OrdinaryDiffEqSymplecticRK
 * VelocityVerlet, called from src/nbody_simulation_result.jl:480-487):
 *   u' = u + dt v + dt^2 (a/2);  vh = v + dt (a/2);  a' = f(u', vh);  v' = vh + dt (a'/2)
 * with thermostats applied per step (Andersen after the step; Langevin as half
 * Ornstein-Uhlenbeck steps around it; Berendsen inside f).  Random numbers come from
 * Philox4x32-10 keyed by `seed` with counter (atom, absolute step), so a run split
 * into pieces reproduces an unsplit one bit for bit.
 * Frames: if save_every > 0, frame 0 is the initial state and frame k the state
 * after k*save_every steps; u_frames/v_frames hold 3n doubles per frame
 * (1 + nsteps/save_every frames).  Pass NULL frames to keep everything in HBM.
 * On return the context state is the final state. */
int nbx_vv_run(nbx_ctx* ctx, double dt, int64_t nsteps, uint64_t seed, int64_t save_every,
               double* u_frames, double* v_frames);
/* Euler-Maruyama run of the Langevin SDE that the reference builds for a potential
 * system with a LangevinThermostat (SciMLBase.SDEProblem, src/nbody_to_ode.jl:567-598,
 * solved by StochasticDiffEq EM through run_simulation, src/nbody_simulation_result.jl:
 * 480-487): with a = f(x) the potential accelerations,
 *   x' = x + dt v;  v' = v + dt (a - gamma v) + sigma sqrt(dt) xi,
 *   sigma = sqrt(2 gamma kb T / m_1)  (the first body's mass, for every atom).
 * Needs the Langevin thermostat (T, gamma); water systems are not supported.  xi comes
 * from Philox4x32-10 keyed by `seed` with counter (atom, absolute step); steps count
 * into nbx_steps_taken like nbx_vv_run.  Frames and the final state as nbx_vv_run. */
int nbx_em_run(nbx_ctx* ctx, double dt, int64_t nsteps, uint64_t seed, int64_t save_every,
               double* u_frames, double* v_frames);
/* Steps taken by all nbx_vv_run / nbx_em_run calls since the last set_state
 * (solution.stats.naccept). */
int64_t nbx_steps_taken(const nbx_ctx* ctx);
/* Nose-Hoover thermostat state of the current step: zeta and its rate xi, which the
 * reference keeps in the extra state column, u[3n+1] and v[3n+1]
 * (gather_bodies_initial_coordinates src/nbody_to_ode.jl:1-20, nosehoover_acceleration!
 * src/thermostats.jl:120-126).  Both are 0 after nbx_set_state and without Nose-Hoover. */
int nbx_get_thermostat_state(const nbx_ctx* ctx, double* zeta, double* xi);

/* Energies of the current state, in FP64 on the device:
 *   ke = kinetic_energy (src/nbody_simulation_result.jl:209-218)
 *   pe = potential_energy with the reference's conventions (:239-403): LJ tail for
 *        far pairs, Coulomb q_j/R beyond R, SPC/Fw bonds and angles; gravity adds
 *        nothing there, so it is returned separately in pe_grav (may be NULL). */
int nbx_energies(nbx_ctx* ctx, double* ke, double* pe, double* pe_grav);
/* md_temperature of the current state (src/thermostats.jl:87-91 with the system's
 * degrees of freedom, src/nbody_simulation_result.jl:152-164). */
int nbx_temperature(nbx_ctx* ctx, double* T);

/* rdf (src/nbody_simulation_result.jl:664-709) over nframes position frames
 * (each 3 x n, host memory), cubic box only; rs and gr receive maxbin values. */
int nbx_rdf(nbx_ctx* ctx, const double* frames, int64_t nframes, int64_t maxbin,
            double* rs, double* gr);

/* No counterpart in the reference (it allocates freely): rdf exactly as nbx_rdf, for
 * callers that cannot allocate device memory -- code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc.  It calls no hipMalloc/hipFree: the histogram is a
 * library-owned device array of NBX_RDF_NOALLOC_MAXBIN + 1 counters (128 KiB per GPU, held
 * until the process ends, cleared by a fill kernel at every call) and each frame is staged in
 * the context's own buffer.  maxbin > NBX_RDF_NOALLOC_MAXBIN fails with NBX_ERR_CAPACITY.
 * Every call on one GPU uses that GPU's one array: calls that overlap on a GPU (from several
 * host threads or contexts) race and must be serialised by the caller. */
#define NBX_RDF_NOALLOC_MAXBIN 16384
int nbx_rdf_noalloc(nbx_ctx* ctx, const double* frames, int64_t nframes, int64_t maxbin,
                    double* rs, double* gr);

/* Measurement support: average duration (ms) of the step's dominant kernel over
 * `iters` back-to-back launches on the context's stream, timed with HIP events on
 * that stream, on the current state (which is left unchanged).  *kernel_name
 * receives a static string naming the kernel. */
int nbx_time_force_kernel(nbx_ctx* ctx, int32_t iters, double* avg_ms,
                          const char** kernel_name);
/* Measurement support (no reference counterpart): with enable != 0, every later
 * mixed-precision nbx_vv_run records a pair of HIP events on the context's stream around
 * each launch of its force kernel (k_grav_tiled; k_nb_list / k_nb_allpairs, whose launch
 * also ends the step and drifts) and runs without captured step graphs; the counters are
 * reset by every call.  nbx_get_kernel_timing returns the average duration (ms) and the
 * number of launches timed since, and the kernel's name (static string). */
int nbx_set_kernel_timing(nbx_ctx* ctx, int32_t enable);
int nbx_get_kernel_timing(nbx_ctx* ctx, double* avg_ms, int64_t* launches,
                          const char** kernel_name);
/* Ordered in-cutoff pairs of the current state (pair-interaction count of one force
 * evaluation; n(n-1) for gravity). */
int nbx_count_pairs(nbx_ctx* ctx, int64_t* pairs);
/* Measurement support (no reference counterpart): average duration (ms) of a plain
 * device-to-device copy kernel moving `bytes` (rounded down to 16 B) over `iters`
 * launches on the context's stream.  Calibrates HBM byte counters (the kernel reads
 * and writes exactly `bytes` each) and gives the achievable bandwidth. */
int nbx_copy_probe(nbx_ctx* ctx, int64_t bytes, int32_t iters, double* avg_ms);
/* Measurement support (no reference counterpart): VALU issue rate, in wave64
 * instructions per SIMD per nanosecond (best of `reps` launches), of a stream of
 * independent instructions of one kind: 0 v_fmac_f32, 1 v_mul_f32, 2 v_rsq_f32.  With
 * the instruction counts of a kernel it gives the issue-limited time of that kernel. */
int nbx_valu_probe(nbx_ctx* ctx, int32_t kind, int32_t reps, double* rate);
/* Autotuner (no reference counterpart): times `trial_steps` mixed-path steps of length dt
 * from the current state for every candidate launch shape and keeps the fastest -- list
 * plans: lanes per atom {2, 4, 8, 16, 32} x Verlet skin {0.06 .. 0.40} x cutoff; gravity:
 * targets per thread {1, 2, 4} x source chunks {1 .. 64}; list plans also try list-build
 * opportunities every 1, 2 or 5 steps.  The state and step
 * count are restored.  out (nout <= 6): [0] lanes | targets per thread, [1] skin | chunks,
 * [2] best ms per step, [3] ms per step of the configuration on entry, [4] candidates
 * timed, [5] steps between list-build opportunities (1 for gravity).
 * Results are kept per (list | gravity plan, n, box side, cutoff, dt, trial_steps, system
 * kind, thermostat, device architecture) in the process and, when the environment variable
 * NBX_TUNE_CACHE names a file, in that file (tab-separated lines, appended); a kept result
 * is applied without trials and reports out[4] = 0. */
int nbx_autotune(nbx_ctx* ctx, double dt, int32_t trial_steps, double* out, int32_t nout);
/* Diagnostics of the mixed path (no reference counterpart): out[0] Verlet-list builds
 * so far, [1] list slots per atom, [2] cells per box side (0: all-pairs search),
 * [3] lanes per atom in the force kernel, [4] steps per captured graph, [5] steps between
 * list-build opportunities, [6] chunks replayed because a list went stale between two
 * opportunities (each such replay sets [5] to 1), [7] spatial re-sorts of the state
 * applied inside nbx_vv_run (NBX_SORT_EVERY).  nout <= 8. */
int nbx_get_stats(nbx_ctx* ctx, int64_t* out, int32_t nout);

/* ---------------------------------------------------------------------------------
 * Batched replicas: many independent Lennard-Jones systems (one NBodySimulation each,
 * src/nbody_simulation.jl:42-54, with PotentialNBodySystem + LennardJonesParameters and
 * CubicPeriodicBoundaryConditions) integrated together by one persistent launch in
 * which each replica is owned by one workgroup.  The reference runs such an ensemble
 * as separate run_simulation calls (src/nbody_simulation_result.jl:468-487); results per
 * replica follow the single-system mixed path (FP32 pair forces with Kahan sums, FP64
 * state, velocity Verlet, Andersen / Langevin).  Replica r draws its random numbers from
 * Philox keyed by seed + r * 0x9E3779B97F4A7C15 with counter (local atom, step), i.e. the
 * streams of a single-system run of replica r with that seed.
 * --------------------------------------------------------------------------------- */
typedef struct nbx_batch nbx_batch;

/* nrep replicas of natoms[r] atoms (<= 4096 each) in cubic boxes of side box_L[r];
 * masses: sum(natoms) values, replicas concatenated.  One LJ law (eps, sigma, R) for all. */
int nbx_batch_create(nbx_batch** out, int32_t nrep, const int64_t* natoms,
                     const double* box_L, const double* masses, double eps, double sigma,
                     double R, double kb);
int nbx_batch_destroy(nbx_batch* b);
const char* nbx_batch_last_error(const nbx_batch* b);
/* per replica: NBX_THERM_NONE, NBX_THERM_ANDERSEN (param nu) or NBX_THERM_LANGEVIN (gamma) */
int nbx_batch_set_thermostat(nbx_batch* b, int32_t replica, int32_t kind, double T,
                             double param);
int nbx_batch_set_skin(nbx_batch* b, double skin);
/* Autotuner (no reference counterpart): times `trial_steps` steps of length dt from the
 * current state for the current skin and for skins {0.2, 0.3, 0.4, 0.5, 0.6, 0.8} x sigma,
 * restoring the state after each trial, and keeps the fastest.  out[3] (or NULL): the
 * chosen skin, its ms per step, the ms per step of the skin on entry. */
int nbx_batch_autotune(nbx_batch* b, double dt, int64_t trial_steps, double* out);
/* u, v: 3 x sum(natoms) column-major, replicas concatenated */
int nbx_batch_set_state(nbx_batch* b, const double* u, const double* v);
int nbx_batch_get_state(const nbx_batch* b, double* u, double* v);
/* nsteps velocity-Verlet steps of every replica; steps_per_launch steps per persistent
 * launch (0: 1000), the launches replayed from a captured hipGraph. */
int nbx_batch_run(nbx_batch* b, double dt, int64_t nsteps, uint64_t seed,
                  int64_t steps_per_launch);
/* per-replica kinetic and potential energy (reference conventions) and in-cutoff
 * ordered pair counts of the current state; each array holds nrep values (or NULL) */
int nbx_batch_energies(nbx_batch* b, double* ke, double* pe, int64_t* pairs);
/* out[0] list builds summed over replicas, [1] steps taken, [2] list slots per atom
 * (nout <= 3) */
int nbx_batch_stats(nbx_batch* b, int64_t* out, int32_t nout);

/* ---------------------------------------------------------------------------------
 * Learned stride (no counterpart in the reference; it accelerates the
 * `solve(SecondOrderODEProblem(sim), VelocityVerlet(); dt)` path of
 * src/nbody_simulation_result.jl:480-487 for a Lennard-Jones fluid in a cubic box).
 *
 * One strided step of length D = stride * dt is a velocity-Verlet step of length D whose
 * drift and kick carry per-atom corrections predicted by a small MLP:
 *     x' = x + D v + D^2 a/2 + dx,   a' = f(x'),   v' = v + D (a + a')/2 + dv
 * (dx, dv with their mean over atoms removed).  Network input per atom (35): 32 Gaussian
 * radial density bins (centres 0.8 .. 2.5 sigma, standard deviation 0.1 sigma, cosine
 * switch over 2.0 .. 2.5 sigma), then v.e1, |v - (v.e1) e1| and |a|, where the force
 * frame is e1 = a/|a|, e2 = the unit part of v normal to e1, e3 = e1 x e2.  Output (6):
 * dx and dv along (e1, e2, e3), each divided by its output scale.  Layers 35-128-128-
 * 128-6 with SiLU, BF16 weights and activations, FP32 accumulation (MFMA), FP32 master
 * weights trained with Adam.  Training targets come from plain velocity-Verlet
 * trajectories at dt: dx = x(t+D) - (x + D v + D^2 a/2), dv = v(t+D) - (v + D (a +
 * a(t+D))/2), in the frame at t.
 * ------------------------------------------------------------------------------- */
typedef struct nbx_model nbx_model;
#define NBX_MLP_INPUTS 35
#define NBX_MLP_OUTPUTS 6
#define NBX_MLP_PARAMS 38406
/* normalization block: feature mean[35], feature inverse scale[35], output scale[6] */
#define NBX_MLP_NORM 76

/* parameters initialised U(-1/sqrt(fan_in), 1/sqrt(fan_in)) from seed; identity
 * normalization; empty training set */
int nbx_model_create(nbx_model** out, uint64_t seed);
int nbx_model_destroy(nbx_model* model);
const char* nbx_model_last_error(const nbx_model* model);
/* flat FP32 parameters (W1[128][35], b1, W2[128][128], b2, W3[128][128], b3, W4[6][128],
 * b4) as doubles, and the normalization block; either pointer may be NULL.  Setting
 * parameters resets the Adam state. */
int nbx_model_set(nbx_model* model, const double* params, const double* norm);
int nbx_model_get(nbx_model* model, double* params, double* norm);
/* network outputs (normalized units) of n raw feature rows: feats n x 35, out n x 6 */
int nbx_model_forward(nbx_model* model, const double* feats, int64_t n, double* out);
/* gradient of mean((out - targets / output_scale)^2) over n x 6 with respect to the flat
 * parameters, and that loss; n a multiple of 256; targets n x 6 in raw units */
int nbx_model_gradient(nbx_model* model, const double* feats, const double* targets,
                       int64_t n, double* grad, double* loss);
/* training set (device resident): append rows, read rows back, clear */
int nbx_model_add_samples(nbx_model* model, const double* feats, const double* targets,
                          int64_t n);
int nbx_model_get_samples(nbx_model* model, int64_t first, int64_t n, double* feats,
                          double* targets, int64_t* count);
int nbx_model_clear_samples(nbx_model* model);
/* normalization fitted to the training set: feature mean / (1 / standard deviation),
 * per-component RMS of the targets */
int nbx_model_fit_normalization(nbx_model* model);
/* iters Adam steps (beta 0.9 / 0.999, eps 1e-8) on minibatches of `batch` rows (a
 * multiple of 256) drawn with replacement (Philox, seed); losses[iters] or NULL */
int nbx_model_train(nbx_model* model, int64_t iters, int32_t batch, double lr, uint64_t seed,
                    double* losses);
/* per-component mean squared error of the outputs on rows [first, first + n) of the
 * training set, and (base6, or NULL) that of predicting zero, in normalized units */
int nbx_model_evaluate(nbx_model* model, int64_t first, int64_t n, double* mse6,
                       double* base6);

/* descriptor of the context's current state: feats n x 35 (raw), frames n x 6 (e1, e2);
 * mixed precision, one Lennard-Jones term, cubic box, no thermostat */
int nbx_features(nbx_ctx* ctx, double* feats, double* frames);
/* kinetic energy and the Lennard-Jones energy shifted to zero at the cutoff, from the
 * Verlet list (the reference's potential_energy equals pe_shifted + n(n-1)/2 * U(R),
 * src/nbody_simulation_result.jl:293-319) */
int nbx_learn_energy(nbx_ctx* ctx, double* ke, double* pe_shifted);
/* training set from plain velocity Verlet at dt: per frame, `atoms_per_frame` atoms
 * drawn with Philox(seed) are described at t, the context advances `stride` steps, and
 * their targets are appended to the model's training set; nframes frames */
int nbx_learn_collect(nbx_ctx* ctx, nbx_model* model, double dt, int32_t stride,
                      int64_t nframes, int32_t atoms_per_frame, uint64_t seed);
/* nsteps strided steps.  Every check_every strided steps the total (shifted) energy is
 * compared with the run's initial value; when |E - E0| / |E0| exceeds drift_tol the
 * state returns to the start of that block and plain velocity Verlet at dt runs the
 * remaining steps.  report[9]: strided steps taken, plain steps taken, tripped (0/1),
 * strided steps before the trip (-1: none), E0, final E, largest relative deviation of
 * the accepted blocks and the final state, Verlet-list builds during the strided blocks,
 * the relative deviation that tripped the monitor (-1: list overflow or NaN; 0: none).  A list overflow inside a strided block
 * (atoms piling up after an unstable step) trips the monitor like an energy excursion. */
int nbx_learned_run(nbx_ctx* ctx, const nbx_model* model, double dt, int32_t stride,
                    int64_t nsteps, double drift_tol, int32_t check_every, double* report);

#ifdef __cplusplus
}
#endif
#endif /* NBX_H */
