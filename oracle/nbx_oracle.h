/*
 * nbx_oracle.h -- FP64 CPU restatement of NBodySimulator.jl's force-evaluation and
 * velocity-Verlet hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product library (libnbx.so) never links, loads or calls it.
 *
 * Conventions: coordinate arrays are Julia column-major 3 x n matrices, i.e.
 * x1 y1 z1 x2 y2 z2 ...; atom indices are 0-based here (Julia's i is i+1).
 * Build: gcc -O2 -ffp-contract=off (no FMA contraction, like Julia without
 * @fastmath/@muladd), so every expression rounds exactly as written.
 */
#ifndef NBX_ORACLE_H
#define NBX_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_BC_INFINITE = 0, ORC_BC_CUBIC = 1, ORC_BC_PERIODIC = 2 };
enum { ORC_TERM_LJ = 1, ORC_TERM_ELEC = 2, ORC_TERM_GRAV = 3, ORC_TERM_RF = 4, ORC_TERM_MAG = 5 };
enum { ORC_THERM_NONE = 0, ORC_THERM_ANDERSEN = 1, ORC_THERM_BERENDSEN = 2,
       ORC_THERM_LANGEVIN = 3, ORC_THERM_NOSEHOOVER = 4 };

typedef struct {
    int32_t kind;     /* ORC_BC_* */
    int32_t pad;
    double b[6];      /* PERIODIC: xmin xmax ymin ymax zmin zmax; CUBIC: b[0] = L */
} orc_bc;

typedef struct {
    int32_t kind;     /* ORC_TERM_* */
    int32_t pad;
    double p[4];      /* LJ: eps sigma R; ELEC: k R; GRAV: G; RF: k R eps_rf; MAG: mu/4pi */
} orc_term;

typedef struct {
    int64_t n;           /* coordinate columns (atoms) */
    int32_t water;       /* 0: PotentialNBodySystem, 1: WaterSPCFw (n = 3 * molecules) */
    int32_t nterms;      /* potential terms, applied in this order (PotentialNBodySystem) */
    const double* ms;    /* n masses */
    const double* qs;    /* n charges (may be NULL without ELEC/RF) */
    orc_bc bc;
    orc_term terms[8];
    /* WaterSPCFw: lj term, electrostatic term (ELEC or RF) and SPC/Fw parameters */
    orc_term water_lj, water_elec;
    double rOH, aHOH, kbond, kangle;
    /* thermostat */
    int32_t therm;       /* ORC_THERM_* */
    int32_t pad2;
    double therm_T, therm_p;  /* Andersen: nu; Berendsen, Nose-Hoover: tau; Langevin: gamma */
    double kb;
    const double* mm;    /* n x 3 magnetic moments (MAG), or NULL */
    double nh_zeta;      /* Nose-Hoover zeta at the start (the reference's u[3n+1] = 0) */
} orc_system;

/* src/boundary_conditions.jl:111-172 */
void orc_interparticle_distance(const double* ri, const double* rj, const orc_bc* bc,
                                double* rij, double* r, double* r2);

/* src/basic_potentials.jl:240-272 (adds into dv[0..2]) */
void orc_lj_accel(double* dv, const double* rs, int64_t i, const int64_t* indxs,
                  int64_t nidx, const double* ms, double eps, double sigma, double R,
                  const orc_bc* bc);
/* src/basic_potentials.jl:274-304; water != 0 excludes the own molecule
 * (src/nbody_to_ode.jl:331-351), otherwise only j == i (:316-329) */
void orc_elec_accel(double* dv, const double* rs, int64_t i, int64_t n, const double* qs,
                    const double* ms, int32_t water, double k, double R, const orc_bc* bc);
/* reaction-field electrostatics (no counterpart in the reference) */
void orc_rf_accel(double* dv, const double* rs, int64_t i, int64_t n, const double* qs,
                  const double* ms, int32_t water, double k, double R, double eps_rf,
                  const orc_bc* bc);
/* src/basic_potentials.jl:306-331 */
void orc_grav_accel(double* dv, const double* rs, int64_t i, int64_t n, const double* ms,
                    double G);
/* src/basic_potentials.jl:367-393 with neighbourhoods of src/nbody_to_ode.jl:263-288 */
void orc_bond_accel(double* dv, const double* rs, int64_t i, const double* ms, double rOH,
                    double kbond);
/* src/basic_potentials.jl:395-433 (a = H1, b = O, c = H2; adds into dv columns) */
void orc_angle_accel(double* dv, const double* rs, int64_t a, int64_t b, int64_t c,
                     const double* ms, double aHOH, double ka);
/* src/thermostats.jl:87-91 */
double orc_md_temperature(const double* vs, const double* ms, double kb, int64_t N,
                          int64_t Nc);
/* src/thermostats.jl:76-83 */
void orc_berendsen(double* dv, const double* v, const double* ms, double kb, int64_t N,
                   int64_t Nc, double T0, double gamma);

/* src/nbody_to_ode.jl:474-488 (potential systems) and :502-532 (water):
 * dv = sum of closures per atom in the reference order, then thermostat terms. */
void orc_rhs(const orc_system* s, const double* u, const double* v, double* dv);

/* velocity Verlet (OrdinaryDiffEqSymplecticRK form, see .c) for nsteps;
 * u, v updated in place; frames (may be NULL) receive every save_every-th state,
 * starting with the initial one. Returns the number of frames written. */
int64_t orc_vv_run(const orc_system* s, double* u, double* v, double dt, int64_t nsteps,
                   uint64_t seed, int64_t save_every, double* u_frames, double* v_frames);
/* orc_vv_run that also returns the Nose-Hoover zeta and its rate xi at the end of the run
 * (nh_out[2], may be NULL): the reference's u[3n+1], v[3n+1] */
int64_t orc_vv_run_nh(const orc_system* s, double* u, double* v, double dt, int64_t nsteps,
                      uint64_t seed, int64_t save_every, double* u_frames, double* v_frames,
                      double* nh_out);

/* Euler-Maruyama run of the Langevin SDE of a potential system (src/nbody_to_ode.jl:
 * 567-598 solved by StochasticDiffEq EM): with a = f(u) (the potential accelerations),
 * u' = u + dt v and v' = v + dt (a - gamma v) + sigma (sqrt(dt) xi), sigma =
 * sqrt(2 gamma kb T / m_1); gamma = therm_p, T = therm_T.  xi from the GPU's Philox
 * stream 3 with counter (atom, step), steps numbered from 1.  Frames as orc_vv_run. */
int64_t orc_em_run(const orc_system* s, double* u, double* v, double dt, int64_t nsteps,
                   uint64_t seed, int64_t save_every, double* u_frames, double* v_frames);

/* src/nbody_simulation_result.jl:209-218 */
double orc_kinetic_energy(const double* v, const double* ms, int64_t n);
/* src/nbody_simulation_result.jl:293-319 */
double orc_lj_potential(const double* rs, const int64_t* indxs, int64_t nidx, double eps,
                        double sigma, double R, const orc_bc* bc);
/* src/nbody_simulation_result.jl:321-351 */
double orc_elec_potential(const double* rs, int64_t n, const double* qs, int32_t water,
                          double k, double R, const orc_bc* bc);
double orc_rf_potential(const double* rs, int64_t n, const double* qs, int32_t water,
                        double k, double R, double eps_rf, const orc_bc* bc);
/* gravitational potential energy (the reference has none) */
double orc_grav_potential(const double* rs, int64_t n, const double* ms, double G);
/* src/nbody_simulation_result.jl:353-372 */
double orc_bond_potential(const double* rs, int64_t nmol, double rOH, double kbond);
/* src/nbody_simulation_result.jl:374-397 */
double orc_angle_potential(const double* rs, int64_t nmol, double aHOH, double ka);
/* src/nbody_simulation_result.jl:239-291 */
double orc_potential_energy(const orc_system* s, const double* u);

/* src/nbody_simulation_result.jl:664-709; frames: nframes x (3 x n) */
void orc_rdf(const double* frames, int64_t nframes, int64_t n, const int64_t* indxs,
             int64_t nidx, double L, int64_t maxbin, double* rs, double* gr);

/* src/basic_potentials.jl:333-365 (adds into dv[0..2]); mm: n x 3 moments.  The reference
 * takes r_i - r_j without boundary conditions. */
void orc_mag_accel(double* dv, const double* rs, int64_t i, int64_t n, const double* ms,
                   const double* mm, double mu_4pi);
/* magnetic dipole-dipole energy mu/4pi sum_{i<j} [m_i.m_j - 3 (m_i.r)(m_j.r)] / |r_ij|^3
 * (r = unit r_ij; the reference has no such energy function) */
double orc_mag_potential(const double* rs, int64_t n, const double* mm, double mu_4pi);

/* Learned-stride descriptor (no reference counterpart): for every atom of a cubic box of
 * side L, 32 Gaussian radial bins (centres 0.8 + 1.7 b / 31, standard deviation 0.1,
 * cosine switch 1 -> 0 over r in [2.0, 2.5), nothing beyond 2.5) over all other atoms at
 * the minimum-image distance, then v.e1, |v - (v.e1) e1|, |a| in the force frame
 * e1 = a/|a| (x axis if a = 0), e2 = unit part of v normal to e1 (if v is parallel to
 * e1: the axis least aligned with e1, made normal to it).  u, a, v: n x 3 rows;
 * feats: n x 35; frames: n x 6 (e1, e2). */
void orc_lr_features(const double* u, const double* a, const double* v, int64_t n, double L,
                     double* feats, double* frames);

/* Philox4x32-10 (Salmon et al., SC'11) and the derived samplers the GPU uses */
void orc_philox4x32(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                    uint32_t k1, uint32_t* out4);
double orc_u01(uint32_t a, uint32_t b);

#ifdef __cplusplus
}
#endif
#endif
