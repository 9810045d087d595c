// nbx_mixed.h -- parameter blocks of the mixed-precision kernels (FP32 pair forces,
// FP64 state) shared by nbx_nonbonded.hip, nbx_gravity.hip and nbx_mixed.cpp.
#pragma once
#include "nbx_internal.h"

// Non-bonded pair law parameters (Lennard-Jones + Coulomb / reaction field)
struct mx_pair_params {
    float lj_s2;      // sigma^2
    float lj_R2;      // LJ cutoff^2 (FP32)
    float el_R2;      // electrostatic cutoff^2 (FP32)
    float krf2;       // 2 k_rf (reaction field), 0 for plain Coulomb
    double lj_c;      // 24 eps     (per-atom coefficient = lj_c / m_i)
    double el_k;      // k          (per-atom coefficient = k q_i / m_i)
    double L, invL;   // cubic box side (list mode)
};

// End-of-step work fused into the force kernels (one thread per atom / molecule)
struct mx_kick_params {
    int do_kick;          // 0: force evaluation only (prologue / timing)
    int therm;            // NBX_THERM_*
    double dt;
    double kb, T;         // thermostat target
    double c1, c2;        // Langevin half-step OU coefficients
    double prob;          // Andersen collision probability nu*dt
    double ber_gamma;     // Berendsen 1/(2 tau)
    double m0;            // first body's mass (Andersen velocity scale, potential systems)
    uint64_t seed;
    int64_t step;         // absolute step number of the step being finished
    const double* d_T;    // Berendsen: temperature of the half-kicked velocities
};

// velocity update of one atom at the end of a step: v = vh + dt (a/2), then the
// Langevin B half-step; Andersen is applied by the caller (per body).
__device__ __forceinline__ void mx_finish_atom(int64_t i, int64_t n, double ax, double ay,
                                               double az, const double* __restrict__ vh,
                                               double* __restrict__ v,
                                               const double* __restrict__ ms,
                                               const mx_kick_params& kp) {
    double vx = vh[i] + kp.dt * (0.5 * ax);
    double vy = vh[n + i] + kp.dt * (0.5 * ay);
    double vz = vh[2 * n + i] + kp.dt * (0.5 * az);
    if (kp.therm == NBX_THERM_LANGEVIN) {
        double sig = kp.c2 * sqrt(kp.kb * kp.T / ms[i]);
        double g[3];
        nbx_rng_normals3(kp.seed, i, kp.step, NBX_RNG_LANGEVIN_B, g);
        vx = kp.c1 * vx + sig * g[0];
        vy = kp.c1 * vy + sig * g[1];
        vz = kp.c1 * vz + sig * g[2];
    }
    v[i] = vx;
    v[n + i] = vy;
    v[2 * n + i] = vz;
}

__device__ __forceinline__ void mx_andersen_atom(int64_t i, int64_t n, double m,
                                                 double* __restrict__ v,
                                                 const mx_kick_params& kp) {
    const double vdev = sqrt(kp.kb * kp.T / m);
    double g[3];
    nbx_rng_normals3(kp.seed, i, kp.step, NBX_RNG_ANDERSEN, g);
    v[i] = vdev * g[0];
    v[n + i] = vdev * g[1];
    v[2 * n + i] = vdev * g[2];
}

// Berendsen term gamma (T0/T - 1) vh (src/thermostats.jl:76-83)
__device__ __forceinline__ double mx_berendsen_coeff(const mx_kick_params& kp) {
    const double T = kp.d_T[0];
    if (1.0 / T == __builtin_inf()) return kp.ber_gamma;
    return kp.ber_gamma * (kp.T / T - 1.0);
}

// launchers (nbx_nonbonded.hip / nbx_gravity.hip)
struct mx_nb_launch {
    int mode_list;        // 1: Verlet list, 0: all pairs
    int lanes;
    int has_lj, el_kind;  // el_kind: 0 none, 1 Coulomb, 2 reaction field
    int water;
};
void mx_launch_nonbonded(hipStream_t s, const mx_nb_launch& L, const nbx_ctx* c,
                         const mx_pair_params& pp, const mx_kick_params& kp,
                         int* clear_flag);
void mx_launch_drift(hipStream_t s, const nbx_ctx* c, double dt, int therm, double c1,
                     double c2, double T, double kb, uint64_t seed, int64_t step,
                     int write_pos_kind, double half_skin2);
void mx_launch_rebuild(hipStream_t s, const nbx_ctx* c, int water);
void mx_launch_water_molecules(hipStream_t s, const nbx_ctx* c, const mx_kick_params& kp);
void mx_launch_temperature(hipStream_t s, const nbx_ctx* c, const double* v, int64_t Nc,
                           double* d_T);
void mx_launch_positions(hipStream_t s, const nbx_ctx* c, int write_pos_kind);
void mx_launch_count_pairs(hipStream_t s, const nbx_ctx* c, const mx_nb_launch& L,
                           const mx_pair_params& pp, unsigned long long* d_count);
void mx_launch_gravity(hipStream_t s, const nbx_ctx* c, double G, const mx_kick_params& kp,
                       int force_only_kernel);
