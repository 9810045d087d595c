// nbx_mixed.h -- parameter blocks and per-atom device steps of the mixed-precision path
// (FP32 pair forces, FP64 state) shared by nbx_nonbonded.hip, nbx_gravity.hip and
// nbx_mixed.cpp.
//
// Step k of a run, with x_k in buffer X[k & 1] and its wrapped copy in P[k & 1]:
//   build(k)  : on build-opportunity steps, if req: Verlet list from the bins B; the last
//               block to finish clears req (and, with per-step bins, the bins)
//   step(k)   : a_k = f(x_k, vh_k); v_k = vh_k + dt a_k / 2 (+ Langevin B half step,
//               Andersen collisions); then the start of step k+1: Langevin A half step
//               (registers only), x_{k+1} = x_k + dt v + dt^2 a_k / 2 into X[(k+1) & 1],
//               vh_{k+1} = v + dt a_k / 2, wrapped copy into P[(k+1) & 1], skin test
//               against the last build -> req (and the stale-list flag when step k+1 has
//               no build opportunity), and the Berendsen temperature of vh_{k+1}
//               (last-block reduction)
//   bin(k+1)  : before a build opportunity, if req: P[(k+1) & 1] binned into B
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

// Thermostat / step parameters
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
    const int64_t* step_base;  // if set, step = *step_base + step_off (graph replays)
    int64_t step_off;
    const double* d_T;    // Berendsen: temperature of the half-kicked velocities
    // if set (water, Berendsen): the molecule kernel adds the block partials of m vh^2 of
    // the drift before it (an int count at T_parts[0], partials from T_parts[1]) and
    // divides by T_denom itself, instead of reading d_T
    const double* T_parts;
    double T_denom;
    // Euler-Maruyama steps of the Langevin SDE (src/nbody_to_ode.jl:567-598) instead of
    // velocity Verlet: x' = x + dt v, v' = v + dt (a - gamma v) + sigma sqrt(dt) xi
    int em;
    double em_gamma, em_sigma, em_sqrt_dt;
    // slot -> user atom index when the state is kept in a sorted order (random numbers are
    // drawn for the user's atom), nullptr for the identity
    const int* ids;
};

__device__ __forceinline__ int64_t mx_atom_id(const mx_kick_params& kp, int64_t i) {
    return kp.ids != nullptr ? (int64_t)kp.ids[i] : i;
}

// Nose-Hoover update at the end of a temperature reduction (src/thermostats.jl:116-128):
// zeta += dt xi (advance), then xi = (1/tau)^2 (T/T0 - (ndf+1)/ndf)
struct mx_nh_params {
    int nh, advance;
    double dt, T0, tau;
    int64_t ndf;
};

// Buffers of one fused step launch
struct mx_io {
    const double* x_in;   // x_k (SoA, unwrapped)
    double* x_out;        // x_{k+1}
    void* pos_out;        // wrapped copy of x_{k+1}: double4 (+q) or float4 (+m wscale)
    int pos_kind;         // 0 none, 1 double4 wrapped (+ fixed-point copy), 2 float4
    double wscale;        // pos_kind 2: factor on the mass in .w (gravity: 1 / m_max)
    double* a;            // a_k out
    double* v;            // v_k out
    double* vh;           // in: vh_k, out: vh_{k+1}
    const double* ms;
    const double* qs;
    int do_drift;         // 0: finish the step only (no x_{k+1})
    // Verlet-list bookkeeping (list mode)
    const double* xref;   // positions at the last build
    double half_skin2;    // squared displacement that requests a rebuild: (skin/2)^2 with a
                          // build opportunity at every step, less with fewer
    double stale2;        // (skin/2)^2: past it the list no longer holds every pair
    int next_opp;         // the next step has a build opportunity
    double guard2;        // NBX_BINS_BUILD: (3 skin/4)^2, past which the bins of the last
                          // build may miss pairs (flagged as an overflow)
    int* req_out;         // rebuild request for step k+1
    int* flags;           // [1] overflow
    double L;             // cubic box side (wrapped copy)
    double invL;          // 1 / L (the fixed-point copy, mx_posq)
    // Berendsen temperature of vh_{k+1}
    int need_T;
    double* red_part;     // one partial per block
    int* red_count;       // number of partials of the last step kernel
    double* T_out;
    double T_denom;       // kb (3N - Nc)
    mx_nh_params nh;      // Nose-Hoover update applied with the temperature
};

// resolve the absolute step number inside a kernel (captured graphs are replayed with a
// device-side step base that the last node of each replay advances)
__device__ __forceinline__ void mx_resolve_step(mx_kick_params& kp) {
    if (kp.step_base != nullptr) kp.step = *kp.step_base + kp.step_off;
}

// Per-atom inputs of the end of a step, loaded together (mx_load_atom) so that a kernel
// can issue them before its force loop: after the loop they would come in several
// dependent rounds, since the step's own stores may alias them.
struct mx_atom {
    double m, q;          // mass, charge (0 without charges)
    double vh[3];         // vh_k
    double x[3];          // x_k (when drifting)
    double xr[3];         // positions at the last list build (list mode, when drifting)
};

__device__ __forceinline__ void mx_load_atom(int64_t i, int64_t n, const mx_io& io,
                                             mx_atom& A) {
    A.m = io.ms[i];
    A.q = io.qs != nullptr ? io.qs[i] : 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) A.vh[k] = io.vh[k * n + i];
    if (io.do_drift) {
#pragma unroll
        for (int k = 0; k < 3; ++k) A.x[k] = io.x_in[k * n + i];
        if (io.xref != nullptr) {
#pragma unroll
            for (int k = 0; k < 3; ++k) A.xr[k] = io.xref[k * n + i];
        }
    }
}

// v = vh + dt (a/2), then the Langevin B half-step of step kp.step (Euler-Maruyama:
// v = vh, the velocity the previous drift computed)
__device__ __forceinline__ void mx_kick_atom(int64_t i, double ax, double ay, double az,
                                             const mx_atom& A, const mx_kick_params& kp,
                                             double* v3) {
    if (kp.em) {  // Euler-Maruyama: vh already holds v of the step being finished
        v3[0] = A.vh[0];
        v3[1] = A.vh[1];
        v3[2] = A.vh[2];
        return;
    }
    v3[0] = A.vh[0] + kp.dt * (0.5 * ax);
    v3[1] = A.vh[1] + kp.dt * (0.5 * ay);
    v3[2] = A.vh[2] + kp.dt * (0.5 * az);
    if (kp.therm == NBX_THERM_LANGEVIN) {
        const double sig = kp.c2 * sqrt(kp.kb * kp.T / A.m);
        double g[3];
        nbx_rng_normals3(kp.seed, mx_atom_id(kp, i), kp.step, NBX_RNG_LANGEVIN_B, g);
        for (int k = 0; k < 3; ++k) v3[k] = kp.c1 * v3[k] + sig * g[k];
    }
}

// Andersen resampling of one atom (src/nbody_simulation_result.jl:523-540)
__device__ __forceinline__ void mx_andersen_atom(int64_t i, double m, const mx_kick_params& kp,
                                                 double* v3) {
    const double vdev = sqrt(kp.kb * kp.T / m);
    double g[3];
    nbx_rng_normals3(kp.seed, mx_atom_id(kp, i), kp.step, NBX_RNG_ANDERSEN, g);
    for (int k = 0; k < 3; ++k) v3[k] = vdev * g[k];
}

// Workgroups of a launch are dispatched round-robin over the 8 XCDs of the chip, each with
// its own L2.  This maps dispatch slot b to a logical block so that XCD x runs the x-th
// contiguous eighth of the logical blocks: atoms that are stored near each other in space
// then share one L2 for the positions they gather.  A bijection on [0, nb).
#define NBX_XCDS 8
__device__ __forceinline__ int mx_xcd_block(int b, int nb) {
    const int q = nb / NBX_XCDS, r = nb % NBX_XCDS;
    const int x = b % NBX_XCDS, s = b / NBX_XCDS;
    return x * q + (x < r ? x : r) + s;
}

__device__ __forceinline__ int mx_cell_coord(double w, double inv_cs, int nc) {
    int cx = (int)(w * inv_cs);
    return cx < 0 ? 0 : (cx >= nc ? nc - 1 : cx);
}

__device__ __forceinline__ double mx_wrap(double p, double L) {
    double w = p - L * floor(p / L);
    return w >= L ? w - L : w;
}

// Verlet-list entries: 16-bit atom indices when the system has at most 65,536 atoms (half
// the list bytes), 32-bit otherwise; `wide` is uniform over a launch
__device__ __forceinline__ int mx_nbr_load(const void* nbr, int wide, size_t k) {
    return wide ? static_cast<const int*>(nbr)[k]
                : (int)static_cast<const uint16_t*>(nbr)[k];
}
__device__ __forceinline__ void mx_nbr_store(void* nbr, int wide, size_t k, int j) {
    if (wide) static_cast<int*>(nbr)[k] = j;
    else static_cast<uint16_t*>(nbr)[k] = (uint16_t)j;
}

// Fixed-point copy of the wrapped positions (list plans): the box fraction of each
// coordinate as an unsigned 32-bit integer and the charge as FP32 bits, stored after the n
// double4 of a position buffer.  The difference of two such words, read as a signed
// integer, is the minimum-image displacement in units of L / 2^32 (exact: the wrap-around
// of the integer subtraction is the periodic image), and a gather moves 16 bytes, not 32.
__device__ __forceinline__ uint4* mx_posq(void* pos, int64_t n) {
    return reinterpret_cast<uint4*>(static_cast<double4*>(pos) + n);
}
__device__ __forceinline__ const uint4* mx_posq(const void* pos, int64_t n) {
    return reinterpret_cast<const uint4*>(static_cast<const double4*>(pos) + n);
}
// box fraction x / L - floor(x / L) in [0, 1) (the subtraction is exact) times 2^32
__device__ __forceinline__ uint32_t mx_frac32(double x, double invL) {
    double t = x * invL;
    t -= floor(t);
    return (uint32_t)(uint64_t)(t * 4294967296.0);
}
__device__ __forceinline__ uint4 mx_fixed_position(double px, double py, double pz,
                                                   double invL, double q) {
    return make_uint4(mx_frac32(px, invL), mx_frac32(py, invL), mx_frac32(pz, invL),
                      __float_as_uint((float)q));
}

// wrapped copy of x_{k+1} and the skin test against the positions of the last build
__device__ __forceinline__ void mx_store_position(int64_t i, int64_t n, double px, double py,
                                                  double pz, const mx_atom& A,
                                                  const mx_io& io) {
    if (io.pos_kind == 1) {
        const double wx = mx_wrap(px, io.L), wy = mx_wrap(py, io.L), wz = mx_wrap(pz, io.L);
        ((double4*)io.pos_out)[i] = make_double4(wx, wy, wz, A.q);
        mx_posq(io.pos_out, n)[i] = mx_fixed_position(px, py, pz, io.invL, A.q);
    } else if (io.pos_kind == 2) {
        ((float4*)io.pos_out)[i] = make_float4((float)px, (float)py, (float)pz,
                                               (float)(A.m * io.wscale));
    }
    if (io.xref != nullptr) {
        const double ex = px - A.xr[0], ey = py - A.xr[1], ez = pz - A.xr[2];
        const double e2 = ex * ex + ey * ey + ez * ez;
        if (e2 > io.half_skin2) io.req_out[0] = 1;
        if (e2 > io.guard2) io.flags[1] = 1;
        if (!io.next_opp && e2 > io.stale2) io.flags[7] = 1;
    }
}

// Euler-Maruyama step kp.step + 1 of the Langevin SDE (src/nbody_to_ode.jl:575-596) from
// (x, v, a = f(x)): x' = x + dt v and v' = v + dt (a - gamma v) + sigma (sqrt(dt) xi),
// sigma = sqrt(2 gamma kb T / m_1) for every atom; v' goes to vh (the next step's kick
// copies it to v).  Returns 0 (no thermostat temperature is needed).
__device__ __forceinline__ double mx_em_atom(int64_t i, int64_t n, double vx, double vy,
                                             double vz, double ax, double ay, double az,
                                             const mx_atom& A, const mx_kick_params& kp,
                                             const mx_io& io) {
    const double dt = kp.dt;
    const double px = A.x[0] + dt * vx;
    const double py = A.x[1] + dt * vy;
    const double pz = A.x[2] + dt * vz;
    io.x_out[i] = px;
    io.x_out[n + i] = py;
    io.x_out[2 * n + i] = pz;
    double g[3];
    nbx_rng_normals3(kp.seed, mx_atom_id(kp, i), kp.step + 1, NBX_RNG_EM, g);
    io.vh[i] = vx + dt * (ax - kp.em_gamma * vx) + kp.em_sigma * (kp.em_sqrt_dt * g[0]);
    io.vh[n + i] = vy + dt * (ay - kp.em_gamma * vy) + kp.em_sigma * (kp.em_sqrt_dt * g[1]);
    io.vh[2 * n + i] = vz + dt * (az - kp.em_gamma * vz) + kp.em_sigma * (kp.em_sqrt_dt * g[2]);
    mx_store_position(i, n, px, py, pz, A, io);
    return 0.0;
}

// Start of step kp.step + 1 for atom i from v (after the previous step's end) and a:
// Langevin A half-step in registers, drift, half kick, wrapped copy, skin test.
// Returns m vh^2 (the Berendsen temperature input).
__device__ __forceinline__ double mx_drift_atom(int64_t i, int64_t n, const double* v3,
                                                double ax, double ay, double az,
                                                const mx_atom& A, const mx_kick_params& kp,
                                                const mx_io& io) {
    double vx = v3[0], vy = v3[1], vz = v3[2];
    const double mi = A.m;
    if (kp.em) return mx_em_atom(i, n, vx, vy, vz, ax, ay, az, A, kp, io);
    if (kp.therm == NBX_THERM_LANGEVIN) {
        const double sig = kp.c2 * sqrt(kp.kb * kp.T / mi);
        double g[3];
        nbx_rng_normals3(kp.seed, mx_atom_id(kp, i), kp.step + 1, NBX_RNG_LANGEVIN_A, g);
        vx = kp.c1 * vx + sig * g[0];
        vy = kp.c1 * vy + sig * g[1];
        vz = kp.c1 * vz + sig * g[2];
    }
    const double dt = kp.dt, dtsq = dt * dt;
    const double px = A.x[0] + dt * vx + dtsq * (0.5 * ax);
    const double py = A.x[1] + dt * vy + dtsq * (0.5 * ay);
    const double pz = A.x[2] + dt * vz + dtsq * (0.5 * az);
    io.x_out[i] = px;
    io.x_out[n + i] = py;
    io.x_out[2 * n + i] = pz;
    const double hx = vx + dt * (0.5 * ax), hy = vy + dt * (0.5 * ay), hz = vz + dt * (0.5 * az);
    io.vh[i] = hx;
    io.vh[n + i] = hy;
    io.vh[2 * n + i] = hz;
    mx_store_position(i, n, px, py, pz, A, io);
    return mi * (hx * hx + hy * hy + hz * hz);
}

// Thermostat force a += c vh: Berendsen c = gamma (T0/T - 1) (src/thermostats.jl:76-83),
// Nose-Hoover c = -zeta (src/thermostats.jl:116-122)
__device__ __forceinline__ bool mx_therm_force(const mx_kick_params& kp) {
    return kp.therm == NBX_THERM_BERENDSEN || kp.therm == NBX_THERM_NOSEHOOVER;
}

__device__ __forceinline__ double mx_berendsen_from_T(const mx_kick_params& kp, double T) {
    if (1.0 / T == __builtin_inf()) return kp.ber_gamma;
    return kp.ber_gamma * (kp.T / T - 1.0);
}

__device__ __forceinline__ double mx_berendsen_coeff(const mx_kick_params& kp) {
    if (kp.therm == NBX_THERM_NOSEHOOVER) return -kp.d_T[NBX_SCAL_ZETA];
    return mx_berendsen_from_T(kp, kp.d_T[0]);
}

// Epilogue of a step kernel, called by every thread of every block (blockDim.x <= 256):
// with io.need_T, the block sum of `val` (m vh^2) goes to red_part[blockIdx.x] and block 0
// records the number of blocks in *red_count.  mx_launch_temperature_finish, launched next
// on the stream, adds the partials in block order, writes the temperature of vh_{k+1} and
// applies the Nose-Hoover update; with Berendsen, the water molecule kernel of the next
// step adds them itself instead (kp.T_parts).  Not a last-block reduction:
// the kernel boundary makes every partial visible, where a device-scope fence per block
// (writing back its XCD's L2) cost more than the launch.  Deterministic: the sum does not
// depend on the order in which blocks finish.
__device__ __forceinline__ void mx_epilogue(double val, const mx_io& io, int n) {
    (void)n;
    if (!io.need_T) return;
    __shared__ double red[256];
    const int t = threadIdx.x;
    const int nt = (int)blockDim.x;
    red[t] = val;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (t < off && t + off < nt) red[t] += red[t + off];
        __syncthreads();
    }
    if (t == 0) {
        io.red_part[blockIdx.x] = red[0];
        if (blockIdx.x == 0) *io.red_count = (int)gridDim.x;
    }
}

// The same with one partial per wavefront of 64 consecutive atoms (grid thread g is atom g,
// blockDim.x a multiple of 64; count ceil(n / 64)), summed by a butterfly: the partials do
// not depend on the kernel's block size, so a drift redone at the start of a run (the drift
// kernel) gives the temperature that the step kernel's drift gave, bit for bit
__device__ __forceinline__ void mx_epilogue_waves(double val, const mx_io& io, int n) {
    if (!io.need_T) return;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) val += __shfl_xor(val, off, 64);
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if ((g & 63) == 0 && g < n) io.red_part[g >> 6] = val;
    if (g == 0) *io.red_count = (n + 63) >> 6;
}

// launchers (nbx_nonbonded.hip / nbx_gravity.hip)
struct mx_nb_launch {
    int mode_list;        // 1: Verlet list, 0: all pairs
    int lanes;
    int has_lj, el_kind;  // el_kind: 0 none, 1 Coulomb, 2 reaction field
    int water;
};
// non-bonded forces at P (list mode) or io.x_in (all pairs); potential systems finish
// the step (kp.do_kick) and drift (io.do_drift); water stores a_nb only.
void mx_launch_nonbonded(hipStream_t s, const mx_nb_launch& L, const nbx_ctx* c,
                         const void* pos_in, const mx_pair_params& pp,
                         const mx_kick_params& kp, const mx_io& io);
void mx_launch_water_molecules(hipStream_t s, const nbx_ctx* c, const mx_kick_params& kp,
                               const mx_io& io);
// start of step kp.step + 1 from (x_in, v, a): drift of every atom
void mx_launch_drift(hipStream_t s, const nbx_ctx* c, const mx_kick_params& kp,
                     const mx_io& io);
// with io.need_T: the temperature of vh_{k+1} from the block partials of the step kernel
// just launched (and the Nose-Hoover update)
void mx_launch_temperature_finish(hipStream_t s, const mx_io& io);
// dst[k * n + s] = src[k * n + map[s]] for k < comps (map: a permutation of [0, n))
// device-side sort of the state's slots by cell (nbx_sort.hip): scratch bytes of the radix
// sort for n slots on a g^3 grid; the keys (cell * n + user index) and slots of the
// current order with the number of descents of the key sequence; and the sort itself
// (map[t] = old slot of new slot t; perm_new / iperm_new = the new slot <-> user maps)
size_t mx_sort_temp_bytes(int64_t n, int g);
int mx_launch_sort_keys(hipStream_t s, const double* x, int64_t n, double L, int g,
                        const int* perm, unsigned long long* keys, int* vals,
                        unsigned* descents);
int mx_launch_sort_apply(hipStream_t s, int64_t n, int g, unsigned long long* keys_in,
                         unsigned long long* keys_out, int* vals_in, int* map, void* temp,
                         size_t temp_bytes, const int* perm, int* perm_new, int* iperm_new);
void mx_launch_iota(hipStream_t s, int* p, int64_t n);
// arrays of n atoms reordered together: n3 of 3 SoA components, n1 of one
struct mx_perm_arrays {
    double* a3[5];
    double* a1[2];
    int n3, n1;
};
// new[s] = old[map[s]] for every array, through tmp ((3 n3 + n1) n doubles)
void mx_launch_permute(hipStream_t s, const int* map, int64_t n, const mx_perm_arrays& A,
                       double* tmp);
void mx_launch_gather(hipStream_t s, const double* src, double* dst, const int* map,
                      int64_t n, int comps);
// Verlet list (gated by req) from bins B; its last block clears req and, with
// clear_bins, the counts of B (NBX_BINS_BUILD: B is the base of the three rotating sets,
// which the build advances)
void mx_launch_build(hipStream_t s, const nbx_ctx* c, int water, const void* pos,
                     const double* x, int* req, int* bin_count, int* bin_atoms,
                     int clear_bins);
// wrapped copy of x into pos, binned into bins; sets *req
void mx_launch_positions(hipStream_t s, const nbx_ctx* c, const double* x, void* pos,
                         int pos_kind, double wscale, int* bin_count, int* bin_atoms, int* req);
void mx_launch_temperature(hipStream_t s, const nbx_ctx* c, const double* v, int64_t Nc,
                           double* d_T);
void mx_launch_count_pairs(hipStream_t s, const nbx_ctx* c, const mx_nb_launch& L,
                           const void* pos, const mx_pair_params& pp,
                           unsigned long long* d_count);
void mx_launch_gravity(hipStream_t s, const nbx_ctx* c, const void* pos, double G,
                       const mx_kick_params& kp, const mx_io& io, int force_only_kernel,
                       hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
// all-pairs magnetic dipoles: force into c->d_part (one chunk), then k_grav_finish
void mx_launch_magnetic(hipStream_t s, const nbx_ctx* c, const void* pos, double mu_4pi,
                        const mx_kick_params& kp, const mx_io& io, int force_only_kernel);
void mx_launch_step_advance(hipStream_t s, int64_t* step_base, int64_t delta, int set);
void mx_launch_fill_int(hipStream_t s, int* p, int64_t count, int value);
// bins of P (gated by *req; the bins were cleared by the build kernel that used them last)
void mx_launch_bin(hipStream_t s, const nbx_ctx* c, const void* pos, const int* req,
                   int* bin_count, int* bin_atoms);
// T = (sum of nparts per-block partials, in order) / denom into scal[0]; with a
// Nose-Hoover thermostat (nh != 0) also zeta += dt xi (advance != 0) and then
// xi = (1/tau)^2 (T/T0 - (ndf+1)/ndf)
void mx_launch_nh_rate(hipStream_t s, double* scal, const mx_nh_params& nh);
