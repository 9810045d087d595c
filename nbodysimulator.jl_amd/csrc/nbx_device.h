// nbx_device.h -- device helpers shared by the gfx950 kernels: boundary-condition
// displacement (src/boundary_conditions.jl:111-172), Philox4x32-10 counter RNG and
// the samplers built on it.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define NBX_WRAP_PASSES 64

struct nbx_bcv {
    int kind;       // NBX_BC_*
    int pad;
    double b[6];    // PERIODIC: xmin xmax ymin ymax zmin zmax; CUBIC: b[0] = L
};

// while x >= radius: x -= size; while x < -radius: x += size
// (src/boundary_conditions.jl:143-160), capped at NBX_WRAP_PASSES passes, after
// which the remainder is removed with floor() (the reference would keep looping).
__device__ __forceinline__ double nbx_wrap_cubic(double x, double size, double radius) {
    int k = 0;
    while (x >= radius && k < NBX_WRAP_PASSES) { x -= size; ++k; }
    while (x < -radius && k < NBX_WRAP_PASSES) { x += size; ++k; }
    if (k >= NBX_WRAP_PASSES && isfinite(x)) x -= size * floor((x + radius) / size);
    return x;
}

// while x < lo: x += hi - lo; while x >= hi: x -= hi - lo
// (src/boundary_conditions.jl:114-131): wraps into [lo, hi), not minimum image.
__device__ __forceinline__ double nbx_wrap_lohi(double x, double lo, double hi) {
    int k = 0;
    while (x < lo && k < NBX_WRAP_PASSES) { x += hi - lo; ++k; }
    while (x >= hi && k < NBX_WRAP_PASSES) { x -= hi - lo; ++k; }
    if (k >= NBX_WRAP_PASSES && isfinite(x)) x -= (hi - lo) * floor((x - lo) / (hi - lo));
    return x;
}

// get_interparticle_distance(ri, rj, bc): rij = ri - rj wrapped per the boundary
// condition, r2 = (x^2 + y^2) + z^2.  (r = sqrt(r2) is taken by callers that use it.)
__device__ __forceinline__ void nbx_displacement(double xi, double yi, double zi, double xj,
                                                 double yj, double zj, const nbx_bcv& bc,
                                                 double& x, double& y, double& z,
                                                 double& r2) {
    x = xi - xj;
    y = yi - yj;
    z = zi - zj;
    if (bc.kind == 1) {
        const double size = bc.b[0];
        const double radius = 0.5 * size;
        x = nbx_wrap_cubic(x, size, radius);
        y = nbx_wrap_cubic(y, size, radius);
        z = nbx_wrap_cubic(z, size, radius);
    } else if (bc.kind == 2) {
        x = nbx_wrap_lohi(x, bc.b[0], bc.b[1]);
        y = nbx_wrap_lohi(y, bc.b[2], bc.b[3]);
        z = nbx_wrap_lohi(z, bc.b[4], bc.b[5]);
    }
    double s = x * x + y * y;
    r2 = s + z * z;
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as easy as
// 1, 2, 3", SC'11).  Counter (c0..c3), key (k0, k1).
__device__ __forceinline__ void nbx_philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                           uint32_t k0, uint32_t k1, uint32_t* out) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0 = __umulhi(M0, c0), lo0 = M0 * c0;
        uint32_t hi1 = __umulhi(M1, c2), lo1 = M1 * c2;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += W0; k1 += W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// 53-bit uniform in [0, 1) from two words
__device__ __forceinline__ double nbx_u01(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

// RNG streams; counter = (atom, step, stream, sub), key = seed
enum { NBX_RNG_ANDERSEN = 0, NBX_RNG_LANGEVIN_A = 1, NBX_RNG_LANGEVIN_B = 2, NBX_RNG_EM = 3 };

__device__ __forceinline__ double nbx_rng_uniform(uint64_t seed, int64_t atom, int64_t step,
                                                  uint32_t stream) {
    uint32_t w[4];
    nbx_philox((uint32_t)atom, (uint32_t)step, stream, 0u, (uint32_t)seed,
               (uint32_t)(seed >> 32), w);
    return nbx_u01(w[0], w[1]);
}

// three standard normals by Box-Muller from two Philox blocks
__device__ __forceinline__ void nbx_rng_normals3(uint64_t seed, int64_t atom, int64_t step,
                                                 uint32_t stream, double* g) {
    uint32_t w[4];
    nbx_philox((uint32_t)atom, (uint32_t)step, stream, 1u, (uint32_t)seed,
               (uint32_t)(seed >> 32), w);
    double u1 = nbx_u01(w[0], w[1]), u2 = nbx_u01(w[2], w[3]);
    double r = sqrt(-2.0 * log(1.0 - u1));
    double th = 6.283185307179586 * u2;
    g[0] = r * cos(th);
    g[1] = r * sin(th);
    nbx_philox((uint32_t)atom, (uint32_t)step, stream, 2u, (uint32_t)seed,
               (uint32_t)(seed >> 32), w);
    u1 = nbx_u01(w[0], w[1]);
    u2 = nbx_u01(w[2], w[3]);
    r = sqrt(-2.0 * log(1.0 - u1));
    th = 6.283185307179586 * u2;
    g[2] = r * cos(th);
}

// Kahan (compensated) FP32 accumulator, Kahan 1965.  The compiler may not
// reassociate it: the HIP sources are built without fast-math.
struct nbx_kahan {
    float s, c;
    __device__ __forceinline__ void init() { s = 0.0f; c = 0.0f; }
    __device__ __forceinline__ void add(float x) {
        float y = x - c;
        float t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ __forceinline__ double value() const { return (double)s - (double)c; }
};
