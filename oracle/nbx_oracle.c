/*
 * nbx_oracle.c -- FP64 CPU restatement of NBodySimulator.jl's hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): loaded by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 *
 * Every function follows the reference expression by expression, in the same
 * evaluation order, so that with -ffp-contract=off it rounds like Julia does
 * (Julia never fuses a*b+c into an FMA without @fastmath/@muladd, and neither
 * appears in the reference).  Julia literal powers x^2 and x^3 are x*x and
 * x*x*x (Base.literal_pow), and are written out that way here.
 *
 * Deviations that the reference leaves open (parity unpinned at the last bit):
 *  - StaticArrays' norm/normalize: restated as sqrt((x*x+y*y)+z*z) and
 *    inv(norm)*a respectively;
 *  - BLAS ddot inside md_temperature/kinetic_energy: restated as a sequential sum;
 *  - libm acos/log/cos/sin: glibc here, OCML on the GPU;
 *  - the OrdinaryDiffEq VelocityVerlet update (not vendored): restated as
 *      u' = u + dt*v + dt^2*(1/2*a);  vh = v + dt*(1/2*a);
 *      a' = f(u', vh);                v' = vh + dt*(1/2*a').
 *  - the minimum-image while loops are capped at 64 passes; beyond that (an atom
 *    more than 64 box lengths away) the remainder is removed with floor(), where
 *    the reference would keep looping (forever for an infinite coordinate).
 */
#include "nbx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define WRAP_PASSES 64

/* ---------------------------------------------------------------------------
 * Boundary conditions: src/boundary_conditions.jl:111-172
 * ------------------------------------------------------------------------- */

/* while x >= radius: x -= size; while x < -radius: x += size  (:143-160) */
static double wrap_cubic(double x, double size, double radius) {
    int k = 0;
    while (x >= radius && k < WRAP_PASSES) { x -= size; ++k; }
    while (x < -radius && k < WRAP_PASSES) { x += size; ++k; }
    if (k >= WRAP_PASSES && isfinite(x)) {
        x -= size * floor((x + radius) / size);
    }
    return x;
}

/* while x < lo: x += hi - lo; while x >= hi: x -= hi - lo  (:114-131) */
static double wrap_lohi(double x, double lo, double hi) {
    int k = 0;
    while (x < lo && k < WRAP_PASSES) { x += hi - lo; ++k; }
    while (x >= hi && k < WRAP_PASSES) { x -= hi - lo; ++k; }
    if (k >= WRAP_PASSES && isfinite(x)) {
        x -= (hi - lo) * floor((x - lo) / (hi - lo));
    }
    return x;
}

void orc_interparticle_distance(const double* ri, const double* rj, const orc_bc* bc,
                                double* rij, double* r, double* r2) {
    double x = ri[0] - rj[0];
    double y = ri[1] - rj[1];
    double z = ri[2] - rj[2];
    if (bc->kind == ORC_BC_CUBIC) {
        double size = bc->b[0];
        double radius = 0.5 * size;
        x = wrap_cubic(x, size, radius);
        y = wrap_cubic(y, size, radius);
        z = wrap_cubic(z, size, radius);
    } else if (bc->kind == ORC_BC_PERIODIC) {
        x = wrap_lohi(x, bc->b[0], bc->b[1]);
        y = wrap_lohi(y, bc->b[2], bc->b[3]);
        z = wrap_lohi(z, bc->b[4], bc->b[5]);
    }
    rij[0] = x; rij[1] = y; rij[2] = z;
    double s = x * x + y * y;
    s = s + z * z;
    *r2 = s;
    *r = sqrt(s);
}

static double norm3(const double* a) {
    double s = a[0] * a[0] + a[1] * a[1];
    s = s + a[2] * a[2];
    return sqrt(s);
}

/* ---------------------------------------------------------------------------
 * Force laws: src/basic_potentials.jl:240-433
 * ------------------------------------------------------------------------- */

void orc_lj_accel(double* dv, const double* rs, int64_t i, const int64_t* indxs,
                  int64_t nidx, const double* ms, double eps, double sigma, double R,
                  const orc_bc* bc) {
    const double s2 = sigma * sigma, R2 = R * R;   /* LennardJonesParameters :67-69 */
    double f1 = 0.0, f2 = 0.0, f3 = 0.0;
    const double* ri = rs + 3 * i;
    for (int64_t jj = 0; jj < nidx; ++jj) {        /* :253 */
        int64_t j = indxs[jj];
        if (j == i) continue;
        double rij[3], r, rij_2;
        orc_interparticle_distance(ri, rs + 3 * j, bc, rij, &r, &rij_2);
        if (rij_2 < R2) {                          /* :258 */
            double q = s2 / rij_2;
            double s6 = q * q * q;                 /* (σ2/r2)^3 */
            double s12 = s6 * s6;                  /* σ_rij_6^2 */
            double factor = (2.0 * s12 - s6) / rij_2;
            f1 += factor * rij[0];
            f2 += factor * rij[1];
            f3 += factor * rij[2];
        }
    }
    double coeff = 24.0 * eps / ms[i];             /* :268 */
    dv[0] += coeff * f1;
    dv[1] += coeff * f2;
    dv[2] += coeff * f3;
}

static int excluded(int64_t i, int64_t j, int32_t water) {
    if (water) return (i / 3) == (j / 3);          /* src/nbody_to_ode.jl:346-348 */
    return i == j;                                 /* src/nbody_to_ode.jl:326 */
}

void orc_elec_accel(double* dv, const double* rs, int64_t i, int64_t n, const double* qs,
                    const double* ms, int32_t water, double k, double R, const orc_bc* bc) {
    const double R2 = R * R;
    double f1 = 0.0, f2 = 0.0, f3 = 0.0;
    const double* ri = rs + 3 * i;
    for (int64_t j = 0; j < n; ++j) {              /* :288 */
        if (excluded(i, j, water)) continue;       /* :289 */
        double rij[3], r, r2;
        orc_interparticle_distance(ri, rs + 3 * j, bc, rij, &r, &r2);
        if (r2 < R2) {                             /* :292 */
            double factor = qs[j] / (r * r2);
            f1 += factor * rij[0];
            f2 += factor * rij[1];
            f3 += factor * rij[2];
        }
    }
    double coeff = k * qs[i] / ms[i];              /* :300 */
    dv[0] += coeff * f1;
    dv[1] += coeff * f2;
    dv[2] += coeff * f3;
}

/* Reaction field (Tironi et al., J. Chem. Phys. 102, 5451 (1995)):
 *   U_ij = k q_i q_j (1/r + k_rf r^2 - c_rf),  r < R
 *   k_rf = (eps_rf - 1) / ((2 eps_rf + 1) R^3)  (1/(2 R^3) for eps_rf = inf)
 *   c_rf = 1/R + k_rf R^2
 *   a_i  = k q_i / m_i * sum_j q_j (1/r^3 - 2 k_rf) r_ij
 * Excluded pairs as for the truncated Coulomb law. */
static double rf_krf(double R, double eps_rf) {
    if (isinf(eps_rf)) return 1.0 / (2.0 * R * R * R);
    return (eps_rf - 1.0) / ((2.0 * eps_rf + 1.0) * (R * R * R));
}

void orc_rf_accel(double* dv, const double* rs, int64_t i, int64_t n, const double* qs,
                  const double* ms, int32_t water, double k, double R, double eps_rf,
                  const orc_bc* bc) {
    const double R2 = R * R;
    const double krf2 = 2.0 * rf_krf(R, eps_rf);
    double f1 = 0.0, f2 = 0.0, f3 = 0.0;
    const double* ri = rs + 3 * i;
    for (int64_t j = 0; j < n; ++j) {
        if (excluded(i, j, water)) continue;
        double rij[3], r, r2;
        orc_interparticle_distance(ri, rs + 3 * j, bc, rij, &r, &r2);
        if (r2 < R2) {
            double factor = qs[j] * (1.0 / (r * r2) - krf2);
            f1 += factor * rij[0];
            f2 += factor * rij[1];
            f3 += factor * rij[2];
        }
    }
    double coeff = k * qs[i] / ms[i];
    dv[0] += coeff * f1;
    dv[1] += coeff * f2;
    dv[2] += coeff * f3;
}

void orc_grav_accel(double* dv, const double* rs, int64_t i, int64_t n, const double* ms,
                    double G) {
    double a1 = 0.0, a2 = 0.0, a3 = 0.0;
    const double* ri = rs + 3 * i;
    for (int64_t j = 0; j < n; ++j) {              /* :317 */
        if (j == i) continue;
        double rij[3] = {ri[0] - rs[3 * j], ri[1] - rs[3 * j + 1], ri[2] - rs[3 * j + 2]};
        double nr = norm3(rij);
        double factor = -G * ms[j] / (nr * nr * nr);   /* :321 */
        a1 += factor * rij[0];
        a2 += factor * rij[1];
        a3 += factor * rij[2];
    }
    dv[0] += a1;
    dv[1] += a2;
    dv[2] += a3;
}

/* neighbourhoods (src/nbody_to_ode.jl:274-284): O -> (H1, H2), H -> (O) */
void orc_bond_accel(double* dv, const double* rs, int64_t i, const double* ms, double rOH,
                    double kbond) {
    int64_t o = 3 * (i / 3);
    int64_t nb[2];
    int nn;
    if (i == o) { nb[0] = o + 1; nb[1] = o + 2; nn = 2; }
    else { nb[0] = o; nn = 1; }
    double f1 = 0.0, f2 = 0.0, f3 = 0.0;
    const double* ri = rs + 3 * i;
    for (int t = 0; t < nn; ++t) {                 /* :378 */
        const double* rj = rs + 3 * nb[t];
        double rij[3] = {ri[0] - rj[0], ri[1] - rj[1], ri[2] - rj[2]};
        double r = norm3(rij);
        double d = r - rOH;
        double factor = -d * kbond / r;            /* :383 */
        f1 += factor * rij[0];
        f2 += factor * rij[1];
        f3 += factor * rij[2];
    }
    double coeff = 1.0 / ms[i];                    /* :389 */
    dv[0] += coeff * f1;
    dv[1] += coeff * f2;
    dv[2] += coeff * f3;
}

static void cross3(const double* a, const double* b, double* c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

static double dot3(const double* a, const double* b) {
    double s = a[0] * b[0] + a[1] * b[1];
    return s + a[2] * b[2];
}

void orc_angle_accel(double* dv, const double* rs, int64_t a, int64_t b, int64_t c,
                     const double* ms, double aHOH, double ka) {
    const double* ra = rs + 3 * a;
    const double* rb = rs + 3 * b;
    const double* rc = rs + 3 * c;
    double rba[3], rbc[3], rcb[3];
    for (int k = 0; k < 3; ++k) {
        rba[k] = ra[k] - rb[k];
        rbc[k] = rc[k] - rb[k];
        rcb[k] = rb[k] - rc[k];
    }
    double rbaXbc[3], ta[3], tc[3], pa[3], pc[3];
    cross3(rba, rbc, rbaXbc);                      /* :412 */
    cross3(rba, rbaXbc, ta);
    cross3(rcb, rbaXbc, tc);
    double ia = 1.0 / norm3(ta), ic = 1.0 / norm3(tc);
    for (int k = 0; k < 3; ++k) { pa[k] = ia * ta[k]; pc[k] = ic * tc[k]; }
    double nba = norm3(rba), nbc = norm3(rbc);
    double cosine = dot3(rba, rbc) / (nba * nbc);  /* :416 */
    if (cosine > 1.0) cosine = 1.0;
    else if (cosine < -1.0) cosine = -1.0;
    double ang = acos(cosine);
    double force = -ka * (ang - aHOH);             /* :425 */
    for (int k = 0; k < 3; ++k) {
        double fa = pa[k] * force / nba;
        double fc = pc[k] * force / nbc;
        double fb = -(fa + fc);
        dv[3 * a + k] += fa / ms[a];               /* :430-432 */
        dv[3 * b + k] += fb / ms[b];
        dv[3 * c + k] += fc / ms[c];
    }
}

/* ---------------------------------------------------------------------------
 * Thermostats: src/thermostats.jl:72-91
 * ------------------------------------------------------------------------- */

double orc_md_temperature(const double* vs, const double* ms, double kb, int64_t N,
                          int64_t Nc) {
    double e_kin = 0.0;
    for (int64_t i = 0; i < N; ++i) {
        const double* vi = vs + 3 * i;
        double s = vi[0] * vi[0] + vi[1] * vi[1];
        s = s + vi[2] * vi[2];
        e_kin += ms[i] * s;
    }
    return e_kin / (kb * (double)(3 * N - Nc));
}

void orc_berendsen(double* dv, const double* v, const double* ms, double kb, int64_t N,
                   int64_t Nc, double T0, double gamma) {
    double T = orc_md_temperature(v, ms, kb, N, Nc);
    if (1.0 / T == INFINITY) {                     /* inv(T) == Inf */
        for (int64_t k = 0; k < 3 * N; ++k) dv[k] += gamma * v[k];
    } else {
        double c = gamma * (T0 / T - 1.0);
        for (int64_t k = 0; k < 3 * N; ++k) dv[k] += c * v[k];
    }
}

/* ---------------------------------------------------------------------------
 * RHS assembly: src/nbody_to_ode.jl:474-488 and :502-532
 * ------------------------------------------------------------------------- */

static void apply_term(const orc_system* s, const orc_term* t, double* dvi, const double* u,
                       int64_t i, const int64_t* all, int64_t nall) {
    switch (t->kind) {
    case ORC_TERM_LJ:
        orc_lj_accel(dvi, u, i, all, nall, s->ms, t->p[0], t->p[1], t->p[2], &s->bc);
        break;
    case ORC_TERM_ELEC:
        orc_elec_accel(dvi, u, i, s->n, s->qs, s->ms, s->water, t->p[0], t->p[1], &s->bc);
        break;
    case ORC_TERM_RF:
        orc_rf_accel(dvi, u, i, s->n, s->qs, s->ms, s->water, t->p[0], t->p[1], t->p[2],
                     &s->bc);
        break;
    case ORC_TERM_GRAV:
        orc_grav_accel(dvi, u, i, s->n, s->ms, t->p[0]);
        break;
    case ORC_TERM_MAG:
        orc_mag_accel(dvi, u, i, s->n, s->ms, s->mm, t->p[0]);
        break;
    default:
        break;
    }
}

/* Nose-Hoover (src/thermostats.jl:116-122): dv -= zeta v; xi = (1/tau)^2 (T/T0 - (ndf+1)/ndf)
 * with T = md_temperature of v; the reference keeps zeta at u[3n+1] and writes xi into
 * v[3n+1] on every right-hand side. */
static void nosehoover(double* dv, const double* v, const orc_system* s, int64_t N, int64_t Nc,
                       double zeta, double* xi_out) {
    for (int64_t k = 0; k < 3 * N; ++k) dv[k] -= zeta * v[k];
    const double T = orc_md_temperature(v, s->ms, s->kb, N, Nc);
    const int64_t ndf = 3 * N - Nc;
    const double it = 1.0 / s->therm_p;
    if (xi_out) *xi_out = (it * it) * (T / s->therm_T - (double)(ndf + 1) / (double)ndf);
}

static void rhs_impl(const orc_system* s, const double* u, const double* v, double* dv,
                     double zeta, double* xi_out);

void orc_rhs(const orc_system* s, const double* u, const double* v, double* dv) {
    rhs_impl(s, u, v, dv, s->nh_zeta, NULL);
}

static void rhs_impl(const orc_system* s, const double* u, const double* v, double* dv,
                     double zeta, double* xi_out) {
    const int64_t n = s->n;
    if (!s->water) {
        int64_t* all = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
        for (int64_t i = 0; i < n; ++i) all[i] = i;   /* obtain_data_for_lj :290-300 */
        for (int64_t i = 0; i < n; ++i) {
            double* dvi = dv + 3 * i;
            dvi[0] = 0.0; dvi[1] = 0.0; dvi[2] = 0.0;
            for (int t = 0; t < s->nterms; ++t) apply_term(s, &s->terms[t], dvi, u, i, all, n);
        }
        free(all);
        if (s->therm == ORC_THERM_BERENDSEN)
            orc_berendsen(dv, v, s->ms, s->kb, n, 0, s->therm_T, 0.5 / s->therm_p);
        if (s->therm == ORC_THERM_NOSEHOOVER) nosehoover(dv, v, s, n, 0, zeta, xi_out);
        return;
    }
    const int64_t nmol = n / 3;
    int64_t* oidx = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nmol > 0 ? nmol : 1));
    for (int64_t m = 0; m < nmol; ++m) oidx[m] = 3 * m;   /* :302-314 */
    /* O atoms: electrostatic, bond, LJ (:503-512, order :542-553) */
    for (int64_t m = 0; m < nmol; ++m) {
        int64_t i = 3 * m;
        double* dvi = dv + 3 * i;
        dvi[0] = 0.0; dvi[1] = 0.0; dvi[2] = 0.0;
        apply_term(s, &s->water_elec, dvi, u, i, NULL, 0);
        orc_bond_accel(dvi, u, i, s->ms, s->rOH, s->kbond);
        orc_lj_accel(dvi, u, i, oidx, nmol, s->ms, s->water_lj.p[0], s->water_lj.p[1],
                     s->water_lj.p[2], &s->bc);
    }
    /* H atoms: electrostatic, bond (:513-522) */
    for (int64_t m = 0; m < nmol; ++m) {
        for (int64_t h = 1; h <= 2; ++h) {
            int64_t i = 3 * m + h;
            double* dvi = dv + 3 * i;
            dvi[0] = 0.0; dvi[1] = 0.0; dvi[2] = 0.0;
            apply_term(s, &s->water_elec, dvi, u, i, NULL, 0);
            orc_bond_accel(dvi, u, i, s->ms, s->rOH, s->kbond);
        }
    }
    /* angle group per molecule (:523-527; indices :256-258) */
    for (int64_t m = 0; m < nmol; ++m)
        orc_angle_accel(dv, u, 3 * m + 1, 3 * m, 3 * m + 2, s->ms, s->aHOH, s->kangle);
    free(oidx);
    if (s->therm == ORC_THERM_BERENDSEN)             /* :401-408: N = 3n, Nc = 2n */
        orc_berendsen(dv, v, s->ms, s->kb, n, 2 * nmol, s->therm_T, 0.5 / s->therm_p);
    if (s->therm == ORC_THERM_NOSEHOOVER)           /* :425-433: N = 3n, Nc = 2n */
        nosehoover(dv, v, s, n, 2 * nmol, zeta, xi_out);
}

/* ---------------------------------------------------------------------------
 * Counter-based RNG: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)
 * ------------------------------------------------------------------------- */

void orc_philox4x32(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                    uint32_t k1, uint32_t* out4) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)M0 * c0;
        uint64_t p1 = (uint64_t)M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n1 = lo1;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        uint32_t n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += W0; k1 += W1;
    }
    out4[0] = c0; out4[1] = c1; out4[2] = c2; out4[3] = c3;
}

/* 53-bit uniform in [0, 1) from two 32-bit words */
double orc_u01(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

/* Box-Muller pair from one Philox block */
static void normal_pair(const uint32_t* w, double* n0, double* n1) {
    double u1 = orc_u01(w[0], w[1]);
    double u2 = orc_u01(w[2], w[3]);
    double r = sqrt(-2.0 * log(1.0 - u1));
    double th = 6.283185307179586 * u2;
    *n0 = r * cos(th);
    *n1 = r * sin(th);
}

/* RNG streams: c = (atom, step, stream, sub), key = seed */
enum { RNG_ANDERSEN = 0, RNG_LANGEVIN_A = 1, RNG_LANGEVIN_B = 2, RNG_EM = 3 };

static void rng_normals3(uint64_t seed, int64_t atom, int64_t step, uint32_t stream,
                         double* out) {
    uint32_t w[4];
    double spare;
    orc_philox4x32((uint32_t)atom, (uint32_t)step, stream, 1u, (uint32_t)seed,
                   (uint32_t)(seed >> 32), w);
    normal_pair(w, &out[0], &out[1]);
    orc_philox4x32((uint32_t)atom, (uint32_t)step, stream, 2u, (uint32_t)seed,
                   (uint32_t)(seed >> 32), w);
    normal_pair(w, &out[2], &spare);
}

static double rng_uniform(uint64_t seed, int64_t atom, int64_t step, uint32_t stream) {
    uint32_t w[4];
    orc_philox4x32((uint32_t)atom, (uint32_t)step, stream, 0u, (uint32_t)seed,
                   (uint32_t)(seed >> 32), w);
    return orc_u01(w[0], w[1]);
}

/* Andersen collision (src/nbody_simulation_result.jl:511-540): probability nu*dt per
 * body per step.  A potential system resamples one atom with the first body's mass
 * (:527); water resamples all three atoms of a molecule with the O and H masses
 * (:535-539).  The uniform and normals of body b come from counter (atom, step). */
static void andersen(const orc_system* s, double* v, double dt, int64_t step, uint64_t seed) {
    const double prob = s->therm_p * dt;
    const int64_t nb = s->water ? s->n / 3 : s->n;
    for (int64_t b = 0; b < nb; ++b) {
        int64_t first = s->water ? 3 * b : b;
        if (!(rng_uniform(seed, first, step, RNG_ANDERSEN) < prob)) continue;
        int64_t natoms = s->water ? 3 : 1;
        for (int64_t t = 0; t < natoms; ++t) {
            int64_t i = first + t;
            double m = s->water ? s->ms[i] : s->ms[0];
            double vdev = sqrt(s->kb * s->therm_T / m);
            double g[3];
            rng_normals3(seed, i, step, RNG_ANDERSEN, g);
            v[3 * i] = vdev * g[0];
            v[3 * i + 1] = vdev * g[1];
            v[3 * i + 2] = vdev * g[2];
        }
    }
}

/* Langevin half O-step: v <- c1 v + sqrt((1 - c1^2) kb T / m_i) xi, c1 = exp(-gamma dt/2).
 * Applied before and after each velocity-Verlet step (the reference instead
 * integrates an SDE with Euler-Maruyama, src/nbody_to_ode.jl:567-598). */
static void langevin_half(const orc_system* s, double* v, double dt, int64_t step,
                          uint64_t seed, uint32_t stream) {
    const double c1 = exp(-s->therm_p * (0.5 * dt));
    const double c2 = sqrt(1.0 - c1 * c1);
    for (int64_t i = 0; i < s->n; ++i) {
        double sig = c2 * sqrt(s->kb * s->therm_T / s->ms[i]);
        double g[3];
        rng_normals3(seed, i, step, stream, g);
        for (int k = 0; k < 3; ++k) v[3 * i + k] = c1 * v[3 * i + k] + sig * g[k];
    }
}

/* nh_out (may be NULL): the Nose-Hoover zeta and xi at the end of the run */
int64_t orc_vv_run_nh(const orc_system* s, double* u, double* v, double dt, int64_t nsteps,
                      uint64_t seed, int64_t save_every, double* u_frames, double* v_frames,
                      double* nh_out) {
    const int64_t n3 = 3 * s->n;
    const size_t bytes = sizeof(double) * (size_t)(n3 > 0 ? n3 : 1);
    double* a = (double*)malloc(bytes);
    double* an = (double*)malloc(bytes);
    double* vh = (double*)malloc(bytes);
    const double half = 0.5, dtsq = dt * dt;
    int64_t nf = 0;
    if (save_every > 0 && u_frames) {
        memcpy(u_frames, u, bytes); memcpy(v_frames, v, bytes); nf = 1;
    }
    /* Nose-Hoover extra coordinate: zeta (position part) and its rate xi (velocity part),
       rewritten by every right-hand side; its acceleration is zero */
    double zeta = s->nh_zeta, xi = 0.0;
    rhs_impl(s, u, v, a, zeta, &xi);
    for (int64_t step = 1; step <= nsteps; ++step) {
        if (s->therm == ORC_THERM_LANGEVIN) langevin_half(s, v, dt, step, seed, RNG_LANGEVIN_A);
        for (int64_t k = 0; k < n3; ++k) {
            u[k] = u[k] + dt * v[k] + dtsq * (half * a[k]);
            vh[k] = v[k] + dt * (half * a[k]);
        }
        zeta = zeta + dt * xi;
        rhs_impl(s, u, vh, an, zeta, &xi);
        for (int64_t k = 0; k < n3; ++k) v[k] = vh[k] + dt * (half * an[k]);
        double* t = a; a = an; an = t;
        if (s->therm == ORC_THERM_LANGEVIN) langevin_half(s, v, dt, step, seed, RNG_LANGEVIN_B);
        if (s->therm == ORC_THERM_ANDERSEN) andersen(s, v, dt, step, seed);
        if (save_every > 0 && u_frames && step % save_every == 0) {
            memcpy(u_frames + nf * n3, u, bytes);
            memcpy(v_frames + nf * n3, v, bytes);
            ++nf;
        }
    }
    free(a); free(an); free(vh);
    if (nh_out) { nh_out[0] = zeta; nh_out[1] = xi; }
    return nf;
}

int64_t orc_vv_run(const orc_system* s, double* u, double* v, double dt, int64_t nsteps,
                   uint64_t seed, int64_t save_every, double* u_frames, double* v_frames) {
    return orc_vv_run_nh(s, u, v, dt, nsteps, seed, save_every, u_frames, v_frames, NULL);
}

/* Euler-Maruyama for the Langevin SDE, src/nbody_to_ode.jl:567-598.  The drift f is
 * du = v, dv = sum of the closures - gamma v (:575-589); the noise g is sigma on the
 * velocity rows only (:592-595).  One EM step from (u, v): u + dt*v (the position rows'
 * noise term is zero) and v + dt*(a - gamma*v) + sigma*dW, dW = sqrt(dt)*xi. */
int64_t orc_em_run(const orc_system* s, double* u, double* v, double dt, int64_t nsteps,
                   uint64_t seed, int64_t save_every, double* u_frames, double* v_frames) {
    const int64_t n3 = 3 * s->n;
    const size_t bytes = sizeof(double) * (size_t)(n3 > 0 ? n3 : 1);
    double* a = (double*)malloc(bytes);
    const double gamma = s->therm_p;
    const double sigma = sqrt(2 * gamma * s->kb * s->therm_T / s->ms[0]);
    const double sqrt_dt = sqrt(dt);
    int64_t nf = 0;
    if (save_every > 0 && u_frames) {
        memcpy(u_frames, u, bytes); memcpy(v_frames, v, bytes); nf = 1;
    }
    double xi = 0.0;
    rhs_impl(s, u, v, a, 0.0, &xi);
    for (int64_t step = 1; step <= nsteps; ++step) {
        for (int64_t i = 0; i < s->n; ++i) {
            double g[3];
            rng_normals3(seed, i, step, RNG_EM, g);
            for (int k = 0; k < 3; ++k) {
                const double vk = v[3 * i + k];
                u[3 * i + k] = u[3 * i + k] + dt * vk;
                v[3 * i + k] = vk + dt * (a[3 * i + k] - gamma * vk) + sigma * (sqrt_dt * g[k]);
            }
        }
        rhs_impl(s, u, v, a, 0.0, &xi);
        if (save_every > 0 && u_frames && step % save_every == 0) {
            memcpy(u_frames + nf * n3, u, bytes);
            memcpy(v_frames + nf * n3, v, bytes);
            ++nf;
        }
    }
    free(a);
    return nf;
}

/* ---------------------------------------------------------------------------
 * Energies: src/nbody_simulation_result.jl:209-446
 * ------------------------------------------------------------------------- */

double orc_kinetic_energy(const double* v, const double* ms, int64_t n) {
    double ke = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        const double* vi = v + 3 * i;
        double s = vi[0] * vi[0] + vi[1] * vi[1];
        s = s + vi[2] * vi[2];
        ke += s * (ms[i] / 2.0);
    }
    return ke;
}

double orc_lj_potential(const double* rs, const int64_t* indxs, int64_t nidx, double eps,
                        double sigma, double R, const orc_bc* bc) {
    const double s2 = sigma * sigma, R2 = R * R;
    double e = 0.0;
    for (int64_t a = 0; a < nidx; ++a) {
        const double* ri = rs + 3 * indxs[a];
        for (int64_t b = a + 1; b < nidx; ++b) {
            double rij[3], r, r2;
            orc_interparticle_distance(ri, rs + 3 * indxs[b], bc, rij, &r, &r2);
            if (!(r2 < R2)) r2 = R2;               /* :308-310 */
            double q = s2 / r2;
            double s6 = q * q * q;
            double s12 = s6 * s6;
            e += (s12 - s6);
        }
    }
    return 4.0 * eps * e;
}

double orc_elec_potential(const double* rs, int64_t n, const double* qs, int32_t water,
                          double k, double R, const orc_bc* bc) {
    const double R2 = R * R;
    double e = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        const double* ri = rs + 3 * i;
        double ei = 0.0;
        for (int64_t j = i + 1; j < n; ++j) {
            if (excluded(i, j, water)) continue;
            double rij[3], r, r2;
            orc_interparticle_distance(ri, rs + 3 * j, bc, rij, &r, &r2);
            if (r2 < R2) ei += qs[j] / r;          /* :340-344 */
            else ei += qs[j] / R;
        }
        e += ei * qs[i];
    }
    return e * k;
}

double orc_rf_potential(const double* rs, int64_t n, const double* qs, int32_t water,
                        double k, double R, double eps_rf, const orc_bc* bc) {
    const double R2 = R * R;
    const double krf = rf_krf(R, eps_rf);
    const double crf = 1.0 / R + krf * R2;
    double e = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        const double* ri = rs + 3 * i;
        double ei = 0.0;
        for (int64_t j = i + 1; j < n; ++j) {
            if (excluded(i, j, water)) continue;
            double rij[3], r, r2;
            orc_interparticle_distance(ri, rs + 3 * j, bc, rij, &r, &r2);
            if (r2 < R2) ei += qs[j] * (1.0 / r + krf * r2 - crf);
        }
        e += ei * qs[i];
    }
    return e * k;
}

double orc_grav_potential(const double* rs, int64_t n, const double* ms, double G) {
    double e = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        const double* ri = rs + 3 * i;
        double ei = 0.0;
        for (int64_t j = i + 1; j < n; ++j) {
            double rij[3] = {ri[0] - rs[3 * j], ri[1] - rs[3 * j + 1], ri[2] - rs[3 * j + 2]};
            ei += ms[j] / norm3(rij);
        }
        e += ei * ms[i];
    }
    return -G * e;
}

double orc_bond_potential(const double* rs, int64_t nmol, double rOH, double kbond) {
    /* every O-H bond is visited from both ends (:361-370), then divided by 4 */
    double e = 0.0;
    for (int64_t i = 0; i < 3 * nmol; ++i) {
        int64_t o = 3 * (i / 3);
        int64_t nb[2];
        int nn;
        if (i == o) { nb[0] = o + 1; nb[1] = o + 2; nn = 2; }
        else { nb[0] = o; nn = 1; }
        for (int t = 0; t < nn; ++t) {
            double rij[3] = {rs[3 * i] - rs[3 * nb[t]], rs[3 * i + 1] - rs[3 * nb[t] + 1],
                             rs[3 * i + 2] - rs[3 * nb[t] + 2]};
            double d = norm3(rij) - rOH;
            e += d * d * kbond;
        }
    }
    return e / 4.0;
}

double orc_angle_potential(const double* rs, int64_t nmol, double aHOH, double ka) {
    double e = 0.0;
    for (int64_t m = 0; m < nmol; ++m) {
        const double* ra = rs + 3 * (3 * m + 1);
        const double* rb = rs + 3 * (3 * m);
        const double* rc = rs + 3 * (3 * m + 2);
        double rba[3], rbc[3];
        for (int k = 0; k < 3; ++k) { rba[k] = ra[k] - rb[k]; rbc[k] = rc[k] - rb[k]; }
        double ang = acos(dot3(rba, rbc) / (norm3(rba) * norm3(rbc)));
        double d = ang - aHOH;
        e += ka * (d * d);
    }
    return e / 2.0;
}

double orc_potential_energy(const orc_system* s, const double* u) {
    double e = 0.0;
    if (!s->water) {
        int64_t* all = (int64_t*)malloc(sizeof(int64_t) * (size_t)(s->n > 0 ? s->n : 1));
        for (int64_t i = 0; i < s->n; ++i) all[i] = i;
        for (int t = 0; t < s->nterms; ++t) {
            const orc_term* tm = &s->terms[t];
            if (tm->kind == ORC_TERM_LJ)
                e += orc_lj_potential(u, all, s->n, tm->p[0], tm->p[1], tm->p[2], &s->bc);
            else if (tm->kind == ORC_TERM_ELEC)
                e += orc_elec_potential(u, s->n, s->qs, 0, tm->p[0], tm->p[1], &s->bc);
            else if (tm->kind == ORC_TERM_RF)
                e += orc_rf_potential(u, s->n, s->qs, 0, tm->p[0], tm->p[1], tm->p[2], &s->bc);
        }
        free(all);
        return e;
    }
    const int64_t nmol = s->n / 3;
    int64_t* oidx = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nmol > 0 ? nmol : 1));
    for (int64_t m = 0; m < nmol; ++m) oidx[m] = 3 * m;
    e += orc_lj_potential(u, oidx, nmol, s->water_lj.p[0], s->water_lj.p[1], s->water_lj.p[2],
                          &s->bc);
    free(oidx);
    if (s->water_elec.kind == ORC_TERM_RF)
        e += orc_rf_potential(u, s->n, s->qs, 1, s->water_elec.p[0], s->water_elec.p[1],
                              s->water_elec.p[2], &s->bc);
    else
        e += orc_elec_potential(u, s->n, s->qs, 1, s->water_elec.p[0], s->water_elec.p[1],
                                &s->bc);
    e += orc_bond_potential(u, nmol, s->rOH, s->kbond);
    e += orc_angle_potential(u, nmol, s->aHOH, s->kangle);
    return e;
}

/* ---------------------------------------------------------------------------
 * RDF: src/nbody_simulation_result.jl:664-709
 * ------------------------------------------------------------------------- */

void orc_rdf(const double* frames, int64_t nframes, int64_t n, const int64_t* indxs,
             int64_t nidx, double L, int64_t maxbin, double* rs, double* gr) {
    orc_bc bc;
    memset(&bc, 0, sizeof bc);
    bc.kind = ORC_BC_CUBIC;
    bc.b[0] = L;
    const double dr = L / (double)maxbin;
    double* hist = (double*)calloc((size_t)maxbin + 1, sizeof(double));
    const double lim = (0.5 * L) * (0.5 * L);
    for (int64_t t = 0; t < nframes; ++t) {
        const double* cc = frames + t * 3 * n;
        for (int64_t a = 0; a < nidx; ++a) {
            const double* ri = cc + 3 * indxs[a];
            for (int64_t b = a + 1; b < nidx; ++b) {
                double rij[3], r, r2;
                orc_interparticle_distance(ri, cc + 3 * indxs[b], &bc, rij, &r, &r2);
                if (r2 < lim) {
                    double fb = ceil(r / dr);
                    if (fb > 1.0 && fb <= (double)maxbin) hist[(int64_t)fb] += 2.0;
                }
            }
        }
    }
    const double c = 4.0 / 3.0 * 3.141592653589793 * (double)nidx / (L * L * L);
    for (int64_t bin = 1; bin <= maxbin; ++bin) {
        double rl = (double)(bin - 1) * dr;
        double ru = rl + dr;
        double nideal = c * (ru * ru * ru - rl * rl * rl);
        gr[bin - 1] = (hist[bin] / (double)(nframes * nidx)) / nideal;
        rs[bin - 1] = rl + dr / 2.0;
    }
    free(hist);
}

/* ---------------------------------------------------------------------------
 * Magnetic dipoles: src/basic_potentials.jl:333-365
 * ------------------------------------------------------------------------- */
static double dot3s(const double* a, const double* b) {
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

void orc_mag_accel(double* dv, const double* rs, int64_t i, int64_t n, const double* ms,
                   const double* mm, double mu_4pi) {
    double f1 = 0.0, f2 = 0.0, f3 = 0.0;
    const double* mi = mm + 3 * i;
    const double* ri = rs + 3 * i;
    for (int64_t j = 0; j < n; ++j) {
        if (j == i) continue;
        const double* mj = mm + 3 * j;
        const double* rj = rs + 3 * j;
        const double rij[3] = {ri[0] - rj[0], ri[1] - rj[1], ri[2] - rj[2]};
        const double d = dot3s(rij, rij);
        const double rij4 = d * d;                                   /* dot(rij, rij)^2 */
        const double nr = norm3(rij);
        const double r[3] = {rij[0] / nr, rij[1] / nr, rij[2] / nr};  /* rij / norm(rij) */
        const double mir = dot3s(mi, r);
        const double mij = dot3s(mj, r);
        const double dmm = dot3s(mi, mj);
        double c[3];
        for (int k = 0; k < 3; ++k)
            c[k] = (((mi[k] * mij + mj[k] * mir) + r[k] * dmm) - ((5.0 * r[k]) * mir) * mij) / rij4;
        f1 += c[0];
        f2 += c[1];
        f3 += c[2];
    }
    const double coeff = 3.0 * mu_4pi / ms[i];
    dv[0] += coeff * f1;
    dv[1] += coeff * f2;
    dv[2] += coeff * f3;
}

double orc_mag_potential(const double* rs, int64_t n, const double* mm, double mu_4pi) {
    double e = 0.0;
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = i + 1; j < n; ++j) {
            const double rij[3] = {rs[3 * i] - rs[3 * j], rs[3 * i + 1] - rs[3 * j + 1],
                                   rs[3 * i + 2] - rs[3 * j + 2]};
            const double nr = norm3(rij);
            const double r[3] = {rij[0] / nr, rij[1] / nr, rij[2] / nr};
            e += (dot3s(mm + 3 * i, mm + 3 * j) -
                  3.0 * dot3s(mm + 3 * i, r) * dot3s(mm + 3 * j, r)) / (nr * nr * nr);
        }
    return mu_4pi * e;
}

/* ---------------------------------------------------------------------------
 * Learned-stride descriptor (definition in nbx_oracle.h; no reference counterpart)
 * ------------------------------------------------------------------------- */
static void lr_frame(const double* a, const double* v, double* e1, double* e2, double* ff) {
    const double an = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    if (an > 0.0) { e1[0] = a[0] / an; e1[1] = a[1] / an; e1[2] = a[2] / an; }
    else { e1[0] = 1.0; e1[1] = 0.0; e1[2] = 0.0; }
    const double vp = v[0] * e1[0] + v[1] * e1[1] + v[2] * e1[2];
    const double w[3] = {v[0] - vp * e1[0], v[1] - vp * e1[1], v[2] - vp * e1[2]};
    const double wn = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    if (wn > 0.0) { e2[0] = w[0] / wn; e2[1] = w[1] / wn; e2[2] = w[2] / wn; }
    else {
        double c[3] = {0.0, 0.0, 0.0};
        const double ux = fabs(e1[0]), uy = fabs(e1[1]), uz = fabs(e1[2]);
        if (ux <= uy && ux <= uz) c[0] = 1.0; else if (uy <= uz) c[1] = 1.0; else c[2] = 1.0;
        const double cp = c[0] * e1[0] + c[1] * e1[1] + c[2] * e1[2];
        const double q[3] = {c[0] - cp * e1[0], c[1] - cp * e1[1], c[2] - cp * e1[2]};
        const double qn = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
        e2[0] = q[0] / qn; e2[1] = q[1] / qn; e2[2] = q[2] / qn;
    }
    ff[0] = vp;
    ff[1] = wn;
    ff[2] = an;
}

void orc_lr_features(const double* u, const double* a, const double* v, int64_t n, double L,
                     double* feats, double* frames) {
    orc_bc bc;
    memset(&bc, 0, sizeof(bc));
    bc.kind = ORC_BC_CUBIC;
    bc.b[0] = L;
    const double pi = 3.14159265358979323846;
    for (int64_t i = 0; i < n; ++i) {
        double g[32];
        for (int b = 0; b < 32; ++b) g[b] = 0.0;
        for (int64_t j = 0; j < n; ++j) {
            if (j == i) continue;
            double rij[3], r, r2;
            orc_interparticle_distance(u + 3 * i, u + 3 * j, &bc, rij, &r, &r2);
            if (!(r2 < 6.25)) continue;
            const double sw = r < 2.0 ? 1.0 : 0.5 * (1.0 + cos(pi * (r - 2.0) / 0.5));
            for (int b = 0; b < 32; ++b) {
                const double e = r - (0.8 + 1.7 * (double)b / 31.0);
                g[b] += sw * exp(-0.5 * e * e / 0.01);
            }
        }
        double e1[3], e2[3], ff[3];
        lr_frame(a + 3 * i, v + 3 * i, e1, e2, ff);
        for (int b = 0; b < 32; ++b) feats[35 * i + b] = g[b];
        feats[35 * i + 32] = ff[0];
        feats[35 * i + 33] = ff[1];
        feats[35 * i + 34] = ff[2];
        for (int k = 0; k < 3; ++k) { frames[6 * i + k] = e1[k]; frames[6 * i + 3 + k] = e2[k]; }
    }
}
