/*
 * oracle.c -- CPU restatement of red.cuda's direct-sum N-body hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker the HIP engine is compared with.
 * Build: gcc -O2 -ffp-contract=off (the reference builds with -fmad=false,
 * Makefile:17, so restating its operation order reproduces its FP64 results).
 *
 * Parity pins (what fixes each part; see tests/test_oracle_*.py):
 *   bounds           red_type.h:670-700      KAT src/red.cuda/test.cu:387-448
 *   tools            tools.cpp:367-507       KATs red_test.cpp:22-125, 626-950
 *   kepler           tools.cpp:715-750       KAT red_test.cpp:1139-1175 (the
 *                    reference solver fails its own 225 deg case; this one
 *                    solves E - e sin E = M by Newton and passes both)
 *   mt19937          distribution.cpp:11-31  std::mt19937 (C++ standard)
 *   gravity (GPU)    pp_disk.cu:374-416      brute-force long double, 2-body
 *                    closed form, Newton-3 invariant, symmetric ring
 *   gravity (CPU)    pp_disk.cu:822-865      vs GPU formula (1-ulp rounding of
 *                    m/(d d^2) vs m*(1/(d d^2)))
 *   drag             pp_disk.cu:60-155,486   quadratic drag law 3 C_D rho_g |u| u
 *                    / (8 R rho_p) with rho_g and v_gas from the disk model in
 *                    long double (inside / outside the taper, off the
 *                    midplane); decrease laws; decr<1e-6 cut; the gas
 *                    parameters integrate over z to the stated Sigma
 *   grid drag        pp_disk.cu:158-193, 778-814: rotate_2D_vector KATs
 *                    (red_test.cpp:170-250); the cell against numpy's
 *                    searchsorted and floor(angle / dalpha) at ring and
 *                    sector centres and on the edges; a uniform grid at rest
 *                    relative to Keplerian rotation equals the analytic
 *                    disk's drag with eta = 0 and a flat midplane density;
 *                    gas velocity from unit vectors in long double
 *   RK4 / RKF78      int_rungekutta4.cu, int_rungekutta8.cu: order of
 *                    convergence on Kepler; the oracle's tableau against the
 *                    published fractions, row sums == c, quadrature order 7
 *   error estimates  int_rungekutta8.cu:250-260, 693-696: lambda |dt| max|E|
 *                    equals the true local error of the 7th-order step from an
 *                    exact Kepler propagation (ratio -> 1, O(dt^8)); RK4's
 *                    embedded estimate O(dt^4)
 *   controller       int_rungekutta8.cu:697-708: err/tol settles at 0.9^7
 *                    (0.9^4 for RK4) on a circular orbit; zero error doubles
 *   events           pp_disk.cu:284-371,1220-1368: constructed collision,
 *                    ejection and hit-centrum scenarios
 *   integrals        tools.cpp:545-700       2-body closed form
 *   disk generator   src/red.cuda.initial/main.cu:251-342 (samplers pinned by
 *                    moments); parity unpinned against redinit output (no
 *                    reference output files exist)
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OR_K   0.01720209895
#define OR_K2  0.0002959122082855911025
#define OR_PI  3.1415926535897932384626
#define OR_TWOPI 6.2831853071795864769253
#define SQR(x) ((x) * (x))
#define CUBE(x) ((x) * (x) * (x))

/* ------------------------------------------------------------------------ */
/* n_objects_t bounds: red_type.h:622-700                                    */
/* which: 0 SI, 1 NSI, 2 NI, 3 GD                                            */
/* ------------------------------------------------------------------------ */
void or_bound(const uint32_t pl[OR_NTYPE], int which, uint32_t out[4])
{
    uint32_t n_si  = pl[OR_STAR] + pl[OR_GIANT] + pl[OR_ROCKY] + pl[OR_PROTO];
    uint32_t n_nsi = pl[OR_SUPERPL] + pl[OR_PL];
    uint32_t n_ni  = pl[OR_TP];
    uint32_t n_mas = n_si + n_nsi;
    switch (which) {
    case 0: out[0] = 0;     out[1] = n_si;          out[2] = 0; out[3] = n_mas; break;
    case 1: out[0] = n_si;  out[1] = n_si + n_nsi;  out[2] = 0; out[3] = n_si;  break;
    case 2: out[0] = n_mas; out[1] = n_mas + n_ni;  out[2] = 0; out[3] = n_mas; break;
    default: out[0] = n_si; out[1] = n_si + n_nsi;  out[2] = 0; out[3] = 0;     break;
    }
}

static uint32_t n_si_of(const uint32_t pl[OR_NTYPE]) { return pl[0] + pl[1] + pl[2] + pl[3]; }
static uint32_t n_nsi_of(const uint32_t pl[OR_NTYPE]) { return pl[4] + pl[5]; }
static uint32_t n_total_of(const uint32_t pl[OR_NTYPE])
{
    uint32_t n = 0;
    for (int i = 0; i < OR_NTYPE; i++) n += pl[i];
    return n;
}

/* ------------------------------------------------------------------------ */
/* tools (src/red.cuda.utility/tools.cpp)                                    */
/* ------------------------------------------------------------------------ */
/* tools.cpp:455-460 */
double or_calc_radius(double m, double density)
{
    const double four_pi_over_three = 4.188790204786391;
    return pow(1.0 / four_pi_over_three * m / density, 1.0 / 3.0);
}
/* tools.cpp:462-471 */
double or_calc_density(double m, double R)
{
    const double four_pi_over_three = 4.188790204786391;
    if (R == 0.0) return 0.0;
    return m / (four_pi_over_three * CUBE(R));
}
/* tools.cpp:473-478 */
double or_calc_mass(double R, double density)
{
    const double four_pi_over_three = 4.188790204786391;
    return four_pi_over_three * CUBE(R) * density;
}
/* tools.cpp:480-487 */
void or_calc_position_after_collision(double m1, double m2, const or_var4* r1, const or_var4* r2, or_var4* r)
{
    const double M = m1 + m2;
    r->x = (m1 * r1->x + m2 * r2->x) / M;
    r->y = (m1 * r1->y + m2 * r2->y) / M;
    r->z = (m1 * r1->z + m2 * r2->z) / M;
}
/* tools.cpp:489-496 */
void or_calc_velocity_after_collision(double m1, double m2, const or_var4* v1, const or_var4* v2, or_var4* v)
{
    const double M = m1 + m2;
    v->x = (m1 * v1->x + m2 * v2->x) / M;
    v->y = (m1 * v1->y + m2 * v2->y) / M;
    v->z = (m1 * v1->z + m2 * v2->z) / M;
}
/* tools.cpp:498-507 */
void or_calc_physical_properties(const or_param* p1, const or_param* p2, or_param* p)
{
    double volume = 4.188790204786391 * (CUBE(p1->radius) + CUBE(p2->radius));
    p->mass    = p1->mass + p2->mass;
    p->density = p->mass / volume;
    p->radius  = or_calc_radius(p->mass, p->density);
    p->cd      = p1->cd;
}
/* tools.cpp:367-382 (j runs downwards) */
double or_get_total_mass(uint32_t n, const or_body_md* bmd, const or_param* p)
{
    double M0 = 0.0;
    for (int j = (int)n - 1; j >= 0; j--) {
        if (0 > bmd[j].id) continue;
        M0 += p[j].mass;
    }
    return M0;
}
/* tools.cpp:384-410 */
void or_calc_bc(uint32_t n, const or_body_md* bmd, const or_param* p, const or_var4* r, const or_var4* v,
                double M0, or_var4* R0, or_var4* V0)
{
    R0->x = R0->y = R0->z = R0->w = 0.0;
    V0->x = V0->y = V0->z = V0->w = 0.0;
    for (int j = (int)n - 1; j >= 0; j--) {
        if (0 > bmd[j].id) continue;
        double m_j = p[j].mass;
        R0->x += m_j * r[j].x; R0->y += m_j * r[j].y; R0->z += m_j * r[j].z;
        V0->x += m_j * v[j].x; V0->y += m_j * v[j].y; V0->z += m_j * v[j].z;
    }
    R0->x /= M0; R0->y /= M0; R0->z /= M0;
    V0->x /= M0; V0->y /= M0; V0->z /= M0;
}
/* tools.cpp:412-441 */
void or_transform_to_bc(uint32_t n, const or_body_md* bmd, const or_param* p, or_var4* r, or_var4* v)
{
    or_var4 R0, V0;
    double M0 = or_get_total_mass(n, bmd, p);
    or_calc_bc(n, bmd, p, r, v, M0, &R0, &V0);
    for (int j = (int)n - 1; j >= 0; j--) {
        if (0 > bmd[j].id) continue;
        r[j].x -= R0.x; r[j].y -= R0.y; r[j].z -= R0.z;
        v[j].x -= V0.x; v[j].y -= V0.y; v[j].z -= V0.z;
    }
}
/* pp_disk.cu:1504-1516 -> tools.cpp:545-565 (C), 591-620 (U), 653-670 (T), 692-695 (E) */
void or_calc_integral(uint32_t n, const or_body_md* bmd, const or_param* p, const or_var4* r, const or_var4* v,
                      or_integral* out)
{
    double M0 = or_get_total_mass(n, bmd, p);
    or_calc_bc(n, bmd, p, r, v, M0, &out->R, &out->V);

    or_var4 C = {0.0, 0.0, 0.0, 0.0};
    for (uint32_t i = 0; i < n; i++) {
        if (0 > bmd[i].id) continue;
        or_var4 c;
        c.x = r[i].y * v[i].z - r[i].z * v[i].y;
        c.y = r[i].z * v[i].x - r[i].x * v[i].z;
        c.z = r[i].x * v[i].y - r[i].y * v[i].x;
        c.x *= p[i].mass; c.y *= p[i].mass; c.z *= p[i].mass;
        C.x += c.x; C.y += c.y; C.z += c.z;
    }
    out->C = C;

    double T = 0.0;
    for (uint32_t i = 0; i < n; i++) {
        if (0 > bmd[i].id) continue;
        T += p[i].mass * (SQR(v[i].x) + SQR(v[i].y) + SQR(v[i].z));
    }
    T = T / 2.0;
    double U = 0.0;
    for (uint32_t i = 0; i < n; i++) {
        if (0 > bmd[i].id) continue;
        for (uint32_t j = 0; j < n; j++) {
            if (i == j) continue;
            if (0 > bmd[j].id) continue;
            double r_ij = sqrt(SQR(r[j].x - r[i].x) + SQR(r[j].y - r[i].y) + SQR(r[j].z - r[i].z));
            U += p[i].mass * p[j].mass / r_ij;
        }
    }
    U = U / 2.0;
    out->E = T - U;
}

/* Kepler's equation E - e sin E = M, Newton iteration from the same starting
 * guess as tools.cpp:724.  For M > pi the symmetry E(M) = 2 pi - E(2 pi - M)
 * is used (tools.cpp:719-723 shifts by pi instead and then discards the
 * solution, tools.cpp:745-748, which is why the reference fails its own
 * 225-degree known-answer test, red_test.cpp:1159-1172). */
int or_kepler_equation_solver(double ecc, double mean, double eps, double* E)
{
    const int n_max_step = 32;
    if (0.0 == ecc || 0.0 == mean || OR_PI == mean) { *E = mean; return OR_OK; }
    int mirror = 0;
    if (OR_PI < mean) { mean = OR_TWOPI - mean; mirror = 1; }
    double x = mean + ecc * (sin(mean)) / (1.0 - sin(mean + ecc) + sin(mean));
    double err;
    int step = 0;
    do {
        double x1 = x - (x - ecc * sin(x) - mean) / (1.0 - ecc * cos(x));
        err = fabs(x1 - x);
        x = x1;
    } while (err > eps && step++ <= n_max_step);
    if (n_max_step < step) return OR_EINVAL;
    *E = mirror ? OR_TWOPI - x : x;
    return OR_OK;
}

/* tools.cpp:789-828 */
int or_calc_phase(double mu, const or_orbelem* oe, or_var4* rv, or_var4* vv)
{
    double ecc = oe->ecc, E = 0.0;
    int rc = or_kepler_equation_solver(ecc, oe->mean, 1.0e-14, &E);
    if (rc) return rc;
    double v = 2.0 * atan(sqrt((1.0 + ecc) / (1.0 - ecc)) * tan(E / 2.0));
    double p = oe->sma * (1.0 - SQR(ecc));
    double r = p / (1.0 + ecc * cos(v));
    double kszi = r * cos(v), eta = r * sin(v);
    double vKszi = -sqrt(mu / p) * sin(v);
    double vEta = sqrt(mu / p) * (ecc + cos(v));
    double cw = cos(oe->peri), sw = sin(oe->peri);
    double cO = cos(oe->node), sO = sin(oe->node);
    double ci = cos(oe->inc), si = sin(oe->inc);
    double Px = cw * cO - sw * sO * ci, Py = cw * sO + sw * cO * ci, Pz = sw * si;
    double Qx = -sw * cO - cw * sO * ci, Qy = -sw * sO + cw * cO * ci, Qz = cw * si;
    rv->x = kszi * Px + eta * Qx; rv->y = kszi * Py + eta * Qy; rv->z = kszi * Pz + eta * Qz; rv->w = 0.0;
    vv->x = vKszi * Px + vEta * Qx; vv->y = vKszi * Py + vEta * Qy; vv->z = vKszi * Pz + vEta * Qz; vv->w = 0.0;
    return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* MT19937 and samplers (src/red.cuda.initial/distribution.cpp)              */
/* ------------------------------------------------------------------------ */
/* distribution.cpp:11-19: standard MT19937 initialisation */
void or_rng_seed(or_rng* g, uint32_t seed)
{
    g->idx = 0;
    g->I[0] = seed;
    for (uint32_t i = 1; i < 624; i++)
        g->I[i] = 1812433253u * (g->I[i - 1] ^ (g->I[i - 1] >> 30)) + i;
}
/* distribution.cpp:24-31: one-at-a-time MT19937 twist + tempering */
uint32_t or_rng_u32(or_rng* g)
{
    uint32_t i = g->idx;
    uint32_t nx = (i + 1 == 624) ? 0 : i + 1;
    uint32_t m = (i + 397) % 624;
    uint32_t y = (g->I[i] & 0x80000000u) | (g->I[nx] & 0x7fffffffu);
    uint32_t s = g->I[m] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    g->I[i] = s;
    g->idx = nx;
    s ^= s >> 11;
    s ^= (s << 7) & 0x9d2c5680u;
    s ^= (s << 15) & 0xefc60000u;
    s ^= s >> 18;
    return s;
}
/* distribution.cpp:38-41: value in (lo, hi] */
double or_rng_uniform(or_rng* g, double lo, double hi)
{
    return lo + (hi - lo) * (or_rng_u32(g) + 1.0) / 4294967296.0;
}
/* distribution.cpp:103-109 */
double or_rng_rayleigh(or_rng* g, double sigma)
{
    double u = (or_rng_u32(g) + 1.0) / 4294967296.0;
    return sigma * sqrt(-2.0 * log(u));
}
/* Rejection sampling under the pdf x^power on [x_min, x_max], with the draw
 * sequence of distribution.cpp:139-162 (x, then y, per trial).  The
 * reference draws y on [min(y), max(y)], which samples the pdf
 * x^power - min(x_min^power, x_max^power) instead; here y is drawn on
 * (0, max(y)] so the accepted x follow x^power. */
double or_rng_power_law(or_rng* g, double x_min, double x_max, double power)
{
    double y_min = power == 0.0 ? 0.0 : pow(x_min, power);
    double y_max = power == 0.0 ? 1.0 : pow(x_max, power);
    if (y_min > y_max) { double t = y_min; y_min = y_max; y_max = t; }
    double d_y = y_max, d_x = x_max - x_min;
    double area_max = d_x * d_y;
    double x, y;
    do {
        x = or_rng_uniform(g, 0.0, area_max) / d_y + x_min;
        y = or_rng_uniform(g, 0.0, y_max);
    } while (y > pow(x, power));
    return x;
}

/* Disk generator (src/red.cuda.initial/main.cu:251-342, 1662-1670):
 * elements -> Cartesian (mu = K^2 (m0 + m)) -> barycentric -> tau units.
 * One MT19937 stream, per body in the order a, e, i, peri, node, mean, size. */
int or_make_disk(const or_disk_spec* s, or_var4* r, or_var4* v, or_param* p, or_body_md* bmd)
{
    const double AuToMeter = 1.495978707e11;
    const double MeterToAu = 1.0 / AuToMeter;
    const double SolarToGram = 1.0e3 * 1.98911e30;
    const double GramPerCm3ToSolarPerAu3 = (1.0 / SolarToGram) / CUBE(1.0e-2 * MeterToAu);
    const double KilometerToAu = 1.0 / (AuToMeter / 1.0e3);
    or_rng g;
    or_rng_seed(&g, s->seed);
    uint32_t n = s->n + 1;
    double rho = s->density_cgs * GramPerCm3ToSolarPerAu3;

    memset(r, 0, n * sizeof(or_var4));
    memset(v, 0, n * sizeof(or_var4));
    p[0].mass = s->m_star; p[0].radius = s->r_star;
    p[0].density = or_calc_density(s->m_star, s->r_star); p[0].cd = 0.0;
    bmd[0].id = 1; bmd[0].body_type = OR_STAR; bmd[0].mig_type = 0; bmd[0].mig_stop_at = 0.0;
    for (uint32_t i = 1; i < n; i++) {
        or_orbelem oe;
        oe.sma  = or_rng_power_law(&g, s->a_min, s->a_max, s->a_power);
        oe.ecc  = or_rng_rayleigh(&g, s->sig_e);
        oe.inc  = or_rng_rayleigh(&g, s->sig_i);
        oe.peri = or_rng_uniform(&g, 0.0, OR_TWOPI);
        oe.node = or_rng_uniform(&g, 0.0, OR_TWOPI);
        oe.mean = or_rng_uniform(&g, 0.0, OR_TWOPI);
        if (oe.peri >= OR_TWOPI) oe.peri -= OR_TWOPI;
        if (oe.node >= OR_TWOPI) oe.node -= OR_TWOPI;
        if (oe.mean >= OR_TWOPI) oe.mean -= OR_TWOPI;
        double u = or_rng_uniform(&g, log10(s->lo), log10(s->hi));
        if (s->size_mode == 0) {
            p[i].mass = pow(10.0, u);
            p[i].density = rho;
            p[i].radius = or_calc_radius(p[i].mass, rho);
        } else {
            p[i].radius = pow(10.0, u) * KilometerToAu;
            p[i].density = rho;
            p[i].mass = or_calc_mass(p[i].radius, rho);
        }
        p[i].cd = s->cd;
        bmd[i].id = (int32_t)(i + 1); bmd[i].body_type = s->body_type; bmd[i].mig_type = 0; bmd[i].mig_stop_at = 0.0;
        double mu = OR_K2 * (p[0].mass + p[i].mass);
        int rc = or_calc_phase(mu, &oe, &r[i], &v[i]);
        if (rc) return rc;
    }
    or_transform_to_bc(n, bmd, p, r, v);
    for (uint32_t i = 0; i < n; i++) {   /* tools.cpp:443-453 */
        v[i].x /= OR_K; v[i].y /= OR_K; v[i].z /= OR_K;
    }
    return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Gravity                                                                   */
/* ------------------------------------------------------------------------ */
/* pp_disk.cu:374-416 for one interaction_bound */
static void grav_gpu_bound(const uint32_t b[4], const or_body_md* bmd, const or_param* p, const or_var4* r, or_var4* a)
{
    for (uint32_t i = b[0]; i < b[1]; i++) {
        a[i].x = a[i].y = a[i].z = a[i].w = 0.0;
        if (0 < bmd[i].id) {
            double ax = 0.0, ay = 0.0, az = 0.0;
            for (uint32_t j = b[2]; j < b[3]; j++) {
                if (i == j || 0 > bmd[j].id) continue;
                double dx = r[j].x - r[i].x;
                double dy = r[j].y - r[i].y;
                double dz = r[j].z - r[i].z;
                double r2 = SQR(dx) + SQR(dy) + SQR(dz);
                double d = sqrt(r2);
                double s = p[j].mass / (d * r2);
                ax += s * dx; ay += s * dy; az += s * dz;
            }
            a[i].x = ax; a[i].y = ay; a[i].z = az;
        }
    }
}
/* pp_disk.cu:1014-1045: SI, NSI, NI launches in that order */
void or_grav_gpu(uint32_t n, const uint32_t pl[OR_NTYPE], const or_body_md* bmd,
                 const or_param* p, const or_var4* r, or_var4* a)
{
    (void)n;
    uint32_t b[4];
    for (int c = 0; c < 3; c++) {
        or_bound(pl, c, b);
        if (b[1] > b[0]) grav_gpu_bound(b, bmd, p, r, a);
    }
}
/* pp_disk.cu:984-1012 (+ 822-865 SI symmetric, 933-967 NI, 974-977 NSI) */
int or_grav_cpu(uint32_t n, const uint32_t pl[OR_NTYPE], const or_body_md* bmd,
                const or_param* p, const or_var4* r, or_var4* a)
{
    uint32_t b[4];
    memset(a, 0, n * sizeof(or_var4));
    if (n_si_of(pl) > 0) {
        or_bound(pl, 0, b);
        for (uint32_t i = b[0]; i < b[1]; i++) {
            if (!(0 < bmd[i].id)) continue;
            for (uint32_t j = i + 1; j < b[3]; j++) {
                if (0 > bmd[j].id) continue;
                double dx = r[j].x - r[i].x;
                double dy = r[j].y - r[i].y;
                double dz = r[j].z - r[i].z;
                double r2 = SQR(dx) + SQR(dy) + SQR(dz);
                double d = sqrt(r2);
                double r_3 = 1.0 / (d * r2);
                double w = p[j].mass * r_3;
                a[i].x += w * dx; a[i].y += w * dy; a[i].z += w * dz;
                w = p[i].mass * r_3;
                a[j].x -= w * dx; a[j].y -= w * dy; a[j].z -= w * dz;
            }
        }
    }
    if (n_nsi_of(pl) > 0) return OR_ENOTIMPL;
    if (pl[OR_TP] > 0) {
        or_bound(pl, 2, b);
        for (uint32_t i = b[0]; i < b[1]; i++) {
            if (!(0 < bmd[i].id)) continue;
            for (uint32_t j = b[2]; j < b[3]; j++) {
                if (i == j || 0 > bmd[j].id) continue;
                double dx = r[j].x - r[i].x;
                double dy = r[j].y - r[i].y;
                double dz = r[j].z - r[i].z;
                double r2 = SQR(dx) + SQR(dy) + SQR(dz);
                double d = sqrt(r2);
                double r_3 = 1.0 / (d * r2);
                double w = p[j].mass * r_3;
                a[i].x += w * dx; a[i].y += w * dy; a[i].z += w * dz;
            }
        }
    }
    return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Gas drag (pp_disk.cu:60-155, 486-521)                                     */
/* ------------------------------------------------------------------------ */
static double reduction_factor(const or_gas* g, double t)   /* pp_disk.cu:60-84 */
{
    switch (g->gas_decrease) {
    case 0: return 1.0;
    case 1:
        if (t <= g->t0) return 1.0;
        else if (g->t0 < t && t <= g->t1 && g->t0 != g->t1) return 1.0 - (t - g->t0) / (g->t1 - g->t0);
        else return 0.0;
    case 2: return exp(-(t - g->t0) / g->e_folding_time);
    default: return 1.0;
    }
}
static double get_density(const or_gas* g, const or_var4* rv)  /* pp_disk.cu:87-105 */
{
    double density = 0.0;
    double r = sqrt(SQR(rv->x) + SQR(rv->y));
    double h = g->sch[0] * pow(r, g->sch[1]);
    double arg = SQR(rv->z / h);
    if (0.1 < r) {
        density = g->rho[0] * pow(r, g->rho[1]) * exp(-arg);
    } else {
        double a = g->rho[0] * pow(0.1, g->rho[1] - 4.0);
        density = a * SQR(SQR(r)) * exp(-arg);
    }
    return density;
}
static or_var4 circular_velocity(double mu, const or_var4* rv)  /* pp_disk.cu:108-142 */
{
    or_var4 res = {0.0, 0.0, 0.0, 0.0};
    double r = sqrt(SQR(rv->x) + SQR(rv->y));
    double vc = sqrt(mu / r);
    double p;
    if (rv->x == 0.0 && rv->y == 0.0) {
        return res;
    } else if (rv->y == 0.0) {
        res.y = rv->x > 0.0 ? vc : -vc;
    } else if (rv->x == 0.0) {
        res.x = rv->y > 0.0 ? -vc : vc;
    } else if (rv->x >= rv->y) {
        p = rv->y / rv->x;
        res.y = rv->x >= 0 ? vc / sqrt(1.0 + SQR(p)) : -vc / sqrt(1.0 + SQR(p));
        res.x = -res.y * p;
    } else {
        p = rv->x / rv->y;
        res.x = rv->y >= 0 ? -vc / sqrt(1.0 + SQR(p)) : vc / sqrt(1.0 + SQR(p));
        res.y = -res.x * p;
    }
    return res;
}
void or_drag(double t, const or_gas* g, int drag_mode, uint32_t n, const uint32_t pl[OR_NTYPE],
             const or_param* p, const or_var4* r, const or_var4* v, or_var4* a)
{
    (void)n;
    double decr = reduction_factor(g, t);
    if (1.0e-6 > decr) return;
    uint32_t lo, hi;
    if (drag_mode == OR_DRAG_NSI) { lo = n_si_of(pl); hi = lo + n_nsi_of(pl); }
    else { lo = pl[OR_STAR]; hi = n_si_of(pl) + n_nsi_of(pl); }
    double m_star = p[0].mass;
    for (uint32_t i = lo; i < hi; i++) {
        or_var4 vg = circular_velocity(m_star, &r[i]);     /* pp_disk.cu:145-155 */
        double rr = sqrt(SQR(r[i].x) + SQR(r[i].y));
        double fac = sqrt(1.0 - 2.0 * g->eta[0] * pow(rr, g->eta[1]));
        vg.x *= fac; vg.y *= fac;
        double ux = vg.x - v[i].x, uy = vg.y - v[i].y, uz = vg.z - v[i].z;
        double u_n = sqrt(SQR(ux) + SQR(uy) + SQR(uz));
        double dens = get_density(g, &r[i]);
        double f = decr * (3.0 * p[i].cd * dens * u_n) / (8.0 * p[i].radius * p[i].density);
        a[i].x += f * ux; a[i].y += f * uy; a[i].z += f * uz;
    }
}

/* ------------------------------------------------------------------------ */
/* FARGO grid drag (pp_disk.cu:158-193, 778-814; util.cu:76-83)              */
/* ------------------------------------------------------------------------ */
or_var4 or_rotate_2d(double theta, or_var4 rv)   /* rotate_2D_vector, util.cu:76-83 */
{
    double ct = cos(theta);
    double st = sin(theta);
    or_var4 res = {ct * rv.x - st * rv.y, st * rv.x + ct * rv.y, 0.0, 0.0};
    return res;
}
/* calc_linear_index (pp_disk.cu:158-193), with the ring found by the
 * reference's linear scan over used_rad[0..n_rad].  Where the reference's
 * index is undefined or past the grid (a negative alpha from atan2(-0, x < 0)
 * converted to unsigned; sector n_sec in the last ring) this returns sector 0
 * and the last cell. */
uint32_t or_grid_cell(const or_grid* g, double x, double y)
{
    const double dalpha = OR_TWOPI / g->n_sec;
    double r = sqrt(SQR(x) + SQR(y));
    if (g->used_rad[0] > r) return 0;
    else if (g->used_rad[g->n_rad] < r) return g->n_rad * g->n_sec - 1;
    uint32_t i_rad = 0, i_sec = 0;
    for (uint32_t k = 0; k < g->n_rad; k++) {
        if (g->used_rad[k] <= r && r < g->used_rad[k + 1]) { i_rad = k; break; }
    }
    double alpha = y >= 0.0 ? atan2(y, x) : OR_TWOPI + atan2(y, x);
    double q = alpha / dalpha;
    i_sec = q >= 0.0 ? (uint32_t)q : 0u;
    uint32_t i_linear = i_rad * g->n_sec + i_sec;
    return i_linear < g->n_rad * g->n_sec ? i_linear : g->n_rad * g->n_sec - 1;
}
/* The reference's CPU FARGO loop (pp_disk.cu:778-814) over the drag range. */
void or_grid_drag(const or_grid* g, int drag_mode, const uint32_t pl[OR_NTYPE], const or_param* p,
                  const or_var4* r, const or_var4* v, or_var4* a)
{
    uint32_t lo, hi;
    if (drag_mode == OR_DRAG_NSI) { lo = n_si_of(pl); hi = lo + n_nsi_of(pl); }
    else { lo = pl[OR_STAR]; hi = n_si_of(pl) + n_nsi_of(pl); }
    double m_star = p[0].mass;
    for (uint32_t i = lo; i < hi; i++) {
        uint32_t idx = or_grid_cell(g, r[i].x, r[i].y);
        double r_norm = sqrt(SQR(r[i].x) + SQR(r[i].y));
        double vc_theta = sqrt(m_star / r_norm);
        double v_g_theta = vc_theta + g->vtheta[idx];
        double v_g_rad = g->vrad[idx];
        or_var4 v_g_polar = {v_g_rad, v_g_theta, 0.0, 0.0};
        double theta = r[i].y >= 0.0 ? atan2(r[i].y, r[i].x) : OR_TWOPI + atan2(r[i].y, r[i].x);
        or_var4 v_g = or_rotate_2d(theta, v_g_polar);
        double density_g = g->density[idx];
        double ux = v_g.x - v[i].x, uy = v_g.y - v[i].y, uz = v_g.z - v[i].z;
        double u_n = sqrt(SQR(ux) + SQR(uy) + SQR(uz));
        double f = (3.0 * p[i].cd * density_g * u_n) / (8.0 * p[i].radius * p[i].density);
        a[i].x += f * ux; a[i].y += f * uy; a[i].z += f * uz;
    }
}

/* ------------------------------------------------------------------------ */
/* pp_disk object                                                            */
/* ------------------------------------------------------------------------ */
struct or_disk {
    uint32_t n;
    uint32_t playing[OR_NTYPE], inactive[OR_NTYPE], removed[OR_NTYPE];
    or_var4 *r, *v, *rout, *vout;
    or_param* p;
    or_body_md* bmd;
    double t;
    double thr[3];
    int has_gas;
    or_gas gas;
    int has_grid;       /* a FARGO grid replaces the analytic disk */
    or_grid grid;       /* owns its arrays: radii, density, vrad, vtheta in one block */
    int drag_mode, formula;
    or_event* events; uint32_t event_counter, event_cap;
    or_event* log; uint32_t n_log, log_cap;
    uint32_t n_event[3];
};

or_disk* or_disk_create(uint32_t n, const uint32_t nbt[OR_NTYPE], const or_var4* r, const or_var4* v,
                        const or_param* p, const or_body_md* bmd, double t, const double thr[3], const or_gas* gas)
{
    or_disk* d = (or_disk*)calloc(1, sizeof(or_disk));
    d->n = n;
    memcpy(d->playing, nbt, sizeof(d->playing));
    d->r = (or_var4*)malloc(n * sizeof(or_var4)); d->v = (or_var4*)malloc(n * sizeof(or_var4));
    d->rout = (or_var4*)calloc(n, sizeof(or_var4)); d->vout = (or_var4*)calloc(n, sizeof(or_var4));
    d->p = (or_param*)malloc(n * sizeof(or_param)); d->bmd = (or_body_md*)malloc(n * sizeof(or_body_md));
    memcpy(d->r, r, n * sizeof(or_var4)); memcpy(d->v, v, n * sizeof(or_var4));
    memcpy(d->p, p, n * sizeof(or_param)); memcpy(d->bmd, bmd, n * sizeof(or_body_md));
    d->t = t;
    memcpy(d->thr, thr, sizeof(d->thr));
    if (gas) { d->has_gas = 1; d->gas = *gas; }
    d->event_cap = n < 16 ? 16 : n;
    d->events = (or_event*)calloc(d->event_cap, sizeof(or_event));
    d->log_cap = 64;
    d->log = (or_event*)calloc(d->log_cap, sizeof(or_event));
    return d;
}
void or_disk_destroy(or_disk* d)
{
    if (!d) return;
    free(d->r); free(d->v); free(d->rout); free(d->vout); free(d->p); free(d->bmd);
    free(d->events); free(d->log); free((void*)d->grid.used_rad); free(d);
}
int or_disk_set_gas_grid(or_disk* d, uint32_t n_rad, uint32_t n_sec, const double* used_rad, const double* density,
                         const double* vrad, const double* vtheta)
{
    free((void*)d->grid.used_rad);
    memset(&d->grid, 0, sizeof(d->grid));
    d->has_grid = 0;
    if (!used_rad) return OR_OK;
    if (n_rad == 0 || n_sec == 0) return OR_EINVAL;
    size_t nr = (size_t)n_rad + 1, nc = (size_t)n_rad * n_sec;
    double* b = (double*)malloc((nr + 3 * nc) * sizeof(double));
    if (!b) return OR_EINVAL;
    memcpy(b, used_rad, nr * sizeof(double));
    memcpy(b + nr, density, nc * sizeof(double));
    memcpy(b + nr + nc, vrad, nc * sizeof(double));
    memcpy(b + nr + 2 * nc, vtheta, nc * sizeof(double));
    d->grid.n_rad = n_rad; d->grid.n_sec = n_sec;
    d->grid.used_rad = b; d->grid.density = b + nr; d->grid.vrad = b + nr + nc; d->grid.vtheta = b + nr + 2 * nc;
    d->has_grid = 1;
    return OR_OK;
}
void or_disk_set_formula(or_disk* d, int f) { d->formula = f; }
void or_disk_set_drag_mode(or_disk* d, int m) { d->drag_mode = m; }
uint32_t or_disk_n(const or_disk* d) { return d->n; }
double or_disk_t(const or_disk* d) { return d->t; }
void or_disk_get(const or_disk* d, or_var4* r, or_var4* v, or_param* p, or_body_md* bmd)
{
    if (r) memcpy(r, d->r, d->n * sizeof(or_var4));
    if (v) memcpy(v, d->v, d->n * sizeof(or_var4));
    if (p) memcpy(p, d->p, d->n * sizeof(or_param));
    if (bmd) memcpy(bmd, d->bmd, d->n * sizeof(or_body_md));
}
void or_disk_counts(const or_disk* d, uint32_t pl[OR_NTYPE], uint32_t ia[OR_NTYPE], uint32_t rm[OR_NTYPE])
{
    if (pl) memcpy(pl, d->playing, sizeof(d->playing));
    if (ia) memcpy(ia, d->inactive, sizeof(d->inactive));
    if (rm) memcpy(rm, d->removed, sizeof(d->removed));
}
void or_disk_integral(const or_disk* d, or_integral* out)
{
    or_calc_integral(d->n, d->bmd, d->p, d->r, d->v, out);
}

/* pp_disk.cu:1450-1502 */
int or_calc_dydx(or_disk* d, uint32_t i, uint32_t rr, double t, const or_var4* r, const or_var4* v, or_var4* dy)
{
    (void)rr;
    if (i == 0) {
        memcpy(dy, v, d->n * sizeof(or_var4));
        return OR_OK;
    }
    if (i != 1) return OR_EINVAL;
    int rc = OR_OK;
    if (d->formula == OR_FORMULA_GPU) or_grav_gpu(d->n, d->playing, d->bmd, d->p, r, dy);
    else rc = or_grav_cpu(d->n, d->playing, d->bmd, d->p, r, dy);
    if (rc) return rc;
    if (d->has_grid) or_grid_drag(&d->grid, d->drag_mode, d->playing, d->p, r, v, dy);
    else if (d->has_gas) or_drag(t, &d->gas, d->drag_mode, d->n, d->playing, d->p, r, v, dy);
    return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Events                                                                    */
/* ------------------------------------------------------------------------ */
/* pp_disk.cu:196-253 (the device printf is omitted) */
static void store_event(int name, double t, double dist, uint32_t idx1, uint32_t idx2, const or_param* p,
                        const or_var4* r, const or_var4* v, const or_body_md* bmd, or_event* e)
{
    memset(e, 0, sizeof(*e));
    e->event_name = name; e->d = dist; e->t = t;
    e->id0 = bmd[0].id; e->id1 = bmd[idx1].id; e->id2 = bmd[idx2].id;
    e->idx0 = 0; e->idx1 = idx1; e->idx2 = idx2;
    e->r0 = r[0]; e->v0 = v[0]; e->r1 = r[idx1]; e->v1 = v[idx1]; e->r2 = r[idx2]; e->v2 = v[idx2];
    e->p0 = p[0];
    if (OR_EV_EJECTION == name) {
        e->p1 = p[idx1]; e->p2 = p[idx2];
        e->rs = e->r1; e->vs = e->v1; e->ps = e->p1;
    }
}
static int push_event(or_disk* d, const or_event* e)
{
    if (d->event_counter >= d->event_cap) return OR_EOVERFLOW;
    d->events[d->event_counter++] = *e;
    return OR_OK;
}
/* GPU kernel semantics pp_disk.cu:284-318 over [0, n_total) */
uint32_t or_check_ejection_hit_centrum(or_disk* d)
{
    double ed2 = (0.0 < d->thr[OR_THR_EJ] ? SQR(d->thr[OR_THR_EJ]) : DBL_MAX);
    double hcd2 = SQR(d->thr[OR_THR_HC]);
    or_event e;
    for (uint32_t i = 0; i < d->n; i++) {
        if (d->bmd[i].id > 0 && d->bmd[i].body_type != OR_STAR) {
            double r2 = SQR(d->r[i].x) + SQR(d->r[i].y) + SQR(d->r[i].z);
            if (ed2 < r2) {
                store_event(OR_EV_EJECTION, d->t, sqrt(r2), 0, i, d->p, d->r, d->v, d->bmd, &e);
                push_event(d, &e);
            }
            if (hcd2 > r2) {
                store_event(OR_EV_HIT_CENTRUM, d->t, sqrt(r2), 0, i, d->p, d->r, d->v, d->bmd, &e);
                push_event(d, &e);
            }
        }
    }
    return d->event_counter;
}
/* GPU kernel semantics pp_disk.cu:321-371 (the star takes part as i);
 * events in canonical (i, j) ascending order */
int or_check_collision(or_disk* d, uint32_t* n_events)
{
    const double f = SQR(d->thr[OR_THR_RAD]);
    or_event e;
    for (uint32_t i = 0; i < d->n; i++) {
        if (!(0 < d->bmd[i].id)) continue;
        for (uint32_t j = i + 1; j < d->n; j++) {
            if (0 > d->bmd[j].id) continue;
            double dx = d->r[j].x - d->r[i].x;
            double dy = d->r[j].y - d->r[i].y;
            double dz = d->r[j].z - d->r[i].z;
            double r2 = SQR(dx) + SQR(dy) + SQR(dz);
            if (r2 < f * SQR((d->p[i].radius + d->p[j].radius))) {
                uint32_t s = i, m = j;
                if (d->p[m].mass > d->p[s].mass) { s = j; m = i; }
                store_event(OR_EV_COLLISION, d->t, sqrt(r2), s, m, d->p, d->r, d->v, d->bmd, &e);
                if (push_event(d, &e)) { *n_events = d->event_counter; return OR_EOVERFLOW; }
            }
        }
    }
    *n_events = d->event_counter;
    return OR_OK;
}
static void log_event(or_disk* d, const or_event* e)
{
    if (d->n_log == d->log_cap) {
        d->log_cap *= 2;
        d->log = (or_event*)realloc(d->log, d->log_cap * sizeof(or_event));
    }
    d->log[d->n_log++] = *e;
}
static void inc_counters(or_disk* d) { for (int i = 0; i < 3; i++) d->n_event[i]++; }
/* pp_disk.cu:1328-1368 */
static int handle_collision_pair(or_disk* d, or_event* c)
{
    uint32_t s = c->idx1, m = c->idx2;
    if (OR_SUPERPL == d->bmd[m].body_type) return OR_ENOTIMPL;
    c->p1 = d->p[s];
    c->p2 = d->p[m];
    or_calc_position_after_collision(c->p1.mass, c->p2.mass, &c->r1, &c->r2, &c->rs);
    or_calc_velocity_after_collision(c->p1.mass, c->p2.mass, &c->v1, &c->v2, &c->vs);
    c->rs.w = 0.0; c->vs.w = 0.0;
    d->r[s] = c->rs;
    d->v[s] = c->vs;
    or_calc_physical_properties(&c->p1, &c->p2, &c->ps);
    d->p[s] = c->ps;
    return OR_OK;
}
/* pp_disk.cu:1250-1288: one entry per (id1, id2) with the earliest t */
static uint32_t populate_sp_events(or_disk* d, or_event* sp)
{
    uint32_t ne = d->event_counter, n = 0;
    char* done = (char*)calloc(ne ? ne : 1, 1);
    for (uint32_t k = 0; k < ne; k++) {
        if (done[k]) continue;
        done[k] = 1;
        sp[n] = d->events[k];
        for (uint32_t i = k + 1; i < ne; i++) {
            if (sp[n].id1 == d->events[i].id1 && sp[n].id2 == d->events[i].id2) {
                done[i] = 1;
                if (sp[n].t > d->events[i].t) sp[n] = d->events[i];
            }
        }
        n++;
    }
    free(done);
    return n;
}
/* pp_disk.cu:1220-1248 */
int or_handle_collision(or_disk* d)
{
    or_event* sp = (or_event*)calloc(d->event_counter ? d->event_counter : 1, sizeof(or_event));
    uint32_t n = populate_sp_events(d, sp);
    for (uint32_t i = 0; i < n; i++) {
        uint32_t ictv = sp[i].idx2;
        int rc = handle_collision_pair(d, &sp[i]);
        if (rc) { free(sp); return rc; }
        inc_counters(d);
        d->bmd[ictv].id *= -1;
        d->inactive[d->bmd[ictv].body_type]++;
        log_event(d, &sp[i]);
    }
    free(sp);
    return OR_OK;
}
/* pp_disk.cu:1290-1326 */
int or_handle_ejection_hit_centrum(or_disk* d)
{
    for (uint32_t i = 0; i < d->event_counter; i++) {
        or_event e = d->events[i];
        uint32_t ictv = e.idx2;
        if (OR_EV_EJECTION != e.event_name) {
            int rc = handle_collision_pair(d, &e);
            if (rc) return rc;
        }
        inc_counters(d);
        d->bmd[ictv].id *= -1;
        d->inactive[d->bmd[ictv].body_type]++;
        log_event(d, &e);
    }
    return OR_OK;
}
void or_clear_event_counter(or_disk* d) { d->event_counter = 0; }
uint32_t or_event_log(const or_disk* d, or_event* out, uint32_t cap)
{
    uint32_t n = d->n_log < cap ? d->n_log : cap;
    if (out) memcpy(out, d->log, n * sizeof(or_event));
    return d->n_log;
}
/* pp_disk.cu:1370-1430 + n_objects_t::update (red_type.h:545-560) */
int or_rebuild_vectors(or_disk* d)
{
    uint32_t k = 0;
    for (uint32_t i = 0; i < d->n; i++) {
        if (0 < d->bmd[i].id) {
            d->r[k] = d->r[i]; d->v[k] = d->v[i]; d->p[k] = d->p[i]; d->bmd[k] = d->bmd[i];
            k++;
        }
    }
    for (int t = 0; t < OR_NTYPE; t++) {
        d->playing[t] -= d->inactive[t];
        d->removed[t] += d->inactive[t];
        d->inactive[t] = 0;
    }
    if (k != n_total_of(d->playing)) return OR_EINVAL;
    d->n = k;
    return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Integrators                                                               */
/* ------------------------------------------------------------------------ */
/* int_rungekutta8.cu:26-48 */
static const double rk8_c[13] = { 0.0, 2.0/27.0, 1.0/9.0, 1.0/6.0, 5.0/12.0, 1.0/2.0, 5.0/6.0, 1.0/6.0, 2.0/3.0, 1.0/3.0, 1.0, 0.0, 1.0 };
static const double rk8_a[79] = {
    0.0,
    2.0/27.0,
    1.0/36.0, 1.0/12.0,
    1.0/24.0, 0.0, 1.0/8.0,
    5.0/12.0, 0.0, -25.0/16.0, 25.0/16.0,
    1.0/20.0, 0.0, 0.0, 1.0/4.0, 1.0/5.0,
    -25.0/108.0, 0.0, 0.0, 125.0/108.0, -65.0/27.0, 125.0/54.0,
    31.0/300.0, 0.0, 0.0, 0.0, 61.0/225.0, -2.0/9.0, 13.0/900.0,
    2.0, 0.0, 0.0, -53.0/6.0, 704.0/45.0, -107.0/9.0, 67.0/90.0, 3.0,
    -91.0/108.0, 0.0, 0.0, 23.0/108.0, -976.0/135.0, 311.0/54.0, -19.0/60.0, 17.0/6.0, -1.0/12.0,
    2383.0/4100.0, 0.0, 0.0, -341.0/164.0, 4496.0/1025.0, -301.0/82.0, 2133.0/4100.0, 45.0/82.0, 45.0/164.0, 18.0/41.0,
    3.0/205.0, 0.0, 0.0, 0.0, 0.0, -6.0/41.0, -3.0/205.0, -3.0/41.0, 3.0/41.0, 6.0/41.0, 0.0,
    -1777.0/4100.0, 0.0, 0.0, -341.0/164.0, 4496.0/1025.0, -289.0/82.0, 2193.0/4100.0, 51.0/82.0, 33.0/164.0, 12.0/41.0, 0.0, 1.0 };
static const double rk8_b[11] = { 41.0/840.0, 0.0, 0.0, 0.0, 0.0, 34.0/105.0, 9.0/35.0, 9.0/35.0, 9.0/280.0, 9.0/280.0, 41.0/840.0 };
/* nonzero a-columns per stage r, int_rungekutta8.cu:414-585 */
static const int rk8_nnz[13] = { 0, 1, 2, 2, 3, 3, 4, 4, 6, 7, 8, 6, 9 };
static const int rk8_cols[13][9] = {
    {0}, {0}, {0, 1}, {0, 2}, {0, 2, 3}, {0, 3, 4}, {0, 3, 4, 5}, {0, 4, 5, 6},
    {0, 3, 4, 5, 6, 7}, {0, 3, 4, 5, 6, 7, 8}, {0, 3, 4, 5, 6, 7, 8, 9},
    {0, 5, 6, 7, 8, 9}, {0, 3, 4, 5, 6, 7, 8, 9, 11} };
static const int rk8_row[13] = { 0, 1, 2, 4, 7, 11, 16, 22, 29, 37, 46, 56, 67 };
void or_rk8_tableau(double c[13], double a[79], double b[11])
{
    memcpy(c, rk8_c, sizeof(rk8_c));
    memcpy(a, rk8_a, sizeof(rk8_a));
    memcpy(b, rk8_b, sizeof(rk8_b));
}
/* int_rungekutta4.cu:26-31 */
static const double rk4_a[8] = { 0.0, 1.0/2.0, 1.0/2.0, 1.0, 1.0/6.0, 1.0/3.0, 1.0/3.0, 1.0/6.0 };
static const double rk4_b[5] = { 1.0/6.0, 1.0/3.0, 1.0/3.0, 1.0/6.0 - 1.0/10.0, 1.0/10.0 };
static const double rk4_c[5] = { 0.0, 1.0/2.0, 1.0/2.0, 1.0, 1.0 };

struct or_integ {
    or_disk* d;
    int method, adaptive, n_stage, order, error_check_for_tp;
    double tol, dt_try, dt_did, dt_next, last_err;
    uint64_t tried, passed, failed;
    uint32_t max_attempts;
    double min_dt;
    int nonfinite, last_bad;
    or_var4* k[2][13];
    or_var4* ytemp[2];
    uint32_t cap;
};

static void integ_alloc(or_integ* it, uint32_t n)
{
    for (int i = 0; i < 2; i++) {
        free(it->ytemp[i]);
        it->ytemp[i] = (or_var4*)calloc(n, sizeof(or_var4));
        for (int r = 0; r < it->n_stage; r++) {
            free(it->k[i][r]);
            it->k[i][r] = (or_var4*)calloc(n, sizeof(or_var4));
        }
    }
    it->cap = n;
}

or_integ* or_integ_create(or_disk* d, int method, double dt, int adaptive, double tol)
{
    or_integ* it = (or_integ*)calloc(1, sizeof(or_integ));
    it->d = d; it->method = method; it->adaptive = adaptive; it->tol = tol;
    it->dt_try = dt; it->dt_next = dt;
    if (method == 4) { it->n_stage = adaptive ? 13 : 11; it->order = 7; }
    else if (method == 2) { it->n_stage = adaptive ? 5 : 4; it->order = 4; }
    /* euler / rungekutta2 construct the base with adaptive = false
       (int_euler.cu:48-54, int_rungekutta2.cu:49-55) */
    else if (method == 1) { it->n_stage = 2; it->order = 2; it->adaptive = 0; }
    else { it->n_stage = 1; it->order = 1; it->adaptive = 0; }
    integ_alloc(it, d->n);
    return it;
}
void or_integ_destroy(or_integ* it)
{
    if (!it) return;
    for (int i = 0; i < 2; i++) {
        free(it->ytemp[i]);
        for (int r = 0; r < 13; r++) free(it->k[i][r]);
    }
    free(it);
}
double or_get_dt_next(const or_integ* it) { return it->dt_next; }
void or_set_dt_next(or_integ* it, double dt) { it->dt_next = dt; }
void or_counters(const or_integ* it, uint64_t out[3]) { out[0] = it->tried; out[1] = it->passed; out[2] = it->failed; }
void or_set_error_check_for_tp(or_integ* it, int on) { it->error_check_for_tp = on; }

static inline double* comp(or_var4* a) { return (double*)a; }
static inline const double* ccomp(const or_var4* a) { return (const double*)a; }
double or_last_error(const or_integ* it) { return it->last_err; }
void or_set_guard(or_integ* it, uint32_t max_attempts, double min_dt, int nonfinite)
{
    it->max_attempts = max_attempts;
    it->min_dt = min_dt;
    it->nonfinite = nonfinite;
}
/* the guard's decision on the attempt just judged (iter counts it) */
static int guard_trips(const or_integ* it, double max_err, int iter)
{
    if (it->nonfinite && (it->last_bad || !isfinite(max_err))) return 1;
    if (max_err > it->tol && it->max_attempts && (uint32_t)iter >= it->max_attempts) return 1;
    return 0;
}
/* the guard's decision on the attempt about to be made (iter made so far) */
static int guard_refuses_dt(or_integ* it, int iter)
{
    if (!(it->adaptive && 0.0 < it->min_dt && fabs(it->dt_try) <= it->min_dt)) return 0;
    it->tried += iter; it->failed += iter;
    return 1;
}
uint32_t or_last_error_terms(const or_integ* it, double* out, uint32_t cap)
{
    if (!it->adaptive || (it->method != 4 && it->method != 2)) return 0;
    const uint32_t nv = 4 * it->d->n;
    for (int i = 0; i < 2; i++)
        for (uint32_t t = 0; t < nv && i * nv + t < cap; t++) {
            double e;
            if (it->method == 4)
                e = fabs(comp(it->k[i][0])[t] + comp(it->k[i][10])[t] - comp(it->k[i][11])[t] - comp(it->k[i][12])[t]);
            else
                e = fabs(comp(it->k[i][3])[t] - comp(it->k[i][4])[t]);
            out[i * nv + t] = e;
        }
    return 2 * nv;
}

/* RKF78 ytemp for stage r: ytemp = y_n + dt * (sum_j a_rj f_j), left to right
 * over the nonzero columns (int_rungekutta8.cu:82-246) */
static void rk8_ytemp(or_integ* it, int r, uint32_t nv)
{
    const double* a = rk8_a + rk8_row[r];
    for (int i = 0; i < 2; i++) {
        const double* y = ccomp(i == 0 ? it->d->r : it->d->v);
        double* yt = comp(it->ytemp[i]);
        for (uint32_t t = 0; t < nv; t++) {
            double s = a[rk8_cols[r][0]] * comp(it->k[i][rk8_cols[r][0]])[t];
            for (int q = 1; q < rk8_nnz[r]; q++) {
                int c = rk8_cols[r][q];
                s = s + a[c] * comp(it->k[i][c])[t];
            }
            yt[t] = y[t] + it->dt_try * s;
        }
    }
}
/* int_rungekutta8.cu:264-274 */
static void rk8_ynp1(or_integ* it, uint32_t nv)
{
    const double* b = rk8_b;
    for (int i = 0; i < 2; i++) {
        const double* y = ccomp(i == 0 ? it->d->r : it->d->v);
        double* yo = comp(i == 0 ? it->d->rout : it->d->vout);
        double* f[11];
        for (int q = 0; q < 11; q++) f[q] = comp(it->k[i][q]);
        for (uint32_t t = 0; t < nv; t++)
            yo[t] = y[t] + it->dt_try * (b[0] * f[0][t] + b[5] * f[5][t] + b[6] * (f[6][t] + f[7][t]) + b[8] * (f[8][t] + f[9][t]) + b[10] * f[10][t]);
    }
}
/* integrator.cu:248-263 over err = |f0 + f10 - f11 - f12| (int_rungekutta8.cu:250-260) */
static double rk8_max_err(or_integ* it, uint32_t nv)
{
    double m[2] = {0.0, 0.0};
    it->last_bad = 0;
    for (int i = 0; i < 2; i++) {
        const double *f0 = comp(it->k[i][0]), *f10 = comp(it->k[i][10]), *f11 = comp(it->k[i][11]), *f12 = comp(it->k[i][12]);
        for (uint32_t t = 0; t < nv; t++) {
            double e = fabs(f0[t] + f10[t] - f11[t] - f12[t]);
            if (m[i] < e) m[i] = e;          /* a NaN term never compares greater */
            if (!(e <= DBL_MAX)) it->last_bad = 1;
        }
    }
    return m[0] > m[1] ? m[0] : m[1];
}
static void swap_state(or_disk* d)  /* pp_disk.cu:1518-1529 */
{
    or_var4* t = d->r; d->r = d->rout; d->rout = t;
    t = d->v; d->v = d->vout; d->vout = t;
}
static int dydx_pair(or_integ* it, int r, double t, const or_var4* rr, const or_var4* vv)
{
    int rc = or_calc_dydx(it->d, 0, r, t, rr, vv, it->k[0][r]);
    if (!rc) rc = or_calc_dydx(it->d, 1, r, t, rr, vv, it->k[1][r]);
    return rc;
}
/* int_rungekutta8.cu:635-716 */
static int rk8_step(or_integ* it, double* dt_did)
{
    or_disk* d = it->d;
    const uint32_t nvt = 4 * d->n;
    if (it->cap < d->n) integ_alloc(it, d->n);
    it->dt_try = it->dt_next;
    double ttemp = d->t + rk8_c[0] * it->dt_try;
    int rc = dydx_pair(it, 0, ttemp, d->r, d->v);
    if (rc) return rc;
    double max_err = 0.0;
    int iter = 0;
    do {
        if (guard_refuses_dt(it, iter)) return OR_ESTATE;
        it->dt_did = it->dt_try;
        for (int r = 1; r <= 10; r++) {
            ttemp = d->t + rk8_c[r] * it->dt_try;
            rk8_ytemp(it, r, nvt);
            if ((rc = dydx_pair(it, r, ttemp, it->ytemp[0], it->ytemp[1]))) return rc;
        }
        rk8_ynp1(it, nvt);
        if (it->adaptive) {
            for (int r = 11; r < it->n_stage; r++) {
                ttemp = d->t + rk8_c[r] * it->dt_try;
                rk8_ytemp(it, r, nvt);
                if ((rc = dydx_pair(it, r, ttemp, it->ytemp[0], it->ytemp[1]))) return rc;
            }
            uint32_t nmass = n_si_of(d->playing) + n_nsi_of(d->playing);
            uint32_t nv = 4 * (it->error_check_for_tp ? d->n : nmass);
            max_err = fabs(41.0 / 840.0 * it->dt_try * rk8_max_err(it, nv));
            it->last_err = max_err;
            if (0.0 < max_err) it->dt_next = it->dt_try *= 0.9 * pow(it->tol / max_err, 1.0 / (it->order));
            else it->dt_next = it->dt_try *= 2.0;
            if (guard_trips(it, max_err, iter + 1)) {
                it->tried += iter + 1; it->failed += iter + 1;
                return OR_ESTATE;
            }
        }
        iter++;
    } while (it->adaptive && max_err > it->tol);
    it->tried += iter; it->failed += iter - 1; it->passed++;
    d->t += it->dt_did;
    swap_state(d);
    *dt_did = it->dt_did;
    return OR_OK;
}
/* int_rungekutta4.cu:173-247 */
static int rk4_step(or_integ* it, double* dt_did)
{
    or_disk* d = it->d;
    const uint32_t nvt = 4 * d->n;
    if (it->cap < d->n) integ_alloc(it, d->n);
    it->dt_try = it->dt_next;
    double ttemp = d->t + rk4_c[0] * it->dt_try;
    int rc = dydx_pair(it, 0, ttemp, d->r, d->v);
    if (rc) return rc;
    double max_err = 0.0;
    int iter = 0;
    do {
        if (guard_refuses_dt(it, iter)) return OR_ESTATE;
        it->dt_did = it->dt_try;
        for (int s = 1; s < 4; s++) {
            ttemp = d->t + rk4_c[s] * it->dt_try;
            for (int i = 0; i < 2; i++) {     /* int_rungekutta4.cu:37-47, 130-148 */
                const double* y = ccomp(i == 0 ? d->r : d->v);
                const double* f = comp(it->k[i][s - 1]);
                double* yt = comp(it->ytemp[i]);
                double fac = rk4_a[s] * it->dt_try;
                for (uint32_t t = 0; t < nvt; t++) yt[t] = y[t] + fac * f[t];
            }
            if ((rc = dydx_pair(it, s, ttemp, it->ytemp[0], it->ytemp[1]))) return rc;
        }
        for (int i = 0; i < 2; i++) {         /* int_rungekutta4.cu:50-60, 150-171 */
            const double* y = ccomp(i == 0 ? d->r : d->v);
            double* yo = comp(i == 0 ? d->rout : d->vout);
            const double *f1 = comp(it->k[i][0]), *f2 = comp(it->k[i][1]), *f3 = comp(it->k[i][2]), *f4 = comp(it->k[i][3]);
            double b0 = rk4_b[0] * it->dt_try, b1 = rk4_b[1] * it->dt_try;
            for (uint32_t t = 0; t < nvt; t++) yo[t] = y[t] + b0 * (f1[t] + f4[t]) + b1 * (f2[t] + f3[t]);
        }
        if (it->adaptive) {
            ttemp = d->t + rk4_c[4] * it->dt_try;
            if ((rc = dydx_pair(it, 4, ttemp, d->rout, d->vout))) return rc;
            uint32_t nmass = n_si_of(d->playing) + n_nsi_of(d->playing);
            uint32_t nv = 4 * (it->error_check_for_tp ? d->n : nmass);
            double m = 0.0;
            it->last_bad = 0;
            for (int i = 0; i < 2; i++) {
                const double *f4 = comp(it->k[i][3]), *f5 = comp(it->k[i][4]);
                for (uint32_t t = 0; t < nv; t++) {
                    double e = fabs(f4[t] - f5[t]);
                    if (m < e) m = e;
                    if (!(e <= DBL_MAX)) it->last_bad = 1;
                }
            }
            max_err = fabs(1.0 / 10.0 * it->dt_try * m);
            it->last_err = max_err;
            if (0.0 < max_err) it->dt_next = it->dt_try *= 0.9 * pow(it->tol / max_err, 1.0 / (it->order));
            else it->dt_next = it->dt_try *= 2.0;
            if (guard_trips(it, max_err, iter + 1)) {
                it->tried += iter + 1; it->failed += iter + 1;
                return OR_ESTATE;
            }
        }
        iter++;
    } while (it->adaptive && max_err > it->tol);
    it->tried += iter; it->failed += iter - 1; it->passed++;
    d->t += it->dt_did;
    swap_state(d);
    *dt_did = it->dt_did;
    return OR_OK;
}
/* y_out = y + fac * f over both halves (the sum_vector kernels of
   int_euler.cu:26-36 and int_rungekutta2.cu:30-40) */
static void sum_halves(or_integ* it, const or_var4* const y[2], or_var4* const out[2], int r, double fac)
{
    const uint32_t nvt = 4 * it->d->n;
    for (int i = 0; i < 2; i++) {
        const double* a = ccomp(y[i]);
        const double* f = comp(it->k[i][r]);
        double* o = comp(out[i]);
        for (uint32_t t = 0; t < nvt; t++) o[t] = a[t] + fac * f[t];
    }
}
/* int_euler.cu:78-105.  The step is always dt_try, the constructor's dt:
   euler::step never takes dt_next, which it sets to dt_try. */
static int euler_step(or_integ* it, double* dt_did)
{
    or_disk* d = it->d;
    if (it->cap < d->n) integ_alloc(it, d->n);
    int rc = dydx_pair(it, 0, d->t, d->r, d->v);
    if (rc) return rc;
    const or_var4* y[2] = {d->r, d->v};
    or_var4* out[2] = {d->rout, d->vout};
    sum_halves(it, y, out, 0, it->dt_try);
    it->dt_did = it->dt_try;
    it->dt_next = it->dt_try;
    it->tried++; it->passed++;
    d->t += it->dt_did;
    swap_state(d);
    *dt_did = it->dt_did;
    return OR_OK;
}
/* int_rungekutta2.cu:99-131: the midpoint rule, a = {0, 1/2}, b = {0, 1},
   c = {0, 1/2} (int_rungekutta2.cu:13-15); dt_try fixed as for Euler. */
static int rk2_step(or_integ* it, double* dt_did)
{
    or_disk* d = it->d;
    if (it->cap < d->n) integ_alloc(it, d->n);
    int rc = dydx_pair(it, 0, d->t + 0.0 * it->dt_try, d->r, d->v);
    if (rc) return rc;
    const or_var4* y[2] = {d->r, d->v};
    or_var4* tmp[2] = {it->ytemp[0], it->ytemp[1]};
    sum_halves(it, y, tmp, 0, (1.0 / 2.0) * it->dt_try);
    if ((rc = dydx_pair(it, 1, d->t + (1.0 / 2.0) * it->dt_try, it->ytemp[0], it->ytemp[1]))) return rc;
    or_var4* out[2] = {d->rout, d->vout};
    sum_halves(it, y, out, 1, 1.0 * it->dt_try);
    it->dt_did = it->dt_try;
    it->dt_next = it->dt_try;
    it->tried++; it->passed++;
    d->t += it->dt_did;
    swap_state(d);
    *dt_did = it->dt_did;
    return OR_OK;
}
int or_step(or_integ* it, double* dt_did)
{
    switch (it->method) {
    case 0: return euler_step(it, dt_did);
    case 1: return rk2_step(it, dt_did);
    case 4: return rk8_step(it, dt_did);
    default: return rk4_step(it, dt_did);
    }
}

/* src/red.cuda/main.cu:470-625 without output, dumps or device switching */
int or_run(or_disk* d, or_integ* it, double length, double output_interval, uint64_t max_steps, uint64_t* steps_done)
{
    double elapsed = 0.0, ps = 0.0, dt = 0.0;
    uint64_t steps = 0;
    int rc;
    for (;;) {
        uint32_t active = 0;
        for (uint32_t i = 0; i < d->n; i++) active += d->bmd[i].id > 0;
        if (!(fabs(elapsed) <= fabs(length) && 1 < active) || steps >= max_steps) break;
        if (0.0 < d->thr[OR_THR_EJ] || 0.0 < d->thr[OR_THR_HC]) {
            if (or_check_ejection_hit_centrum(d)) {
                if ((rc = or_handle_ejection_hit_centrum(d))) return rc;
                or_clear_event_counter(d);
            }
        }
        if ((rc = or_step(it, &dt))) return rc;
        steps++;
        elapsed += dt; ps += dt;
        if (fabs(output_interval) < fabs(it->dt_next))
            it->dt_next = 0.0 < dt ? output_interval : -output_interval;
        if (0.0 < d->thr[OR_THR_RAD]) {
            uint32_t ne = 0;
            if ((rc = or_check_collision(d, &ne))) return rc;
            if (ne) {
                if ((rc = or_handle_collision(d))) return rc;
                or_clear_event_counter(d);
            }
        }
        if (output_interval <= fabs(ps)) ps = 0.0;
        if (16 <= d->n_event[1]) {
            d->n_event[1] = 0;
            if ((rc = or_rebuild_vectors(d))) return rc;
        }
    }
    if (steps_done) *steps_done = steps;
    return OR_OK;
}
