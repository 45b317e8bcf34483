// ic.cpp -- seeded planetesimal disks (rd_make_disk).
//
// Follows redinit's pipeline (src/red.cuda.initial/main.cu:251-342,
// 1662-1670): orbital elements from MT19937-driven samplers
// (src/red.cuda.initial/distribution.cpp:11-162), Cartesian state with
// mu = K^2 (m0 + m) (tools.cpp:789-828), barycentric transform
// (tools.cpp:412-441), velocities to tau units (tools.cpp:443-453).
#include <cmath>
#include <cstring>
#include <random>
#include <string>

#include "../../include/red_abi.h"

namespace {

constexpr double kK = 0.01720209895;                  // red_macro.h:13
constexpr double kK2 = 0.0002959122082855911025;      // red_macro.h:14
constexpr double kPi = 3.1415926535897932384626;
constexpr double kTwoPi = 6.2831853071795864769253;

// MT19937 as in distribution.cpp:11-31 (identical to std::mt19937).
class Mt {
  public:
    explicit Mt(uint32_t seed) : g_(seed) {}
    uint32_t u32() { return (uint32_t)g_(); }
    // value in (lo, hi], distribution.cpp:38-41
    double uniform(double lo, double hi) { return lo + (hi - lo) * (u32() + 1.0) / 4294967296.0; }
    // distribution.cpp:103-109
    double rayleigh(double sigma) { return sigma * std::sqrt(-2.0 * std::log((u32() + 1.0) / 4294967296.0)); }
    // Rejection sampling under x^power with the trial sequence of
    // distribution.cpp:139-162; y is drawn on (0, max(y)] so the accepted x
    // follow x^power (the reference's y range [min(y), max(y)] samples
    // x^power - min(y) instead).
    double power_law(double x_min, double x_max, double power)
    {
        double y_min = power == 0.0 ? 0.0 : std::pow(x_min, power);
        double y_max = power == 0.0 ? 1.0 : std::pow(x_max, power);
        if (y_min > y_max) std::swap(y_min, y_max);
        const double d_y = y_max, d_x = x_max - x_min, area = d_x * d_y;
        double x, y;
        do {
            x = uniform(0.0, area) / d_y + x_min;
            y = uniform(0.0, y_max);
        } while (y > std::pow(x, power));
        return x;
    }

  private:
    std::mt19937 g_;
};

double radius_of(double m, double density)   // tools.cpp:455-460
{
    const double four_pi_over_three = 4.188790204786391;
    return std::pow(1.0 / four_pi_over_three * m / density, 1.0 / 3.0);
}
double mass_of(double R, double density)     // tools.cpp:473-478
{
    const double four_pi_over_three = 4.188790204786391;
    return four_pi_over_three * (R * R * R) * density;
}
double density_of(double m, double R)        // tools.cpp:462-471
{
    const double four_pi_over_three = 4.188790204786391;
    return R == 0.0 ? 0.0 : m / (four_pi_over_three * (R * R * R));
}

// E - e sin E = M by Newton from the starting guess of tools.cpp:724, with
// E(M) = 2 pi - E(2 pi - M) for M > pi.
bool kepler(double ecc, double mean, double* E)
{
    if (0.0 == ecc || 0.0 == mean || kPi == mean) { *E = mean; return true; }
    bool mirror = false;
    if (kPi < mean) { mean = kTwoPi - mean; mirror = true; }
    double x = mean + ecc * std::sin(mean) / (1.0 - std::sin(mean + ecc) + std::sin(mean));
    double err;
    int step = 0;
    do {
        const double x1 = x - (x - ecc * std::sin(x) - mean) / (1.0 - ecc * std::cos(x));
        err = std::fabs(x1 - x);
        x = x1;
    } while (err > 1.0e-14 && step++ <= 32);
    if (step > 32) return false;
    *E = mirror ? kTwoPi - x : x;
    return true;
}

// orbital elements -> Cartesian, tools.cpp:789-828
bool phase(double mu, double a, double e, double inc, double peri, double node, double mean, rd_var4* r, rd_var4* v)
{
    double E;
    if (!kepler(e, mean, &E)) return false;
    const double nu = 2.0 * std::atan(std::sqrt((1.0 + e) / (1.0 - e)) * std::tan(E / 2.0));
    const double p = a * (1.0 - e * e);
    const double rr = p / (1.0 + e * std::cos(nu));
    const double kszi = rr * std::cos(nu), eta = rr * std::sin(nu);
    const double vk = -std::sqrt(mu / p) * std::sin(nu), ve = std::sqrt(mu / p) * (e + std::cos(nu));
    const double cw = std::cos(peri), sw = std::sin(peri), cO = std::cos(node), sO = std::sin(node);
    const double ci = std::cos(inc), si = std::sin(inc);
    const double Px = cw * cO - sw * sO * ci, Py = cw * sO + sw * cO * ci, Pz = sw * si;
    const double Qx = -sw * cO - cw * sO * ci, Qy = -sw * sO + cw * cO * ci, Qz = cw * si;
    *r = rd_var4{kszi * Px + eta * Qx, kszi * Py + eta * Qy, kszi * Pz + eta * Qz, 0.0};
    *v = rd_var4{vk * Px + ve * Qx, vk * Py + ve * Qy, vk * Pz + ve * Qz, 0.0};
    return true;
}

}  // namespace

extern "C" int rd_make_disk(const rd_disk_spec* s, rd_var4* r, rd_var4* v, rd_param* p, rd_body_md* bmd)
{
    if (!s || !r || !v || !p || !bmd || !(s->a_min > 0.0) || !(s->a_max >= s->a_min) || !(s->lo > 0.0) ||
        !(s->hi >= s->lo) || !(s->density_cgs > 0.0))
        return RD_EINVAL;
    const double AuToMeter = 1.495978707e11;          // red_constants.h:58
    const double MeterToAu = 1.0 / AuToMeter;
    const double SolarToGram = 1.0e3 * 1.98911e30;   // red_constants.h:50-51
    const double cm = 1.0e-2 * MeterToAu;
    const double GramPerCm3ToSolarPerAu3 = (1.0 / SolarToGram) / (cm * cm * cm);   // red_constants.h:80
    const double KilometerToAu = 1.0 / (AuToMeter / 1.0e3);
    const double rho = s->density_cgs * GramPerCm3ToSolarPerAu3;
    const uint32_t n = s->n + 1;
    Mt g(s->seed);
    std::memset(r, 0, n * sizeof(rd_var4));
    std::memset(v, 0, n * sizeof(rd_var4));
    p[0] = rd_param{s->m_star, s->r_star, density_of(s->m_star, s->r_star), 0.0};
    bmd[0] = rd_body_md{1, RD_BODY_STAR, 0, 0.0};
    for (uint32_t i = 1; i < n; ++i) {
        const double a = g.power_law(s->a_min, s->a_max, s->a_power);
        const double e = g.rayleigh(s->sig_e);
        const double inc = g.rayleigh(s->sig_i);
        double peri = g.uniform(0.0, kTwoPi), node = g.uniform(0.0, kTwoPi), mean = g.uniform(0.0, kTwoPi);
        if (peri >= kTwoPi) peri -= kTwoPi;
        if (node >= kTwoPi) node -= kTwoPi;
        if (mean >= kTwoPi) mean -= kTwoPi;
        const double u = g.uniform(std::log10(s->lo), std::log10(s->hi));
        if (s->size_mode == 0) {
            p[i].mass = std::pow(10.0, u);
            p[i].density = rho;
            p[i].radius = radius_of(p[i].mass, rho);
        } else {
            p[i].radius = std::pow(10.0, u) * KilometerToAu;
            p[i].density = rho;
            p[i].mass = mass_of(p[i].radius, rho);
        }
        p[i].cd = s->cd;
        bmd[i] = rd_body_md{(int32_t)(i + 1), s->body_type, 0, 0.0};
        if (!phase(kK2 * (p[0].mass + p[i].mass), a, e, inc, peri, node, mean, &r[i], &v[i])) return RD_EINVAL;
    }
    // barycentre (tools.cpp:367-441, j runs downwards)
    double M0 = 0.0;
    for (int j = (int)n - 1; j >= 0; --j) M0 += p[j].mass;
    double R[3] = {0, 0, 0}, V[3] = {0, 0, 0};
    for (int j = (int)n - 1; j >= 0; --j) {
        const double m = p[j].mass;
        R[0] += m * r[j].x; R[1] += m * r[j].y; R[2] += m * r[j].z;
        V[0] += m * v[j].x; V[1] += m * v[j].y; V[2] += m * v[j].z;
    }
    for (int c = 0; c < 3; ++c) { R[c] /= M0; V[c] /= M0; }
    for (int j = (int)n - 1; j >= 0; --j) {
        r[j].x -= R[0]; r[j].y -= R[1]; r[j].z -= R[2];
        v[j].x -= V[0]; v[j].y -= V[1]; v[j].z -= V[2];
    }
    for (uint32_t i = 0; i < n; ++i) { v[i].x /= kK; v[i].y /= kK; v[i].z /= kK; }
    return RD_OK;
}
