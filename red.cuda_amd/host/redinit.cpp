// redinit.cpp -- red's initial-condition generator (src/red.cuda.initial):
// named presets, seeded the way the reference seeds them, written as red's
// data, data-info, orbital-element, seed and start files.
//
//   redinit -preset <name> [-seed <n>] [-odir <dir>] [-f <file name>] [-l]
//
// The reference picks its scenario at compile time and seeds C's rand() from
// the clock (main.cu:382-1463, 3223-3290); this tool takes the scenario by
// name and the seed as an option (the clock when absent), so the seed a
// reference run wrote to <file>.seed.txt regenerates its bodies.
//
// Generation follows the reference step by step:
//   - srand(seed), then each distribution of the preset takes one rand() as
//     the seed of its own MT19937, in the order the preset builds them;
//   - populate_disk (main.cu:251-314): per body, the physical properties
//     (mass, radius, density, cd from their distributions, the missing one of
//     mass / radius / density derived), then the six orbital elements (the
//     range check of generate_oe, util_init.cpp:56-66, never rejects: its
//     condition x_min > x && x_max < x cannot hold, and this keeps that);
//   - for Dvorak, the protoplanet masses scaled to the MMSN solid mass
//     (nebula.cpp) and radii recomputed (main.cu:1841-1880); the other
//     drivers' own steps (equal MMSN shares, the real Jupiter, Saturn and
//     Uranus in place of the first giants, the 0.4 AU floor on a);
//   - calculate_phase (main.cu:317-342) with tools::calc_phase and its Kepler
//     solver as written (tools.cpp:715-828, including the shortcut that
//     returns E = M for M > pi), dt = shortest period / 1000;
//   - barycentric positions and velocities, velocities and times to k-day
//     units (tools.cpp:412-453).
// rand() is glibc's (the reference's Makefile builds redinit with gcc): the
// additive feedback generator r[i] = r[i-3] + r[i-31], seeded through the
// Park-Miller LCG, the first 310 outputs discarded, each output r >> 1.
#include <sys/stat.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <random>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

namespace {

// ---- constants (red.type/red_constants.h, red_macro.h) ------------------------
namespace cst {
constexpr double PI = 3.1415926535897932384626;
constexpr double K2 = 0.0002959122082855911025;
constexpr double Gauss = 1.720209895e-2;
constexpr double DegreeToRadian = 1.745329251994329509e-2;
constexpr double SolarToKilogram = 1.98911e30;
constexpr double SolarToGram = 1.0e3 * SolarToKilogram;
constexpr double GramToSolar = 1.0 / SolarToGram;
constexpr double MoonToSolar = 1.0 / 2.7071155e7;
constexpr double EarthToSolar = 1.0 / 3.3294605e5;
constexpr double SaturnToSolar = 1.0 / 3.497898e3;
constexpr double MercuryToSolar = 1.0 / 6.023600e6;
constexpr double VenusToSolar = 1.0 / 4.0852371e5;
constexpr double MarsToSolar = 1.0 / 3.098708e6;
constexpr double JupiterToSolar = 1.0 / 1.0473486e3;
constexpr double UranusToSolar = 1.0 / 2.290298e4;
constexpr double NeptuneToSolar = 1.0 / 1.941224e4;
constexpr double CeresToSolar = 1.0 / 2.109342e9;
constexpr double MoonToEarth = 1.0 / 8.130059e1;
constexpr double AuToMeter = 1.495978707e11;
constexpr double AuToKilometer = AuToMeter / 1.0e3;
constexpr double MeterToAu = 1.0 / AuToMeter;
constexpr double KilometerToAu = 1.0 / AuToKilometer;
constexpr double SolarRadiusToAu = 1.0 / 215.094;
constexpr double GramPerCm2ToSolarPerAu2 = GramToSolar / ((1.0e-2 * MeterToAu) * (1.0e-2 * MeterToAu));
constexpr double GramPerCm3ToSolarPerAu3 =
    GramToSolar / ((1.0e-2 * MeterToAu) * (1.0e-2 * MeterToAu) * (1.0e-2 * MeterToAu));
}  // namespace cst

enum BodyType { STAR = 0, GIANT = 1, ROCKY = 2, PROTO = 3, SUPERPL = 4, PL = 5, TESTP = 6, N_TYPE = 7 };
enum Slot { SMA, ECC, INC, PERI, NODE, MEAN, MASS, RADIUS, DENSITY, CD, N_SLOT };

// ---- C rand() as glibc implements it ----------------------------------------------
class CRand {
  public:
    void seed(uint32_t s)
    {
        int32_t w = (int32_t)(s == 0 ? 1u : s);
        uint32_t r[34];
        r[0] = (uint32_t)w;
        for (int i = 1; i < 31; ++i) {
            // 16807 w mod (2^31 - 1) by Schrage's method, on the signed seed
            const int32_t hi = w / 127773, lo = w % 127773;
            w = 16807 * lo - 2836 * hi;
            if (w < 0) w += 2147483647;
            r[i] = (uint32_t)w;
        }
        for (int i = 31; i < 34; ++i) r[i] = r[i - 31];
        for (int i = 0; i < 34; ++i) ring_[i] = r[i];
        pos_ = 34;
        for (int i = 34; i < 344; ++i) next();
    }
    int32_t rand() { return (int32_t)(next() >> 1); }

  private:
    uint32_t next()
    {
        const uint32_t x = ring_[(pos_ - 31) % kRing] + ring_[(pos_ - 3) % kRing];
        ring_[pos_ % kRing] = x;
        ++pos_;
        return x;
    }
    static constexpr uint32_t kRing = 64;   // > 31: holds every lag the recurrence reads
    uint32_t ring_[kRing] = {0};
    uint32_t pos_ = 0;
};

// ---- distributions (distribution.cpp:11-204) ---------------------------------------
// red_random is MT19937 (its twist done one word at a time yields the same
// stream as std::mt19937).
struct Dist {
    enum Kind { UNIFORM, RAYLEIGH, POWER_LAW, LOGNORMAL, EXPONENTIAL, NORMAL } kind = UNIFORM;
    double a = 0.0, b = 0.0, c = 0.0, d = 0.0;
    std::mt19937 mt;

    double u() { return (mt() + 1.0) / 4294967296.0; }
    double u(double lo, double hi) { return lo + (hi - lo) * (mt() + 1.0) / 4294967296.0; }
    double lognormal_pdf(double x) const
    {
        const double mu = c, sigma = d;
        return 1.0 / (std::sqrt(2.0 * cst::PI) * sigma * x) *
               std::exp(-((std::log(x) - mu) * (std::log(x) - mu)) / (2.0 * (sigma * sigma)));
    }
    double next()
    {
        switch (kind) {
        case UNIFORM: return u(a, b);
        case RAYLEIGH: return a * std::sqrt(-2.0 * std::log(u()));
        case EXPONENTIAL: return -1.0 / a * std::log(u());
        case NORMAL: {
            const double v = u();
            return a + b * std::sqrt(-2.0 * std::log(v)) * std::cos(1.4629180792671596E-9 * (mt() + 1.0));
        }
        case POWER_LAW: {
            // a = x_min, b = x_max, c = power; rejection under x^p with the
            // ordinate drawn on [min y, max y] (distribution.cpp:139-162)
            double y_min = c == 0.0 ? 0.0 : std::pow(a, c);
            double y_max = c == 0.0 ? 1.0 : std::pow(b, c);
            if (y_min > y_max) std::swap(y_min, y_max);
            const double d_y = y_max - y_min, d_x = b - a, area_max = d_x * d_y;
            double x, y;
            do {
                x = u(0.0, area_max) / d_y + a;
                y = u(y_min, y_max);
            } while (y > std::pow(x, c));
            return x;
        }
        case LOGNORMAL: {
            // a = x_min, b = x_max, c = mu, d = sigma (distribution.cpp:165-204)
            double y_min = 0.0;
            const double max_loc = std::exp(c - d * d);
            double y_max = lognormal_pdf(max_loc);
            if (y_min > y_max) std::swap(y_min, y_max);
            const double d_y = y_max - y_min, d_x = b - a, area_max = d_x * d_y;
            double x, y;
            do {
                x = u(0.0, area_max) / d_y + a;
                y = u(y_min, y_max);
            } while (y > lognormal_pdf(x));
            return x;
        }
        }
        return 0.0;
    }
};

struct Entry {
    Slot slot;
    Dist::Kind kind;
    double a, b, c, d;
};

struct TypeSpec {
    BodyType type;
    uint32_t n;
    int name_from;               // create_name(i + name_from, type)
    std::vector<Entry> entries;  // in the order the preset constructs them
};

struct Nebula {                  // nebula.cpp: solid component of a MMSN
    double r_1, r_2, r_SL, f_neb, Sigma_1, f_ice, p;
    double solid_mass() const
    {
        const double p2 = p + 2.0;
        const double C = p2 != 0.0 ? 2.0 * cst::PI * f_neb * Sigma_1 / p2 : 2.0 * cst::PI * f_neb * Sigma_1;
        double I;
        if (r_SL <= r_1)
            I = p2 != 0.0 ? f_ice * (std::pow(r_2, p2) - std::pow(r_1, p2)) : f_ice * (std::log(r_2) - std::log(r_1));
        else if (r_SL <= r_2)
            I = p2 != 0.0 ? std::pow(r_SL, p2) - std::pow(r_1, p2) + f_ice * (std::pow(r_2, p2) - std::pow(r_SL, p2))
                          : (std::log(r_SL) - std::log(r_1)) + f_ice * ((std::log(r_2) - std::log(r_SL)));
        else
            I = p2 != 0.0 ? std::pow(r_2, p2) - std::pow(r_1, p2) : (std::log(r_2) - std::log(r_1));
        return C * I;
    }
};

struct Vec { double x, y, z; };
struct Oe { double sma, ecc, inc, peri, node, mean; };

// Osculating elements of the planets on 2015-May-11 (JPL HORIZONS; the values
// of main.cu:41-83), as extract_from_horizon_output reads them from their
// fixed columns (util_init.cpp:327-366): e, i, Omega, omega, M in degrees, a
// in AU.  The degrees become radians as there.
struct Horizons {
    double ecc, inc, node, peri, mean, sma;
};
constexpr Horizons kMercury{2.056283170656704E-01, 7.004033081870772E+00, 4.831135275416084E+01,
                            2.917018579009503E+01, 8.665226794919347E+01, 3.870990540148295E-01};
constexpr Horizons kVenus{6.761247965610893E-03, 3.394467689974794E+00, 7.663950855912078E+01,
                          5.463888630600258E+01, 3.625130289912041E+01, 7.233287920706470E-01};
constexpr Horizons kEarth{1.706151645376220E-02, 2.832604859699792E-03, 1.991276585462631E+02,
                          2.644902808333621E+02, 1.245850770307868E+02, 9.999423845001243E-01};
constexpr Horizons kMars{9.345598920376896E-02, 1.848403968432629E+00, 4.951276588949865E+01,
                         2.865318030175783E+02, 7.844645806929105E+01, 1.523589291796774E+00};
constexpr Horizons kJupiter{4.895926045495331E-02, 1.303927595561599E+00, 1.005220570364609E+02,
                            2.736235124666692E+02, 1.262463884135522E+02, 5.201472759810757E+00};
constexpr Horizons kSaturn{5.370451604891303E-02, 2.486273622454862E+00, 1.136214918787423E+02,
                           3.402344971420399E+02, 1.437153605146956E+02, 9.553843040430950E+00};
constexpr Horizons kUranus{4.928987553618186E-02, 7.717997843738144E-01, 7.400078543435393E+01,
                           9.677390970582711E+01, 2.083879289018873E+02, 1.915532588905842E+01};
constexpr Horizons kNeptune{8.258758489207289E-03, 1.765657591232885E+00, 1.317164628477979E+02,
                            2.932421741299731E+02, 2.734992366503029E+02, 2.997852567031729E+01};

Oe oe_of(const Horizons& h)
{
    return Oe{h.sma, h.ecc, h.inc * cst::DegreeToRadian, h.peri * cst::DegreeToRadian,
              h.node * cst::DegreeToRadian, h.mean * cst::DegreeToRadian};
}

struct Disk;

enum MmsnMass { MMSN_NONE = 0, MMSN_SCALE = 1, MMSN_SHARE = 2 };

// A giant the driver replaces with a real planet before calculate_phase:
// name, mass, radius and the HORIZONS elements of 2015-May-11 (the density
// follows from mass and radius).
struct GiantSub {
    const char* name;
    double mass, radius_km;
    const Horizons* oe;
};

struct Preset {
    Preset(std::string n, std::vector<TypeSpec> t = {}, bool rc = false, int mm = MMSN_NONE, Nebula nb = Nebula{})
        : name(std::move(n)), types(std::move(t)), recompute(rc), mmsn(mm), neb(nb)
    {
    }
    std::string name;
    std::vector<TypeSpec> types;
    bool recompute = false;      // the driver's "physical quantities with the new mass" pass
    int mmsn = MMSN_NONE;        // protoplanet masses: scaled to the solid mass (Dvorak), or its equal shares
    Nebula neb{};
    std::vector<std::string> names;   // the reference's name list when it is not the types' own
    void (*post)(Disk&) = nullptr;    // the driver's own changes before calculate_phase
    bool table = false;          // bodies from a table (no distributions): the solar system
    std::vector<GiantSub> giants;     // the first giants, in order, become these planets
    double sma_floor = 0.0;      // a < sma_floor is raised to it (bodies 1..n-1), after the giants
};

// The drivers' own changes between populate_disk and calculate_phase.
void colliding_post(Disk& D);
std::string create_name(int i, int type);

// The presets (set_parameters:: in main.cu): body counts and distributions.
std::vector<Preset> presets()
{
    using K = Dist::Kind;
    const double PI = cst::PI, R_sun = cst::SolarRadiusToAu;
    const Entry star_mass{MASS, K::UNIFORM, 1.0, 1.0, 0, 0}, star_radius{RADIUS, K::UNIFORM, R_sun, R_sun, 0, 0},
        star_cd{CD, K::UNIFORM, 0.0, 0.0, 0, 0};
    const TypeSpec star{STAR, 1, 0, {star_mass, star_radius, star_cd}};
    auto generic_oe = [&](std::vector<Entry> pp) {   // n_gp, n_spl, n_pl, n_tp (main.cu:897-1169)
        std::vector<Entry> e = {{SMA, K::UNIFORM, 1.0, 10.0, 0, 0},
                                {ECC, K::UNIFORM, 0.0, 0.1, 0, 0},
                                {INC, K::UNIFORM, 0.0, 10.0 * cst::DegreeToRadian, 0, 0},
                                {PERI, K::UNIFORM, 0.0, 2.0 * PI, 0, 0},
                                {NODE, K::UNIFORM, 0.0, 2.0 * PI, 0, 0},
                                {MEAN, K::UNIFORM, 0.0, 2.0 * PI, 0, 0}};
        e.insert(e.end(), pp.begin(), pp.end());
        return e;
    };
    const double rho = cst::GramPerCm3ToSolarPerAu3;
    std::vector<Preset> v;
    // Two_body (main.cu:846-895)
    v.push_back({"two_body",
                 {star,
                  {ROCKY, 1, 0,
                   {{SMA, K::UNIFORM, 1.0, 1.0, 0, 0}, {ECC, K::UNIFORM, 0.0, 0.0, 0, 0},
                    {INC, K::UNIFORM, 0.0, 0.0, 0, 0}, {PERI, K::UNIFORM, 0.0, 0.0, 0, 0},
                    {NODE, K::UNIFORM, 0.0, 0.0, 0, 0}, {MEAN, K::UNIFORM, 0.0, 0.0, 0, 0},
                    {MASS, K::UNIFORM, cst::EarthToSolar, cst::EarthToSolar, 0, 0},
                    {DENSITY, K::UNIFORM, 2.7 * rho, 2.7 * rho, 0, 0},
                    {CD, K::UNIFORM, 1.0, 1.0, 0, 0}}}},
                 false});
    // n_gp (main.cu:897-948)
    v.push_back({"n_gp",
                 {star,
                  {GIANT, 10, 1,
                   generic_oe({{MASS, K::UNIFORM, 0.1 * cst::SaturnToSolar, 1.0 * cst::SaturnToSolar, 0, 0},
                               {DENSITY, K::UNIFORM, 1.0 * rho, 2.0 * rho, 0, 0},
                               {CD, K::UNIFORM, 0.0, 0.0, 0, 0}})}},
                 true});
    // n_pp (main.cu:950-1001)
    v.push_back({"n_pp",
                 {star,
                  {PROTO, 500, 1,
                   {{SMA, K::UNIFORM, 0.6, 2.6, 0, 0}, {ECC, K::UNIFORM, 0.0, 0.1, 0, 0},
                    {INC, K::UNIFORM, 0.0, 0.0 * cst::DegreeToRadian, 0, 0}, {PERI, K::UNIFORM, 0.0, 2.0 * PI, 0, 0},
                    {NODE, K::UNIFORM, 0.0, 0.0 * PI, 0, 0}, {MEAN, K::UNIFORM, 0.0, 2.0 * PI, 0, 0},
                    {MASS, K::UNIFORM, 0.5 * cst::MoonToSolar, 2.0 * cst::MoonToSolar, 0, 0},
                    {DENSITY, K::UNIFORM, 2.7 * rho, 2.7 * rho, 0, 0},
                    {CD, K::UNIFORM, 0.0, 0.0, 0, 0}}}},
                 true});
    // n_spl (main.cu:1003-1057)
    v.push_back({"n_spl",
                 {star,
                  {SUPERPL, 10, 1,
                   generic_oe({{MASS, K::UNIFORM, 1.0e-10, 1.0e-9, 0, 0},
                               {RADIUS, K::UNIFORM, 10.0 * cst::KilometerToAu, 30.0 * cst::KilometerToAu, 0, 0},
                               {DENSITY, K::UNIFORM, 1.0 * rho, 3.0 * rho, 0, 0},
                               {CD, K::UNIFORM, 1.0, 2.7, 0, 0}})}},
                 false});
    // n_pl (main.cu:1059-1112)
    v.push_back({"n_pl",
                 {star,
                  {PL, 10, 1,
                   generic_oe({{MASS, K::UNIFORM, 1.0e-3 * cst::CeresToSolar, 1.0e-2 * cst::CeresToSolar, 0, 0},
                               {DENSITY, K::UNIFORM, 1.0 * rho, 3.0 * rho, 0, 0},
                               {CD, K::UNIFORM, 0.0, 0.0, 0, 0}})}},
                 true});
    // n_tp (main.cu:1114-1168)
    v.push_back({"n_tp",
                 {star,
                  {TESTP, 10, 1,
                   generic_oe({{MASS, K::UNIFORM, 0.0, 0.0, 0, 0}, {RADIUS, K::UNIFORM, 0.0, 0.0, 0, 0},
                               {DENSITY, K::UNIFORM, 0.0, 0.0, 0, 0}, {CD, K::UNIFORM, 0.0, 0.0, 0, 0}})}},
                 false});
    // set_parameters::Dvorak (main.cu:571-625): 300 protoplanets on the MMSN
    // solid disk `neb` (the power law of a takes its r_1, r_2 and p)
    const double sigma = cst::GramPerCm2ToSolarPerAu2;
    const GiantSub jupiter{"Jupiter", cst::JupiterToSolar, 71492.0, &kJupiter},
        saturn{"Saturn", cst::SaturnToSolar, 60268.0, &kSaturn},
        uranus{"Uranus", cst::UranusToSolar, 25362.0, &kUranus};
    auto dvorak = [&](const std::string& name, const Nebula& neb) {
        Preset p{name, {}, true, MMSN_SCALE, neb};
        p.types = {star,
                   {PROTO, 300, 1,
                    {{SMA, K::POWER_LAW, neb.r_1, neb.r_2, neb.p, 0},
                     {ECC, K::RAYLEIGH, 0.01, 0, 0, 0},
                     {INC, K::UNIFORM, 0.0, 0.0, 0, 0},
                     {PERI, K::UNIFORM, 0.0, 2.0 * PI, 0, 0},
                     {NODE, K::UNIFORM, 0.0, 0.0, 0, 0},
                     {MEAN, K::UNIFORM, 0.0, 2.0 * PI, 0, 0},
                     {MASS, K::LOGNORMAL, 0.1, 3.0, 0.0, 0.25},
                     {DENSITY, K::UNIFORM, 2.7 * rho, 2.7 * rho, 0, 0},
                     {CD, K::UNIFORM, 0.0, 0.0, 0, 0}}}};
        return p;
    };
    // Dvorak (driver main.cu:1804-1917): 0.9-2.7 AU, Sigma_1 = 7 g/cm^2, f_neb = 3, p = -3/2
    v.push_back(dvorak("Dvorak", Nebula{0.9, 2.7, 2.7, 3.0, 7.0 * sigma, 4.2, -3.0 / 2.0}));
    // Chambers2001 (main.cu:1467-1688).  The disk it prints is the last one it
    // builds: Dvorak's parameters on a 0.9-2.7 AU solid disk with Sigma_1 =
    // 8 g/cm^2 and f_neb = 1.  Its earlier set_parameters::Chambers2001 call
    // (main.cu:382-436) seeds rand() with the same clock second, so Dvorak's
    // setter draws the same rand() values; but it pushed its own 154 names
    // first, so the printed names are those 154 followed by Dvorak's.
    {
        Preset p = dvorak("Chambers2001", Nebula{0.9, 2.7, 2.7, 1.0, 8.0 * sigma, 4.2, -3.0 / 2.0});
        for (int round = 0; round < 2; ++round) {
            p.names.push_back("star");
            for (int i = 1; i <= (round == 0 ? 153 : 300); ++i) p.names.push_back(create_name(i, PROTO));
        }
        v.push_back(p);
    }
    // coll_stat_run (set_parameters main.cu:505-569, driver 1690-1802): 1000
    // Ceres-mass protoplanets on 2.5-3 AU, the MMSN solid mass (Sigma_1 =
    // 7 g/cm^2, f_neb = 1) shared equally among them
    {
        Preset p{"coll_stat_run", {}, true, MMSN_SHARE, Nebula{2.5, 3.0, 2.7, 1.0, 7.0 * sigma, 4.2, -3.0 / 2.0}};
        p.types = {star,
                   {PROTO, 1000, 1,
                    {{SMA, K::POWER_LAW, 2.5, 3.0, -3.0 / 2.0, 0},
                     {ECC, K::RAYLEIGH, 0.02, 0, 0, 0},
                     {INC, K::RAYLEIGH, 0.04, 0, 0, 0},
                     {PERI, K::UNIFORM, 0.0, 2.0 * PI, 0, 0},
                     {NODE, K::UNIFORM, 0.0, 2.0 * PI, 0, 0},
                     {MEAN, K::UNIFORM, 0.0, 2.0 * PI, 0, 0},
                     {MASS, K::UNIFORM, cst::CeresToSolar, cst::CeresToSolar, 0, 0},
                     {DENSITY, K::UNIFORM, 2.5 * rho, 2.5 * rho, 0, 0},
                     {CD, K::UNIFORM, 0.0, 0.0, 0, 0}}}};
        v.push_back(p);
    }
    // Hansen_2009 (set_parameters main.cu:626-700, driver 1917-2003): a
    // Jupiter at 5.2 AU (replaced by the real Jupiter) and 400 protoplanets on
    // 0.7-1 AU
    {
        Preset p{"Hansen_2009",
                 {star,
                  {GIANT, 1, 1,
                   {{SMA, K::UNIFORM, 5.2, 5.2, 0, 0}, {ECC, K::UNIFORM, 0.05, 0.05, 0, 0},
                    {INC, K::UNIFORM, 0.0, 0.0, 0, 0}, {PERI, K::UNIFORM, 0.0, 2.0 * PI, 0, 0},
                    {NODE, K::UNIFORM, 0.0, 0.0, 0, 0}, {MEAN, K::UNIFORM, 0.0, 2.0 * PI, 0, 0},
                    {MASS, K::UNIFORM, cst::JupiterToSolar, cst::JupiterToSolar, 0, 0},
                    {DENSITY, K::UNIFORM, 1.326 * rho, 1.326 * rho, 0, 0},
                    {CD, K::UNIFORM, 0.0, 0.0, 0, 0}}},
                  {PROTO, 400, 1,
                   {{SMA, K::UNIFORM, 0.7, 1.0, 0, 0}, {ECC, K::UNIFORM, 0.0, 0.0, 0, 0},
                    {INC, K::UNIFORM, 0.0, 0.5 * cst::DegreeToRadian, 0, 0}, {PERI, K::UNIFORM, 0.0, 0.0, 0, 0},
                    {NODE, K::UNIFORM, 0.0, 2.0 * PI, 0, 0}, {MEAN, K::UNIFORM, 0.0, 2.0 * PI, 0, 0},
                    {MASS, K::UNIFORM, 5.0e-3 * cst::EarthToSolar, 5.0e-3 * cst::EarthToSolar, 0, 0},
                    {DENSITY, K::UNIFORM, 2.7 * rho, 2.7 * rho, 0, 0},
                    {CD, K::UNIFORM, 0.0, 0.0, 0, 0}}}}};
        p.giants = {jupiter};
        v.push_back(p);
    }
    // n_pl_to_test_anal_gd (set_parameters main.cu:456-503, driver
    // 2863-2911): 10 planetesimals of 1 m radius on 1-10 AU circles, named
    // from pl0
    v.push_back({"n_pl_to_test_anal_gd",
                 {star,
                  {PL, 10, 0,
                   {{SMA, K::UNIFORM, 1.0, 10.0, 0, 0}, {ECC, K::UNIFORM, 0.0, 0.0, 0, 0},
                    {INC, K::UNIFORM, 0.0, 0.0, 0, 0}, {PERI, K::UNIFORM, 0.0, 0.0, 0, 0},
                    {NODE, K::UNIFORM, 0.0, 0.0, 0, 0}, {MEAN, K::UNIFORM, 0.0, 0.0, 0, 0},
                    {RADIUS, K::UNIFORM, cst::MeterToAu, cst::MeterToAu, 0, 0},
                    {DENSITY, K::UNIFORM, 2.7 * rho, 2.7 * rho, 0, 0},
                    {CD, K::UNIFORM, 1.0, 1.0, 0, 0}}}}});
    // solar_system (set_parameters main.cu:438-454, driver 2913-2961,
    // populate_solar_system 85-249): the Sun and the eight planets from a table
    {
        Preset p{"solar_system", {}};
        p.table = true;
        v.push_back(p);
    }
    // Colliding_bodies (set_parameters main.cu:1412-1461, driver 3055-3153):
    // two Ceres-mass protoplanets, a target and a projectile 0.01 deg behind
    // it on the same orbit
    {
        Preset p{"Colliding_bodies",
                 {star,
                  {PROTO, 2, 0,
                   {{SMA, K::UNIFORM, 1.0, 2.0, 0, 0}, {ECC, K::UNIFORM, 0.0, 0.0, 0, 0},
                    {INC, K::UNIFORM, 0.0, 0.0, 0, 0}, {PERI, K::UNIFORM, 0.0, 0.0, 0, 0},
                    {NODE, K::UNIFORM, 0.0, 0.0, 0, 0}, {MEAN, K::UNIFORM, 0.0, 0.0, 0, 0},
                    {MASS, K::UNIFORM, cst::CeresToSolar, cst::CeresToSolar, 0, 0},
                    {DENSITY, K::UNIFORM, 2.7 * rho, 2.7 * rho, 0, 0},
                    {CD, K::UNIFORM, 0.0, 0.0, 0, 0}}}}};
        p.post = colliding_post;
        v.push_back(p);
    }
    const double deg = cst::DegreeToRadian;
    // Birgit_2016 (set_parameters main.cu:702-775, driver 2005-2099): Jupiter and
    // Saturn, 1000 tenth-Moon protoplanets on 0.3-2.4 AU
    {
        Preset p{"Birgit_2016",
                 {star,
                  {GIANT, 2, 1,
                   {{SMA, K::UNIFORM, 5.2, 5.2, 0, 0}, {ECC, K::UNIFORM, 0.05, 0.05, 0, 0},
                    {INC, K::UNIFORM, 0.0, 0.0, 0, 0}, {PERI, K::UNIFORM, 0.0, 2.0 * PI, 0, 0},
                    {NODE, K::UNIFORM, 0.0, 0.0, 0, 0}, {MEAN, K::UNIFORM, 0.0, 2.0 * PI, 0, 0},
                    {MASS, K::UNIFORM, cst::JupiterToSolar, cst::JupiterToSolar, 0, 0},
                    {DENSITY, K::UNIFORM, 1.326 * rho, 1.326 * rho, 0, 0},
                    {CD, K::UNIFORM, 0.0, 0.0, 0, 0}}},
                  {PROTO, 1000, 1,
                   {{SMA, K::UNIFORM, 0.3, 2.4, 0, 0}, {ECC, K::UNIFORM, 0.02, 0.02, 0, 0},
                    {INC, K::UNIFORM, 1.0 * deg, 1.0 * deg, 0, 0}, {PERI, K::UNIFORM, 0.0, 2.0 * PI, 0, 0},
                    {NODE, K::UNIFORM, 0.0, 2.0 * PI, 0, 0}, {MEAN, K::UNIFORM, 0.0, 2.0 * PI, 0, 0},
                    {MASS, K::UNIFORM, 0.1 * cst::MoonToSolar, 0.1 * cst::MoonToSolar, 0, 0},
                    {DENSITY, K::UNIFORM, 2.0 * rho, 2.0 * rho, 0, 0},
                    {CD, K::UNIFORM, 0.0, 0.0, 0, 0}}}}};
        p.giants = {jupiter, saturn};
        v.push_back(p);
    }
    // Birgit_scenario (set_parameters main.cu:777-844, driver 2101-2202): Jupiter,
    // a Saturn of radius 58232 km, and 100 comets (test particles, e 0.99-0.9999)
    {
        Preset p{"Birgit_scenario",
                 {star,
                  {GIANT, 2, 1,
                   {{SMA, K::UNIFORM, 1.0, 2.0, 0, 0}, {ECC, K::UNIFORM, 0.0, 0.05, 0, 0},
                    {INC, K::UNIFORM, 0.0, 1.0 * deg, 0, 0}, {PERI, K::UNIFORM, 0.0, 2.0 * PI, 0, 0},
                    {NODE, K::UNIFORM, 0.0, 2.0 * PI, 0, 0}, {MEAN, K::UNIFORM, 0.0, 2.0 * PI, 0, 0},
                    {MASS, K::UNIFORM, cst::JupiterToSolar, cst::JupiterToSolar, 0, 0},
                    {DENSITY, K::UNIFORM, 2.7 * rho, 2.7 * rho, 0, 0},
                    {CD, K::UNIFORM, 0.0, 0.0, 0, 0}}},
                  {TESTP, 100, 1,
                   {{SMA, K::UNIFORM, 50.0, 80.0, 0, 0}, {ECC, K::UNIFORM, 0.99, 0.9999, 0, 0},
                    {INC, K::UNIFORM, 0.0, 0.5, 0, 0}, {PERI, K::UNIFORM, 0.0, 2.0 * PI, 0, 0},
                    {NODE, K::UNIFORM, 0.0, 2.0 * PI, 0, 0}, {MEAN, K::UNIFORM, 0.0, 1.0 * PI, 0, 0}}}}};
        p.giants = {jupiter, GiantSub{"Saturn", cst::SaturnToSolar, 58232.0, &kSaturn}};
        v.push_back(p);
    }
    // GT_scenario (set_parameters main.cu:1170-1245, driver 2204-2345) and
    // GT_scenario_mod (1247-1322, 2347-2495): giants on 1-2 AU replaced by the
    // real ones, 2000 half-Moon protoplanets, a < 0.4 AU raised to 0.4.  The
    // giants' mass and every density are set up as uniform(x, 0), which
    // distribution.cpp's assert(x_min <= x_max) rejects in a build without
    // NDEBUG (the Makefile's); with it (the Visual Studio Release builds) the
    // draw is x + (0 - x) u, as here.
    auto gt = [&](const std::string& name, uint32_t n_giant, Entry sma) {
        const double m_pp = 1.0 / 2.0 * cst::MoonToEarth * cst::EarthToSolar;
        Preset p{name,
                 {star,
                  {GIANT, n_giant, 1,
                   {{SMA, K::UNIFORM, 1.0, 2.00, 0, 0}, {ECC, K::UNIFORM, 0.0, 0.05, 0, 0},
                    {INC, K::UNIFORM, 0.0, 1.0 * deg, 0, 0}, {PERI, K::UNIFORM, 0.0, 2.0 * PI, 0, 0},
                    {NODE, K::UNIFORM, 0.0, 2.0 * PI, 0, 0}, {MEAN, K::UNIFORM, 0.0, 2.0 * PI, 0, 0},
                    {MASS, K::UNIFORM, 1.0 * cst::JupiterToSolar, 0.0, 0, 0},
                    {DENSITY, K::UNIFORM, 2.7 * rho, 0.0, 0, 0},
                    {CD, K::UNIFORM, 0.0, 0.0, 0, 0}}},
                  {PROTO, 2000, 1,
                   {sma, {ECC, K::UNIFORM, 0.0, 0.05, 0, 0},
                    {INC, K::UNIFORM, 0.0, 1.0 * deg, 0, 0}, {PERI, K::UNIFORM, 0.0, 2.0 * PI, 0, 0},
                    {NODE, K::UNIFORM, 0.0, 2.0 * PI, 0, 0}, {MEAN, K::UNIFORM, 0.0, 2.0 * PI, 0, 0},
                    {MASS, K::UNIFORM, m_pp, m_pp, 0, 0},
                    {DENSITY, K::UNIFORM, 2.7 * rho, 0.0, 0, 0},
                    {CD, K::UNIFORM, 0.0, 0.0, 0, 0}}}},
                 true};
        p.sma_floor = 0.4;
        return p;
    };
    {
        Preset p = gt("GT_scenario", 2, Entry{SMA, K::NORMAL, 1.45, 0.35, 0, 0});
        p.giants = {jupiter, saturn};
        v.push_back(p);
    }
    {
        Preset p = gt("GT_scenario_mod", 3, Entry{SMA, K::EXPONENTIAL, 0.5, 0, 0, 0});
        p.giants = {jupiter, saturn, uranus};
        p.names.push_back("Sun");
        v.push_back(p);
    }
    // Elke_binary_test (set_parameters main.cu:1324-1410, driver 2963-3053): a
    // second solar-mass star on a 100 AU orbit (e = 0.1), Jupiter, and 10000
    // Ceres-mass protoplanets on 0.5-4 AU.  The eccentricity's range [0, 0.2]
    // redraws nothing: util_init.cpp:63 repeats while x_min > e && x_max < e,
    // which no e satisfies.
    {
        Preset p{"Elke_binary_test",
                 {{STAR, 2, 0,
                   {{SMA, K::UNIFORM, 100.0, 100.0, 0, 0}, {ECC, K::UNIFORM, 0.1, 0.1, 0, 0},
                    {INC, K::UNIFORM, 0.0, 0.0, 0, 0}, {PERI, K::UNIFORM, 0.0, 0.0 * PI, 0, 0},
                    {NODE, K::UNIFORM, 0.0, 0.0, 0, 0}, {MEAN, K::UNIFORM, 0.0, 0.0 * PI, 0, 0},
                    star_mass, star_radius, star_cd}},
                  {GIANT, 1, 1,
                   {{SMA, K::UNIFORM, 5.2, 5.2, 0, 0}, {ECC, K::UNIFORM, 0.0, 0.0, 0, 0},
                    {INC, K::UNIFORM, 0.0, 0.0, 0, 0}, {PERI, K::UNIFORM, 0.0, 2.0 * PI, 0, 0},
                    {NODE, K::UNIFORM, 0.0, 0.0, 0, 0}, {MEAN, K::UNIFORM, 0.0, 2.0 * PI, 0, 0},
                    {MASS, K::UNIFORM, 1.0 * cst::JupiterToSolar, 1.0 * cst::JupiterToSolar, 0, 0},
                    {DENSITY, K::UNIFORM, 1.326 * rho, 1.326 * rho, 0, 0},
                    {CD, K::UNIFORM, 0.0, 0.0, 0, 0}}},
                  {PROTO, 10000, 1,
                   {{SMA, K::UNIFORM, 0.5, 4.0, 0, 0}, {ECC, K::RAYLEIGH, 0.02, 0, 0, 0},
                    {INC, K::UNIFORM, 0.0, 0.0, 0, 0}, {PERI, K::UNIFORM, 0.0, 2.0 * PI, 0, 0},
                    {NODE, K::UNIFORM, 0.0, 0.0, 0, 0}, {MEAN, K::UNIFORM, 0.0, 2.0 * PI, 0, 0},
                    {MASS, K::UNIFORM, 1.0 * cst::CeresToSolar, 1.0 * cst::CeresToSolar, 0, 0},
                    {DENSITY, K::UNIFORM, 2.0 * rho, 2.0 * rho, 0, 0},
                    {CD, K::UNIFORM, 0.0, 0.0, 0, 0}}}}};
        p.names = {"primary_star", "secondary_star", "Jupiter"};
        p.giants = {jupiter};
        p.sma_floor = 0.4;
        v.push_back(p);
    }
    return v;
}

// ---- tools.cpp ---------------------------------------------------------------------
constexpr double kFourPiOverThree = 4.188790204786391;
double calc_radius(double m, double density) { return std::pow(1.0 / kFourPiOverThree * m / density, 1.0 / 3.0); }
double calc_density(double m, double R) { return R == 0.0 ? 0.0 : m / (kFourPiOverThree * (R * R * R)); }
double calc_mass(double R, double density) { return kFourPiOverThree * (R * R * R) * density; }
double orbital_period(double mu, double a) { return 6.283185307179586 * std::sqrt((a * a * a) / mu); }

// kepler_equation_solver as written (tools.cpp:715-750)
double kepler(double ecc, double mean, double eps)
{
    if (0.0 == ecc || 0.0 == mean || cst::PI == mean) return mean;
    bool s = false;
    if (cst::PI < mean) {
        mean -= cst::PI;
        s = true;
    }
    double E = mean + ecc * (std::sin(mean)) / (1.0 - std::sin(mean + ecc) + std::sin(mean));
    double error;
    uint16_t step = 0;
    do {
        const double E1 = E - (E - ecc * std::sin(E) - mean) / (1.0 - ecc * std::cos(E));
        error = std::fabs(E1 - E);
        E = E1;
    } while (error > eps && step++ <= 32);
    if (32 < step) throw std::string("The kepler_equation_solver() failed: solution did not converge.");
    if (s) E += (mean - E) + cst::PI;
    return E;
}


// calc_phase (tools.cpp:789-828)
void calc_phase(double mu, const Oe& oe, Vec& r, Vec& v)
{
    const double ecc = oe.ecc;
    const double E = kepler(ecc, oe.mean, 1.0e-8);
    const double nu = 2.0 * std::atan(std::sqrt((1.0 + ecc) / (1.0 - ecc)) * std::tan(E / 2.0));
    const double p = oe.sma * (1.0 - ecc * ecc);
    const double rr = p / (1.0 + ecc * std::cos(nu));
    const double kszi = rr * std::cos(nu), eta = rr * std::sin(nu);
    const double vKszi = -std::sqrt(mu / p) * std::sin(nu), vEta = std::sqrt(mu / p) * (ecc + std::cos(nu));
    const double cw = std::cos(oe.peri), sw = std::sin(oe.peri), cO = std::cos(oe.node), sO = std::sin(oe.node),
                 ci = std::cos(oe.inc), si = std::sin(oe.inc);
    const Vec P{cw * cO - sw * sO * ci, cw * sO + sw * cO * ci, sw * si};
    const Vec Q{-sw * cO - cw * sO * ci, -sw * sO + cw * cO * ci, cw * si};
    r = Vec{kszi * P.x + eta * Q.x, kszi * P.y + eta * Q.y, kszi * P.z + eta * Q.z};
    v = Vec{vKszi * P.x + vEta * Q.x, vKszi * P.y + vEta * Q.y, vKszi * P.z + vEta * Q.z};
}

struct Body {
    std::string name;
    int32_t id, type;
    double mass, radius, density, cd;
    Oe oe;
    Vec r, v;
};

std::string create_name(int i, int type)   // util_init.cpp:19-31
{
    static const char* names[] = {"star", "giant", "rocky", "proto", "superpl", "pl", "testp"};
    return std::string(names[type]) + std::to_string(i);
}

// ---- the generation pipeline ---------------------------------------------------------
struct Disk {
    std::vector<Body> bodies;
    uint32_t n_type[N_TYPE] = {0};
    double t0 = 0.0, dt = 0.0;
};

Disk solar_system_table();
Disk finish(Disk D);

Disk generate(const Preset& P, uint32_t seed)
{
    if (P.table) return finish(solar_system_table());
    CRand crand;
    crand.seed(seed);
    // distributions, one rand() each, in construction order
    std::vector<std::vector<Dist*>> slots(P.types.size(), std::vector<Dist*>(N_SLOT, nullptr));
    std::vector<Dist> pool;
    size_t n_entries = 0;
    for (const TypeSpec& T : P.types) n_entries += T.entries.size();
    pool.reserve(n_entries);
    for (size_t k = 0; k < P.types.size(); ++k)
        for (const Entry& e : P.types[k].entries) {
            Dist d;
            d.kind = e.kind;
            d.a = e.a; d.b = e.b; d.c = e.c; d.d = e.d;
            d.mt.seed((uint32_t)crand.rand());
            pool.push_back(d);
            slots[k][e.slot] = &pool.back();
        }
    Disk D;
    int32_t id = 1;
    for (size_t k = 0; k < P.types.size(); ++k) {
        const TypeSpec& T = P.types[k];
        const auto& S = slots[k];
        D.n_type[T.type] += T.n;
        for (uint32_t i = 0; i < T.n; ++i, ++id) {
            Body b{};
            b.name = T.type == STAR && i == 0 ? std::string("star") : create_name((int)i + T.name_from, T.type);
            b.id = id;
            b.type = T.type;
            // generate_pp (util_init.cpp:120-145)
            if (S[MASS]) b.mass = S[MASS]->next();
            if (S[RADIUS]) b.radius = S[RADIUS]->next();
            if (S[DENSITY]) b.density = S[DENSITY]->next();
            if (S[CD]) b.cd = S[CD]->next();
            if (T.type == TESTP) {
                b.mass = b.radius = b.density = b.cd = 0.0;
            } else if (T.type != SUPERPL) {
                if (!S[MASS] && S[RADIUS] && S[DENSITY]) b.mass = calc_mass(b.radius, b.density);
                else if (!S[RADIUS] && S[MASS] && S[DENSITY]) b.radius = calc_radius(b.mass, b.density);
                else if (!S[DENSITY] && S[MASS] && S[RADIUS]) b.density = calc_density(b.mass, b.radius);
                else throw std::string("Missing physical parameters!");
            }
            // generate_oe (util_init.cpp:56-118); the primary star sits at the origin
            if (!(T.type == STAR && id == 1)) {
                double* oe[6] = {&b.oe.sma, &b.oe.ecc, &b.oe.inc, &b.oe.peri, &b.oe.node, &b.oe.mean};
                for (int s = SMA; s <= MEAN; ++s) *oe[s] = S[s] ? S[s]->next() : 0.0;
            }
            D.bodies.push_back(b);
        }
    }
    const uint32_t n = (uint32_t)D.bodies.size();
    if (P.mmsn != MMSN_NONE) {
        // tools::get_total_mass over the protoplanets, descending (tools.cpp:351-365)
        const double m_solid = P.neb.solid_mass();
        auto total = [&]() {
            double m = 0.0;
            for (int j = (int)n - 1; j >= 0; --j)
                if (D.bodies[j].type == PROTO && D.bodies[j].id > 0) m += D.bodies[j].mass;
            return m;
        };
        if (P.mmsn == MMSN_SCALE) {   // Dvorak, Chambers2001 (main.cu:1841-1858)
            const double f = m_solid / total();
            for (Body& b : D.bodies)
                if (b.type == PROTO) b.mass *= f;
        } else {                      // coll_stat_run (main.cu:1734-1748): equal shares
            const double m_pp = m_solid / D.n_type[PROTO];
            for (Body& b : D.bodies)
                if (b.type == PROTO) b.mass = m_pp;
        }
        if (std::fabs(total() - m_solid) > 1.0e-15) {
            std::cerr << "The required mass was not reached." << std::endl;
            std::exit(0);   // as the reference does, writing nothing
        }
    }
    if (P.recompute) {
        size_t k = 0, left = P.types[0].n;
        for (Body& b : D.bodies) {
            while (left == 0) left = P.types[++k].n;
            --left;
            const auto& S = slots[k];
            if (b.mass > 0.0) {
                if (!S[DENSITY] && S[RADIUS]) b.density = calc_density(b.mass, b.radius);
                if (!S[RADIUS] && S[DENSITY]) b.radius = calc_radius(b.mass, b.density);
            }
        }
    }
    for (size_t i = 0; i < P.names.size() && i < D.bodies.size(); ++i) D.bodies[i].name = P.names[i];
    if (P.post) P.post(D);
    // the drivers' loop before calculate_phase (e.g. main.cu:1955-1965, 2265-2292):
    // the first giants become the listed planets, then the floor on a
    size_t g = 0;
    for (size_t i = 1; i < D.bodies.size(); ++i) {
        Body& b = D.bodies[i];
        if (b.type == GIANT && g < P.giants.size()) {
            const GiantSub& G = P.giants[g++];
            b.name = G.name;
            b.mass = 1.0 * G.mass;
            b.radius = G.radius_km * cst::KilometerToAu;
            b.density = calc_density(b.mass, b.radius);
            b.oe = oe_of(*G.oe);
        }
        if (P.sma_floor > b.oe.sma) b.oe.sma = P.sma_floor;
    }
    return finish(std::move(D));
}

// calculate_phase (main.cu:317-342), the barycentre and red's units
Disk finish(Disk D)
{
    const uint32_t n = (uint32_t)D.bodies.size();
    double min_P = 1.7976931348623157e308;
    const double m0 = D.bodies[0].mass;
    D.bodies[0].r = Vec{0, 0, 0};
    D.bodies[0].v = Vec{0, 0, 0};
    for (uint32_t i = 1; i < n; ++i) {
        const double mu = cst::K2 * (m0 + D.bodies[i].mass);
        calc_phase(mu, D.bodies[i].oe, D.bodies[i].r, D.bodies[i].v);
        const double Pi = orbital_period(mu, D.bodies[i].oe.sma);
        if (min_P > Pi) min_P = Pi;
    }
    D.dt = min_P / 1000.0;
    // transform_to_bc (tools.cpp:412-441): mass and barycentre summed from the last body
    double M0 = 0.0;
    for (int j = (int)n - 1; j >= 0; --j)
        if (D.bodies[j].id >= 0) M0 += D.bodies[j].mass;
    Vec R0{0, 0, 0}, V0{0, 0, 0};
    for (int j = (int)n - 1; j >= 0; --j) {
        const Body& b = D.bodies[j];
        if (b.id < 0) continue;
        R0.x += b.mass * b.r.x; R0.y += b.mass * b.r.y; R0.z += b.mass * b.r.z;
        V0.x += b.mass * b.v.x; V0.y += b.mass * b.v.y; V0.z += b.mass * b.v.z;
    }
    R0.x /= M0; R0.y /= M0; R0.z /= M0;
    V0.x /= M0; V0.y /= M0; V0.z /= M0;
    for (int j = (int)n - 1; j >= 0; --j) {
        Body& b = D.bodies[j];
        if (b.id < 0) continue;
        b.r.x -= R0.x; b.r.y -= R0.y; b.r.z -= R0.z;
        b.v.x -= V0.x; b.v.y -= V0.y; b.v.z -= V0.z;
    }
    // transform_velocity, t0 and dt to k days (tools.cpp:443-453, main.cu:2597-2598)
    for (Body& b : D.bodies) {
        b.v.x /= cst::Gauss; b.v.y /= cst::Gauss; b.v.z /= cst::Gauss;
    }
    D.t0 *= cst::Gauss;
    D.dt *= cst::Gauss;
    return D;
}

// Colliding_bodies (main.cu:3090-3115): the first two protoplanets on one
// orbit, the second 0.01 degree behind in mean anomaly
void colliding_post(Disk& D)
{
    int k = 0;
    for (size_t i = 1; i < D.bodies.size() && k < 2; ++i) {
        Body& b = D.bodies[i];
        if (b.type != PROTO) continue;
        b.name = k == 0 ? "target" : "projectile";
        b.oe = Oe{2.7, 0.05, 0.0, 0.0, 0.0, k == 0 ? 0.0 : 0.01 * cst::DegreeToRadian};
        ++k;
    }
}

// populate_solar_system (main.cu:85-249): the Sun, the rocky planets, then the
// giants, in that order (rocky before giant, not sorted by body type);
// parameters from red's constants and radii, the remaining one of mass,
// radius and density computed, cd = 0
Disk solar_system_table()
{
    struct Row {
        const char* name;
        int type;
        double mass, radius_km;
        const Horizons* oe;
    };
    const Row rows[] = {{"Mercury", ROCKY, cst::MercuryToSolar, 2439.7, &kMercury},
                        {"Venus", ROCKY, cst::VenusToSolar, 6051.8, &kVenus},
                        {"Earth", ROCKY, cst::EarthToSolar, 6371.0, &kEarth},
                        {"Mars", ROCKY, cst::MarsToSolar, 3389.5, &kMars},
                        {"Jupiter", GIANT, cst::JupiterToSolar, 71492.0, &kJupiter},
                        {"Saturn", GIANT, cst::SaturnToSolar, 60268.0, &kSaturn},
                        {"Uranus", GIANT, cst::UranusToSolar, 25559.0, &kUranus},
                        {"Neptune", GIANT, cst::NeptuneToSolar, 24766.0, &kNeptune}};
    Disk D;
    Body sun{};
    sun.name = "Sun";
    sun.id = 1;
    sun.type = STAR;
    sun.mass = 1.0;
    sun.radius = 1.0 * cst::SolarRadiusToAu;
    sun.density = calc_density(sun.mass, sun.radius);
    D.bodies.push_back(sun);
    int32_t id = 2;
    for (const Row& r : rows) {
        Body b{};
        b.name = r.name;
        b.id = id++;
        b.type = r.type;
        b.mass = 1.0 * r.mass;
        b.radius = r.radius_km * cst::KilometerToAu;
        b.density = calc_density(b.mass, b.radius);
        b.oe = oe_of(*r.oe);
        D.bodies.push_back(b);
    }
    D.n_type[STAR] = 1;
    D.n_type[ROCKY] = 4;
    D.n_type[GIANT] = 4;
    return D;
}

// ---- output (main.cu:344-380; file_util.cpp:325-391, 536-560) ----------------------
std::string combine_path(const std::string& dir, const std::string& file)
{
    if (dir.empty()) return file;
    return dir.back() == '/' ? dir + file : dir + "/" + file;
}

std::ofstream open_out(const std::string& path)
{
    std::ofstream o(path.c_str(), std::ios::out);
    if (!o) throw std::string("Cannot open " + path + "!");
    return o;
}

void write_all(const std::string& dir, const std::string& file, uint32_t seed, const Disk& D)
{
    const char SEP = ' ';
    {
        std::ofstream o = open_out(combine_path(dir, file) + ".seed.txt");
        o << seed;
    }
    const std::string fn_info = file + ".info.txt", fn_data = file + ".data.txt";
    {
        std::ofstream o = open_out(combine_path(dir, fn_info));
        o.precision(16);
        o.setf(std::ios::right);
        o.setf(std::ios::scientific);
        o << std::setw(25) << D.t0 << SEP << std::setw(25) << D.dt << SEP << std::setw(10) << 0 << SEP;
        for (int t = 0; t < N_TYPE; ++t) o << std::setw(8) << D.n_type[t] << SEP;
    }
    {
        std::ofstream o = open_out(combine_path(dir, fn_data));
        o.precision(16);
        o.setf(std::ios::right);
        o.setf(std::ios::scientific);
        for (const Body& b : D.bodies)
            o << std::setw(30) << b.name << SEP << std::setw(8) << b.id << SEP << std::setw(2) << b.type << SEP
              << std::setw(2) << 0 << SEP << std::setw(25) << 0.0 << SEP << std::setw(25) << b.mass << SEP
              << std::setw(25) << b.radius << SEP << std::setw(25) << b.density << SEP << std::setw(25) << b.cd
              << SEP << std::setw(25) << b.r.x << SEP << std::setw(25) << b.r.y << SEP << std::setw(25) << b.r.z
              << SEP << std::setw(25) << b.v.x << SEP << std::setw(25) << b.v.y << SEP << std::setw(25) << b.v.z
              << std::endl;
    }
    {
        std::ofstream o = open_out(combine_path(dir, file) + ".oe.txt");
        o.precision(6);
        o.setf(std::ios::right);
        o.setf(std::ios::scientific);
        for (const Body& b : D.bodies)
            o << std::setw(15) << D.t0 << SEP << std::setw(7) << b.id << SEP << std::setw(2) << b.type << SEP
              << std::setw(15) << b.mass << SEP << std::setw(15) << b.radius << SEP << std::setw(15) << b.density
              << SEP << std::setw(15) << b.cd << SEP << std::setw(15) << b.oe.sma << SEP << std::setw(15) << b.oe.ecc
              << SEP << std::setw(15) << b.oe.inc << SEP << std::setw(15) << b.oe.peri << SEP << std::setw(15)
              << b.oe.node << SEP << std::setw(15) << b.oe.mean << std::endl;
    }
    {
        std::ofstream o = open_out(combine_path(dir, "start_files.txt"));
        o << fn_info << std::endl << fn_data << std::endl;
    }
}

void usage(const std::vector<Preset>& P)
{
    std::cout << "Usage: redinit -preset <name> [-seed <n>] [-odir <dir>] [-f <file name>] [-l]\n"
                 "  -preset <name>  scenario to generate\n"
                 "  -seed <n>       srand() seed (default: the clock, as the reference)\n"
                 "  -odir <dir>     output directory (default: .)\n"
                 "  -f <file name>  stem of the output files (default: the preset name)\n"
                 "  -l              list the presets\n"
                 "Presets:";
    for (const Preset& p : P) std::cout << ' ' << p.name;
    std::cout << std::endl;
}

}  // namespace

int main(int argc, char** argv)
{
    const std::vector<Preset> P = presets();
    try {
        std::string preset, dir, file;
        bool have_seed = false;
        uint32_t seed = 0;
        for (int i = 1; i < argc; ++i) {
            const std::string a = argv[i];
            auto need = [&]() -> std::string {
                if (i + 1 >= argc) throw std::string("Missing value after " + a + ".");
                return argv[++i];
            };
            if (a == "-preset") preset = need();
            else if (a == "-seed") {
                const std::string s = need();
                char* end = nullptr;
                const unsigned long v = std::strtoul(s.c_str(), &end, 10);
                if (s.empty() || *end != '\0' || v > 0xffffffffUL) throw std::string("Invalid number at: -seed");
                seed = (uint32_t)v;
                have_seed = true;
            } else if (a == "-odir") dir = need();
            else if (a == "-f") file = need();
            else if (a == "-l" || a == "-h") {
                usage(P);
                return EXIT_SUCCESS;
            } else throw std::string("Invalid switch '" + a + "' on command-line.");
        }
        const Preset* chosen = nullptr;
        for (const Preset& p : P)
            if (p.name == preset) chosen = &p;
        if (!chosen) throw std::string(preset.empty() ? "No -preset given." : "Unknown preset '" + preset + "'.");
        if (!have_seed) seed = (uint32_t)time(nullptr);
        std::cout << "The seed number is " << seed << std::endl;
        if (file.empty()) file = chosen->name;
        const Disk D = generate(*chosen, seed);
        write_all(dir, file, seed, D);
        std::cout << "Wrote " << D.bodies.size() << " bodies to " << combine_path(dir, file) << ".data.txt" << std::endl;
    } catch (const std::string& msg) {
        std::cerr << "Error: " << msg << std::endl;
        return EXIT_FAILURE;
    }
    return EXIT_SUCCESS;
}
