// red_cli.cpp -- red's command-line driver (src/red.cuda/main.cu,
// options.cpp, parameter.cpp) over the engine: the same switches, parameter
// and start files, the same text or binary data and data-info files in, and
// the same data, dump, event, integral, info and log files out, so that an
// existing red run directory can be continued with this build.
//
//   red -gpu -idir <dir> -p <parameter file> -i <start files> [-ga <gas disk file> | -gf <FARGO file>]
//       [-odir <dir>] [-id_dev <n>] [-v] [-pts] [-pe] [-ef] [-t] [-h]
//       [-prec fp64|mixed]
//
// -gpu runs the engine; the reference's CPU path (-cpu, -n_chg) is not part of
// this build and is refused.  -gf reads the FARGO run in -idir (fargo_gas_disk.hpp
// says how that differs from the reference).  -prec is this build's addition
// (default fp64: the reference's GPU arithmetic, bit for bit).
#include <sys/stat.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "fargo_gas_disk.hpp"
#include "red_host.hpp"

using namespace redhost;

namespace {

// Physical constants and unit conversions of red (red.type/red_constants.h):
// the same published values, so that files convert identically.
namespace cst {
constexpr double Gauss = 1.720209895e-2;
constexpr double Gauss2 = 2.959122082855911025e-4;
constexpr double YearToDay = 365.25;
constexpr double SolarToKilogram = 1.98911e30;
constexpr double SolarToGram = 1.0e3 * SolarToKilogram;
constexpr double KilogramToSolar = 1.0 / SolarToKilogram;
constexpr double GramToSolar = 1.0 / SolarToGram;
constexpr double AuToMeter = 1.495978707e11;
constexpr double MeterToAu = 1.0 / AuToMeter;
constexpr double CmToAu = 1.0e-2 * MeterToAu;
constexpr double GramPerCm3ToSolarPerAu3 = GramToSolar / (CmToAu * CmToAu * CmToAu);
constexpr double SecondToDay = 1.0 / 86400.0;
constexpr double Boltzman_SI = 1.3806488e-23;
constexpr double ProtonMass_SI = 1.672621777e-27;
constexpr double Boltzman_CMU = Boltzman_SI * (KilogramToSolar * MeterToAu * MeterToAu) / (SecondToDay * SecondToDay);
constexpr double ProtonMass_CMU = ProtonMass_SI * KilogramToSolar;
constexpr double PI = 3.14159265358979323846;
}  // namespace cst

constexpr int kOrdinalWidth = 9;   // OUTPUT_ORDINAL_NUMBER_WIDTH (red_macro.h:11)
constexpr char SEP = ' ';

std::string trim(const std::string& s)
{
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}
std::string lower(std::string s)
{
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return (char)std::tolower(c); });
    return s;
}
bool is_number(const std::string& s)
{
    if (s.empty()) return false;
    char* end = nullptr;
    std::strtod(s.c_str(), &end);
    return end && *end == '\0';
}
double number(const std::string& key, const std::string& value)
{
    if (!is_number(value)) throw std::string("Invalid number at: " + key);
    return std::atof(value.c_str());
}
std::string combine_path(const std::string& dir, const std::string& file)
{
    if (dir.empty()) return file;
    return dir.back() == '/' ? dir + file : dir + "/" + file;
}
std::string extension(const std::string& path)
{
    const size_t p = path.find_last_of('.');
    return p == std::string::npos ? "" : path.substr(p + 1);
}
std::string read_text(const std::string& path)
{
    std::ifstream in(path.c_str(), std::ios::in);
    if (!in) throw std::string("Cannot open " + path + ".");
    std::stringstream ss;
    ss << in.rdbuf();
    return ss.str();
}
std::string ordinal(uint32_t n)
{
    std::ostringstream o;
    o << std::setw(kOrdinalWidth) << std::setfill('0') << n;
    return o.str();
}
std::string time_stamp()
{
    const time_t now = time(nullptr);
    char buf[64];
    std::strftime(buf, sizeof(buf), "%Y-%m-%d %H:%M:%S", std::localtime(&now));
    return buf;
}

// key = value lines; a line without a value is an error (parameter.cpp:51-90)
template <typename F>
void parse_key_values(const std::string& text, F set)
{
    std::istringstream in(text);
    std::string line;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (trim(line).empty()) continue;
        std::vector<std::string> tok;
        std::string t;
        std::istringstream ls(line);
        while (std::getline(ls, t, '='))
            if (!t.empty()) tok.push_back(t);
        if (tok.size() < 2) throw std::string("Invalid key/value pair: " + line + ".");
        std::string key = lower(trim(tok[0])), value = trim(tok[1]);
        set(key, value);
    }
}

// ---- parameter file (src/red.cuda/parameter.cpp:36-252) ---------------------
struct Parameters {
    std::string text, name, desc;
    bool binary = false;
    int method = RD_RKF78;
    bool adaptive = true;
    bool error_check_for_tp = false;
    double tolerance = 1.0e-10;
    double length = 0.0, output_interval = 0.0;   // [k day] after loading
    double info_dt = 5.0, dump_dt = 3600.0;        // [s]
    double thr[RD_THRESHOLD_N] = {0.0, 0.0, 0.0};
};

Parameters load_parameters(const std::string& path)
{
    Parameters P;
    P.text = read_text(path);
    parse_key_values(P.text, [&](const std::string& key, std::string value) {
        if (key == "name") P.name = value;
        else if (key == "description") P.desc = value;
        else if (key == "output data representation" || key == "odr") {
            value = lower(value);
            if (value == "t" || value == "text") P.binary = false;
            else if (value == "b" || value == "binary") P.binary = true;
            else throw std::string("Invalid data type: " + value);
        } else if (key == "integrator") {
            value = lower(value);
            if (value == "euler" || value == "e") P.method = RD_EULER;
            else if (value == "runge-kutta2" || value == "rk2") P.method = RD_RK2;
            else if (value == "runge-kutta4" || value == "rk4") P.method = RD_RK4;
            else if (value == "runge-kutta-fehlberg8" || value == "rkf8") P.method = RD_RKF78;
            else throw std::string("Invalid integrator type: " + value);
        } else if (key == "tolerance") {
            P.adaptive = true;
            P.tolerance = number(key, value);
        } else if (key == "error check for tp") {
            if (value == "true") P.error_check_for_tp = true;
            else if (value == "false") P.error_check_for_tp = false;
            else throw std::string("Invalid value at: " + key);
        } else if (key == "length") P.length = number(key, value) * cst::YearToDay;
        else if (key == "output interval") P.output_interval = number(key, value) * cst::YearToDay;
        else if (key == "info dt") P.info_dt = number(key, value);
        else if (key == "dump dt" || key == "ddt") P.dump_dt = number(key, value);
        else if (key == "ejection") P.thr[RD_THRESHOLD_EJECTION_DISTANCE] = number(key, value);
        else if (key == "hit centrum") P.thr[RD_THRESHOLD_HIT_CENTRUM_DISTANCE] = number(key, value);
        else if (key == "radii enhance factor" || key == "collision factor")
            P.thr[RD_THRESHOLD_RADII_ENHANCE_FACTOR] = number(key, value);
        else throw std::string("Invalid parameter: " + key + ".");
    });
    P.length *= cst::Gauss;            // parameter::transform_time
    P.output_interval *= cst::Gauss;
    return P;
}

// ---- analytic gas disk file (src/red.cuda/analytic_gas_disk.cu:90-284) --------
rd_gas_params load_gas(const std::string& path, double m_star)
{
    rd_gas_params g;
    std::memset(&g, 0, sizeof(g));
    g.gas_decrease = 0;   // GAS_DENSITY_CONSTANT
    parse_key_values(read_text(path), [&](const std::string& key, const std::string& value) {
        if (key == "name" || key == "description") return;
        if (key == "time_dependence") {
            if (value == "constant" || value == "const") g.gas_decrease = 0;
            else if (value == "linear" || value == "lin") g.gas_decrease = 1;
            else if (value == "exponential" || value == "exp") g.gas_decrease = 2;
            else throw std::string("Invalid value at: " + key + ".");
        } else if (key == "t0") g.t0 = number(key, value) * cst::YearToDay;
        else if (key == "t1") g.t1 = number(key, value) * cst::YearToDay;
        else if (key == "e_folding_time") g.e_folding_time = number(key, value) * cst::YearToDay;
        else if (key == "alpha") g.alpha = number(key, value);
        else if (key == "mean_molecular_weight" || key == "mmw")
            g.mean_molecular_weight = number(key, value) * cst::GramToSolar;
        else if (key == "particle_diameter" || key == "diameter")
            g.particle_diameter = number(key, value) * cst::MeterToAu;
        else if (key == "eta_c") g.eta[0] = number(key, value);
        else if (key == "eta_p") g.eta[1] = number(key, value);
        else if (key == "rho_c") g.rho[0] = number(key, value) * cst::GramPerCm3ToSolarPerAu3;
        else if (key == "rho_p") g.rho[1] = number(key, value);
        else if (key == "sch_c") g.sch[0] = number(key, value);
        else if (key == "sch_p") g.sch[1] = number(key, value);
        else if (key == "tau_c") g.tau[0] = number(key, value);
        else if (key == "tau_p") g.tau[1] = number(key, value);
        else throw std::string("Invalid parameter :" + key + ".");
    });
    g.t0 *= cst::Gauss;   // transform_time
    g.t1 *= cst::Gauss;
    g.e_folding_time *= cst::Gauss;
    // analytic_gas_disk::calc (analytic_gas_disk.cu:258-267)
    if (g.mean_molecular_weight > 0.0) {
        g.c_vth = std::sqrt((8.0 * cst::Boltzman_CMU) / (cst::PI * g.mean_molecular_weight * cst::ProtonMass_CMU));
        if (g.particle_diameter > 0.0 && g.rho[0] > 0.0)
            g.mfp[0] = g.mean_molecular_weight * cst::ProtonMass_CMU /
                       (std::sqrt(2.0) * cst::PI * g.particle_diameter * g.particle_diameter * g.rho[0]);
        g.temp[0] = g.sch[0] * g.sch[0] * cst::Gauss2 * m_star * g.mean_molecular_weight * cst::ProtonMass_CMU /
                    cst::Boltzman_CMU;
    }
    g.mfp[1] = -g.rho[1];
    g.temp[1] = 2.0 * g.sch[1] - 3.0;
    return g;
}

// ---- data and data-info files (src/red.cuda.utility/file_util.cpp) ----------
struct Bodies {
    std::vector<std::string> name;
    std::vector<rd_var4> r, v;
    std::vector<rd_param> p;
    std::vector<rd_body_md> b;
};
struct Info {
    double t = 0.0, dt = 0.0;
    uint32_t dt_cpu = 0;
    uint32_t n[RD_BODY_TYPE_N] = {0};
};

// load_data_info_record_ascii/_binary (file_util.cpp:562-591)
Info load_info(const std::string& path, bool binary)
{
    Info I;
    std::ifstream in(path.c_str(), binary ? std::ios::in | std::ios::binary : std::ios::in);
    if (!in) throw std::string("Cannot open " + path + ".");
    if (binary) {
        in.read((char*)&I.t, sizeof(double));
        in.read((char*)&I.dt, sizeof(double));
        in.read((char*)&I.dt_cpu, sizeof(uint32_t));
        in.read((char*)I.n, sizeof(I.n));
    } else {
        in >> I.t >> I.dt >> I.dt_cpu;
        for (int t = 0; t < RD_BODY_TYPE_N; ++t) in >> I.n[t];
    }
    if (!in) throw std::string("Cannot read " + path + ".");
    return I;
}

// load_data_record_ascii/_binary (file_util.cpp:593-647)
Bodies load_data(const std::string& path, bool binary, uint32_t n)
{
    Bodies B;
    B.name.resize(n);
    B.r.assign(n, rd_var4{0, 0, 0, 0});
    B.v.assign(n, rd_var4{0, 0, 0, 0});
    B.p.resize(n);
    B.b.resize(n);
    std::ifstream in(path.c_str(), binary ? std::ios::in | std::ios::binary : std::ios::in);
    if (!in) throw std::string("Cannot open " + path + ".");
    for (uint32_t i = 0; i < n; ++i) {
        if (binary) {
            char buf[31] = {0};
            in.read(buf, 30);
            B.name[i] = buf;
            in.read((char*)&B.b[i], sizeof(rd_body_md));
            in.read((char*)&B.p[i], sizeof(rd_param));
            in.read((char*)&B.r[i], 3 * sizeof(double));
            in.read((char*)&B.v[i], 3 * sizeof(double));
        } else {
            std::string nm;
            int type = 0, mig = 0;
            in >> nm >> B.b[i].id >> type >> mig >> B.b[i].mig_stop_at;
            B.name[i] = nm.substr(0, 30);
            B.b[i].body_type = type;
            B.b[i].mig_type = mig;
            in >> B.p[i].mass >> B.p[i].radius >> B.p[i].density >> B.p[i].cd;
            in >> B.r[i].x >> B.r[i].y >> B.r[i].z >> B.v[i].x >> B.v[i].y >> B.v[i].z;
        }
        if (!in) throw std::string("Cannot read body " + std::to_string(i) + " of " + path + ".");
    }
    return B;
}

void set_sci(std::ostream& o)
{
    o.precision(16);
    o.setf(std::ios::right);
    o.setf(std::ios::scientific);
}

// print_data_info_record_ascii_RED/_binary_RED (file_util.cpp:325-354)
void write_info(const std::string& path, bool binary, double t, double dt, uint32_t dt_cpu,
                const uint32_t active[RD_BODY_TYPE_N])
{
    std::ofstream o(path.c_str(), binary ? std::ios::out | std::ios::binary : std::ios::out);
    if (!o) throw std::string("Cannot open " + path + ".");
    if (binary) {
        o.write((const char*)&t, sizeof(double));
        o.write((const char*)&dt, sizeof(double));
        o.write((const char*)&dt_cpu, sizeof(uint32_t));
        o.write((const char*)active, RD_BODY_TYPE_N * sizeof(uint32_t));
    } else {
        set_sci(o);
        o << std::setw(25) << t << SEP << std::setw(25) << dt << SEP << std::setw(10) << dt_cpu << SEP;
        for (int q = 0; q < RD_BODY_TYPE_N; ++q) o << std::setw(8) << active[q] << SEP;
    }
}

// print_body_record_ascii_RED/_binary_RED (file_util.cpp:356-390); inactive
// bodies are skipped (pp_disk.cu:1855-1880)
void write_data(const std::string& path, bool binary, const Bodies& B)
{
    std::ofstream o(path.c_str(), binary ? std::ios::out | std::ios::binary : std::ios::out);
    if (!o) throw std::string("Cannot open " + path + ".");
    if (!binary) set_sci(o);
    for (size_t i = 0; i < B.r.size(); ++i) {
        if (B.b[i].id < 0) continue;
        if (binary) {
            char buf[30] = {0};
            std::memcpy(buf, B.name[i].data(), std::min<size_t>(30, B.name[i].size()));
            o.write(buf, 30);
            o.write((const char*)&B.b[i], sizeof(rd_body_md));
            o.write((const char*)&B.p[i], sizeof(rd_param));
            o.write((const char*)&B.r[i], 3 * sizeof(double));
            o.write((const char*)&B.v[i], 3 * sizeof(double));
        } else {
            o << std::setw(30) << B.name[i] << SEP << std::setw(8) << B.b[i].id << SEP << std::setw(2)
              << B.b[i].body_type << SEP << std::setw(2) << B.b[i].mig_type << SEP << std::setw(25)
              << B.b[i].mig_stop_at << SEP << std::setw(25) << B.p[i].mass << SEP << std::setw(25) << B.p[i].radius
              << SEP << std::setw(25) << B.p[i].density << SEP << std::setw(25) << B.p[i].cd << SEP << std::setw(25)
              << B.r[i].x << SEP << std::setw(25) << B.r[i].y << SEP << std::setw(25) << B.r[i].z << SEP
              << std::setw(25) << B.v[i].x << SEP << std::setw(25) << B.v[i].y << SEP << std::setw(25) << B.v[i].z
              << std::endl;
        }
    }
}

// pp_disk::print_integral_data (pp_disk.cu:2100-2127)
void append_integral(const std::string& path, double t, const rd_integral& I)
{
    std::ofstream o(path.c_str(), std::ios::out | std::ios::app);
    if (!o) throw std::string("Cannot open " + path + ".");
    set_sci(o);
    const double w[11] = {t / cst::Gauss, I.R.x, I.R.y, I.R.z, I.V.x * cst::Gauss, I.V.y * cst::Gauss,
                          I.V.z * cst::Gauss, I.C.x * cst::Gauss, I.C.y * cst::Gauss, I.C.z * cst::Gauss,
                          I.E * cst::Gauss2};
    for (double x : w) o << std::setw(25) << x << SEP;
    o << std::endl;
}

// pp_disk::print_event_data (pp_disk.cu:2008-2098)
void append_events(const std::string& path, std::ostream* log, const std::vector<rd_event>& ev, size_t from)
{
    static const char* names[] = {"NONE", "HIT_CENTRUM", "EJECTION", "COLLISION"};
    std::ofstream o(path.c_str(), std::ios::out | std::ios::app);
    if (!o) throw std::string("Cannot open " + path + ".");
    set_sci(o);
    for (size_t q = from; q < ev.size(); ++q) {
        const rd_event& e = ev[q];
        const char* nm = (e.event_name >= 0 && e.event_name <= 3) ? names[e.event_name] : "NONE";
        o << std::setw(16) << nm << SEP << std::setw(25) << e.t << SEP << std::setw(25) << e.d << SEP << std::setw(8)
          << e.id0 << SEP << std::setw(8) << e.id1 << SEP << std::setw(8) << e.id2 << SEP;
        const rd_param* ps[4] = {&e.p0, &e.p1, &e.p2, &e.ps};
        const rd_var4* rs[4] = {&e.r0, &e.r1, &e.r2, &e.rs};
        const rd_var4* vs[4] = {&e.v0, &e.v1, &e.v2, &e.vs};
        for (int u = 0; u < 4; ++u)
            o << std::setw(25) << ps[u]->mass << SEP << std::setw(25) << ps[u]->density << SEP << std::setw(25)
              << ps[u]->radius << SEP << std::setw(25) << rs[u]->x << SEP << std::setw(25) << rs[u]->y << SEP
              << std::setw(25) << rs[u]->z << SEP << std::setw(25) << vs[u]->x << SEP << std::setw(25) << vs[u]->y
              << SEP << std::setw(25) << vs[u]->z << SEP;
        o << std::endl;
        if (log) {
            set_sci(*log);
            *log << time_stamp() << SEP << nm << SEP << std::setw(8) << e.id1 << SEP << std::setw(8) << e.id2 << SEP
                 << std::endl;
        }
    }
}

// ---- command line (src/red.cuda/options.cpp:56-175) ---------------------------
struct Options {
    bool test = false, verbose = false, print_to_screen = false, print_dbe = false, ef = false;
    int id_dev = 0;
    std::string idir, odir, start_files, param_file, gas_file;
    bool fargo = false;   // gas_file names a FARGO parameter file (-gf), not an analytic disk (-ga)
    int precision = RD_PREC_FP64;
};

void usage()
{
    std::cout << "Usage: red [options] -i <file> -p <file>\n"
                 "Options:\n"
                 "  -t                Run self-tests\n"
                 "  -v                Verbose mode (log all event during the execution to the log file)\n"
                 "  -pe               Print the data before the event is detected\n"
                 "  -pts              Verbose mode and print everything to the stdout too\n\n"
                 "  -id_dev <number>  The id of the device which will execute the code (default value is 0)\n"
                 "  -gpu              Execute the code on the graphics processing unit (GPU)\n"
                 "  -prec <fp64|mixed> Pair-force arithmetic (default fp64: the reference's GPU arithmetic)\n\n"
                 "  -idir <directory> The directory containing the input files\n"
                 "  -odir <directory> The directory where the output files will be stored (default: -idir)\n\n"
                 "  -i <file>         The input file containing the filenames of the input data\n"
                 "  -p <file>         The input file containing the parameters of the simulation\n"
                 "  -ga <file>        The input file containing the parameters of an analytic gas disk\n"
                 "  -gf <file>        The input file containing the details of the gas disk resulted from FARGO simulations\n"
                 "                    (in -idir with used_rad.dat and frame 0, taken in AU, Msun/AU^3 and AU/tau)\n"
                 "  -ef               Prefix the output files with the run configuration\n\n"
                 "  -h                Display this information\n\n"
                 "The CPU path (-cpu, -n_chg) is not part of this build.\n";
}

Options parse_options(int argc, char** argv)
{
    Options o;
    auto need = [&](int& i, const std::string& sw) -> std::string {
        if (i + 1 >= argc) throw std::string("Missing value after " + sw + ".");
        return argv[++i];
    };
    for (int i = 1; i < argc; ++i) {
        const std::string p = argv[i];
        if (p == "-t") o.test = true;
        else if (p == "-v") o.verbose = true;
        else if (p == "-pts") o.verbose = o.print_to_screen = true;
        else if (p == "-pe") o.print_dbe = true;
        else if (p == "-ef") o.ef = true;
        else if (p == "-id_dev") {
            const std::string v = need(i, p);
            if (!is_number(v)) throw std::string("Invalid number at: " + p);
            o.id_dev = std::atoi(v.c_str());
        } else if (p == "-gpu") {
        } else if (p == "-cpu" || p == "-n_chg") {
            throw std::string("The CPU path (" + p + ") is the reference's; this build runs the GPU path (-gpu).");
        } else if (p == "-prec") {
            const std::string v = lower(need(i, p));
            if (v == "fp64") o.precision = RD_PREC_FP64;
            else if (v == "mixed") o.precision = RD_PREC_MIXED;
            else throw std::string("Invalid value at: -prec (fp64 or mixed).");
        } else if (p == "-idir") o.idir = need(i, p);
        else if (p == "-odir") o.odir = need(i, p);
        else if (p == "-i") o.start_files = need(i, p);
        else if (p == "-p") o.param_file = need(i, p);
        else if (p == "-ga") {
            o.gas_file = need(i, p);
            o.fargo = false;
        } else if (p == "-gf") {
            o.gas_file = need(i, p);
            o.fargo = true;
        }
        else if (p == "-h") {
            usage();
            std::exit(EXIT_SUCCESS);
        } else throw std::string("Invalid switch on command-line: " + p + ".");
    }
    if (o.odir.empty()) o.odir = o.idir;
    return o;
}

// ---- the run -----------------------------------------------------------------
struct Run {
    Options opt;
    Parameters par;
    std::string prefix, ext, data_fn;
    std::string path_info, path_integral, path_integral_event, path_event;
    std::vector<std::string> names;   // per body of the engine's arrays (compacted with them)
    std::ofstream* slog = nullptr;
    uint32_t dt_cpu_offset = 0;
    double t_cpu = 0.0;               // [s] CPU time of this run's steps
    double dt_file = 0.0;             // pp_disk::dt: the step read from the input, written unchanged
    size_t n_logged_events = 0;

    void log(const std::string& msg)
    {
        if (slog) *slog << time_stamp() << SEP << msg << std::endl;
        if (opt.print_to_screen) std::cout << time_stamp() << SEP << msg << std::endl;
    }

    Bodies state(pp_disk& ppd, double* t)
    {
        Bodies B;
        ppd.copy_to_host(B.r, B.v, B.p, B.b, *t);
        B.name = names;
        return B;
    }
    uint32_t dt_cpu() const { return dt_cpu_offset + (uint32_t)t_cpu; }

    void print_data_files(pp_disk& ppd, integrator& intgr, const std::string& stem, bool write_start_files)
    {
        double t;
        Bodies B = state(ppd, &t);
        const std::string fn_data = prefix + stem + "." + ext;
        const std::string fn_info = prefix + stem + ".info." + ext;
        write_data(combine_path(opt.odir, fn_data), par.binary, B);
        const n_objects n = ppd.counts();
        uint32_t active[RD_BODY_TYPE_N];
        for (int q = 0; q < RD_BODY_TYPE_N; ++q) active[q] = n.playing[q] - n.inactive[q];
        (void)intgr;
        write_info(combine_path(opt.odir, fn_info), par.binary, t, dt_file, dt_cpu(), active);
        if (write_start_files) {
            const std::string path = combine_path(opt.odir, "start_files.txt");
            std::ofstream s(path.c_str(), std::ios::out);
            if (!s) throw std::string("Cannot open " + path + "!");
            s << fn_info << std::endl << fn_data << std::endl;
        }
    }

    // main.cu print_data: snapshot n_print, start files, integral
    void print_data(pp_disk& ppd, integrator& intgr, uint32_t& n_print)
    {
        print_data_files(ppd, intgr, "data_" + ordinal(n_print), true);
        n_print++;
        rd_integral I;
        ppd.calc_integral(I);
        append_integral(path_integral, ppd.t(), I);
    }
    void print_dump(pp_disk& ppd, integrator& intgr)
    {
        print_data_files(ppd, intgr, "dump", true);
        if (opt.verbose) log("Dump file was created");
    }

    // main.cu print_info: one line to the screen, one CSV row to info.txt
    void print_info(pp_disk& ppd, integrator& intgr, double dt, double dT)
    {
        static bool first = true;
        const n_objects n = ppd.counts();
        const auto c = ppd.counters_of_events();
        const double t = ppd.t() / cst::Gauss;
        const uint64_t np = std::max<uint64_t>(1, intgr.get_n_passed_step());
        uint32_t n_inactive = 0, n_removed = 0;
        for (int q = 0; q < RD_BODY_TYPE_N; ++q) {
            n_inactive += n.inactive[q];
            n_removed += n.removed[q];
        }
        std::cout.setf(std::ios::right);
        std::cout.setf(std::ios::scientific);
        std::cout << "[GPU] " << time_stamp() << " t: " << std::setprecision(4) << std::setw(10) << t
                  << ", dt: " << std::setw(10) << dt / cst::Gauss << " (" << std::setw(10) << t / np << ") [d]"
                  << ", dT: " << std::setw(10) << dT << " (" << std::setw(10) << t_cpu / np << ") [s]"
                  << ", Nc: " << std::setw(5) << c.n_collision[0] << ", Ne: " << std::setw(5) << c.n_ejection[0]
                  << ", Nh: " << std::setw(5) << c.n_hit_centrum[0] << ", N : " << std::setw(6)
                  << n.get_n_total_playing() << "(" << std::setw(3) << n_inactive << ", " << std::setw(5) << n_removed
                  << ")" << ", nt: " << std::setw(11) << intgr.get_n_tried_step() << ", np: " << std::setw(11)
                  << intgr.get_n_passed_step() << ", nf: " << std::setw(11) << intgr.get_n_failed_step()
                  << std::endl;
        std::ofstream s(path_info.c_str(), std::ios::out | std::ios::app);
        if (!s) throw std::string("Cannot open " + path_info + ".");
        if (first) {
            first = false;
            s << "dev,date      ,time    ,t [d]      ,dt [d]     ,dt_avg [d] ,dT [s]     ,dT_avg [s] ,Nc   ,Ne   ,"
                 "Nh   ,nb_p ,nb_i,nb_r ,ns_t    ,ns_p    ,ns_f    ,ns_a,ns_r  ,ngp_a,ngp_r,nrp_a,nrp_r,npp_a,"
                 "npp_r,nspl_a,nspl_r,npl_a,npl_r,ntp_a,ntp_r"
              << std::endl;
        }
        s.setf(std::ios::right);
        s.setf(std::ios::scientific);
        const char cvs = ',';
        const std::string ts = time_stamp();
        s << "GPU" << cvs << ts.substr(0, 10) << cvs << ts.substr(11) << cvs << std::setprecision(4) << t << cvs
          << dt / cst::Gauss << cvs << t / np << cvs << dT << cvs << t_cpu / np << cvs << c.n_collision[0] << cvs
          << c.n_ejection[0] << cvs << c.n_hit_centrum[0] << cvs << n.get_n_total_playing() << cvs << n_inactive
          << cvs << n_removed << cvs << intgr.get_n_tried_step() << cvs << intgr.get_n_passed_step() << cvs
          << intgr.get_n_failed_step() << cvs;
        for (int q = 0; q < RD_BODY_TYPE_N; ++q)
            s << n.playing[q] - n.inactive[q] << cvs << n.removed[q] << (q < RD_BODY_TYPE_N - 1 ? "," : "");
        s << std::endl;
    }

    // events just handled: integral before/after, event records, dump
    void after_events(pp_disk& ppd, integrator& intgr, const rd_integral& before)
    {
        append_integral(path_integral_event, ppd.t(), before);
        rd_integral I;
        ppd.calc_integral(I);
        append_integral(path_integral_event, ppd.t(), I);
        const std::vector<rd_event> ev = ppd.event_log();
        append_events(path_event, slog, ev, n_logged_events);
        n_logged_events = ev.size();
        ppd.clear_event_counter();
        print_dump(ppd, intgr);
    }

    // pp_disk::rebuild_vectors keeps the active bodies in order and their names
    // with them (pp_disk.cu:1370-1420): rebuild the engine, then the names by
    // the ids read before it
    void rebuild(pp_disk& ppd)
    {
        double t;
        std::vector<rd_var4> r, v;
        std::vector<rd_param> p;
        std::vector<rd_body_md> b;
        ppd.copy_to_host(r, v, p, b, t);
        ppd.rebuild_vectors();
        std::vector<std::string> kept;
        for (size_t i = 0; i < b.size() && i < names.size(); ++i)
            if (b[i].id > 0) kept.push_back(names[i]);
        names.swap(kept);
    }
};

std::string create_prefix(const Options& o, const Parameters& p)
{
    if (!o.ef) return "";
    // main.cu:100-131 create_prefix: configuration, device, adaptive/fixed, integrator
    static const char* short_name[] = {"E", "RK2", "RK4", "RK5", "RKF7"};
    return std::string("R_gpu_") + (p.adaptive ? "as" : "fs") + "_" + short_name[p.method] + "_";
}

int self_test(int id_dev)
{
    // A Kepler orbit (G = 1, r = 1, v = 1 perpendicular, masses 1 and 1e-9):
    // after 100 RKF7(8) steps at tolerance 1e-12 the relative position must lie
    // within 1e-8 of the circle at angle n t, n^2 = mu / a^3 with mu = 1 + 1e-9
    // and a from vis-viva (the eccentricity, ~1e-9, sets the bound), and the
    // energy must hold to 1e-9.
    std::vector<rd_var4> r = {{0, 0, 0, 0}, {1.0, 0, 0, 0}}, v = {{0, 0, 0, 0}, {0, 1.0, 0, 0}};
    std::vector<rd_param> p = {{1.0, 0.00465, 1.0, 0.0}, {1e-9, 1e-5, 1.0, 0.0}};
    std::vector<rd_body_md> b(2);
    b[0] = rd_body_md{1, RD_BODY_STAR, 0, 0.0};
    b[1] = rd_body_md{2, RD_BODY_ROCKYPLANET, 0, 0.0};
    const uint32_t nbt[RD_BODY_TYPE_N] = {1, 0, 1, 0, 0, 0, 0};
    const double thr[RD_THRESHOLD_N] = {0, 0, 0};
    pp_disk ppd(r, v, p, b, nbt, 0.0, thr, nullptr, id_dev);
    rungekutta8 it(&ppd, 0.01, true, 1e-12);
    rd_integral I0, I1;
    ppd.calc_integral(I0);
    for (int s = 0; s < 100; ++s) it.step();
    ppd.calc_integral(I1);
    double t;
    ppd.copy_to_host(r, v, p, b, t);
    const double mu = 1.0 + 1e-9, a = 1.0 / (2.0 - 1.0 / mu), n = std::sqrt(mu / (a * a * a));
    const double dx = r[1].x - r[0].x - std::cos(n * t), dy = r[1].y - r[0].y - std::sin(n * t);
    const double pos = std::sqrt(dx * dx + dy * dy + (r[1].z - r[0].z) * (r[1].z - r[0].z));
    const double en = std::fabs((I1.E - I0.E) / I0.E);
    const bool ok = pos < 1e-8 && en < 1e-9 && t > 1.0;
    std::cout << "self-test: two-body RKF7(8) to t = " << t << ": distance from the Kepler orbit " << pos
              << ", relative energy error " << en << (ok ? " ok" : " FAILED") << std::endl;
    return ok ? EXIT_SUCCESS : EXIT_FAILURE;
}

}  // namespace

int main(int argc, char** argv)
{
    const time_t start = time(nullptr);
    Run run;
    try {
        run.opt = parse_options(argc, argv);
        if (run.opt.test) return self_test(run.opt.id_dev);
        if (run.opt.param_file.empty()) throw std::string("The -p <file> option was not provided.");
        if (run.opt.start_files.empty()) throw std::string("The -i <file> option was not provided.");
        // parameter.cpp:11-28: a file name with a directory is used as given
        const std::string ppath = run.opt.param_file.find('/') != std::string::npos
                                      ? run.opt.param_file
                                      : combine_path(run.opt.idir, run.opt.param_file);
        run.par = load_parameters(ppath);
        run.prefix = create_prefix(run.opt, run.par);
        run.ext = run.par.binary ? "dat" : "txt";
        run.path_info = combine_path(run.opt.odir, run.prefix + "info.txt");
        run.path_integral = combine_path(run.opt.odir, run.prefix + "integral.txt");
        run.path_integral_event = combine_path(run.opt.odir, run.prefix + "integral.event.txt");
        run.path_event = combine_path(run.opt.odir, run.prefix + "event.txt");
        const std::string path_log = combine_path(run.opt.odir, run.prefix + "log.txt");
        run.slog = new std::ofstream(path_log.c_str(), std::ios::out | std::ios::app);
        if (!*run.slog) throw std::string("Cannot open " + path_log + ".");
        {
            std::ostringstream cmd;
            for (int i = 0; i < argc; ++i) cmd << (i ? " " : "") << argv[i];
            *run.slog << time_stamp() << SEP << "Command line: " << cmd.str() << std::endl;
            *run.slog << "Parameters:" << std::endl << run.par.text << std::endl;
        }

        // options::create_pp_disk (options.cpp:175-251): start files -> data info + data
        std::vector<std::string> sf;
        {
            std::istringstream in(read_text(combine_path(run.opt.idir, run.opt.start_files)));
            std::string line;
            while (std::getline(in, line)) {
                if (!line.empty() && line.back() == '\r') line.pop_back();
                sf.push_back(trim(line));
            }
        }
        if (sf.size() < 2) throw std::string("The start file must name the data-info and the data file.");
        run.data_fn = sf[1];
        const std::string ext = extension(sf[1]);
        bool binary;
        if (ext == "txt") binary = false;
        else if (ext == "dat") binary = true;
        else throw std::string("The extension of the input data must be either 'txt' or 'dat'.");
        const Info info = load_info(combine_path(run.opt.idir, sf[0]), binary);
        uint32_t n = 0;
        for (int q = 0; q < RD_BODY_TYPE_N; ++q) n += info.n[q];
        Bodies B = load_data(combine_path(run.opt.idir, sf[1]), binary, n);
        run.names = B.name;
        run.dt_cpu_offset = info.dt_cpu;

        rd_gas_params gas;
        const bool has_gas = !run.opt.gas_file.empty() && !run.opt.fargo;
        if (has_gas) gas = load_gas(combine_path(run.opt.idir, run.opt.gas_file), n ? B.p[0].mass : 1.0);
        // the FARGO files are read before the GPU is touched, as the other inputs
        std::unique_ptr<fargo_gas_disk> fgd;
        if (run.opt.fargo) fgd.reset(new fargo_gas_disk(run.opt.idir, run.opt.gas_file));
        pp_disk ppd(B.r, B.v, B.p, B.b, info.n, info.t, run.par.thr, has_gas ? &gas : nullptr, run.opt.id_dev);
        if (fgd) ppd.set_gas_grid(fgd->grid());
        if (run.opt.precision == RD_PREC_MIXED) ppd.set_precision(RD_PREC_MIXED);
        double dt = info.dt;
        if (run.par.length < 0.0 && dt > 0.0) dt = -dt;   // backward integration
        run.dt_file = dt;
        // options::create_integrator (options.cpp:253-282)
        integrator* intgr = nullptr;
        switch (run.par.method) {
        case RD_EULER: intgr = new euler(&ppd, dt); break;
        case RD_RK2: intgr = new rungekutta2(&ppd, dt); break;
        case RD_RK4: intgr = new rungekutta4(&ppd, dt, run.par.adaptive, run.par.tolerance); break;
        default: intgr = new rungekutta8(&ppd, dt, run.par.adaptive, run.par.tolerance); break;
        }
        if (run.par.error_check_for_tp) intgr->set_error_check_for_tp(true);

        // run_simulation (main.cu:403-640)
        double it_t = 0.0, ps = 0.0, step_dt = 0.0, dT = 0.0;
        clock_t last_info = clock(), last_dump = clock();
        uint32_t n_print = 0, n_event_files = 0;
        if (4 <= run.data_fn.size() && run.data_fn.compare(0, 4, "data") == 0) {
            const size_t pos = run.data_fn.find_first_of('_');
            if (pos != std::string::npos) n_print = (uint32_t)std::atoi(run.data_fn.substr(pos + 1, kOrdinalWidth).c_str());
            n_print++;
        }
        if (0 == n_print && run.data_fn.compare(0, 4, "dump") != 0) run.print_data(ppd, *intgr, n_print);
        if (run.opt.precision == RD_PREC_MIXED) {
            const rd_tuning t = ppd.benchmark();
            if (run.opt.verbose)
                run.log("Launch shape set to " + std::to_string(t.sinks_per_lane) + " sinks/lane, chunk " +
                        std::to_string(t.chunk) + ", " + std::to_string(t.blocks_per_cu) + " blocks/CU");
        }
        Bodies before_step;   // -pe: the state the colliding step started from
        const double* thr = run.par.thr;
        while (std::fabs(it_t) <= std::fabs(run.par.length) && 1 < ppd.counts().get_n_total_active()) {
            if (0.0 < thr[RD_THRESHOLD_EJECTION_DISTANCE] || 0.0 < thr[RD_THRESHOLD_HIT_CENTRUM_DISTANCE]) {
                if (ppd.check_for_ejection_hit_centrum()) {
                    rd_integral I;
                    ppd.calc_integral(I);
                    ppd.handle_ejection_hit_centrum();
                    run.after_events(ppd, *intgr, I);
                    last_dump = clock();
                }
            }
            double tb = 0.0;
            if (run.opt.print_dbe) before_step = run.state(ppd, &tb);
            const clock_t c0 = clock();
            step_dt = intgr->step();
            dT = (double)(clock() - c0) / CLOCKS_PER_SEC;
            run.t_cpu += dT;
            it_t += step_dt;
            ps += step_dt;
            if (std::fabs(run.par.output_interval) < std::fabs(intgr->get_dt_next()))
                intgr->set_dt_next(0.0 < step_dt ? run.par.output_interval : -run.par.output_interval);
            if (0.0 < thr[RD_THRESHOLD_RADII_ENHANCE_FACTOR]) {
                if (ppd.check_for_collision()) {
                    if (run.opt.print_dbe) {
                        const std::string stem = "event_" + ordinal(n_event_files++);
                        write_data(combine_path(run.opt.odir, run.prefix + stem + "." + run.ext), run.par.binary,
                                   before_step);
                        const n_objects nb = ppd.counts();
                        uint32_t active[RD_BODY_TYPE_N];
                        for (int q = 0; q < RD_BODY_TYPE_N; ++q) active[q] = nb.playing[q] - nb.inactive[q];
                        write_info(combine_path(run.opt.odir, run.prefix + stem + ".info." + run.ext),
                                   run.par.binary, tb, run.dt_file, run.dt_cpu(), active);
                    }
                    rd_integral I;
                    ppd.calc_integral(I);
                    ppd.handle_collision();
                    run.after_events(ppd, *intgr, I);
                    last_dump = clock();
                }
            }
            if (run.par.output_interval <= std::fabs(ps)) {
                ps = 0.0;
                run.print_data(ppd, *intgr, n_print);
                if (run.opt.verbose) run.log("Data file was created");
            }
            if (16 <= ppd.counters_of_events().n_event[RD_EVENT_COUNTER_LAST_CLEAR]) {
                ppd.set_event_counter(RD_EVENT_COUNTER_LAST_CLEAR, 0);
                run.rebuild(ppd);
                if (run.opt.verbose)
                    run.log("Rebuild the vectors (" + std::to_string(ppd.counts().get_n_total_playing()) +
                            " bodies left at t: " + std::to_string(ppd.t() / cst::Gauss) + " [d])");
                if (run.opt.precision == RD_PREC_MIXED) ppd.benchmark();
            }
            if (run.par.dump_dt < (double)(clock() - last_dump) / CLOCKS_PER_SEC) {
                last_dump = clock();
                run.print_dump(ppd, *intgr);
            }
            if (run.par.info_dt < (double)(clock() - last_info) / CLOCKS_PER_SEC) {
                last_info = clock();
                run.print_info(ppd, *intgr, step_dt, dT);
            }
        }
        run.print_info(ppd, *intgr, step_dt, dT);
        if (0.0 < std::fabs(ps)) {
            ps = 0.0;
            run.print_data(ppd, *intgr, n_print);
        }
        delete intgr;
    } catch (const std::string& msg) {
        std::cerr << "Error: " << msg << std::endl;
        if (run.slog) *run.slog << time_stamp() << SEP << "Error: " << msg << std::endl;
        delete run.slog;
        return EXIT_FAILURE;
    }
    const time_t total = run.dt_cpu_offset + time(nullptr) - start;
    std::cout << "Total time: " << total << " s" << std::endl;
    if (run.slog) *run.slog << time_stamp() << SEP << "Total time: " << total << " s" << std::endl;
    delete run.slog;
    return EXIT_SUCCESS;
}
