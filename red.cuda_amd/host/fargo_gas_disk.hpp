// fargo_gas_disk.hpp -- the reference's fargo_gas_disk (src/red.cuda/
// fargo_gas_disk.h:8-79, fargo_gas_disk.cu:15-555) over the C ABI: read a
// FARGO run's parameter file, its ring radii (used_rad.dat) and frame 0
// (gasdens0.dat, gasvrad0.dat, gasvtheta0.dat), and hand the grid to
// rd_disk_set_gas_grid.
//
// Files, keys and error texts are the reference's.  Three things differ:
//   - the directory is the one given to the constructor; the reference's
//     options pass it the data file's name (options.cpp:233), so no path it
//     builds names a FARGO file;
//   - the frames are taken as they are, in the solver's units (density
//     [Msun/AU^3], velocities [AU/tau]), where the reference's transform_time,
//     transform_velocity and transform_density throw "not yet implemented"
//     (fargo_gas_disk.cu:252-274).  FARGO's own units with r0 = 1 AU and
//     M = 1 Msun make the time and velocity transforms the identity; the
//     density must be the midplane volume density.
//   - used_rad.dat must hold n_rad + 1 radii (FARGO writes the n_rad + 1 ring
//     interfaces; the reference stores them in an array of n_rad).
#pragma once

#include <algorithm>
#include <cctype>
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/red_abi.h"

namespace redhost {

// fargo_gas_disk_params_t (red.type/red_type.h): what a FARGO parameter file sets
struct fargo_gas_disk_params {
    double aspect_ratio = 0.0, sigma_0 = 0.0, alpha_viscosity = 0.0, sigma_slope = 0.0, flaring_index = 0.0;
    bool exclude_hill = true;
    double thickness_smoothing = 0.0, omega_frame = 0.0;
    int n_rad = 0, n_sec = 0;
    double r_min = 0.0, r_max = 0.0;
    int n_tot = 0, n_interm = 0;
    double dT = 0.0;
    double visc_mod_r1 = 0.0, visc_mod_delta_r1 = 0.0, visc_mod_r2 = 0.0, visc_mod_delta_r2 = 0.0, visc_mod = 0.0;
};

class fargo_gas_disk {
  public:
    fargo_gas_disk(const std::string& dir, const std::string& filename) : dir_(dir)
    {
        parse(load_ascii(path_of(filename)));
        if (params.n_rad < 1 || params.n_sec < 1)
            throw std::string("The FARGO parameter file does not set Nrad and Nsec.");
        const size_t n = (size_t)params.n_rad * (size_t)params.n_sec;
        load_used_rad(path_of("used_rad.dat"));
        density_ = load_frame(path_of("gasdens0.dat"), n);
        vrad_ = load_frame(path_of("gasvrad0.dat"), n);
        vtheta_ = load_frame(path_of("gasvtheta0.dat"), n);
    }

    // the grid for rd_disk_set_gas_grid (points into this object)
    rd_gas_grid grid() const
    {
        rd_gas_grid g;
        g.n_rad = (uint32_t)params.n_rad;
        g.n_sec = (uint32_t)params.n_sec;
        g.used_rad = used_rad_.data();
        g.density = density_.data();
        g.vrad = vrad_.data();
        g.vtheta = vtheta_.data();
        return g;
    }

    fargo_gas_disk_params params;
    // string-valued keys, kept as read
    std::string exclude_hill, planet_config, transport, inner_boundary, outer_boundary, disk, frame, indirect_term,
        radial_spacing, output_dir;

  private:
    std::string dir_;
    std::vector<double> used_rad_, density_, vrad_, vtheta_;

    std::string path_of(const std::string& file) const
    {
        if (dir_.empty()) return file;
        return dir_.back() == '/' ? dir_ + file : dir_ + "/" + file;
    }
    // file::load_ascii_file (file_util.cpp:101-130): each line cut at the first
    // '#', trimmed of white space, dropped when empty, ended by '\n'
    static std::string load_ascii(const std::string& path)
    {
        std::ifstream in(path.c_str(), std::ios::in);
        if (!in) throw std::string("The file '" + path + "' could not opened!\r\n");
        std::string out, line;
        while (std::getline(in, line)) {
            const size_t c = line.find('#');
            if (c != std::string::npos) line.erase(c);
            line = trimmed(line);
            if (line.empty()) continue;
            out += line;
            out.push_back('\n');
        }
        return out;
    }
    // tools::is_number (tools.cpp:29-43): digits, e, E, '.', '-' and '+' only
    static bool ref_number(const std::string& s)
    {
        return std::all_of(s.begin(), s.end(), [](char c) {
            return std::isdigit((unsigned char)c) || c == 'e' || c == 'E' || c == '.' || c == '-' || c == '+';
        });
    }
    static std::string trimmed(const std::string& s)
    {
        const char* ws = " \t\n\r\f\v";
        const size_t a = s.find_first_not_of(ws);
        if (a == std::string::npos) return "";
        return s.substr(a, s.find_last_not_of(ws) - a + 1);
    }
    // Lines split at '\r' / '\n' (empty ones skipped), each "key value ..." with
    // blanks or tabs between; words past the second are ignored
    // (fargo_gas_disk::parse, fargo_gas_disk.cu:290-328).
    void parse(const std::string& text)
    {
        size_t pos = 0;
        while (pos < text.size()) {
            const size_t end = text.find_first_of("\r\n", pos);
            const std::string line = text.substr(pos, end == std::string::npos ? std::string::npos : end - pos);
            pos = end == std::string::npos ? text.size() : end + 1;
            if (line.empty()) continue;
            std::vector<std::string> words;
            size_t p = 0;
            while (words.size() < 2) {
                const size_t a = line.find_first_not_of(" \t", p);
                if (a == std::string::npos) break;
                const size_t b = line.find_first_of(" \t", a);
                words.push_back(line.substr(a, b == std::string::npos ? std::string::npos : b - a));
                if (b == std::string::npos) break;
                p = b;
            }
            if (words.size() < 2) throw std::string("Invalid key/value pair: " + line + ".");
            set_param(words[0], words[1]);
        }
    }
    // fargo_gas_disk::set_param (fargo_gas_disk.cu:330-555)
    void set_param(std::string key, std::string value)
    {
        key = trimmed(key);
        value = trimmed(value);
        std::transform(key.begin(), key.end(), key.begin(), [](unsigned char c) { return (char)std::tolower(c); });
        auto real = [&](double& dst) {
            if (!ref_number(value)) throw std::string("Invalid number at: " + key + ".");
            dst = std::atof(value.c_str());
        };
        auto integer = [&](int& dst) {
            if (!ref_number(value)) throw std::string("Invalid number at: " + key + ".");
            dst = std::atoi(value.c_str());
        };
        fargo_gas_disk_params& P = params;
        if (key == "aspectratio") real(P.aspect_ratio);
        else if (key == "sigma0") real(P.sigma_0);
        else if (key == "alphaviscosity") real(P.alpha_viscosity);
        else if (key == "sigmaslope") real(P.sigma_slope);
        else if (key == "flaringindex") real(P.flaring_index);
        else if (key == "excludehill") {
            exclude_hill = value;
            P.exclude_hill = value == "YES";
        } else if (key == "planetconfig") planet_config = value;
        else if (key == "thicknesssmoothing") real(P.thickness_smoothing);
        else if (key == "transport") transport = value;
        else if (key == "innerboundary") inner_boundary = value;
        else if (key == "outerboundary") outer_boundary = value;
        else if (key == "disk") disk = value;
        else if (key == "omegaframe") real(P.omega_frame);
        else if (key == "frame") frame = value;
        else if (key == "indirectterm") indirect_term = value;
        else if (key == "nrad") integer(P.n_rad);
        else if (key == "nsec") integer(P.n_sec);
        else if (key == "rmin") real(P.r_min);
        else if (key == "rmax") real(P.r_max);
        else if (key == "radialspacing") radial_spacing = value;
        else if (key == "ntot") integer(P.n_tot);
        else if (key == "ninterm") integer(P.n_interm);
        else if (key == "dt") real(P.dT);
        else if (key == "outputdir") output_dir = value;
        else if (key == "viscmodr1") real(P.visc_mod_r1);
        else if (key == "viscmoddeltar1") real(P.visc_mod_delta_r1);
        else if (key == "viscmodr2") real(P.visc_mod_r2);
        else if (key == "viscmoddeltar2") real(P.visc_mod_delta_r2);
        else if (key == "viscmod") real(P.visc_mod);
        else if (key == "damprmin" || key == "damprmax") {
        } else throw std::string("Invalid parameter :" + key + ".");
    }
    // One radius per line, at most 30 characters each; a line longer than
    // that is read in pieces of 30 with the character between them dropped,
    // as the reference's reader does (fargo_gas_disk.cu:225-250).
    void load_used_rad(const std::string& path)
    {
        const std::string text = load_ascii(path);
        used_rad_.clear();
        for (size_t i = 0; i < text.size(); i++) {
            std::string num;
            while (i < text.size() && num.size() < 30 && text[i] != '\n') num += text[i++];
            if (!ref_number(num)) throw std::string("Invalid number (" + num + ") in file '" + path + ".");
            used_rad_.push_back(std::atof(num.c_str()));
        }
        if (used_rad_.size() != (size_t)params.n_rad + 1)
            throw std::string("The file '" + path + "' holds " + std::to_string(used_rad_.size()) +
                              " radii; Nrad + 1 = " + std::to_string(params.n_rad + 1) + " expected.");
    }
    // file::load_binary_file (file_util.cpp:132-153): n doubles, nothing more
    static std::vector<double> load_frame(const std::string& path, size_t n)
    {
        std::ifstream in(path.c_str(), std::ios::in | std::ios::binary);
        if (!in) throw std::string("The file '" + path + "' could not opened!\r\n");
        in.seekg(0, in.end);
        const size_t bytes = (size_t)in.tellg();
        in.seekg(0, in.beg);
        if (bytes != n * sizeof(double))
            throw std::string("The file '" + path + "' has different number of data than expected!\r\n");
        std::vector<double> v(n);
        in.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(double)));
        return v;
    }
};

}  // namespace redhost
