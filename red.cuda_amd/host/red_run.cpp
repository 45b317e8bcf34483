// red_run.cpp -- driver loop of red (src/red.cuda/main.cu:403-640) over the
// C++ host classes, on a seeded planetesimal disk, without file output.
//
//   red_run [n] [seed] [method rk4|rkf78] [tol] [steps] [precision fp64|mixed]
//           [hit_centrum] [ejection] [radii_factor]
//
// Prints one JSON line with the step counters, events and energy drift.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "red_host.hpp"

using namespace redhost;

int main(int argc, char** argv)
{
    const uint32_t n = argc > 1 ? (uint32_t)atoi(argv[1]) : 1024;
    const uint32_t seed = argc > 2 ? (uint32_t)atoi(argv[2]) : 1;
    const std::string method = argc > 3 ? argv[3] : "rkf78";
    const double tol = argc > 4 ? atof(argv[4]) : 1e-10;
    const uint64_t max_steps = argc > 5 ? (uint64_t)atoll(argv[5]) : 100;
    const std::string prec = argc > 6 ? argv[6] : "mixed";
    double thr[RD_THRESHOLD_N] = {argc > 7 ? atof(argv[7]) : 0.0, argc > 8 ? atof(argv[8]) : 0.0,
                                  argc > 9 ? atof(argv[9]) : 0.0};
    try {
        rd_disk_spec s;
        std::memset(&s, 0, sizeof(s));
        s.n = n; s.seed = seed; s.body_type = RD_BODY_PROTOPLANET;
        s.a_min = 0.5; s.a_max = 5.0; s.a_power = -0.5; s.sig_e = 0.01; s.sig_i = 0.005;
        s.size_mode = 0; s.lo = 1e-10; s.hi = 1e-8; s.density_cgs = 2.0; s.cd = 0.0;
        s.m_star = 1.0; s.r_star = 1.0 / 215.094;
        std::vector<rd_var4> r(n + 1), v(n + 1);
        std::vector<rd_param> p(n + 1);
        std::vector<rd_body_md> b(n + 1);
        rd_check(rd_make_disk(&s, r.data(), v.data(), p.data(), b.data()), "rd_make_disk");
        uint32_t nbt[RD_BODY_TYPE_N] = {1, 0, 0, n, 0, 0, 0};
        pp_disk ppd(r, v, p, b, nbt, 0.0, thr, nullptr, 0);
        if (prec == "mixed") {
            ppd.set_precision(RD_PREC_MIXED);
            ppd.benchmark();
        }
        const double dt0 = 0.1 * 0.01720209895;   // 0.1 day in tau units
        rungekutta4* rk4 = nullptr;
        rungekutta8* rk8 = nullptr;
        integrator* intgr;
        if (method == "rk4") intgr = rk4 = new rungekutta4(&ppd, dt0, false, tol);
        else intgr = rk8 = new rungekutta8(&ppd, dt0, true, tol);
        rd_integral I0, I1;
        ppd.calc_integral(I0);
        const double output_interval = 1e30;
        uint64_t steps = 0;
        while (steps < max_steps && 1 < ppd.counts().get_n_total_active()) {
            if (0.0 < thr[RD_THRESHOLD_EJECTION_DISTANCE] || 0.0 < thr[RD_THRESHOLD_HIT_CENTRUM_DISTANCE]) {
                if (ppd.check_for_ejection_hit_centrum()) {
                    ppd.handle_ejection_hit_centrum();
                    ppd.clear_event_counter();
                }
            }
            const double dt = intgr->step();
            ++steps;
            if (std::fabs(output_interval) < std::fabs(intgr->get_dt_next()))
                intgr->set_dt_next(0.0 < dt ? output_interval : -output_interval);
            if (0.0 < thr[RD_THRESHOLD_RADII_ENHANCE_FACTOR]) {
                if (ppd.check_for_collision()) {
                    ppd.handle_collision();
                    ppd.clear_event_counter();
                }
            }
        }
        ppd.calc_integral(I1);
        std::printf("{\"n\": %u, \"method\": \"%s\", \"precision\": \"%s\", \"steps\": %llu, \"tried\": %llu, "
                    "\"failed\": %llu, \"t\": %.17g, \"events\": %zu, \"rel_energy_error\": %.6g}\n",
                    n, method.c_str(), prec.c_str(), (unsigned long long)steps,
                    (unsigned long long)intgr->get_n_tried_step(), (unsigned long long)intgr->get_n_failed_step(),
                    ppd.t(), ppd.event_log().size(), std::fabs((I1.E - I0.E) / I0.E));
        delete rk4;
        delete rk8;
    } catch (const std::string& msg) {
        std::fprintf(stderr, "Error: %s\n", msg.c_str());
        return 1;
    }
    return 0;
}
