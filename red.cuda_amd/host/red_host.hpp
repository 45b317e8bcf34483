// red_host.hpp -- C++ classes with the reference's pp_disk / integrator
// interface, implemented over the C ABI of librdamd.so (include/red_abi.h).
//
// Method names, argument meaning and error behaviour follow the reference:
// every failing ABI call throws std::string, as the reference's
// CUDA_SAFE_CALL / CUDA_CHECK_ERROR do (red.type/red_macro.h:37-95).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/red_abi.h"

namespace redhost {

inline void rd_check(int rc, const char* what)
{
    if (rc != RD_OK) throw std::string(what) + ": " + rd_last_error();
}

// n_objects_t (red.type/red_type.h:522-712): counts and interaction bounds.
struct n_objects {
    uint32_t playing[RD_BODY_TYPE_N] = {0};
    uint32_t inactive[RD_BODY_TYPE_N] = {0};
    uint32_t removed[RD_BODY_TYPE_N] = {0};
    uint32_t get_n_SI() const { return playing[0] + playing[1] + playing[2] + playing[3]; }
    uint32_t get_n_NSI() const { return playing[4] + playing[5]; }
    uint32_t get_n_NI() const { return playing[6]; }
    uint32_t get_n_massive() const { return get_n_SI() + get_n_NSI(); }
    uint32_t get_n_total_playing() const { return get_n_massive() + get_n_NI(); }
    uint32_t get_n_total_active() const
    {
        uint32_t n = 0;
        for (int t = 0; t < RD_BODY_TYPE_N; ++t) n += playing[t] - inactive[t];
        return n;
    }
};

// pp_disk (src/red.cuda/pp_disk.h:10-210) for one disk on one GPU.
class pp_disk {
  public:
    pp_disk(const std::vector<rd_var4>& r, const std::vector<rd_var4>& v, const std::vector<rd_param>& p,
            const std::vector<rd_body_md>& bmd, const uint32_t n_by_type[RD_BODY_TYPE_N], double t0,
            const double threshold[RD_THRESHOLD_N], const rd_gas_params* gas, int id_dev)
    {
        // the structs of this header must match the library's
        if (rd_version() != RD_ABI_VERSION)
            throw std::string("librdamd.so ABI version ") + std::to_string(rd_version()) + ", header " +
                std::to_string(RD_ABI_VERSION) + ": rebuild";
        rd_check(rd_disk_create((uint32_t)r.size(), n_by_type, r.data(), v.data(), p.data(), bmd.data(), t0, threshold,
                                gas, id_dev, &h_),
                 "pp_disk");
    }
    ~pp_disk() { rd_disk_destroy(h_); }
    pp_disk(const pp_disk&) = delete;
    pp_disk& operator=(const pp_disk&) = delete;

    rd_disk* handle() { return h_; }
    void set_precision(int prec) { rd_check(rd_disk_set_precision(h_, prec), "set_precision"); }
    void set_drag_mode(int mode) { rd_check(rd_disk_set_drag_mode(h_, mode), "set_drag_mode"); }
    // GAS_DISK_MODEL_FARGO: drag through a FARGO grid (fargo_gas_disk.hpp) in
    // place of the analytic disk; the grid is copied to the device
    void set_gas_grid(const rd_gas_grid& g) { rd_check(rd_disk_set_gas_grid(h_, &g), "set_gas_grid"); }
    // pp_disk::benchmark (pp_disk.cu:553-611): picks the launch shape by timing
    rd_tuning benchmark()
    {
        rd_tuning t;
        rd_check(rd_autotune(h_, &t), "benchmark");
        return t;
    }

    n_objects counts() const
    {
        n_objects n;
        rd_check(rd_disk_n_objects(h_, 0, n.playing, n.inactive, n.removed), "n_objects");
        return n;
    }
    double t() const
    {
        double tt;
        rd_check(rd_get_state(h_, 0, &tt, nullptr, nullptr, nullptr, nullptr), "t");
        return tt;
    }
    // pp_disk::copy_to_host (pp_disk.cu:1686-1699)
    void copy_to_host(std::vector<rd_var4>& r, std::vector<rd_var4>& v, std::vector<rd_param>& p,
                      std::vector<rd_body_md>& bmd, double& t) const
    {
        uint32_t n;
        rd_check(rd_disk_n_total(h_, 0, &n), "n_total");
        r.resize(n); v.resize(n); p.resize(n); bmd.resize(n);
        rd_check(rd_get_state(h_, 0, &t, r.data(), v.data(), p.data(), bmd.data()), "copy_to_host");
    }
    // pp_disk::calc_dydx (pp_disk.cu:1450-1502) on host arrays
    void calc_dydx(uint32_t i, uint32_t rr, double curr_t, const rd_var4* r, const rd_var4* v, rd_var4* dy)
    {
        rd_check(rd_calc_dydx(h_, 0, i, rr, curr_t, r, v, dy), "calc_dydx");
    }
    // ... on device arrays of the disk's device (no PCIe copies)
    void calc_dydx_device(uint32_t i, uint32_t rr, double curr_t, const rd_var4* r_dev, const rd_var4* v_dev,
                          rd_var4* dy_dev)
    {
        rd_check(rd_calc_dydx_device(h_, 0, i, rr, curr_t, r_dev, v_dev, dy_dev), "calc_dydx_device");
    }
    // event hooks (pp_disk.h:83-102)
    bool check_for_ejection_hit_centrum()
    {
        uint32_t n;
        rd_check(rd_check_ejection_hit_centrum(h_, 0, &n), "check_for_ejection_hit_centrum");
        return n > 0;
    }
    bool check_for_collision()
    {
        uint32_t n;
        rd_check(rd_check_collision(h_, 0, &n), "check_for_collision");
        return n > 0;
    }
    void handle_collision() { rd_check(rd_handle_collision(h_, 0), "handle_collision"); }
    void handle_ejection_hit_centrum() { rd_check(rd_handle_ejection_hit_centrum(h_, 0), "handle_ejection_hit_centrum"); }
    void clear_event_counter() { rd_check(rd_clear_event_counter(h_, 0), "clear_event_counter"); }
    // n_event / n_collision / n_ejection / n_hit_centrum (pp_disk.h:164-167), each indexed by
    // RD_EVENT_COUNTER_TOTAL / _LAST_CLEAR / _LAST_STEP
    struct event_counters {
        uint32_t n_event[3], n_collision[3], n_ejection[3], n_hit_centrum[3];
    };
    event_counters counters_of_events() const
    {
        event_counters c;
        rd_check(rd_event_counters(h_, 0, c.n_event, c.n_collision, c.n_ejection, c.n_hit_centrum),
                 "event_counters");
        return c;
    }
    // pp_disk::set_event_counter (pp_disk.cu:1442-1448)
    void set_event_counter(int field, uint32_t value)
    {
        rd_check(rd_set_event_counter(h_, 0, field, value), "set_event_counter");
    }
    void rebuild_vectors() { rd_check(rd_rebuild_vectors(h_, 0), "rebuild_vectors"); }
    std::vector<rd_event> event_log() const
    {
        uint32_t n = 0;
        rd_check(rd_event_log(h_, 0, nullptr, 0, &n), "event_log");
        std::vector<rd_event> out(n);
        rd_check(rd_event_log(h_, 0, out.data(), n, &n), "event_log");
        return out;
    }
    // pp_disk::calc_integral (pp_disk.cu:1504-1516)
    void calc_integral(rd_integral& I) { rd_check(rd_calc_integral(h_, 0, &I), "calc_integral"); }

  private:
    rd_disk* h_ = nullptr;
};

// The learned step proposer (no reference counterpart): a 12 -> 64 -> 64 -> 1
// BF16 MLP trained on the GPU from the engine's own logged steps.
class mlp {
  public:
    explicit mlp(int id_dev = 0, uint32_t seed = 0) { rd_check(rd_mlp_create(id_dev, seed, &h_), "mlp"); }
    ~mlp() { rd_mlp_destroy(h_); }
    mlp(const mlp&) = delete;
    mlp& operator=(const mlp&) = delete;

    rd_mlp* handle() { return h_; }
    // standardisation of the 12 features and of the target log h
    void set_norm(const float mean[RD_MLP_IN], const float stdev[RD_MLP_IN], float t_mean, float t_std)
    {
        rd_check(rd_mlp_set_norm(h_, mean, stdev, t_mean, t_std), "mlp::set_norm");
    }
    // targets as offsets from feature base_feature (-1: none); quantile 0 fits the mean
    void set_target(int base_feature, float quantile)
    {
        rd_check(rd_mlp_set_target(h_, base_feature, quantile), "mlp::set_target");
    }
    // fits the samples (x: n rows of RD_MLP_IN features, t: n targets); returns the last epoch's loss
    float train(const std::vector<float>& x, const std::vector<float>& t, uint32_t epochs, uint32_t batch, float lr)
    {
        if (x.size() != t.size() * RD_MLP_IN) throw std::string("mlp::train: x must hold 12 features per target");
        float loss = 0.0f;
        rd_check(rd_mlp_train(h_, x.data(), t.data(), (uint32_t)t.size(), epochs, batch, lr, &loss), "mlp::train");
        return loss;
    }
    std::vector<float> params()
    {
        std::vector<float> w(RD_MLP_NPARAM);
        rd_check(rd_mlp_get_params(h_, w.data(), nullptr), "mlp::params");
        return w;
    }
    void set_params(const std::vector<float>& w)
    {
        if (w.size() != RD_MLP_NPARAM) throw std::string("mlp::set_params: wrong parameter count");
        rd_check(rd_mlp_set_params(h_, w.data()), "mlp::set_params");
    }

  private:
    rd_mlp* h_ = nullptr;
};

// integrator (src/red.cuda/integrator.h:10-84)
class integrator {
  public:
    integrator(pp_disk* ppd, int method, double dt, bool adaptive, double tolerance) : ppd_(ppd)
    {
        rd_check(rd_integrator_create(ppd->handle(), method, dt, adaptive ? 1 : 0, tolerance, &h_), "integrator");
    }
    virtual ~integrator() { rd_integrator_destroy(h_); }
    integrator(const integrator&) = delete;
    integrator& operator=(const integrator&) = delete;

    double step()
    {
        double dt;
        rd_check(rd_step(h_, &dt), "step");
        return dt;
    }
    double get_dt_next() const
    {
        double dt;
        rd_check(rd_get_dt_next(h_, 0, &dt), "get_dt_next");
        return dt;
    }
    void set_dt_next(double dt) { rd_check(rd_set_dt_next(h_, 0, dt), "set_dt_next"); }
    uint64_t get_n_tried_step() const { return counter(0); }
    uint64_t get_n_passed_step() const { return counter(1); }
    uint64_t get_n_failed_step() const { return counter(2); }
    void set_error_check_for_tp(bool on) { rd_check(rd_set_error_check_for_tp(h_, on ? 1 : 0), "error_check_for_tp"); }
    // opt-in bound on the retry loop (red_v2 int_rungekutta7.cu:176-185); 0 turns a bound off
    void set_guard(uint32_t max_attempts, double min_dt, bool nonfinite)
    {
        rd_check(rd_integrator_set_guard(h_, max_attempts, min_dt, nonfinite ? 1 : 0), "set_guard");
    }
    // rd_steps overlaps the host's step-size control with the device (on by default)
    void set_overlap(bool on) { rd_check(rd_integrator_set_overlap(h_, on ? 1 : 0), "set_overlap"); }
    rd_integrator* handle() { return h_; }

    // learned step proposer: first try of each step from m (nullptr: reference controller only)
    void set_proposer(mlp* m, float safety = 1.0f)
    {
        rd_check(rd_integrator_set_proposer(h_, m ? m->handle() : nullptr, safety), "set_proposer");
    }
    // records (features -> target log h) of every accepted step
    void set_logging(bool on) { rd_check(rd_integrator_set_logging(h_, on ? 1 : 0), "set_logging"); }
    void log(std::vector<float>& x, std::vector<float>& t)
    {
        uint32_t n = 0;
        rd_check(rd_integrator_log(h_, nullptr, nullptr, 0, &n), "log");
        x.resize((size_t)n * RD_MLP_IN);
        t.resize(n);
        rd_check(rd_integrator_log(h_, x.data(), t.data(), n, &n), "log");
    }

  protected:
    uint64_t counter(int k) const
    {
        uint64_t c[3];
        rd_check(rd_counters(h_, 0, &c[0], &c[1], &c[2]), "counters");
        return c[k];
    }
    pp_disk* ppd_;
    rd_integrator* h_ = nullptr;
};

// euler (src/red.cuda/int_euler.h): fixed step, first order
class euler : public integrator {
  public:
    euler(pp_disk* ppd, double dt) : integrator(ppd, RD_EULER, dt, false, 0.0) {}
};

// rungekutta2, the midpoint rule (src/red.cuda/int_rungekutta2.h): fixed step
class rungekutta2 : public integrator {
  public:
    rungekutta2(pp_disk* ppd, double dt) : integrator(ppd, RD_RK2, dt, false, 0.0) {}
};

// rungekutta4 (src/red.cuda/int_rungekutta4.h:7-28)
class rungekutta4 : public integrator {
  public:
    rungekutta4(pp_disk* ppd, double dt, bool adaptive, double tolerance)
        : integrator(ppd, RD_RK4, dt, adaptive, tolerance) {}
};

// rungekutta8, Runge-Kutta-Fehlberg 7(8) (src/red.cuda/int_rungekutta8.h:10-47)
class rungekutta8 : public integrator {
  public:
    rungekutta8(pp_disk* ppd, double dt, bool adaptive, double tolerance)
        : integrator(ppd, RD_RKF78, dt, adaptive, tolerance) {}
};

}  // namespace redhost
