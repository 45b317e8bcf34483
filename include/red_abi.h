/*
 * red_abi.h -- C ABI of the MI355X-native direct-sum N-body engine.
 *
 * This is the drop-in boundary for red.cuda's GPU path: the pp_disk body
 * arrays, the pp_disk::calc_dydx operator, the RK4 / RKF7(8) integrators and
 * the collision / ejection / hit-centrum event hooks.  Plain C types only;
 * no exceptions cross it; every call returns an RD_* status and
 * rd_last_error() describes the last failure of the calling thread.
 *
 * Reference citations are relative to the red.cuda repository root.
 *
 * Threading: one HIP stream per rd_disk; calls on one handle are not
 * thread-safe, distinct handles may be used from distinct threads.
 */
#ifndef RED_ABI_H
#define RED_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 2: rd_tuning gained sym_chunk_min; 3: rd_gas_grid, rd_disk_set_gas_grid, rd_integrator_set_overlap,
 * rd_mlp_set_target, rd_calc_dydx_device */
#define RD_ABI_VERSION 3

/* Status codes.  The reference throws std::string (red.type/red_macro.h:37-95,
 * src/red.cuda/pp_disk.cu:688, 974-982); here each throw site maps to a code. */
enum {
    RD_OK = 0,
    RD_EINVAL = 1,     /* bad argument or inconsistent body arrays           */
    RD_EDEVICE = 2,    /* HIP runtime error                                  */
    RD_ENOTIMPL = 3,   /* the reference also refuses (e.g. super-planetesimal merge, pp_disk.cu:1333-1337) */
    RD_EOVERFLOW = 4,  /* event buffer full (the reference does not guard, pp_disk.cu:1623) */
    RD_ESTATE = 5,     /* call not valid in the current state                */
    RD_ENOMEM = 6      /* device or host allocation failed                   */
};

/* body_type_t, red.type/red_type.h:169-179 (arrays are sorted by it) */
enum {
    RD_BODY_STAR = 0, RD_BODY_GIANTPLANET, RD_BODY_ROCKYPLANET, RD_BODY_PROTOPLANET,
    RD_BODY_SUPERPLANETESIMAL, RD_BODY_PLANETESIMAL, RD_BODY_TESTPARTICLE, RD_BODY_TYPE_N
};
/* threshold_t, red.type/red_type.h:128-134 */
enum { RD_THRESHOLD_HIT_CENTRUM_DISTANCE = 0, RD_THRESHOLD_EJECTION_DISTANCE = 1,
       RD_THRESHOLD_RADII_ENHANCE_FACTOR = 2, RD_THRESHOLD_N = 3 };
/* event_name_t, red.type/red_type.h:145-152 */
enum { RD_EVENT_NONE = 0, RD_EVENT_HIT_CENTRUM = 1, RD_EVENT_EJECTION = 2, RD_EVENT_COLLISION = 3 };
/* gas_decrease_t, red.type/red_type.h:105-111 */
enum { RD_GAS_DENSITY_CONSTANT = 0, RD_GAS_DENSITY_DECREASE_LINEAR = 1, RD_GAS_DENSITY_DECREASE_EXPONENTIAL = 2 };
/* integrator_type_t, red.type/red_type.h:136-143 (all but RKF5(6) are built).
 * Euler and RK2 are fixed-step: their step is always the dt they were created
 * with, and rd_set_dt_next does not change it (int_euler.cu:78-105,
 * int_rungekutta2.cu:99-131). */
enum { RD_EULER = 0, RD_RK2 = 1, RD_RK4 = 2, RD_RKF78 = 4 };

/* Force precision of the pair sum. */
enum {
    RD_PREC_FP64 = 0,   /* reference formula m_j/(d*d^2) in FP64, j ascending per sink:
                           bit-identical to kernel_pp_disk::calc_grav_accel_naive
                           (src/red.cuda/pp_disk.cu:374-416)                      */
    RD_PREC_MIXED = 1   /* star (and giant-planet) sources in FP64; all other pairs use
                           FP64 offsets rounded to FP32, FP32 rsqrt and FP32 tile sums
                           accumulated in FP64                                      */
};
/* Which bodies feel the gas.  The reference drags the NSI block only
 * (n_objects_t::get_bound_GD, red.type/red_type.h:694-700). */
enum { RD_DRAG_NSI = 0, RD_DRAG_ALL_MASSIVE = 1 };

/* ---- body arrays, byte-compatible with the reference ------------------- */
typedef struct rd_var4 { double x, y, z, w; } rd_var4;                 /* var4_t red_type.h:206-212 */
typedef struct rd_param { double mass, radius, density, cd; } rd_param; /* pp_disk_t::param_t red_type.h:347-353 */
typedef struct rd_body_md {                                             /* pp_disk_t::body_metadata_t red_type.h:356-362 */
    int32_t id;          /* > 0; negated when the body becomes inactive */
    int32_t body_type;
    int32_t mig_type;
    double  mig_stop_at;
} rd_body_md;
typedef struct rd_integral { rd_var4 R, V, C; double E; } rd_integral;  /* pp_disk_t::integral_t red_type.h:375-381 */
typedef struct rd_event {                                               /* pp_disk_t::event_data_t red_type.h:415-461 */
    int32_t event_name;
    double  t, d;
    int32_t id0; uint32_t idx0; rd_param p0; rd_var4 r0, v0;
    int32_t id1; uint32_t idx1; rd_param p1; rd_var4 r1, v1;
    int32_t id2; uint32_t idx2; rd_param p2; rd_var4 r2, v2;
    rd_param ps; rd_var4 rs, vs;
} rd_event;
typedef struct rd_gas_params {                                          /* analytic_gas_disk_params_t red_type.h:464-485 */
    double rho[2], sch[2], eta[2], tau[2], mfp[2], temp[2];
    double c_vth;
    int32_t gas_decrease;
    double t0, t1, e_folding_time;
    double alpha, mean_molecular_weight, particle_diameter;
} rd_gas_params;
/* Frame 0 of a FARGO gas run (fargo_gas_disk.h:8-60; the frame the reference
 * loads, fargo_gas_disk.cu:190-205): n_rad rings x n_sec sectors, the cell of
 * ring k and sector s at k * n_sec + s.  used_rad holds the n_rad + 1 ring
 * interface radii [AU], increasing (FARGO's used_rad.dat).  Values are taken in
 * the solver's units -- density [Msun/AU^3], velocities [AU/tau] -- since the
 * reference's unit transforms are not implemented (fargo_gas_disk.cu:252-274);
 * vtheta is the gas's azimuthal velocity on top of the circular one. */
typedef struct rd_gas_grid {
    uint32_t n_rad, n_sec;
    const double* used_rad;   /* n_rad + 1 */
    const double* density;    /* n_rad * n_sec */
    const double* vrad;       /* n_rad * n_sec */
    const double* vtheta;     /* n_rad * n_sec */
} rd_gas_grid;

/* Launch shapes of the pair kernels (chosen by rd_autotune or by hand). */
typedef struct rd_tuning {
    uint32_t sinks_per_lane;   /* mixed kernel: 1, 2 or 4 sinks per lane          */
    uint32_t chunk;            /* mixed kernel: sources per work item (mult. of 256) */
    uint32_t blocks_per_cu;    /* persistent grid size = blocks_per_cu * CUs      */
    uint32_t use_graph;        /* capture one RK attempt in a hipGraph            */
    uint32_t pair_mode;        /* mixed kernel, SI x SI pairs: RD_PAIRS_*         */
    uint32_t sym_chunk_min;    /* symmetric kernel: 256-body blocks per chunk at  */
                               /* least (1..8; 0 reads as 1).  Fewer partial-sum  */
                               /* slots per body, fewer and larger work items.    */
} rd_tuning;

/* Mixed precision: RD_PAIRS_SYMMETRIC evaluates each SI x SI pair once and
 * applies it to both bodies (the reference CPU path's i < j scheme,
 * pp_disk.cu:822-865); RD_PAIRS_ONE_SIDED evaluates every ordered pair (the
 * reference GPU kernel's scheme, pp_disk.cu:374-416). */
enum { RD_PAIRS_ONE_SIDED = 0, RD_PAIRS_SYMMETRIC = 1 };

typedef struct rd_disk rd_disk;
typedef struct rd_integrator rd_integrator;

/* ---- library ------------------------------------------------------------ */
const char* rd_last_error(void);
int rd_version(void);
int rd_device_count(int* n);

/* ---- pp_disk: body arrays, state, options ------------------------------- */
/* pp_disk(n_objects_t*, gas_disk_model_t, id_dev, comp_dev) + data load
 * (src/red.cuda/pp_disk.h:14-15, pp_disk.cu:1534-1575).  n_by_type are the
 * n_objects_t counts (red_type.h:524-541); bodies must be sorted by
 * body_type with those counts.  gas may be NULL (GAS_DISK_MODEL_NONE).  The
 * thresholds are the dc_threshold values (pp_disk.cu:24, 1681). */
int rd_disk_create(uint32_t n_total, const uint32_t n_by_type[RD_BODY_TYPE_N],
                   const rd_var4* r, const rd_var4* v, const rd_param* p, const rd_body_md* bmd,
                   double t, const double threshold[RD_THRESHOLD_N], const rd_gas_params* gas,
                   int device, rd_disk** out);
/* Several independent disks in one engine (no reference counterpart: the
 * reference runs one disk per process).  Arrays are the per-disk arrays
 * concatenated; n_by_type holds n_disk*RD_BODY_TYPE_N counts; t may be NULL. */
int rd_batch_create(uint32_t n_disk, const uint32_t* n_by_type,
                    const rd_var4* r, const rd_var4* v, const rd_param* p, const rd_body_md* bmd,
                    const double* t, const double threshold[RD_THRESHOLD_N], const rd_gas_params* gas,
                    int device, rd_disk** out);
/* ~pp_disk (pp_disk.cu:1577-1587) */
int rd_disk_destroy(rd_disk* d);
int rd_disk_set_precision(rd_disk* d, int precision);
int rd_disk_set_drag_mode(rd_disk* d, int mode);
/* GAS_DISK_MODEL_FARGO (options.cpp:149-153, 232-234; drag pp_disk.cu:778-814):
 * copy the grid to the device and drag the gas-feeling bodies through it
 * (every disk of a batch, the same grid), in place of the analytic disk.
 * NULL removes the grid; the disk returns to the gas model it was created
 * with.  The arrays are copied, so the caller may free them on return. */
int rd_disk_set_gas_grid(rd_disk* d, const rd_gas_grid* grid);
int rd_disk_set_tuning(rd_disk* d, const rd_tuning* tuning);
int rd_disk_get_tuning(rd_disk* d, rd_tuning* tuning);
/* pp_disk::benchmark picks n_tpb by timing (pp_disk.cu:553-632); this times
 * the mixed-precision pair kernel over its tuning space and keeps the best. */
int rd_autotune(rd_disk* d, rd_tuning* best);
int rd_disk_count(rd_disk* d, uint32_t* n_disk);
/* playing body count of disk k (n_objects_t::get_n_total_playing) */
int rd_disk_n_total(rd_disk* d, uint32_t k, uint32_t* n);
/* n_objects_t playing / inactive / removed (red_type.h:702-705) */
int rd_disk_n_objects(rd_disk* d, uint32_t k, uint32_t playing[RD_BODY_TYPE_N],
                      uint32_t inactive[RD_BODY_TYPE_N], uint32_t removed[RD_BODY_TYPE_N]);
/* pp_disk::copy_to_host + sim_data h_y / h_p / h_body_md (pp_disk.cu:1686-1699) and pp_disk::t */
int rd_get_state(rd_disk* d, uint32_t k, double* t, rd_var4* r, rd_var4* v, rd_param* p, rd_body_md* bmd);
/* pp_disk::copy_to_device (pp_disk.cu:1667-1684) of the positions and velocities of disk k */
int rd_set_state(rd_disk* d, uint32_t k, double t, const rd_var4* r, const rd_var4* v);
int rd_sync(rd_disk* d);

/* ---- the operator -------------------------------------------------------- */
/* pp_disk::calc_dydx(i, rr, t, r, v, dy) (pp_disk.h:114, pp_disk.cu:1450-1502)
 * on host arrays of disk k: i = 0 gives dy = v, i = 1 gravity (+ drag).
 * Copies in and out over PCIe; the integrators never go through it. */
int rd_calc_dydx(rd_disk* d, uint32_t k, uint32_t i, uint32_t rr, double t,
                 const rd_var4* r, const rd_var4* v, rd_var4* dy);
/* The same on device arrays (the reference's calc_dydx works on the GPU copies
 * it owns, pp_disk.cu:1450-1502): r, v and dy point to device memory of the
 * disk's device, each holding disk k's bodies as rd_var4 records (w of dy set
 * to 0).  Pointers that are not device memory of that device, or whose
 * allocation ends before the last record, are refused (RD_EINVAL) before any
 * kernel runs.  Returns when dy is written. */
int rd_calc_dydx_device(rd_disk* d, uint32_t k, uint32_t i, uint32_t rr, double t,
                        const rd_var4* r, const rd_var4* v, rd_var4* dy);
/* Time n_iter launches of the pair kernel on the current state of all disks
 * (pp_disk::benchmark_calc_grav_accel, pp_disk.cu:613-632).  Returns the mean
 * kernel time in ms (HIP events on the engine's stream) and the ordered
 * sink-source pair count of one launch. */
int rd_bench_force(rd_disk* d, uint32_t n_iter, double* ms_per_launch, double* pairs_per_launch);

/* Measurement helper without a reference counterpart: time n_iter launches
 * of a plain device copy of `bytes` bytes (read + write), used to scale
 * hardware byte counters. */
int rd_bench_copy(int device, size_t bytes, uint32_t n_iter, double* ms_per_launch);
/* Measurement helper without a reference counterpart: mean time of one pair
 * kernel launch plus the stage kernel that folds its partial sums (what
 * rd_autotune minimises). */
int rd_bench_force_stage(rd_disk* d, uint32_t n_iter, double* ms);
/* Measurement helper without a reference counterpart: mean time of the stage
 * kernels of one adaptive attempt of `method` (RD_RKF78: k_comb1 and the 13
 * stage kernels; RD_RK4: k_comb1 and 5), with dt = 0 so the state is kept,
 * folding the partial sums of one pair-kernel launch made first.  The HIP
 * events bracket the launches on the engine's stream. */
int rd_bench_stages(rd_disk* d, int method, uint32_t n_iter, double* ms);

/* ---- integrators --------------------------------------------------------- */
/* rungekutta4 / rungekutta8 constructors (int_rungekutta4.h:15, int_rungekutta8.h:18) */
int rd_integrator_create(rd_disk* d, int method, double dt, int adaptive, double tol, rd_integrator** out);
int rd_integrator_destroy(rd_integrator* it);
/* integrator::step (integrator.h:31): one accepted step of every disk;
 * dt_did of disk 0 is returned. */
int rd_step(rd_integrator* it, double* dt_did);
/* n accepted steps per disk without host event handling in between */
int rd_steps(rd_integrator* it, uint32_t n_steps);
/* integrator::get_dt_next / set_dt_next (integrator.h:23-24) for disk k */
int rd_get_dt_next(rd_integrator* it, uint32_t k, double* dt);
int rd_set_dt_next(rd_integrator* it, uint32_t k, double dt);
/* integrator n_tried/n_passed/n_failed (integrator.h:26-28) for disk k */
int rd_counters(rd_integrator* it, uint32_t k, uint64_t* tried, uint64_t* passed, uint64_t* failed);
/* integrator::error_check_for_tp (integrator.h:32) */
int rd_set_error_check_for_tp(rd_integrator* it, int on);
/* Opt-in bound on the adaptive retry loop.  red.cuda's loop runs until
 * err <= tol (int_rungekutta8.cu:645-709); src/red_v2/int_rungekutta7.cu:
 * 176-185 bounds it by an iteration count and a minimum step.  A step fails
 * when it is still rejected after max_attempts attempts, or when an attempt
 * would use |dt| <= min_dt (the attempt is not made; this also stops a run
 * whose step has collapsed to 0, which the reference accepts over and over:
 * zero step, zero error), or (nonfinite != 0) when an error term is NaN or
 * infinite -- under the reference's max (integrator.cu:248-263) a NaN term is
 * never the maximum, so a NaN state is accepted with its step doubled.  0
 * turns each bound off; all off is the default.  A failing disk keeps y_n and
 * counts the attempts it made as failed; the other disks of a batch finish
 * their step, then rd_step / rd_steps / rd_run return RD_ESTATE naming the
 * first failing disk. */
int rd_integrator_set_guard(rd_integrator* it, uint32_t max_attempts, double min_dt, int nonfinite);
/* No reference counterpart (an execution option, results unchanged): in
 * rd_steps, after each adaptive attempt the device makes the accept decision
 * with the host's arithmetic (|lambda h max| > tol rejects,
 * int_rungekutta8.cu:693-708), commits accepted disks and starts the next
 * step's stage-0 pair pass while the host reads the error and computes the
 * next step size, which it still does with its own pow.  On by default; 0
 * restores the host-only sequence.  rd_step and rd_run, which may handle
 * events between steps, never overlap. */
int rd_integrator_set_overlap(rd_integrator* it, int on);
/* Number of pair-kernel launches and the ordered pairs they summed so far. */
int rd_work_counters(rd_integrator* it, uint64_t* force_evals, double* pairs);

/* ---- driver loop ---------------------------------------------------------- */
/* run_simulation's loop (src/red.cuda/main.cu:470-625): ejection/hit-centrum
 * check and handling, step, dt clamp to output_interval, collision check and
 * handling, rebuild after 16 events; until |elapsed| > length or at most
 * max_steps steps per disk.  No files are written. */
int rd_run(rd_integrator* it, double length, double output_interval, uint64_t max_steps);

/* ---- events ---------------------------------------------------------------- */
/* pp_disk::check_for_ejection_hit_centrum (pp_disk.cu:1062-1076, 284-318) and
 * check_for_collision (1078-1092, 321-371) on disk k; n = events found.  The
 * event list is in canonical order: (idx, kind) for ejection / hit-centrum
 * and (lower index, higher index) for collisions -- the reference's order is
 * that of atomicAdd (pp_disk.cu:309, 357). */
int rd_check_ejection_hit_centrum(rd_disk* d, uint32_t k, uint32_t* n);
int rd_check_collision(rd_disk* d, uint32_t k, uint32_t* n);
/* copy of the pending events of disk k (pp_disk::copy_event_data_to_host, pp_disk.cu:1710-1713) */
int rd_get_events(rd_disk* d, uint32_t k, rd_event* out, uint32_t cap, uint32_t* n);
/* pp_disk::handle_collision (pp_disk.cu:1220-1248) and
 * handle_ejection_hit_centrum (1290-1326) on disk k's pending events */
int rd_handle_collision(rd_disk* d, uint32_t k);
int rd_handle_ejection_hit_centrum(rd_disk* d, uint32_t k);
/* pp_disk::clear_event_counter (pp_disk.cu:1715-1722) */
int rd_clear_event_counter(rd_disk* d, uint32_t k);
/* all events handled so far on disk k, in handling order (what print_event_data writes) */
int rd_event_log(rd_disk* d, uint32_t k, rd_event* out, uint32_t cap, uint32_t* n);
/* event counters of disk k, each indexed by RD_EVENT_COUNTER_* (pp_disk.h:164-167:
 * n_event, n_collision, n_ejection, n_hit_centrum); any pointer may be NULL */
enum { RD_EVENT_COUNTER_TOTAL = 0, RD_EVENT_COUNTER_LAST_CLEAR = 1, RD_EVENT_COUNTER_LAST_STEP = 2 };
int rd_event_counters(rd_disk* d, uint32_t k, uint32_t n_event[3], uint32_t n_collision[3], uint32_t n_ejection[3],
                      uint32_t n_hit_centrum[3]);
/* pp_disk::set_event_counter (pp_disk.cu:1442-1448): field of every counter set to value */
int rd_set_event_counter(rd_disk* d, uint32_t k, int field, uint32_t value);
/* pp_disk::rebuild_vectors (pp_disk.cu:1370-1430) on disk k */
int rd_rebuild_vectors(rd_disk* d, uint32_t k);

/* ---- integrals -------------------------------------------------------------- */
/* pp_disk::calc_integral (pp_disk.cu:1504-1516): R, V, C and E of disk k,
 * computed on the GPU in FP64 with compensated sums. */
int rd_calc_integral(rd_disk* d, uint32_t k, rd_integral* out);
/* No counterpart in the reference: rd_calc_integral for callers that cannot
 * allocate device memory (code that must not call hipMalloc, such as tight
 * serving or monitoring loops and harnesses that forbid it).  Same kernels and
 * results, bit for bit; the partial sums live in a file-scope device array of
 * 2^21 doubles (16 MiB per GPU, held until the process ends) instead of a
 * buffer the disk allocates on first use, so it takes disks of at most 65536
 * bodies (RD_EINVAL above that).  Calls that overlap on one GPU share that
 * array and race. */
int rd_calc_integral_noalloc(rd_disk* d, uint32_t k, rd_integral* out);

/* ---- learned step-size proposer (no reference counterpart) ------------------- */
/* A 12 -> 64 -> 64 -> 1 MLP (tanh hidden layers, BF16 weights, FP32
 * accumulation) that proposes log h for the next RKF7(8) step from per-disk
 * features; the Fehlberg estimate still accepts or rejects every attempt and
 * a rejection falls back to the reference controller
 * (int_rungekutta8.cu:693-708).  Features, in order:
 *   0 log h of the last accepted step      6 log h of the step before
 *   1 log(err/tol) of that step            7 log(err/tol) of the step before
 *   2 log min_{i<j} |r_ij|/|v_ij|          8 rejected attempts in the last step
 *   3 log min periapsis about the star     9 log h proposed by the reference controller
 *   4 log N                               10 log tol
 *   5 log10 relative drift of the Keplerian energy since the start
 *                                         11 change of feature 2 since the last step */
#define RD_MLP_IN 12
#define RD_MLP_HIDDEN 64
#define RD_MLP_NPARAM (RD_MLP_IN * RD_MLP_HIDDEN + RD_MLP_HIDDEN + RD_MLP_HIDDEN * RD_MLP_HIDDEN + RD_MLP_HIDDEN + RD_MLP_HIDDEN + 1)
typedef struct rd_mlp rd_mlp;
int rd_mlp_create(int device, uint32_t seed, rd_mlp** out);
int rd_mlp_destroy(rd_mlp* m);
/* feature / target standardisation applied inside the network */
int rd_mlp_set_norm(rd_mlp* m, const float mean[RD_MLP_IN], const float stdev[RD_MLP_IN], float t_mean, float t_std);
/* What the network is trained on (default: base -1, quantile 0).  The output
 * y' becomes y' t_std + t_mean + x[base_feature] (no offset for -1), so a
 * network whose training targets are offsets from a feature -- feature 9, the
 * reference controller's proposal -- still returns log h.  quantile 0 fits
 * the mean (loss 0.5 mean e^2 in standardised units); q in (0, 1) fits the
 * q-quantile (pinball loss max(q u, (q - 1) u), u = t' - y'), so about a
 * fraction q of the targets fall below the output. */
int rd_mlp_set_target(rd_mlp* m, int base_feature, float quantile);
/* FP32 master parameters (layout W1[12][64] b1[64] W2[64][64] b2[64] w3[64] b3) */
int rd_mlp_set_params(rd_mlp* m, const float* params);
int rd_mlp_get_params(rd_mlp* m, float* params, uint16_t* params_bf16);
/* y = log h for n raw feature rows */
int rd_mlp_forward(rd_mlp* m, const float* x, uint32_t n, float* y);
/* the loss set by rd_mlp_set_target (default 0.5 mean((y - t)/t_std)^2) and
 * its gradient w.r.t. the FP32 parameters */
int rd_mlp_loss_grad(rd_mlp* m, const float* x, const float* t, uint32_t n, float* loss, float* grad);
/* Adam on the GPU over shuffled minibatches */
int rd_mlp_train(rd_mlp* m, const float* x, const float* t, uint32_t n, uint32_t epochs, uint32_t batch, float lr,
                 float* final_loss);
/* use m to propose the first attempt of each step (NULL: reference controller
 * only): h = safety * exp(y), clamped to [h_c / 8, 8 h_c] around the reference
 * controller's proposal h_c.  m must live on the disk's device and outlive its
 * use by the integrator; the integrator must be adaptive. */
int rd_integrator_set_proposer(rd_integrator* it, rd_mlp* m, float safety);
/* record (features -> target log h) of every accepted step; the target is
 * log(h * (tol/err)^(1/8)), the step the Fehlberg estimate says would have
 * just met the tolerance.  The log is host memory that grows by 13 floats per
 * accepted step and disk until rd_integrator_clear_log. */
int rd_integrator_set_logging(rd_integrator* it, int on);
int rd_integrator_log(rd_integrator* it, float* x, float* t, uint32_t cap, uint32_t* n);
int rd_integrator_clear_log(rd_integrator* it);

/* ---- seeded disks ------------------------------------------------------------- */
/* Star + n planetesimals from orbital elements, barycentric, tau units
 * (src/red.cuda.initial/main.cu:251-342, 1662-1670; samplers
 * src/red.cuda.initial/distribution.cpp:11-162).  size_mode 0: log-uniform
 * mass on [lo, hi] Msun; 1: log-uniform radius on [lo, hi] km. */
typedef struct rd_disk_spec {
    uint32_t n, seed;
    int32_t  body_type;
    double   a_min, a_max, a_power;
    double   sig_e, sig_i;
    int32_t  size_mode;
    double   lo, hi;
    double   density_cgs, cd;
    double   m_star, r_star;
} rd_disk_spec;
int rd_make_disk(const rd_disk_spec* spec, rd_var4* r, rd_var4* v, rd_param* p, rd_body_md* bmd);

#ifdef __cplusplus
}
#endif
#endif
