/*
 * oracle.h -- CPU restatement of red.cuda's direct-sum N-body hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the HIP engine is
 * compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product never does.
 *
 * Every function restates the reference algorithm with the same operation
 * order (compile with -ffp-contract=off, as the reference builds with
 * -fmad=false, Makefile:17), citing the reference file:line it follows.
 * Paths are relative to the reference repository root.
 *
 * Data layout is the reference's own AoS FP64 layout:
 *   var4_t      red.type/red_type.h:206-212     (32 B, w = padding)
 *   param_t     red.type/red_type.h:347-353     (32 B)
 *   body_md_t   red.type/red_type.h:356-362     (24 B on Linux)
 *   event_t     red.type/red_type.h:415-461     (432 B)
 *   gas_t       red.type/red_type.h:464-485     (analytic gas disk params)
 *
 * Parity status of each part is listed in the header comment of oracle.c.
 */
#ifndef RED_ORACLE_H
#define RED_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double x, y, z, w; } or_var4;
typedef struct { double mass, radius, density, cd; } or_param;
typedef struct { int32_t id, body_type, mig_type; double mig_stop_at; } or_body_md;
typedef struct { or_var4 R, V, C; double E; } or_integral;

typedef struct {
    int32_t event_name;
    double  t, d;
    int32_t id0; uint32_t idx0; or_param p0; or_var4 r0, v0;
    int32_t id1; uint32_t idx1; or_param p1; or_var4 r1, v1;
    int32_t id2; uint32_t idx2; or_param p2; or_var4 r2, v2;
    or_param ps; or_var4 rs, vs;
} or_event;

typedef struct {
    double rho[2], sch[2], eta[2], tau[2], mfp[2], temp[2];
    double c_vth;
    int32_t gas_decrease;
    double t0, t1, e_folding_time;
    double alpha, mean_molecular_weight, particle_diameter;
} or_gas;

typedef struct { double sma, ecc, inc, peri, node, mean; } or_orbelem;

enum { OR_STAR = 0, OR_GIANT, OR_ROCKY, OR_PROTO, OR_SUPERPL, OR_PL, OR_TP, OR_NTYPE };
enum { OR_THR_HC = 0, OR_THR_EJ = 1, OR_THR_RAD = 2 };
enum { OR_EV_NONE = 0, OR_EV_HIT_CENTRUM, OR_EV_EJECTION, OR_EV_COLLISION };
enum { OR_FORMULA_GPU = 0, OR_FORMULA_CPU_SYM = 1 };
enum { OR_DRAG_NSI = 0, OR_DRAG_ALL = 1 };
enum { OR_OK = 0, OR_EINVAL = 1, OR_ENOTIMPL = 3, OR_EOVERFLOW = 4, OR_ESTATE = 5 };

/* ---- n_objects_t interaction bounds (red_type.h:622-700) ---------------- */
void or_bound(const uint32_t playing[OR_NTYPE], int which, uint32_t out[4]);

/* ---- unit helpers / tools (src/red.cuda.utility/tools.cpp) ------------- */
double or_calc_radius(double m, double density);
double or_calc_density(double m, double R);
double or_calc_mass(double R, double density);
void   or_calc_physical_properties(const or_param* p1, const or_param* p2, or_param* p);
void   or_calc_position_after_collision(double m1, double m2, const or_var4* r1, const or_var4* r2, or_var4* r);
void   or_calc_velocity_after_collision(double m1, double m2, const or_var4* v1, const or_var4* v2, or_var4* v);
double or_get_total_mass(uint32_t n, const or_body_md* bmd, const or_param* p);
void   or_calc_bc(uint32_t n, const or_body_md* bmd, const or_param* p, const or_var4* r, const or_var4* v, double M0, or_var4* R0, or_var4* V0);
void   or_transform_to_bc(uint32_t n, const or_body_md* bmd, const or_param* p, or_var4* r, or_var4* v);
void   or_calc_integral(uint32_t n, const or_body_md* bmd, const or_param* p, const or_var4* r, const or_var4* v, or_integral* out);
int    or_kepler_equation_solver(double ecc, double mean, double eps, double* E);
int    or_calc_phase(double mu, const or_orbelem* oe, or_var4* r, or_var4* v);

/* ---- MT19937 + samplers (src/red.cuda.initial/distribution.cpp) -------- */
typedef struct { uint32_t I[624]; uint32_t idx; } or_rng;
void     or_rng_seed(or_rng* g, uint32_t seed);
uint32_t or_rng_u32(or_rng* g);
double   or_rng_uniform(or_rng* g, double lo, double hi);
double   or_rng_rayleigh(or_rng* g, double sigma);
double   or_rng_power_law(or_rng* g, double lo, double hi, double power);

/* Seeded planetesimal disk (star + n bodies) in tau units, barycentric. */
typedef struct {
    uint32_t n;            /* bodies besides the star */
    uint32_t seed;
    int32_t  body_type;    /* type given to the disk bodies (3 = protoplanet: all-pairs) */
    double   a_min, a_max, a_power;   /* dN/da ~ a^a_power */
    double   sig_e, sig_i;            /* Rayleigh sigmas */
    int32_t  size_mode;               /* 0: log-uniform mass; 1: log-uniform radius */
    double   lo, hi;                  /* mass [Msun] or radius [km] range */
    double   density_cgs;             /* body density [g/cm^3] */
    double   cd;                      /* drag coefficient */
    double   m_star, r_star;          /* [Msun], [AU] */
} or_disk_spec;
int or_make_disk(const or_disk_spec* s, or_var4* r, or_var4* v, or_param* p, or_body_md* bmd);

/* ---- forces (src/red.cuda/pp_disk.cu) ----------------------------------- */
/* Reference GPU formula: per class (SI, NSI, NI), zero a_i, sum j ascending,
 * a += (m_j/(d*d^2)) * d_vec  (pp_disk.cu:374-416, 1014-1045). */
void or_grav_gpu(uint32_t n, const uint32_t playing[OR_NTYPE], const or_body_md* bmd,
                 const or_param* p, const or_var4* r, or_var4* a);
/* Reference CPU path: memset, SI symmetric (Newton 3), NSI throws, NI
 * (pp_disk.cu:822-865, 933-967, 974-1012).  Returns OR_ENOTIMPL with NSI. */
int  or_grav_cpu(uint32_t n, const uint32_t playing[OR_NTYPE], const or_body_md* bmd,
                 const or_param* p, const or_var4* r, or_var4* a);
/* Gas drag, quadratic regime (pp_disk.cu:60-155, 486-521, 701-777). */
void or_drag(double t, const or_gas* g, int drag_mode, uint32_t n, const uint32_t playing[OR_NTYPE],
             const or_param* p, const or_var4* r, const or_var4* v, or_var4* a);

/* FARGO gas grid, frame 0 (fargo_gas_disk.h:8-60): n_rad + 1 ring radii,
 * n_rad * n_sec cells of density, radial and extra azimuthal velocity. */
typedef struct or_grid {
    uint32_t n_rad, n_sec;
    const double *used_rad, *density, *vrad, *vtheta;
} or_grid;
or_var4  or_rotate_2d(double theta, or_var4 r);                 /* util.cu:76-83 */
uint32_t or_grid_cell(const or_grid* g, double x, double y);   /* pp_disk.cu:158-193 */
void     or_grid_drag(const or_grid* g, int drag_mode, const uint32_t playing[OR_NTYPE], const or_param* p,
                      const or_var4* r, const or_var4* v, or_var4* a);   /* pp_disk.cu:778-814 */

/* ---- disk + integrator objects ------------------------------------------ */
typedef struct or_disk or_disk;
typedef struct or_integ or_integ;

or_disk* or_disk_create(uint32_t n, const uint32_t n_by_type[OR_NTYPE],
                        const or_var4* r, const or_var4* v, const or_param* p, const or_body_md* bmd,
                        double t, const double thr[3], const or_gas* gas);
void     or_disk_destroy(or_disk* d);
void     or_disk_set_formula(or_disk* d, int formula);
void     or_disk_set_drag_mode(or_disk* d, int mode);
/* copy a FARGO grid into the disk; used_rad NULL removes it */
int      or_disk_set_gas_grid(or_disk* d, uint32_t n_rad, uint32_t n_sec, const double* used_rad,
                              const double* density, const double* vrad, const double* vtheta);
uint32_t or_disk_n(const or_disk* d);
double   or_disk_t(const or_disk* d);
void     or_disk_get(const or_disk* d, or_var4* r, or_var4* v, or_param* p, or_body_md* bmd);
void     or_disk_counts(const or_disk* d, uint32_t playing[OR_NTYPE], uint32_t inactive[OR_NTYPE], uint32_t removed[OR_NTYPE]);
int      or_calc_dydx(or_disk* d, uint32_t i, uint32_t rr, double t, const or_var4* r, const or_var4* v, or_var4* dy);
void     or_disk_integral(const or_disk* d, or_integral* out);

/* events (pp_disk.cu:284-371, 1096-1368) */
uint32_t or_check_ejection_hit_centrum(or_disk* d);
int      or_check_collision(or_disk* d, uint32_t* n_events);
int      or_handle_ejection_hit_centrum(or_disk* d);
int      or_handle_collision(or_disk* d);
void     or_clear_event_counter(or_disk* d);
int      or_rebuild_vectors(or_disk* d);
uint32_t or_event_log(const or_disk* d, or_event* out, uint32_t cap);

/* integrators (int_euler.cu, int_rungekutta2.cu, int_rungekutta4.cu,
   int_rungekutta8.cu, integrator.cu); method as integrator_type_t
   (red_type.h:136-143): 0 Euler, 1 RK2, 2 RK4, 4 RKF7(8).  Euler and RK2 are
   fixed-step whatever `adaptive` says, as their constructors are. */
or_integ* or_integ_create(or_disk* d, int method, double dt, int adaptive, double tol);
void      or_integ_destroy(or_integ* it);
int       or_step(or_integ* it, double* dt_did);
double    or_get_dt_next(const or_integ* it);
void      or_set_dt_next(or_integ* it, double dt);
void      or_counters(const or_integ* it, uint64_t out[3]);   /* tried, passed, failed */
void      or_set_error_check_for_tp(or_integ* it, int on);
/* Opt-in bound on the retry loop (src/red_v2/int_rungekutta7.cu:176-185 has
 * such a bound; red.cuda has none): or_step returns OR_ESTATE, keeping y_n and
 * counting the step's attempts as failed, when a step is still rejected after
 * max_attempts attempts, when an attempt would use |dt| <= min_dt (it is not
 * made), or (nonfinite) when an error term is NaN or infinite.  0 = off. */
void      or_set_guard(or_integ* it, uint32_t max_attempts, double min_dt, int nonfinite);
/* The controller's error of the last attempt, |lambda * dt * max_i |E_i||
 * (int_rungekutta8.cu:693-696, int_rungekutta4.cu:226-229), and the terms
 * |E_i| it is the maximum of, for every component of every body (positions
 * then velocities, 4 doubles per body): E = k0 + k10 - k11 - k12 for RKF7(8),
 * E = k3 - k4 for adaptive RK4.  Returns the number of terms (8 n), or 0 for
 * a method without an error estimate.  Test hooks: they read what step() kept. */
double    or_last_error(const or_integ* it);
/* The Fehlberg 7(8) tableau the RKF7(8) step uses: c[13], the packed lower
 * triangle a[79] (a[0] = 0, then row r >= 1 holds its r entries from index
 * 1 + r (r - 1) / 2) and the 7th-order b[11]. */
void      or_rk8_tableau(double c[13], double a[79], double b[11]);
uint32_t  or_last_error_terms(const or_integ* it, double* out, uint32_t cap);

/* driver loop semantics (src/red.cuda/main.cu:470-625) without file output */
int or_run(or_disk* d, or_integ* it, double length, double output_interval, uint64_t max_steps, uint64_t* steps_done);

#ifdef __cplusplus
}
#endif
#endif
