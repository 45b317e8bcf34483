// engine.hpp -- internal types shared by the kernels and the host engine.
//
// Device layout (HBM), all disks of a batch concatenated:
//   state y_n      : x, y, z, vx, vy, vz          FP64 SoA (24 B per body per half)
//   state y_{n+1}  : xo, yo, zo, vxo, vyo, vzo    FP64 SoA
//   stage positions: P[2][3]                      FP64 SoA, double-buffered by stage parity
//   stage k        : kv[r][3] (= velocity at stage r, the position derivative)
//                    ka[r][3] (= acceleration at stage r)            r < 13
//   pair partials  : part[(c*3 + comp) * n_all + i]: FP64 from the FP64 pair
//                    kernel (one slot), FP32 chunk totals in mixed precision
//   sources        : m_ex (FP64 mass, -1 if inactive), m32 (FP32 mass, 0 if
//                    inactive or summed in FP64)
//   gas grid       : FARGO cells as 32-byte records (density, v_rad, v_theta, 0),
//                    ring-major, then the n_rad + 1 ring radii
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rdamd {

constexpr int kMaxStage = 13;
constexpr int kBlock = 256;          // threads per block of the pair kernels
constexpr int kBodyBlock = 64;       // per-body kernels: one wave per block, so small N still spreads over all CUs
constexpr int kTile = 256;           // sources staged in LDS per tile
// sinks per work item of the exact (FP64, reference-order) pair kernel
constexpr uint32_t kExactSinks = 16;
// ... and 4 for disks whose 16-sink items would leave CUs idle
constexpr uint32_t kExactSinksSmall = 4;
// bodies per block of the symmetric pair kernel (64 lanes x 4 sinks)
constexpr uint32_t kSymBlockBodies = 256;
// largest chunk (in blocks) of a symmetric-kernel work item; its reaction
// accumulators take 6 KB of LDS per block (8 KB + 8 x 6 KB = 56 KB at most)
constexpr uint32_t kSymMaxR = 8;
// the symmetric kernel groups blocks into R-block chunks, R = blocks / kSymChunks
// (at least 1, at most kSymMaxR): a disk has kSymChunks..2 kSymChunks - 1
// chunks (partial-sum slots per body) once it has that many blocks
constexpr uint32_t kSymChunks = 128;
// the symmetric kernel takes the diagonal tiles of a disk of at most this many
// 256-body blocks (2 / (blocks + 1) of its tiles, 8% or more) with reactions,
// the others one-sided: the kernel instance that holds the reaction sweep
// compiles its main loop in another order, and the sweep saves a third of a
// diagonal tile's work (measured on one MI355X: a batch of 512 disks of
// 64..4096 bodies, 16% diagonal tiles, runs 2.7% faster with it; a single
// 16384-body disk, 3%, ran 2% slower before the source tile was doubled in LDS
// and runs at the same speed since)
constexpr uint32_t kSymSelfMaxBlocks = 24;

// Per-disk control flags (DiskCtl::flags).
enum : uint32_t {
    F_ACTIVE = 1u,   // disk takes part in this attempt
    F_NEED_K0 = 2u,  // k0 = f(t, y_n) must be (re)computed
    F_COMMIT = 4u,   // copy y_{n+1} -> y_n
    F_CHECK = 8u,    // event detection runs on this disk
    F_K0_FORCE = 16u,  // the attempt's own stage-0 pair pass runs (not done ahead of it)
    F_MORE = 32u,      // the disk goes on with another step if this attempt is accepted
    F_SPEC_K0 = 64u,   // set on the device: accepted, committed, next k0 pass runs now
};

// Per-disk static layout.  Indices are local to the disk except off.
struct DiskInfo {
    uint32_t off, n;               // global offset, playing body count
    uint32_t n_si, n_massive;      // class bounds (red_type.h:670-700)
    uint32_t n_fp64;               // leading sources summed in FP64 (mixed mode)
    uint32_t drag_lo, drag_hi;     // bodies that feel the gas
    uint32_t err_hi;               // bodies [0, err_hi) enter the error norm
    uint32_t nchunk[3];            // partial-sum slots per class (SI, NSI, NI)
    uint32_t pad;
};

// Per-disk dynamic control, written by the host before each attempt.
struct DiskCtl {
    double t;       // time at the start of the step
    double h;       // dt_try
    uint32_t flags;
    uint32_t pad;
};

// One block of a pair kernel.  Global body indices.
struct WorkItem {
    uint32_t disk;
    uint32_t sink_lo, sink_hi;
    uint32_t src_lo, src_hi;
    uint32_t slot;       // partial-sum slot of the sinks
    uint32_t src_slot;   // symmetric kernel: partial-sum slot of the sources
    uint32_t flags;      // symmetric kernel: kItemSelf
};
// diagonal item whose full-block diagonal tiles take each pair once with its
// reaction (a disk of at most kSymSelfMaxBlocks blocks)
constexpr uint32_t kItemSelf = 1u;

// Gas disk, device copy: the analytic model (analytic_gas_disk_params_t
// red_type.h:464-485) or frame 0 of a FARGO grid (fargo_gas_disk.h:8-60).
enum : int32_t { kGasNone = 0, kGasAnalytic = 1, kGasGrid = 2 };
struct GasDev {
    double rho0, rho1, sch0, sch1, eta0, eta1;
    double t0, t1, e_folding_time;
    int32_t gas_decrease;
    int32_t on;                  // kGasNone, kGasAnalytic or kGasGrid
    // grid: n_rad rings x n_sec sectors, cell (ring, sector) at ring * n_sec + sector
    const double* g_rad;         // n_rad + 1 ring interface radii, increasing
    // per cell (density, radial velocity, azimuthal velocity on top of the
    // circular one, 0): one 32-byte record, so a lookup reads one segment
    const double4* g_cell;
    uint32_t n_rad, n_sec;
};

struct StageArgs {
    const double* pos[3];        // y_n positions
    const double* vel[3];        // y_n velocities (= stage-0 position derivative)
    double* out[6];              // y_{n+1}: x, y, z, vx, vy, vz
    const double* fpos[3];       // force-input positions of this stage
    double* npos[3];             // next-stage positions (the other P buffer)
    double* K;                   // [13][6][n_all]: velocity (0-2) and acceleration (3-5) per stage
    const double* part;
    uint32_t n_all;
    uint32_t err_all;            // error_check_for_tp
    const uint32_t* disk_of;
    const DiskInfo* info;
    const DiskCtl* ctl;
    unsigned long long* err;     // per disk, bits of the raw max |E|
    unsigned long long* errf;    // per disk, nonzero if an error term was NaN or infinite
    const double *mass, *radius, *density, *cd;
    const int32_t* id;
    GasDev gas;
    int mixed;
    int adaptive;
    int fuse_comb1;              // RK4 stage 0 also forms stage 1 (fixed step: k0 is never reused)
    int commit;                  // fused fixed-step RK4 stage 3: -1 stores y_{n+1} in out; 0 or 1
                                 // commits it in place (k_commit's copy) and, for 1, advances t
    int one_disk;                // every body belongs to disk 0
    DiskInfo info0;              // info[0] when one_disk (kernel arguments are scalar loads)
    double* trash;               // [3][kBodyBlock]: stores of a single-disk stage that does not apply
    unsigned long long* errw;    // single disk: error bits per wave of the last stage, [3][blocks]
                                 // (kErrNonFinite set if a term of the wave was NaN or infinite)
};

constexpr unsigned long long kErrNonFinite = 1ull << 63;

// The per-body arrays an event record is read from and written back to, in
// record order: x, y, z, vx, vy, vz, mass, radius, density, cd.
constexpr uint32_t kBodyRecord = 10;
struct BodyArrays {
    double* a[kBodyRecord];
};

// Launchers (kernels.hip).
// lane: 0 term-parallel 16-sink items, 1 one sink per lane, 2 term-parallel 4-sink items
hipError_t launch_force_exact(hipStream_t s, int lane, const WorkItem* items, uint32_t n_items, uint32_t grid,
                              const double* x, const double* y, const double* z, const double* m_ex,
                              const DiskCtl* ctl, uint32_t need, double* part, uint32_t n_all);
// The FP64 pair pass (4-sink items) of single-disk, gas-free, fixed-step RK4
// stage r (0-3) with the stage combination done by the kernel's summing lanes
// (no stage kernel, no partial-sum slot).
hipError_t launch_force_exact_rk4(hipStream_t s, int r, const WorkItem* items, uint32_t n_items, uint32_t grid,
                                  const double* x, const double* y, const double* z, const double* m_ex,
                                  const DiskCtl* ctl, uint32_t need, const StageArgs& a);
hipError_t launch_force_mixed(hipStream_t s, int sinks_per_lane, const WorkItem* items, uint32_t n_items,
                              uint32_t grid, const double* x, const double* y, const double* z,
                              const float* m32, const DiskCtl* ctl, uint32_t need_mask, float* part,
                              uint32_t n_all);
// Symmetric pair kernel over (chunk a, chunk c) items: sink_lo/hi = chunk a,
// src_lo/hi = chunk c (256-body blocks, R per chunk), slot = c (sink side),
// src_slot = a (source side).
// self: some item carries kItemSelf (the kernel instance with the diagonal
// reaction sweep).
hipError_t launch_force_sym(hipStream_t s, uint32_t R, bool self, const WorkItem* items, uint32_t n_items,
                            uint32_t grid, const double* x, const double* y, const double* z, const float* m32,
                            const DiskCtl* ctl, uint32_t need_mask, float* part, uint32_t n_all);
size_t force_sym_lds_bytes(uint32_t R);
hipError_t launch_stage(hipStream_t s, int method, int r, const StageArgs& a);
hipError_t launch_comb1(hipStream_t s, int method, const StageArgs& a);
hipError_t launch_commit(hipStream_t s, const StageArgs& a, int advance_time, DiskCtl* ctl_rw);
// The host's accept decision made on the device: errw (nw per-wave slots) for
// one disk, else err[k] / err[n_disk + k]; sets F_COMMIT, and F_SPEC_K0 where
// F_MORE was set, on accepted disks; clears every other flag.
hipError_t launch_decide(hipStream_t s, const unsigned long long* errw, uint32_t nw, const unsigned long long* err,
                         uint32_t n_disk, DiskCtl* ctl, double lambda, double tol, int nonfinite_guard);
hipError_t launch_operator(hipStream_t s, const StageArgs& a, double t, uint32_t disk, double* ax, double* ay,
                           double* az, const double* vx, const double* vy, const double* vz);
// var4 records (x, y, z, w) <-> component arrays, for calc_dydx on device arrays
hipError_t launch_var4_split(hipStream_t s, uint32_t n, const double* r, const double* v, double* const p[3],
                             double* const vel[3]);
hipError_t launch_var4_join(hipStream_t s, uint32_t n, const double* const a[3], double* dy);
hipError_t launch_collision_grid(hipStream_t s, uint32_t n_all, const double* x, const double* y, const double* z,
                                 const double* radius, const int32_t* id, const uint32_t* disk_of,
                                 const DiskCtl* ctl, double f, uint32_t table, uint32_t* head, uint32_t* next,
                                 long long* cell, unsigned long long* rmax, uint32_t* pairs, uint32_t* counter,
                                 uint32_t cap);
hipError_t launch_ejection(hipStream_t s, uint32_t n_all, const double* x, const double* y, const double* z,
                           const int32_t* id, const int32_t* type, const uint32_t* disk_of, const DiskCtl* ctl,
                           double ed2, double hcd2, uint32_t* out, uint32_t* counter, uint32_t cap);
hipError_t launch_integral(hipStream_t s, uint32_t lo, uint32_t hi, const double* x, const double* y,
                           const double* z, const double* vx, const double* vy, const double* vz,
                           const double* mass, const int32_t* id, double* scratch, size_t scratch_len,
                           double* out /* 11 doubles: M, Rx,Ry,Rz, Vx,Vy,Vz, Cx,Cy,Cz, T */,
                           double* out_u);
hipError_t launch_copy(hipStream_t s, const void* src, void* dst, size_t bytes);
// Step-proposer features of disks flagged F_CHECK, from state (x..vz):
// feat[3*d + 0] = min over pairs of r^2 / v^2 (float bits, atomicMin; the
// caller fills feat with 0xff first), feat[3*d + 1] = min periapsis (same),
// esum[d] = Keplerian energy of the disk (fixed-order sum of eterm[n_all]).
hipError_t launch_features(hipStream_t s, const WorkItem* items, uint32_t n_items, uint32_t grid, uint32_t n_all,
                           uint32_t n_disk, const double* x, const double* y, const double* z, const double* vx,
                           const double* vy, const double* vz, const double* mass, const int32_t* id,
                           const uint32_t* disk_of, const DiskInfo* info, const DiskCtl* ctl, uint32_t* feat,
                           double* eterm, double* esum);
// rec[kBodyRecord q + f] = b.a[f][idx[q]] for q < n
hipError_t launch_gather_bodies(hipStream_t s, uint32_t n, const uint32_t* idx, const BodyArrays& b, double* rec);
// the inverse, plus id[idx[q]] = ids[q]
hipError_t launch_scatter_bodies(hipStream_t s, uint32_t n, const uint32_t* idx, const double* rec,
                                 const int32_t* ids, const BodyArrays& b, int32_t* id);
hipError_t launch_set_sources(hipStream_t s, uint32_t n_all, const double* mass, const int32_t* id,
                              const uint32_t* disk_of, const DiskInfo* info, double* m_ex, float* m32);

}  // namespace rdamd
