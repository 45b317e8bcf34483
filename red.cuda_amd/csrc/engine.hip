// engine.hip -- host side of the engine behind include/red_abi.h.
//
// One rd_disk holds one or more independent planetesimal disks on one GPU.
// An integrator "attempt" is a fixed launch sequence over all disks:
//   force(0) + stage(0)   [disks that start a new step: k0 = f(t, y_n)]
//   comb1                  [stage-1 positions/velocities from k0]
//   force(r) + stage(r)    r = 1 .. last stage
// where every stage kernel finishes k_r for its body and forms the next
// stage's inputs (one pass over HBM per stage), and the last stage folds the
// Fehlberg error into a per-disk max.  The host reads one double per disk,
// applies the reference step controller (int_rungekutta8.cu:693-708) with the
// host's libm pow, and uploads the next per-disk (t, h, flags).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/red_abi.h"
#include "engine.hpp"
#include "device_scope.hpp"
#include "mlp.hpp"

using namespace rdamd;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

}  // namespace

namespace rdamd {
void set_last_error(const std::string& msg) { g_err = msg; }
}  // namespace rdamd

namespace {

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return fail(RD_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));       \
    } while (0)

constexpr int kNType = RD_BODY_TYPE_N;

struct HostDisk {
    uint32_t off = 0, n = 0;
    uint32_t playing[kNType] = {0}, inactive[kNType] = {0}, removed[kNType] = {0};
    double t = 0.0;
    std::vector<rd_event> pending;
    std::vector<rd_event> log;
    uint32_t n_event[3] = {0, 0, 0};   // TOTAL, LAST_CLEAR, LAST_STEP (red_type.h:154-160)
    // the same counters per kind: collision, ejection, hit-centrum (pp_disk.h:164-166)
    uint32_t n_kind[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};

    uint32_t n_si() const { return playing[0] + playing[1] + playing[2] + playing[3]; }
    uint32_t n_nsi() const { return playing[4] + playing[5]; }
    uint32_t n_massive() const { return n_si() + n_nsi(); }
    uint32_t n_active() const
    {
        uint32_t a = 0;
        for (int k = 0; k < kNType; ++k) a += playing[k] - inactive[k];
        return a;
    }
};

template <typename T>
int dev_alloc(T** p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)p, count * sizeof(T));
    if (e != hipSuccess) return fail(RD_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    return RD_OK;
}
template <typename T>
void dev_free(T*& p)
{
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

// Timing events released on every return path.
struct EventPair {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t create()
    {
        hipError_t e = hipEventCreate(&e0);
        return e == hipSuccess ? hipEventCreate(&e1) : e;
    }
    // ms between the two events, after waiting for the second
    hipError_t elapsed(float* ms)
    {
        hipError_t e = hipEventSynchronize(e1);
        return e == hipSuccess ? hipEventElapsedTime(ms, e0, e1) : e;
    }
    ~EventPair()
    {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
};

}  // namespace

struct rd_disk {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_err = nullptr;   // the error words of an attempt are on the host
    int n_cu = 256;
    uint32_t n_disk = 0, n_all = 0;
    std::vector<HostDisk> hd;

    // host copies of the per-body data that the device does not need
    std::vector<rd_body_md> bmd;

    // device state
    double *x = nullptr, *y = nullptr, *z = nullptr, *vx = nullptr, *vy = nullptr, *vz = nullptr;
    double *xo = nullptr, *yo = nullptr, *zo = nullptr, *vxo = nullptr, *vyo = nullptr, *vzo = nullptr;
    double* P[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
    double* K = nullptr;           // [13][6][n_all]
    double* part = nullptr;
    size_t part_len = 0;
    double *mass = nullptr, *radius = nullptr, *density = nullptr, *cd = nullptr;
    int32_t *id = nullptr, *type = nullptr;
    double* m_ex = nullptr;
    float* m32 = nullptr;
    uint32_t* disk_of = nullptr;
    DiskInfo* info = nullptr;
    DiskCtl* ctl = nullptr;
    unsigned long long* err = nullptr;
    double* scratch = nullptr;
    size_t scratch_len = 0;
    double* red_out = nullptr;     // 12 doubles for integrals
    double* trash = nullptr;       // [3][kBodyBlock], written by single-disk stages that do not apply
    unsigned long long* errw = nullptr;     // single disk: per-wave error bits of the last stage
    unsigned long long* h_errw = nullptr;   // pinned copy of errw
    uint32_t* feat = nullptr;      // step-proposer features, 3 per disk
    double* esum = nullptr;        // Keplerian energy per disk
    double* eterm = nullptr;       // per-body Keplerian energy [n_all]
    float* fx = nullptr;           // proposer input rows [n_disk][12]
    float* fy = nullptr;           // proposer output [n_disk]

    // work lists
    WorkItem* items_force = nullptr;
    uint32_t n_items_force = 0, grid_force = 0;
    WorkItem* items_sym = nullptr;   // mixed precision, symmetric SI x SI pairs
    uint32_t n_items_sym = 0, grid_sym = 0, sym_R = 1;
    bool sym_self = false;           // diagonal tiles with reactions (launch_force_sym)
    // collision grid: bucket heads, per-body list links and cells, largest radius
    uint32_t* coll_head = nullptr;
    uint32_t* coll_next = nullptr;
    long long* coll_cell = nullptr;
    unsigned long long* coll_rmax = nullptr;
    uint32_t coll_table = 0;
    WorkItem* items_feat = nullptr;
    uint32_t n_items_feat = 0;
    bool exact_lane = false;       // FP64 pair kernel: one sink per lane (see build_lists)
    bool exact_small = false;      // FP64 pair kernel: 4-sink items (small disks, see build_lists)
    bool lists_valid = false;
    double pairs_per_eval = 0.0;

    // events
    uint32_t* ev_buf = nullptr;
    uint32_t* ev_counter = nullptr;
    uint32_t ev_cap = 0;
    // records of the bodies an event touches, staged for one gather / scatter
    uint32_t* ev_sidx = nullptr;
    int32_t* ev_sids = nullptr;
    double* ev_srec = nullptr;
    uint32_t ev_scap = 0;
    // integrators over this disk: their graphs are released before the stream
    std::vector<rd_integrator*> integrators;

    // pinned host mirrors
    DiskCtl* h_ctl = nullptr;
    unsigned long long* h_err = nullptr;
    uint32_t* h_feat = nullptr;    // [3 * n_disk]
    double* h_esum = nullptr;      // [n_disk]
    float* h_fx = nullptr;         // [n_disk][12]
    float* h_fy = nullptr;         // [n_disk]

    std::vector<DiskInfo> h_info;

    int precision = RD_PREC_FP64;
    int drag_mode = RD_DRAG_NSI;
    rd_tuning tuning{2, 1024, 4, 0, 1, 1};
    double thr[3] = {0, 0, 0};
    GasDev gas{};
    int32_t gas_created = kGasNone;  // the model given at creation (a grid's removal returns to it)
    // FARGO grid on the device: cell records (density, v_rad, v_theta, 0), then ring radii
    double* grid = nullptr;
    uint64_t generation = 0;       // bumps when buffers / lists change (invalidates graphs)
};

struct rd_integrator {
    rd_disk* d = nullptr;
    int device = -1;   // the disk's device, kept so that destroy never reads the disk
    bool fixed_dt = false;   // Euler, RK2: every step is dt_try, dt_next follows it
    int method = RD_RKF78;
    int adaptive = 1;
    double tol = 0.0;
    int n_stage = 13, order = 7;
    int err_all = 0;
    std::vector<double> dt_try, dt_did, dt_next;
    std::vector<uint64_t> tried, passed, failed;
    uint64_t force_evals = 0;
    double pairs = 0.0;
    // opt-in bound on the retry loop (0 = off: the reference's unbounded loop)
    uint32_t max_attempts = 0;
    double min_dt = 0.0;
    bool reject_nonfinite = false;
    // rd_steps: accept, commit and the next k0 pass on the device while the
    // host reads the error (same arithmetic, so the same results)
    bool overlap = true;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    uint64_t graph_gen = ~0ull;
    int graph_kind = -1;
    // learned step proposer
    struct Hist {
        double h1 = 0, h2 = 0, e1 = 1, e2 = 1, tenc_prev = 0, ek0 = 0;
        uint32_t rej = 0;
        bool ek0_set = false;
        int accepted = 0;
        bool has_feat = false;
        float feat[RD_MLP_IN];
    };
    std::vector<Hist> hist;
    rd_mlp* proposer = nullptr;
    float safety = 1.0f;
    bool logging = false;
    std::vector<float> log_x, log_t;
    uint64_t proposals = 0;
};

namespace {

// --------------------------------------------------------------------------
// layout helpers
// --------------------------------------------------------------------------
// An integrator's captured attempt is destroyed only once the stream it was
// launched on has drained, and always before that stream is destroyed.
// An integrator whose disk was destroyed refuses every call that needs the disk.
int check_live(const rd_integrator* it)
{
    if (!it) return fail(RD_EINVAL, "NULL integrator");
    if (!it->d) return fail(RD_ESTATE, "the integrator's disk was destroyed");
    return RD_OK;
}

void release_graph(rd_integrator* it)
{
    if (!it->gexec && !it->graph) return;
    if (it->d) (void)hipStreamSynchronize(it->d->stream);
    if (it->gexec) (void)hipGraphExecDestroy(it->gexec);
    if (it->graph) (void)hipGraphDestroy(it->graph);
    it->gexec = nullptr;
    it->graph = nullptr;
}

double* kv_ptr(rd_disk* d, int r, int c) { return d->K + ((size_t)r * 6 + c) * d->n_all; }
double* ka_ptr(rd_disk* d, int r, int c) { return d->K + ((size_t)r * 6 + 3 + c) * d->n_all; }

void free_buffers(rd_disk* d)
{
    double** dp[] = {&d->x, &d->y, &d->z, &d->vx, &d->vy, &d->vz, &d->xo, &d->yo, &d->zo, &d->vxo, &d->vyo,
                     &d->vzo, &d->K, &d->part, &d->mass, &d->radius, &d->density, &d->cd, &d->m_ex, &d->scratch,
                     &d->red_out, &d->eterm, &d->trash};
    for (auto p : dp) dev_free(*p);
    for (int b = 0; b < 2; ++b)
        for (int c = 0; c < 3; ++c) dev_free(d->P[b][c]);
    dev_free(d->id);
    dev_free(d->type);
    dev_free(d->m32);
    dev_free(d->disk_of);
    dev_free(d->errw);
    dev_free(d->info);
    dev_free(d->ctl);
    dev_free(d->err);
    dev_free(d->items_force);
    dev_free(d->items_sym);
    dev_free(d->coll_head);
    dev_free(d->coll_next);
    dev_free(d->coll_cell);
    dev_free(d->coll_rmax);
    d->coll_table = 0;
    dev_free(d->items_feat);
    dev_free(d->ev_buf);
    dev_free(d->ev_counter);
    dev_free(d->ev_sidx);
    dev_free(d->ev_sids);
    dev_free(d->ev_srec);
    d->ev_scap = 0;
    dev_free(d->feat);
    dev_free(d->esum);
    dev_free(d->fx);
    dev_free(d->fy);
    dev_free(d->grid);
    for (void** h : {(void**)&d->h_ctl, (void**)&d->h_err, (void**)&d->h_feat, (void**)&d->h_esum, (void**)&d->h_fx,
                     (void**)&d->h_fy, (void**)&d->h_errw}) {
        if (*h) (void)hipHostFree(*h);
        *h = nullptr;
    }
    d->part_len = 0;
    d->scratch_len = 0;
}

// Waves of a per-body launch (3 components): the single-disk error slots.
size_t n_errw(const rd_disk* d) { return 3 * (((size_t)d->n_all + kBodyBlock - 1) / kBodyBlock) + 1; }

int alloc_buffers(rd_disk* d)
{
    const size_t n = d->n_all;
    double** dp[] = {&d->x, &d->y, &d->z, &d->vx, &d->vy, &d->vz, &d->xo, &d->yo, &d->zo, &d->vxo, &d->vyo,
                     &d->vzo, &d->mass, &d->radius, &d->density, &d->cd, &d->m_ex, &d->eterm};
    for (auto p : dp)
        if (int rc = dev_alloc(p, n)) return rc;
    for (int b = 0; b < 2; ++b)
        for (int c = 0; c < 3; ++c)
            if (int rc = dev_alloc(&d->P[b][c], n)) return rc;
    if (int rc = dev_alloc(&d->K, (size_t)kMaxStage * 6 * n)) return rc;
    if (int rc = dev_alloc(&d->id, n)) return rc;
    if (int rc = dev_alloc(&d->type, n)) return rc;
    if (int rc = dev_alloc(&d->m32, n)) return rc;
    if (int rc = dev_alloc(&d->disk_of, n)) return rc;
    if (int rc = dev_alloc(&d->info, d->n_disk)) return rc;
    if (int rc = dev_alloc(&d->ctl, d->n_disk)) return rc;
    if (int rc = dev_alloc(&d->err, 2 * (size_t)d->n_disk)) return rc;   // max bits, then non-finite flags
    if (int rc = dev_alloc(&d->red_out, 16)) return rc;
    if (int rc = dev_alloc(&d->trash, 3 * (size_t)kBodyBlock)) return rc;
    if (int rc = dev_alloc(&d->errw, n_errw(d))) return rc;
    if (int rc = dev_alloc(&d->feat, 3 * (size_t)d->n_disk)) return rc;
    if (int rc = dev_alloc(&d->esum, d->n_disk)) return rc;
    if (int rc = dev_alloc(&d->fx, (size_t)RD_MLP_IN * d->n_disk)) return rc;
    if (int rc = dev_alloc(&d->fy, d->n_disk)) return rc;
    d->ev_cap = std::max<uint32_t>(4096u, (uint32_t)n);
    if (int rc = dev_alloc(&d->ev_buf, 2 * (size_t)d->ev_cap)) return rc;
    if (int rc = dev_alloc(&d->ev_counter, 1)) return rc;
    d->coll_table = 1024;
    while (d->coll_table < 2 * (size_t)n) d->coll_table *= 2;
    if (int rc = dev_alloc(&d->coll_head, d->coll_table)) return rc;
    if (int rc = dev_alloc(&d->coll_next, n)) return rc;
    if (int rc = dev_alloc(&d->coll_cell, 3 * n)) return rc;
    if (int rc = dev_alloc(&d->coll_rmax, 1)) return rc;
    HIP_TRY(hipHostMalloc((void**)&d->h_ctl, sizeof(DiskCtl) * std::max<uint32_t>(1, d->n_disk)));
    HIP_TRY(hipHostMalloc((void**)&d->h_err, 2 * sizeof(unsigned long long) * std::max<uint32_t>(1, d->n_disk)));
    HIP_TRY(hipHostMalloc((void**)&d->h_errw, sizeof(unsigned long long) * n_errw(d)));
    HIP_TRY(hipHostMalloc((void**)&d->h_feat, 3 * sizeof(uint32_t) * std::max<uint32_t>(1, d->n_disk)));
    HIP_TRY(hipHostMalloc((void**)&d->h_esum, sizeof(double) * std::max<uint32_t>(1, d->n_disk)));
    HIP_TRY(hipHostMalloc((void**)&d->h_fx, RD_MLP_IN * sizeof(float) * std::max<uint32_t>(1, d->n_disk)));
    HIP_TRY(hipHostMalloc((void**)&d->h_fy, sizeof(float) * std::max<uint32_t>(1, d->n_disk)));
    HIP_TRY(hipMemsetAsync(d->K, 0, (size_t)kMaxStage * 6 * n * sizeof(double), d->stream));
    HIP_TRY(hipMemsetAsync(d->ctl, 0, sizeof(DiskCtl) * d->n_disk, d->stream));
    d->generation++;
    return RD_OK;
}

// Host AoS arrays <-> device SoA for disk range [off, off + n).
int upload_bodies(rd_disk* d, uint32_t off, uint32_t n, const rd_var4* r, const rd_var4* v, const rd_param* p,
                  const rd_body_md* bmd)
{
    if (n == 0) return RD_OK;
    std::vector<double> buf((size_t)n);
    auto put = [&](double* dst, auto get) -> int {
        for (uint32_t i = 0; i < n; ++i) buf[i] = get(i);
        HIP_TRY(hipMemcpyAsync(dst + off, buf.data(), n * sizeof(double), hipMemcpyHostToDevice, d->stream));
        HIP_TRY(hipStreamSynchronize(d->stream));
        return RD_OK;
    };
    int rc = RD_OK;
    if (r) {
        if ((rc = put(d->x, [&](uint32_t i) { return r[i].x; }))) return rc;
        if ((rc = put(d->y, [&](uint32_t i) { return r[i].y; }))) return rc;
        if ((rc = put(d->z, [&](uint32_t i) { return r[i].z; }))) return rc;
    }
    if (v) {
        if ((rc = put(d->vx, [&](uint32_t i) { return v[i].x; }))) return rc;
        if ((rc = put(d->vy, [&](uint32_t i) { return v[i].y; }))) return rc;
        if ((rc = put(d->vz, [&](uint32_t i) { return v[i].z; }))) return rc;
    }
    if (p) {
        if ((rc = put(d->mass, [&](uint32_t i) { return p[i].mass; }))) return rc;
        if ((rc = put(d->radius, [&](uint32_t i) { return p[i].radius; }))) return rc;
        if ((rc = put(d->density, [&](uint32_t i) { return p[i].density; }))) return rc;
        if ((rc = put(d->cd, [&](uint32_t i) { return p[i].cd; }))) return rc;
    }
    if (bmd) {
        std::vector<int32_t> ib(n), tb(n);
        for (uint32_t i = 0; i < n; ++i) {
            ib[i] = bmd[i].id;
            tb[i] = bmd[i].body_type;
        }
        HIP_TRY(hipMemcpyAsync(d->id + off, ib.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, d->stream));
        HIP_TRY(hipMemcpyAsync(d->type + off, tb.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, d->stream));
        HIP_TRY(hipStreamSynchronize(d->stream));
    }
    return RD_OK;
}

int download_bodies(rd_disk* d, uint32_t off, uint32_t n, rd_var4* r, rd_var4* v, rd_param* p)
{
    if (n == 0) return RD_OK;
    std::vector<double> a((size_t)n), b((size_t)n), c((size_t)n), e((size_t)n);
    auto get3 = [&](const double* s0, const double* s1, const double* s2, rd_var4* out) -> int {
        HIP_TRY(hipMemcpyAsync(a.data(), s0 + off, n * sizeof(double), hipMemcpyDeviceToHost, d->stream));
        HIP_TRY(hipMemcpyAsync(b.data(), s1 + off, n * sizeof(double), hipMemcpyDeviceToHost, d->stream));
        HIP_TRY(hipMemcpyAsync(c.data(), s2 + off, n * sizeof(double), hipMemcpyDeviceToHost, d->stream));
        HIP_TRY(hipStreamSynchronize(d->stream));
        for (uint32_t i = 0; i < n; ++i) out[i] = rd_var4{a[i], b[i], c[i], 0.0};
        return RD_OK;
    };
    int rc;
    if (r && (rc = get3(d->x, d->y, d->z, r))) return rc;
    if (v && (rc = get3(d->vx, d->vy, d->vz, v))) return rc;
    if (p) {
        HIP_TRY(hipMemcpyAsync(a.data(), d->mass + off, n * sizeof(double), hipMemcpyDeviceToHost, d->stream));
        HIP_TRY(hipMemcpyAsync(b.data(), d->radius + off, n * sizeof(double), hipMemcpyDeviceToHost, d->stream));
        HIP_TRY(hipMemcpyAsync(c.data(), d->density + off, n * sizeof(double), hipMemcpyDeviceToHost, d->stream));
        HIP_TRY(hipMemcpyAsync(e.data(), d->cd + off, n * sizeof(double), hipMemcpyDeviceToHost, d->stream));
        HIP_TRY(hipStreamSynchronize(d->stream));
        for (uint32_t i = 0; i < n; ++i) p[i] = rd_param{a[i], b[i], c[i], e[i]};
    }
    return RD_OK;
}

// --------------------------------------------------------------------------
// disk info + work lists
// --------------------------------------------------------------------------
void fill_info(rd_disk* d)
{
    d->h_info.resize(d->n_disk);
    for (uint32_t k = 0; k < d->n_disk; ++k) {
        const HostDisk& h = d->hd[k];
        DiskInfo& di = d->h_info[k];
        di = DiskInfo{};
        di.off = h.off;
        di.n = h.n;
        di.n_si = h.n_si();
        di.n_massive = h.n_massive();
        // mixed precision sums the leading stars and giants in FP64 (the
        // types' counts when the SI class is sorted by type)
        uint32_t lead = 0;
        while (lead < h.n && (d->bmd[h.off + lead].body_type == RD_BODY_STAR ||
                              d->bmd[h.off + lead].body_type == RD_BODY_GIANTPLANET))
            ++lead;
        di.n_fp64 = d->precision == RD_PREC_MIXED ? lead : 0;
        if (d->drag_mode == RD_DRAG_NSI) {
            di.drag_lo = h.n_si();
            di.drag_hi = h.n_si() + h.n_nsi();
        } else {
            di.drag_lo = h.playing[0];
            di.drag_hi = h.n_massive();
        }
        di.err_hi = h.n_massive();
    }
}

// (sink range, source range) of the three reference classes, local indices
// (n_objects_t::get_bound_SI/NSI/NI, red_type.h:670-692).
void class_bounds(const HostDisk& h, int cls, uint32_t b[4])
{
    const uint32_t si = h.n_si(), nsi = h.n_nsi(), mas = h.n_massive(), ni = h.playing[6];
    if (cls == 0) { b[0] = 0; b[1] = si; b[2] = 0; b[3] = mas; }
    else if (cls == 1) { b[0] = si; b[1] = si + nsi; b[2] = 0; b[3] = si; }
    else { b[0] = mas; b[1] = mas + ni; b[2] = 0; b[3] = mas; }
}

int build_lists(rd_disk* d)
{
    d->lists_valid = false;   // until this build completes: a failure part-way forces a rebuild
    fill_info(d);
    std::vector<WorkItem> items;
    std::vector<std::pair<double, uint32_t>> order;   // (cost, disk)
    for (uint32_t k = 0; k < d->n_disk; ++k) order.push_back({(double)d->hd[k].n * d->hd[k].n, k});
    std::stable_sort(order.begin(), order.end(), [](auto& a, auto& b) { return a.first > b.first; });
    size_t max_chunks = 1;
    double pairs = 0.0;
    const uint32_t S = d->tuning.sinks_per_lane;
    const uint32_t C = std::max<uint32_t>(kTile, d->tuning.chunk / kTile * kTile);
    // FP64: one sink per lane when 256-sink items give >= 4 blocks per CU,
    // else the term-parallel kernel on 16-sink items, or on 4-sink items when
    // 16-sink items would be fewer than the CUs (same bits in every case)
    if (d->precision == RD_PREC_FP64) {
        size_t groups = 0, groups16 = 0;
        for (auto& o : order)
            for (int cls = 0; cls < 3; ++cls) {
                uint32_t b[4];
                class_bounds(d->hd[o.second], cls, b);
                if (b[1] > b[0]) {
                    groups += (b[1] - b[0] + kBlock - 1) / kBlock;
                    groups16 += (b[1] - b[0] + kExactSinks - 1) / kExactSinks;
                }
            }
        d->exact_lane = groups >= 4 * (size_t)d->n_cu;
        d->exact_small = !d->exact_lane && groups16 < (size_t)d->n_cu;
    }
    const uint32_t exact_group = d->exact_lane ? (uint32_t)kBlock : (d->exact_small ? kExactSinksSmall : kExactSinks);
    const bool sym = d->precision == RD_PREC_MIXED && d->tuning.pair_mode == RD_PAIRS_SYMMETRIC;
    std::vector<std::pair<uint32_t, WorkItem>> sym_items;   // (tiles, item)
    d->sym_R = 1;
    for (auto& o : order) {
        const uint32_t k = o.second;
        const HostDisk& h = d->hd[k];
        DiskInfo& di = d->h_info[k];
        for (int cls = 0; cls < 3; ++cls) {
            uint32_t b[4];
            class_bounds(h, cls, b);
            di.nchunk[cls] = 0;
            if (b[1] <= b[0]) continue;
            const uint32_t nsink = b[1] - b[0];
            pairs += (double)nsink * (b[3] - b[2]) - (cls == 0 ? nsink : 0);
            if (d->precision == RD_PREC_FP64) {
                di.nchunk[cls] = 1;
                for (uint32_t s = b[0]; s < b[1]; s += exact_group)
                    items.push_back(WorkItem{k, h.off + s, h.off + std::min(b[1], s + exact_group), h.off + b[2],
                                             h.off + b[3], 0, 0, 0});
            } else {
                // symmetric: SI x SI pairs once each (the FP64 prefix takes part
                // with FP32 mass 0: its own sink sums are kept, it pulls no one);
                // the SI sinks' NSI sources stay one-sided, in slots after them
                uint32_t base = 0, lo = std::max(b[2], di.n_fp64);
                if (sym && cls == 0) {
                    const uint32_t nblk = (b[1] + kSymBlockBodies - 1) / kSymBlockBodies;
                    // kSymChunks..2 kSymChunks - 1 chunks once the disk is that large:
                    // enough items to fill the device, few partial-sum slots; at least
                    // sym_chunk_min blocks (many small disks fill the device by themselves)
                    const uint32_t rmin = std::max<uint32_t>(1, d->tuning.sym_chunk_min);
                    const uint32_t R = std::min<uint32_t>(kSymMaxR, std::max<uint32_t>(rmin, nblk / kSymChunks));
                    const uint32_t nc = (nblk + R - 1) / R, span = R * kSymBlockBodies;
                    d->sym_R = std::max(d->sym_R, R);
                    // a decision of the disk alone: results never depend on the batch
                    const uint32_t self = nblk <= kSymSelfMaxBlocks ? kItemSelf : 0u;
                    for (uint32_t a = 0; a < nc; ++a)
                        for (uint32_t c = a; c < nc; ++c) {
                            const uint32_t ia = std::min(nblk - a * R, R), ic = std::min(nblk - c * R, R);
                            const uint32_t tiles = a == c ? ia * (ia + 1) / 2 : ia * ic;
                            sym_items.push_back({tiles, WorkItem{k, h.off + a * span, h.off + std::min(b[1], (a + 1) * span),
                                                                 h.off + c * span, h.off + std::min(b[1], (c + 1) * span),
                                                                 c, a, a == c ? self : 0u}});
                        }
                    base = nc;
                    lo = std::max(lo, b[1]);
                }
                const uint32_t hi = b[3];
                const uint32_t nsrc = hi > lo ? hi - lo : 0;
                const uint32_t nch = (nsrc + C - 1) / C;
                di.nchunk[cls] = base + nch;
                max_chunks = std::max<size_t>(max_chunks, base + nch);
                const uint32_t G = 64u * S;
                for (uint32_t s = b[0]; s < b[1]; s += G)
                    for (uint32_t c = 0; c < nch; ++c)
                        items.push_back(WorkItem{k, h.off + s, h.off + std::min(b[1], s + G),
                                                 h.off + lo + c * C, h.off + std::min(hi, lo + (c + 1) * C), base + c,
                                                 0, 0});
            }
        }
    }
    d->pairs_per_eval = pairs;
    // proposer-feature work: the pairs j > i in (128 sinks) x (source chunk)
    // tiles, the chunk cut so that the launch has about 16 items per CU
    std::vector<WorkItem> feat;
    {
        const uint32_t G = 128;   // 2 sinks per lane x 64 lanes (k_feat_pairs)
        size_t groups = 0;
        for (auto& o : order) groups += (d->hd[o.second].n + G - 1) / G;
        const size_t want = 16 * (size_t)d->n_cu;
        for (auto& o : order) {
            const HostDisk& h = d->hd[o.second];
            if (h.n == 0) continue;
            const uint32_t max_ch = (h.n + kTile - 1) / kTile;
            const uint32_t nch = (uint32_t)std::min<size_t>(max_ch, std::max<size_t>(1, 2 * want / groups));
            const uint32_t C = ((h.n + nch - 1) / nch + kTile - 1) / kTile * kTile;
            for (uint32_t s = 0; s < h.n; s += G)
                for (uint32_t j = 0; j < h.n; j += C) {
                    const uint32_t j_hi = std::min(h.n, j + C);
                    if (j_hi <= s + 1) continue;   // no pair j > i in this tile
                    const uint32_t j_lo = std::max(j, (s + 1) / kTile * kTile);
                    feat.push_back(WorkItem{o.second, h.off + s, h.off + std::min(h.n, s + G), h.off + j_lo,
                                            h.off + j_hi, 0, 0, 0});
                }
        }
    }
    // largest items first; ties keep the disk order
    std::stable_sort(sym_items.begin(), sym_items.end(), [](auto& a, auto& b) { return a.first > b.first; });
    d->sym_self = false;
    for (auto& q : sym_items) d->sym_self |= (q.second.flags & kItemSelf) != 0;
    std::vector<WorkItem> symv;
    for (auto& q : sym_items) symv.push_back(q.second);
    dev_free(d->items_force);
    dev_free(d->items_sym);
    dev_free(d->items_feat);
    if (int rc = dev_alloc(&d->items_sym, symv.size())) return rc;
    if (int rc = dev_alloc(&d->items_force, items.size())) return rc;
    if (int rc = dev_alloc(&d->items_feat, feat.size())) return rc;
    d->n_items_force = (uint32_t)items.size();
    d->n_items_sym = (uint32_t)symv.size();
    if (!symv.empty())
        HIP_TRY(hipMemcpy(d->items_sym, symv.data(), symv.size() * sizeof(WorkItem), hipMemcpyHostToDevice));
    d->n_items_feat = (uint32_t)feat.size();
    if (!feat.empty())
        HIP_TRY(hipMemcpy(d->items_feat, feat.data(), feat.size() * sizeof(WorkItem), hipMemcpyHostToDevice));
    if (!items.empty())
        HIP_TRY(hipMemcpy(d->items_force, items.data(), items.size() * sizeof(WorkItem), hipMemcpyHostToDevice));
    const uint32_t maxgrid = std::max<uint32_t>(1, d->tuning.blocks_per_cu) * (uint32_t)d->n_cu;
    // the exact kernel's blocks use 26 KB of LDS: 6 fit on a CU
    const uint32_t grid_cap = d->precision == RD_PREC_FP64 && !d->exact_lane ? 6u * (uint32_t)d->n_cu : maxgrid;
    d->grid_force = std::max<uint32_t>(1, std::min<uint32_t>(d->n_items_force, grid_cap));
    d->grid_sym = std::max<uint32_t>(1, std::min<uint32_t>(d->n_items_sym, maxgrid));
    const size_t need = max_chunks * 3 * (size_t)d->n_all;
    if (need > d->part_len) {
        dev_free(d->part);
        d->part_len = 0;   // a failed allocation leaves no stale size behind
        if (int rc = dev_alloc(&d->part, need)) return rc;
        d->part_len = need;
        HIP_TRY(hipMemsetAsync(d->part, 0, need * sizeof(double), d->stream));
    }
    HIP_TRY(hipMemcpy(d->info, d->h_info.data(), d->n_disk * sizeof(DiskInfo), hipMemcpyHostToDevice));
    HIP_TRY(launch_set_sources(d->stream, d->n_all, d->mass, d->id, d->disk_of, d->info, d->m_ex, d->m32));
    HIP_TRY(hipStreamSynchronize(d->stream));
    d->lists_valid = true;
    d->generation++;
    return RD_OK;
}

int ensure_lists(rd_disk* d) { return d->lists_valid ? RD_OK : build_lists(d); }

int upload_disk_of(rd_disk* d)
{
    std::vector<uint32_t> dof(d->n_all);
    for (uint32_t k = 0; k < d->n_disk; ++k)
        for (uint32_t i = 0; i < d->hd[k].n; ++i) dof[d->hd[k].off + i] = k;
    if (d->n_all) HIP_TRY(hipMemcpy(d->disk_of, dof.data(), d->n_all * sizeof(uint32_t), hipMemcpyHostToDevice));
    return RD_OK;
}

// --------------------------------------------------------------------------
// launch helpers
// --------------------------------------------------------------------------
StageArgs make_args(rd_disk* d, int r, int adaptive, int err_all)
{
    StageArgs a{};
    a.pos[0] = d->x; a.pos[1] = d->y; a.pos[2] = d->z;
    a.vel[0] = d->vx; a.vel[1] = d->vy; a.vel[2] = d->vz;
    a.out[0] = d->xo; a.out[1] = d->yo; a.out[2] = d->zo; a.out[3] = d->vxo; a.out[4] = d->vyo; a.out[5] = d->vzo;
    for (int c = 0; c < 3; ++c) {
        a.fpos[c] = r == 0 ? a.pos[c] : d->P[r & 1][c];
        a.npos[c] = d->P[(r + 1) & 1][c];
    }
    a.K = d->K;
    a.part = d->part;
    a.n_all = d->n_all;
    a.err_all = err_all;
    a.disk_of = d->disk_of;
    a.info = d->info;
    a.ctl = d->ctl;
    a.err = d->err;
    a.errf = d->err + d->n_disk;
    a.mass = d->mass; a.radius = d->radius; a.density = d->density; a.cd = d->cd;
    a.id = d->id;
    a.gas = d->gas;
    a.mixed = d->precision == RD_PREC_MIXED;
    a.adaptive = adaptive;
    a.commit = -1;
    a.one_disk = d->n_disk == 1;
    if (a.one_disk && !d->h_info.empty()) a.info0 = d->h_info[0];
    a.trash = d->trash;
    a.errw = d->errw;
    return a;
}

hipError_t enqueue_force(rd_disk* d, const double* fx, const double* fy, const double* fz, uint32_t need)
{
    if (d->precision == RD_PREC_FP64)
        return launch_force_exact(d->stream, d->exact_lane ? 1 : (d->exact_small ? 2 : 0), d->items_force,
                                  d->n_items_force, d->grid_force,
                                  fx, fy, fz, d->m_ex, d->ctl, need, d->part, d->n_all);
    // mixed precision: FP32 slot totals, in float units of the same buffer
    float* part32 = reinterpret_cast<float*>(d->part);
    if (hipError_t e = launch_force_sym(d->stream, d->sym_R, d->sym_self, d->items_sym, d->n_items_sym, d->grid_sym,
                                        fx, fy, fz, d->m32, d->ctl, need, part32, d->n_all))
        return e;
    return launch_force_mixed(d->stream, (int)d->tuning.sinks_per_lane, d->items_force, d->n_items_force,
                              d->grid_force, fx, fy, fz, d->m32, d->ctl, need, part32, d->n_all);
}

int last_stage(const rd_integrator* it) { return it->n_stage - 1; }

// Uploads ctl with F_CHECK set on the disks in `on` (t and h kept).
int flag_check_only(rd_disk* d, const std::vector<char>& on)
{
    for (uint32_t q = 0; q < d->n_disk; ++q) {
        d->h_ctl[q].t = d->hd[q].t;
        d->h_ctl[q].flags = on[q] ? F_CHECK : 0u;
    }
    HIP_TRY(hipMemcpyAsync(d->ctl, d->h_ctl, d->n_disk * sizeof(DiskCtl), hipMemcpyHostToDevice, d->stream));
    return RD_OK;
}

// One attempt for every disk whose ctl flags say so; with commit_advance >= 0
// (fixed-step chains) the commit follows in the same sequence.
hipError_t enqueue_attempt(rd_integrator* it, int commit_advance)
{
    rd_disk* d = it->d;
    hipError_t e;
    // FP64, one small gas-free disk (4-sink items: launches, not pairs, set
    // the time), fixed-step RK4: each stage is done by the summing lanes of
    // its FP64 pair pass (launch_force_exact_rk4), which hold every body's
    // complete sum; the same operations in the same order as pair pass + stage
    // kernel, one launch per stage instead of two
    if (d->precision == RD_PREC_FP64 && !d->exact_lane && d->exact_small && d->n_disk == 1 &&
        d->gas.on == kGasNone && it->method == RD_RK4 && !it->adaptive) {
        // the last stage also commits (k_commit's copy and time advance): its
        // summing lanes cover every body, and no pair term of that launch reads y_n
        for (int r = 0; r <= last_stage(it); ++r) {
            StageArgs a = make_args(d, r, it->adaptive, it->err_all);
            a.fuse_comb1 = r == 0;   // stage 0 forms stage 1 itself, as in the separate kernel
            if (r == last_stage(it)) a.commit = commit_advance;
            if ((e = launch_force_exact_rk4(d->stream, r, d->items_force, d->n_items_force, d->grid_force,
                                            a.fpos[0], a.fpos[1], a.fpos[2], d->m_ex, d->ctl,
                                            r == 0 ? F_K0_FORCE : F_ACTIVE, a)) != hipSuccess)
                return e;
        }
        return hipSuccess;
    }
    {
        StageArgs a = make_args(d, 0, it->adaptive, it->err_all);
        // fixed-step RK4 never retries with k0 kept, so stage 0 forms stage 1 itself
        a.fuse_comb1 = it->method == RD_RK4 && !it->adaptive;
        if ((e = enqueue_force(d, d->x, d->y, d->z, F_K0_FORCE)) != hipSuccess) return e;
        if ((e = launch_stage(d->stream, it->method, 0, a)) != hipSuccess) return e;
        if (it->n_stage > 1 && !a.fuse_comb1) {
            StageArgs a1 = make_args(d, 0, it->adaptive, it->err_all);   // comb1 writes P[1]
            if ((e = launch_comb1(d->stream, it->method, a1)) != hipSuccess) return e;
        }
    }
    for (int r = 1; r <= last_stage(it); ++r) {
        StageArgs a = make_args(d, r, it->adaptive, it->err_all);
        if ((e = enqueue_force(d, a.fpos[0], a.fpos[1], a.fpos[2], F_ACTIVE)) != hipSuccess) return e;
        if ((e = launch_stage(d->stream, it->method, r, a)) != hipSuccess) return e;
    }
    if (commit_advance >= 0) {
        StageArgs a = make_args(d, 0, it->adaptive, it->err_all);
        if ((e = launch_commit(d->stream, a, commit_advance, d->ctl)) != hipSuccess) return e;
    }
    return hipSuccess;
}

int run_attempt(rd_integrator* it, bool use_graph, int commit_advance = -1)
{
    rd_disk* d = it->d;
    if (use_graph) {
        const int kind = ((it->method * 16 + it->adaptive * 2 + it->err_all) * 4) + (commit_advance + 1);
        if (!it->gexec || it->graph_gen != d->generation || it->graph_kind != kind) {
            release_graph(it);
            HIP_TRY(hipStreamBeginCapture(d->stream, hipStreamCaptureModeThreadLocal));
            hipError_t e = enqueue_attempt(it, commit_advance);
            hipGraph_t g = nullptr;
            hipError_t e2 = hipStreamEndCapture(d->stream, &g);
            if (e != hipSuccess || e2 != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                if (e != hipSuccess) return fail(RD_EDEVICE, std::string("capture: ") + hipGetErrorString(e));
                return fail(RD_EDEVICE, std::string("end capture: ") + hipGetErrorString(e2));
            }
            it->graph = g;
            HIP_TRY(hipGraphInstantiate(&it->gexec, g, nullptr, nullptr, 0));
            it->graph_gen = d->generation;
            it->graph_kind = kind;
        }
        HIP_TRY(hipGraphLaunch(it->gexec, d->stream));
    } else {
        HIP_TRY(enqueue_attempt(it, commit_advance));
    }
    return RD_OK;
}

}  // namespace

// ============================================================================
// C ABI
// ============================================================================
extern "C" {

const char* rd_last_error(void) { return g_err.c_str(); }
int rd_version(void) { return RD_ABI_VERSION; }

int rd_device_count(int* n)
{
    if (!n) return fail(RD_EINVAL, "n is NULL");
    HIP_TRY(hipGetDeviceCount(n));
    return RD_OK;
}

// Each interaction class -- SI (star, giant, rocky, protoplanet), NSI
// (super-planetesimal, planetesimal), NI (test particle) -- must be one
// contiguous block, in that class order, and the bodies' types must count up
// to n_by_type.  Within a class the order is free: the reference's bounds are
// class ranges (red_type.h:670-700), and its own solar-system generator
// writes the rocky planets before the giants (src/red.cuda.initial/main.cu:
// 85-249).
static int validate_disk(const uint32_t* nbt, const rd_body_md* bmd, uint32_t n)
{
    uint32_t tot = 0;
    for (int t = 0; t < kNType; ++t) tot += nbt[t];
    if (tot != n) return fail(RD_EINVAL, "n_by_type does not sum to the body count");
    if (n == 0) return fail(RD_EINVAL, "a disk needs at least one body");
    uint32_t cnt[kNType] = {0};
    int prev = 0;
    for (uint32_t q = 0; q < n; ++q) {
        const int t = bmd[q].body_type;
        if (t < 0 || t >= kNType) return fail(RD_EINVAL, "invalid body_type (body " + std::to_string(q) + ")");
        const int cls = t <= RD_BODY_PROTOPLANET ? 0 : (t <= RD_BODY_PLANETESIMAL ? 1 : 2);
        if (cls < prev)
            return fail(RD_EINVAL, "bodies must be grouped by interaction class (SI, NSI, NI), in that order (body " +
                                       std::to_string(q) + ")");
        prev = cls;
        cnt[t]++;
    }
    for (int t = 0; t < kNType; ++t)
        if (cnt[t] != nbt[t])
            return fail(RD_EINVAL, "n_by_type does not match the body types (type " + std::to_string(t) + ")");
    for (uint32_t q = 0; q < n; ++q)
        if (bmd[q].id == 0) return fail(RD_EINVAL, "body id 0 is not allowed (ids are > 0, negative = inactive)");
    return RD_OK;
}

int rd_batch_create(uint32_t n_disk, const uint32_t* n_by_type, const rd_var4* r, const rd_var4* v,
                    const rd_param* p, const rd_body_md* bmd, const double* t, const double threshold[3],
                    const rd_gas_params* gas, int device, rd_disk** out)
{
    DeviceScope dev_scope(device);
    if (!out || !n_by_type || !r || !v || !p || !bmd || n_disk == 0) return fail(RD_EINVAL, "NULL argument");
    *out = nullptr;
    rd_disk* d = new rd_disk();
    d->n_disk = n_disk;
    d->hd.resize(n_disk);
    uint32_t off = 0;
    for (uint32_t k = 0; k < n_disk; ++k) {
        HostDisk& h = d->hd[k];
        uint32_t n = 0;
        for (int q = 0; q < kNType; ++q) {
            h.playing[q] = n_by_type[k * kNType + q];
            n += h.playing[q];
        }
        if (int rc = validate_disk(h.playing, bmd + off, n)) {
            delete d;
            return rc;
        }
        h.off = off;
        h.n = n;
        h.t = t ? t[k] : 0.0;
        off += n;
    }
    d->n_all = off;
    d->bmd.assign(bmd, bmd + off);
    if (threshold) std::memcpy(d->thr, threshold, sizeof(d->thr));
    if (gas) {
        d->gas.on = d->gas_created = kGasAnalytic;
        d->gas.rho0 = gas->rho[0]; d->gas.rho1 = gas->rho[1];
        d->gas.sch0 = gas->sch[0]; d->gas.sch1 = gas->sch[1];
        d->gas.eta0 = gas->eta[0]; d->gas.eta1 = gas->eta[1];
        d->gas.t0 = gas->t0; d->gas.t1 = gas->t1; d->gas.e_folding_time = gas->e_folding_time;
        d->gas.gas_decrease = gas->gas_decrease;
    }
    d->device = device;
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) { delete d; return fail(RD_EDEVICE, std::string("hipSetDevice: ") + hipGetErrorString(e)); }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        d->n_cu = prop.multiProcessorCount;
    e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete d; return fail(RD_EDEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e)); }
    int rc = alloc_buffers(d);
    if (!rc) rc = upload_disk_of(d);
    if (!rc) rc = upload_bodies(d, 0, d->n_all, r, v, p, bmd);
    if (!rc) rc = build_lists(d);
    if (rc) {
        std::string keep = g_err;
        free_buffers(d);
        (void)hipStreamDestroy(d->stream);
        delete d;
        g_err = keep;
        return rc;
    }
    *out = d;
    return RD_OK;
}

int rd_disk_create(uint32_t n_total, const uint32_t n_by_type[RD_BODY_TYPE_N], const rd_var4* r, const rd_var4* v,
                   const rd_param* p, const rd_body_md* bmd, double t, const double threshold[3],
                   const rd_gas_params* gas, int device, rd_disk** out)
{
    if (!n_by_type) return fail(RD_EINVAL, "n_by_type is NULL");
    uint32_t tot = 0;
    for (int q = 0; q < kNType; ++q) tot += n_by_type[q];
    if (tot != n_total) return fail(RD_EINVAL, "n_by_type does not sum to n_total");
    return rd_batch_create(1, n_by_type, r, v, p, bmd, &t, threshold, gas, device, out);
}

int rd_disk_destroy(rd_disk* d)
{
    DeviceScope dev_scope(d ? d->device : -1);
    if (!d) return RD_OK;
    // integrators that outlive the disk keep no graph and no disk pointer
    for (rd_integrator* it : d->integrators) {
        release_graph(it);
        it->d = nullptr;
    }
    d->integrators.clear();
    (void)hipStreamSynchronize(d->stream);
    free_buffers(d);
    if (d->ev_err) (void)hipEventDestroy(d->ev_err);
    (void)hipStreamDestroy(d->stream);
    delete d;
    return RD_OK;
}

int rd_disk_set_precision(rd_disk* d, int precision)
{
    DeviceScope dev_scope(d ? d->device : -1);
    if (!d || (precision != RD_PREC_FP64 && precision != RD_PREC_MIXED)) return fail(RD_EINVAL, "bad precision");
    d->precision = precision;
    d->lists_valid = false;
    return build_lists(d);
}

int rd_disk_set_drag_mode(rd_disk* d, int mode)
{
    DeviceScope dev_scope(d ? d->device : -1);
    if (!d || (mode != RD_DRAG_NSI && mode != RD_DRAG_ALL_MASSIVE)) return fail(RD_EINVAL, "bad drag mode");
    d->drag_mode = mode;
    d->lists_valid = false;
    return build_lists(d);
}

int rd_disk_set_gas_grid(rd_disk* d, const rd_gas_grid* g)
{
    DeviceScope dev_scope(d ? d->device : -1);
    if (!d) return fail(RD_EINVAL, "NULL disk");
    if (g) {
        if (!g->used_rad || !g->density || !g->vrad || !g->vtheta) return fail(RD_EINVAL, "NULL grid array");
        if (g->n_rad == 0 || g->n_sec == 0) return fail(RD_EINVAL, "the grid needs n_rad >= 1 and n_sec >= 1");
        if ((uint64_t)g->n_rad * g->n_sec > (uint64_t)INT32_MAX)
            return fail(RD_EINVAL, "the grid has more than 2^31 - 1 cells");
        for (uint32_t k = 0; k <= g->n_rad; ++k) {
            if (!std::isfinite(g->used_rad[k])) return fail(RD_EINVAL, "used_rad holds a value that is not finite");
            if (k > 0 && g->used_rad[k] < g->used_rad[k - 1]) return fail(RD_EINVAL, "used_rad is not increasing");
        }
    }
    // kernels queued on the disk's stream may still read the old grid
    HIP_TRY(hipStreamSynchronize(d->stream));
    dev_free(d->grid);
    d->gas.g_rad = nullptr;
    d->gas.g_cell = nullptr;
    d->gas.n_rad = d->gas.n_sec = 0;
    d->gas.on = d->gas_created;
    d->generation++;
    if (!g) return RD_OK;
    // one block: the cell records (32 B each, 32-byte aligned), then the radii
    const size_t nr = (size_t)g->n_rad + 1, nc = (size_t)g->n_rad * g->n_sec;
    std::vector<double> h(4 * nc + nr);
    for (size_t c = 0; c < nc; ++c) {
        h[4 * c + 0] = g->density[c];
        h[4 * c + 1] = g->vrad[c];
        h[4 * c + 2] = g->vtheta[c];
        h[4 * c + 3] = 0.0;
    }
    std::memcpy(h.data() + 4 * nc, g->used_rad, nr * sizeof(double));
    if (int rc = dev_alloc(&d->grid, h.size())) return rc;
    HIP_TRY(hipMemcpy(d->grid, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    d->gas.g_cell = reinterpret_cast<const double4*>(d->grid);
    d->gas.g_rad = d->grid + 4 * nc;
    d->gas.n_rad = g->n_rad;
    d->gas.n_sec = g->n_sec;
    d->gas.on = kGasGrid;
    return RD_OK;
}

int rd_disk_set_tuning(rd_disk* d, const rd_tuning* t)
{
    DeviceScope dev_scope(d ? d->device : -1);
    if (!d || !t) return fail(RD_EINVAL, "NULL argument");
    if (t->sinks_per_lane != 1 && t->sinks_per_lane != 2 && t->sinks_per_lane != 4)
        return fail(RD_EINVAL, "sinks_per_lane must be 1, 2 or 4");
    if (t->chunk < (uint32_t)kTile || t->chunk % kTile) return fail(RD_EINVAL, "chunk must be a multiple of 256");
    if (t->blocks_per_cu < 1 || t->blocks_per_cu > 64) return fail(RD_EINVAL, "blocks_per_cu must be in [1, 64]");
    if (t->pair_mode != RD_PAIRS_ONE_SIDED && t->pair_mode != RD_PAIRS_SYMMETRIC)
        return fail(RD_EINVAL, "pair_mode must be RD_PAIRS_ONE_SIDED or RD_PAIRS_SYMMETRIC");
    if (t->sym_chunk_min > kSymMaxR) return fail(RD_EINVAL, "sym_chunk_min must be in [0, 8]");
    d->tuning = *t;
    if (d->tuning.sym_chunk_min == 0) d->tuning.sym_chunk_min = 1;
    d->lists_valid = false;
    return build_lists(d);
}

int rd_disk_get_tuning(rd_disk* d, rd_tuning* t)
{
    DeviceScope dev_scope(d ? d->device : -1);
    if (!d || !t) return fail(RD_EINVAL, "NULL argument");
    *t = d->tuning;
    return RD_OK;
}

int rd_disk_count(rd_disk* d, uint32_t* n)
{
    DeviceScope dev_scope(d ? d->device : -1);
    if (!d || !n) return fail(RD_EINVAL, "NULL argument");
    *n = d->n_disk;
    return RD_OK;
}

int rd_disk_n_total(rd_disk* d, uint32_t k, uint32_t* n)
{
    DeviceScope dev_scope(d ? d->device : -1);
    if (!d || !n || k >= d->n_disk) return fail(RD_EINVAL, "bad argument");
    *n = d->hd[k].n;
    return RD_OK;
}

int rd_disk_n_objects(rd_disk* d, uint32_t k, uint32_t playing[RD_BODY_TYPE_N], uint32_t inactive[RD_BODY_TYPE_N],
                      uint32_t removed[RD_BODY_TYPE_N])
{
    DeviceScope dev_scope(d ? d->device : -1);
    if (!d || k >= d->n_disk) return fail(RD_EINVAL, "bad argument");
    if (playing) std::memcpy(playing, d->hd[k].playing, sizeof(d->hd[k].playing));
    if (inactive) std::memcpy(inactive, d->hd[k].inactive, sizeof(d->hd[k].inactive));
    if (removed) std::memcpy(removed, d->hd[k].removed, sizeof(d->hd[k].removed));
    return RD_OK;
}

int rd_sync(rd_disk* d)
{
    DeviceScope dev_scope(d ? d->device : -1);
    if (!d) return fail(RD_EINVAL, "NULL disk");
    HIP_TRY(hipStreamSynchronize(d->stream));
    return RD_OK;
}

int rd_get_state(rd_disk* d, uint32_t k, double* t, rd_var4* r, rd_var4* v, rd_param* p, rd_body_md* bmd)
{
    DeviceScope dev_scope(d ? d->device : -1);
    if (!d || k >= d->n_disk) return fail(RD_EINVAL, "bad argument");
    const HostDisk& h = d->hd[k];
    if (t) *t = h.t;
    if (int rc = download_bodies(d, h.off, h.n, r, v, p)) return rc;
    if (bmd) std::memcpy(bmd, d->bmd.data() + h.off, h.n * sizeof(rd_body_md));
    return RD_OK;
}

int rd_set_state(rd_disk* d, uint32_t k, double t, const rd_var4* r, const rd_var4* v)
{
    DeviceScope dev_scope(d ? d->device : -1);
    if (!d || k >= d->n_disk) return fail(RD_EINVAL, "bad argument");
    d->hd[k].t = t;
    return upload_bodies(d, d->hd[k].off, d->hd[k].n, r, v, nullptr, nullptr);
}

// ---- operator ---------------------------------------------------------------
int rd_calc_dydx(rd_disk* d, uint32_t k, uint32_t i, uint32_t rr, double t, const rd_var4* r, const rd_var4* v,
                 rd_var4* dy)
{
    DeviceScope dev_scope(d ? d->device : -1);
    (void)rr;
    if (!d || k >= d->n_disk || !r || !v || !dy) return fail(RD_EINVAL, "bad argument");
    const HostDisk& h = d->hd[k];
    if (i == 0) {
        std::memcpy(dy, v, h.n * sizeof(rd_var4));
        return RD_OK;
    }
    if (i != 1) return fail(RD_EINVAL, "Parameter 'i' is out of range.");
    if (int rc = ensure_lists(d)) return rc;
    // positions into P[0], velocities into the stage-1 velocity slot
    std::vector<double> buf(h.n);
    double* dst[6] = {d->P[0][0], d->P[0][1], d->P[0][2], kv_ptr(d, 1, 0), kv_ptr(d, 1, 1), kv_ptr(d, 1, 2)};
    for (int c = 0; c < 6; ++c) {
        const rd_var4* src = c < 3 ? r : v;
        for (uint32_t q = 0; q < h.n; ++q) buf[q] = c % 3 == 0 ? src[q].x : (c % 3 == 1 ? src[q].y : src[q].z);
        HIP_TRY(hipMemcpyAsync(dst[c] + h.off, buf.data(), h.n * sizeof(double), hipMemcpyHostToDevice, d->stream));
        HIP_TRY(hipStreamSynchronize(d->stream));
    }
    for (uint32_t q = 0; q < d->n_disk; ++q) d->h_ctl[q] = DiskCtl{d->hd[q].t, 0.0, q == k ? F_NEED_K0 : 0u, 0};
    HIP_TRY(hipMemcpyAsync(d->ctl, d->h_ctl, d->n_disk * sizeof(DiskCtl), hipMemcpyHostToDevice, d->stream));
    HIP_TRY(enqueue_force(d, d->P[0][0], d->P[0][1], d->P[0][2], F_NEED_K0));
    StageArgs a = make_args(d, 2, 0, 0);   // fx = P[0]
    HIP_TRY(launch_operator(d->stream, a, t, k, ka_ptr(d, 0, 0), ka_ptr(d, 0, 1), ka_ptr(d, 0, 2), kv_ptr(d, 1, 0),
                            kv_ptr(d, 1, 1), kv_ptr(d, 1, 2)));
    std::vector<double> ax(h.n), ay(h.n), az(h.n);
    HIP_TRY(hipMemcpyAsync(ax.data(), ka_ptr(d, 0, 0) + h.off, h.n * sizeof(double), hipMemcpyDeviceToHost, d->stream));
    HIP_TRY(hipMemcpyAsync(ay.data(), ka_ptr(d, 0, 1) + h.off, h.n * sizeof(double), hipMemcpyDeviceToHost, d->stream));
    HIP_TRY(hipMemcpyAsync(az.data(), ka_ptr(d, 0, 2) + h.off, h.n * sizeof(double), hipMemcpyDeviceToHost, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    for (uint32_t q = 0; q < h.n; ++q) dy[q] = rd_var4{ax[q], ay[q], az[q], 0.0};
    return RD_OK;
}

// p holds n rd_var4 records in device memory of the disk's device
static int check_device_var4(const rd_disk* d, const void* p, uint32_t n, const char* what)
{
    hipPointerAttribute_t at{};
    if (hipPointerGetAttributes(&at, p) != hipSuccess) {
        (void)hipGetLastError();   // leave no error for the next call to report
        return fail(RD_EINVAL, std::string(what) + " is not device memory");
    }
    if ((at.type != hipMemoryTypeDevice && at.type != hipMemoryTypeManaged) || at.device != d->device)
        return fail(RD_EINVAL, std::string(what) + " is not device memory of the disk's device");
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, const_cast<void*>(p)) != hipSuccess) {
        (void)hipGetLastError();
        return fail(RD_EINVAL, std::string(what) + ": its allocation could not be found");
    }
    if ((const char*)p + (size_t)n * sizeof(rd_var4) > (const char*)base + size)
        return fail(RD_EINVAL, std::string(what) + " holds fewer records than the disk has bodies");
    return RD_OK;
}

int rd_calc_dydx_device(rd_disk* d, uint32_t k, uint32_t i, uint32_t rr, double t, const rd_var4* r,
                        const rd_var4* v, rd_var4* dy)
{
    DeviceScope dev_scope(d ? d->device : -1);
    (void)rr;
    if (!d || k >= d->n_disk || !r || !v || !dy) return fail(RD_EINVAL, "bad argument");
    const HostDisk& h = d->hd[k];
    if (i != 0 && i != 1) return fail(RD_EINVAL, "Parameter 'i' is out of range.");
    if (int rc = check_device_var4(d, v, h.n, "v")) return rc;
    if (int rc = check_device_var4(d, dy, h.n, "dy")) return rc;
    if (i == 0) {
        HIP_TRY(hipMemcpyAsync(dy, v, h.n * sizeof(rd_var4), hipMemcpyDeviceToDevice, d->stream));
        HIP_TRY(hipStreamSynchronize(d->stream));
        return RD_OK;
    }
    if (int rc = check_device_var4(d, r, h.n, "r")) return rc;
    if (int rc = ensure_lists(d)) return rc;
    // positions into P[0], velocities into the stage-1 velocity slot, as rd_calc_dydx
    double* const p[3] = {d->P[0][0] + h.off, d->P[0][1] + h.off, d->P[0][2] + h.off};
    double* const vel[3] = {kv_ptr(d, 1, 0) + h.off, kv_ptr(d, 1, 1) + h.off, kv_ptr(d, 1, 2) + h.off};
    HIP_TRY(launch_var4_split(d->stream, h.n, reinterpret_cast<const double*>(r), reinterpret_cast<const double*>(v),
                              p, vel));
    for (uint32_t q = 0; q < d->n_disk; ++q) d->h_ctl[q] = DiskCtl{d->hd[q].t, 0.0, q == k ? F_NEED_K0 : 0u, 0};
    HIP_TRY(hipMemcpyAsync(d->ctl, d->h_ctl, d->n_disk * sizeof(DiskCtl), hipMemcpyHostToDevice, d->stream));
    HIP_TRY(enqueue_force(d, d->P[0][0], d->P[0][1], d->P[0][2], F_NEED_K0));
    StageArgs a = make_args(d, 2, 0, 0);   // fx = P[0]
    HIP_TRY(launch_operator(d->stream, a, t, k, ka_ptr(d, 0, 0), ka_ptr(d, 0, 1), ka_ptr(d, 0, 2), kv_ptr(d, 1, 0),
                            kv_ptr(d, 1, 1), kv_ptr(d, 1, 2)));
    const double* const acc[3] = {ka_ptr(d, 0, 0) + h.off, ka_ptr(d, 0, 1) + h.off, ka_ptr(d, 0, 2) + h.off};
    HIP_TRY(launch_var4_join(d->stream, h.n, acc, reinterpret_cast<double*>(dy)));
    HIP_TRY(hipStreamSynchronize(d->stream));
    return RD_OK;
}

int rd_bench_force(rd_disk* d, uint32_t n_iter, double* ms, double* pairs)
{
    DeviceScope dev_scope(d ? d->device : -1);
    if (!d || n_iter == 0) return fail(RD_EINVAL, "bad argument");
    if (int rc = ensure_lists(d)) return rc;
    for (uint32_t q = 0; q < d->n_disk; ++q) d->h_ctl[q] = DiskCtl{d->hd[q].t, 0.0, F_NEED_K0 | F_ACTIVE, 0};
    HIP_TRY(hipMemcpyAsync(d->ctl, d->h_ctl, d->n_disk * sizeof(DiskCtl), hipMemcpyHostToDevice, d->stream));
    HIP_TRY(enqueue_force(d, d->x, d->y, d->z, F_NEED_K0));   // warm-up
    EventPair ev;
    HIP_TRY(ev.create());
    HIP_TRY(hipEventRecord(ev.e0, d->stream));
    for (uint32_t q = 0; q < n_iter; ++q) HIP_TRY(enqueue_force(d, d->x, d->y, d->z, F_NEED_K0));
    HIP_TRY(hipEventRecord(ev.e1, d->stream));
    float el = 0.0f;
    HIP_TRY(ev.elapsed(&el));
    if (ms) *ms = (double)el / n_iter;
    if (pairs) *pairs = d->pairs_per_eval;
    return RD_OK;
}

// Mean time of one pair-kernel launch followed by the stage kernel that folds
// its partial sums (stage 1 of the current method's tableau, h = 0 so the
// state is untouched).
static int time_force_stage(rd_disk* d, uint32_t n_iter, double* ms)
{
    for (uint32_t q = 0; q < d->n_disk; ++q) d->h_ctl[q] = DiskCtl{d->hd[q].t, 0.0, F_NEED_K0 | F_ACTIVE, 0};
    HIP_TRY(hipMemcpyAsync(d->ctl, d->h_ctl, d->n_disk * sizeof(DiskCtl), hipMemcpyHostToDevice, d->stream));
    StageArgs a = make_args(d, 1, 0, 0);
    a.fpos[0] = d->x; a.fpos[1] = d->y; a.fpos[2] = d->z;
    HIP_TRY(enqueue_force(d, d->x, d->y, d->z, F_NEED_K0));
    HIP_TRY(launch_stage(d->stream, RD_RKF78, 1, a));
    EventPair ev;
    HIP_TRY(ev.create());
    HIP_TRY(hipEventRecord(ev.e0, d->stream));
    for (uint32_t q = 0; q < n_iter; ++q) {
        HIP_TRY(enqueue_force(d, d->x, d->y, d->z, F_NEED_K0));
        HIP_TRY(launch_stage(d->stream, RD_RKF78, 1, a));
    }
    HIP_TRY(hipEventRecord(ev.e1, d->stream));
    float el = 0.0f;
    HIP_TRY(ev.elapsed(&el));
    *ms = (double)el / n_iter;
    return RD_OK;
}

int rd_bench_copy(int device, size_t bytes, uint32_t n_iter, double* ms)
{
    DeviceScope dev_scope(device);
    if (bytes < 32 || n_iter == 0) return fail(RD_EINVAL, "bad argument");
    HIP_TRY(hipSetDevice(device));
    struct Res {   // released on every return path
        hipStream_t st = nullptr;
        void *a = nullptr, *b = nullptr;
        ~Res()
        {
            if (st) (void)hipStreamSynchronize(st);
            if (a) (void)hipFree(a);
            if (b) (void)hipFree(b);
            if (st) (void)hipStreamDestroy(st);
        }
    } res;
    HIP_TRY(hipStreamCreateWithFlags(&res.st, hipStreamNonBlocking));
    HIP_TRY(hipMalloc(&res.a, bytes));
    HIP_TRY(hipMalloc(&res.b, bytes));
    HIP_TRY(hipMemsetAsync(res.a, 0, bytes, res.st));
    HIP_TRY(launch_copy(res.st, res.a, res.b, bytes));
    EventPair ev;
    HIP_TRY(ev.create());
    HIP_TRY(hipEventRecord(ev.e0, res.st));
    for (uint32_t q = 0; q < n_iter; ++q) HIP_TRY(launch_copy(res.st, res.a, res.b, bytes));
    HIP_TRY(hipEventRecord(ev.e1, res.st));
    float el = 0.0f;
    HIP_TRY(ev.elapsed(&el));
    if (ms) *ms = (double)el / n_iter;
    return RD_OK;
}

int rd_bench_stages(rd_disk* d, int method, uint32_t n_iter, double* ms)
{
    DeviceScope dev_scope(d ? d->device : -1);
    if (!d || n_iter == 0 || !ms || (method != RD_RKF78 && method != RD_RK4)) return fail(RD_EINVAL, "bad argument");
    if (int rc = ensure_lists(d)) return rc;
    const int n_stage = method == RD_RKF78 ? 13 : 5;
    // h = 0: every stage rewrites y_n's values, the state is untouched
    for (uint32_t q = 0; q < d->n_disk; ++q) d->h_ctl[q] = DiskCtl{d->hd[q].t, 0.0, F_NEED_K0 | F_ACTIVE, 0};
    HIP_TRY(hipMemcpyAsync(d->ctl, d->h_ctl, d->n_disk * sizeof(DiskCtl), hipMemcpyHostToDevice, d->stream));
    HIP_TRY(enqueue_force(d, d->x, d->y, d->z, F_NEED_K0));   // partial sums for the stages to fold
    auto attempt = [&]() -> hipError_t {
        const StageArgs a0 = make_args(d, 0, 1, 0);
        if (hipError_t e = launch_stage(d->stream, method, 0, a0)) return e;
        if (hipError_t e = launch_comb1(d->stream, method, a0)) return e;
        for (int r = 1; r < n_stage; ++r)
            if (hipError_t e = launch_stage(d->stream, method, r, make_args(d, r, 1, 0))) return e;
        return hipSuccess;
    };
    HIP_TRY(attempt());
    EventPair ev;
    HIP_TRY(ev.create());
    HIP_TRY(hipEventRecord(ev.e0, d->stream));
    for (uint32_t q = 0; q < n_iter; ++q) HIP_TRY(attempt());
    HIP_TRY(hipEventRecord(ev.e1, d->stream));
    float el = 0.0f;
    HIP_TRY(ev.elapsed(&el));
    *ms = (double)el / n_iter;
    return RD_OK;
}

int rd_bench_force_stage(rd_disk* d, uint32_t n_iter, double* ms)
{
    DeviceScope dev_scope(d ? d->device : -1);
    if (!d || n_iter == 0 || !ms) return fail(RD_EINVAL, "bad argument");
    if (int rc = ensure_lists(d)) return rc;
    return time_force_stage(d, n_iter, ms);
}

int rd_autotune(rd_disk* d, rd_tuning* best)
{
    DeviceScope dev_scope(d ? d->device : -1);
    if (!d) return fail(RD_EINVAL, "NULL disk");
    if (d->precision != RD_PREC_MIXED) {
        if (best) *best = d->tuning;
        return RD_OK;
    }
    const rd_tuning keep = d->tuning;
    rd_tuning win = keep;
    double best_ms = 1e300;
    const uint32_t S_opts[] = {1, 2, 4};
    const uint32_t C_opts[] = {512, 1024, 2048, 4096, 8192};
    const uint32_t B_opts[] = {1, 2, 4, 8};
    uint32_t max_n = 0;
    bool one_sided_rest = false;   // pairs the symmetric kernel leaves to the one-sided one
    bool small_disks = false;      // a disk whose chunking sym_chunk_min changes
    for (auto& h : d->hd) {
        max_n = std::max(max_n, h.n);
        one_sided_rest |= h.n_massive() > h.n_si() || h.n > h.n_massive();
        small_disks |= (h.n_si() + kSymBlockBodies - 1) / kSymBlockBodies < 2 * kSymChunks && h.n_si() > kSymBlockBodies;
    }
    auto trial = [&](const rd_tuning& t) -> int {
        d->tuning = t;
        if (int rc = build_lists(d)) { d->tuning = keep; build_lists(d); return rc; }
        double ms = 0.0;
        if (int rc = time_force_stage(d, 3, &ms)) { d->tuning = keep; build_lists(d); return rc; }
        if (ms < best_ms) { best_ms = ms; win = t; }
        return RD_OK;
    };
    for (uint32_t mode : {(uint32_t)RD_PAIRS_SYMMETRIC, (uint32_t)RD_PAIRS_ONE_SIDED})
        for (uint32_t S : S_opts)
            for (uint32_t C : C_opts) {
                if (C > 2 * std::max<uint32_t>(max_n, kTile) && C != 512) continue;
                // with every pair symmetric, S and C shape nothing: one pass
                if (mode == RD_PAIRS_SYMMETRIC && !one_sided_rest && (S != keep.sinks_per_lane || C != keep.chunk))
                    continue;
                for (uint32_t B : B_opts)
                    for (uint32_t M : {1u, 2u}) {
                        if (M != 1 && (mode != RD_PAIRS_SYMMETRIC || !small_disks)) continue;
                        rd_tuning t = keep;
                        t.sinks_per_lane = S;
                        t.chunk = C;
                        t.blocks_per_cu = B;
                        t.pair_mode = mode;
                        t.sym_chunk_min = M;
                        if (int rc = trial(t)) return rc;
                    }
            }
    d->tuning = win;
    if (int rc = build_lists(d)) return rc;
    if (best) *best = win;
    return RD_OK;
}

// ---- integrators ---------------------------------------------------------------
int rd_integrator_create(rd_disk* d, int method, double dt, int adaptive, double tol, rd_integrator** out)
{
    DeviceScope dev_scope(d ? d->device : -1);
    if (!d || !out) return fail(RD_EINVAL, "NULL argument");
    if (method != RD_EULER && method != RD_RK2 && method != RD_RK4 && method != RD_RKF78)
        return fail(RD_ENOTIMPL, "Euler, RK2, RK4 and RKF78 are built (RKF5(6) is not)");
    rd_integrator* it = new rd_integrator();
    it->d = d;
    it->device = d->device;
    it->method = method;
    it->adaptive = adaptive ? 1 : 0;
    it->tol = tol;
    if (method == RD_RKF78) { it->n_stage = adaptive ? 13 : 11; it->order = 7; }
    else if (method == RD_RK4) { it->n_stage = adaptive ? 5 : 4; it->order = 4; }
    else {
        // euler / rungekutta2 construct the base with adaptive = false and step
        // with the constructor's dt (int_euler.cu:48-54, int_rungekutta2.cu:49-55)
        it->adaptive = 0;
        it->fixed_dt = true;
        it->n_stage = method == RD_RK2 ? 2 : 1;
        it->order = method == RD_RK2 ? 2 : 1;
    }
    it->dt_try.assign(d->n_disk, dt);
    it->dt_did.assign(d->n_disk, 0.0);
    it->dt_next.assign(d->n_disk, dt);
    it->tried.assign(d->n_disk, 0);
    it->passed.assign(d->n_disk, 0);
    it->failed.assign(d->n_disk, 0);
    it->hist.assign(d->n_disk, rd_integrator::Hist{});
    d->integrators.push_back(it);
    *out = it;
    return RD_OK;
}

int rd_integrator_destroy(rd_integrator* it)
{
    // may run after the disk is gone (destroy order is the caller's; garbage
    // collectors pick one): the disk then already released the graph and
    // cleared it->d
    DeviceScope dev_scope(it ? it->device : -1);
    if (!it) return RD_OK;
    release_graph(it);
    if (it->d) {
        auto& v = it->d->integrators;
        v.erase(std::remove(v.begin(), v.end(), it), v.end());
    }
    delete it;
    return RD_OK;
}

// Features of the disks that just accepted a step (state y_{n+1}), the
// logged sample of the step just taken, and the proposer's first try for the
// next step.
static int after_accept(rd_integrator* it, const std::vector<char>& acc, const std::vector<double>& err_ratio,
                        const std::vector<int>& iters)
{
    rd_disk* d = it->d;
    const uint32_t nd = d->n_disk;
    for (uint32_t k = 0; k < nd; ++k) {
        if (!acc[k]) continue;
        rd_integrator::Hist& H = it->hist[k];
        // sample of the step just accepted: features at its start -> the step
        // that would just have met the tolerance
        if (it->logging && H.has_feat) {
            const double er = std::max(err_ratio[k], 1.0e-12);
            const double target = std::log(std::fabs(it->dt_did[k])) + std::log(1.0 / er) / 8.0;
            it->log_x.insert(it->log_x.end(), H.feat, H.feat + RD_MLP_IN);
            it->log_t.push_back((float)target);
        }
        H.h2 = H.accepted ? H.h1 : it->dt_did[k];
        H.e2 = H.accepted ? H.e1 : err_ratio[k];
        H.h1 = it->dt_did[k];
        H.e1 = err_ratio[k];
        H.rej = (uint32_t)(iters[k] - 1);
        H.accepted++;
    }
    if (!it->proposer && !it->logging) return RD_OK;
    if (int rc = flag_check_only(d, acc)) return rc;
    HIP_TRY(hipMemsetAsync(d->feat, 0xff, 3 * (size_t)nd * sizeof(uint32_t), d->stream));
    HIP_TRY(launch_features(d->stream, d->items_feat, d->n_items_feat,
                            std::min<uint32_t>(d->n_items_feat, 16u * (uint32_t)d->n_cu), d->n_all, nd, d->xo, d->yo,
                            d->zo, d->vxo, d->vyo, d->vzo, d->mass, d->id, d->disk_of, d->info, d->ctl, d->feat,
                            d->eterm, d->esum));
    HIP_TRY(hipMemcpyAsync(d->h_feat, d->feat, 3 * (size_t)nd * sizeof(uint32_t), hipMemcpyDeviceToHost, d->stream));
    HIP_TRY(hipMemcpyAsync(d->h_esum, d->esum, nd * sizeof(double), hipMemcpyDeviceToHost, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    const uint32_t* f = d->h_feat;
    const double* es = d->h_esum;
    std::vector<uint32_t> who;
    for (uint32_t k = 0; k < nd; ++k) {
        if (!acc[k]) continue;
        rd_integrator::Hist& H = it->hist[k];
        float r2v2, qmin;
        std::memcpy(&r2v2, &f[3 * k], 4);
        std::memcpy(&qmin, &f[3 * k + 1], 4);
        if (!(r2v2 < 1e38f)) r2v2 = 1e38f;
        if (!(qmin < 1e38f)) qmin = 1e38f;
        if (!H.ek0_set) {
            H.ek0 = es[k];
            H.ek0_set = true;
        }
        const double tenc = 0.5 * std::log(std::max((double)r2v2, 1e-60));
        // step sizes enter as magnitudes, so a backward run (dt < 0) sees the
        // features of the forward one; the sign goes back on the proposal
        float* F = H.feat;
        F[0] = (float)std::log(std::fabs(H.h1));
        F[1] = (float)std::log(std::max(H.e1, 1e-30));
        F[2] = (float)tenc;
        F[3] = (float)std::log(std::max((double)qmin, 1e-30));
        F[4] = (float)std::log((double)d->hd[k].n);
        F[5] = (float)std::log10(std::fabs(es[k] - H.ek0) / std::max(std::fabs(H.ek0), 1e-300) + 1e-20);
        F[6] = (float)std::log(std::fabs(H.h2));
        F[7] = (float)std::log(std::max(H.e2, 1e-30));
        F[8] = (float)H.rej;
        F[9] = (float)std::log(std::fabs(it->dt_next[k]));
        F[10] = (float)std::log(it->tol);
        F[11] = (float)(H.tenc_prev != 0.0 ? tenc - H.tenc_prev : 0.0);
        H.tenc_prev = tenc;
        H.has_feat = true;
        std::memcpy(d->h_fx + (size_t)RD_MLP_IN * who.size(), F, RD_MLP_IN * sizeof(float));
        who.push_back(k);
    }
    if (it->proposer && !who.empty()) {
        const uint32_t m = (uint32_t)who.size();
        HIP_TRY(hipMemcpyAsync(d->fx, d->h_fx, (size_t)m * RD_MLP_IN * sizeof(float), hipMemcpyHostToDevice,
                               d->stream));
        if (int rc = mlp_forward_dev(it->proposer, d->stream, d->fx, m, d->fy))
            return fail(rc, "proposer forward launch");
        HIP_TRY(hipMemcpyAsync(d->h_fy, d->fy, m * sizeof(float), hipMemcpyDeviceToHost, d->stream));
        HIP_TRY(hipStreamSynchronize(d->stream));
        for (uint32_t q = 0; q < m; ++q) {
            const uint32_t k = who[q];
            const double hc = it->dt_next[k];
            double hp = it->safety * std::exp((double)d->h_fy[q]);
            if (!(hp > 0.0) || !std::isfinite(hp)) continue;
            hp = std::min(std::max(hp, std::fabs(hc) / 8.0), 8.0 * std::fabs(hc));
            it->dt_next[k] = hc < 0.0 ? -hp : hp;
            it->proposals++;
        }
    }
    return RD_OK;
}

// Advance every disk in `want` by one accepted step (reference step(),
// int_rungekutta8.cu:635-716 / int_rungekutta4.cu:173-247). `want[k]` is the
// number of accepted steps disk k takes in this call. Every round makes one
// attempt for each disk that still owes steps, so a disk that accepts starts
// its next step in the next round instead of waiting for the disks that
// rejected; each disk's own sequence of attempts is the same either way.
// With `overlap` (rd_steps), an attempt is followed on the stream by the
// device's own accept decision (k_decide, the host's arithmetic), the commit
// of accepted disks and, for disks that go on, the next step's stage-0 pair
// pass; the host waits only for the error words, so its step-size control
// runs while that pass does.  The next attempt then skips its stage-0 pair
// pass for those disks (F_K0_FORCE clear).  The operations and their inputs
// are the same, so the results are too.
static int step_disks(rd_integrator* it, const std::vector<uint32_t>& want, bool overlap = false)
{
    rd_disk* d = it->d;
    if (int rc = ensure_lists(d)) return rc;
    const uint32_t nd = d->n_disk;
    const bool early = overlap && it->overlap && it->adaptive;
    if (early && !d->ev_err) HIP_TRY(hipEventCreateWithFlags(&d->ev_err, hipEventDisableTiming));
    std::vector<uint32_t> remaining(want);
    std::vector<char> pending(nd);
    for (uint32_t k = 0; k < nd; ++k) pending[k] = remaining[k] > 0;
    std::vector<int> iter(nd, 0);
    std::vector<char> need_k0(nd, 1);
    std::vector<char> k0_ahead(nd, 0);   // the stage-0 pair pass already ran on the device
    for (uint32_t k = 0; k < nd; ++k)
        if (pending[k] && !it->fixed_dt) it->dt_try[k] = it->dt_next[k];
    const double lambda = it->method == RD_RKF78 ? 41.0 / 840.0 : 1.0 / 10.0;
    uint32_t commit_flags_pending = 0;
    bool any_acc = false;
    std::vector<uint32_t> commit(nd, 0);
    std::vector<char> acc(nd, 0);
    std::vector<double> err_ratio(nd, 0.0);
    std::string guard_msg;   // first disk whose step the guard failed
    for (;;) {
        bool any = false;
        for (uint32_t k = 0; k < nd; ++k) {
            // opt-in guard: an attempt with |dt| <= min_dt is not made
            if (pending[k] && it->adaptive && 0.0 < it->min_dt && std::fabs(it->dt_try[k]) <= it->min_dt) {
                it->tried[k] += iter[k];
                it->failed[k] += iter[k];
                pending[k] = 0;
                remaining[k] = 0;
                if (guard_msg.empty())
                    guard_msg = "disk " + std::to_string(k) + ": The stepsize is smaller than the limit.";
            }
            uint32_t fl = commit[k];
            if (pending[k]) {
                fl |= F_ACTIVE | (need_k0[k] ? F_NEED_K0 : 0u) | (need_k0[k] && !k0_ahead[k] ? F_K0_FORCE : 0u) |
                      (early && remaining[k] > 1 ? F_MORE : 0u);
                it->dt_did[k] = it->dt_try[k];
                any = true;
            }
            k0_ahead[k] = 0;
            d->h_ctl[k] = DiskCtl{d->hd[k].t, it->dt_try[k], fl, 0};
        }
        HIP_TRY(hipMemcpyAsync(d->ctl, d->h_ctl, nd * sizeof(DiskCtl), hipMemcpyHostToDevice, d->stream));
        if (commit_flags_pending) {
            StageArgs a = make_args(d, 0, it->adaptive, it->err_all);
            HIP_TRY(launch_commit(d->stream, a, 0, d->ctl));
            commit_flags_pending = 0;
            std::fill(commit.begin(), commit.end(), 0u);
        }
        if (!any) break;
        if (int rc = run_attempt(it, d->tuning.use_graph != 0)) return rc;
        uint32_t n_new = 0, n_act = 0;
        for (uint32_t k = 0; k < nd; ++k) {
            if (!pending[k]) continue;
            n_act++;
            n_new += need_k0[k];
        }
        it->force_evals += (uint64_t)(last_stage(it)) + (n_new ? 1 : 0);
        const size_t nw = 3 * (((size_t)d->n_all + kBodyBlock - 1) / kBodyBlock);
        if (it->adaptive && nd == 1)   // single disk: one error slot per wave of the last stage
            HIP_TRY(hipMemcpyAsync(d->h_errw, d->errw, nw * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                   d->stream));
        else if (it->adaptive)   // max bits and non-finite flags of every disk
            HIP_TRY(hipMemcpyAsync(d->h_err, d->err, 2 * nd * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                   d->stream));
        if (early) {
            HIP_TRY(hipEventRecord(d->ev_err, d->stream));
            HIP_TRY(launch_decide(d->stream, nd == 1 ? d->errw : nullptr, (uint32_t)nw, d->err, nd, d->ctl, lambda,
                                  it->tol, it->reject_nonfinite ? 1 : 0));
            StageArgs ca = make_args(d, 0, it->adaptive, it->err_all);
            HIP_TRY(launch_commit(d->stream, ca, 0, d->ctl));
            HIP_TRY(enqueue_force(d, d->x, d->y, d->z, F_SPEC_K0));
            HIP_TRY(hipEventSynchronize(d->ev_err));
        } else {
            HIP_TRY(hipStreamSynchronize(d->stream));
        }
        if (it->adaptive && nd == 1) {
            unsigned long long m = 0, bad = 0;
            for (size_t w = 0; w < nw; ++w) {
                m = std::max(m, d->h_errw[w] & ~kErrNonFinite);
                bad |= d->h_errw[w] & kErrNonFinite;
            }
            d->h_err[0] = m;
            d->h_err[1] = bad;
        }
        // pair count of this attempt: active disks' full evaluation per stage
        for (uint32_t k = 0; k < nd; ++k) {
            if (!pending[k]) continue;
            const HostDisk& h = d->hd[k];
            double per = 0.0;
            for (int cls = 0; cls < 3; ++cls) {
                uint32_t b[4];
                class_bounds(h, cls, b);
                if (b[1] > b[0]) per += (double)(b[1] - b[0]) * (b[3] - b[2]) - (cls == 0 ? (b[1] - b[0]) : 0);
            }
            it->pairs += per * (last_stage(it) + (need_k0[k] ? 1 : 0));
            double max_err = 0.0;
            if (it->adaptive) {
                double raw;
                std::memcpy(&raw, &d->h_err[k], sizeof(double));
                max_err = std::fabs(lambda * it->dt_try[k] * raw);
                if (0.0 < max_err) it->dt_next[k] = it->dt_try[k] *= 0.9 * std::pow(it->tol / max_err, 1.0 / (it->order));
                else it->dt_next[k] = it->dt_try[k] *= 2.0;
            }
            iter[k]++;
            err_ratio[k] = it->adaptive ? max_err / it->tol : 1.0;
            // opt-in guard: the step of disk k fails instead of retrying
            // without end or accepting a non-finite estimate; y_n is kept
            const char* why = nullptr;
            if (it->adaptive && it->reject_nonfinite && (d->h_err[nd + k] || !std::isfinite(max_err)))
                why = "The error estimate is not finite.";
            else if (it->adaptive && max_err > it->tol && it->max_attempts && (uint32_t)iter[k] >= it->max_attempts)
                why = "The number of iteration exceeded the limit.";
            if (why) {
                it->tried[k] += iter[k];
                it->failed[k] += iter[k];
                pending[k] = 0;
                remaining[k] = 0;
                if (guard_msg.empty()) guard_msg = "disk " + std::to_string(k) + ": " + why;
                continue;
            }
            if (it->adaptive && max_err > it->tol) {
                need_k0[k] = 0;       // k0 is reused (int_rungekutta8.cu:648-656)
            } else {
                it->tried[k] += iter[k];
                it->failed[k] += iter[k] - 1;
                it->passed[k]++;
                if (it->fixed_dt) it->dt_next[k] = it->dt_try[k];
                d->hd[k].t += it->dt_did[k];
                pending[k] = 0;
                remaining[k]--;
                if (early) {
                    // committed on the device; its next k0 pass ran if it goes on
                    k0_ahead[k] = remaining[k] > 0;
                } else {
                    commit[k] = F_COMMIT;
                    commit_flags_pending = 1;
                }
                acc[k] = 1;
                any_acc = true;
            }
        }
        (void)n_act;
        const bool accepted_some = any_acc;
        any_acc = false;
        if (accepted_some && it->adaptive) {
            if (int rc = after_accept(it, acc, err_ratio, iter)) return rc;
            for (uint32_t k = 0; k < nd; ++k) {
                if (!acc[k] || remaining[k] == 0) continue;
                // next step of this disk: committed at the start of the round
                pending[k] = 1;
                need_k0[k] = 1;
                iter[k] = 0;
                it->dt_try[k] = it->dt_next[k];
            }
            std::fill(acc.begin(), acc.end(), 0);
        }
    }
    HIP_TRY(hipStreamSynchronize(d->stream));
    if (!guard_msg.empty()) return fail(RD_ESTATE, guard_msg);
    return RD_OK;
}

int rd_step(rd_integrator* it, double* dt_did)
{
    DeviceScope dev_scope(it ? it->device : -1);
    if (int rc = check_live(it)) return rc;
    std::vector<uint32_t> want(it->d->n_disk, 1u);
    if (int rc = step_disks(it, want)) return rc;
    if (dt_did) *dt_did = it->dt_did[0];
    return RD_OK;
}

int rd_steps(rd_integrator* it, uint32_t n_steps)
{
    DeviceScope dev_scope(it ? it->device : -1);
    if (int rc = check_live(it)) return rc;
    rd_disk* d = it->d;
    if (!it->adaptive) {
        // Fixed step: every attempt is accepted, so attempts are chained on the
        // device with the time advanced by the commit kernel.
        if (int rc = ensure_lists(d)) return rc;
        const uint32_t nd = d->n_disk;
        for (uint32_t k = 0; k < nd; ++k) {
            if (!it->fixed_dt) it->dt_try[k] = it->dt_next[k];
            it->dt_next[k] = it->dt_try[k];
            it->dt_did[k] = it->dt_try[k];
            d->h_ctl[k] = DiskCtl{d->hd[k].t, it->dt_try[k], F_ACTIVE | F_NEED_K0 | F_K0_FORCE | F_COMMIT, 0};
        }
        HIP_TRY(hipMemcpyAsync(d->ctl, d->h_ctl, nd * sizeof(DiskCtl), hipMemcpyHostToDevice, d->stream));
        for (uint32_t s = 0; s < n_steps; ++s) {
            // the attempt and its commit (t += h on the device) as one sequence
            if (int rc = run_attempt(it, d->tuning.use_graph != 0, 1)) return rc;
            it->force_evals += (uint64_t)it->n_stage;
            // at most 64 attempts (each a graph launch and a commit) queued
            // between host waits
            if ((s & 63u) == 63u) HIP_TRY(hipStreamSynchronize(d->stream));
        }
        for (uint32_t k = 0; k < nd; ++k) {
            const HostDisk& h = d->hd[k];
            double per = 0.0;
            for (int cls = 0; cls < 3; ++cls) {
                uint32_t b[4];
                class_bounds(h, cls, b);
                if (b[1] > b[0]) per += (double)(b[1] - b[0]) * (b[3] - b[2]) - (cls == 0 ? (b[1] - b[0]) : 0);
            }
            it->pairs += per * it->n_stage * n_steps;
            for (uint32_t s = 0; s < n_steps; ++s) d->hd[k].t += it->dt_did[k];
            it->tried[k] += n_steps;
            it->passed[k] += n_steps;
        }
        HIP_TRY(hipStreamSynchronize(d->stream));
        return RD_OK;
    }
    std::vector<uint32_t> want(d->n_disk, n_steps);
    return step_disks(it, want, true);
}

int rd_get_dt_next(rd_integrator* it, uint32_t k, double* dt)
{
    DeviceScope dev_scope(it ? it->device : -1);
    if (int rc = check_live(it)) return rc;
    if (!dt || k >= it->d->n_disk) return fail(RD_EINVAL, "bad argument");
    *dt = it->dt_next[k];
    return RD_OK;
}

int rd_set_dt_next(rd_integrator* it, uint32_t k, double dt)
{
    DeviceScope dev_scope(it ? it->device : -1);
    if (int rc = check_live(it)) return rc;
    if (k >= it->d->n_disk) return fail(RD_EINVAL, "bad argument");
    it->dt_next[k] = dt;
    return RD_OK;
}

int rd_counters(rd_integrator* it, uint32_t k, uint64_t* tried, uint64_t* passed, uint64_t* failed)
{
    DeviceScope dev_scope(it ? it->device : -1);
    if (int rc = check_live(it)) return rc;
    if (k >= it->d->n_disk) return fail(RD_EINVAL, "bad argument");
    if (tried) *tried = it->tried[k];
    if (passed) *passed = it->passed[k];
    if (failed) *failed = it->failed[k];
    return RD_OK;
}

int rd_set_error_check_for_tp(rd_integrator* it, int on)
{
    DeviceScope dev_scope(it ? it->device : -1);
    if (!it) return fail(RD_EINVAL, "NULL integrator");
    it->err_all = on ? 1 : 0;
    return RD_OK;
}

int rd_integrator_set_guard(rd_integrator* it, uint32_t max_attempts, double min_dt, int nonfinite)
{
    DeviceScope dev_scope(it ? it->device : -1);
    if (!it) return fail(RD_EINVAL, "NULL integrator");
    if (!(min_dt >= 0.0)) return fail(RD_EINVAL, "min_dt must be >= 0");
    it->max_attempts = max_attempts;
    it->min_dt = min_dt;
    it->reject_nonfinite = nonfinite != 0;
    return RD_OK;
}

int rd_integrator_set_overlap(rd_integrator* it, int on)
{
    if (!it) return fail(RD_EINVAL, "NULL integrator");
    it->overlap = on != 0;
    return RD_OK;
}

int rd_integrator_set_proposer(rd_integrator* it, rd_mlp* m, float safety)
{
    DeviceScope dev_scope(it ? it->device : -1);
    if (int rc = check_live(it)) return rc;
    if (!(safety > 0.0f)) return fail(RD_EINVAL, "bad argument");
    if (m && !it->adaptive) return fail(RD_ESTATE, "the proposer needs an adaptive integrator");
    if (m && mlp_device(m) != it->d->device) return fail(RD_EINVAL, "proposer and disk are on different devices");
    it->proposer = m;
    it->safety = safety;
    return RD_OK;
}

int rd_integrator_set_logging(rd_integrator* it, int on)
{
    DeviceScope dev_scope(it ? it->device : -1);
    if (!it) return fail(RD_EINVAL, "NULL integrator");
    if (on && !it->adaptive) return fail(RD_ESTATE, "logging needs an adaptive integrator");
    it->logging = on != 0;
    return RD_OK;
}

int rd_integrator_log(rd_integrator* it, float* x, float* t, uint32_t cap, uint32_t* n)
{
    DeviceScope dev_scope(it ? it->device : -1);
    if (!it) return fail(RD_EINVAL, "NULL integrator");
    const size_t cnt = it->log_t.size();
    const size_t m = std::min<size_t>(cap, cnt);
    if (x) std::memcpy(x, it->log_x.data(), m * RD_MLP_IN * sizeof(float));
    if (t) std::memcpy(t, it->log_t.data(), m * sizeof(float));
    if (n) *n = (uint32_t)cnt;
    return RD_OK;
}

int rd_integrator_clear_log(rd_integrator* it)
{
    DeviceScope dev_scope(it ? it->device : -1);
    if (!it) return fail(RD_EINVAL, "NULL integrator");
    it->log_x.clear();
    it->log_t.clear();
    return RD_OK;
}

int rd_work_counters(rd_integrator* it, uint64_t* evals, double* pairs)
{
    DeviceScope dev_scope(it ? it->device : -1);
    if (!it) return fail(RD_EINVAL, "NULL integrator");
    if (evals) *evals = it->force_evals;
    if (pairs) *pairs = it->pairs;
    return RD_OK;
}

}  // extern "C"

// ============================================================================
// events, integrals, driver loop
// ============================================================================
namespace {

struct HostView {
    std::vector<rd_var4> r, v;
    std::vector<rd_param> p;
};

int fetch(rd_disk* d, uint32_t k, HostView& hv)
{
    const HostDisk& h = d->hd[k];
    hv.r.resize(h.n);
    hv.v.resize(h.n);
    hv.p.resize(h.n);
    return download_bodies(d, h.off, h.n, hv.r.data(), hv.v.data(), hv.p.data());
}

// The records (r, v, p) of the few bodies an event touches, by local index.
// Events move only these over PCIe: one gather kernel and one copy out, one
// copy in and one scatter kernel back (the reference copies the three records
// of each merge, pp_disk.cu:1358-1363, and one metadata record, 1240).
struct BodyRec {
    rd_var4 r, v;
    rd_param p;
};
struct SparseView {
    std::vector<uint32_t> idx;   // local indices, ascending
    std::vector<BodyRec> rec;
    BodyRec& at(uint32_t i) { return rec[std::lower_bound(idx.begin(), idx.end(), i) - idx.begin()]; }
    const BodyRec& at(uint32_t i) const { return rec[std::lower_bound(idx.begin(), idx.end(), i) - idx.begin()]; }
};

BodyArrays body_arrays(rd_disk* d)
{
    return BodyArrays{{d->x, d->y, d->z, d->vx, d->vy, d->vz, d->mass, d->radius, d->density, d->cd}};
}

int ensure_stage(rd_disk* d, size_t n)
{
    if (n <= d->ev_scap) return RD_OK;
    const uint32_t cap = (uint32_t)std::max<size_t>(256, 2 * n);
    dev_free(d->ev_sidx);
    dev_free(d->ev_sids);
    dev_free(d->ev_srec);
    d->ev_scap = 0;
    if (int rc = dev_alloc(&d->ev_sidx, cap)) return rc;
    if (int rc = dev_alloc(&d->ev_sids, cap)) return rc;
    if (int rc = dev_alloc(&d->ev_srec, (size_t)kBodyRecord * cap)) return rc;
    d->ev_scap = cap;
    return RD_OK;
}

// Records of disk k's bodies `want` (local indices, any order, repeats allowed).
int fetch_records(rd_disk* d, uint32_t k, std::vector<uint32_t> want, SparseView& sv)
{
    std::sort(want.begin(), want.end());
    want.erase(std::unique(want.begin(), want.end()), want.end());
    const uint32_t off = d->hd[k].off, n = (uint32_t)want.size();
    sv.idx = want;
    sv.rec.assign(n, BodyRec{});
    if (n == 0) return RD_OK;
    if (int rc = ensure_stage(d, n)) return rc;
    std::vector<uint32_t> g(n);
    for (uint32_t q = 0; q < n; ++q) g[q] = off + want[q];
    std::vector<double> buf((size_t)kBodyRecord * n);
    HIP_TRY(hipMemcpyAsync(d->ev_sidx, g.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, d->stream));
    HIP_TRY(launch_gather_bodies(d->stream, n, d->ev_sidx, body_arrays(d), d->ev_srec));
    HIP_TRY(hipMemcpyAsync(buf.data(), d->ev_srec, buf.size() * sizeof(double), hipMemcpyDeviceToHost, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    for (uint32_t q = 0; q < n; ++q) {
        const double* b = &buf[(size_t)kBodyRecord * q];
        sv.rec[q] = BodyRec{rd_var4{b[0], b[1], b[2], 0.0}, rd_var4{b[3], b[4], b[5], 0.0},
                            rd_param{b[6], b[7], b[8], b[9]}};
    }
    return RD_OK;
}

// Writes the records of sv and the bodies' current ids back to disk k, and
// refreshes the source masses the pair kernels read (a device pass).
int push_records(rd_disk* d, uint32_t k, const SparseView& sv)
{
    const uint32_t off = d->hd[k].off, n = (uint32_t)sv.idx.size();
    if (n == 0) return RD_OK;
    if (int rc = ensure_stage(d, n)) return rc;
    std::vector<uint32_t> g(n);
    std::vector<int32_t> ids(n);
    std::vector<double> buf((size_t)kBodyRecord * n);
    for (uint32_t q = 0; q < n; ++q) {
        g[q] = off + sv.idx[q];
        ids[q] = d->bmd[off + sv.idx[q]].id;
        const BodyRec& b = sv.rec[q];
        const double v[kBodyRecord] = {b.r.x, b.r.y, b.r.z, b.v.x, b.v.y, b.v.z, b.p.mass, b.p.radius, b.p.density,
                                       b.p.cd};
        std::memcpy(&buf[(size_t)kBodyRecord * q], v, sizeof(v));
    }
    HIP_TRY(hipMemcpyAsync(d->ev_sidx, g.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, d->stream));
    HIP_TRY(hipMemcpyAsync(d->ev_sids, ids.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, d->stream));
    HIP_TRY(hipMemcpyAsync(d->ev_srec, buf.data(), buf.size() * sizeof(double), hipMemcpyHostToDevice, d->stream));
    HIP_TRY(launch_scatter_bodies(d->stream, n, d->ev_sidx, d->ev_srec, d->ev_sids, body_arrays(d), d->id));
    HIP_TRY(launch_set_sources(d->stream, d->n_all, d->mass, d->id, d->disk_of, d->info, d->m_ex, d->m32));
    HIP_TRY(hipStreamSynchronize(d->stream));
    return RD_OK;
}

// pp_disk_utility::store_event_data (pp_disk.cu:196-253) without the printf.
rd_event make_event(int name, double t, double dist, uint32_t idx1, uint32_t idx2, const SparseView& sv,
                    const rd_body_md* bmd)
{
    rd_event e;
    std::memset(&e, 0, sizeof(e));
    e.event_name = name;
    e.d = dist;
    e.t = t;
    e.id0 = bmd[0].id; e.id1 = bmd[idx1].id; e.id2 = bmd[idx2].id;
    e.idx0 = 0; e.idx1 = idx1; e.idx2 = idx2;
    e.r0 = sv.at(0).r; e.v0 = sv.at(0).v;
    e.r1 = sv.at(idx1).r; e.v1 = sv.at(idx1).v;
    e.r2 = sv.at(idx2).r; e.v2 = sv.at(idx2).v;
    e.p0 = sv.at(0).p;
    if (RD_EVENT_EJECTION == name) {
        e.p1 = sv.at(idx1).p;
        e.p2 = sv.at(idx2).p;
        e.rs = e.r1;
        e.vs = e.v1;
        e.ps = e.p1;
    }
    return e;
}

double calc_radius(double m, double density)   // tools.cpp:455-460
{
    const double four_pi_over_three = 4.188790204786391;
    return std::pow(1.0 / four_pi_over_three * m / density, 1.0 / 3.0);
}

// The reference's refusal (pp_disk.cu:1333-1337), checked for every pair of
// an event batch before any of them is merged, so a refused batch changes
// nothing (the reference throws part-way through, after earlier merges).
int refuse_superpl(const std::vector<rd_event>& ev, const rd_body_md* bmd, bool skip_ejections)
{
    for (const rd_event& e : ev) {
        if (skip_ejections && RD_EVENT_EJECTION == e.event_name) continue;
        if (RD_BODY_SUPERPLANETESIMAL == bmd[e.idx2].body_type)
            return fail(RD_ENOTIMPL, "Collision between a massive body and a super-planetesimal is not yet implemented.");
    }
    return RD_OK;
}

// handle_collision_pair (pp_disk.cu:1328-1368) with tools.cpp:480-507
void merge_pair(SparseView& sv, rd_event& c)
{
    const uint32_t s = c.idx1, m = c.idx2;
    c.p1 = sv.at(s).p;
    c.p2 = sv.at(m).p;
    const double m1 = c.p1.mass, m2 = c.p2.mass, M = m1 + m2;
    c.rs.x = (m1 * c.r1.x + m2 * c.r2.x) / M;
    c.rs.y = (m1 * c.r1.y + m2 * c.r2.y) / M;
    c.rs.z = (m1 * c.r1.z + m2 * c.r2.z) / M;
    c.rs.w = 0.0;
    c.vs.x = (m1 * c.v1.x + m2 * c.v2.x) / M;
    c.vs.y = (m1 * c.v1.y + m2 * c.v2.y) / M;
    c.vs.z = (m1 * c.v1.z + m2 * c.v2.z) / M;
    c.vs.w = 0.0;
    sv.at(s).r = c.rs;
    sv.at(s).v = c.vs;
    const double R1 = c.p1.radius, R2 = c.p2.radius;
    const double volume = 4.188790204786391 * (R1 * R1 * R1 + R2 * R2 * R2);
    c.ps.mass = c.p1.mass + c.p2.mass;
    c.ps.density = c.ps.mass / volume;
    c.ps.radius = calc_radius(c.ps.mass, c.ps.density);
    c.ps.cd = c.p1.cd;
    sv.at(s).p = c.ps;
}

// pp_disk::increment_event_counter (pp_disk.cu:1431-1440); kind 0 collision,
// 1 ejection, 2 hit-centrum
void inc_counters(HostDisk& h, int kind)
{
    for (int q = 0; q < 3; ++q) {
        h.n_event[q]++;
        h.n_kind[kind][q]++;
    }
}

}  // namespace

extern "C" {

}  // extern "C"

namespace {

// Uploads ctl with F_CHECK on the disks in `want` (t and h kept).
int flag_check(rd_disk* d, const std::vector<char>& want)
{
    for (uint32_t q = 0; q < d->n_disk; ++q) {
        d->h_ctl[q].t = d->hd[q].t;
        d->h_ctl[q].flags = want[q] ? F_CHECK : 0u;
    }
    HIP_TRY(hipMemcpyAsync(d->ctl, d->h_ctl, d->n_disk * sizeof(DiskCtl), hipMemcpyHostToDevice, d->stream));
    HIP_TRY(hipMemsetAsync(d->ev_counter, 0, sizeof(uint32_t), d->stream));
    return RD_OK;
}

int read_pairs(rd_disk* d, std::vector<std::pair<uint32_t, uint32_t>>& pr)
{
    uint32_t cnt = 0;
    HIP_TRY(hipMemcpyAsync(&cnt, d->ev_counter, sizeof(uint32_t), hipMemcpyDeviceToHost, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    if (cnt > d->ev_cap)
        return fail(RD_EOVERFLOW, "event buffer overflow: more than " + std::to_string(d->ev_cap) + " events");
    pr.clear();
    if (cnt) {
        std::vector<uint32_t> buf(2 * (size_t)cnt);
        HIP_TRY(hipMemcpy(buf.data(), d->ev_buf, buf.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
        for (uint32_t q = 0; q < cnt; ++q) pr.push_back({buf[2 * q], buf[2 * q + 1]});
    }
    return RD_OK;
}

uint32_t disk_of_body(const rd_disk* d, uint32_t i)
{
    uint32_t lo = 0, hi = d->n_disk;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (d->hd[mid].off <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

// pp_disk::check_for_ejection_hit_centrum on every disk in `want`
// (pp_disk.cu:284-318, 1186-1201); events are appended to each disk's pending
// list in canonical order (body index, ejection before hit-centrum).
int detect_ejections(rd_disk* d, const std::vector<char>& want)
{
    const double ej = d->thr[RD_THRESHOLD_EJECTION_DISTANCE], hc = d->thr[RD_THRESHOLD_HIT_CENTRUM_DISTANCE];
    const double ed2 = (0.0 < ej ? ej * ej : DBL_MAX);
    const double hcd2 = hc * hc;
    if (int rc = flag_check(d, want)) return rc;
    HIP_TRY(launch_ejection(d->stream, d->n_all, d->x, d->y, d->z, d->id, d->type, d->disk_of, d->ctl, ed2, hcd2,
                            d->ev_buf, d->ev_counter, d->ev_cap));
    std::vector<std::pair<uint32_t, uint32_t>> ev;
    if (int rc = read_pairs(d, ev)) return rc;
    std::sort(ev.begin(), ev.end(), [](auto& a, auto& b) {
        return a.first != b.first ? a.first < b.first : a.second > b.second;
    });
    for (size_t q0 = 0; q0 < ev.size();) {   // one disk's events at a time
        const uint32_t k = disk_of_body(d, ev[q0].first);
        HostDisk& h = d->hd[k];
        size_t q1 = q0;
        std::vector<uint32_t> who{0};
        for (; q1 < ev.size() && ev[q1].first < h.off + h.n; ++q1) who.push_back(ev[q1].first - h.off);
        SparseView sv;
        if (int rc = fetch_records(d, k, who, sv)) return rc;
        for (size_t q = q0; q < q1; ++q) {
            const uint32_t li = ev[q].first - h.off;
            const rd_var4& r = sv.at(li).r;
            const double r2 = r.x * r.x + r.y * r.y + r.z * r.z;
            h.pending.push_back(make_event((int)ev[q].second, h.t, std::sqrt(r2), 0, li, sv, d->bmd.data() + h.off));
        }
        q0 = q1;
    }
    return RD_OK;
}

// pp_disk::check_for_collision on every disk in `want` (pp_disk.cu:321-371,
// 1203-1218); events per disk in (i, j) order.
int detect_collisions(rd_disk* d, const std::vector<char>& want)
{
    if (int rc = ensure_lists(d)) return rc;
    const double f = d->thr[RD_THRESHOLD_RADII_ENHANCE_FACTOR];
    if (int rc = flag_check(d, want)) return rc;
    HIP_TRY(launch_collision_grid(d->stream, d->n_all, d->x, d->y, d->z, d->radius, d->id, d->disk_of, d->ctl, f,
                                  d->coll_table, d->coll_head, d->coll_next, d->coll_cell, d->coll_rmax, d->ev_buf,
                                  d->ev_counter, d->ev_cap));
    std::vector<std::pair<uint32_t, uint32_t>> pr;
    if (int rc = read_pairs(d, pr)) return rc;
    std::sort(pr.begin(), pr.end());
    for (size_t q0 = 0; q0 < pr.size();) {   // one disk's pairs at a time
        const uint32_t k = disk_of_body(d, pr[q0].first);
        HostDisk& h = d->hd[k];
        size_t q1 = q0;
        std::vector<uint32_t> who{0};
        for (; q1 < pr.size() && pr[q1].first < h.off + h.n; ++q1) {
            who.push_back(pr[q1].first - h.off);
            who.push_back(pr[q1].second - h.off);
        }
        SparseView sv;
        if (int rc = fetch_records(d, k, who, sv)) return rc;
        for (size_t q = q0; q < q1; ++q) {
            uint32_t s = pr[q].first - h.off, m = pr[q].second - h.off;
            const rd_var4 &rs = sv.at(s).r, &rm = sv.at(m).r;
            const double dx = rm.x - rs.x, dy = rm.y - rs.y, dz = rm.z - rs.z;
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (sv.at(m).p.mass > sv.at(s).p.mass) std::swap(s, m);
            h.pending.push_back(make_event(RD_EVENT_COLLISION, h.t, std::sqrt(r2), s, m, sv, d->bmd.data() + h.off));
        }
        q0 = q1;
    }
    return RD_OK;
}

}  // namespace

extern "C" {

int rd_check_ejection_hit_centrum(rd_disk* d, uint32_t k, uint32_t* n)
{
    DeviceScope dev_scope(d ? d->device : -1);
    if (!d || k >= d->n_disk) return fail(RD_EINVAL, "bad argument");
    std::vector<char> want(d->n_disk, 0);
    want[k] = 1;
    if (int rc = detect_ejections(d, want)) return rc;
    if (n) *n = (uint32_t)d->hd[k].pending.size();
    return RD_OK;
}

int rd_check_collision(rd_disk* d, uint32_t k, uint32_t* n)
{
    DeviceScope dev_scope(d ? d->device : -1);
    if (!d || k >= d->n_disk) return fail(RD_EINVAL, "bad argument");
    std::vector<char> want(d->n_disk, 0);
    want[k] = 1;
    if (int rc = detect_collisions(d, want)) return rc;
    if (n) *n = (uint32_t)d->hd[k].pending.size();
    return RD_OK;
}

int rd_get_events(rd_disk* d, uint32_t k, rd_event* out, uint32_t cap, uint32_t* n)
{
    DeviceScope dev_scope(d ? d->device : -1);
    if (!d || k >= d->n_disk) return fail(RD_EINVAL, "bad argument");
    const auto& p = d->hd[k].pending;
    if (out) std::memcpy(out, p.data(), std::min<size_t>(cap, p.size()) * sizeof(rd_event));
    if (n) *n = (uint32_t)p.size();
    return RD_OK;
}

int rd_event_log(rd_disk* d, uint32_t k, rd_event* out, uint32_t cap, uint32_t* n)
{
    DeviceScope dev_scope(d ? d->device : -1);
    if (!d || k >= d->n_disk) return fail(RD_EINVAL, "bad argument");
    const auto& p = d->hd[k].log;
    if (out) std::memcpy(out, p.data(), std::min<size_t>(cap, p.size()) * sizeof(rd_event));
    if (n) *n = (uint32_t)p.size();
    return RD_OK;
}

int rd_clear_event_counter(rd_disk* d, uint32_t k)
{
    DeviceScope dev_scope(d ? d->device : -1);
    if (!d || k >= d->n_disk) return fail(RD_EINVAL, "bad argument");
    d->hd[k].pending.clear();
    return RD_OK;
}

int rd_handle_collision(rd_disk* d, uint32_t k)
{
    DeviceScope dev_scope(d ? d->device : -1);
    if (!d || k >= d->n_disk) return fail(RD_EINVAL, "bad argument");
    HostDisk& h = d->hd[k];
    if (h.pending.empty()) return RD_OK;
    rd_body_md* bmd = d->bmd.data() + h.off;
    // populate_sp_events (pp_disk.cu:1250-1288)
    std::vector<rd_event> sp;
    std::vector<char> done(h.pending.size(), 0);
    for (size_t q = 0; q < h.pending.size(); ++q) {
        if (done[q]) continue;
        done[q] = 1;
        rd_event e = h.pending[q];
        for (size_t u = q + 1; u < h.pending.size(); ++u)
            if (e.id1 == h.pending[u].id1 && e.id2 == h.pending[u].id2) {
                done[u] = 1;
                if (e.t > h.pending[u].t) e = h.pending[u];
            }
        sp.push_back(e);
    }
    if (int rc = refuse_superpl(sp, bmd, false)) return rc;
    std::vector<uint32_t> who;
    for (auto& e : sp) {
        who.push_back(e.idx1);
        who.push_back(e.idx2);
    }
    SparseView sv;
    if (int rc = fetch_records(d, k, who, sv)) return rc;
    for (auto& e : sp) {
        const uint32_t ictv = e.idx2;
        merge_pair(sv, e);
        inc_counters(h, 0);
        bmd[ictv].id *= -1;
        h.inactive[bmd[ictv].body_type]++;
        h.log.push_back(e);
    }
    h.n_kind[0][2] = 0;   // pp_disk.cu:1247
    return push_records(d, k, sv);
}

int rd_handle_ejection_hit_centrum(rd_disk* d, uint32_t k)
{
    DeviceScope dev_scope(d ? d->device : -1);
    if (!d || k >= d->n_disk) return fail(RD_EINVAL, "bad argument");
    HostDisk& h = d->hd[k];
    if (h.pending.empty()) return RD_OK;
    rd_body_md* bmd = d->bmd.data() + h.off;
    if (int rc = refuse_superpl(h.pending, bmd, true)) return rc;
    std::vector<uint32_t> who;
    for (auto& e : h.pending) {
        who.push_back(e.idx1);
        who.push_back(e.idx2);
    }
    SparseView sv;
    if (int rc = fetch_records(d, k, who, sv)) return rc;
    for (auto e : h.pending) {
        const uint32_t ictv = e.idx2;
        if (RD_EVENT_EJECTION != e.event_name) merge_pair(sv, e);
        inc_counters(h, RD_EVENT_EJECTION == e.event_name ? 1 : 2);
        bmd[ictv].id *= -1;
        h.inactive[bmd[ictv].body_type]++;
        h.log.push_back(e);
    }
    h.n_kind[1][2] = h.n_kind[2][2] = 0;   // pp_disk.cu:1324-1325
    return push_records(d, k, sv);
}

int rd_event_counters(rd_disk* d, uint32_t k, uint32_t n_event[3], uint32_t n_collision[3], uint32_t n_ejection[3],
                      uint32_t n_hit_centrum[3])
{
    if (!d || k >= d->n_disk) return fail(RD_EINVAL, "bad argument");
    const HostDisk& h = d->hd[k];
    for (int q = 0; q < 3; ++q) {
        if (n_event) n_event[q] = h.n_event[q];
        if (n_collision) n_collision[q] = h.n_kind[0][q];
        if (n_ejection) n_ejection[q] = h.n_kind[1][q];
        if (n_hit_centrum) n_hit_centrum[q] = h.n_kind[2][q];
    }
    return RD_OK;
}

int rd_set_event_counter(rd_disk* d, uint32_t k, int field, uint32_t value)
{
    if (!d || k >= d->n_disk || field < 0 || field > 2) return fail(RD_EINVAL, "bad argument");
    HostDisk& h = d->hd[k];
    h.n_event[field] = value;
    for (int q = 0; q < 3; ++q) h.n_kind[q][field] = value;
    return RD_OK;
}

int rd_rebuild_vectors(rd_disk* d, uint32_t k)
{
    DeviceScope dev_scope(d ? d->device : -1);
    if (!d || k >= d->n_disk) return fail(RD_EINVAL, "bad argument");
    // download everything, compact disk k, re-lay all disks
    std::vector<HostView> all(d->n_disk);
    for (uint32_t q = 0; q < d->n_disk; ++q)
        if (int rc = fetch(d, q, all[q])) return rc;
    std::vector<rd_body_md> nb;
    HostDisk& hk = d->hd[k];
    HostView cv;
    for (uint32_t i = 0; i < hk.n; ++i) {
        if (0 < d->bmd[hk.off + i].id) {
            cv.r.push_back(all[k].r[i]);
            cv.v.push_back(all[k].v[i]);
            cv.p.push_back(all[k].p[i]);
        }
    }
    for (int t = 0; t < kNType; ++t) {   // n_objects_t::update (red_type.h:545-556)
        hk.playing[t] -= hk.inactive[t];
        hk.removed[t] += hk.inactive[t];
        hk.inactive[t] = 0;
    }
    uint32_t off = 0;
    std::vector<rd_body_md> bm;
    for (uint32_t q = 0; q < d->n_disk; ++q) {
        HostDisk& h = d->hd[q];
        for (uint32_t i = 0; i < h.n; ++i)
            if (q != k || 0 < d->bmd[h.off + i].id) bm.push_back(d->bmd[h.off + i]);
        if (q == k) {
            all[q] = cv;
            h.n = (uint32_t)cv.r.size();
        }
        h.off = off;
        off += h.n;
    }
    uint32_t chk = 0;
    for (int t = 0; t < kNType; ++t) chk += hk.playing[t];
    if (chk != hk.n) return fail(RD_EINVAL, "Number of copied bodies does not equal to the number of active bodies.");
    d->bmd = bm;
    d->n_all = off;
    free_buffers(d);
    if (int rc = alloc_buffers(d)) return rc;
    if (int rc = upload_disk_of(d)) return rc;
    for (uint32_t q = 0; q < d->n_disk; ++q)
        if (int rc = upload_bodies(d, d->hd[q].off, d->hd[q].n, all[q].r.data(), all[q].v.data(), all[q].p.data(),
                                   d->bmd.data() + d->hd[q].off))
            return rc;
    d->lists_valid = false;
    return build_lists(d);
}

}  // extern "C"

namespace {

// The integrals of disk k from the kernels of launch_integral on `scr` (len
// doubles of partial sums) with the 12 reduced values in `red`.
int integral_with(rd_disk* d, uint32_t k, double* scr, size_t len, double* red, rd_integral* out)
{
    const HostDisk& h = d->hd[k];
    HIP_TRY(launch_integral(d->stream, h.off, h.off + h.n, d->x, d->y, d->z, d->vx, d->vy, d->vz, d->mass, d->id,
                            scr, len, red, red + 11));
    double o[12];
    HIP_TRY(hipMemcpyAsync(o, red, sizeof(o), hipMemcpyDeviceToHost, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    const double M0 = o[0];
    out->R = rd_var4{o[1] / M0, o[2] / M0, o[3] / M0, 0.0};
    out->V = rd_var4{o[4] / M0, o[5] / M0, o[6] / M0, 0.0};
    out->C = rd_var4{o[7], o[8], o[9], 0.0};
    out->E = o[10] / 2.0 - o[11] / 2.0;
    return RD_OK;
}

// doubles of partial sums launch_integral needs for n bodies
size_t integral_scratch(uint32_t n)
{
    const size_t nch = (n + 2047) / 2048;
    return std::max<size_t>(11 * (size_t)n, nch * n);
}

// Scratch of rd_calc_integral_noalloc: a file-scope device array, so the call
// allocates nothing.  Every GPU the library runs on holds its own copy (16 MiB
// of partial sums plus 16 doubles of results) until the process ends; calls on
// one GPU share it, so calls that overlap on one GPU race.  Every value a call
// reads was written by the same call: nothing needs clearing.
constexpr size_t kNoAllocScratch = (size_t)1 << 21;
static __device__ __align__(256) double scratch[kNoAllocScratch + 16];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per device
// (host code must never use the symbol's name as an address).
int scratch_on(int device, double** out)
{
    if (device < 0 || device >= kMaxDevices) return fail(RD_EINVAL, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(RD_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[device];
    return RD_OK;
}

}  // namespace

extern "C" {

int rd_calc_integral(rd_disk* d, uint32_t k, rd_integral* out)
{
    DeviceScope dev_scope(d ? d->device : -1);
    if (!d || !out || k >= d->n_disk) return fail(RD_EINVAL, "bad argument");
    const size_t need = integral_scratch(d->hd[k].n);
    if (need > d->scratch_len) {
        dev_free(d->scratch);
        d->scratch_len = 0;
        if (int rc = dev_alloc(&d->scratch, need)) return rc;
        d->scratch_len = need;
    }
    return integral_with(d, k, d->scratch, d->scratch_len, d->red_out, out);
}

int rd_calc_integral_noalloc(rd_disk* d, uint32_t k, rd_integral* out)
{
    DeviceScope dev_scope(d ? d->device : -1);
    if (!d || !out || k >= d->n_disk) return fail(RD_EINVAL, "bad argument");
    const size_t need = integral_scratch(d->hd[k].n);
    if (need > kNoAllocScratch)
        return fail(RD_EINVAL, "rd_calc_integral_noalloc: disk of " + std::to_string(d->hd[k].n) +
                                   " bodies needs " + std::to_string(need) + " doubles of scratch, more than the " +
                                   std::to_string(kNoAllocScratch) + " it holds (at most 65536 bodies)");
    double* scr = nullptr;
    if (int rc = scratch_on(d->device, &scr)) return rc;
    return integral_with(d, k, scr, kNoAllocScratch, scr + kNoAllocScratch, out);
}

// run_simulation's loop (src/red.cuda/main.cu:470-625) per disk, all disks in
// lock step: each iteration every running disk takes one accepted step.
int rd_run(rd_integrator* it, double length, double output_interval, uint64_t max_steps)
{
    DeviceScope dev_scope(it ? it->device : -1);
    if (int rc = check_live(it)) return rc;
    rd_disk* d = it->d;
    const uint32_t nd = d->n_disk;
    std::vector<double> elapsed(nd, 0.0), ps(nd, 0.0);
    std::vector<uint64_t> steps(nd, 0);
    for (;;) {
        std::vector<char> want(nd, 0);
        bool any = false;
        for (uint32_t k = 0; k < nd; ++k) {
            if (std::fabs(elapsed[k]) <= std::fabs(length) && 1 < d->hd[k].n_active() && steps[k] < max_steps) {
                want[k] = 1;
                any = true;
            }
        }
        if (!any) break;
        if (0.0 < d->thr[RD_THRESHOLD_EJECTION_DISTANCE] || 0.0 < d->thr[RD_THRESHOLD_HIT_CENTRUM_DISTANCE]) {
            if (int rc = detect_ejections(d, want)) return rc;
            for (uint32_t k = 0; k < nd; ++k) {
                if (!want[k] || d->hd[k].pending.empty()) continue;
                if (int rc = rd_handle_ejection_hit_centrum(d, k)) return rc;
                rd_clear_event_counter(d, k);
            }
        }
        if (int rc = step_disks(it, std::vector<uint32_t>(want.begin(), want.end()))) return rc;
        for (uint32_t k = 0; k < nd; ++k) {
            if (!want[k]) continue;
            const double dt = it->dt_did[k];
            steps[k]++;
            elapsed[k] += dt;
            ps[k] += dt;
            if (std::fabs(output_interval) < std::fabs(it->dt_next[k]))
                it->dt_next[k] = 0.0 < dt ? output_interval : -output_interval;
        }
        if (0.0 < d->thr[RD_THRESHOLD_RADII_ENHANCE_FACTOR]) {
            if (int rc = detect_collisions(d, want)) return rc;
            for (uint32_t k = 0; k < nd; ++k) {
                if (!want[k] || d->hd[k].pending.empty()) continue;
                if (int rc = rd_handle_collision(d, k)) return rc;
                rd_clear_event_counter(d, k);
            }
        }
        for (uint32_t k = 0; k < nd; ++k) {
            if (!want[k]) continue;
            if (output_interval <= std::fabs(ps[k])) ps[k] = 0.0;
            if (16 <= d->hd[k].n_event[1]) {
                rd_set_event_counter(d, k, RD_EVENT_COUNTER_LAST_CLEAR, 0);
                if (int rc = rd_rebuild_vectors(d, k)) return rc;
            }
        }
    }
    return RD_OK;
}

}  // extern "C"
