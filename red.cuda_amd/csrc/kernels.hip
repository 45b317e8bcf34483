// kernels.hip -- HIP kernels of the direct-sum engine for gfx950 (CDNA4).
//
// Compiled with -ffp-contract=off: every FP64 expression that restates a
// reference formula keeps the reference's rounding (the reference builds with
// -fmad=false, Makefile:17).  Fused multiply-adds appear only where written
// explicitly (fmaf in the mixed-precision pair loop).
#include "engine.hpp"

#include <float.h>

#include <algorithm>

namespace rdamd {

#define SQR(x) ((x) * (x))

// 2 pi rounded to double, the value of the reference's TWOPI (red_macro.h:17)
constexpr double kTwoPi = 6.2831853071795864769253;

// ---------------------------------------------------------------------------
// Runge-Kutta tableaux
// ---------------------------------------------------------------------------
struct RowDef {
    int n;
    int col[9];
    double a[9];
};

// Fehlberg 7(8), int_rungekutta8.cu:30-48; nonzero columns as used by
// calc_ytemp_for_f1..f12 (int_rungekutta8.cu:82-246, 414-585).
__host__ __device__ constexpr RowDef rk8_row(int r)
{
    switch (r) {
    case 1: return {1, {0}, {2.0 / 27.0}};
    case 2: return {2, {0, 1}, {1.0 / 36.0, 1.0 / 12.0}};
    case 3: return {2, {0, 2}, {1.0 / 24.0, 1.0 / 8.0}};
    case 4: return {3, {0, 2, 3}, {5.0 / 12.0, -25.0 / 16.0, 25.0 / 16.0}};
    case 5: return {3, {0, 3, 4}, {1.0 / 20.0, 1.0 / 4.0, 1.0 / 5.0}};
    case 6: return {4, {0, 3, 4, 5}, {-25.0 / 108.0, 125.0 / 108.0, -65.0 / 27.0, 125.0 / 54.0}};
    case 7: return {4, {0, 4, 5, 6}, {31.0 / 300.0, 61.0 / 225.0, -2.0 / 9.0, 13.0 / 900.0}};
    case 8: return {6, {0, 3, 4, 5, 6, 7}, {2.0, -53.0 / 6.0, 704.0 / 45.0, -107.0 / 9.0, 67.0 / 90.0, 3.0}};
    case 9:
        return {7, {0, 3, 4, 5, 6, 7, 8},
                {-91.0 / 108.0, 23.0 / 108.0, -976.0 / 135.0, 311.0 / 54.0, -19.0 / 60.0, 17.0 / 6.0, -1.0 / 12.0}};
    case 10:
        return {8, {0, 3, 4, 5, 6, 7, 8, 9},
                {2383.0 / 4100.0, -341.0 / 164.0, 4496.0 / 1025.0, -301.0 / 82.0, 2133.0 / 4100.0, 45.0 / 82.0,
                 45.0 / 164.0, 18.0 / 41.0}};
    case 11:
        return {6, {0, 5, 6, 7, 8, 9},
                {3.0 / 205.0, -6.0 / 41.0, -3.0 / 205.0, -3.0 / 41.0, 3.0 / 41.0, 6.0 / 41.0}};
    case 12:
        return {9, {0, 3, 4, 5, 6, 7, 8, 9, 11},
                {-1777.0 / 4100.0, -341.0 / 164.0, 4496.0 / 1025.0, -289.0 / 82.0, 2193.0 / 4100.0, 51.0 / 82.0,
                 33.0 / 164.0, 12.0 / 41.0, 1.0}};
    default: return {0, {0}, {0.0}};
    }
}
__host__ __device__ constexpr double rk8_c(int r)
{
    switch (r) {
    case 1: return 2.0 / 27.0;  case 2: return 1.0 / 9.0;  case 3: return 1.0 / 6.0;
    case 4: return 5.0 / 12.0;  case 5: return 1.0 / 2.0;  case 6: return 5.0 / 6.0;
    case 7: return 1.0 / 6.0;   case 8: return 2.0 / 3.0;  case 9: return 1.0 / 3.0;
    case 10: return 1.0;        case 11: return 0.0;       case 12: return 1.0;
    default: return 0.0;
    }
}
// Classical RK4, int_rungekutta4.cu:26-31
__host__ __device__ constexpr double rk4_a(int r) { return r == 3 ? 1.0 : (r == 0 ? 0.0 : 1.0 / 2.0); }
__host__ __device__ constexpr double rk4_c(int r) { return (r == 1 || r == 2) ? 1.0 / 2.0 : (r == 0 ? 0.0 : 1.0); }

// ---------------------------------------------------------------------------
// Pair kernels
// ---------------------------------------------------------------------------
// Reference formula (pp_disk.cu:374-416): sources j ascending, skip i == j
// and inactive j (m_ex < 0), a += (m_j / (d * d^2)) * d_vec, every operation
// rounded separately (the reference builds -fmad=false).
//
// The sum per (sink, component) is a chain of FP64 adds in j order, so it
// cannot be split; the pair terms feeding it can.  A block takes kS sinks
// (16, or 4 for disks whose 16-sink items would not reach every CU) and has
// five waves.  Waves 0-3 compute, for each LDS tile of kTile sources, the
// kS x kTile terms (m_j / (d * d^2)) * dx (sqrt and divide are the costly
// part) into LDS; 3 kS lanes of wave 4 -- one per (sink, component) -- add a
// tile's terms in j order to their running sums.  Terms and sources are
// double-buffered: wave 4 adds tile t - 1 while waves 0-3 compute tile t, one
// barrier per tile, so a tile costs the longer of the two, not their sum.
// Same terms, same order: the result is bit-identical to one lane summing
// alone, for either kS, while the term work spreads over N / kS blocks instead
// of N lanes.  Threads of sinks past the item's end (a disk's last item)
// compute nothing.
constexpr int kExactBlock = kBlock + 64;   // 4 term waves + 1 summing wave

// m / (sqrt(r2) * r2) for r2 and m in the range exact_scale_in_range accepts,
// bit for bit what the IEEE square root and division give: the instruction
// sequences the compiler emits for them (v_rsq_f64 and two Goldschmidt steps
// for the root; v_rcp_f64, two Newton steps and a residual step for the
// quotient) without their range handling, which does nothing in that range (no
// operand is scaled by v_ldexp or v_div_scale, and v_div_fixup and the
// zero/infinity test of the root return the computed value).  About a fifth of
// a pair term's instructions.
__device__ __forceinline__ bool exact_scale_in_range(double r2, double m)
{
    return r2 >= 0x1p-200 && r2 <= 0x1p200 && m >= 0x1p-300 && m <= 0x1p300;
}
__device__ __forceinline__ double exact_scale_fast(double r2, double m)
{
    const double y = __builtin_amdgcn_rsq(r2);
    double g = r2 * y, h = y * 0.5;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    double e = __builtin_fma(-g, g, r2);
    h = __builtin_fma(h, r, h);
    g = __builtin_fma(e, h, g);
    e = __builtin_fma(-g, g, r2);
    g = __builtin_fma(e, h, g);   // sqrt(r2)
    const double den = g * r2;
    double q = __builtin_amdgcn_rcp(den);
    double t = __builtin_fma(-den, q, 1.0);
    q = __builtin_fma(q, t, q);
    t = __builtin_fma(-den, q, 1.0);
    q = __builtin_fma(q, t, q);   // 1 / den
    const double a = m * q;
    const double rem = __builtin_fma(-den, a, m);
    return __builtin_fma(rem, q, a);
}

// Single-disk classical RK4 stage R for body i, component q (defined with the
// stage kernels below).
template <int R, bool kFused>
__device__ void stage_rk4_body(const StageArgs& a, uint32_t i, int q, bool in, double ps_fused);
struct NoStage {};

// FR < 0: each summing lane stores its sum in the partial-sum slot.  FR >= 0
// (a single gas-free disk, fixed-step RK4, FP64): the lane holds body si's complete
// component-sc_c sum -- the value the stage kernel would read back from that
// slot -- and runs stage FR for the body itself: it reads only that body's
// state and k rows and writes the other position buffer, which no pair term
// of this launch reads.
template <uint32_t kS, int FR, typename SA>
__device__ __forceinline__ void force_exact_body(const WorkItem* __restrict__ items, uint32_t n_items,
                                                 const double* __restrict__ x, const double* __restrict__ y,
                                                 const double* __restrict__ z, const double* __restrict__ m_ex,
                                                 const DiskCtl* __restrict__ ctl, uint32_t need,
                                                 double* __restrict__ part, uint32_t n_all, const SA& sa)
{
    constexpr uint32_t kLanesPerSink = kBlock / kS;              // term threads sharing one sink
    // sources per tile: 32 for 16-sink items (two term buffers in 26 KB of
    // LDS), one per term thread of a sink otherwise
    constexpr uint32_t kTile = kS >= 16 ? 32u : kLanesPerSink;
    constexpr uint32_t kT = kTile / kLanesPerSink;               // terms per thread per tile
    static_assert(kS * kTile == kT * kBlock && kT >= 1 && 3 * kS <= 64 && kTile <= kBlock, "tile shape");
    __shared__ double4 src[2][kTile];
    __shared__ double term[2][3][kS][kTile + 1];   // +1: summing lanes read with an odd stride
    const uint32_t tid = threadIdx.x;
    const bool termer = tid < (uint32_t)kBlock;   // waves 0-3
    const uint32_t g = tid / kLanesPerSink;   // term threads: sink of this thread's terms
    const uint32_t k_lo = tid % kLanesPerSink;    //            its sources: k_lo + kLanesPerSink q, q < kT
    const uint32_t sl = tid - (uint32_t)kBlock;   // wave 4: summing lane (sl < 3 kS)
    const uint32_t sc_c = sl / kS % 3u;      //         component
    const uint32_t sc_g = sl % kS;           //         sink
    for (uint32_t it = blockIdx.x; it < n_items; it += gridDim.x) {
        const WorkItem w = items[it];
        if (!(ctl[w.disk].flags & need)) continue;
        const uint32_t i = w.sink_lo + g;
        const bool valid = termer && i < w.sink_hi;
        const double xi = valid ? x[i] : 0.0, yi = valid ? y[i] : 0.0, zi = valid ? z[i] : 0.0;
        const uint32_t si = w.sink_lo + sc_g;  // sink of a summing lane
        const bool summer = !termer && sl < 3 * kS && si < w.sink_hi;
        const uint32_t n_tiles = (w.src_hi - w.src_lo + kTile - 1) / kTile;
        auto load_src = [&](uint32_t j) {
            return j < w.src_hi ? make_double4(x[j], y[j], z[j], m_ex[j]) : make_double4(0.0, 0.0, 0.0, -1.0);
        };
        double acc = 0.0;
        // tile 0 into LDS, tile 1 into registers; the loop stages tile t + 1
        // and loads tile t + 2 while tile t's terms are computed
        double4 nxt = make_double4(0.0, 0.0, 0.0, -1.0);
        if (tid < kTile) {
            src[0][tid] = load_src(w.src_lo + tid);
            nxt = load_src(w.src_lo + kTile + tid);
        }
        __syncthreads();
        for (uint32_t t = 0; t <= n_tiles; ++t) {
            const uint32_t b = t & 1u;
            if (termer) {
                if (t < n_tiles) {
                    const uint32_t t0 = w.src_lo + t * kTile;
                    if (valid) {
#pragma unroll
                        for (uint32_t q = 0; q < kT; ++q) {
                            const uint32_t k = k_lo + kLanesPerSink * q;
                            const double4 sj = src[b][k];
                            const double dx = sj.x - xi;
                            const double dy = sj.y - yi;
                            const double dz = sj.z - zi;
                            const double r2 = SQR(dx) + SQR(dy) + SQR(dz);
                            const bool skip = sj.w < 0.0 || t0 + k == i;
                            double sc;
                            if constexpr (kS >= 16) {
                                // throughput-bound (large disks): the short sequence, with
                                // the library's for the rare pair outside its range
                                sc = exact_scale_fast(r2, sj.w);
                                if (!skip && !exact_scale_in_range(r2, sj.w)) sc = sj.w / (sqrt(r2) * r2);
                            } else {
                                // latency-bound (small disks): the branch would cost more
                                // than the shorter sequence saves
                                sc = sj.w / (sqrt(r2) * r2);
                            }
                            // a skipped pair (self, inactive or padding source) adds 0 * d,
                            // a zero of either sign: the running sum starts at +0 and
                            // round-to-nearest never makes it -0, so the sum is unchanged
                            sc = skip ? 0.0 : sc;
                            term[b][0][g][k] = sc * dx;
                            term[b][1][g][k] = sc * dy;
                            term[b][2][g][k] = sc * dz;
                        }
                    }
                    if (tid < kTile) {
                        src[b ^ 1u][tid] = nxt;
                        nxt = load_src(w.src_lo + (t + 2) * kTile + tid);
                    }
                }
            } else if (t > 0 && summer) {
                const double* tr = &term[b ^ 1u][sc_c][sc_g][0];
#pragma unroll
                for (uint32_t k = 0; k < kTile; ++k) acc += tr[k];
            }
            __syncthreads();
        }
        if constexpr (FR >= 0) {
            if (summer) stage_rk4_body<FR, true>(sa, si, (int)sc_c, true, acc);
        } else {
            if (summer) part[(size_t)(w.slot * 3 + sc_c) * n_all + si] = acc;
        }
    }
}

template <uint32_t kS>
__global__ __launch_bounds__(kExactBlock) void k_force_exact(const WorkItem* __restrict__ items, uint32_t n_items,
                                                             const double* __restrict__ x,
                                                             const double* __restrict__ y,
                                                             const double* __restrict__ z,
                                                             const double* __restrict__ m_ex,
                                                             const DiskCtl* __restrict__ ctl, uint32_t need,
                                                             double* __restrict__ part, uint32_t n_all)
{
    force_exact_body<kS, -1>(items, n_items, x, y, z, m_ex, ctl, need, part, n_all, NoStage{});
}

// The FP64 pair pass of RK4 stage R with the stage itself done by the summing lanes.
template <uint32_t kS, int R>
__global__ __launch_bounds__(kExactBlock) void k_force_exact_rk4(const WorkItem* __restrict__ items,
                                                                 uint32_t n_items, const double* __restrict__ x,
                                                                 const double* __restrict__ y,
                                                                 const double* __restrict__ z,
                                                                 const double* __restrict__ m_ex,
                                                                 const DiskCtl* __restrict__ ctl, uint32_t need,
                                                                 StageArgs sa)
{
    force_exact_body<kS, R>(items, n_items, x, y, z, m_ex, ctl, need, nullptr, sa.n_all, sa);
}

// The same sum with one sink per lane, each lane adding its terms in j order
// as it computes them (the reference kernel's structure): cheaper per pair,
// but only one lane per sink, so it is used when 256-sink work items alone
// fill the device (large batches); bit-identical to k_force_exact.
__global__ __launch_bounds__(kBlock) void k_force_exact_lane(const WorkItem* __restrict__ items, uint32_t n_items,
                                                             const double* __restrict__ x,
                                                             const double* __restrict__ y,
                                                             const double* __restrict__ z,
                                                             const double* __restrict__ m_ex,
                                                             const DiskCtl* __restrict__ ctl, uint32_t need,
                                                             double* __restrict__ part, uint32_t n_all)
{
    __shared__ double4 tile[kTile];
    const uint32_t tid = threadIdx.x;
    for (uint32_t it = blockIdx.x; it < n_items; it += gridDim.x) {
        const WorkItem w = items[it];
        if (!(ctl[w.disk].flags & need)) continue;
        const uint32_t i = w.sink_lo + tid;
        const bool valid = i < w.sink_hi;
        const double xi = valid ? x[i] : 0.0, yi = valid ? y[i] : 0.0, zi = valid ? z[i] : 0.0;
        double ax = 0.0, ay = 0.0, az = 0.0;
        for (uint32_t t0 = w.src_lo; t0 < w.src_hi; t0 += kTile) {
            __syncthreads();
            const uint32_t j = t0 + tid;
            if (j < w.src_hi) tile[tid] = make_double4(x[j], y[j], z[j], m_ex[j]);
            __syncthreads();
            const uint32_t nt = min((uint32_t)kTile, w.src_hi - t0);
            for (uint32_t k = 0; k < nt; ++k) {
                const double4 sj = tile[k];
                if (sj.w < 0.0 || t0 + k == i) continue;
                const double dx = sj.x - xi;
                const double dy = sj.y - yi;
                const double dz = sj.z - zi;
                const double r2 = SQR(dx) + SQR(dy) + SQR(dz);
                double sc = exact_scale_fast(r2, sj.w);
                if (!exact_scale_in_range(r2, sj.w)) sc = sj.w / (sqrt(r2) * r2);
                ax += sc * dx;
                ay += sc * dy;
                az += sc * dz;
            }
        }
        if (valid) {
            part[(size_t)(w.slot * 3 + 0) * n_all + i] = ax;
            part[(size_t)(w.slot * 3 + 1) * n_all + i] = ay;
            part[(size_t)(w.slot * 3 + 2) * n_all + i] = az;
        }
        __syncthreads();
    }
}

struct alignas(16) SrcF {
    double x, y, z;
    float m, pad;
};

// Softening floor that keeps the self pair (d = 0) and zero-mass padding
// finite: r^-3 at r^2 = 1e-24 AU^2 is 1e36, inside FP32 range, and the term is
// multiplied by d_vec = 0.  Real separations are > 1e-8 AU (two 1 km bodies
// touching), so the floor changes r^2 by < 1e-8 ulp.
constexpr float kEps2 = 1.0e-24f;

// Mixed precision pair kernel.  A work item is 64*S sinks (S per lane) against
// a chunk of sources; the four waves of the block split each 256-source LDS
// tile in quarters.  Per pair: FP64 offset rounded to FP32, r^2 by FMA, one
// v_rsq_f32, FP32 partial sums per 64-source quarter tile folded into FP64.
// The quarter sums of the four waves are combined in fixed order, so the
// result is deterministic.
template <int S>
__global__ __launch_bounds__(kBlock) void k_force_mixed(const WorkItem* __restrict__ items, uint32_t n_items,
                                                        const double* __restrict__ x, const double* __restrict__ y,
                                                        const double* __restrict__ z,
                                                        const float* __restrict__ m32,
                                                        const DiskCtl* __restrict__ ctl, uint32_t need,
                                                        float* __restrict__ part, uint32_t n_all)
{
    __shared__ SrcF tile[kTile];
    __shared__ double red[3][S][3][64];
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    for (uint32_t it = blockIdx.x; it < n_items; it += gridDim.x) {
        const WorkItem w = items[it];
        if (!(ctl[w.disk].flags & need)) continue;
        double xi[S], yi[S], zi[S], ax[S], ay[S], az[S];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const uint32_t i = w.sink_lo + lane + 64u * s;
            const bool valid = i < w.sink_hi;
            xi[s] = valid ? x[i] : 0.0;
            yi[s] = valid ? y[i] : 0.0;
            zi[s] = valid ? z[i] : 0.0;
            ax[s] = ay[s] = az[s] = 0.0;
        }
        for (uint32_t t0 = w.src_lo; t0 < w.src_hi; t0 += kTile) {
            __syncthreads();
            {
                const uint32_t j = t0 + tid;
                SrcF r;
                if (j < w.src_hi) {
                    r.x = x[j]; r.y = y[j]; r.z = z[j]; r.m = m32[j];
                } else {
                    r.x = 0.0; r.y = 0.0; r.z = 0.0; r.m = 0.0f;
                }
                r.pad = 0.0f;
                tile[tid] = r;
            }
            __syncthreads();
            const uint32_t nt = min((uint32_t)kTile, w.src_hi - t0);
            const uint32_t k0 = wave * 64u;
            const uint32_t k1 = min(k0 + 64u, (nt + 3u) & ~3u);
            float fx[S], fy[S], fz[S];
#pragma unroll
            for (int s = 0; s < S; ++s) fx[s] = fy[s] = fz[s] = 0.0f;
            for (uint32_t k = k0; k < k1; k += 4) {
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const SrcF r = tile[k + u];
#pragma unroll
                    for (int s = 0; s < S; ++s) {
                        const float dx = (float)(r.x - xi[s]);
                        const float dy = (float)(r.y - yi[s]);
                        const float dz = (float)(r.z - zi[s]);
                        float r2 = __builtin_fmaf(dz, dz, kEps2);
                        r2 = __builtin_fmaf(dy, dy, r2);
                        r2 = __builtin_fmaf(dx, dx, r2);
                        const float ri = __builtin_amdgcn_rsqf(r2);
                        const float ri2 = ri * ri;
                        const float sc = (r.m * ri) * ri2;
                        fx[s] = __builtin_fmaf(sc, dx, fx[s]);
                        fy[s] = __builtin_fmaf(sc, dy, fy[s]);
                        fz[s] = __builtin_fmaf(sc, dz, fz[s]);
                    }
                }
            }
#pragma unroll
            for (int s = 0; s < S; ++s) {
                ax[s] += (double)fx[s];
                ay[s] += (double)fy[s];
                az[s] += (double)fz[s];
            }
        }
        if (wave > 0) {
#pragma unroll
            for (int s = 0; s < S; ++s) {
                red[wave - 1][s][0][lane] = ax[s];
                red[wave - 1][s][1][lane] = ay[s];
                red[wave - 1][s][2][lane] = az[s];
            }
        }
        __syncthreads();
        if (wave == 0) {
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const uint32_t i = w.sink_lo + lane + 64u * s;
                if (i < w.sink_hi) {
                    const double sx = ((ax[s] + red[0][s][0][lane]) + red[1][s][0][lane]) + red[2][s][0][lane];
                    const double sy = ((ay[s] + red[0][s][1][lane]) + red[1][s][1][lane]) + red[2][s][1][lane];
                    const double sz = ((az[s] + red[0][s][2][lane]) + red[1][s][2][lane]) + red[2][s][2][lane];
                    part[(size_t)(w.slot * 3 + 0) * n_all + i] = (float)sx;
                    part[(size_t)(w.slot * 3 + 1) * n_all + i] = (float)sy;
                    part[(size_t)(w.slot * 3 + 2) * n_all + i] = (float)sz;
                }
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Symmetric mixed-precision pair kernel: every unordered pair of the SI class
// is evaluated once and applied to both bodies (the CPU path's i < j scheme,
// pp_disk.cu:822-865), halving the pair work of the one-sided kernel.
//
// Bodies are cut in blocks of kSymBlock = 256 (64 lanes x 4 sinks).  A work
// item is a pair of block chunks (a, c), a <= c, R blocks each; it runs the
// tiles (I, J) of chunk a x chunk c with J >= I.  In a tile, each of the four
// waves holds all 256 sinks of block I and sweeps one 64-source quarter of
// block J: at step u lane l takes source (l - u) mod 64, so the 64 lanes
// touch 64 different sources, and the reaction sum of each source rides one
// lane up per step (a DPP wave rotation folded into the accumulating
// subtract).  After 64 steps lane l holds the reaction on source (l + 1) mod 64
// over all 256 sinks.  Diagonal tiles (I = J) are one-sided, except, in a
// disk of at most kSymSelfMaxBlocks blocks, those of full blocks: they take
// each pair once with its reaction through a per-wave choice of rows
// (sym_sweep_self).
//
// Partial-sum slots: a body of chunk k gets one slot per partner chunk p:
// as a sink of item (k, p) for p >= k, as a source of item (p, k) for p < k;
// in the diagonal item both land in slot k.  Every slot is written by exactly
// one item in a fixed order, so results do not depend on the schedule.
// Sink sums: FP32 per 64-source quarter, folded into FP64, waves combined in
// fixed order.  Reactions: FP32 over 16 sweep steps (64 terms), folded into
// FP64, accumulated over the chunk's I blocks in FP64 (LDS).
constexpr uint32_t kSymBlock = kSymBlockBodies;
constexpr uint32_t kSymS = kSymBlock / 64;
// LDS source tile: each 64-source quarter q twice in a row at entry 128 q, so
// that the source of lane l at step u, (l - u) mod 64, is entry
// 128 q + 64 + l - u: one address per lane, stepped down by one entry, and no
// wrap to compute at every step
constexpr uint32_t kSymTile = 2 * kSymBlock;

struct alignas(16) SrcZM {
    double z;
    float m, pad;
};

// acc taken from lane l - 1 (wave_ror:1), minus v
__device__ inline float rot_sub(float acc, float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(acc), 0x13C, 0xF, 0xF, false)) - v;
}

// One wave's sweep of its 64-source quarter over the first kS sink rows
// (lane l, row s: sink l + 64 s; kb: lane l's tile entry at step 0,
// 128 q + 64 + l for quarter q): sink sums into f*, reactions (when kReact)
// into r* (lane l: source (l + 1) mod 64 of the quarter).  The FP32 reaction
// sums are folded into FP64 every 16 steps (64 terms, as on the sink side):
// the FP64 running sum is rotated 16 lanes at each fold so that it stays with
// the sources the FP32 sum in the same lane belongs to.
template <bool kReact, int kS>
__device__ inline void sym_sweep(const double2* __restrict__ txy, const SrcZM* __restrict__ tzm, uint32_t kb,
                                 uint32_t lane, const double* xi, const double* yi, const double* zi,
                                 const float* mi, float* fx, float* fy, float* fz, double& rx, double& ry,
                                 double& rz)
{
    rx = ry = rz = 0.0;
    const int from16 = (int)((lane - 16u) & 63u);
    for (uint32_t u0 = 0; u0 < 64; u0 += 16) {
        float ax = 0.0f, ay = 0.0f, az = 0.0f;
        // Each phase runs over the S sinks before the next starts, so that the
        // S independent chains interleave (the FP64 subtract, the conversion
        // and the rsq each have latency the next phase would otherwise wait for).
#pragma unroll 1
        for (uint32_t u = u0; u < u0 + 16; ++u) {
            const uint32_t k = kb - u;
            const double2 a = txy[k];
            const SrcZM b = tzm[k];
            float dx[kS], dy[kS], dz[kS], r2[kS], ri3[kS];
#pragma unroll
            for (int s = 0; s < kS; ++s) {
                dx[s] = (float)(a.x - xi[s]);
                dy[s] = (float)(a.y - yi[s]);
                dz[s] = (float)(b.z - zi[s]);
            }
#pragma unroll
            for (int s = 0; s < kS; ++s) r2[s] = __builtin_fmaf(dz[s], dz[s], kEps2);
#pragma unroll
            for (int s = 0; s < kS; ++s) r2[s] = __builtin_fmaf(dy[s], dy[s], r2[s]);
#pragma unroll
            for (int s = 0; s < kS; ++s) r2[s] = __builtin_fmaf(dx[s], dx[s], r2[s]);
#pragma unroll
            for (int s = 0; s < kS; ++s) ri3[s] = __builtin_amdgcn_rsqf(r2[s]);
#pragma unroll
            for (int s = 0; s < kS; ++s) ri3[s] = (ri3[s] * ri3[s]) * ri3[s];
#pragma unroll
            for (int s = 0; s < kS; ++s) {
                const float scj = b.m * ri3[s];
                fx[s] = __builtin_fmaf(scj, dx[s], fx[s]);
                fy[s] = __builtin_fmaf(scj, dy[s], fy[s]);
                fz[s] = __builtin_fmaf(scj, dz[s], fz[s]);
            }
            if constexpr (kReact) {
                float sx = 0.0f, sy = 0.0f, sz = 0.0f;
#pragma unroll
                for (int s = 0; s < kS; ++s) {
                    const float sci = mi[s] * ri3[s];
                    sx = __builtin_fmaf(sci, dx[s], sx);
                    sy = __builtin_fmaf(sci, dy[s], sy);
                    sz = __builtin_fmaf(sci, dz[s], sz);
                }
                ax = rot_sub(ax, sx);
                ay = rot_sub(ay, sy);
                az = rot_sub(az, sz);
            }
        }
        if constexpr (kReact) {
            // lane l's FP32 sum belongs to source (l - u0 - 15) mod 64, the
            // running sum of lane l - 16 to the same source
            rx = __shfl(rx, from16) + (double)ax;
            ry = __shfl(ry, from16) + (double)ay;
            rz = __shfl(rz, from16) + (double)az;
        }
    }
}

// sym_sweep over the first nS sink rows: rows past them hold no valid sink
// (a disk's last block), and skipping them only drops terms that are exactly 0.
template <bool kReact>
__device__ inline void sym_sweep_rows(uint32_t nS, const double2* __restrict__ txy, const SrcZM* __restrict__ tzm,
                                      uint32_t kb, uint32_t lane, const double* xi, const double* yi,
                                      const double* zi, const float* mi, float* fx, float* fy, float* fz,
                                      double& rx, double& ry, double& rz)
{
    switch (nS) {
    case 1: sym_sweep<kReact, 1>(txy, tzm, kb, lane, xi, yi, zi, mi, fx, fy, fz, rx, ry, rz); break;
    case 2: sym_sweep<kReact, 2>(txy, tzm, kb, lane, xi, yi, zi, mi, fx, fy, fz, rx, ry, rz); break;
    case 3: sym_sweep<kReact, 3>(txy, tzm, kb, lane, xi, yi, zi, mi, fx, fy, fz, rx, ry, rz); break;
    default: sym_sweep<kReact, kSymS>(txy, tzm, kb, lane, xi, yi, zi, mi, fx, fy, fz, rx, ry, rz); break;
    }
}

// One step of the diagonal-tile sweep for wave W: source (a, b) against row
// slots 0..kN-1 (rows kRow[]) with reaction and, when kDiag, slot 2 on the
// sink side only.  The arithmetic per pair is sym_sweep's.
template <uint32_t W, int kN, bool kDiag>
__device__ inline void self_step(const double2& a, const SrcZM& b, const double* xi, const double* yi,
                                 const double* zi, const float* mi, float* gx, float* gy, float* gz, float& ax,
                                 float& ay, float& az)
{
    constexpr uint32_t kRow[3] = {(W + 3u) & 3u, (W + 2u) & 3u, W};
    constexpr int kR = kN + (kDiag ? 1 : 0);
    static_assert(kN >= 1 && kR <= 3, "row slots");
    float dx[kR], dy[kR], dz[kR], r2[kR], ri3[kR];
#pragma unroll
    for (int t = 0; t < kR; ++t) {
        const uint32_t r = kRow[t < kN ? t : 2];
        dx[t] = (float)(a.x - xi[r]);
        dy[t] = (float)(a.y - yi[r]);
        dz[t] = (float)(b.z - zi[r]);
    }
#pragma unroll
    for (int t = 0; t < kR; ++t) r2[t] = __builtin_fmaf(dz[t], dz[t], kEps2);
#pragma unroll
    for (int t = 0; t < kR; ++t) r2[t] = __builtin_fmaf(dy[t], dy[t], r2[t]);
#pragma unroll
    for (int t = 0; t < kR; ++t) r2[t] = __builtin_fmaf(dx[t], dx[t], r2[t]);
#pragma unroll
    for (int t = 0; t < kR; ++t) ri3[t] = __builtin_amdgcn_rsqf(r2[t]);
#pragma unroll
    for (int t = 0; t < kR; ++t) ri3[t] = (ri3[t] * ri3[t]) * ri3[t];
#pragma unroll
    for (int t = 0; t < kR; ++t) {
        const int s = t < kN ? t : 2;
        const float scj = b.m * ri3[t];
        gx[s] = __builtin_fmaf(scj, dx[t], gx[s]);
        gy[s] = __builtin_fmaf(scj, dy[t], gy[s]);
        gz[s] = __builtin_fmaf(scj, dz[t], gz[s]);
    }
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
#pragma unroll
    for (int t = 0; t < kN; ++t) {
        const float sci = mi[kRow[t]] * ri3[t];
        sx = __builtin_fmaf(sci, dx[t], sx);
        sy = __builtin_fmaf(sci, dy[t], sy);
        sz = __builtin_fmaf(sci, dz[t], sz);
    }
    ax = rot_sub(ax, sx);
    ay = rot_sub(ay, sy);
    az = rot_sub(az, sz);
}

// The diagonal tile (I = J) of a full block, with reactions, for wave W.  Its
// 16 (sink row, source quarter) blocks hold every unordered pair twice and
// every body once with itself, so wave W sweeps its quarter against three
// rows only:
//   slot 0, row W - 1 (mod 4), steps 0..63: the pairs of blocks W - 1 and W;
//   slot 2, row W, steps 1..31 and, on the sink side only, step 32: the pairs
//     inside block W (lanes l and l - 32 meet at step 32 from both ends);
//   slot 1, row W + 2 (mod 4), steps 0..31 when W >= 2 and 1..32 when W < 2:
//     of the pairs of blocks W and W + 2 (lanes p and q), wave W + 2 (W < 2)
//     takes those with (p - q) mod 64 < 32 and wave W the others.
// Each unordered pair is taken once, with its reaction, and each wave runs 127
// row-steps with reaction and one without, where the one-sided sweep of the
// four rows runs 256 without.  Sink sums go to f* by row (0 for row W + 1),
// reactions to r* as in sym_sweep: the steps run in order 0..63, so the
// rotation and the 16-step folds are sym_sweep's.
template <uint32_t W>
__device__ inline void sym_sweep_self(const double2* __restrict__ txy, const SrcZM* __restrict__ tzm,
                                      uint32_t lane, const double* xi, const double* yi, const double* zi,
                                      const float* mi, float* fx, float* fy, float* fz, double& rx, double& ry,
                                      double& rz)
{
    const uint32_t kb = W * 128u + 64u + lane;   // tile entry at step 0 (kSymTile)
    constexpr uint32_t kRow[3] = {(W + 3u) & 3u, (W + 2u) & 3u, W};
    constexpr bool kHi = W >= 2u;
    float gx[3] = {0.0f, 0.0f, 0.0f}, gy[3] = {0.0f, 0.0f, 0.0f}, gz[3] = {0.0f, 0.0f, 0.0f};
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    rx = ry = rz = 0.0;
    const int from16 = (int)((lane - 16u) & 63u);
    auto fold = [&]() {
        rx = __shfl(rx, from16) + (double)ax;
        ry = __shfl(ry, from16) + (double)ay;
        rz = __shfl(rz, from16) + (double)az;
        ax = ay = az = 0.0f;
    };
#define SELF_STEP(U, N, D)                                                                          \
    {                                                                                               \
        const uint32_t k = kb - (U);                                                                \
        self_step<W, N, D>(txy[k], tzm[k], xi, yi, zi, mi, gx, gy, gz, ax, ay, az);                 \
    }
    SELF_STEP(0u, kHi ? 2 : 1, false)
#pragma unroll 1
    for (uint32_t u = 1; u < 16; ++u) SELF_STEP(u, 3, false)
    fold();
#pragma unroll 1
    for (uint32_t u = 16; u < 32; ++u) SELF_STEP(u, 3, false)
    fold();
    SELF_STEP(32u, kHi ? 1 : 2, true)
#pragma unroll 1
    for (uint32_t u = 33; u < 48; ++u) SELF_STEP(u, 1, false)
    fold();
#pragma unroll 1
    for (uint32_t u = 48; u < 64; ++u) SELF_STEP(u, 1, false)
    fold();
#undef SELF_STEP
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        fx[kRow[t]] = gx[t];
        fy[kRow[t]] = gy[t];
        fz[kRow[t]] = gz[t];
    }
}

// Dynamic LDS: the source tile (xy, zm, each quarter twice: 16 KB; between
// tiles it also carries one wave's sink sums, 6 KB, for the combine) and the
// reaction accumulators (R x 6 KB).
__host__ __device__ constexpr size_t sym_lds_bytes(uint32_t R)
{
    return kSymTile * (sizeof(double2) + sizeof(SrcZM)) + (size_t)R * 3 * kSymBlock * sizeof(double);
}
static_assert(kSymS * 3 * 64 * sizeof(double) <= kSymTile * (sizeof(double2) + sizeof(SrcZM)),
              "the wave-combine buffer must fit in the tile");

// kSelf: the diagonal tiles of full blocks of items flagged kItemSelf go
// through sym_sweep_self; without it (and in unflagged items) they are
// one-sided.  Which diagonal tiles take reactions is a property of the item
// alone, so both instances give the same results; the engine launches this
// one only when an item is flagged (the compiler schedules the main loop of
// the instance with the extra sweep less well).
template <bool kSelf>
__global__ __launch_bounds__(kBlock) void k_force_sym(const WorkItem* __restrict__ items, uint32_t n_items,
                                                      const double* __restrict__ x, const double* __restrict__ y,
                                                      const double* __restrict__ z, const float* __restrict__ m32,
                                                      const DiskCtl* __restrict__ ctl, uint32_t need,
                                                      float* __restrict__ part, uint32_t n_all)
{
    extern __shared__ __align__(16) unsigned char sym_lds[];
    double2* txy = reinterpret_cast<double2*>(sym_lds);
    SrcZM* tzm = reinterpret_cast<SrcZM*>(txy + kSymTile);
    double* red = reinterpret_cast<double*>(sym_lds);          // [S][3][64], aliases the tile
    double* acc = reinterpret_cast<double*>(tzm + kSymTile);   // [R][3][256]
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t kb = wave * 128u + 64u + lane;   // this lane's tile entry at step 0 (kSymTile)
    for (uint32_t it = blockIdx.x; it < n_items; it += gridDim.x) {
        const WorkItem w = items[it];
        if (!(ctl[w.disk].flags & need)) continue;
        const bool diag = w.sink_lo == w.src_lo;
        const uint32_t nI = (w.sink_hi - w.sink_lo + kSymBlock - 1) / kSymBlock;
        const uint32_t nJ = (w.src_hi - w.src_lo + kSymBlock - 1) / kSymBlock;
        for (uint32_t q = tid; q < nJ * 3 * kSymBlock; q += kBlock) acc[q] = 0.0;
        for (uint32_t bi = 0; bi < nI; ++bi) {
            const uint32_t i0 = w.sink_lo + bi * kSymBlock;
            // sink rows holding at least one valid sink (fewer only in a disk's last block)
            const uint32_t nS = min(kSymS, (min(kSymBlock, w.sink_hi - i0) + 63u) / 64u);
            double xi[kSymS], yi[kSymS], zi[kSymS], ax[kSymS], ay[kSymS], az[kSymS];
            float mi[kSymS];
#pragma unroll
            for (int s = 0; s < (int)kSymS; ++s) {
                const uint32_t i = i0 + lane + 64u * s;
                const bool valid = i < w.sink_hi;
                xi[s] = valid ? x[i] : 0.0;
                yi[s] = valid ? y[i] : 0.0;
                zi[s] = valid ? z[i] : 0.0;
                mi[s] = valid ? m32[i] : 0.0f;
                ax[s] = ay[s] = az[s] = 0.0;
            }
            for (uint32_t bj = diag ? bi : 0; bj < nJ; ++bj) {
                const uint32_t j0 = w.src_lo + bj * kSymBlock;
                __syncthreads();
                {
                    const uint32_t j = j0 + tid;
                    double2 a = make_double2(0.0, 0.0);
                    SrcZM b{0.0, 0.0f, 0.0f};
                    if (j < w.src_hi) {
                        a = make_double2(x[j], y[j]);
                        b.z = z[j];
                        b.m = m32[j];
                    }
                    const uint32_t e = (tid & ~63u) * 2u + (tid & 63u);   // quarter q: 128 q + l and + 64
                    txy[e] = a;
                    txy[e + 64u] = a;
                    tzm[e] = b;
                    tzm[e + 64u] = b;
                }
                __syncthreads();
                float fx[kSymS], fy[kSymS], fz[kSymS];
#pragma unroll
                for (int s = 0; s < (int)kSymS; ++s) fx[s] = fy[s] = fz[s] = 0.0f;
                double rx = 0.0, ry = 0.0, rz = 0.0;
                const bool self = diag && bj == bi;
                // the diagonal tile of a block with all four sink rows takes each
                // pair once with its reaction; with fewer rows (a disk's last
                // block) the one-sided sweep of the valid rows is the cheaper one
                bool react = false;
                // a wave whose quarter holds no source (a disk's last block) has
                // nothing to add: its sources are padding with mass 0
                if constexpr (kSelf) {
                    react = !self || ((w.flags & kItemSelf) && nS == kSymS);
                    if (j0 + wave * 64u < w.src_hi) {
                        if (!self)
                            sym_sweep_rows<true>(nS, txy, tzm, kb, lane, xi, yi, zi, mi, fx, fy, fz, rx, ry, rz);
                        else if (!react)
                            sym_sweep_rows<false>(nS, txy, tzm, kb, lane, xi, yi, zi, mi, fx, fy, fz, rx, ry, rz);
                        else
                            switch (wave) {
                            case 0: sym_sweep_self<0>(txy, tzm, lane, xi, yi, zi, mi, fx, fy, fz, rx, ry, rz); break;
                            case 1: sym_sweep_self<1>(txy, tzm, lane, xi, yi, zi, mi, fx, fy, fz, rx, ry, rz); break;
                            case 2: sym_sweep_self<2>(txy, tzm, lane, xi, yi, zi, mi, fx, fy, fz, rx, ry, rz); break;
                            default: sym_sweep_self<3>(txy, tzm, lane, xi, yi, zi, mi, fx, fy, fz, rx, ry, rz); break;
                            }
                    }
                } else if (j0 + wave * 64u < w.src_hi) {
                    if (self)
                        sym_sweep_rows<false>(nS, txy, tzm, kb, lane, xi, yi, zi, mi, fx, fy, fz, rx, ry, rz);
                    else
                        sym_sweep_rows<true>(nS, txy, tzm, kb, lane, xi, yi, zi, mi, fx, fy, fz, rx, ry, rz);
                }
#pragma unroll
                for (int s = 0; s < (int)kSymS; ++s) {
                    ax[s] += (double)fx[s];
                    ay[s] += (double)fy[s];
                    az[s] += (double)fz[s];
                }
                if (kSelf ? react : !self) {
                    const uint32_t k = wave * 64u + ((lane + 1u) & 63u);
                    double* ab = acc + (size_t)bj * 3 * kSymBlock;
                    ab[k] += rx;
                    ab[kSymBlock + k] += ry;
                    ab[2 * kSymBlock + k] += rz;
                }
            }
            // sinks of block I: ((w0 + w1) + w2) + w3 through the tile buffer
            for (uint32_t src = 1; src < 4; ++src) {
                __syncthreads();
                if (wave == src) {
#pragma unroll
                    for (int s = 0; s < (int)kSymS; ++s) {
                        double* r = red + (size_t)s * 3 * 64;
                        r[lane] = ax[s];
                        r[64 + lane] = ay[s];
                        r[128 + lane] = az[s];
                    }
                }
                __syncthreads();
                if (wave == 0) {
#pragma unroll
                    for (int s = 0; s < (int)kSymS; ++s) {
                        const double* r = red + (size_t)s * 3 * 64;
                        ax[s] += r[lane];
                        ay[s] += r[64 + lane];
                        az[s] += r[128 + lane];
                    }
                }
            }
            if (wave == 0) {
#pragma unroll
                for (int s = 0; s < (int)kSymS; ++s) {
                    const double sx = ax[s], sy = ay[s], sz = az[s];
                    const uint32_t k = lane + 64u * s;
                    if (diag) {
                        // same slot as the reactions this block got from I' < I
                        double* ab = acc + (size_t)bi * 3 * kSymBlock;
                        ab[k] += sx;
                        ab[kSymBlock + k] += sy;
                        ab[2 * kSymBlock + k] += sz;
                    } else if (i0 + k < w.sink_hi) {
                        part[(size_t)(w.slot * 3 + 0) * n_all + i0 + k] = (float)sx;
                        part[(size_t)(w.slot * 3 + 1) * n_all + i0 + k] = (float)sy;
                        part[(size_t)(w.slot * 3 + 2) * n_all + i0 + k] = (float)sz;
                    }
                }
            }
        }
        // reactions (and, on the diagonal, sink sums) of chunk c -> slot a
        __syncthreads();
        for (uint32_t q = tid; q < nJ * kSymBlock; q += kBlock) {
            const uint32_t j = w.src_lo + q;
            if (j < w.src_hi) {
                const double* ab = acc + (size_t)(q / kSymBlock) * 3 * kSymBlock + q % kSymBlock;
                part[(size_t)(w.src_slot * 3 + 0) * n_all + j] = (float)ab[0];
                part[(size_t)(w.src_slot * 3 + 1) * n_all + j] = (float)ab[kSymBlock];
                part[(size_t)(w.src_slot * 3 + 2) * n_all + j] = (float)ab[2 * kSymBlock];
            }
        }
        __syncthreads();
    }
}

hipError_t launch_force_sym(hipStream_t st, uint32_t R, bool self, const WorkItem* items, uint32_t n_items,
                            uint32_t grid, const double* x, const double* y, const double* z, const float* m32,
                            const DiskCtl* ctl, uint32_t need, float* part, uint32_t n_all)
{
    if (n_items == 0) return hipSuccess;
    if (R < 1 || R > kSymMaxR) return hipErrorInvalidValue;
    if (self)
        hipLaunchKernelGGL(k_force_sym<true>, dim3(grid), dim3(kBlock), sym_lds_bytes(R), st, items, n_items, x, y, z,
                           m32, ctl, need, part, n_all);
    else
        hipLaunchKernelGGL(k_force_sym<false>, dim3(grid), dim3(kBlock), sym_lds_bytes(R), st, items, n_items, x, y,
                           z, m32, ctl, need, part, n_all);
    return hipGetLastError();
}

size_t force_sym_lds_bytes(uint32_t R) { return sym_lds_bytes(R); }

hipError_t launch_force_exact(hipStream_t st, int lane, const WorkItem* items, uint32_t n_items, uint32_t grid,
                              const double* x, const double* y, const double* z, const double* m_ex,
                              const DiskCtl* ctl, uint32_t need, double* part, uint32_t n_all)
{
    if (n_items == 0) return hipSuccess;
    if (lane == 1)
        hipLaunchKernelGGL(k_force_exact_lane, dim3(grid), dim3(kBlock), 0, st, items, n_items, x, y, z, m_ex, ctl,
                           need, part, n_all);
    else if (lane == 2)
        hipLaunchKernelGGL(k_force_exact<kExactSinksSmall>, dim3(grid), dim3(kExactBlock), 0, st, items, n_items, x,
                           y, z, m_ex, ctl, need, part, n_all);
    else
        hipLaunchKernelGGL(k_force_exact<kExactSinks>, dim3(grid), dim3(kExactBlock), 0, st, items, n_items, x, y,
                           z, m_ex, ctl, need, part, n_all);
    return hipGetLastError();
}

hipError_t launch_force_exact_rk4(hipStream_t st, int r, const WorkItem* items, uint32_t n_items, uint32_t grid,
                                  const double* x, const double* y, const double* z, const double* m_ex,
                                  const DiskCtl* ctl, uint32_t need, const StageArgs& a)
{
    if (n_items == 0) return hipSuccess;
#define EXACT_RK4(R)                                                                                            \
    hipLaunchKernelGGL((k_force_exact_rk4<kExactSinksSmall, R>), dim3(grid), dim3(kExactBlock), 0, st, items,   \
                       n_items, x, y, z, m_ex, ctl, need, a)
    switch (r) {
    case 0: EXACT_RK4(0); break;
    case 1: EXACT_RK4(1); break;
    case 2: EXACT_RK4(2); break;
    case 3: EXACT_RK4(3); break;
    default: return hipErrorInvalidValue;
    }
#undef EXACT_RK4
    return hipGetLastError();
}

hipError_t launch_force_mixed(hipStream_t st, int S, const WorkItem* items, uint32_t n_items, uint32_t grid,
                              const double* x, const double* y, const double* z, const float* m32,
                              const DiskCtl* ctl, uint32_t need, float* part, uint32_t n_all)
{
    if (n_items == 0) return hipSuccess;
    switch (S) {
    case 1: hipLaunchKernelGGL(k_force_mixed<1>, dim3(grid), dim3(kBlock), 0, st, items, n_items, x, y, z, m32, ctl, need, part, n_all); break;
    case 2: hipLaunchKernelGGL(k_force_mixed<2>, dim3(grid), dim3(kBlock), 0, st, items, n_items, x, y, z, m32, ctl, need, part, n_all); break;
    case 4: hipLaunchKernelGGL(k_force_mixed<4>, dim3(grid), dim3(kBlock), 0, st, items, n_items, x, y, z, m32, ctl, need, part, n_all); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Per-body stage kernels.  One thread per (body, component): every RK
// combination is separable by component, so the grid is 3 x ceil(n/64) and a
// wave holds 64 consecutive bodies of one component (coalesced).  All loads
// are issued before the first store.
// ---------------------------------------------------------------------------
// pp_disk.cu:60-84
__device__ double reduction_factor(const GasDev& g, double t)
{
    switch (g.gas_decrease) {
    case 0: return 1.0;
    case 1:
        if (t <= g.t0) return 1.0;
        else if (g.t0 < t && t <= g.t1 && g.t0 != g.t1) return 1.0 - (t - g.t0) / (g.t1 - g.t0);
        else return 0.0;
    case 2: return exp(-(t - g.t0) / g.e_folding_time);
    default: return 1.0;
    }
}

// Quadratic drag a += f u with the analytic gas disk (pp_disk.cu:87-155,
// 486-521); returns component q of f u.
__device__ double drag_component(const GasDev& g, double decr, double m_star, double rx, double ry, double rz,
                                 double vx, double vy, double vz, double cd, double radius, double density, int q)
{
    // circular_velocity (pp_disk.cu:108-142)
    double gx = 0.0, gy = 0.0;
    const double r = sqrt(SQR(rx) + SQR(ry));
    const double vc = sqrt(m_star / r);
    if (rx == 0.0 && ry == 0.0) {
    } else if (ry == 0.0) {
        gy = rx > 0.0 ? vc : -vc;
    } else if (rx == 0.0) {
        gx = ry > 0.0 ? -vc : vc;
    } else if (rx >= ry) {
        const double p = ry / rx;
        gy = rx >= 0 ? vc / sqrt(1.0 + SQR(p)) : -vc / sqrt(1.0 + SQR(p));
        gx = -gy * p;
    } else {
        const double p = rx / ry;
        gx = ry >= 0 ? -vc / sqrt(1.0 + SQR(p)) : vc / sqrt(1.0 + SQR(p));
        gy = -gx * p;
    }
    // get_velocity (pp_disk.cu:145-155)
    const double fac = sqrt(1.0 - 2.0 * g.eta0 * pow(r, g.eta1));
    gx *= fac;
    gy *= fac;
    const double ux = gx - vx, uy = gy - vy, uz = 0.0 - vz;
    const double u_n = sqrt(SQR(ux) + SQR(uy) + SQR(uz));
    // get_density (pp_disk.cu:87-105)
    const double h = g.sch0 * pow(r, g.sch1);
    const double arg = SQR(rz / h);
    double dens;
    if (0.1 < r) {
        dens = g.rho0 * pow(r, g.rho1) * exp(-arg);
    } else {
        const double a = g.rho0 * pow(0.1, g.rho1 - 4.0);
        dens = a * SQR(SQR(r)) * exp(-arg);
    }
    const double f = decr * (3.0 * cd * dens * u_n) / (8.0 * radius * density);
    return f * (q == 0 ? ux : (q == 1 ? uy : uz));
}

// Cell of the FARGO grid under (x, y) (calc_linear_index, pp_disk.cu:158-193),
// from r = sqrt(x^2 + y^2) and alpha = atan2(y, x) taken into [0, 2 pi] as in
// the reference.  Inside the grid, ring k holds g_rad[k] <= r < g_rad[k + 1]
// (bisection over the increasing radii finds the k of the reference's linear
// scan) and the sector is floor(alpha / (2 pi / n_sec)).  r below the first
// interface reads cell 0, above the last reads the last cell; r equal to the
// last interface (or NaN) fits no half-open ring, and the reference's scan
// then leaves ring 0.  A sector of n_sec (alpha rounded to 2 pi) reads the
// next ring's first cell, as the flat index does; the one index the
// reference would read past the grid (there, and a negative alpha from
// atan2(-0, x < 0), which it converts with undefined behaviour) is held
// inside it: the last cell and sector 0.
__device__ uint32_t grid_cell(const GasDev& g, double r, double alpha)
{
    const uint32_t n_cell = g.n_rad * g.n_sec;
    if (g.g_rad[0] > r) return 0;
    if (g.g_rad[g.n_rad] < r) return n_cell - 1;
    uint32_t ring = 0;
    if (g.g_rad[g.n_rad] != r) {
        uint32_t lo = 0, hi = g.n_rad;   // g_rad[lo] <= r < g_rad[hi]
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (g.g_rad[mid] <= r) lo = mid;
            else hi = mid;
        }
        ring = lo;
    }
    const double q = alpha / (kTwoPi / g.n_sec);
    const uint32_t sec = q >= 0.0 ? (uint32_t)q : 0u;
    const uint32_t cell = ring * g.n_sec + sec;
    return cell < n_cell ? cell : n_cell - 1;
}

// Quadratic drag a += f u in a FARGO gas grid, component q (the reference's
// CPU loop, pp_disk.cu:778-814; its GPU kernel reads only the analytic
// disk's parameters).  The gas moves at the circular speed sqrt(m_star / r)
// plus the cell's azimuthal velocity, with the cell's radial velocity,
// rotated from (r, theta) to (x, y) (rotate_2D_vector, util.cu:76-83); no
// reduction factor applies.
__device__ double grid_drag_component(const GasDev& g, double m_star, double rx, double ry, double vx, double vy,
                                      double vz, double cd, double radius, double density, int q)
{
    const double r = sqrt(SQR(rx) + SQR(ry));
    const double at = atan2(ry, rx);
    const double theta = ry >= 0.0 ? at : kTwoPi + at;
    const double4 cell = g.g_cell[grid_cell(g, r, theta)];
    const double vc_theta = sqrt(m_star / r);
    const double vg_theta = vc_theta + cell.z;
    const double vg_rad = cell.y;
    const double ct = cos(theta), st = sin(theta);
    const double gx = ct * vg_rad - st * vg_theta;
    const double gy = st * vg_rad + ct * vg_theta;
    const double rho = cell.x;
    const double ux = gx - vx, uy = gy - vy, uz = 0.0 - vz;
    const double u_n = sqrt(SQR(ux) + SQR(uy) + SQR(uz));
    const double f = (3.0 * cd * rho * u_n) / (8.0 * radius * density);
    return f * (q == 0 ? ux : (q == 1 ? uy : uz));
}

__device__ inline const double* kv_of(const StageArgs& a, int r, int q)
{
    return r == 0 ? a.vel[q] : a.K + ((size_t)r * 6 + q) * a.n_all;
}
__device__ inline const double* ka_of(const StageArgs& a, int r, int q)
{
    return a.K + ((size_t)r * 6 + 3 + q) * a.n_all;
}
__device__ inline double* kv_w(const StageArgs& a, int r, int q) { return a.K + ((size_t)r * 6 + q) * a.n_all; }
__device__ inline double* ka_w(const StageArgs& a, int r, int q) { return a.K + ((size_t)r * 6 + 3 + q) * a.n_all; }

// Sum of the pair partials of body i, component q, in slot order, in FP64.
// The FP64 pair kernel leaves one FP64 slot; the mixed-precision kernels leave
// one FP32 slot per source chunk (the chunk's total, formed in FP64 and
// rounded once), at the same index in float units.  Loads go out in batches
// of 128 bytes per lane (16 doubles or 32 floats); a slot index past the last
// is clamped to it, so every load is unconditional and in bounds and only the
// sum skips it (a guarded load would become a branch and a wait per slot).
template <typename T>
__device__ double partial_sum_t(const T* __restrict__ p, size_t st, uint32_t nc)
{
    constexpr uint32_t kB = 128 / sizeof(T);
    double s = 0.0;
    for (uint32_t c = 0; c < nc; c += kB) {
        T v[kB];
#pragma unroll
        for (uint32_t u = 0; u < kB; ++u) v[u] = p[(size_t)min(c + u, nc - 1) * st];
#pragma unroll
        for (uint32_t u = 0; u < kB; ++u)
            s = (c + u >= nc) ? s : ((c + u == 0) ? (double)v[u] : s + (double)v[u]);
    }
    return s;
}

__device__ double partial_sum(const StageArgs& a, uint32_t i, int q, uint32_t nc)
{
    const size_t off = (size_t)q * a.n_all + i, st = (size_t)3 * a.n_all;
    if (a.mixed) return partial_sum_t(reinterpret_cast<const float*>(a.part) + off, st, nc);
    return partial_sum_t(a.part + off, st, nc);
}

// partial_sum for a launch over one disk: the loop runs to ncmax, the
// disk's largest slot count (uniform, from the kernel arguments), and each
// lane adds its own nc slots in slot order.  The first batch of loads then
// waits for nothing.  Slots in [nc, ncmax) lie inside the partial buffer,
// which holds the largest slot count of any body.  A disk's stage kernels
// are few waves (one per 64 bodies and component), so a batch as wide as the
// register budget allows costs no occupancy and saves round trips.
template <typename T, uint32_t kB>
__device__ double partial_sum_upto_t(const T* __restrict__ p, size_t st, uint32_t nc, uint32_t ncmax)
{
    double s = 0.0;
    for (uint32_t c = 0; c < ncmax; c += kB) {
        T v[kB];
#pragma unroll
        for (uint32_t u = 0; u < kB; ++u) v[u] = p[(size_t)min(c + u, ncmax - 1) * st];
#pragma unroll
        for (uint32_t u = 0; u < kB; ++u)
            s = (c + u >= nc) ? s : ((c + u == 0) ? (double)v[u] : s + (double)v[u]);
    }
    return s;
}

__device__ double partial_sum_upto(const StageArgs& a, uint32_t i, int q, uint32_t nc, uint32_t ncmax)
{
    const size_t off = (size_t)q * a.n_all + i, st = (size_t)3 * a.n_all;
    // one batch of 72 FP32 slots covers a 65-chunk disk (16385 bodies) at once
    if (a.mixed) return partial_sum_upto_t<float, 72>(reinterpret_cast<const float*>(a.part) + off, st, nc, ncmax);
    return partial_sum_upto_t<double, 16>(a.part + off, st, nc, ncmax);
}

// Mixed mode: the FP64 prefix sources' (the star block's) acceleration on
// body i, component q, reference formula and order; 0 in FP64 mode.
// pos: force-input positions; s0, if given: the first prefix source's
// position, mass and id, loaded by the caller.
struct PrefixBody {
    double x, y, z, m;
    int id;
};

__device__ double prefix_accel(const StageArgs& a, uint32_t i, int q, const DiskInfo& di,
                               const double* __restrict__ const* pos, const PrefixBody* s0 = nullptr)
{
    if (!a.mixed) return 0.0;
    const uint32_t li = i - di.off;
    const int cls = li < di.n_si ? 0 : (li < di.n_massive ? 1 : 2);
    const double xi = pos[0][i], yi = pos[1][i], zi = pos[2][i];
    const uint32_t src_hi = cls == 1 ? di.n_si : di.n_massive;
    const uint32_t hi = min(di.n_fp64, src_hi);
    double b = 0.0;
    for (uint32_t lj = 0; lj < hi; ++lj) {
        const uint32_t j = di.off + lj;
        const bool pre = s0 && lj == 0;
        if (j == i || (pre ? s0->id : a.id[j]) < 0) continue;
        const double dx = (pre ? s0->x : pos[0][j]) - xi;
        const double dy = (pre ? s0->y : pos[1][j]) - yi;
        const double dz = (pre ? s0->z : pos[2][j]) - zi;
        const double r2 = SQR(dx) + SQR(dy) + SQR(dz);
        const double d = sqrt(r2);
        const double sc = (pre ? s0->m : a.mass[j]) / (d * r2);
        b += sc * (q == 0 ? dx : (q == 1 ? dy : dz));
    }
    return b;
}

// Component q of the acceleration of body i at this stage from the FP64
// prefix part b and the summed partials s: b + s in mixed mode, zero for
// inactive sinks, plus drag.  pos: force-input positions; vel: velocities of
// this stage.
__device__ double stage_accel(const StageArgs& a, uint32_t i, int q, const DiskInfo& di, double t, double b, double s,
                              const double* __restrict__ const* pos, const double* __restrict__ const* vel)
{
    const uint32_t li = i - di.off;
    if (a.mixed) s = b + s;
    if (!(a.id[i] > 0)) s = 0.0;
    if (a.gas.on == kGasAnalytic && li >= di.drag_lo && li < di.drag_hi) {
        const double decr = reduction_factor(a.gas, t);
        if (!(1.0e-6 > decr))
            s += drag_component(a.gas, decr, a.mass[di.off], pos[0][i], pos[1][i], pos[2][i], vel[0][i], vel[1][i],
                                vel[2][i], a.cd[i], a.radius[i], a.density[i], q);
    } else if (a.gas.on == kGasGrid && li >= di.drag_lo && li < di.drag_hi) {
        s += grid_drag_component(a.gas, a.mass[di.off], pos[0][i], pos[1][i], vel[0][i], vel[1][i], vel[2][i],
                                 a.cd[i], a.radius[i], a.density[i], q);
    }
    return s;
}

// Pins already-issued loads above the per-disk control branch (the compiler
// would otherwise sink them into it and serialise them behind the control
// loads).  The asm is empty: it only marks the values as used here.
__device__ inline void pin(double& v) { asm volatile("" : "+v"(v)); }

__device__ inline uint32_t nchunk_of(const DiskInfo& di, uint32_t i)
{
    const uint32_t li = i - di.off;
    return li < di.n_si ? di.nchunk[0] : (li < di.n_massive ? di.nchunk[1] : di.nchunk[2]);
}

// Column values of stage row R (int_rungekutta8.cu:82-246) for body i,
// component q; the column R_DONE (just computed) is filled in later.
template <int R, int R_DONE>
__device__ void rk8_load_row(const StageArgs& a, uint32_t i, int q, double* kv, double* ka)
{
    constexpr RowDef row = rk8_row(R);
#pragma unroll
    for (int c = 0; c < row.n; ++c) {
        if (row.col[c] != R_DONE) {
            kv[c] = kv_of(a, row.col[c], q)[i];
            ka[c] = ka_of(a, row.col[c], q)[i];
        }
    }
}
// y + h (sum_c a_c k_c), left to right over the nonzero columns.
template <int R, int R_DONE>
__device__ void rk8_comb_q(double y, double v, double h, double* kv, double* ka, double kRv, double kRa,
                           double& pos, double& vel)
{
    constexpr RowDef row = rk8_row(R);
#pragma unroll
    for (int c = 0; c < row.n; ++c)
        if (row.col[c] == R_DONE) {
            kv[c] = kRv;
            ka[c] = kRa;
        }
    double sv = row.a[0] * kv[0], sa = row.a[0] * ka[0];
#pragma unroll
    for (int c = 1; c < row.n; ++c) {
        sv = sv + row.a[c] * kv[c];
        sa = sa + row.a[c] * ka[c];
    }
    pos = y + h * sv;
    vel = v + h * sa;
}

// Wave-level max into the per-disk error slot; waves whose lanes span two
// disks fall back to one atomic per lane.  The max takes a term only if it
// compares greater, as the reference's max loop does (integrator.cu:248-263),
// so a NaN term is never the maximum; NaN and infinite terms instead raise the
// disk's non-finite flag a.errf[d] (read only by the opt-in step guard).
__device__ void error_reduce(const StageArgs& a, bool valid, uint32_t d, double e)
{
    unsigned long long bits = (valid && e == e) ? (unsigned long long)__double_as_longlong(e) : 0ull;
    const bool bad = valid && !(e <= DBL_MAX);
    const uint32_t d0 = __builtin_amdgcn_readfirstlane(valid ? d : 0xffffffffu);
    const bool same = __all(!valid || d == d0);
    if (same) {
        for (int off = 32; off > 0; off >>= 1) {
            const unsigned long long o = __shfl_xor(bits, off, 64);
            bits = o > bits ? o : bits;
        }
        const bool any_bad = __any(bad);
        if ((threadIdx.x & 63u) == 0 && d0 != 0xffffffffu) {
            if (bits) atomicMax(&a.err[d0], bits);
            if (any_bad) atomicOr(&a.errf[d0], 1ull);
        }
    } else if (valid) {
        if (bits) atomicMax(&a.err[d], bits);
        if (bad) atomicOr(&a.errf[d], 1ull);
    }
}

// Single-disk form of error_reduce: the wave's max goes to its own slot,
// a.errw[blockIdx.y * gridDim.x + blockIdx.x] (one wave per block), with the
// non-finite flag in bit 63 (free: a non-negative double's bits are below it),
// and the host takes the max of the slots.  One same-address atomic per wave
// would serialise hundreds of waves on one word.
__device__ void error_store_wave(const StageArgs& a, bool valid, double e)
{
    unsigned long long bits = (valid && e == e) ? (unsigned long long)__double_as_longlong(e) : 0ull;
    const bool bad = valid && !(e <= DBL_MAX);
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(bits, off, 64);
        bits = o > bits ? o : bits;
    }
    if (__any(bad)) bits |= kErrNonFinite;
    if ((threadIdx.x & 63u) == 0) a.errw[blockIdx.y * gridDim.x + blockIdx.x] = bits;
}

// Stage kernels over a single disk (ONE = true).  The disk's info comes in
// the kernel arguments and its control load goes out with the per-body
// loads; every value is then computed without a branch, and a stage that
// does not apply to the disk (its flags) stores into a.trash instead of the
// state.  With no branch to sink them into, the per-body loads, the control
// load and the first batch of partials share one memory round trip (a branch
// on the flags would put the control load's round trip in front of them).
// Stage 0 is the exception: it returns early when k0 is reused after a
// rejected attempt, since it then has nothing to do.

// Summed partials ps and FP64 prefix part b of body i's acceleration, for a
// single-disk stage.  Called after the body's own loads are issued, so that
// the first batch of partials goes out with them.
__device__ __forceinline__ void one_disk_sums(const StageArgs& a, uint32_t i, int q, double& ps, double& b)
{
    const DiskInfo& di = a.info0;
    // slot count of body i's class, as arithmetic on the uniform counts (a
    // dynamic index into nchunk would put the info in LDS)
    const uint32_t li = i - di.off;
    const uint32_t nc = di.nchunk[0] + (li >= di.n_si ? di.nchunk[1] - di.nchunk[0] : 0u) +
                        (li >= di.n_massive ? di.nchunk[2] - di.nchunk[1] : 0u);
    const uint32_t ncmax = max(di.nchunk[0], max(di.nchunk[1], di.nchunk[2]));
    const uint32_t j0 = di.off;
    const PrefixBody s0{a.fpos[0][j0], a.fpos[1][j0], a.fpos[2][j0], a.mass[j0], a.id[j0]};
    ps = partial_sum_upto(a, i, q, nc, ncmax);
    b = prefix_accel(a, i, q, di, a.fpos, &s0);
}

// RKF7(8) stage R over a single disk.
template <int R>
__device__ __forceinline__ void stage_rk8_one(const StageArgs& a)
{
    const uint32_t i0 = blockIdx.x * kBodyBlock + threadIdx.x;
    const bool in = i0 < a.n_all;
    const uint32_t i = in ? i0 : a.n_all - 1;
    const int q = blockIdx.y;
    const DiskInfo& di = a.info0;
    const DiskCtl c = a.ctl[0];
    if constexpr (R == 0)
        if (!(c.flags & F_NEED_K0)) return;
    const double yy = a.pos[q][i], vv = a.vel[q][i];
    const double kk = kv_of(a, R, q)[i];
    constexpr int NR = (R >= 1 && R <= 9) ? R + 1 : (R == 10 ? 11 : (R == 11 ? 12 : 1));
    double kv[9], ka[9];
    if constexpr (R >= 1 && R <= 11) rk8_load_row<NR, R>(a, i, q, kv, ka);
    double y10v[6], y10a[6];
    if constexpr (R == 10) {
        const int cols[6] = {0, 5, 6, 7, 8, 9};
#pragma unroll
        for (int u = 0; u < 6; ++u) {
            y10v[u] = kv_of(a, cols[u], q)[i];
            y10a[u] = ka_of(a, cols[u], q)[i];
        }
    }
    double e0v = 0, e10v = 0, e11v = 0, e0a = 0, e10a = 0, e11a = 0;
    if constexpr (R == 12) {
        e0v = kv_of(a, 0, q)[i]; e10v = kv_of(a, 10, q)[i]; e11v = kv_of(a, 11, q)[i];
        e0a = ka_of(a, 0, q)[i]; e10a = ka_of(a, 10, q)[i]; e11a = ka_of(a, 11, q)[i];
    }
    double ps, b;
    one_disk_sums(a, i, q, ps, b);
    const bool valid = in && (c.flags & (R == 0 ? F_NEED_K0 : F_ACTIVE)) != 0;
    const double h = c.h;
    const double t = c.t + rk8_c(R) * h;
    const double* vR[3] = {kv_of(a, R, 0), kv_of(a, R, 1), kv_of(a, R, 2)};
    const double kRa = stage_accel(a, i, q, di, t, b, ps, a.fpos, vR);
    double* const trash = a.trash + (size_t)q * kBodyBlock + threadIdx.x;
    auto at = [&](double* p) { return valid ? p + i : trash; };
    *at(ka_w(a, R, q)) = kRa;
    if constexpr (R >= 1 && R <= 9) {
        double p, w;
        rk8_comb_q<NR, R>(yy, vv, h, kv, ka, kk, kRa, p, w);
        *at(a.npos[q]) = p;
        *at(kv_w(a, NR, q)) = w;
    } else if constexpr (R == 10) {
        // y_{n+1} = y + h (b0 k0 + b5 k5 + b6 (k6 + k7) + b8 (k8 + k9) + b10 k10) (int_rungekutta8.cu:264-274)
        constexpr double b0 = 41.0 / 840.0, b5 = 34.0 / 105.0, b6 = 9.0 / 35.0, b8 = 9.0 / 280.0,
                         b10 = 41.0 / 840.0;
        *at(a.out[q]) = yy + h * (b0 * y10v[0] + b5 * y10v[1] + b6 * (y10v[2] + y10v[3]) +
                                 b8 * (y10v[4] + y10v[5]) + b10 * kk);
        *at(a.out[3 + q]) = vv + h * (b0 * y10a[0] + b5 * y10a[1] + b6 * (y10a[2] + y10a[3]) +
                                     b8 * (y10a[4] + y10a[5]) + b10 * kRa);
        if (a.adaptive) {
            double p, w;
            rk8_comb_q<11, R>(yy, vv, h, kv, ka, kk, kRa, p, w);
            *at(a.npos[q]) = p;
            *at(kv_w(a, 11, q)) = w;
        }
    } else if constexpr (R == 11) {
        double p, w;
        rk8_comb_q<12, R>(yy, vv, h, kv, ka, kk, kRa, p, w);
        *at(a.npos[q]) = p;
        *at(kv_w(a, 12, q)) = w;
    } else if constexpr (R == 12) {
        const uint32_t li = i - di.off;
        double e = 0.0;
        // |f0 + f10 - f11 - f12| (int_rungekutta8.cu:250-260)
        if (valid && (a.err_all || li < di.err_hi))
            e = fmax(fabs(e0v + e10v - e11v - kk), fabs(e0a + e10a - e11a - kRa));
        error_store_wave(a, valid, e);
    }
}

// RKF7(8) stage kernel for stage R: finishes k_R (partials, FP64 prefix,
// drag) and forms what the next stage needs from body i's own k values.
// Loads whose address depends only on i are issued before the per-disk
// control is known.
template <int R, bool ONE>
__global__ __launch_bounds__(kBodyBlock) void k_stage_rk8(StageArgs a)
{
    if constexpr (ONE) {
        stage_rk8_one<R>(a);
        return;
    }
    const uint32_t i0 = blockIdx.x * kBodyBlock + threadIdx.x;
    const bool in = i0 < a.n_all;
    const uint32_t i = in ? i0 : a.n_all - 1;
    const int q = blockIdx.y;
    // first the loads whose address depends on i only
    const uint32_t d = a.disk_of[i];
    const double y = a.pos[q][i], v = a.vel[q][i];
    const double kRv = kv_of(a, R, q)[i];
    constexpr int NR = (R >= 1 && R <= 9) ? R + 1 : (R == 10 ? 11 : (R == 11 ? 12 : 1));
    double kv[9], ka[9];
    if constexpr (R >= 1 && R <= 11) rk8_load_row<NR, R>(a, i, q, kv, ka);
    double y10v[6], y10a[6];
    if constexpr (R == 10) {
        const int cols[6] = {0, 5, 6, 7, 8, 9};
#pragma unroll
        for (int u = 0; u < 6; ++u) {
            y10v[u] = kv_of(a, cols[u], q)[i];
            y10a[u] = ka_of(a, cols[u], q)[i];
        }
    }
    double e0v = 0, e10v = 0, e11v = 0, e0a = 0, e10a = 0, e11a = 0;
    if constexpr (R == 12) {
        e0v = kv_of(a, 0, q)[i]; e10v = kv_of(a, 10, q)[i]; e11v = kv_of(a, 11, q)[i];
        e0a = ka_of(a, 0, q)[i]; e10a = ka_of(a, 10, q)[i]; e11a = ka_of(a, 11, q)[i];
        pin(e0v); pin(e10v); pin(e11v); pin(e0a); pin(e10a); pin(e11a);
    }
    if constexpr (R >= 1 && R <= 11) {
        constexpr RowDef row = rk8_row(NR);
#pragma unroll
        for (int c = 0; c < row.n; ++c)
            if (row.col[c] != R) {
                pin(kv[c]);
                pin(ka[c]);
            }
    }
    if constexpr (R == 10) {
#pragma unroll
        for (int u = 0; u < 6; ++u) {
            pin(y10v[u]);
            pin(y10a[u]);
        }
    }
    double yy = y, vv = v, kk = kRv;
    pin(yy); pin(vv); pin(kk);
    // then the per-disk control
    const DiskCtl c = a.ctl[d];
    const DiskInfo di = a.info[d];
    const bool valid = in && (c.flags & (R == 0 ? F_NEED_K0 : F_ACTIVE)) != 0;
    double e = 0.0;
    if (valid) {
        const double h = c.h;
        const double t = c.t + rk8_c(R) * h;
        const double* vR[3] = {kv_of(a, R, 0), kv_of(a, R, 1), kv_of(a, R, 2)};
        const double b = prefix_accel(a, i, q, di, a.fpos);
        const double kRa = stage_accel(a, i, q, di, t, b, partial_sum(a, i, q, nchunk_of(di, i)), a.fpos, vR);
        ka_w(a, R, q)[i] = kRa;
        if constexpr (R >= 1 && R <= 9) {
            double p, w;
            rk8_comb_q<NR, R>(yy, vv, h, kv, ka, kk, kRa, p, w);
            a.npos[q][i] = p;
            kv_w(a, NR, q)[i] = w;
        } else if constexpr (R == 10) {
            // y_{n+1} = y + h (b0 k0 + b5 k5 + b6 (k6 + k7) + b8 (k8 + k9) + b10 k10) (int_rungekutta8.cu:264-274)
            constexpr double b0 = 41.0 / 840.0, b5 = 34.0 / 105.0, b6 = 9.0 / 35.0, b8 = 9.0 / 280.0,
                             b10 = 41.0 / 840.0;
            a.out[q][i] = yy + h * (b0 * y10v[0] + b5 * y10v[1] + b6 * (y10v[2] + y10v[3]) +
                                   b8 * (y10v[4] + y10v[5]) + b10 * kk);
            a.out[3 + q][i] = vv + h * (b0 * y10a[0] + b5 * y10a[1] + b6 * (y10a[2] + y10a[3]) +
                                       b8 * (y10a[4] + y10a[5]) + b10 * kRa);
            if (a.adaptive) {
                double p, w;
                rk8_comb_q<11, R>(yy, vv, h, kv, ka, kk, kRa, p, w);
                a.npos[q][i] = p;
                kv_w(a, 11, q)[i] = w;
            }
        } else if constexpr (R == 11) {
            double p, w;
            rk8_comb_q<12, R>(yy, vv, h, kv, ka, kk, kRa, p, w);
            a.npos[q][i] = p;
            kv_w(a, 12, q)[i] = w;
        } else if constexpr (R == 12) {
            const uint32_t li = i - di.off;
            // |f0 + f10 - f11 - f12| (int_rungekutta8.cu:250-260)
            if (a.err_all || li < di.err_hi) e = fmax(fabs(e0v + e10v - e11v - kk), fabs(e0a + e10a - e11a - kRa));
        }
    }
    if constexpr (R == 12) error_reduce(a, valid, d, e);
}

// Classical RK4 stage R over a single disk, for body i and component q.
// kFused: called by the FP64 pair kernel's summing lane with the body's
// complete FP64 sum, for a body of a gas-free disk (fixed step: no stage 4),
// and returns early where the stage does not apply; stores go to the arrays.
// Otherwise a stage-kernel thread: the sums come from the partial-sum slots,
// and a thread past the end or outside the attempt stores into its discard
// slot, so no branch sits between the loads and the stores.
template <int R, bool kFused>
__device__ __forceinline__ void stage_rk4_body(const StageArgs& a, uint32_t i, int q, bool in, double ps_fused)
{
    static_assert(!kFused || R <= 3, "fused stages are the fixed-step ones");
    const DiskInfo& di = a.info0;
    const DiskCtl c = a.ctl[0];
    if constexpr (R == 0)
        if (!(c.flags & F_NEED_K0)) return;
    const double yy = a.pos[q][i], vv = a.vel[q][i];
    const double kk = kv_of(a, R, q)[i];
    double k0v = 0, k1v = 0, k2v = 0, k0a = 0, k1a = 0, k2a = 0, k3v = 0, k3a = 0;
    if constexpr (R == 3) {
        k0v = kv_of(a, 0, q)[i]; k1v = kv_of(a, 1, q)[i]; k2v = kv_of(a, 2, q)[i];
        k0a = ka_of(a, 0, q)[i]; k1a = ka_of(a, 1, q)[i]; k2a = ka_of(a, 2, q)[i];
    }
    if constexpr (R == 4) {
        k3v = kv_of(a, 3, q)[i];
        k3a = ka_of(a, 3, q)[i];
    }
    double ps, b;
    if constexpr (kFused) {
        ps = ps_fused;   // the FP64 kernel's one slot: no prefix in FP64 mode
        b = 0.0;
    } else {
        one_disk_sums(a, i, q, ps, b);
    }
    const bool valid = in && (c.flags & (R == 0 ? F_NEED_K0 : F_ACTIVE)) != 0;
    if constexpr (kFused)
        if (!valid) return;
    const double h = c.h;
    const double t = c.t + rk4_c(R) * h;
    const double* vR[3] = {kv_of(a, R, 0), kv_of(a, R, 1), kv_of(a, R, 2)};
    // the fused form runs only without gas (its drag code would take the
    // registers the pair pass needs): the reference's sum, zero for an
    // inactive sink (stage_accel without the drag)
    double kRa;
    if constexpr (kFused) {
        (void)vR;
        (void)t;
        kRa = a.id[i] > 0 ? ps : 0.0;
    } else {
        kRa = stage_accel(a, i, q, di, t, b, ps, a.fpos, vR);
    }
    auto at = [&](double* p) -> double* {
        if constexpr (kFused) return p + i;
        else return valid ? p + i : a.trash + (size_t)q * kBodyBlock + threadIdx.x;
    };
    *at(ka_w(a, R, q)) = kRa;
    if constexpr (R == 0) {
        if (a.fuse_comb1) {   // k_comb1's stage-1 combination, same operands
            const double fac = rk4_a(1) * h;
            *at(a.npos[q]) = yy + fac * vv;
            *at(kv_w(a, 1, q)) = vv + fac * kRa;
        }
    } else if constexpr (R == 1 || R == 2) {
        const double fac = rk4_a(R + 1) * h;
        *at(a.npos[q]) = yy + fac * kk;
        *at(kv_w(a, R + 1, q)) = vv + fac * kRa;
    } else if constexpr (R == 3) {
        // y + b0 (f1 + f4) + b1 (f2 + f3), b pre-multiplied by h (int_rungekutta4.cu:50-60, 150-171)
        const double b0 = (1.0 / 6.0) * h, b1 = (1.0 / 3.0) * h;
        const double yo = yy + b0 * (k0v + kk) + b1 * (k1v + k2v);
        const double vo = vv + b0 * (k0a + kRa) + b1 * (k1a + k2a);
        if constexpr (kFused) {
            // k_commit's work for a committed attempt: y_n <- y_{n+1} in place (this
            // lane alone reads body i's y_n, and no pair term of the launch does)
            // and t += h (no stage of a gas-free fused step reads t)
            if (a.commit >= 0 && (c.flags & F_COMMIT)) {
                const_cast<double*>(a.pos[q])[i] = yo;
                const_cast<double*>(a.vel[q])[i] = vo;
                if (a.commit == 1 && q == 0 && i == di.off) const_cast<DiskCtl*>(a.ctl)[0].t = c.t + c.h;
                return;
            }
        }
        *at(a.out[q]) = yo;
        *at(a.out[3 + q]) = vo;
        if (a.adaptive) {
            *at(a.npos[q]) = yo;
            *at(kv_w(a, 4, q)) = vo;
        }
    } else if constexpr (R == 4) {
        const uint32_t li = i - di.off;
        double e = 0.0;
        if (valid && (a.err_all || li < di.err_hi)) e = fmax(fabs(k3v - kk), fabs(k3a - kRa));   // |f4 - f5|
        error_store_wave(a, valid, e);
    }
}

template <int R>
__device__ __forceinline__ void stage_rk4_one(const StageArgs& a)
{
    const uint32_t i0 = blockIdx.x * kBodyBlock + threadIdx.x;
    const bool in = i0 < a.n_all;
    stage_rk4_body<R, false>(a, in ? i0 : a.n_all - 1, (int)blockIdx.y, in, 0.0);
}

// Classical RK4 stage kernel (int_rungekutta4.cu:37-73, 130-247).
template <int R, bool ONE>
__global__ __launch_bounds__(kBodyBlock) void k_stage_rk4(StageArgs a)
{
    if constexpr (ONE) {
        stage_rk4_one<R>(a);
        return;
    }
    const uint32_t i0 = blockIdx.x * kBodyBlock + threadIdx.x;
    const bool in = i0 < a.n_all;
    const uint32_t i = in ? i0 : a.n_all - 1;
    const int q = blockIdx.y;
    const uint32_t d = a.disk_of[i];
    const double y = a.pos[q][i], v = a.vel[q][i];
    const double kRv = kv_of(a, R, q)[i];
    double k0v = 0, k1v = 0, k2v = 0, k0a = 0, k1a = 0, k2a = 0, k3v = 0, k3a = 0;
    if constexpr (R == 3) {
        k0v = kv_of(a, 0, q)[i]; k1v = kv_of(a, 1, q)[i]; k2v = kv_of(a, 2, q)[i];
        k0a = ka_of(a, 0, q)[i]; k1a = ka_of(a, 1, q)[i]; k2a = ka_of(a, 2, q)[i];
    }
    if constexpr (R == 4) {
        k3v = kv_of(a, 3, q)[i];
        k3a = ka_of(a, 3, q)[i];
    }
    pin(k0v); pin(k1v); pin(k2v); pin(k0a); pin(k1a); pin(k2a); pin(k3v); pin(k3a);
    double yy = y, vv = v, kk = kRv;
    pin(yy); pin(vv); pin(kk);
    const DiskCtl c = a.ctl[d];
    const DiskInfo di = a.info[d];
    const bool valid = in && (c.flags & (R == 0 ? F_NEED_K0 : F_ACTIVE)) != 0;
    double e = 0.0;
    if (valid) {
        const double h = c.h;
        const double t = c.t + rk4_c(R) * h;
        const double* vR[3] = {kv_of(a, R, 0), kv_of(a, R, 1), kv_of(a, R, 2)};
        const double b = prefix_accel(a, i, q, di, a.fpos);
        const double kRa = stage_accel(a, i, q, di, t, b, partial_sum(a, i, q, nchunk_of(di, i)), a.fpos, vR);
        ka_w(a, R, q)[i] = kRa;
        if constexpr (R == 0) {
            if (a.fuse_comb1) {   // k_comb1's stage-1 combination, same operands
                const double fac = rk4_a(1) * h;
                a.npos[q][i] = yy + fac * vv;
                kv_w(a, 1, q)[i] = vv + fac * kRa;
            }
        } else if constexpr (R == 1 || R == 2) {
            const double fac = rk4_a(R + 1) * h;
            a.npos[q][i] = yy + fac * kk;
            kv_w(a, R + 1, q)[i] = vv + fac * kRa;
        } else if constexpr (R == 3) {
            // y + b0 (f1 + f4) + b1 (f2 + f3), b pre-multiplied by h (int_rungekutta4.cu:50-60, 150-171)
            const double b0 = (1.0 / 6.0) * h, b1 = (1.0 / 3.0) * h;
            const double yo = yy + b0 * (k0v + kk) + b1 * (k1v + k2v);
            const double vo = vv + b0 * (k0a + kRa) + b1 * (k1a + k2a);
            a.out[q][i] = yo;
            a.out[3 + q][i] = vo;
            if (a.adaptive) {
                a.npos[q][i] = yo;
                kv_w(a, 4, q)[i] = vo;
            }
        } else if constexpr (R == 4) {
            const uint32_t li = i - di.off;
            if (a.err_all || li < di.err_hi) e = fmax(fabs(k3v - kk), fabs(k3a - kRa));   // |f4 - f5|
        }
    }
    if constexpr (R == 4) error_reduce(a, valid, d, e);
}

// Euler (int_euler.cu:56-105): y_{n+1} = y_n + h f(t, y_n), the position's
// derivative being the velocity.
template <int R>
__global__ __launch_bounds__(kBodyBlock) void k_stage_euler(StageArgs a)
{
    static_assert(R == 0, "Euler has one stage");
    const uint32_t i = blockIdx.x * kBodyBlock + threadIdx.x;
    const int q = blockIdx.y;
    if (i >= a.n_all) return;
    const uint32_t d = a.disk_of[i];
    const DiskCtl c = a.ctl[d];
    if (!(c.flags & F_NEED_K0)) return;
    const DiskInfo di = a.info[d];
    const double y = a.pos[q][i], v = a.vel[q][i];
    const double b = prefix_accel(a, i, q, di, a.fpos);
    const double kRa = stage_accel(a, i, q, di, c.t, b, partial_sum(a, i, q, nchunk_of(di, i)), a.fpos, a.vel);
    ka_w(a, 0, q)[i] = kRa;
    a.out[q][i] = y + c.h * v;
    a.out[3 + q][i] = v + c.h * kRa;
}

// Midpoint RK2 (int_rungekutta2.cu:13-15, 58-131): stage 0 is k0 (k_comb1<2>
// then forms y + (h/2) k0, RK4's first combination); stage 1 evaluates there,
// at t + h/2, and y_{n+1} = y_n + (1 h) k1.
template <int R>
__global__ __launch_bounds__(kBodyBlock) void k_stage_rk2(StageArgs a)
{
    const uint32_t i = blockIdx.x * kBodyBlock + threadIdx.x;
    const int q = blockIdx.y;
    if (i >= a.n_all) return;
    const uint32_t d = a.disk_of[i];
    const DiskCtl c = a.ctl[d];
    if (!(c.flags & (R == 0 ? F_NEED_K0 : F_ACTIVE))) return;
    const DiskInfo di = a.info[d];
    const double h = c.h;
    const double t = c.t + (R == 0 ? 0.0 : 1.0 / 2.0) * h;
    const double* vR[3] = {kv_of(a, R, 0), kv_of(a, R, 1), kv_of(a, R, 2)};
    const double b = prefix_accel(a, i, q, di, a.fpos);
    const double kRa = stage_accel(a, i, q, di, t, b, partial_sum(a, i, q, nchunk_of(di, i)), a.fpos,
                                   R == 0 ? a.vel : vR);
    ka_w(a, R, q)[i] = kRa;
    if constexpr (R == 1) {
        const double fac = 1.0 * h;
        a.out[q][i] = a.pos[q][i] + fac * kv_of(a, 1, q)[i];
        a.out[3 + q][i] = a.vel[q][i] + fac * kRa;
    }
}

// First combination of an attempt (stage 1 from k0), and clears the error
// slot of each active disk.
template <int M>
__global__ __launch_bounds__(kBodyBlock) void k_comb1(StageArgs a)
{
    const uint32_t i = blockIdx.x * kBodyBlock + threadIdx.x;
    const int q = blockIdx.y;
    if (i >= a.n_all) return;
    const uint32_t d = a.disk_of[i];
    const double y = a.pos[q][i], v = a.vel[q][i];
    const double k0v = a.vel[q][i], k0a = ka_of(a, 0, q)[i];
    const DiskCtl c = a.ctl[d];
    if (!(c.flags & F_ACTIVE)) return;
    if (q == 0 && i == a.info[d].off) {
        a.err[d] = 0ull;
        a.errf[d] = 0ull;
    }
    if constexpr (M == 4) {
        double kv[9], ka[9], p, w;
        rk8_comb_q<1, 0>(y, v, c.h, kv, ka, k0v, k0a, p, w);
        a.npos[q][i] = p;
        kv_w(a, 1, q)[i] = w;
    } else {
        const double fac = rk4_a(1) * c.h;
        a.npos[q][i] = y + fac * k0v;
        kv_w(a, 1, q)[i] = v + fac * k0a;
    }
}

// y_n <- y_{n+1} for disks flagged F_COMMIT; optionally t += h on the device
// (fixed-step runs that never return to the host between steps).
__global__ __launch_bounds__(kBodyBlock) void k_commit(StageArgs a, int advance, DiskCtl* ctl)
{
    const uint32_t i = blockIdx.x * kBodyBlock + threadIdx.x;
    const int q = blockIdx.y;
    if (i >= a.n_all) return;
    const uint32_t d = a.disk_of[i];
    if (!(ctl[d].flags & F_COMMIT)) return;
    const_cast<double*>(a.pos[q])[i] = a.out[q][i];
    const_cast<double*>(a.vel[q])[i] = a.out[3 + q][i];
    if (advance && q == 0 && i == a.info[d].off) ctl[d].t = ctl[d].t + ctl[d].h;
}

// The accept decision of an adaptive attempt, on the device, with the host's
// arithmetic: max_err = |(lambda h) raw|, rejected when max_err > tol
// (int_rungekutta8.cu:693-708), refused by the opt-in guard when a term was
// not finite.  Only comparisons and two products, so the outcome is the
// host's bit for bit.  Accepted disks get F_COMMIT, and F_SPEC_K0 when they go
// on (F_MORE): k_commit and the next step's stage-0 pair pass then run while
// the host is still reading the error.  One block: a single disk reduces its
// per-wave slots; a batch takes one thread per disk.
__global__ __launch_bounds__(256) void k_decide(const unsigned long long* __restrict__ errw, uint32_t nw,
                                                const unsigned long long* __restrict__ err, uint32_t n_disk,
                                                DiskCtl* __restrict__ ctl, double lambda, double tol,
                                                int nonfinite_guard)
{
    __shared__ unsigned long long smax[256], sbad[256];
    const uint32_t t = threadIdx.x;
    if (errw) {
        unsigned long long m = 0, bad = 0;
        for (uint32_t w = t; w < nw; w += blockDim.x) {
            const unsigned long long v = errw[w];
            m = max(m, v & ~kErrNonFinite);
            bad |= v & kErrNonFinite;
        }
        smax[t] = m;
        sbad[t] = bad;
        __syncthreads();
        for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1) {
            if (t < s) {
                smax[t] = max(smax[t], smax[t + s]);
                sbad[t] |= sbad[t + s];
            }
            __syncthreads();
        }
    }
    for (uint32_t k = t; k < n_disk; k += blockDim.x) {
        const DiskCtl c = ctl[k];
        uint32_t fl = 0;
        if (c.flags & F_ACTIVE) {
            const unsigned long long raw_bits = errw ? smax[0] : err[k];
            const bool bad = errw ? sbad[0] != 0 : err[n_disk + k] != 0;
            const double max_err = fabs(lambda * c.h * __longlong_as_double((long long)raw_bits));
            const bool refused = nonfinite_guard && (bad || !isfinite(max_err));
            if (!refused && !(max_err > tol)) fl = F_COMMIT | ((c.flags & F_MORE) ? F_SPEC_K0 : 0u);
        }
        ctl[k].flags = fl;
    }
}

hipError_t launch_decide(hipStream_t s, const unsigned long long* errw, uint32_t nw, const unsigned long long* err,
                         uint32_t n_disk, DiskCtl* ctl, double lambda, double tol, int nonfinite_guard)
{
    hipLaunchKernelGGL(k_decide, dim3(1), dim3(256), 0, s, errw, nw, err, n_disk, ctl, lambda, tol, nonfinite_guard);
    return hipGetLastError();
}

// pp_disk::calc_dydx(1, ...) for one disk: acceleration at time t from the
// partials, written to acc[q].
__global__ __launch_bounds__(kBodyBlock) void k_operator(StageArgs a, double t, uint32_t disk, double* ax,
                                                         double* ay, double* az, const double* vx, const double* vy,
                                                         const double* vz)
{
    const DiskInfo di = a.info[disk];
    const uint32_t li = blockIdx.x * kBodyBlock + threadIdx.x;
    const int q = blockIdx.y;
    if (li >= di.n) return;
    const uint32_t i = di.off + li;
    const double* vel[3] = {vx, vy, vz};
    const double b = prefix_accel(a, i, q, di, a.fpos);
    const double s = stage_accel(a, i, q, di, t, b, partial_sum(a, i, q, nchunk_of(di, i)), a.fpos, vel);
    (q == 0 ? ax : (q == 1 ? ay : az))[i] = s;
}

#define STAGE_CASE(M, R)                                                                   \
    case R:                                                                                \
        hipLaunchKernelGGL(k_stage_##M<R>, grid, dim3(kBodyBlock), 0, st, a);              \
        break;

#define STAGE_CASE_ONE(M, R)                                                               \
    case R:                                                                                \
        if (a.one_disk)                                                                    \
            hipLaunchKernelGGL((k_stage_##M<R, true>), grid, dim3(kBodyBlock), 0, st, a);  \
        else                                                                               \
            hipLaunchKernelGGL((k_stage_##M<R, false>), grid, dim3(kBodyBlock), 0, st, a); \
        break;

static dim3 body_grid(uint32_t n) { return dim3((n + kBodyBlock - 1) / kBodyBlock, 3); }

hipError_t launch_stage(hipStream_t st, int method, int r, const StageArgs& a)
{
    if (a.n_all == 0) return hipSuccess;
    const dim3 grid = body_grid(a.n_all);
    if (method == 4) {
        switch (r) {
            STAGE_CASE_ONE(rk8, 0) STAGE_CASE_ONE(rk8, 1) STAGE_CASE_ONE(rk8, 2) STAGE_CASE_ONE(rk8, 3)
            STAGE_CASE_ONE(rk8, 4) STAGE_CASE_ONE(rk8, 5) STAGE_CASE_ONE(rk8, 6) STAGE_CASE_ONE(rk8, 7)
            STAGE_CASE_ONE(rk8, 8) STAGE_CASE_ONE(rk8, 9) STAGE_CASE_ONE(rk8, 10) STAGE_CASE_ONE(rk8, 11)
            STAGE_CASE_ONE(rk8, 12)
        default: return hipErrorInvalidValue;
        }
    } else if (method == 0) {
        switch (r) {
            STAGE_CASE(euler, 0)
        default: return hipErrorInvalidValue;
        }
    } else if (method == 1) {
        switch (r) {
            STAGE_CASE(rk2, 0) STAGE_CASE(rk2, 1)
        default: return hipErrorInvalidValue;
        }
    } else {
        switch (r) {
            STAGE_CASE_ONE(rk4, 0) STAGE_CASE_ONE(rk4, 1) STAGE_CASE_ONE(rk4, 2) STAGE_CASE_ONE(rk4, 3)
            STAGE_CASE_ONE(rk4, 4)
        default: return hipErrorInvalidValue;
        }
    }
    return hipGetLastError();
}

hipError_t launch_comb1(hipStream_t st, int method, const StageArgs& a)
{
    if (a.n_all == 0) return hipSuccess;
    if (method == 4) hipLaunchKernelGGL(k_comb1<4>, body_grid(a.n_all), dim3(kBodyBlock), 0, st, a);
    else hipLaunchKernelGGL(k_comb1<2>, body_grid(a.n_all), dim3(kBodyBlock), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_commit(hipStream_t st, const StageArgs& a, int advance, DiskCtl* ctl)
{
    if (a.n_all == 0) return hipSuccess;
    hipLaunchKernelGGL(k_commit, body_grid(a.n_all), dim3(kBodyBlock), 0, st, a, advance, ctl);
    return hipGetLastError();
}

hipError_t launch_operator(hipStream_t st, const StageArgs& a, double t, uint32_t disk, double* ax, double* ay,
                           double* az, const double* vx, const double* vy, const double* vz)
{
    if (a.n_all == 0) return hipSuccess;
    hipLaunchKernelGGL(k_operator, body_grid(a.n_all), dim3(kBodyBlock), 0, st, a, t, disk, ax, ay, az, vx, vy, vz);
    return hipGetLastError();
}

// calc_dydx on device arrays: the caller's var4 records (x, y, z, w; 32 B)
// split into the engine's component arrays, and the accelerations joined back
// into var4 records with w = 0.  Scalar loads and stores: a caller's record
// array need only be 8-byte aligned.
__global__ __launch_bounds__(kBodyBlock) void k_var4_split(uint32_t n, const double* __restrict__ r,
                                                           const double* __restrict__ v, double* __restrict__ px,
                                                           double* __restrict__ py, double* __restrict__ pz,
                                                           double* __restrict__ vx, double* __restrict__ vy,
                                                           double* __restrict__ vz)
{
    const uint32_t q = blockIdx.x * kBodyBlock + threadIdx.x;
    if (q >= n) return;
    px[q] = r[4 * (size_t)q + 0];
    py[q] = r[4 * (size_t)q + 1];
    pz[q] = r[4 * (size_t)q + 2];
    vx[q] = v[4 * (size_t)q + 0];
    vy[q] = v[4 * (size_t)q + 1];
    vz[q] = v[4 * (size_t)q + 2];
}

__global__ __launch_bounds__(kBodyBlock) void k_var4_join(uint32_t n, const double* __restrict__ ax,
                                                          const double* __restrict__ ay,
                                                          const double* __restrict__ az, double* __restrict__ dy)
{
    const uint32_t q = blockIdx.x * kBodyBlock + threadIdx.x;
    if (q >= n) return;
    dy[4 * (size_t)q + 0] = ax[q];
    dy[4 * (size_t)q + 1] = ay[q];
    dy[4 * (size_t)q + 2] = az[q];
    dy[4 * (size_t)q + 3] = 0.0;
}

hipError_t launch_var4_split(hipStream_t st, uint32_t n, const double* r, const double* v, double* const p[3],
                             double* const vel[3])
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_var4_split, dim3((n + kBodyBlock - 1) / kBodyBlock), dim3(kBodyBlock), 0, st, n, r, v, p[0],
                       p[1], p[2], vel[0], vel[1], vel[2]);
    return hipGetLastError();
}

hipError_t launch_var4_join(hipStream_t st, uint32_t n, const double* const a[3], double* dy)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_var4_join, dim3((n + kBodyBlock - 1) / kBodyBlock), dim3(kBodyBlock), 0, st, n, a[0], a[1],
                       a[2], dy);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Sources: FP64 mass with -1 for inactive bodies, FP32 mass with 0 for
// inactive bodies and for the FP64 prefix.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_set_sources(uint32_t n_all, const double* mass, const int32_t* id,
                                                        const uint32_t* disk_of, const DiskInfo* info, double* m_ex,
                                                        float* m32)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n_all) return;
    const DiskInfo di = info[disk_of[i]];
    const bool act = id[i] > 0;
    m_ex[i] = act ? mass[i] : -1.0;
    m32[i] = (act && (i - di.off) >= di.n_fp64) ? (float)mass[i] : 0.0f;
}

hipError_t launch_set_sources(hipStream_t st, uint32_t n_all, const double* mass, const int32_t* id,
                              const uint32_t* disk_of, const DiskInfo* info, double* m_ex, float* m32)
{
    const uint32_t grid = (n_all + kBlock - 1) / kBlock;
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL(k_set_sources, dim3(grid), dim3(kBlock), 0, st, n_all, mass, id, disk_of, info, m_ex, m32);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Body records of a few bodies (events): gather into / scatter from a packed
// buffer of kBodyRecord doubles per body, so an event moves O(events) bytes
// over PCIe instead of whole disks.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_gather_bodies(uint32_t n, const uint32_t* __restrict__ idx, BodyArrays b,
                                                          double* __restrict__ rec)
{
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n * kBodyRecord) return;
    const uint32_t q = t / kBodyRecord, f = t % kBodyRecord;
    rec[t] = b.a[f][idx[q]];
}

__global__ __launch_bounds__(kBlock) void k_scatter_bodies(uint32_t n, const uint32_t* __restrict__ idx,
                                                           const double* __restrict__ rec,
                                                           const int32_t* __restrict__ ids, BodyArrays b,
                                                           int32_t* __restrict__ id)
{
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n * (kBodyRecord + 1)) return;
    const uint32_t q = t / (kBodyRecord + 1), f = t % (kBodyRecord + 1);
    if (f < kBodyRecord) b.a[f][idx[q]] = rec[q * kBodyRecord + f];
    else id[idx[q]] = ids[q];
}

hipError_t launch_gather_bodies(hipStream_t st, uint32_t n, const uint32_t* idx, const BodyArrays& b, double* rec)
{
    const uint32_t grid = (n * kBodyRecord + kBlock - 1) / kBlock;
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_bodies, dim3(grid), dim3(kBlock), 0, st, n, idx, b, rec);
    return hipGetLastError();
}

hipError_t launch_scatter_bodies(hipStream_t st, uint32_t n, const uint32_t* idx, const double* rec,
                                 const int32_t* ids, const BodyArrays& b, int32_t* id)
{
    const uint32_t grid = (n * (kBodyRecord + 1) + kBlock - 1) / kBlock;
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL(k_scatter_bodies, dim3(grid), dim3(kBlock), 0, st, n, idx, rec, ids, b, id);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Events
// ---------------------------------------------------------------------------
// Collision test d^2 < f^2 (R_i + R_j)^2 for active i < j of one disk
// (pp_disk.cu:321-371; the star takes part as i, as in the reference kernel),
// on a hashed grid.  A colliding pair is closer than f (R_i + R_j) <= 2 f R_max,
// so with cells at least that wide the two bodies lie in the same or in
// neighbouring cells.  The active bodies of the disks flagged F_CHECK are
// hashed by (disk, cell) into per-bucket linked lists; each body then walks the
// lists of its 27 neighbouring cells and tests the bodies j > i of its own disk
// found in those cells (a body is found once: in the list of its own cell).
// Every pair within reach gets the reference's test, in O(N) rather than the
// reference's O(N^2).  Pairs are appended in any order; the host sorts them
// into (i, j) order.

// Largest radius of a checked active body, as the bits of a non-negative
// double (their unsigned order is the numeric order).
__global__ __launch_bounds__(kBlock) void k_coll_rmax(uint32_t n_all, const double* __restrict__ radius,
                                                      const int32_t* __restrict__ id,
                                                      const uint32_t* __restrict__ disk_of,
                                                      const DiskCtl* __restrict__ ctl, unsigned long long* rmax)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    double r = 0.0;
    if (i < n_all && id[i] > 0 && (ctl[disk_of[i]].flags & F_CHECK)) r = fmax(radius[i], 0.0);
    unsigned long long bits = (unsigned long long)__double_as_longlong(r);
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(bits, off, 64);
        bits = o > bits ? o : bits;
    }
    if ((threadIdx.x & 63u) == 0 && bits) atomicMax(rmax, bits);
}

// Cell width 2 f R_max, widened by 1/1024 so that the rounding of x / width
// (below 1e-4 of a cell for |x| / width < 1e12) cannot put two bodies closer
// than the contact distance two cells apart.  0: no pair can collide.
__device__ inline double coll_cell_inv(const unsigned long long* rmax, double f)
{
    const double w = 2.0 * f * __longlong_as_double((long long)*rmax) * (1.0 + 1.0 / 1024.0);
    return w > 0.0 ? 1.0 / w : 0.0;
}

__device__ inline uint32_t coll_hash(uint32_t disk, long long cx, long long cy, long long cz, uint32_t mask)
{
    uint64_t h = (uint64_t)cx * 0x9E3779B97F4A7C15ull ^ (uint64_t)cy * 0xC2B2AE3D27D4EB4Full ^
                 (uint64_t)cz * 0x165667B19E3779F9ull ^ (uint64_t)disk * 0xD6E8FEB86659FD93ull;
    h ^= h >> 31;
    h *= 0xBF58476D1CE4E5B9ull;
    h ^= h >> 29;
    return (uint32_t)h & mask;
}

__global__ __launch_bounds__(kBlock) void k_coll_insert(uint32_t n_all, const double* __restrict__ x,
                                                        const double* __restrict__ y, const double* __restrict__ z,
                                                        const int32_t* __restrict__ id,
                                                        const uint32_t* __restrict__ disk_of,
                                                        const DiskCtl* __restrict__ ctl,
                                                        const unsigned long long* rmax, double f, uint32_t mask,
                                                        uint32_t* head, uint32_t* next, long long* cell)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n_all) return;
    const uint32_t d = disk_of[i];
    if (!(id[i] > 0 && (ctl[d].flags & F_CHECK))) return;
    const double inv = coll_cell_inv(rmax, f);
    if (inv == 0.0) return;
    const long long cx = (long long)floor(x[i] * inv), cy = (long long)floor(y[i] * inv),
                    cz = (long long)floor(z[i] * inv);
    cell[3 * (size_t)i] = cx;
    cell[3 * (size_t)i + 1] = cy;
    cell[3 * (size_t)i + 2] = cz;
    next[i] = atomicExch(&head[coll_hash(d, cx, cy, cz, mask)], i);
}

__global__ __launch_bounds__(kBlock) void k_coll_query(uint32_t n_all, const double* __restrict__ x,
                                                       const double* __restrict__ y, const double* __restrict__ z,
                                                       const double* __restrict__ radius,
                                                       const int32_t* __restrict__ id,
                                                       const uint32_t* __restrict__ disk_of,
                                                       const DiskCtl* __restrict__ ctl,
                                                       const unsigned long long* rmax, double f, double f2,
                                                       uint32_t mask, const uint32_t* __restrict__ head,
                                                       const uint32_t* __restrict__ next,
                                                       const long long* __restrict__ cell, uint32_t* pairs,
                                                       uint32_t* counter, uint32_t cap)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n_all) return;
    const uint32_t d = disk_of[i];
    if (!(id[i] > 0 && (ctl[d].flags & F_CHECK))) return;
    if (coll_cell_inv(rmax, f) == 0.0) return;
    const double xi = x[i], yi = y[i], zi = z[i], ri = radius[i];
    const long long cx = cell[3 * (size_t)i], cy = cell[3 * (size_t)i + 1], cz = cell[3 * (size_t)i + 2];
    for (int oz = -1; oz <= 1; ++oz)
        for (int oy = -1; oy <= 1; ++oy)
            for (int ox = -1; ox <= 1; ++ox) {
                const long long nx = cx + ox, ny = cy + oy, nz = cz + oz;
                for (uint32_t j = head[coll_hash(d, nx, ny, nz, mask)]; j != ~0u; j = next[j]) {
                    if (j <= i || disk_of[j] != d) continue;
                    if (cell[3 * (size_t)j] != nx || cell[3 * (size_t)j + 1] != ny || cell[3 * (size_t)j + 2] != nz)
                        continue;
                    const double dx = x[j] - xi;
                    const double dy = y[j] - yi;
                    const double dz = z[j] - zi;
                    const double r2 = SQR(dx) + SQR(dy) + SQR(dz);
                    if (r2 < f2 * SQR((ri + radius[j]))) {
                        const uint32_t slot = atomicAdd(counter, 1u);
                        if (slot < cap) {
                            pairs[2 * slot] = i;
                            pairs[2 * slot + 1] = j;
                        }
                    }
                }
            }
}

hipError_t launch_collision_grid(hipStream_t st, uint32_t n_all, const double* x, const double* y, const double* z,
                                 const double* radius, const int32_t* id, const uint32_t* disk_of,
                                 const DiskCtl* ctl, double f, uint32_t table, uint32_t* head, uint32_t* next,
                                 long long* cell, unsigned long long* rmax, uint32_t* pairs, uint32_t* counter,
                                 uint32_t cap)
{
    if (n_all == 0) return hipSuccess;
    if (table == 0 || (table & (table - 1))) return hipErrorInvalidValue;
    const uint32_t grid = (n_all + kBlock - 1) / kBlock;
    if (hipError_t e = hipMemsetAsync(rmax, 0, sizeof(unsigned long long), st)) return e;
    if (hipError_t e = hipMemsetAsync(head, 0xff, (size_t)table * sizeof(uint32_t), st)) return e;
    hipLaunchKernelGGL(k_coll_rmax, dim3(grid), dim3(kBlock), 0, st, n_all, radius, id, disk_of, ctl, rmax);
    hipLaunchKernelGGL(k_coll_insert, dim3(grid), dim3(kBlock), 0, st, n_all, x, y, z, id, disk_of, ctl, rmax, f,
                       table - 1, head, next, cell);
    hipLaunchKernelGGL(k_coll_query, dim3(grid), dim3(kBlock), 0, st, n_all, x, y, z, radius, id, disk_of, ctl, rmax,
                       f, f * f, table - 1, head, next, cell, pairs, counter, cap);
    return hipGetLastError();
}

// Ejection / hit-centrum test on r^2 from the origin (pp_disk.cu:284-318) for
// bodies of disks flagged F_CHECK.  out[2k] = body index, out[2k+1] = event name.
__global__ __launch_bounds__(kBlock) void k_ejection(uint32_t n_all, const double* x, const double* y,
                                                     const double* z, const int32_t* id, const int32_t* type,
                                                     const uint32_t* disk_of, const DiskCtl* ctl, double ed2,
                                                     double hcd2, uint32_t* out, uint32_t* counter, uint32_t cap)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n_all) return;
    if (!(ctl[disk_of[i]].flags & F_CHECK)) return;
    if (!(id[i] > 0 && type[i] != 0)) return;
    const double r2 = SQR(x[i]) + SQR(y[i]) + SQR(z[i]);
    if (ed2 < r2) {
        const uint32_t k = atomicAdd(counter, 1u);
        if (k < cap) { out[2 * k] = i; out[2 * k + 1] = 2u; }
    }
    if (hcd2 > r2) {
        const uint32_t k = atomicAdd(counter, 1u);
        if (k < cap) { out[2 * k] = i; out[2 * k + 1] = 1u; }
    }
}

hipError_t launch_ejection(hipStream_t st, uint32_t n_all, const double* x, const double* y, const double* z,
                           const int32_t* id, const int32_t* type, const uint32_t* disk_of, const DiskCtl* ctl,
                           double ed2, double hcd2, uint32_t* out, uint32_t* counter, uint32_t cap)
{
    if (n_all == 0) return hipSuccess;
    const uint32_t grid = (n_all + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_ejection, dim3(grid), dim3(kBlock), 0, st, n_all, x, y, z, id, type, disk_of, ctl, ed2, hcd2,
                       out, counter, cap);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Integrals (tools.cpp:367-410, 545-700) in FP64 with Neumaier summation.
// ---------------------------------------------------------------------------
struct KSum {
    double s, c;
    __device__ void add(double v)
    {
        const double t = s + v;
        if (fabs(s) >= fabs(v)) c += (s - t) + v;
        else c += (v - t) + s;
        s = t;
    }
    __device__ double value() const { return s + c; }
};

constexpr int kIntChunk = 2048;

// Potential partials: one value per (sink, source chunk).
__global__ __launch_bounds__(kBlock) void k_potential(uint32_t lo, uint32_t hi, const double* __restrict__ x,
                                                      const double* __restrict__ y, const double* __restrict__ z,
                                                      const double* __restrict__ mass,
                                                      const int32_t* __restrict__ id, double* __restrict__ out)
{
    __shared__ double4 tile[kTile];
    const uint32_t n = hi - lo;
    const uint32_t li = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t c = blockIdx.y;
    const uint32_t j0 = lo + c * kIntChunk, j1 = min(hi, j0 + kIntChunk);
    const uint32_t i = lo + li;
    const bool valid = li < n && id[i] > 0;
    const double xi = valid ? x[i] : 0.0, yi = valid ? y[i] : 0.0, zi = valid ? z[i] : 0.0;
    const double mi = valid ? mass[i] : 0.0;
    KSum s{0.0, 0.0};
    for (uint32_t t0 = j0; t0 < j1; t0 += kTile) {
        __syncthreads();
        const uint32_t j = t0 + threadIdx.x;
        if (j < j1) tile[threadIdx.x] = make_double4(x[j], y[j], z[j], id[j] > 0 ? mass[j] : -1.0);
        __syncthreads();
        const uint32_t nt = min((uint32_t)kTile, j1 - t0);
        if (valid) {
            for (uint32_t k = 0; k < nt; ++k) {
                const double4 q = tile[k];
                if (q.w < 0.0 || t0 + k == i) continue;
                const double r = sqrt(SQR(q.x - xi) + SQR(q.y - yi) + SQR(q.z - zi));
                s.add(mi * q.w / r);
            }
        }
    }
    if (li < n) out[(size_t)c * n + li] = s.value();
}

// Per-body terms of M, R, V, C and T.
__global__ __launch_bounds__(kBlock) void k_body_terms(uint32_t lo, uint32_t hi, const double* x, const double* y,
                                                       const double* z, const double* vx, const double* vy,
                                                       const double* vz, const double* mass, const int32_t* id,
                                                       double* out)
{
    const uint32_t n = hi - lo;
    const uint32_t li = blockIdx.x * kBlock + threadIdx.x;
    if (li >= n) return;
    const uint32_t i = lo + li;
    const bool act = id[i] > 0;
    const double m = act ? mass[i] : 0.0;
    const double cx = y[i] * vz[i] - z[i] * vy[i];
    const double cy = z[i] * vx[i] - x[i] * vz[i];
    const double cz = x[i] * vy[i] - y[i] * vx[i];
    const double t[11] = {m, m * x[i], m * y[i], m * z[i], m * vx[i], m * vy[i], m * vz[i],
                          cx * m, cy * m, cz * m, m * (SQR(vx[i]) + SQR(vy[i]) + SQR(vz[i]))};
#pragma unroll
    for (int k = 0; k < 11; ++k) out[(size_t)k * n + li] = act ? t[k] : 0.0;
}

// Sums rows of a [rows][len] array, one block per row, fixed order.
__global__ __launch_bounds__(kBlock) void k_row_sum(const double* in, size_t len, double* out)
{
    __shared__ double ss[kBlock], cs[kBlock];
    const double* row = in + (size_t)blockIdx.x * len;
    KSum s{0.0, 0.0};
    for (size_t k = threadIdx.x; k < len; k += kBlock) s.add(row[k]);
    ss[threadIdx.x] = s.s;
    cs[threadIdx.x] = s.c;
    __syncthreads();
    if (threadIdx.x == 0) {
        KSum t{0.0, 0.0};
        for (int k = 0; k < kBlock; ++k) t.add(ss[k]);
        for (int k = 0; k < kBlock; ++k) t.add(cs[k]);
        out[blockIdx.x] = t.value();
    }
}

hipError_t launch_integral(hipStream_t st, uint32_t lo, uint32_t hi, const double* x, const double* y,
                           const double* z, const double* vx, const double* vy, const double* vz,
                           const double* mass, const int32_t* id, double* scratch, size_t scratch_len, double* out,
                           double* out_u)
{
    const uint32_t n = hi - lo;
    if (n == 0) return hipSuccess;
    const uint32_t nchunk = (n + kIntChunk - 1) / kIntChunk;
    if (scratch_len < (size_t)11 * n || scratch_len < (size_t)nchunk * n) return hipErrorInvalidValue;
    const uint32_t gb = (n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_body_terms, dim3(gb), dim3(kBlock), 0, st, lo, hi, x, y, z, vx, vy, vz, mass, id, scratch);
    hipLaunchKernelGGL(k_row_sum, dim3(11), dim3(kBlock), 0, st, scratch, (size_t)n, out);
    hipLaunchKernelGGL(k_potential, dim3(gb, nchunk), dim3(kBlock), 0, st, lo, hi, x, y, z, mass, id, scratch);
    hipLaunchKernelGGL(k_row_sum, dim3(1), dim3(kBlock), 0, st, scratch, (size_t)nchunk * n, out_u);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Step-proposer features
// ---------------------------------------------------------------------------
// min over active pairs i < j of |r_ij|^2 / |v_ij|^2, per disk flagged
// F_CHECK. A feature, not a force: FP32 throughout, positions and velocities
// taken relative to the disk's star (body 0 of the disk) before rounding.
// Layout as the mixed pair kernel: a work item is kFeatS * 64 sinks (kFeatS
// per lane) against a source chunk on or above the diagonal; the four waves
// take quarters of each 256-source LDS tile; j > i is tested per pair.  The
// minimum does not depend on the order, so the result is deterministic.
constexpr uint32_t kFeatS = 2;
__global__ __launch_bounds__(kBlock) void k_feat_pairs(const WorkItem* __restrict__ items, uint32_t n_items,
                                                       const double* __restrict__ x, const double* __restrict__ y,
                                                       const double* __restrict__ z, const double* __restrict__ vx,
                                                       const double* __restrict__ vy, const double* __restrict__ vz,
                                                       const int32_t* __restrict__ id,
                                                       const DiskInfo* __restrict__ info,
                                                       const DiskCtl* __restrict__ ctl, uint32_t* __restrict__ feat)
{
    __shared__ float4 tp[kTile];
    __shared__ float4 tv[kTile];
    __shared__ float red[kBlock / 64];
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    for (uint32_t it = blockIdx.x; it < n_items; it += gridDim.x) {
        const WorkItem w = items[it];
        if (!(ctl[w.disk].flags & F_CHECK)) continue;
        const uint32_t s = info[w.disk].off;
        const double sx = x[s], sy = y[s], sz = z[s], svx = vx[s], svy = vy[s], svz = vz[s];
        // Inactive or padding bodies sit at far sentinels, sinks at -3e19 and
        // sources at +3e19 AU: the squares overflow, r^2 is infinite and never
        // the minimum, so the inner loop needs no activity test.  Only tiles
        // that reach back to the sinks' own indices test j > i.
        uint32_t si[kFeatS];
        float xi[kFeatS], yi[kFeatS], zi[kFeatS], ui[kFeatS], vi[kFeatS], wi[kFeatS];
#pragma unroll
        for (uint32_t u = 0; u < kFeatS; ++u) {
            const uint32_t i = w.sink_lo + lane + 64u * u;
            const bool valid = i < w.sink_hi && id[i] > 0;
            si[u] = valid ? i : 0xffffffffu;     // an invalid sink matches no j > i
            xi[u] = valid ? (float)(x[i] - sx) : -3.0e19f;
            yi[u] = valid ? (float)(y[i] - sy) : -3.0e19f;
            zi[u] = valid ? (float)(z[i] - sz) : -3.0e19f;
            ui[u] = valid ? (float)(vx[i] - svx) : 0.0f;
            vi[u] = valid ? (float)(vy[i] - svy) : 0.0f;
            wi[u] = valid ? (float)(vz[i] - svz) : 0.0f;
        }
        float best = 3.0e38f;
        auto pair = [&](const float4& pj, const float4& qj, uint32_t u) {
            const float dx = pj.x - xi[u], dy = pj.y - yi[u], dz = pj.z - zi[u];
            const float du = qj.x - ui[u], dv = qj.y - vi[u], dw = qj.z - wi[u];
            const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
            const float u2 = __builtin_fmaf(dw, dw, __builtin_fmaf(dv, dv, __builtin_fmaf(du, du, 1.0e-30f)));
            return r2 * __builtin_amdgcn_rcpf(u2);
        };
        for (uint32_t t0 = w.src_lo; t0 < w.src_hi; t0 += kTile) {
            __syncthreads();
            const uint32_t j = t0 + tid;
            if (j < w.src_hi) {
                const bool on = id[j] > 0;
                tp[tid] = on ? make_float4((float)(x[j] - sx), (float)(y[j] - sy), (float)(z[j] - sz), 0.0f)
                             : make_float4(3.0e19f, 3.0e19f, 3.0e19f, 0.0f);
                tv[tid] = make_float4((float)(vx[j] - svx), (float)(vy[j] - svy), (float)(vz[j] - svz), 0.0f);
            }
            __syncthreads();
            const uint32_t nt = min((uint32_t)kTile, w.src_hi - t0);
            const uint32_t k0 = wave * 64u, k1 = min(k0 + 64u, nt);
            if (t0 < w.sink_hi) {
                // the tile may hold j <= i: keep i < j only
#pragma unroll 4
                for (uint32_t k = k0; k < k1; ++k) {
                    const float4 pj = tp[k], qj = tv[k];
                    const uint32_t jj = t0 + k;
#pragma unroll
                    for (uint32_t u = 0; u < kFeatS; ++u) {
                        const float qq = pair(pj, qj, u);
                        best = jj > si[u] ? fminf(best, qq) : best;
                    }
                }
            } else {
#pragma unroll 4
                for (uint32_t k = k0; k < k1; ++k) {
                    const float4 pj = tp[k], qj = tv[k];
#pragma unroll
                    for (uint32_t u = 0; u < kFeatS; ++u) best = fminf(best, pair(pj, qj, u));
                }
            }
        }
        // positive floats order like their bit patterns
        for (int off = 32; off > 0; off >>= 1) best = fminf(best, __shfl_xor(best, off, 64));
        if (lane == 0) red[wave] = best;
        __syncthreads();
        if (tid == 0) {
            const float m = fminf(fminf(red[0], red[1]), fminf(red[2], red[3]));
            if (m < 3.0e38f) atomicMin(&feat[3 * w.disk + 0], __float_as_uint(m));
        }
        __syncthreads();   // red[] is rewritten by the next item
    }
}

// Min periapsis about the disk's star (float bits, atomicMin) and the
// per-body Keplerian energy m (v^2/2 - mu/r), mu = m_star + m, relative to the
// star (body 0 of the disk), written to eterm[i] (0 for the star, inactive
// bodies and disks not flagged F_CHECK).
__global__ __launch_bounds__(kBlock) void k_feat_bodies(uint32_t n_all, const double* x, const double* y,
                                                        const double* z, const double* vx, const double* vy,
                                                        const double* vz, const double* mass, const int32_t* id,
                                                        const uint32_t* disk_of, const DiskInfo* info,
                                                        const DiskCtl* ctl, uint32_t* feat, double* eterm)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const bool in = i < n_all;
    const uint32_t d = in ? disk_of[i] : 0u;
    const uint32_t s = info[d].off;
    const bool valid = in && (ctl[d].flags & F_CHECK) && i != s && id[i] > 0;
    float q = 3.0e38f;
    double et = 0.0;
    if (valid) {
        const double rx = x[i] - x[s], ry = y[i] - y[s], rz = z[i] - z[s];
        const double ux = vx[i] - vx[s], uy = vy[i] - vy[s], uz = vz[i] - vz[s];
        const double mu = mass[s] + mass[i];
        const double r = sqrt(rx * rx + ry * ry + rz * rz);
        const double v2 = ux * ux + uy * uy + uz * uz;
        const double eps = 0.5 * v2 - mu / r;
        const double hx = ry * uz - rz * uy, hy = rz * ux - rx * uz, hz = rx * uy - ry * ux;
        const double h2 = hx * hx + hy * hy + hz * hz;
        const double e = sqrt(fmax(0.0, 1.0 + 2.0 * eps * h2 / (mu * mu)));
        q = (float)(h2 / (mu * (1.0 + e)));
        et = mass[i] * eps;
    }
    if (in) eterm[i] = et;
    // one atomic per wave when the wave lies in one disk, else one per lane
    const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
    if (__all(d == d0)) {
        for (int off = 32; off > 0; off >>= 1) q = fminf(q, __shfl_xor(q, off, 64));
        if ((threadIdx.x & 63u) == 0 && q < 3.0e38f) atomicMin(&feat[3 * d0 + 1], __float_as_uint(q));
    } else if (valid) {
        atomicMin(&feat[3 * d + 1], __float_as_uint(q));
    }
}

// esum[d] = sum of eterm over the disk's bodies, one block per disk, in a
// fixed order (so the feature is identical run to run).
__global__ __launch_bounds__(kBlock) void k_feat_esum(const double* __restrict__ eterm,
                                                      const DiskInfo* __restrict__ info,
                                                      const DiskCtl* __restrict__ ctl, double* __restrict__ esum)
{
    __shared__ double red[kBlock];
    const uint32_t d = blockIdx.x;
    if (!(ctl[d].flags & F_CHECK)) return;
    const uint32_t lo = info[d].off, hi = lo + info[d].n;
    double acc = 0.0;
    uint32_t i = lo + threadIdx.x;
    for (; i + 7 * kBlock < hi; i += 8 * kBlock) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = eterm[i + u * kBlock];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += v[u];
    }
    for (; i < hi; i += kBlock) acc += eterm[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) esum[d] = red[0];
}

hipError_t launch_features(hipStream_t st, const WorkItem* items, uint32_t n_items, uint32_t grid, uint32_t n_all,
                           uint32_t n_disk, const double* x, const double* y, const double* z, const double* vx,
                           const double* vy, const double* vz, const double* mass, const int32_t* id,
                           const uint32_t* disk_of, const DiskInfo* info, const DiskCtl* ctl, uint32_t* feat,
                           double* eterm, double* esum)
{
    if (n_items)
        hipLaunchKernelGGL(k_feat_pairs, dim3(grid), dim3(kBlock), 0, st, items, n_items, x, y, z, vx, vy, vz, id,
                           info, ctl, feat);
    if (n_all) {
        hipLaunchKernelGGL(k_feat_bodies, dim3((n_all + kBlock - 1) / kBlock), dim3(kBlock), 0, st, n_all, x, y, z,
                           vx, vy, vz, mass, id, disk_of, info, ctl, feat, eterm);
        hipLaunchKernelGGL(k_feat_esum, dim3(n_disk), dim3(kBlock), 0, st, eterm, info, ctl, esum);
    }
    return hipGetLastError();
}

// Plain device copy used to calibrate hardware byte counters (FETCH_SIZE,
// WRITE_SIZE) against a known traffic volume.
__global__ __launch_bounds__(kBlock) void k_copy(const double4* __restrict__ src, double4* __restrict__ dst, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) dst[i] = src[i];
}

hipError_t launch_copy(hipStream_t st, const void* src, void* dst, size_t bytes)
{
    const size_t n = bytes / sizeof(double4);
    if (n == 0) return hipSuccess;
    const size_t grid = std::min<size_t>((n + kBlock - 1) / kBlock, 4096);
    hipLaunchKernelGGL(k_copy, dim3((uint32_t)grid), dim3(kBlock), 0, st, (const double4*)src, (double4*)dst, n);
    return hipGetLastError();
}

}  // namespace rdamd
