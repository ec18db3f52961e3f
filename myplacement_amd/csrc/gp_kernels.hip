// gp_kernels.hip — gfx950 kernels of the global-placement evaluate and controller passes.
//
// Reference counterparts (src/placement/cuda/CudaPlacementBackend.cu unless noted):
//   k_wire      buildPinPositions + computeNetBounds + computeNetSums + computePinGradient (:194-337),
//               CPU form AdaptiveGlobalPlacer.cpp:29-82
//   k_dhist / k_dscan / k_dplace / k_ddeposit / k_dbins
//               depositParticlesKernel + buildDensityMetricsKernel + 10 reductions (:339-398,
//               :1022-1038); CPU DensityMap::addRectangle / metrics, Metrics.cpp:74-119
//   k_wire_dbins
//               k_dbins and k_wire of the same evaluation chain in one launch
//   k_finalize_wire
//               k_finalize of a read-back and k_wire evaluated ahead in one launch
//   k_gemm      forwardDct / solveSpectralKernel / inverse2d (:547-567, :679-742); CPU
//               ElectrostaticField.cpp:205-237 (FFTW REDFT10 / REDFT01 / RODFT01 as dense products)
//   k_gather    sampleDensityGradientKernel (:575-591) + the gradient combine of
//               AdaptiveGlobalPlacer.cpp:308-329
//   controller  proposeStep / estimateStep / momentumAlignment / look-ahead (AdaptiveGlobalPlacer.cpp
//               :470-519, :610-617, :776-780)
//
// Every reduction writes one partial row per workgroup; k_finalize, launched once before each
// read-back, reduces them in index order, so the scalars are deterministic for given partials.
// No hot kernel issues a device-scope fence or atomic: on MI355X those reach across the eight
// XCDs' L2 caches and cost more than the kernels' own work.

#include "gp_kernels.hpp"

#include <algorithm>
#include <array>
#include <utility>
#include <climits>
#include <cmath>
#include <stdexcept>
#include <string>

namespace mpgp {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kFinalizeThreads = 256;
constexpr int kGemmGroup = 4;
constexpr double kEps = 1e-9;
constexpr double kPi = 3.141592653589793238462643383279502884;

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Block coordinates and grid extent a kernel body works with: those of its own launch, or, in a
// launch batched over independent jobs, those of the workgroup within its job's entry.
struct Blk {
    int x, y, z, gx, gy, gz;
};
__device__ inline Blk nativeBlk() {
    return Blk{static_cast<int>(blockIdx.x), static_cast<int>(blockIdx.y), static_cast<int>(blockIdx.z),
               static_cast<int>(gridDim.x), static_cast<int>(gridDim.y), static_cast<int>(gridDim.z)};
}

// (evaluation, block) of a launch batched over the y extent.  With two evaluations the even
// workgroups serve the first and the odd ones the second: workgroups go to the eight XCDs round
// robin, so each XCD's L2 then holds the gathered data of one evaluation only.
__device__ inline int2 batchBlock(const Blk& b) {
    if (b.gy == 1) return make_int2(0, b.x);
    const int lin = b.x + b.y * b.gx;
    return make_int2(lin % b.gy, lin / b.gy);
}
// Bijection of [0, n) that gives the workgroups landing on one XCD (pid % X) one contiguous range.
__device__ inline int xcdChunk(int pid, int n, int X) {
    const int per = (n + X - 1) / X;
    const int tall = n % X == 0 ? X : n % X;
    const int xcd = pid % X, local = pid / X;
    return xcd < tall ? xcd * per + local : tall * per + (xcd - tall) * (per - 1) + local;
}
// batchBlock with each XCD's blocks of an evaluation made one contiguous range: for the density
// deposit, whose row-ordered particles then put each XCD's adds into one band of bin rows (its L2
// holds that band).  Measured: the deposit at 1024 x 1024 bins and 1.35M particles 98 -> 79 us;
// the wirelength and gather kernels are slower with it (their blocks differ in length, or their
// gathers span the whole design anyway).
__device__ inline int2 batchBlockChunked(const Blk& b) {
    if (b.gy == 1) return make_int2(0, xcdChunk(b.x, b.gx, 8));
    const int lin = b.x + b.y * b.gx;
    return make_int2(lin % b.gy, xcdChunk(lin / b.gy, b.gx, 8 / b.gy));
}

// Flattened block a workgroup of a launch batched over jobs runs.  The hardware hands workgroups
// to the eight XCDs round robin (workgroup i to XCD i mod 8); this bijection gives each XCD one
// contiguous range of the flattened blocks, i.e. whole jobs, so that a job's positions, pins and
// grids are cached in one XCD's L2 instead of in all eight.
__device__ inline int batchSlot() {
    return xcdChunk(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x), 8);
}

// Entry of a launch batched over jobs: entry e owns the flattened blocks [first[e], first[e + 1]),
// first[0] = 0.  Every thread tests entries against the workgroup's flattened block `slot`; the
// count of entries starting at or before it is the entry's index + 1.
template <int THREADS>
__device__ inline int batchEntry(const int* __restrict__ first, int entries, int slot) {
    int count = 0;
    for (int k0 = 0; k0 < entries; k0 += THREADS) {
        const int k = k0 + static_cast<int>(threadIdx.x);
        count += __syncthreads_count(k < entries && first[k] <= slot);
    }
    return count - 1;
}
// Local block coordinates within an entry of extent gx x gy x gz.
__device__ inline Blk entryBlk(int local, int gx, int gy, int gz) {
    return Blk{local % gx, (local / gx) % gy, local / (gx * gy), gx, gy, gz};
}

__device__ inline double waveSum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ inline double waveMax(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}
__device__ inline double waveMin(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
    return v;
}
__device__ inline float waveSumF(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Sum (or max, bit k of max_mask) of K values over a 256-thread block; result in thread 0.
template <int K, int NW = kWaves>
__device__ void blockReduce(double (&v)[K], unsigned max_mask, double* lds) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = ((max_mask >> k) & 1U) ? waveMax(v[k]) : waveSum(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) lds[wave * K + k] = v[k];
    __syncthreads();
    // thread k < K combines value k over the wavefronts (in wave order), then thread 0 collects
    if (threadIdx.x < K) {
        const int k = threadIdx.x;
        double a = lds[k];
        for (int w = 1; w < NW; ++w) a = ((max_mask >> k) & 1U) ? fmax(a, lds[w * K + k]) : a + lds[w * K + k];
        lds[NW * K + k] = a;
    }
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] = lds[NW * K + k];
    __syncthreads();
}

// Writes this block's K partials (one row per workgroup).  No fence: the read happens in a later
// kernel (k_finalize), after the kernel boundary has made the writes visible to every XCD.
template <int K, int NW = kWaves>
__device__ void publishPartials(double (&v)[K], unsigned max_mask, int block, double* partials, double* lds) {
    blockReduce<K, NW>(v, max_mask, lds);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) partials[static_cast<size_t>(block) * K + k] = v[k];
}

// Inclusive scan of one int per thread over a block of NW wavefronts; returns the thread's
// inclusive prefix, *total receives the block sum.
template <int NW>
__device__ int blockScan(int v, int* lds, int* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(v, o);
        if (lane >= o) v += u;
    }
    if (lane == 63) lds[wave] = v;
    __syncthreads();
    int before = 0, sum = 0;
    for (int w = 0; w < NW; ++w) {
        const int t = lds[w];
        if (w < wave) before += t;
        sum += t;
    }
    __syncthreads();
    *total = sum;
    return v + before;
}

// Slot of pin p's gradient term: the region-major slot the set-up assigned, or pin order.
__device__ inline int pinSlot(const WireArgs& a, int p) { return a.pin_dest ? a.pin_dest[p] : p; }

__device__ inline double2 pinPosition(const double2* pos, const int* pin_part, const double2* pin_off, int p) {
    const int part = pin_part[p];
    const double2 o = pin_off[p];
    if (part < 0) return o;
    const double2 c = pos[part];
    return make_double2(c.x + o.x, c.y + o.y);
}

// ---------------------------------------------------------------------------------------------
// Wirelength.  Nets arrive ordered by degree class (the host permutes them once): nets of degree
// 2..8 one thread each with their pins held in registers; 9..32 a quarter-wavefront (16 lanes,
// two pins per lane) each; larger nets one wavefront each.  Max/min and positions in FP64; exponentials, their sums
// and the log in FP32; per-net HPWL and smooth WL accumulated in FP64.

__device__ inline float ex32(double d, double inv_gamma) { return __expf(static_cast<float>(d * inv_gamma)); }

template <int W>
__device__ inline double groupMax(double v) {
#pragma unroll
    for (int o = W / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, W));
    return v;
}
template <int W>
__device__ inline double groupMin(double v) {
#pragma unroll
    for (int o = W / 2; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, W));
    return v;
}
template <int W>
__device__ inline float groupSum(float v) {
#pragma unroll
    for (int o = W / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, W);
    return v;
}

// One net on a group of W lanes (W = 16 or 64); returns (hpwl, smooth) in lane 0 of the group.
// Lane `sub` owns pins b + sub + W i.  Nets of up to W * PER pins keep their pins in registers
// from one load pass (all loads of a lane in flight together); larger nets stream them three times.
// The gradient pass recomputes the pin exponentials rather than holding them (71 instead of 78
// VGPRs: 7 waves per SIMD instead of 6 for this gather-latency-bound kernel), and stores each pin's
// term at the slot pin_dest gives (issuing those slot loads with the pin loads instead takes 78
// VGPRs and is slower).
template <int W, int PER>
__device__ void wireGroup(const WireArgs& a, const double2* pos, float2* pin_grad, int net, int sub, double inv_gamma,
                          float gamma_f, double& hp, double& wl) {
    const int b = a.net_off[net], e = a.net_off[net + 1];
    const double w = a.net_w[net];
    double mx = -INFINITY, nx = INFINITY, my = -INFINITY, ny = INFINITY;
    float spx = 0.f, snx = 0.f, spy = 0.f, sny = 0.f;
    if (e - b <= W * PER) {
        double qx[PER], qy[PER];
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int p = b + sub + W * i;
            if (p < e) {
                const double2 q = pinPosition(pos, a.pin_part, a.pin_off, p);
                qx[i] = q.x;
                qy[i] = q.y;
                mx = fmax(mx, q.x); nx = fmin(nx, q.x);
                my = fmax(my, q.y); ny = fmin(ny, q.y);
            }
        }
        mx = groupMax<W>(mx); nx = groupMin<W>(nx); my = groupMax<W>(my); ny = groupMin<W>(ny);
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            if (b + sub + W * i < e) {
                spx += ex32(qx[i] - mx, inv_gamma);
                snx += ex32(nx - qx[i], inv_gamma);
                spy += ex32(qy[i] - my, inv_gamma);
                sny += ex32(ny - qy[i], inv_gamma);
            }
        }
        spx = groupSum<W>(spx); snx = groupSum<W>(snx); spy = groupSum<W>(spy); sny = groupSum<W>(sny);
        if (a.want_grad) {
            const float wf = static_cast<float>(w);
            const float ipx = wf / spx, inx = wf / snx, ipy = wf / spy, iny = wf / sny;
#pragma unroll
            for (int i = 0; i < PER; ++i) {
                const int p = b + sub + W * i;
                const int d = p < e ? pinSlot(a, p) : -1;
                if (d >= 0)
                    pin_grad[d] = make_float2(ex32(qx[i] - mx, inv_gamma) * ipx - ex32(nx - qx[i], inv_gamma) * inx,
                                              ex32(qy[i] - my, inv_gamma) * ipy - ex32(ny - qy[i], inv_gamma) * iny);
            }
        }
    } else {
        for (int p = b + sub; p < e; p += W) {
            const double2 q = pinPosition(pos, a.pin_part, a.pin_off, p);
            mx = fmax(mx, q.x); nx = fmin(nx, q.x);
            my = fmax(my, q.y); ny = fmin(ny, q.y);
        }
        mx = groupMax<W>(mx); nx = groupMin<W>(nx); my = groupMax<W>(my); ny = groupMin<W>(ny);
        for (int p = b + sub; p < e; p += W) {
            const double2 q = pinPosition(pos, a.pin_part, a.pin_off, p);
            spx += ex32(q.x - mx, inv_gamma);
            snx += ex32(nx - q.x, inv_gamma);
            spy += ex32(q.y - my, inv_gamma);
            sny += ex32(ny - q.y, inv_gamma);
        }
        spx = groupSum<W>(spx); snx = groupSum<W>(snx); spy = groupSum<W>(spy); sny = groupSum<W>(sny);
        if (a.want_grad) {
            const float wf = static_cast<float>(w);
            const float ipx = wf / spx, inx = wf / snx, ipy = wf / spy, iny = wf / sny;
            for (int p = b + sub; p < e; p += W) {
                const int d = pinSlot(a, p);
                if (d < 0) continue;
                const double2 q = pinPosition(pos, a.pin_part, a.pin_off, p);
                pin_grad[d] = make_float2(ex32(q.x - mx, inv_gamma) * ipx - ex32(nx - q.x, inv_gamma) * inx,
                                          ex32(q.y - my, inv_gamma) * ipy - ex32(ny - q.y, inv_gamma) * iny);
            }
        }
    }
    if (sub == 0) {
        const double lse = static_cast<double>(gamma_f * (__logf(spx) + __logf(snx) + __logf(spy) + __logf(sny)));
        wl = w * ((mx - nx) + (my - ny) + lse);
        hp = w * ((mx - nx) + (my - ny));
    }
}

// Workgroups in the order large nets (one wavefront each, largest first), medium nets (a
// quarter-wavefront each), small nets (a thread each): the longest-running workgroups are
// dispatched first.
constexpr int kLargePins = 8;  // pins per lane held in registers for a large net (512 per net)
// The controller's smoothing law with exponent 1, operation for operation as the host evaluates it
// (clampd = max(lo, min(v, hi)); no contraction), so that both get the same bits.
__device__ inline double aheadSmoothing(const WireArgs& a, double overflow) {
#pragma clang fp contract(off)
    const double r = overflow / a.ahead_denominator;
    const double m = 1.0 < r ? 1.0 : r;
    const double ratio = 0.0 < m ? m : 0.0;
    return (a.ahead_final + a.ahead_span * ratio) * a.ahead_scale;
}

// The optimizer's normalized overflow (DensityMap::metrics, Metrics.cpp:101-119) from a metrics
// reduction's partial rows, formed as k_finalize forms it: thread t adds rows t, t + 256, ... in
// that order, blockReduce combines the threads, thread 0 divides; the same bits as the finalized
// value.  Every thread of the block calls it and gets the value.
constexpr int kMetricsWidth = 9;
__device__ double metricOverflow(const double* __restrict__ p, int rows, double* lds) {
    static_assert(kThreads == kFinalizeThreads, "the overflow's sums follow k_finalize's thread layout");
    double v[2] = {0.0, 0.0};
    for (int b = threadIdx.x; b < rows; b += kThreads) {
        v[0] += p[static_cast<size_t>(b) * kMetricsWidth + 0];  // total overflow area
        v[1] += p[static_cast<size_t>(b) * kMetricsWidth + 4];  // normalization area
    }
    blockReduce<2>(v, 0U, lds);
    if (threadIdx.x == 0) lds[0] = v[0] / fmax(v[1], kEps);
    __syncthreads();
    const double r = lds[0];
    __syncthreads();
    return r;
}

__device__ __forceinline__ void wireBody(const WireArgs& a, int large_blocks, int medium_blocks, const Blk& bk) {
    __shared__ double lds[(kWaves + 1) * 2];
    double wl = 0.0, hp = 0.0;
    double overflow = 0.0;
    if (a.ahead_metrics)  // uniform: every thread of the block takes the reduction
        overflow = metricOverflow(a.ahead_metrics, a.ahead_metric_rows, lds);
    else if (a.ahead_overflow)
        overflow = *a.ahead_overflow;
    const double gamma = a.ahead_overflow ? aheadSmoothing(a, overflow) : a.gamma;
    const double inv_gamma = 1.0 / gamma;
    const float gamma_f = static_cast<float>(gamma);
    const int2 eb = batchBlock(bk);
    const int block = eb.y;
    if (a.ahead_overflow && block == 0 && threadIdx.x == 0) a.partials[eb.x][2 * a.ahead_rows] = gamma;
    const double2* pos = a.pos[eb.x];
    float2* pin_grad = a.pin_grad[eb.x];
    if (block < large_blocks) {
        const int net = a.n_small + a.n_medium + block * kWaves + (threadIdx.x >> 6);
        if (net < a.n_small + a.n_medium + a.n_large)
            wireGroup<64, kLargePins>(a, pos, pin_grad, net, threadIdx.x & 63, inv_gamma, gamma_f, hp, wl);
    } else if (block < large_blocks + medium_blocks) {
        const int lane = threadIdx.x & 63;
        const int net = a.n_small + (block - large_blocks) * (4 * kWaves) + 4 * (threadIdx.x >> 6) + (lane >> 4);
        if (net < a.n_small + a.n_medium) wireGroup<16, 2>(a, pos, pin_grad, net, lane & 15, inv_gamma, gamma_f, hp, wl);
    } else {
        const int net = (block - large_blocks - medium_blocks) * kThreads + threadIdx.x;
        if (net < a.n_small) {
            // degree group of this net (groups are contiguous; empty ones repeat the next start)
            int deg = 2;
#pragma unroll
            for (int d = 3; d <= kSmallDegree; ++d) deg = net >= a.small_net[d] ? d : deg;
            const int size = a.small_net[deg + 1] - a.small_net[deg];
            const int b = a.small_pin[deg] + (net - a.small_net[deg]);  // pin j at b + j * size
            double qx[kSmallDegree], qy[kSmallDegree];
            double mx = -INFINITY, nx = INFINITY, my = -INFINITY, ny = INFINITY;
#pragma unroll
            for (int i = 0; i < kSmallDegree; ++i) {
                if (i < deg) {
                    const double2 q = pinPosition(pos, a.pin_part, a.pin_off, b + i * size);
                    qx[i] = q.x;
                    qy[i] = q.y;
                    mx = fmax(mx, q.x); nx = fmin(nx, q.x);
                    my = fmax(my, q.y); ny = fmin(ny, q.y);
                }
            }
            float spx = 0.f, snx = 0.f, spy = 0.f, sny = 0.f;
#pragma unroll
            for (int i = 0; i < kSmallDegree; ++i) {
                if (i < deg) {
                    spx += ex32(qx[i] - mx, inv_gamma);
                    snx += ex32(nx - qx[i], inv_gamma);
                    spy += ex32(qy[i] - my, inv_gamma);
                    sny += ex32(ny - qy[i], inv_gamma);
                }
            }
            const double w = a.net_w[net];
            const double lse = static_cast<double>(gamma_f * (__logf(spx) + __logf(snx) + __logf(spy) + __logf(sny)));
            wl = w * ((mx - nx) + (my - ny) + lse);
            hp = w * ((mx - nx) + (my - ny));
            if (a.want_grad) {
                const float wf = static_cast<float>(w);
                const float ipx = wf / spx, inx = wf / snx, ipy = wf / spy, iny = wf / sny;
#pragma unroll
                for (int i = 0; i < kSmallDegree; ++i) {
                    const int d = i < deg ? pinSlot(a, b + i * size) : -1;
                    if (d >= 0)
                        pin_grad[d] = make_float2(ex32(qx[i] - mx, inv_gamma) * ipx - ex32(nx - qx[i], inv_gamma) * inx,
                                                  ex32(qy[i] - my, inv_gamma) * ipy - ex32(ny - qy[i], inv_gamma) * iny);
                }
            }
        }
    }
    double v[2] = {hp, wl};
    publishPartials<2>(v, 0U, block, a.partials[eb.x], lds);
}
__global__ __launch_bounds__(kThreads) void k_wire(WireArgs a, int large_blocks, int medium_blocks) {
    wireBody(a, large_blocks, medium_blocks, nativeBlk());
}

// ---------------------------------------------------------------------------------------------
// Density.  The exact rectangle / bin overlaps (FP64 geometry, bin index math of Metrics.cpp:78-85)
// are added as unsigned 2^40-per-bin-area fixed point, so the density does not depend on the order
// of the additions and is bit-reproducible.  The particles are visited in an order sorted by the
// first bin row they cover (a counting sort, rebuilt only when positions may have moved far): the
// particles of one deposit workgroup then cover a few bin rows, which it accumulates in an LDS
// window before adding the window to the global layers.  The order only affects speed, never the
// result.  Particles taller than two bin rows (movable macros) are not sorted; their own
// workgroups add them straight to the global layers.

__device__ inline void clippedSpan(double2 c, double2 sz, const Region& r, double& left, double& right,
                                   double& bottom, double& top, int& c0, int& c1, int& r0, int& r1, bool& empty) {
    left = fmax(c.x - sz.x * 0.5, r.llx);
    right = fmin(c.x + sz.x * 0.5, r.urx);
    bottom = fmax(c.y - sz.y * 0.5, r.lly);
    top = fmin(c.y + sz.y * 0.5, r.ury);
    empty = fmax(0.0, right - left) * fmax(0.0, top - bottom) <= kEps;
    c0 = max(0, min(r.nx - 1, static_cast<int>(floor((left - r.llx) / r.bw))));
    c1 = max(0, min(r.nx - 1, static_cast<int>(ceil((right - r.llx) / r.bw) - 1.0)));
    r0 = max(0, min(r.ny - 1, static_cast<int>(floor((bottom - r.lly) / r.bh))));
    r1 = max(0, min(r.ny - 1, static_cast<int>(ceil((top - r.lly) / r.bh) - 1.0)));
}

__device__ inline unsigned long long toFixed(double area, double fix) {
    return static_cast<unsigned long long>(__double2ll_rn(area * fix));
}

// key: first covered row (clamped into the grid), or ny for the tall particles the bands visit
// directly.  Particles whose clipped area is empty keep their row key: they may become non-empty
// while the order is reused, and the bands test emptiness at the positions they deposit.
__device__ __forceinline__ void dhistBody(const DensityArgs& a, const Blk& bk) {
    extern __shared__ int hist[];
    const int ny = a.r.ny;
    for (int k = threadIdx.x; k < ny; k += kSortThreads) hist[k] = 0;
    __syncthreads();
    const int p = bk.x * kSortThreads + threadIdx.x;
    if (p < a.n) {
        double l, rr, bo, to;
        int c0, c1, r0, r1;
        bool empty;
        clippedSpan(a.pos[0][p], a.size[p], a.r, l, rr, bo, to, c0, c1, r0, r1, empty);
        const int key = a.tall[p] ? ny : r0;
        a.key[p] = key;
        if (key < ny) atomicAdd(&hist[key], 1);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < ny; k += kSortThreads) a.hist[static_cast<size_t>(bk.x) * ny + k] = hist[k];
}
__global__ __launch_bounds__(kSortThreads) void k_dhist(DensityArgs a) { dhistBody(a, nativeBlk()); }

// one workgroup per key (bin row): exclusive prefix of that key's count over the histogram rows
__device__ __forceinline__ void dscanBody(const DensityArgs& a, int rows, const Blk& bk) {
    __shared__ int lds[kWaves];
    const int ny = a.r.ny, k = bk.x;
    int carry = 0;
    for (int h0 = 0; h0 < rows; h0 += kThreads) {
        const int h = h0 + threadIdx.x;
        const size_t i = static_cast<size_t>(h) * ny + k;
        const int c = h < rows ? a.hist[i] : 0;
        int total;
        const int incl = blockScan<kWaves>(c, lds, &total);
        if (h < rows) a.colprefix[i] = carry + incl - c;
        carry += total;
    }
    if (threadIdx.x == 0) a.keytotal[k] = carry;
}
__global__ __launch_bounds__(kThreads) void k_dscan(DensityArgs a, int rows) { dscanBody(a, rows, nativeBlk()); }

// row starts from the key totals (every workgroup, in LDS), then each particle's slot in row order
__device__ __forceinline__ void dplaceBody(const DensityArgs& a, const Blk& bk) {
    extern __shared__ int sh[];
    int* start = sh;             // [ny + 1]
    int* rank = sh + a.r.ny + 1; // [ny]
    __shared__ int lds[kSortThreads / 64];
    const int ny = a.r.ny;
    int carry = 0;
    for (int k0 = 0; k0 < ny; k0 += kSortThreads) {
        const int k = k0 + threadIdx.x;
        const int c = k < ny ? a.keytotal[k] : 0;
        int total;
        const int incl = blockScan<kSortThreads / 64>(c, lds, &total);
        if (k < ny) {
            start[k] = carry + incl - c;
            rank[k] = 0;
        }
        carry += total;
    }
    if (threadIdx.x == 0) start[ny] = carry;
    __syncthreads();
    if (bk.x == 0)
        for (int k = threadIdx.x; k <= ny; k += kSortThreads) a.keystart[k] = start[k];
    const int p = bk.x * kSortThreads + threadIdx.x;
    if (p >= a.n) return;
    const int key = a.key[p];
    if (key >= ny) return;
    const int slot = start[key] + a.colprefix[static_cast<size_t>(bk.x) * ny + key] + atomicAdd(&rank[key], 1);
    a.perm[slot] = static_cast<unsigned>(p) | (static_cast<unsigned>(a.layer[p]) << kPermLayerShift);
    a.sorted_size[slot] = a.size[p];
}
__global__ __launch_bounds__(kSortThreads) void k_dplace(DensityArgs a) { dplaceBody(a, nativeBlk()); }
// Deposit: one workgroup per kDepositThreads consecutive particles of the row order (then the tall
// particles).  The workgroup finds the bin rows its particles cover at the current positions and
// walks them in slabs that fit its LDS window: accumulate the slab in LDS, then add its non-zero
// cells to the global fixed-point layers.  Integer addition keeps the result independent of the
// order in which workgroups and lanes arrive.
template <int T, int WIN>
__device__ __forceinline__ void ddepositBody(const DensityArgs& a, int sorted_blocks, const Blk& bk) {
    __shared__ unsigned long long win[2 * WIN];
    __shared__ int lds[2 * (T / 64)];
    const int2 eb = batchBlockChunked(bk);
    const Region& r = a.r;
    const int nx = r.nx;
    const size_t G = static_cast<size_t>(nx) * r.ny;
    const double2* pos = a.pos[eb.x];
    unsigned long long* grid = a.grid[eb.x];
    int p = -1, layer = 0;
    double2 sz = make_double2(0.0, 0.0);
    if (eb.y < sorted_blocks) {
        const int q = eb.y * T + threadIdx.x;
        if (q < a.keystart[r.ny]) {
            const unsigned e = a.perm[q];
            p = static_cast<int>(e & kPermIdMask);
            layer = static_cast<int>(e >> kPermLayerShift);
            sz = a.sorted_size[q];
        }
    } else {
        const int k = (eb.y - sorted_blocks) * T + threadIdx.x;
        if (k < a.n_tall) {
            p = a.tall_ids[k];
            sz = a.size[p];
            layer = a.layer[p];
        }
    }
    double left = 0.0, right = 0.0, bottom = 0.0, top = 0.0;
    int c0 = 0, c1 = -1, r0 = 0, r1 = -1;
    bool live = false;
    if (p >= 0) {
        bool empty;
        clippedSpan(pos[p], sz, r, left, right, bottom, top, c0, c1, r0, r1, empty);
        live = !empty;
    }
    // rows covered by the workgroup
    int lo = live ? r0 : INT_MAX, hi = live ? r1 : INT_MIN;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        lo = min(lo, __shfl_xor(lo, o));
        hi = max(hi, __shfl_xor(hi, o));
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        lds[wave] = lo;
        lds[T / 64 + wave] = hi;
    }
    __syncthreads();
    lo = INT_MAX;
    hi = INT_MIN;
    for (int w = 0; w < T / 64; ++w) {
        lo = min(lo, lds[w]);
        hi = max(hi, lds[T / 64 + w]);
    }
    if (lo > hi) return;  // nothing to deposit (uniform across the workgroup)
    const double scale = layer == 1 ? a.t : 1.0;
    const int L = layer == 2 ? 1 : 0;
    auto deposit = [&](int ra, int rb, unsigned long long* dst, int dst_row0, size_t layer_stride) {
        for (int row = ra; row <= rb; ++row) {
            const double bin_bottom = r.lly + static_cast<double>(row) * r.bh;
            const double oy = fmax(0.0, fmin(top, bin_bottom + r.bh) - fmax(bottom, bin_bottom));
            for (int col = c0; col <= c1; ++col) {
                const double bin_left = r.llx + static_cast<double>(col) * r.bw;
                const double ox = fmax(0.0, fmin(right, bin_left + r.bw) - fmax(left, bin_left));
                const double area = ox * oy;
                if (area <= kEps) continue;
                atomicAdd(dst + L * layer_stride + static_cast<size_t>(row - dst_row0) * nx + col,
                          toFixed(scale * area, a.fix));
            }
        }
    };
    const int slab_rows = WIN / nx;
    if (slab_rows == 0) {  // grids wider than the window: straight to the global layers
        if (live) deposit(r0, r1, grid, 0, G);
        return;
    }
    // the covered rows in slabs of the window's height: accumulate in LDS, add non-zero cells
    for (int s0 = lo; s0 <= hi; s0 += slab_rows) {
        const int s1 = min(hi, s0 + slab_rows - 1);
        const int cells = (s1 - s0 + 1) * nx;
        for (int i = threadIdx.x; i < 2 * cells; i += T) win[i] = 0ULL;
        __syncthreads();
        if (live && r0 <= s1 && r1 >= s0) deposit(max(r0, s0), min(r1, s1), win, s0, static_cast<size_t>(cells));
        __syncthreads();
        const size_t base = static_cast<size_t>(s0) * nx;
        for (int i = threadIdx.x; i < cells; i += T) {
            const unsigned long long m = win[i], f = win[cells + i];
            if (m) atomicAdd(grid + base + i, m);
            if (f) atomicAdd(grid + G + base + i, f);
        }
        __syncthreads();
    }
}
__global__ __launch_bounds__(kDepositThreads) void k_ddeposit(DensityArgs a, int sorted_blocks) {
    ddepositBody<kDepositThreads, kWindowCells>(a, sorted_blocks, nativeBlk());
}
// the batch driver's shape (its jobs' grids are small; the sums are the same integers)
__global__ __launch_bounds__(kBatchDepositThreads) void k_ddeposit_small(DensityArgs a, int sorted_blocks) {
    ddepositBody<kBatchDepositThreads, kBatchWindowCells>(a, sorted_blocks, nativeBlk());
}

// Bin metrics (Metrics.cpp:101-119 per bin) and the FP32 deviation grid for the DCT from the
// fixed-point layers, which it clears for the next evaluation.  Each thread takes kBinsPerThread
// consecutive bins (loads of all of them in flight before the arithmetic), so one block reduction
// covers kBinsPerThread * kThreads bins.
constexpr int kBinsPerThread = 4;
__device__ __forceinline__ void dbinsBody(const DensityArgs& a, const Blk& bk) {
    __shared__ double lds[(kWaves + 1) * 9];
    const int2 eb = batchBlock(bk);
    const Region& r = a.r;
    const int nx = r.nx;
    const size_t G = static_cast<size_t>(nx) * r.ny;
    unsigned long long* grid = a.grid[eb.x];
    double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const size_t i0 = (static_cast<size_t>(eb.y) * kThreads + threadIdx.x) * kBinsPerThread;
    unsigned long long gm[kBinsPerThread], gf[kBinsPerThread];
    double sc[kBinsPerThread];
    float* rho = a.rho[eb.x];
    // A thread's bins share one row and every array is 16-byte aligned at them when the row width,
    // the transform's leading dimension and the bases allow: then 16-byte loads and stores
    // (uniform test; the per-bin path below otherwise).
    const bool wide = (nx % kBinsPerThread) == 0 && (a.ldrho % 4) == 0 &&
                      ((reinterpret_cast<uintptr_t>(grid) | reinterpret_cast<uintptr_t>(a.static_charge) |
                        reinterpret_cast<uintptr_t>(rho)) & 15) == 0;
    static_assert(kBinsPerThread == 4, "the 16-byte path moves four bins");
    if (wide) {  // uniform: every thread takes the same path to the one reduction below
        if (i0 < G) {
            const ulonglong2* m2 = reinterpret_cast<const ulonglong2*>(grid + i0);
            const ulonglong2* f2 = reinterpret_cast<const ulonglong2*>(grid + G + i0);
            const double2* s2 = reinterpret_cast<const double2*>(a.static_charge + i0);
            const ulonglong2 m01 = m2[0], m23 = m2[1], f01 = f2[0], f23 = f2[1];
            const double2 s01 = s2[0], s23 = s2[1];
            gm[0] = m01.x; gm[1] = m01.y; gm[2] = m23.x; gm[3] = m23.y;
            gf[0] = f01.x; gf[1] = f01.y; gf[2] = f23.x; gf[3] = f23.y;
            sc[0] = s01.x; sc[1] = s01.y; sc[2] = s23.x; sc[3] = s23.y;
            ulonglong2* z = reinterpret_cast<ulonglong2*>(grid + i0);
            ulonglong2* zf = reinterpret_cast<ulonglong2*>(grid + G + i0);
            z[0] = make_ulonglong2(0ULL, 0ULL);
            z[1] = make_ulonglong2(0ULL, 0ULL);
            zf[0] = make_ulonglong2(0ULL, 0ULL);
            zf[1] = make_ulonglong2(0ULL, 0ULL);
            const int row = static_cast<int>(i0 / nx), col0 = static_cast<int>(i0 - static_cast<size_t>(row) * nx);
            const double bottom = r.lly + static_cast<double>(row) * r.bh;
            const double height = fmax(0.0, (bottom + r.bh) - bottom);
            float out[kBinsPerThread];
#pragma unroll
            for (int q = 0; q < kBinsPerThread; ++q) {
                const double left = r.llx + static_cast<double>(col0 + q) * r.bw;
                const double area = fmax(0.0, (left + r.bw) - left) * height;
                const double mov = static_cast<double>(gm[q]) * a.inv_fix;
                const double fil = static_cast<double>(gf[q]) * a.inv_fix;
                const double q_wo = mov + sc[q];
                const double q_w = q_wo + a.t * fil;
                const double cap = a.t * area;
                const double inv_area = 1.0 / fmax(area, kEps);
                const double d_w = q_w * inv_area, d_wo = q_wo * inv_area;
                v[0] += fmax(0.0, q_w - cap);
                v[1] += fmax(0.0, q_wo - cap);
                v[2] += q_w;
                v[3] += q_wo;
                v[4] += mov;
                v[5] += d_w;
                v[6] += d_wo;
                v[7] = fmax(v[7], d_w);
                v[8] = fmax(v[8], d_wo);
                out[q] = static_cast<float>(d_w - a.t);
            }
            *reinterpret_cast<float4*>(rho + static_cast<size_t>(row) * a.ldrho + col0) =
                make_float4(out[0], out[1], out[2], out[3]);
        }
    } else {
#pragma unroll
        for (int q = 0; q < kBinsPerThread; ++q) {
            const size_t i = i0 + q;
            gm[q] = i < G ? grid[i] : 0ULL;
            gf[q] = i < G ? grid[G + i] : 0ULL;
            sc[q] = i < G ? a.static_charge[i] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < kBinsPerThread; ++q) {
            const size_t i = i0 + q;
            if (i >= G) break;
            const int row = static_cast<int>(i / nx), col = static_cast<int>(i - static_cast<size_t>(row) * nx);
            const double left = r.llx + static_cast<double>(col) * r.bw;
            const double bottom = r.lly + static_cast<double>(row) * r.bh;
            const double area = fmax(0.0, (left + r.bw) - left) * fmax(0.0, (bottom + r.bh) - bottom);
            const double mov = static_cast<double>(gm[q]) * a.inv_fix;
            const double fil = static_cast<double>(gf[q]) * a.inv_fix;
            grid[i] = 0ULL;
            grid[G + i] = 0ULL;
            const double q_wo = mov + sc[q];
            const double q_w = q_wo + a.t * fil;
            const double cap = a.t * area;
            const double inv_area = 1.0 / fmax(area, kEps);
            const double d_w = q_w * inv_area, d_wo = q_wo * inv_area;
            v[0] += fmax(0.0, q_w - cap);
            v[1] += fmax(0.0, q_wo - cap);
            v[2] += q_w;
            v[3] += q_wo;
            v[4] += mov;
            v[5] += d_w;
            v[6] += d_wo;
            v[7] = fmax(v[7], d_w);
            v[8] = fmax(v[8], d_wo);
            rho[static_cast<size_t>(row) * a.ldrho + col] = static_cast<float>(d_w - a.t);
        }
    }
    publishPartials<9>(v, (1U << 7) | (1U << 8), eb.y, a.partials[eb.x], lds);
}
__global__ __launch_bounds__(kThreads) void k_dbins(DensityArgs a) { dbinsBody(a, nativeBlk()); }

// The bin pass and the wirelength of the same evaluations in one launch: the bin workgroups first,
// then the wirelength ones, each finding its block coordinates in its own launch shape.  The two
// are independent (the bins read the deposit's layers, the wirelength the positions and pins) and
// both write their own partial rows, so the results are those of the two launches; the short,
// latency-bound bin pass (one wavefront per SIMD) runs inside the wirelength's time instead of
// after it, and the launch between them is gone.  The bin body takes fewer VGPRs than the
// wirelength body, so the wirelength keeps its occupancy.
__global__ __launch_bounds__(kThreads) void k_wire_dbins(WireArgs w, int large_blocks, int medium_blocks, int wire_gx,
                                                          int wire_gy, DensityArgs d, int bins_gx, int bins_gy) {
    const int b = static_cast<int>(blockIdx.x), nb = bins_gx * bins_gy;
    if (b < nb) {
        dbinsBody(d, Blk{b % bins_gx, b / bins_gx, 0, bins_gx, bins_gy, 1});
    } else {
        const int v = b - nb;
        wireBody(w, large_blocks, medium_blocks, Blk{v % wire_gx, v / wire_gx, 0, wire_gx, wire_gy, 1});
    }
}

// ---------------------------------------------------------------------------------------------
// C = A * B on v_mfma_f32_16x16x4_f32.  Workgroup tile 32x32, four wavefronts split K in
// 16-wide chunks and meet in LDS.  Lane layout of the 16x16x4 instruction: A[i][k] from lane
// i + 16k, B[k][j] from lane j + 16k, D[4*(lane/16) + v][lane%16] in accumulator element v.
typedef float floatx4 __attribute__((ext_vector_type(4)));

template <int EPI>
__device__ __forceinline__ void gemmBody(const GemmBatch& g, int K, int lda, int ldb, int ldc, const SpectralArgs& sp,
                                         const Blk& bk) {
    __shared__ float red[kWaves][32][33];
    __shared__ double lds[kWaves + 1];
    const int z = bk.z;
    const float* __restrict__ A = g.A[z];
    const float* __restrict__ B = g.B[z];
    float* __restrict__ C = g.C[z];
    const int m0 = bk.y * 32, n0 = bk.x * 32;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, r16 = lane & 15, grp = lane >> 4;
    floatx4 acc00 = {0.f, 0.f, 0.f, 0.f}, acc01 = acc00, acc10 = acc00, acc11 = acc00;
    const float* a_row0 = A + static_cast<size_t>(m0 + r16) * lda + 4 * grp;
    const float* a_row1 = a_row0 + static_cast<size_t>(16) * lda;
    const float* b_col = B + static_cast<size_t>(4 * grp) * ldb + n0 + r16;
    // Register pipelining: a wave owns every fourth 16-wide K chunk and keeps two groups of
    // kGemmGroup chunks in registers, loading one group while the MFMAs of the other run.
    // Chunks past K load as zeros and add nothing.
    constexpr int step = 16 * kWaves;
    float4 ca0[kGemmGroup], ca1[kGemmGroup], na0[kGemmGroup], na1[kGemmGroup];
    float cb0[kGemmGroup][4], cb1[kGemmGroup][4], nb0[kGemmGroup][4], nb1[kGemmGroup][4];
#define MP_GEMM_LOAD(A0, A1, B0, B1, KK)                                                   \
    {                                                                                   \
        const int kk_ = (KK);                                                           \
        if (kk_ < K) {                                                                  \
            A0 = *reinterpret_cast<const float4*>(a_row0 + kk_);                        \
            A1 = *reinterpret_cast<const float4*>(a_row1 + kk_);                        \
            const float* bk_ = b_col + static_cast<size_t>(kk_) * ldb;                  \
            _Pragma("unroll") for (int j = 0; j < 4; ++j) {                             \
                B0[j] = bk_[static_cast<size_t>(j) * ldb];                              \
                B1[j] = bk_[static_cast<size_t>(j) * ldb + 16];                         \
            }                                                                           \
        } else {                                                                        \
            A0 = make_float4(0.f, 0.f, 0.f, 0.f);                                       \
            A1 = A0;                                                                    \
            _Pragma("unroll") for (int j = 0; j < 4; ++j) B0[j] = B1[j] = 0.f;          \
        }                                                                               \
    }
    int k0 = wave * 16;
#pragma unroll
    for (int g = 0; g < kGemmGroup; ++g) MP_GEMM_LOAD(ca0[g], ca1[g], cb0[g], cb1[g], k0 + g * step)
    for (; k0 < K; k0 += kGemmGroup * step) {
#pragma unroll
        for (int g = 0; g < kGemmGroup; ++g)
            MP_GEMM_LOAD(na0[g], na1[g], nb0[g], nb1[g], k0 + (kGemmGroup + g) * step)
#pragma unroll
        for (int g = 0; g < kGemmGroup; ++g) {
            const float av0[4] = {ca0[g].x, ca0[g].y, ca0[g].z, ca0[g].w};
            const float av1[4] = {ca1[g].x, ca1[g].y, ca1[g].z, ca1[g].w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                acc00 = __builtin_amdgcn_mfma_f32_16x16x4f32(av0[j], cb0[g][j], acc00, 0, 0, 0);
                acc01 = __builtin_amdgcn_mfma_f32_16x16x4f32(av0[j], cb1[g][j], acc01, 0, 0, 0);
                acc10 = __builtin_amdgcn_mfma_f32_16x16x4f32(av1[j], cb0[g][j], acc10, 0, 0, 0);
                acc11 = __builtin_amdgcn_mfma_f32_16x16x4f32(av1[j], cb1[g][j], acc11, 0, 0, 0);
            }
        }
#pragma unroll
        for (int g = 0; g < kGemmGroup; ++g) {
            ca0[g] = na0[g];
            ca1[g] = na1[g];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                cb0[g][j] = nb0[g][j];
                cb1[g][j] = nb1[g][j];
            }
        }
    }
#undef MP_GEMM_LOAD
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        red[wave][4 * grp + v][r16] = acc00[v];
        red[wave][4 * grp + v][16 + r16] = acc01[v];
        red[wave][16 + 4 * grp + v][r16] = acc10[v];
        red[wave][16 + 4 * grp + v][16 + r16] = acc11[v];
    }
    __syncthreads();
    double energy = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int e = tid + kThreads * q;
        const int row = e >> 5, col = e & 31;
        const float val = red[0][row][col] + red[1][row][col] + red[2][row][col] + red[3][row][col];
        float out = val;
        if (EPI == kEpiSpectral) {
            const int l = m0 + row, k = n0 + col;
            out = 0.f;
            if (l < sp.ny && k < sp.nx) {
                const double wx = kPi * static_cast<double>(k) / sp.W;
                const double wy = kPi * static_cast<double>(l) / sp.H;
                const double den = wx * wx + wy * wy;
                if (den > kEps) {
                    const double chat = static_cast<double>(val);
                    const double phi = chat * sp.normalization / den;
                    out = static_cast<float>(phi);
                    const double weight = (k == 0 || l == 0) ? 0.5 : 1.0;
                    energy += weight * chat * phi;
                }
            }
        }
        if (EPI == kEpiInterleave) C[2 * (static_cast<size_t>(m0 + row) * ldc + n0 + col)] = out;
        else C[static_cast<size_t>(m0 + row) * ldc + n0 + col] = out;
    }
    if (EPI == kEpiSpectral) {
        double v[1] = {energy};
        publishPartials<1>(v, 0U, bk.y * bk.gx + bk.x, sp.partials[z], lds);
    }
}
template <int EPI>
__global__ __launch_bounds__(kThreads) void k_gemm(GemmBatch g, int K, int lda, int ldb, int ldc, SpectralArgs sp) {
    gemmBody<EPI>(g, K, lda, ldb, ldc, sp, nativeBlk());
}

// ---------------------------------------------------------------------------------------------
// Half-length transforms (modes in gp_kernels.hpp).  A workgroup of eight wavefronts owns a 32 x 32
// tile of the half-size output for both parities p (two accumulator sets); the wavefronts split
// the contraction in chunks, fold or de-interleave the operands as they load them, and meet in
// LDS, where the epilogue combines the parities into the full-length outputs.
//
// Precision: FP32 runs v_mfma_f32_16x16x4_f32 over chunks of 16 (KPL = 4 values per lane);
// split-BF16 runs v_mfma_f32_16x16x32_bf16 over chunks of 32 (KPL = 8), each operand split as
// x = hi + lo with hi = bf16(x), lo = bf16(x - hi), and accumulates lo*hi + hi*lo + hi*hi in FP32
// (the lo*lo term, below 2^-16 of the product, is dropped).  The bf16 instruction takes A[i][k]
// from lane i + 16 (k / 8), element k % 8, and B[k][j] from lane j + 16 (k / 8).
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

template <int N>
__device__ inline void loadFloats(const float* p, float (&v)[N]) {
#pragma unroll
    for (int q = 0; q < N / 4; ++q) {
        const float4 x = *reinterpret_cast<const float4*>(p + 4 * q);
        v[4 * q] = x.x;
        v[4 * q + 1] = x.y;
        v[4 * q + 2] = x.z;
        v[4 * q + 3] = x.w;
    }
}

// Operands of one chunk for this lane: a[p][row half][j] = A-side value at contraction index t + j
// for rows r16 (+16), b[p][col half][j] likewise for columns r16 (+16), j < KPL.
template <int MODE, int KPL, int RH, int CH>
__device__ inline void dctLoad(const float* __restrict__ A, const float* __restrict__ B, int n, int h, int lda,
                               int ldb, int row0, int col0, int t, int r16, float (&a)[2][RH][KPL],
                               float (&b)[2][CH][KPL]) {
    if (MODE == kFoldCols) {
#pragma unroll
        for (int rh = 0; rh < RH; ++rh) {
            const float* ar = A + static_cast<size_t>(row0 + 16 * rh + r16) * lda;
            float x[KPL], y[KPL];
            loadFloats<KPL>(ar + t, x);
            loadFloats<KPL>(ar + n - KPL - t, y);  // mirror of t + j is y[KPL - 1 - j]
#pragma unroll
            for (int j = 0; j < KPL; ++j) {
                a[0][rh][j] = x[j] + y[KPL - 1 - j];
                a[1][rh][j] = x[j] - y[KPL - 1 - j];
            }
        }
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int ch = 0; ch < CH; ++ch)
#pragma unroll
                for (int j = 0; j < KPL; ++j)
                    b[p][ch][j] = B[static_cast<size_t>(p) * h * h + static_cast<size_t>(t + j) * h + col0 + 16 * ch + r16];
    } else if (MODE == kFoldRows) {
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int rh = 0; rh < RH; ++rh)
                loadFloats<KPL>(A + static_cast<size_t>(p) * h * h + static_cast<size_t>(row0 + 16 * rh + r16) * h + t, a[p][rh]);
#pragma unroll
        for (int ch = 0; ch < CH; ++ch)
#pragma unroll
            for (int j = 0; j < KPL; ++j) {
                const int col = col0 + 16 * ch + r16;
                const float u = B[static_cast<size_t>(t + j) * ldb + col];
                const float v = B[static_cast<size_t>(n - 1 - t - j) * ldb + col];
                b[0][ch][j] = u + v;
                b[1][ch][j] = u - v;
            }
    } else if (MODE == kMirrorCols) {
#pragma unroll
        for (int rh = 0; rh < RH; ++rh) {
            float x[2 * KPL];
            loadFloats<2 * KPL>(A + static_cast<size_t>(row0 + 16 * rh + r16) * lda + 2 * t, x);
#pragma unroll
            for (int j = 0; j < KPL; ++j) {
                a[0][rh][j] = x[2 * j];
                a[1][rh][j] = x[2 * j + 1];
            }
        }
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int ch = 0; ch < CH; ++ch)
#pragma unroll
                for (int j = 0; j < KPL; ++j)
                    b[p][ch][j] = B[static_cast<size_t>(p) * h * n + static_cast<size_t>(t + j) * n + col0 + 16 * ch + r16];
    } else {  // kMirrorRows
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int rh = 0; rh < RH; ++rh)
                loadFloats<KPL>(A + static_cast<size_t>(p) * h * h + static_cast<size_t>(row0 + 16 * rh + r16) * h + t, a[p][rh]);
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int ch = 0; ch < CH; ++ch)
#pragma unroll
                for (int j = 0; j < KPL; ++j)
                    b[p][ch][j] = B[static_cast<size_t>(2 * (t + j) + p) * ldb + col0 + 16 * ch + r16];
    }
}

template <int RH, int CH>
__device__ inline void dctMma(const float (&a)[2][RH][4], const float (&b)[2][CH][4], floatx4 (&acc)[2][RH][CH]) {
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int rh = 0; rh < RH; ++rh)
#pragma unroll
                for (int ch = 0; ch < CH; ++ch)
                    acc[p][rh][ch] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[p][rh][j], b[p][ch][j], acc[p][rh][ch], 0, 0, 0);
}

__device__ inline void splitBf16(const float (&x)[8], bf16x8& hi, bf16x8& lo) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const __bf16 h = static_cast<__bf16>(x[j]);
        hi[j] = h;
        lo[j] = static_cast<__bf16>(x[j] - static_cast<float>(h));
    }
}

template <int RH, int CH>
__device__ inline void dctMma(const float (&a)[2][RH][8], const float (&b)[2][CH][8], floatx4 (&acc)[2][RH][CH]) {
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        bf16x8 ah[RH], al[RH], bh[CH], bl[CH];
#pragma unroll
        for (int q = 0; q < RH; ++q) splitBf16(a[p][q], ah[q], al[q]);
#pragma unroll
        for (int q = 0; q < CH; ++q) splitBf16(b[p][q], bh[q], bl[q]);
#pragma unroll
        for (int rh = 0; rh < RH; ++rh)
#pragma unroll
            for (int ch = 0; ch < CH; ++ch) {
                floatx4 c = acc[p][rh][ch];
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[rh], bh[ch], c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rh], bl[ch], c, 0, 0, 0);
                acc[p][rh][ch] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[rh], bh[ch], c, 0, 0, 0);
            }
    }
}

// Epilogue of one output element of the half-size product: (row, col) in half-size coordinates, ev /
// od the two parities' sums.
template <int MODE>
__device__ inline void dctStore(const DctBatch& d, int z, float* __restrict__ C, int n, int ldc, int step,
                                const SpectralArgs& sp, int row, int col, int tile_col0, float ev, float od,
                                double& energy) {
    const int h = n >> 1;
    if (MODE == kFoldCols) {
        *reinterpret_cast<float2*>(C + static_cast<size_t>(row) * ldc + 2 * col) = make_float2(ev, od);
    } else if (MODE == kFoldRows) {
        const int k = col;
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const int l = 2 * row + p;
            float out = 0.f;
            if (l < sp.ny && k < sp.nx) {
                const double wx = kPi * static_cast<double>(k) / sp.W;
                const double wy = kPi * static_cast<double>(l) / sp.H;
                const double den = wx * wx + wy * wy;
                if (den > kEps) {
                    const double chat = static_cast<double>(p ? od : ev);
                    const double phi = chat * sp.normalization / den;
                    out = static_cast<float>(phi);
                    const double weight = (k == 0 || l == 0) ? 0.5 : 1.0;
                    energy += weight * chat * phi;
                }
            }
            C[static_cast<size_t>(l) * ldc + k] = out;
        }
    } else if (MODE == kMirrorCols) {
        const bool sine = tile_col0 >= h;  // a tile lies wholly in the cosine or the sine half
        const int c = col - (sine ? h : 0);
        float* out = C + static_cast<size_t>(row) * ldc + (sine ? n : 0);
        out[c] = ev + od;
        out[n - 1 - c] = sine ? od - ev : ev - od;
    } else {  // kMirrorRows
        C[static_cast<size_t>(step) * (static_cast<size_t>(row) * ldc + col)] = ev + od;
        C[static_cast<size_t>(step) * (static_cast<size_t>(n - 1 - row) * ldc + col)] = d.sigma[z] * (ev - od);
    }
}

// A workgroup of W x S wavefronts holds W tiles of TR x TC half-size outputs side by side, both
// parities.  Tile t's contraction is split over its S wavefronts: wavefront s takes chunks s, s + S,
// s + 2S, ... in order (two in flight), and with S > 1 the partial tiles meet in LDS, summed in
// wavefront order, before the epilogue; with S = 1 the epilogue runs from the accumulators.  The
// per-element summation order depends on S only, never on TR, TC, W or F (F: the rotating operand
// prefetch).
template <int MODE, int KPL, int TR, int TC, int W, int S, int F>
__device__ __forceinline__ void dctBody(const DctBatch& d, int n, int lda, int ldb, int ldc, int step,
                                        const SpectralArgs& sp, const Blk& bk) {
    constexpr int RH = TR / 16, CH = TC / 16;
    constexpr int kWaves = W * S;
    __shared__ float red[S > 1 ? 2 * kWaves * TR * (TC + 1) : 1];
    __shared__ double lds[kWaves + 1];
    const int z = bk.z;
    const float* __restrict__ A = d.A[z];
    const float* __restrict__ B = d.B[z];
    float* __restrict__ C = d.C[z];
    const int h = n >> 1;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, r16 = lane & 15, grp = lane >> 4;
    const int tile = wave / S, part = wave % S;
    const int row0 = bk.y * TR, col0 = (bk.x * W + tile) * TC;
    floatx4 acc[2][RH][CH];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int rh = 0; rh < RH; ++rh)
#pragma unroll
            for (int ch = 0; ch < CH; ++ch) acc[p][rh][ch] = floatx4{0.f, 0.f, 0.f, 0.f};
    constexpr int chunk = 4 * KPL;
    constexpr int stride = chunk * S;  // between this wavefront's chunks
    // Chunks enter the accumulators in order k0, k0 + stride, k0 + 2 stride, ... either way.
    if constexpr (F) {
        // two operand buffers rotate: each one's next load is issued as soon as its products have
        // been issued, so a chunk's loads fly while the previous chunk's products run
        float a0[2][RH][KPL], b0[2][CH][KPL], a1[2][RH][KPL], b1[2][CH][KPL];
        int k0 = part * chunk;
        if (k0 < h) dctLoad<MODE, KPL>(A, B, n, h, lda, ldb, row0, col0, k0 + KPL * grp, r16, a0, b0);
        if (k0 + stride < h)
            dctLoad<MODE, KPL>(A, B, n, h, lda, ldb, row0, col0, k0 + stride + KPL * grp, r16, a1, b1);
        for (; k0 < h; k0 += 2 * stride) {
            dctMma(a0, b0, acc);
            if (k0 + 2 * stride < h)
                dctLoad<MODE, KPL>(A, B, n, h, lda, ldb, row0, col0, k0 + 2 * stride + KPL * grp, r16, a0, b0);
            if (k0 + stride < h) {
                dctMma(a1, b1, acc);
                if (k0 + 3 * stride < h)
                    dctLoad<MODE, KPL>(A, B, n, h, lda, ldb, row0, col0, k0 + 3 * stride + KPL * grp, r16, a1, b1);
            }
        }
    } else {
        // two chunks loaded together, then their products
        for (int k0 = part * chunk; k0 < h; k0 += 2 * stride) {
            float a0[2][RH][KPL], b0[2][CH][KPL], a1[2][RH][KPL], b1[2][CH][KPL];
            const bool second = k0 + stride < h;
            dctLoad<MODE, KPL>(A, B, n, h, lda, ldb, row0, col0, k0 + KPL * grp, r16, a0, b0);
            if (second) dctLoad<MODE, KPL>(A, B, n, h, lda, ldb, row0, col0, k0 + stride + KPL * grp, r16, a1, b1);
            dctMma(a0, b0, acc);
            if (second) dctMma(a1, b1, acc);
        }
    }
    double energy = 0.0;
    if constexpr (S == 1) {
#pragma unroll
        for (int rh = 0; rh < RH; ++rh)
#pragma unroll
            for (int ch = 0; ch < CH; ++ch)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    dctStore<MODE>(d, z, C, n, ldc, step, sp, row0 + 16 * rh + 4 * grp + v, col0 + 16 * ch + r16, col0,
                                   acc[0][rh][ch][v], acc[1][rh][ch][v], energy);
    } else {
        // red[p][wave][row][col], row-major within a tile, padded by one column
        auto at = [&](int p, int w, int row, int col) -> float& {
            return red[((p * kWaves + w) * TR + row) * (TC + 1) + col];
        };
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int rh = 0; rh < RH; ++rh)
#pragma unroll
                for (int ch = 0; ch < CH; ++ch)
#pragma unroll
                    for (int v = 0; v < 4; ++v) at(p, wave, 16 * rh + 4 * grp + v, 16 * ch + r16) = acc[p][rh][ch][v];
        __syncthreads();
        constexpr int per = W * TR * TC / (64 * kWaves);  // elements per thread
#pragma unroll
        for (int q = 0; q < per; ++q) {
            const int e = tid + 64 * kWaves * q;
            const int t = e / (TR * TC), row = (e / TC) % TR, col = e % TC;
            float ev = 0.f, od = 0.f;
#pragma unroll
            for (int w = 0; w < S; ++w) {
                ev += at(0, t * S + w, row, col);
                od += at(1, t * S + w, row, col);
            }
            const int tc0 = (bk.x * W + t) * TC;
            dctStore<MODE>(d, z, C, n, ldc, step, sp, row0 + row, tc0 + col, tc0, ev, od, energy);
        }
    }
    if (MODE == kFoldRows) {
        double v[1] = {energy};
        publishPartials<1, kWaves>(v, 0U, bk.y * bk.gx + bk.x, sp.partials[z], lds);
    }
}
template <int MODE, int KPL, int TR, int TC, int W, int S, int F>
__global__ __launch_bounds__(64 * W * S) void k_dct(DctBatch d, int n, int lda, int ldb, int ldc, int step,
                                                     SpectralArgs sp) {
    dctBody<MODE, KPL, TR, TC, W, S, F>(d, n, lda, ldb, ldc, step, sp, nativeBlk());
}

// ---------------------------------------------------------------------------------------------
// Field gather with the Neumann mirror parity (ElectrostaticField.cpp:289-320).
__device__ inline int reflectIndex(int index, int size, bool odd, bool& negate) {
    while (index < 0 || index >= size) {
        index = index < 0 ? -index - 1 : 2 * size - index - 1;
        if (odd) negate = !negate;
    }
    return index;
}

// Bilinear sample of the interleaved (Ex, Ey) grid at (x, y).  Taps outside the grid reflect into
// it; a reflection across a vertical edge flips Ex, across a horizontal edge flips Ey.
__device__ inline double2 sampleField(const float2* __restrict__ f, int ld, double x, double y, const Region& r) {
    x = fmax(r.llx, fmin(x, r.urx));
    y = fmax(r.lly, fmin(y, r.ury));
    const double gx = (x - r.llx) / r.bw - 0.5;
    const double gy = (y - r.lly) / r.bh - 0.5;
    const int left = static_cast<int>(floor(gx));
    const int bottom = static_cast<int>(floor(gy));
    const double tx = gx - static_cast<double>(left);
    const double ty = gy - static_cast<double>(bottom);
    double vx[4], vy[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        bool flip_x = false, flip_y = false;
        const int cc = reflectIndex(left + (q & 1), r.nx, true, flip_x);
        const int rr = reflectIndex(bottom + (q >> 1), r.ny, true, flip_y);
        const float2 v = f[static_cast<size_t>(rr) * ld + cc];
        vx[q] = flip_x ? -static_cast<double>(v.x) : static_cast<double>(v.x);
        vy[q] = flip_y ? -static_cast<double>(v.y) : static_cast<double>(v.y);
    }
    const double lx = vx[0] * (1.0 - tx) + vx[1] * tx, ux = vx[2] * (1.0 - tx) + vx[3] * tx;
    const double ly = vy[0] * (1.0 - tx) + vy[1] * tx, uy = vy[2] * (1.0 - tx) + vy[3] * tx;
    return make_double2(lx * (1.0 - ty) + ux * ty, ly * (1.0 - ty) + uy * ty);
}

__device__ __forceinline__ void gatherBody(const GatherArgs& a, const Blk& bk) {
    __shared__ double lds[(kWaves + 1) * 7];
    const int2 eb = batchBlock(bk);
    const int e = eb.x;
    const int p = eb.y * kThreads + threadIdx.x;
    // wire L1, density L1, |raw|^2, |pre|^2, V.pre, |pos - px|^2, |pre - pg|^2
    double v[7] = {0, 0, 0, 0, 0, 0, 0};
    if (p < a.n) {
        const double2 c = a.pos[e][p];
        const double q = a.charge[p];
        // the particle's pins first (the longest chain of dependent loads), eight index and gradient
        // loads in flight at a time; the field sample after them
        double2 wg = make_double2(0.0, 0.0);
        const int jb = a.ppin_off[p], je = a.ppin_off[p + 1];
        const float2* pg = a.pin_grad[e];
        for (int j0 = jb; j0 < je; j0 += 8) {
            int idx[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) idx[u] = j0 + u < je ? a.ppin_idx[j0 + u] : -1;
            float2 g[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) g[u] = idx[u] >= 0 ? pg[idx[u]] : make_float2(0.f, 0.f);
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (idx[u] >= 0) {
                    wg.x += static_cast<double>(g[u].x);
                    wg.y += static_cast<double>(g[u].y);
                }
        }
        const double2 f = sampleField(a.field[e], a.ldf, c.x, c.y, a.r);
        const double2 dg = make_double2(-q * f.x, -q * f.y);
        v[0] = fabs(wg.x) + fabs(wg.y);
        v[1] = fabs(dg.x) + fabs(dg.y);
        if (a.wire_out) {
            a.wire_out[p] = wg;
            a.density_out[p] = dg;
        }
        if (a.raw_out[e]) {
            const double2 raw = make_double2(wg.x + dg.x * a.penalty, wg.y + dg.y * a.penalty);
            const double s = a.precond[p] * a.inv_strength;
            const double2 pre = make_double2(raw.x * s, raw.y * s);
            a.raw_out[e][p] = raw;
            a.pre_out[e][p] = pre;
            v[2] = raw.x * raw.x + raw.y * raw.y;
            v[3] = pre.x * pre.x + pre.y * pre.y;
            if (a.prev_pos[e]) {  // momentumAlignment, AdaptiveGlobalPlacer.cpp:610-617 (V = X - X_prev)
                const double2 o = a.prev_pos[e][p];
                v[4] = (c.x - o.x) * pre.x + (c.y - o.y) * pre.y;
            }
            if (a.px) {  // estimateStep sums, :480-486
                const double2 b = a.px[p], g = a.pg[p];
                const double dx = c.x - b.x, dy = c.y - b.y, gx = pre.x - g.x, gy = pre.y - g.y;
                v[5] = dx * dx + dy * dy;
                v[6] = gx * gx + gy * gy;
            }
        }
        if (a.ck_out[e]) a.ck_out[e][p] = c;
    }
    publishPartials<7>(v, 0U, eb.y, a.partials[e], lds);
}
static_assert(kThreads == kGatherParticles, "the pin-term regions follow the gather's workgroups");
__global__ __launch_bounds__(kThreads) void k_gather(GatherArgs a) { gatherBody(a, nativeBlk()); }

// ---------------------------------------------------------------------------------------------
// Controller vector passes.
__device__ inline double2 clampCenter(double2 c, double2 sz, const Region& r) {  // GlobalPlacementSupport.cpp:11-19
    const double lx = r.llx + sz.x * 0.5, ux = r.urx - sz.x * 0.5;
    const double ly = r.lly + sz.y * 0.5, uy = r.ury - sz.y * 0.5;
    c.x = lx <= ux ? fmax(lx, fmin(c.x, ux)) : (r.llx + r.urx) * 0.5;
    c.y = ly <= uy ? fmax(ly, fmin(c.y, uy)) : (r.lly + r.ury) * 0.5;
    return c;
}

// B = clamp(X + beta * (X - X_prev)): the look-ahead of :776-780 with the velocity of :861-863
__device__ __forceinline__ void lookaheadBody(const LookaheadArgs& l, const Blk& bk) {
    const int p = bk.x * kThreads + threadIdx.x;
    if (p >= l.n) return;
    const double2 a = l.x[p], o = l.x_prev[p];
    l.out[p] = clampCenter(make_double2(a.x + (a.x - o.x) * l.beta, a.y + (a.y - o.y) * l.beta), l.c.size[p], l.c.r);
}
__global__ __launch_bounds__(kThreads) void k_lookahead(LookaheadArgs l) { lookaheadBody(l, nativeBlk()); }

__device__ double3 decideStep(const DecideArgs& a);
__device__ __forceinline__ void proposeBody(const ProposeArgs& pa, const Blk& bk) {
    __shared__ double lds[(kWaves + 1) * 2];
    // with decide_here every thread takes the same decision from the same device-memory scalars
    const double3 dec = pa.decide_here ? decideStep(pa.decide) : make_double3(0.0, pa.step, pa.max_disp);
    if (pa.decide_here && bk.x == 0 && threadIdx.x == 0) {  // for the host's check
        pa.decide.mirror[0] = dec.x;
        pa.decide.mirror[1] = dec.y;
        pa.decide.mirror[2] = dec.z;
    }
    const bool look = dec.x != 0.0;
    const double2* b = look ? pa.b_look : pa.b;
    const double2* g = look ? pa.g_look : pa.g;
    const double2* raw = look ? pa.raw_look : pa.raw;
    const double step = dec.y, max_disp = dec.z;
    double2* y = pa.y;
    double2* b_copy = pa.b_copy;
    double2* g_copy = pa.g_copy;
    const ClampArgs& c = pa.c;
    const int p = bk.x * kThreads + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (p < pa.n) {
        const double2 base = b[p], gr = g[p];
        double dx = gr.x * (-step), dy = gr.y * (-step);
        const double dn = sqrt(dx * dx + dy * dy);
        if (dn > max_disp) {
            const double s = max_disp / dn;
            dx *= s;
            dy *= s;
        }
        const double2 out = clampCenter(make_double2(base.x + dx, base.y + dy), c.size[p], c.r);
        y[p] = out;
        if (b_copy) {  // staged copies of the step base, kept if the step is accepted (:886-888)
            b_copy[p] = base;
            g_copy[p] = gr;
        }
        const double ax = base.x - out.x, ay = base.y - out.y;
        const double2 r = raw[p];
        v[0] = r.x * ax + r.y * ay;
        v[1] = sqrt(ax * ax + ay * ay);
    }
    publishPartials<2>(v, 2U, bk.x, pa.partials, lds);
}
__global__ __launch_bounds__(kThreads) void k_propose(ProposeArgs pa) { proposeBody(pa, nativeBlk()); }

// The host law's first decision of an iteration, operation for operation (no contraction,
// correctly rounded square root and division), from the finalized scalars: (1 for the look-ahead
// base, step, displacement cap).
__device__ double3 decideStep(const DecideArgs& a) {
#pragma clang fp contract(off)
    const double cur_objective = a.cur[kSmooth] + a.penalty * a.cur[kEnergy];
    bool restarted = a.restarted != 0, use_look = false;
    const double* base = a.cur;
    if (a.try_momentum) {
        if (a.cur[kAlignment] <= 0.0) {
            restarted = true;
        } else {
            const double look_objective = a.look[kSmooth] + a.penalty * a.look[kEnergy];
            if (look_objective > cur_objective * (1.0 + a.growth_threshold)) {
                restarted = true;
            } else {
                use_look = true;
                base = a.look;
            }
        }
    }
    if (restarted) {
        use_look = false;
        base = a.cur;
    }
    double step = fmax(a.last_step, kEps);
    if (a.has_previous) {
        const double dp = base[kCurvDp], dg = base[kCurvDg];
        if (!(dp <= kEps || dg <= kEps)) {
            const double curvature = __dsqrt_rn(__ddiv_rn(dg, dp));
            step = fmax(step * 0.25, fmin(__ddiv_rn(1.0, fmax(curvature, kEps)), step * 4.0));
        }
    }
    const double ratio = fmax(0.0, fmin(__ddiv_rn(base[kOptDensity + 1], fmax(a.initial_overflow, kEps)), 1.0));
    return make_double3(use_look ? 1.0 : 0.0, step, a.move_cap * (0.35 + 0.65 * ratio));
}

__global__ __launch_bounds__(kThreads) void k_zero(double* p, size_t count) {
    for (size_t i = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x; i < count;
         i += static_cast<size_t>(gridDim.x) * kThreads)
        p[i] = 0.0;
}

// ---------------------------------------------------------------------------------------------
// One workgroup per pending reduction: sums (or maxima) of the partial rows in index order, then
// the reference's post-processing of each scalar (DensityMap::metrics Metrics.cpp:101-119, energy
// clamp .cu:1008, alignment sign AdaptiveGlobalPlacer.cpp:610-617).
// Widths of the partial rows by kind: kRedWire 2, kRedMetrics 9, kRedEnergy 1, kRedGather 7,
// kRedPropose 2 (redWidth on the host).
// Accumulates the partial rows tid, tid + T, tid + 2T, ... (T threads) of a K-wide reduction in
// that order.  Rows are read kFinalizeBatch at a time with every load issued before the first
// add: the partials come from another XCD's kernel, so each load waits on the fabric.
// (B rows in flight; the rows a thread adds, and their order, do not depend on B.)
constexpr int kFinalizeBatch = 4;
template <int K, int B = kFinalizeBatch>
__device__ void accumulateRows(const double* __restrict__ p, int rows, unsigned max_mask, double (&acc)[9]) {
    for (int b0 = threadIdx.x; b0 < rows; b0 += B * kFinalizeThreads) {
        double x[B][K];
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const int b = b0 + u * kFinalizeThreads;
#pragma unroll
            for (int k = 0; k < K; ++k) x[u][k] = b < rows ? p[static_cast<size_t>(b) * K + k] : 0.0;
        }
        // rows past the end add 0 and take max with 0: every accumulator starts at 0
#pragma unroll
        for (int u = 0; u < B; ++u)
#pragma unroll
            for (int k = 0; k < K; ++k) acc[k] = ((max_mask >> k) & 1U) ? fmax(acc[k], x[u][k]) : acc[k] + x[u][k];
    }
}

// LEAN: one row in flight for the wide reductions (metrics, gather), so that the body takes no more
// VGPRs than the wirelength's beside which k_finalize_wire runs it; the same sums.
template <bool LEAN = false>
__device__ __forceinline__ void finalizeBody(const RedJob& j, volatile unsigned* flag, unsigned seq) {
    __shared__ double lds[(kFinalizeThreads / 64 + 1) * 9];
    constexpr int BW = LEAN ? 1 : kFinalizeBatch;
    const unsigned max_mask = j.kind == kRedMetrics ? ((1U << 7) | (1U << 8)) : (j.kind == kRedPropose ? 2U : 0U);
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    switch (j.kind) {
        case kRedWire: accumulateRows<2>(j.partials, j.blocks, max_mask, acc); break;
        case kRedMetrics: accumulateRows<9, BW>(j.partials, j.blocks, max_mask, acc); break;
        case kRedEnergy: accumulateRows<1>(j.partials, j.blocks, max_mask, acc); break;
        case kRedGather: accumulateRows<7, BW>(j.partials, j.blocks, max_mask, acc); break;
        default: accumulateRows<2>(j.partials, j.blocks, max_mask, acc); break;
    }
    blockReduce<9, kFinalizeThreads / 64>(acc, max_mask, lds);
    if (threadIdx.x != 0) return;
    // results to the host-mapped bank and, when given, to its copy in device memory (which later
    // kernels read at device speed)
    double* o = j.out;
    double* od = j.out_dev;
    auto put = [&](int slot, double v) {
        o[slot] = v;
        if (od) od[slot] = v;
    };
    switch (j.kind) {
        case kRedWire:
            put(kHpwl, acc[0]);
            put(kSmooth, acc[1]);
            // evaluated ahead: the smoothing it used (after the rows) against the host's (p0)
            if (j.p1 != 0.0) put(kWireGammaMismatch, j.partials[2 * static_cast<size_t>(j.blocks)] == j.p0 ? 0.0 : 1.0);
            break;
        case kRedMetrics:
            for (int f = 0; f < 2; ++f) {
                const int m = f == 0 ? kOptDensity : kDesignDensity;
                put(m + 0, acc[f]);                        // total_overflow_area
                put(m + 5, acc[4]);                        // normalization_area
                put(m + 1, acc[f] / fmax(acc[4], kEps));   // normalized_overflow
                put(m + 2, acc[7 + f]);                    // maximum_density (starts at 0)
                put(m + 3, acc[5 + f] / j.p0);             // average_density
                put(m + 4, acc[2 + f]);                    // total_charge_area
                put(m + 6, j.p2);                          // placeable_area
                put(m + 7, j.p1);                          // dark_area
            }
            break;
        case kRedEnergy:
            put(kEnergy, fmax(0.0, j.p0 * acc[0]));
            break;
        case kRedGather:
            put(kWireL1, acc[0]);
            put(kDensityL1, acc[1]);
            put(kRawNorm2, acc[2]);
            put(kPreNorm2, acc[3]);
            put(kAlignment, -acc[4]);
            put(kCurvDp, acc[5]);
            put(kCurvDg, acc[6]);
            break;
        default:  // kRedPropose
            put(kPredicted, acc[0]);
            put(kMaxDisp, acc[1]);
            break;
    }
    // completion word for a host that polls the scalar block instead of waiting for the stream:
    // written after the results, which the system-scope fence orders before it
    if (flag) {
        __threadfence_system();
        *flag = seq;
    }
}
__global__ __launch_bounds__(kFinalizeThreads) void k_finalize(RedList l) {
    finalizeBody(l.job[blockIdx.x], l.flag ? l.flag + blockIdx.x : nullptr, l.seq);
}

// The read-back's finalize and the wirelength evaluated ahead in one launch: workgroups [0, l.n)
// finalize the pending reductions and store their completion words, the rest run the wirelength
// gradient (one evaluation) at the smoothing law of the optimizer overflow that the finalize is
// forming in the same launch, which each of them forms itself from the same partial rows
// (w.ahead_metrics, metricOverflow).  The host sees the scalars as before; the finalize's launch
// and its gap before the wirelength are gone.
__global__ __launch_bounds__(kThreads) void k_finalize_wire(RedList l, WireArgs w, int large_blocks,
                                                             int medium_blocks, int wire_gx) {
    const int b = static_cast<int>(blockIdx.x);
    if (b < l.n)
        finalizeBody<true>(l.job[b], l.flag ? l.flag + b : nullptr, l.seq);
    else
        wireBody(w, large_blocks, medium_blocks, Blk{b - l.n, 0, 0, wire_gx, 1, 1});
}

int blocksFor(long long n) { return static_cast<int>((n + kThreads - 1) / kThreads); }

// ---------------------------------------------------------------------------------------------
// The same stages batched over independent jobs: one launch runs the stage for every entry, each
// workgroup finding its entry and its block coordinates within it.
__global__ __launch_bounds__(kThreads) void k_wire_batch(const WireEntry* e, const int* first, int n) {
    const int v = batchSlot();
    const int j = batchEntry<kThreads>(first, n, v);
    const WireEntry& E = e[j];
    wireBody(E.a, E.large_blocks, E.medium_blocks, entryBlk(v - first[j], E.gx, E.gy, 1));
}
__global__ __launch_bounds__(kSortThreads) void k_dhist_batch(const SortEntry* e, const int* first, int n) {
    const int v = batchSlot();
    const int j = batchEntry<kSortThreads>(first, n, v);
    const SortEntry& E = e[j];
    dhistBody(E.a, entryBlk(v - first[j], E.gx, 1, 1));
}
__global__ __launch_bounds__(kThreads) void k_dscan_batch(const SortEntry* e, const int* first, int n) {
    const int v = batchSlot();
    const int j = batchEntry<kThreads>(first, n, v);
    const SortEntry& E = e[j];
    dscanBody(E.a, E.rows, entryBlk(v - first[j], E.gx, 1, 1));
}
__global__ __launch_bounds__(kSortThreads) void k_dplace_batch(const SortEntry* e, const int* first, int n) {
    const int v = batchSlot();
    const int j = batchEntry<kSortThreads>(first, n, v);
    const SortEntry& E = e[j];
    dplaceBody(E.a, entryBlk(v - first[j], E.gx, 1, 1));
}
__global__ __launch_bounds__(kBatchDepositThreads) void k_ddeposit_batch(const DensityEntry* e, const int* first, int n) {
    const int v = batchSlot();
    const int j = batchEntry<kBatchDepositThreads>(first, n, v);
    const DensityEntry& E = e[j];
    ddepositBody<kBatchDepositThreads, kBatchWindowCells>(E.a, E.sorted_blocks, entryBlk(v - first[j], E.gx, E.gy, 1));
}
__global__ __launch_bounds__(kThreads) void k_dbins_batch(const DensityEntry* e, const int* first, int n) {
    const int v = batchSlot();
    const int j = batchEntry<kThreads>(first, n, v);
    const DensityEntry& E = e[j];
    dbinsBody(E.a, entryBlk(v - first[j], E.gx, E.gy, 1));
}
template <int EPI>
__global__ __launch_bounds__(kThreads) void k_gemm_batch(const GemmEntry* e, const int* first, int n) {
    const int v = batchSlot();
    const int j = batchEntry<kThreads>(first, n, v);
    const GemmEntry& E = e[j];
    gemmBody<EPI>(E.g, E.K, E.lda, E.ldb, E.ldc, E.sp, entryBlk(v - first[j], E.gx, E.gy, E.gz));
}
template <int MODE, int KPL, int TR, int TC, int W, int S, int F>
__global__ __launch_bounds__(64 * W * S) void k_dct_batch(const DctEntry* e, const int* first, int n) {
    const int v = batchSlot();
    const int j = batchEntry<64 * W * S>(first, n, v);
    const DctEntry& E = e[j];
    dctBody<MODE, KPL, TR, TC, W, S, F>(E.d, E.n, E.lda, E.ldb, E.ldc, E.step, E.sp,
                                        entryBlk(v - first[j], E.gx, E.gy, E.gz));
}
__global__ __launch_bounds__(kThreads) void k_gather_batch(const GatherEntry* e, const int* first, int n) {
    const int v = batchSlot();
    const int j = batchEntry<kThreads>(first, n, v);
    const GatherEntry& E = e[j];
    gatherBody(E.a, entryBlk(v - first[j], E.gx, E.gy, 1));
}
__global__ __launch_bounds__(kThreads) void k_lookahead_batch(const LookaheadEntry* e, const int* first, int n) {
    const int v = batchSlot();
    const int j = batchEntry<kThreads>(first, n, v);
    const LookaheadEntry& E = e[j];
    lookaheadBody(E.l, entryBlk(v - first[j], E.gx, 1, 1));
}
__global__ __launch_bounds__(kThreads) void k_propose_batch(const ProposeEntry* e, const int* first, int n) {
    const int v = batchSlot();
    const int j = batchEntry<kThreads>(first, n, v);
    const ProposeEntry& E = e[j];
    proposeBody(E.p, entryBlk(v - first[j], E.gx, 1, 1));
}
__global__ __launch_bounds__(kThreads) void k_copy_batch(const CopyEntry* e, const int* first, int n) {
    const int v = batchSlot();
    const int j = batchEntry<kThreads>(first, n, v);
    const CopyEntry& E = e[j];
    const int p = (v - first[j]) * kThreads + static_cast<int>(threadIdx.x);
    if (p < E.n) E.dst[p] = E.src[p];
}
__global__ __launch_bounds__(kThreads) void k_zero_batch(const ZeroEntry* e, const int* first, int n) {
    const int v = batchSlot();
    const int j = batchEntry<kThreads>(first, n, v);
    const ZeroEntry& E = e[j];
    const int p = (v - first[j]) * kThreads + static_cast<int>(threadIdx.x);
    if (p < E.n) E.p[p] = 0.0;
}
__global__ __launch_bounds__(kFinalizeThreads) void k_finalize_batch(const RedJob* jobs, volatile unsigned* flags,
                                                                    unsigned seq) {
    finalizeBody(jobs[blockIdx.x], flags + blockIdx.x, seq);
}
// reductions finalized inside a round, read by a later kernel of the round (no completion words)
__global__ __launch_bounds__(kFinalizeThreads) void k_finalize_entries(const RedJob* e, const int* first, int n) {
    const int v = batchSlot();
    const int j = batchEntry<kFinalizeThreads>(first, n, v);
    finalizeBody(e[j], nullptr, 0U);
}

}  // namespace

int wireBlocks(int n_small, int n_medium, int n_large) {
    return blocksFor(n_small) + (n_medium + 4 * kWaves - 1) / (4 * kWaves) + (n_large + kWaves - 1) / kWaves;
}

void launchWire(const WireArgs& a, int count, hipStream_t s) {
    const int large_blocks = (a.n_large + kWaves - 1) / kWaves;
    const int medium_blocks = (a.n_medium + 4 * kWaves - 1) / (4 * kWaves);
    const int blocks = wireBlocks(a.n_small, a.n_medium, a.n_large);
    if (blocks == 0) return;
    if (count < 1 || count > kMaxEval) throw std::invalid_argument("bad evaluation batch");
    if (a.small_net[2] != 0 || a.small_net[kSmallDegree + 1] != a.n_small)
        throw std::invalid_argument("small-net degree groups do not cover the small nets");
    hipLaunchKernelGGL(k_wire, dim3(blocks, count), dim3(kThreads), 0, s, a, large_blocks, medium_blocks);
    check(hipGetLastError(), "launch wirelength kernel");
}

int densityHistBlocks(int n) { return std::max(1, (n + kSortThreads - 1) / kSortThreads); }

void launchDensitySort(const DensityArgs& a, hipStream_t s) {
    const int hb = densityHistBlocks(a.n);
    if (a.n == 0) return;
    hipLaunchKernelGGL(k_dhist, dim3(hb), dim3(kSortThreads), static_cast<size_t>(a.r.ny) * sizeof(int), s, a);
    check(hipGetLastError(), "launch density histogram kernel");
    hipLaunchKernelGGL(k_dscan, dim3(a.r.ny), dim3(kThreads), 0, s, a, hb);
    check(hipGetLastError(), "launch density scan kernel");
    hipLaunchKernelGGL(k_dplace, dim3(hb), dim3(kSortThreads), (2 * static_cast<size_t>(a.r.ny) + 1) * sizeof(int), s, a);
    check(hipGetLastError(), "launch density placement kernel");
}

int densityBinBlocks(const Region& r) {
    const long long per_block = static_cast<long long>(kThreads) * kBinsPerThread;
    return static_cast<int>((static_cast<long long>(r.nx) * r.ny + per_block - 1) / per_block);
}

void launchDeposit(const DensityArgs& a, int count, hipStream_t s) {
    if (count < 1 || count > kMaxEval) throw std::invalid_argument("bad evaluation batch");
    const int sorted_blocks = (a.n + kDepositThreads - 1) / kDepositThreads;
    const int tall_blocks = (a.n_tall + kDepositThreads - 1) / kDepositThreads;
    if (sorted_blocks + tall_blocks > 0) {
        hipLaunchKernelGGL(k_ddeposit, dim3(sorted_blocks + tall_blocks, count), dim3(kDepositThreads), 0, s, a,
                           sorted_blocks);
        check(hipGetLastError(), "launch density deposit kernel");
    }
}

void launchDensity(const DensityArgs& a, int count, hipStream_t s) {
    launchDeposit(a, count, s);
    hipLaunchKernelGGL(k_dbins, dim3(densityBinBlocks(a.r), count), dim3(kThreads), 0, s, a);
    check(hipGetLastError(), "launch density bin kernel");
}

void launchWireBins(const WireArgs& w, int nw, const DensityArgs& d, int nd, hipStream_t s) {
    if (nw < 1 || nw > kMaxEval || nd < 1 || nd > kMaxEval) throw std::invalid_argument("bad evaluation batch");
    const int wire_gx = wireBlocks(w.n_small, w.n_medium, w.n_large);
    if (wire_gx == 0) throw std::invalid_argument("no nets for the wirelength");
    if (w.small_net[2] != 0 || w.small_net[kSmallDegree + 1] != w.n_small)
        throw std::invalid_argument("small-net degree groups do not cover the small nets");
    const int large_blocks = (w.n_large + kWaves - 1) / kWaves;
    const int medium_blocks = (w.n_medium + 4 * kWaves - 1) / (4 * kWaves);
    const int bins_gx = densityBinBlocks(d.r);
    hipLaunchKernelGGL(k_wire_dbins, dim3(bins_gx * nd + wire_gx * nw), dim3(kThreads), 0, s, w, large_blocks,
                       medium_blocks, wire_gx, nw, d, bins_gx, nd);
    check(hipGetLastError(), "launch density bin and wirelength kernel");
}

void launchGemm(const GemmBatch& g, int batch, int M, int N, int K, int lda, int ldb, int ldc, int epilogue,
                const SpectralArgs& sp, int precision, hipStream_t s) {
    if (M % 32 || N % 32 || K % 16) throw std::invalid_argument("DCT product dimensions are not tile multiples");
    if (precision != 0) throw std::invalid_argument("only the FP32 DCT products are built");
    if (batch < 1 || batch > kMaxGemmBatch) throw std::invalid_argument("bad DCT product batch");
    const dim3 grid(N / 32, M / 32, batch);
    if (epilogue == kEpiSpectral)
        hipLaunchKernelGGL(k_gemm<kEpiSpectral>, grid, dim3(kThreads), 0, s, g, K, lda, ldb, ldc, sp);
    else if (epilogue == kEpiInterleave)
        hipLaunchKernelGGL(k_gemm<kEpiInterleave>, grid, dim3(kThreads), 0, s, g, K, lda, ldb, ldc, sp);
    else
        hipLaunchKernelGGL(k_gemm<kEpiStore>, grid, dim3(kThreads), 0, s, g, K, lda, ldb, ldc, sp);
    check(hipGetLastError(), "launch DCT product kernel");
}

bool foldedDct(int nx, int ny) { return nx > 0 && ny > 0 && nx % 64 == 0 && ny % 64 == 0; }

// ---- tile variants of the half-length products ---------------------------------------------------
// Variant index v = (tile * 4 + mode) * 2 + (split-BF16 ? 1 : 0); every variant is instantiated for
// the native and the batched launch, dispatched through tables built at compile time.

namespace {

constexpr int kDctVariants = kDctTileCount * 8;

template <int V>
struct DctVariant {
    static constexpr int mode = (V / 2) % 4, kpl = (V % 2) ? 8 : 4;
    static constexpr DctTile t = kDctTiles[V / 8];
};
using DctLaunchFn = void (*)(dim3, const DctBatch&, int, int, int, int, int, const SpectralArgs&, hipStream_t);
using DctBatchFn = void (*)(dim3, const DctEntry*, const int*, int, hipStream_t);

template <int V>
void launchDctV(dim3 grid, const DctBatch& d, int n, int lda, int ldb, int ldc, int step, const SpectralArgs& sp,
                hipStream_t s) {
    using X = DctVariant<V>;
    hipLaunchKernelGGL((k_dct<X::mode, X::kpl, X::t.tr, X::t.tc, X::t.w, X::t.s, X::t.f>), grid, dim3(64 * X::t.w * X::t.s), 0,
                       s, d, n, lda, ldb, ldc, step, sp);
}
template <int V>
void launchDctBatchV(dim3 grid, const DctEntry* e, const int* first, int n, hipStream_t s) {
    using X = DctVariant<V>;
    hipLaunchKernelGGL((k_dct_batch<X::mode, X::kpl, X::t.tr, X::t.tc, X::t.w, X::t.s, X::t.f>), grid,
                       dim3(64 * X::t.w * X::t.s), 0, s, e, first, n);
}
template <int... V>
constexpr std::array<DctLaunchFn, sizeof...(V)> dctLaunchTable(std::integer_sequence<int, V...>) {
    return {{&launchDctV<V>...}};
}
template <int... V>
constexpr std::array<DctBatchFn, sizeof...(V)> dctBatchTable(std::integer_sequence<int, V...>) {
    return {{&launchDctBatchV<V>...}};
}
constexpr std::array<DctLaunchFn, kDctVariants> kDctLaunch = dctLaunchTable(std::make_integer_sequence<int, kDctVariants>{});
constexpr std::array<DctBatchFn, kDctVariants> kDctBatchLaunch = dctBatchTable(std::make_integer_sequence<int, kDctVariants>{});

// Half-size output extent of a product: rows x cols of (row, col) indices the tiles cover.
void dctHalfExtent(int mode, int rows, int cols, int& hr, int& hc) {
    switch (mode) {
        case kFoldCols: hr = rows; hc = cols / 2; break;
        case kFoldRows: hr = rows / 2; hc = cols; break;
        case kMirrorCols: hr = rows; hc = cols; break;
        case kMirrorRows: hr = rows / 2; hc = cols; break;
        default: throw std::invalid_argument("unknown DCT product mode");
    }
}

// The tuned table: {mode, precision, rows, cols, tile}, written by myplacement_amd/autotune.py.
struct DctTileChoice {
    int mode, precision, rows, cols, tile;
};
constexpr DctTileChoice kDctTuned[] = {
#include "dct_tiles.inc"
};

// Transform length of a product (the contraction is half of it).
int dctLength(int mode, int rows, int cols) { return (mode == kFoldRows || mode == kMirrorRows) ? rows : cols; }

}  // namespace

bool dctTileValid(int tile, int mode, int rows, int cols) {
    if (tile < 0 || tile >= kDctTileCount || rows <= 0 || cols <= 0 || rows % 64 || cols % 64) return false;
    int hr = 0, hc = 0;
    dctHalfExtent(mode, rows, cols, hr, hc);
    const DctTile& t = kDctTiles[tile];
    return hr % t.tr == 0 && hc % (t.w * t.tc) == 0;
}

int dctTileFor(int mode, int precision, int rows, int cols) {
    for (const DctTileChoice& c : kDctTuned)
        if (c.mode == mode && c.precision == precision && c.rows == rows && c.cols == cols &&
            dctTileValid(c.tile, mode, rows, cols))
            return c.tile;
    return 0;  // the eight-way split of 32 x 32 tiles, valid for every folded grid
}

int dctWorkgroups(int tile, int mode, int rows, int cols) {
    int hr = 0, hc = 0;
    dctHalfExtent(mode, rows, cols, hr, hc);
    const DctTile& t = kDctTiles[tile];
    return (hc / (t.w * t.tc)) * (hr / t.tr);
}


void launchDctTile(int tile, int mode, const DctBatch& d, int batch, int rows, int cols, int lda, int ldb, int ldc,
                   int step, int precision, const SpectralArgs& sp, hipStream_t s) {
    if (rows <= 0 || cols <= 0 || rows % 64 || cols % 64)
        throw std::invalid_argument("half-length DCT products need both grid sides to be multiples of 64");
    if (batch < 1 || batch > kMaxGemmBatch) throw std::invalid_argument("bad DCT product batch");
    if (mode == kMirrorRows && step != 1 && step != 2) throw std::invalid_argument("bad DCT output element stride");
    if (precision != 0 && precision != 1) throw std::invalid_argument("unknown DCT product precision");
    if (!dctTileValid(tile, mode, rows, cols)) throw std::invalid_argument("DCT tile does not divide the product");
    int hr = 0, hc = 0;
    dctHalfExtent(mode, rows, cols, hr, hc);
    const DctTile& t = kDctTiles[tile];
    const dim3 grid(hc / (t.w * t.tc), hr / t.tr, batch);
    kDctLaunch[(tile * 4 + mode) * 2 + precision](grid, d, dctLength(mode, rows, cols), lda, ldb, ldc,
                                                  mode == kMirrorRows ? step : 1, sp, s);
    check(hipGetLastError(), "launch half-length DCT product kernel");
}

void launchDct(int mode, const DctBatch& d, int batch, int rows, int cols, int lda, int ldb, int ldc, int step,
               int precision, const SpectralArgs& sp, hipStream_t s) {
    launchDctTile(dctTileFor(mode, precision, rows, cols), mode, d, batch, rows, cols, lda, ldb, ldc, step, precision,
                  sp, s);
}

int particleBlocks(int n) { return std::max(1, blocksFor(n)); }

void launchGather(const GatherArgs& a, int count, hipStream_t s) {
    if (count < 1 || count > kMaxEval) throw std::invalid_argument("bad evaluation batch");
    hipLaunchKernelGGL(k_gather, dim3(particleBlocks(a.n), count), dim3(kThreads), 0, s, a);
    check(hipGetLastError(), "launch field gather kernel");
}

void launchLookahead(const LookaheadArgs& l, hipStream_t s) {
    if (l.n == 0) return;
    hipLaunchKernelGGL(k_lookahead, dim3(blocksFor(l.n)), dim3(kThreads), 0, s, l);
    check(hipGetLastError(), "launch look-ahead kernel");
}

void launchPropose(const ProposeArgs& p, hipStream_t s) {
    hipLaunchKernelGGL(k_propose, dim3(particleBlocks(p.n)), dim3(kThreads), 0, s, p);
    check(hipGetLastError(), "launch propose kernel");
}

int redWidth(int kind) {
    static const int w[kRedKinds] = {2, 9, 1, 7, 2};
    return w[kind];
}

void launchFinalize(const RedList& l, hipStream_t s) {
    if (l.n == 0) return;
    hipLaunchKernelGGL(k_finalize, dim3(l.n), dim3(kFinalizeThreads), 0, s, l);
    check(hipGetLastError(), "launch reduction finalize kernel");
}

void launchFinalizeWire(const RedList& l, const WireArgs& w, hipStream_t s) {
    const int wire_gx = wireBlocks(w.n_small, w.n_medium, w.n_large);
    if (l.n < 1 || wire_gx == 0 || !w.ahead_metrics || !w.ahead_overflow)
        throw std::invalid_argument("the fused finalize needs pending reductions and a wirelength evaluated ahead");
    if (w.small_net[2] != 0 || w.small_net[kSmallDegree + 1] != w.n_small)
        throw std::invalid_argument("small-net degree groups do not cover the small nets");
    const int large_blocks = (w.n_large + kWaves - 1) / kWaves;
    const int medium_blocks = (w.n_medium + 4 * kWaves - 1) / (4 * kWaves);
    hipLaunchKernelGGL(k_finalize_wire, dim3(l.n + wire_gx), dim3(kThreads), 0, s, l, w, large_blocks, medium_blocks,
                       wire_gx);
    check(hipGetLastError(), "launch reduction finalize and wirelength kernel");
}

void loadDeviceCode(hipStream_t s) {
    // an empty launch (count 0: no thread touches memory) makes the runtime load this library's
    // device code now instead of inside the first job's first launch
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(kThreads), 0, s, static_cast<double*>(nullptr), size_t{0});
    check(hipGetLastError(), "load device code");
    check(hipStreamSynchronize(s), "load device code");
}

void launchZero(double* p, size_t count, hipStream_t s) {
    if (count == 0) return;
    const int blocks = static_cast<int>(std::min<size_t>(1024, (count + kThreads - 1) / kThreads));
    hipLaunchKernelGGL(k_zero, dim3(blocks), dim3(kThreads), 0, s, p, count);
    check(hipGetLastError(), "launch zero kernel");
}

// ---- entries and launches batched over jobs ---------------------------------------------------------
WireEntry wireEntry(const WireArgs& a, int count) {
    WireEntry e{a, (a.n_large + kWaves - 1) / kWaves, (a.n_medium + 4 * kWaves - 1) / (4 * kWaves),
                wireBlocks(a.n_small, a.n_medium, a.n_large), count};
    if (a.small_net[2] != 0 || a.small_net[kSmallDegree + 1] != a.n_small)
        throw std::invalid_argument("small-net degree groups do not cover the small nets");
    return e;
}
SortEntry sortEntry(const DensityArgs& a, int stage) {
    const int hb = densityHistBlocks(a.n);
    return SortEntry{a, hb, stage == 1 ? a.r.ny : hb};
}
DensityEntry depositEntry(const DensityArgs& a, int count) {
    const int sorted_blocks = (a.n + kBatchDepositThreads - 1) / kBatchDepositThreads;
    const int tall_blocks = (a.n_tall + kBatchDepositThreads - 1) / kBatchDepositThreads;
    return DensityEntry{a, sorted_blocks, sorted_blocks + tall_blocks, count};
}
DensityEntry binsEntry(const DensityArgs& a, int count) { return DensityEntry{a, 0, densityBinBlocks(a.r), count}; }
GemmEntry gemmEntry(const GemmBatch& g, int batch, int M, int N, int K, int lda, int ldb, int ldc, int epilogue,
                    const SpectralArgs& sp) {
    if (M % 32 || N % 32 || K % 16) throw std::invalid_argument("DCT product dimensions are not tile multiples");
    if (batch < 1 || batch > kMaxGemmBatch) throw std::invalid_argument("bad DCT product batch");
    return GemmEntry{g, sp, K, lda, ldb, ldc, epilogue, N / 32, M / 32, batch};
}
DctEntry dctEntry(int mode, const DctBatch& d, int batch, int rows, int cols, int lda, int ldb, int ldc, int step,
                  int precision, const SpectralArgs& sp) {
    if (rows <= 0 || cols <= 0 || rows % 64 || cols % 64)
        throw std::invalid_argument("half-length DCT products need both grid sides to be multiples of 64");
    if (batch < 1 || batch > kMaxGemmBatch) throw std::invalid_argument("bad DCT product batch");
    if (precision != 0 && precision != 1) throw std::invalid_argument("unknown DCT product precision");
    if (mode == kMirrorRows && step != 1 && step != 2) throw std::invalid_argument("bad DCT output element stride");
    const int tile = dctTileFor(mode, precision, rows, cols);
    int hr = 0, hc = 0;
    dctHalfExtent(mode, rows, cols, hr, hc);
    const DctTile& t = kDctTiles[tile];
    DctEntry e{d, sp, mode, precision == 0 ? 4 : 8, dctLength(mode, rows, cols), lda, ldb, ldc,
               mode == kMirrorRows ? step : 1, hc / (t.w * t.tc), hr / t.tr, batch, tile};
    return e;
}
int dctVariant(const DctEntry& e) { return (e.tile * 4 + e.mode) * 2 + (e.kpl == 8 ? 1 : 0); }
GatherEntry gatherEntry(const GatherArgs& a, int count) { return GatherEntry{a, particleBlocks(a.n), count}; }
LookaheadEntry lookaheadEntry(const LookaheadArgs& l) { return LookaheadEntry{l, blocksFor(l.n)}; }
ProposeEntry proposeEntry(const ProposeArgs& p) { return ProposeEntry{p, particleBlocks(p.n)}; }
CopyEntry copyEntry(double2* dst, const double2* src, int n) { return CopyEntry{dst, src, n, blocksFor(n)}; }
ZeroEntry zeroEntry(double* p, int n) { return ZeroEntry{p, n, blocksFor(n)}; }

void launchBatched(BatchKernel k, const void* entries, const int* first, int n, int total_blocks, std::size_t dyn_lds,
                   hipStream_t s) {
    if (n <= 0 || total_blocks <= 0) return;
    const dim3 grid(total_blocks);
    switch (k) {
        case kBatchWire:
            hipLaunchKernelGGL(k_wire_batch, grid, dim3(kThreads), 0, s, static_cast<const WireEntry*>(entries), first, n);
            break;
        case kBatchSortHist:
            hipLaunchKernelGGL(k_dhist_batch, grid, dim3(kSortThreads), dyn_lds, s, static_cast<const SortEntry*>(entries),
                               first, n);
            break;
        case kBatchSortScan:
            hipLaunchKernelGGL(k_dscan_batch, grid, dim3(kThreads), 0, s, static_cast<const SortEntry*>(entries), first, n);
            break;
        case kBatchSortPlace:
            hipLaunchKernelGGL(k_dplace_batch, grid, dim3(kSortThreads), dyn_lds, s,
                               static_cast<const SortEntry*>(entries), first, n);
            break;
        case kBatchDeposit:
            hipLaunchKernelGGL(k_ddeposit_batch, grid, dim3(kBatchDepositThreads), 0, s,
                               static_cast<const DensityEntry*>(entries), first, n);
            break;
        case kBatchBins:
            hipLaunchKernelGGL(k_dbins_batch, grid, dim3(kThreads), 0, s, static_cast<const DensityEntry*>(entries), first, n);
            break;
        case kBatchGemmStore:
            hipLaunchKernelGGL(k_gemm_batch<kEpiStore>, grid, dim3(kThreads), 0, s,
                               static_cast<const GemmEntry*>(entries), first, n);
            break;
        case kBatchGemmSpectral:
            hipLaunchKernelGGL(k_gemm_batch<kEpiSpectral>, grid, dim3(kThreads), 0, s,
                               static_cast<const GemmEntry*>(entries), first, n);
            break;
        case kBatchGemmInterleave:
            hipLaunchKernelGGL(k_gemm_batch<kEpiInterleave>, grid, dim3(kThreads), 0, s,
                               static_cast<const GemmEntry*>(entries), first, n);
            break;
        case kBatchGather:
            hipLaunchKernelGGL(k_gather_batch, grid, dim3(kThreads), 0, s, static_cast<const GatherEntry*>(entries), first, n);
            break;
        case kBatchLookahead:
            hipLaunchKernelGGL(k_lookahead_batch, grid, dim3(kThreads), 0, s, static_cast<const LookaheadEntry*>(entries),
                               first, n);
            break;
        case kBatchPropose:
            hipLaunchKernelGGL(k_propose_batch, grid, dim3(kThreads), 0, s, static_cast<const ProposeEntry*>(entries),
                               first, n);
            break;
        case kBatchCopy:
            hipLaunchKernelGGL(k_copy_batch, grid, dim3(kThreads), 0, s, static_cast<const CopyEntry*>(entries), first, n);
            break;
        case kBatchZero:
            hipLaunchKernelGGL(k_zero_batch, grid, dim3(kThreads), 0, s, static_cast<const ZeroEntry*>(entries), first, n);
            break;
        case kBatchFinalize:
            hipLaunchKernelGGL(k_finalize_entries, grid, dim3(kFinalizeThreads), 0, s, static_cast<const RedJob*>(entries),
                               first, n);
            break;
        default: {
            // half-length products: kBatchDct + dctVariant(entry)
            const int v = static_cast<int>(k) - kBatchDct;
            if (v < 0 || v >= kDctVariants) throw std::invalid_argument("unknown batched kernel");
            kDctBatchLaunch[v](grid, static_cast<const DctEntry*>(entries), first, n, s);
        }
    }
    check(hipGetLastError(), "launch batched kernel");
}

void launchEntry(BatchKernel k, const void* entry, std::size_t dyn_lds, hipStream_t s) {
    switch (k) {
        case kBatchWire: {
            const WireEntry& E = *static_cast<const WireEntry*>(entry);
            hipLaunchKernelGGL(k_wire, dim3(E.gx, E.gy), dim3(kThreads), 0, s, E.a, E.large_blocks, E.medium_blocks);
            break;
        }
        case kBatchSortHist: {
            const SortEntry& E = *static_cast<const SortEntry*>(entry);
            hipLaunchKernelGGL(k_dhist, dim3(E.gx), dim3(kSortThreads), dyn_lds, s, E.a);
            break;
        }
        case kBatchSortScan: {
            const SortEntry& E = *static_cast<const SortEntry*>(entry);
            hipLaunchKernelGGL(k_dscan, dim3(E.gx), dim3(kThreads), 0, s, E.a, E.rows);
            break;
        }
        case kBatchSortPlace: {
            const SortEntry& E = *static_cast<const SortEntry*>(entry);
            hipLaunchKernelGGL(k_dplace, dim3(E.gx), dim3(kSortThreads), dyn_lds, s, E.a);
            break;
        }
        case kBatchDeposit: {
            const DensityEntry& E = *static_cast<const DensityEntry*>(entry);
            hipLaunchKernelGGL(k_ddeposit_small, dim3(E.gx, E.gy), dim3(kBatchDepositThreads), 0, s, E.a, E.sorted_blocks);
            break;
        }
        case kBatchBins: {
            const DensityEntry& E = *static_cast<const DensityEntry*>(entry);
            hipLaunchKernelGGL(k_dbins, dim3(E.gx, E.gy), dim3(kThreads), 0, s, E.a);
            break;
        }
        case kBatchGemmStore:
        case kBatchGemmSpectral:
        case kBatchGemmInterleave: {
            const GemmEntry& E = *static_cast<const GemmEntry*>(entry);
            const dim3 grid(E.gx, E.gy, E.gz);
            if (k == kBatchGemmSpectral)
                hipLaunchKernelGGL(k_gemm<kEpiSpectral>, grid, dim3(kThreads), 0, s, E.g, E.K, E.lda, E.ldb, E.ldc, E.sp);
            else if (k == kBatchGemmInterleave)
                hipLaunchKernelGGL(k_gemm<kEpiInterleave>, grid, dim3(kThreads), 0, s, E.g, E.K, E.lda, E.ldb, E.ldc, E.sp);
            else
                hipLaunchKernelGGL(k_gemm<kEpiStore>, grid, dim3(kThreads), 0, s, E.g, E.K, E.lda, E.ldb, E.ldc, E.sp);
            break;
        }
        case kBatchGather: {
            const GatherEntry& E = *static_cast<const GatherEntry*>(entry);
            hipLaunchKernelGGL(k_gather, dim3(E.gx, E.gy), dim3(kThreads), 0, s, E.a);
            break;
        }
        case kBatchLookahead: {
            const LookaheadEntry& E = *static_cast<const LookaheadEntry*>(entry);
            hipLaunchKernelGGL(k_lookahead, dim3(E.gx), dim3(kThreads), 0, s, E.l);
            break;
        }
        case kBatchPropose: {
            const ProposeEntry& E = *static_cast<const ProposeEntry*>(entry);
            hipLaunchKernelGGL(k_propose, dim3(E.gx), dim3(kThreads), 0, s, E.p);
            break;
        }
        case kBatchCopy: {
            const CopyEntry& E = *static_cast<const CopyEntry*>(entry);
            check(hipMemcpyAsync(E.dst, E.src, static_cast<size_t>(E.n) * sizeof(double2), hipMemcpyDeviceToDevice, s),
                  "copy positions");
            return;
        }
        case kBatchZero: {
            const ZeroEntry& E = *static_cast<const ZeroEntry*>(entry);
            hipLaunchKernelGGL(k_zero, dim3(E.gx), dim3(kThreads), 0, s, E.p, static_cast<size_t>(E.n));
            break;
        }
        case kBatchFinalize: {
            RedList l{};
            l.job[0] = *static_cast<const RedJob*>(entry);
            l.n = 1;
            hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kFinalizeThreads), 0, s, l);
            break;
        }

        default: {
            const int v = static_cast<int>(k) - kBatchDct;
            if (v < 0 || v >= kDctVariants) throw std::invalid_argument("unknown batched kernel");
            const DctEntry& E = *static_cast<const DctEntry*>(entry);
            kDctLaunch[v](dim3(E.gx, E.gy, E.gz), E.d, E.n, E.lda, E.ldb, E.ldc, E.step, E.sp, s);
        }
    }
    check(hipGetLastError(), "launch kernel from its entry");
}

void launchFinalizeBatch(const RedJob* jobs, int n, volatile unsigned* flags, unsigned seq, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_finalize_batch, dim3(n), dim3(kFinalizeThreads), 0, s, jobs, flags, seq);
    check(hipGetLastError(), "launch batched reduction finalize kernel");
}

}  // namespace mpgp
