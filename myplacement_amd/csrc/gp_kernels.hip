// gp_kernels.hip — gfx950 kernels of the global-placement evaluate and controller passes.
//
// Reference counterparts (src/placement/cuda/CudaPlacementBackend.cu unless noted):
//   k_wire      buildPinPositions + computeNetBounds + computeNetSums + computePinGradient (:194-337),
//               CPU form AdaptiveGlobalPlacer.cpp:29-82
//   k_scatter   depositParticlesKernel (:339-375), CPU DensityMap::addRectangle Metrics.cpp:74-99
//   k_metrics   buildDensityMetricsKernel + 10 reductions (:377-398, :1022-1038), Metrics.cpp:101-119
//   k_gemm      forwardDct / solveSpectralKernel / inverse2d (:547-567, :679-742); CPU
//               ElectrostaticField.cpp:205-237 (FFTW REDFT10 / REDFT01 / RODFT01 as dense products)
//   k_gather    sampleDensityGradientKernel (:575-591) + the gradient combine of
//               AdaptiveGlobalPlacer.cpp:308-329
//   controller  proposeStep / estimateStep / momentumAlignment / look-ahead (AdaptiveGlobalPlacer.cpp
//               :470-519, :610-617, :776-780)
//
// Every reduction writes one partial per workgroup; the last workgroup to finish (atomic ticket)
// reduces them in index order, so the scalars are deterministic for a given set of partials.

#include "gp_kernels.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

namespace mpgp {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr double kEps = 1e-9;
constexpr double kPi = 3.141592653589793238462643383279502884;

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
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

__device__ inline double loadVolatile(const double* p) { return *reinterpret_cast<const volatile double*>(p); }

// Sum (or max, bit k of max_mask) of K values over a 256-thread block; result in thread 0.
template <int K>
__device__ void blockReduce(double (&v)[K], unsigned max_mask, double* lds) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = ((max_mask >> k) & 1U) ? waveMax(v[k]) : waveSum(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) lds[wave * K + k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double a = lds[k];
            for (int w = 1; w < kWaves; ++w) a = ((max_mask >> k) & 1U) ? fmax(a, lds[w * K + k]) : a + lds[w * K + k];
            v[k] = a;
        }
    __syncthreads();
}

// Publishes this block's K partials; the last block reduces all of them (index order) into tot.
// Returns true in every thread of the last block; tot is valid in thread 0.
template <int K>
__device__ bool finishReduction(double (&v)[K], unsigned max_mask, double init_max, int block, int blocks,
                                double* partials, unsigned* counter, double* lds, double (&tot)[K]) {
    __shared__ int last;
    blockReduce<K>(v, max_mask, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) partials[static_cast<size_t>(block) * K + k] = v[k];
        __threadfence();
        const unsigned ticket = atomicAdd(counter, 1U);
        last = ticket == static_cast<unsigned>(blocks - 1);
    }
    __syncthreads();
    if (!last) return false;
    __threadfence();
#pragma unroll
    for (int k = 0; k < K; ++k) tot[k] = ((max_mask >> k) & 1U) ? init_max : 0.0;
    for (int b = threadIdx.x; b < blocks; b += kThreads)
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const double x = loadVolatile(partials + static_cast<size_t>(b) * K + k);
            tot[k] = ((max_mask >> k) & 1U) ? fmax(tot[k], x) : tot[k] + x;
        }
    blockReduce<K>(tot, max_mask, lds);
    if (threadIdx.x == 0) *counter = 0U;
    return true;
}

__device__ inline double2 pinPosition(const double2* pos, const int* pin_part, const double2* pin_off, int p) {
    const int part = pin_part[p];
    const double2 o = pin_off[p];
    if (part < 0) return o;
    const double2 c = pos[part];
    return make_double2(c.x + o.x, c.y + o.y);
}

// ---------------------------------------------------------------------------------------------
// Wirelength: small nets one thread each, large nets one wavefront each (same launch).
__global__ __launch_bounds__(kThreads) void k_wire(WireArgs a, int small_blocks) {
    __shared__ double lds[kWaves * 2];
    double wl = 0.0, hp = 0.0;
    const double inv_gamma = 1.0 / a.gamma;
    const float gamma_f = static_cast<float>(a.gamma);
    if (static_cast<int>(blockIdx.x) < small_blocks) {
        const int idx = blockIdx.x * kThreads + threadIdx.x;
        if (idx < a.n_small) {
            const int net = a.small_nets[idx];
            const int b = a.net_off[net], e = a.net_off[net + 1];
            double mx = -INFINITY, nx = INFINITY, my = -INFINITY, ny = INFINITY;
            for (int p = b; p < e; ++p) {
                const double2 q = pinPosition(a.pos, a.pin_part, a.pin_off, p);
                mx = fmax(mx, q.x); nx = fmin(nx, q.x);
                my = fmax(my, q.y); ny = fmin(ny, q.y);
            }
            float spx = 0.f, snx = 0.f, spy = 0.f, sny = 0.f;
            for (int p = b; p < e; ++p) {
                const double2 q = pinPosition(a.pos, a.pin_part, a.pin_off, p);
                spx += __expf(static_cast<float>((q.x - mx) * inv_gamma));
                snx += __expf(static_cast<float>((nx - q.x) * inv_gamma));
                spy += __expf(static_cast<float>((q.y - my) * inv_gamma));
                sny += __expf(static_cast<float>((ny - q.y) * inv_gamma));
            }
            const double w = a.net_w[net];
            const double lse = static_cast<double>(gamma_f * (__logf(spx) + __logf(snx) + __logf(spy) + __logf(sny)));
            wl = w * ((mx - nx) + (my - ny) + lse);
            hp = w * ((mx - nx) + (my - ny));
            if (a.want_grad) {
                const float wf = static_cast<float>(w);
                const float ipx = 1.f / spx, inx = 1.f / snx, ipy = 1.f / spy, iny = 1.f / sny;
                for (int p = b; p < e; ++p) {
                    const double2 q = pinPosition(a.pos, a.pin_part, a.pin_off, p);
                    const float gx = __expf(static_cast<float>((q.x - mx) * inv_gamma)) * ipx -
                                     __expf(static_cast<float>((nx - q.x) * inv_gamma)) * inx;
                    const float gy = __expf(static_cast<float>((q.y - my) * inv_gamma)) * ipy -
                                     __expf(static_cast<float>((ny - q.y) * inv_gamma)) * iny;
                    a.pin_grad[p] = make_float2(wf * gx, wf * gy);
                }
            }
        }
    } else {
        const int lane = threadIdx.x & 63;
        const int item = (blockIdx.x - small_blocks) * kWaves + (threadIdx.x >> 6);
        if (item < a.n_large) {
            const int net = a.large_nets[item];
            const int b = a.net_off[net], e = a.net_off[net + 1];
            double mx = -INFINITY, nx = INFINITY, my = -INFINITY, ny = INFINITY;
            for (int p = b + lane; p < e; p += 64) {
                const double2 q = pinPosition(a.pos, a.pin_part, a.pin_off, p);
                mx = fmax(mx, q.x); nx = fmin(nx, q.x);
                my = fmax(my, q.y); ny = fmin(ny, q.y);
            }
            mx = waveMax(mx); nx = waveMin(nx); my = waveMax(my); ny = waveMin(ny);
            float spx = 0.f, snx = 0.f, spy = 0.f, sny = 0.f;
            for (int p = b + lane; p < e; p += 64) {
                const double2 q = pinPosition(a.pos, a.pin_part, a.pin_off, p);
                spx += __expf(static_cast<float>((q.x - mx) * inv_gamma));
                snx += __expf(static_cast<float>((nx - q.x) * inv_gamma));
                spy += __expf(static_cast<float>((q.y - my) * inv_gamma));
                sny += __expf(static_cast<float>((ny - q.y) * inv_gamma));
            }
            spx = waveSumF(spx); snx = waveSumF(snx); spy = waveSumF(spy); sny = waveSumF(sny);
            const double w = a.net_w[net];
            if (lane == 0) {
                const double lse = static_cast<double>(gamma_f * (__logf(spx) + __logf(snx) + __logf(spy) + __logf(sny)));
                wl = w * ((mx - nx) + (my - ny) + lse);
                hp = w * ((mx - nx) + (my - ny));
            }
            if (a.want_grad) {
                const float wf = static_cast<float>(w);
                const float ipx = 1.f / spx, inx = 1.f / snx, ipy = 1.f / spy, iny = 1.f / sny;
                for (int p = b + lane; p < e; p += 64) {
                    const double2 q = pinPosition(a.pos, a.pin_part, a.pin_off, p);
                    const float gx = __expf(static_cast<float>((q.x - mx) * inv_gamma)) * ipx -
                                     __expf(static_cast<float>((nx - q.x) * inv_gamma)) * inx;
                    const float gy = __expf(static_cast<float>((q.y - my) * inv_gamma)) * ipy -
                                     __expf(static_cast<float>((ny - q.y) * inv_gamma)) * iny;
                    a.pin_grad[p] = make_float2(wf * gx, wf * gy);
                }
            }
        }
    }
    double v[2] = {hp, wl};
    double tot[2];
    if (finishReduction<2>(v, 0U, 0.0, blockIdx.x, gridDim.x, a.partials, a.counter, lds, tot) &&
        threadIdx.x == 0) {
        a.scal[kHpwl] = tot[0];
        a.scal[kSmooth] = tot[1];
    }
}

// ---------------------------------------------------------------------------------------------
// Density scatter: exact rectangle / bin overlap in FP64 (bin index math as Metrics.cpp:78-85).
__global__ __launch_bounds__(kThreads) void k_scatter(ScatterArgs a) {
    const int p = blockIdx.x * kThreads + threadIdx.x;
    if (p >= a.n) return;
    const double2 c = a.pos[p];
    const double2 sz = a.size[p];
    const Region& r = a.r;
    const double left = fmax(c.x - sz.x * 0.5, r.llx);
    const double right = fmin(c.x + sz.x * 0.5, r.urx);
    const double bottom = fmax(c.y - sz.y * 0.5, r.lly);
    const double top = fmin(c.y + sz.y * 0.5, r.ury);
    if (fmax(0.0, right - left) * fmax(0.0, top - bottom) <= kEps) return;
    const int c0 = max(0, min(r.nx - 1, static_cast<int>(floor((left - r.llx) / r.bw))));
    const int c1 = max(0, min(r.nx - 1, static_cast<int>(ceil((right - r.llx) / r.bw) - 1.0)));
    const int r0 = max(0, min(r.ny - 1, static_cast<int>(floor((bottom - r.lly) / r.bh))));
    const int r1 = max(0, min(r.ny - 1, static_cast<int>(ceil((top - r.lly) / r.bh) - 1.0)));
    const int layer = a.layer[p];
    const double scale = layer == 1 ? a.t : 1.0;
    double* dst = layer == 2 ? a.filler : a.movable;
    for (int row = r0; row <= r1; ++row) {
        const double bin_bottom = r.lly + static_cast<double>(row) * r.bh;
        const double oy = fmax(0.0, fmin(top, bin_bottom + r.bh) - fmax(bottom, bin_bottom));
        for (int col = c0; col <= c1; ++col) {
            const double bin_left = r.llx + static_cast<double>(col) * r.bw;
            const double ox = fmax(0.0, fmin(right, bin_left + r.bw) - fmax(left, bin_left));
            const double area = ox * oy;
            if (area <= kEps) continue;
            atomicAdd(dst + static_cast<size_t>(row) * r.nx + col, scale * area);
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Per-bin charge, overflow, density and the FP32 density deviation for the DCT; clears the
// dynamic layers for the next scatter.
__global__ __launch_bounds__(kThreads) void k_metrics(MetricsArgs a) {
    __shared__ double lds[kWaves * 9];
    const Region& r = a.r;
    const int G = r.nx * r.ny;
    // sums: ovf_w, ovf_wo, q_w, q_wo, norm, dens_w, dens_wo ; max: dmax_w, dmax_wo
    double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < G; i += gridDim.x * kThreads) {
        const int row = i / r.nx, col = i - row * r.nx;
        const double left = r.llx + static_cast<double>(col) * r.bw;
        const double bottom = r.lly + static_cast<double>(row) * r.bh;
        const double area = fmax(0.0, (left + r.bw) - left) * fmax(0.0, (bottom + r.bh) - bottom);
        const double mov = a.movable[i];
        const double fil = a.filler[i];
        a.movable[i] = 0.0;
        a.filler[i] = 0.0;
        const double q_wo = mov + a.static_charge[i];
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
        a.rho[static_cast<size_t>(row) * a.ldrho + col] = static_cast<float>(d_w - a.t);
    }
    double tot[9];
    const unsigned max_mask = (1U << 7) | (1U << 8);
    if (finishReduction<9>(v, max_mask, 0.0, blockIdx.x, gridDim.x, a.partials, a.counter, lds, tot) &&
        threadIdx.x == 0) {
        const double Gd = static_cast<double>(G);
        for (int f = 0; f < 2; ++f) {
            double* m = a.scal + (f == 0 ? kOptDensity : kDesignDensity);
            m[0] = tot[f];                                     // total_overflow_area
            m[5] = tot[4];                                     // normalization_area
            m[1] = m[0] / fmax(m[5], kEps);                    // normalized_overflow
            m[2] = tot[7 + f];                                 // maximum_density (starts at 0)
            m[3] = tot[5 + f] / Gd;                            // average_density
            m[4] = tot[2 + f];                                 // total_charge_area
            m[6] = a.placeable_total;                          // placeable_area
            m[7] = a.dark_total;                               // dark_area
        }
    }
}

// ---------------------------------------------------------------------------------------------
// C = A * B on v_mfma_f32_16x16x4_f32.  Workgroup tile 32x32, four wavefronts split K in
// 16-wide chunks and meet in LDS.  Lane layout of the 16x16x4 instruction: A[i][k] from lane
// i + 16k, B[k][j] from lane j + 16k, D[4*(lane/16) + v][lane%16] in accumulator element v.
typedef float floatx4 __attribute__((ext_vector_type(4)));

template <int EPI>
__global__ __launch_bounds__(kThreads) void k_gemm(GemmBatch g, int K, int lda, int ldb, int ldc, SpectralArgs sp) {
    __shared__ float red[kWaves][32][33];
    __shared__ double lds[kWaves];
    const int z = blockIdx.z;
    const float* __restrict__ A = g.A[z];
    const float* __restrict__ B = g.B[z];
    float* __restrict__ C = g.C[z];
    const int m0 = blockIdx.y * 32, n0 = blockIdx.x * 32;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, r16 = lane & 15, grp = lane >> 4;
    floatx4 acc00 = {0.f, 0.f, 0.f, 0.f}, acc01 = acc00, acc10 = acc00, acc11 = acc00;
    const float* a_row0 = A + static_cast<size_t>(m0 + r16) * lda + 4 * grp;
    const float* a_row1 = a_row0 + static_cast<size_t>(16) * lda;
    const float* b_col = B + static_cast<size_t>(4 * grp) * ldb + n0 + r16;
    for (int k0 = wave * 16; k0 < K; k0 += 16 * kWaves) {
        const float4 a0 = *reinterpret_cast<const float4*>(a_row0 + k0);
        const float4 a1 = *reinterpret_cast<const float4*>(a_row1 + k0);
        const float* bk = b_col + static_cast<size_t>(k0) * ldb;
        float b0[4], b1[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            b0[j] = bk[static_cast<size_t>(j) * ldb];
            b1[j] = bk[static_cast<size_t>(j) * ldb + 16];
        }
        const float av0[4] = {a0.x, a0.y, a0.z, a0.w};
        const float av1[4] = {a1.x, a1.y, a1.z, a1.w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            acc00 = __builtin_amdgcn_mfma_f32_16x16x4f32(av0[j], b0[j], acc00, 0, 0, 0);
            acc01 = __builtin_amdgcn_mfma_f32_16x16x4f32(av0[j], b1[j], acc01, 0, 0, 0);
            acc10 = __builtin_amdgcn_mfma_f32_16x16x4f32(av1[j], b0[j], acc10, 0, 0, 0);
            acc11 = __builtin_amdgcn_mfma_f32_16x16x4f32(av1[j], b1[j], acc11, 0, 0, 0);
        }
    }
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
        C[static_cast<size_t>(m0 + row) * ldc + n0 + col] = out;
    }
    if (EPI == kEpiSpectral) {
        double v[1] = {energy};
        double tot[1];
        const int block = blockIdx.y * gridDim.x + blockIdx.x;
        if (finishReduction<1>(v, 0U, 0.0, block, gridDim.x * gridDim.y, sp.partials, sp.counter, lds, tot) &&
            threadIdx.x == 0)
            sp.scal[kEnergy] = fmax(0.0, sp.half_bin_area * tot[0]);
    }
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

__device__ inline double sampleParity(const float* __restrict__ v, int ld, double x, double y, const Region& r,
                                      bool odd_x, bool odd_y) {
    x = fmax(r.llx, fmin(x, r.urx));
    y = fmax(r.lly, fmin(y, r.ury));
    const double gx = (x - r.llx) / r.bw - 0.5;
    const double gy = (y - r.lly) / r.bh - 0.5;
    const int left = static_cast<int>(floor(gx));
    const int bottom = static_cast<int>(floor(gy));
    const double tx = gx - static_cast<double>(left);
    const double ty = gy - static_cast<double>(bottom);
    double val[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        bool neg = false;
        const int cc = reflectIndex(left + (q & 1), r.nx, odd_x, neg);
        const int rr = reflectIndex(bottom + (q >> 1), r.ny, odd_y, neg);
        const double f = static_cast<double>(v[static_cast<size_t>(rr) * ld + cc]);
        val[q] = neg ? -f : f;
    }
    const double lower = val[0] * (1.0 - tx) + val[1] * tx;
    const double upper = val[2] * (1.0 - tx) + val[3] * tx;
    return lower * (1.0 - ty) + upper * ty;
}

__global__ __launch_bounds__(kThreads) void k_gather(GatherArgs a) {
    __shared__ double lds[kWaves * 4];
    const int p = blockIdx.x * kThreads + threadIdx.x;
    double v[4] = {0, 0, 0, 0};  // wire L1, density L1, |raw|^2, |pre|^2
    if (p < a.n) {
        const double2 c = a.pos[p];
        const double q = a.charge[p];
        const double ex = sampleParity(a.ex, a.ldf, c.x, c.y, a.r, true, false);
        const double ey = sampleParity(a.ey, a.ldf, c.x, c.y, a.r, false, true);
        const double2 dg = make_double2(-q * ex, -q * ey);
        double2 wg = make_double2(0.0, 0.0);
        for (int j = a.ppin_off[p]; j < a.ppin_off[p + 1]; ++j) {
            const float2 g = a.pin_grad[a.ppin_idx[j]];
            wg.x += static_cast<double>(g.x);
            wg.y += static_cast<double>(g.y);
        }
        v[0] = fabs(wg.x) + fabs(wg.y);
        v[1] = fabs(dg.x) + fabs(dg.y);
        if (a.wire_out) {
            a.wire_out[p] = wg;
            a.density_out[p] = dg;
        }
        if (a.raw_out) {
            const double2 raw = make_double2(wg.x + dg.x * a.penalty, wg.y + dg.y * a.penalty);
            const double s = a.precond[p] * a.inv_strength;
            const double2 pre = make_double2(raw.x * s, raw.y * s);
            a.raw_out[p] = raw;
            a.pre_out[p] = pre;
            v[2] = raw.x * raw.x + raw.y * raw.y;
            v[3] = pre.x * pre.x + pre.y * pre.y;
        }
    }
    double tot[4];
    if (finishReduction<4>(v, 0U, 0.0, blockIdx.x, gridDim.x, a.partials, a.counter, lds, tot) &&
        threadIdx.x == 0) {
        a.scal[kWireL1] = tot[0];
        a.scal[kDensityL1] = tot[1];
        a.scal[kRawNorm2] = tot[2];
        a.scal[kPreNorm2] = tot[3];
    }
}

// ---------------------------------------------------------------------------------------------
// Controller vector passes.
__device__ inline double2 clampCenter(double2 c, double2 sz, const Region& r) {  // GlobalPlacementSupport.cpp:11-19
    const double lx = r.llx + sz.x * 0.5, ux = r.urx - sz.x * 0.5;
    const double ly = r.lly + sz.y * 0.5, uy = r.ury - sz.y * 0.5;
    c.x = lx <= ux ? fmax(lx, fmin(c.x, ux)) : (r.llx + r.urx) * 0.5;
    c.y = ly <= uy ? fmax(ly, fmin(c.y, uy)) : (r.lly + r.ury) * 0.5;
    return c;
}

__global__ __launch_bounds__(kThreads) void k_lookahead(const double2* x, const double2* v, double beta,
                                                        double2* out, ClampArgs c, int n) {
    const int p = blockIdx.x * kThreads + threadIdx.x;
    if (p >= n) return;
    const double2 a = x[p], b = v[p];
    out[p] = clampCenter(make_double2(a.x + b.x * beta, a.y + b.y * beta), c.size[p], c.r);
}

__global__ __launch_bounds__(kThreads) void k_dot(const double2* a, const double2* b, double sign, int slot, int n,
                                                  double* partials, unsigned* counter, double* scal) {
    __shared__ double lds[kWaves];
    const int p = blockIdx.x * kThreads + threadIdx.x;
    double v[1] = {0.0};
    if (p < n) {
        const double2 x = a[p], y = b[p];
        v[0] = x.x * y.x + x.y * y.y;
    }
    double tot[1];
    if (finishReduction<1>(v, 0U, 0.0, blockIdx.x, gridDim.x, partials, counter, lds, tot) && threadIdx.x == 0)
        scal[slot] = sign * tot[0];
}

__global__ __launch_bounds__(kThreads) void k_curvature(const double2* b, const double2* pb, const double2* g,
                                                        const double2* pg, int n, int slot, double* partials,
                                                        unsigned* counter, double* scal) {
    __shared__ double lds[kWaves * 2];
    const int p = blockIdx.x * kThreads + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (p < n) {
        const double2 x = b[p], px = pb[p], y = g[p], py = pg[p];
        const double dx = x.x - px.x, dy = x.y - px.y, gx = y.x - py.x, gy = y.y - py.y;
        v[0] = dx * dx + dy * dy;
        v[1] = gx * gx + gy * gy;
    }
    double tot[2];
    if (finishReduction<2>(v, 0U, 0.0, blockIdx.x, gridDim.x, partials, counter, lds, tot) && threadIdx.x == 0) {
        scal[slot] = tot[0];
        scal[slot + 1] = tot[1];
    }
}

__global__ __launch_bounds__(kThreads) void k_propose(const double2* b, const double2* g, const double2* raw, double step,
                                                      double max_disp, double2* y, ClampArgs c, int n,
                                                      double* partials, unsigned* counter, double* scal) {
    __shared__ double lds[kWaves * 2];
    const int p = blockIdx.x * kThreads + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (p < n) {
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
        const double ax = base.x - out.x, ay = base.y - out.y;
        const double2 r = raw[p];
        v[0] = r.x * ax + r.y * ay;
        v[1] = sqrt(ax * ax + ay * ay);
    }
    double tot[2];
    if (finishReduction<2>(v, 2U, 0.0, blockIdx.x, gridDim.x, partials, counter, lds, tot) && threadIdx.x == 0) {
        scal[kPredicted] = tot[0];
        scal[kMaxDisp] = tot[1];
    }
}

__global__ __launch_bounds__(kThreads) void k_difference(const double2* y, const double2* x, double2* v, int n) {
    const int p = blockIdx.x * kThreads + threadIdx.x;
    if (p >= n) return;
    const double2 a = y[p], b = x[p];
    v[p] = make_double2(a.x - b.x, a.y - b.y);
}

__global__ __launch_bounds__(kThreads) void k_zero(double* p, size_t count) {
    for (size_t i = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x; i < count;
         i += static_cast<size_t>(gridDim.x) * kThreads)
        p[i] = 0.0;
}

int blocksFor(long long n) { return static_cast<int>((n + kThreads - 1) / kThreads); }

}  // namespace

int wireBlocks(int n_small, int n_large) { return blocksFor(n_small) + (n_large + kWaves - 1) / kWaves; }

void launchWire(const WireArgs& a, hipStream_t s) {
    const int small_blocks = blocksFor(a.n_small);
    const int blocks = wireBlocks(a.n_small, a.n_large);
    if (blocks == 0) return;
    hipLaunchKernelGGL(k_wire, dim3(blocks), dim3(kThreads), 0, s, a, small_blocks);
    check(hipGetLastError(), "launch wirelength kernel");
}

void launchScatter(const ScatterArgs& a, hipStream_t s) {
    if (a.n == 0) return;
    hipLaunchKernelGGL(k_scatter, dim3(blocksFor(a.n)), dim3(kThreads), 0, s, a);
    check(hipGetLastError(), "launch density scatter kernel");
}

int metricsBlocks(int G) { return std::min(blocksFor(G), 1024); }

void launchMetrics(const MetricsArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_metrics, dim3(metricsBlocks(a.r.nx * a.r.ny)), dim3(kThreads), 0, s, a);
    check(hipGetLastError(), "launch density metrics kernel");
}

void launchGemm(const GemmBatch& g, int batch, int M, int N, int K, int lda, int ldb, int ldc, int epilogue,
                const SpectralArgs& sp, int precision, hipStream_t s) {
    if (M % 32 || N % 32 || K % 16) throw std::invalid_argument("DCT product dimensions are not tile multiples");
    if (precision != 0) throw std::invalid_argument("only the FP32 DCT products are built");
    const dim3 grid(N / 32, M / 32, batch);
    if (epilogue == kEpiSpectral)
        hipLaunchKernelGGL(k_gemm<kEpiSpectral>, grid, dim3(kThreads), 0, s, g, K, lda, ldb, ldc, sp);
    else
        hipLaunchKernelGGL(k_gemm<kEpiStore>, grid, dim3(kThreads), 0, s, g, K, lda, ldb, ldc, sp);
    check(hipGetLastError(), "launch DCT product kernel");
}

int particleBlocks(int n) { return std::max(1, blocksFor(n)); }

void launchGather(const GatherArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_gather, dim3(particleBlocks(a.n)), dim3(kThreads), 0, s, a);
    check(hipGetLastError(), "launch field gather kernel");
}

void launchLookahead(const double2* x, const double2* v, double beta, double2* out, const ClampArgs& c, int n,
                     hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_lookahead, dim3(blocksFor(n)), dim3(kThreads), 0, s, x, v, beta, out, c, n);
    check(hipGetLastError(), "launch look-ahead kernel");
}

void launchDot(const double2* a, const double2* b, double sign, int slot, int n, double* partials, unsigned* counter,
               double* scal, hipStream_t s) {
    hipLaunchKernelGGL(k_dot, dim3(particleBlocks(n)), dim3(kThreads), 0, s, a, b, sign, slot, n, partials, counter,
                       scal);
    check(hipGetLastError(), "launch dot kernel");
}

void launchCurvature(const double2* b, const double2* pb, const double2* g, const double2* pg, int n, int slot,
                     double* partials, unsigned* counter, double* scal, hipStream_t s) {
    hipLaunchKernelGGL(k_curvature, dim3(particleBlocks(n)), dim3(kThreads), 0, s, b, pb, g, pg, n, slot, partials,
                       counter, scal);
    check(hipGetLastError(), "launch curvature kernel");
}

void launchPropose(const double2* b, const double2* g, const double2* raw, double step, double max_disp, double2* y,
                   const ClampArgs& c, int n, double* partials, unsigned* counter, double* scal, hipStream_t s) {
    hipLaunchKernelGGL(k_propose, dim3(particleBlocks(n)), dim3(kThreads), 0, s, b, g, raw, step, max_disp, y, c, n,
                       partials, counter, scal);
    check(hipGetLastError(), "launch propose kernel");
}

void launchDifference(const double2* y, const double2* x, double2* v, int n, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_difference, dim3(blocksFor(n)), dim3(kThreads), 0, s, y, x, v, n);
    check(hipGetLastError(), "launch difference kernel");
}

void launchZero(double* p, size_t count, hipStream_t s) {
    if (count == 0) return;
    const int blocks = static_cast<int>(std::min<size_t>(1024, (count + kThreads - 1) / kThreads));
    hipLaunchKernelGGL(k_zero, dim3(blocks), dim3(kThreads), 0, s, p, count);
    check(hipGetLastError(), "launch zero kernel");
}

}  // namespace mpgp
