// gp_kernels.hpp — launch interface of the gfx950 kernels of the global-placement evaluate.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mpgp {

// Slots of the device scalar block (FP64).  Each producing kernel's last workgroup writes its
// slots after reducing the per-workgroup partials in a fixed order.
enum ScalarSlot : int {
    kHpwl = 0,
    kSmooth = 1,
    kEnergy = 2,
    kOptDensity = 3,      // 8 slots: DensityMetrics with fillers
    kDesignDensity = 11,  // 8 slots: without fillers
    kWireL1 = 19,
    kDensityL1 = 20,
    kRawNorm2 = 21,
    kPreNorm2 = 22,
    kPredicted = 23,
    kMaxDisp = 24,
    kAlignment = 25,
    kCurvDp = 26,
    kCurvDg = 27,
    kCurvDp2 = 28,
    kCurvDg2 = 29,
    kScalarCount = 32
};
// The scalar block holds kScalarBanks banks of kScalarCount slots, so several evaluations can be
// enqueued back to back and read with one copy.
constexpr int kScalarBanks = 4;

// Reductions: every reducing kernel writes one partial row per workgroup (no fences, no atomics);
// one k_finalize launch before each read-back reduces all pending partial sets in index order.
enum RedKind : int { kRedWire = 0, kRedMetrics, kRedEnergy, kRedGather, kRedPropose, kRedDot, kRedCurv, kRedCurv2,
                     kRedKinds };
struct RedJob {
    const double* partials;
    int blocks;
    int kind;
    double* out;     // scalar bank
    int slot;        // dot / curvature output slot
    double p0, p1, p2;  // dot: p0 sign; energy: p0 half bin area; metrics: p0 bin count, p1 dark, p2 placeable
};
constexpr int kMaxRedJobs = 12;
struct RedList {
    RedJob job[kMaxRedJobs];
    int n;
};
int redWidth(int kind);
void launchFinalize(const RedList& l, hipStream_t s);

struct Region {
    double llx, lly, urx, ury;  // placement region (core)
    double bw, bh;              // bin width / height
    int nx, ny;                 // bins
};

// ---- wirelength ------------------------------------------------------------------------------
// Nets (degree >= 2 only) in the backend's internal order: n_small nets of degree 2..kSmallDegree,
// then n_medium of degree ..kMediumDegree, then n_large; pins follow the same order.
struct WireArgs {
    const double2* pos;
    const int* net_off;
    const double* net_w;
    const int* pin_part;
    const double2* pin_off;
    int n_small, n_medium, n_large;
    double gamma;
    int want_grad;
    float2* pin_grad;  // per-pin d(WL)/d(pin), internal pin order
    double* partials;
};
constexpr int kSmallDegree = 8;
constexpr int kMediumDegree = 32;
int wireBlocks(int n_small, int n_medium, int n_large);
void launchWire(const WireArgs& a, hipStream_t s);

// ---- density: counting sort by first bin row, band accumulation in LDS, bin metrics ------------
constexpr int kSortThreads = 1024;   // particles per histogram / placement workgroup
constexpr int kBandThreads = 1024;
constexpr int kBandLookback = 4;     // bin rows a band looks back for particles that reach into it
struct DensityArgs {
    const double2* pos;
    const double2* size;   // (w, h)
    const int* layer;      // 0 std, 1 macro, 2 filler
    const unsigned char* tall;  // 1: taller than the band look-back, deposited through the residual layer
    int n;
    Region r;
    double t;
    double fix, inv_fix;   // fixed-point scale: 2^40 per bin area
    int band_rows, bands, lookback, has_tall;
    int* key;              // [P]
    int* hist;             // [histogram blocks][ny]
    int* colprefix;        // [histogram blocks][ny]
    int* keytotal;         // [ny]
    int* keystart;         // [ny + 1]
    double4* sorted;       // [P] clipped (left, right, bottom, top) in row order
    int4* sorted_span;     // [P] (c0 | layer << 28, c1, r0, r1)
    unsigned long long* residual;  // [2][G] fixed point, movable + t*macro then filler
    const double* static_charge;   // [G] t*(fixed + dark)
    float* rho;            // [nyp*nxp] density deviation, FP32, padded
    int ldrho;
    double* partials;      // one metrics row per band
};
int densityHistBlocks(int n);
void launchDensity(const DensityArgs& a, hipStream_t s);

// ---- DCT products on MFMA (FP32) ---------------------------------------------------------------
enum Epilogue : int { kEpiStore = 0, kEpiSpectral = 1 };
struct GemmBatch {
    const float* A[3];
    const float* B[3];
    float* C[3];
};
struct SpectralArgs {
    int nx, ny;
    double W, H;          // region width / height
    double normalization; // 1 / (4 nx ny)
    double* partials;     // energy partials of the spectral epilogue
};
// C[MxN] = A[MxK] * B[KxN], all row-major, M, N multiples of 32, K multiple of 16.
void launchGemm(const GemmBatch& g, int batch, int M, int N, int K, int lda, int ldb, int ldc, int epilogue,
                const SpectralArgs& sp, int precision, hipStream_t s);

// ---- field gather + gradient assembly ---------------------------------------------------------
struct GatherArgs {
    const double2* pos;
    const double* charge;
    const float* ex;
    const float* ey;
    int ldf;
    Region r;
    const int* ppin_off;   // [P+1] particle -> pin CSR
    const int* ppin_idx;   // internal pin index of each CSR entry
    const float2* pin_grad;
    int n;
    // backend mode: write the two gradients
    double2* wire_out;
    double2* density_out;
    // session mode: raw = wire + penalty * density, pre = raw * base / strength
    const double* precond;
    double penalty, inv_strength;
    double2* raw_out;
    double2* pre_out;
    double* partials;
};
int particleBlocks(int n);
void launchGather(const GatherArgs& a, hipStream_t s);

// ---- controller vector passes (AdaptiveGlobalPlacer.cpp O(P) loops) ---------------------------
struct ClampArgs {
    const double2* size;
    Region r;
};
// out = clamp(x + beta * v)
void launchLookahead(const double2* x, const double2* v, double beta, double2* out, const ClampArgs& c, int n,
                     hipStream_t s);
// partials of sum(a . b)
void launchDot(const double2* a, const double2* b, int n, double* partials, hipStream_t s);
// partials of sum |b - pb|^2 and sum |g - pg|^2
void launchCurvature(const double2* b, const double2* pb, const double2* g, const double2* pg, int n,
                     double* partials, hipStream_t s);
// y = clamp(b + cap(-step * g, max_disp)); partials of sum raw.(b - y) and max |b - y|
void launchPropose(const double2* b, const double2* g, const double2* raw, double step, double max_disp,
                   double2* y, const ClampArgs& c, int n, double* partials, hipStream_t s);
// v = y - x
void launchDifference(const double2* y, const double2* x, double2* v, int n, hipStream_t s);
// fill with zero (used instead of memset so the sequence can be captured)
void launchZero(double* p, size_t count, hipStream_t s);

}  // namespace mpgp
