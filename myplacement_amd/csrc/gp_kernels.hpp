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

// Reduction counters (one per reducing kernel; reset by that kernel's last workgroup).
enum Counter : int { kCtrWire = 0, kCtrMetrics = 1, kCtrEnergy = 2, kCtrGather = 3, kCtrPropose = 4,
                     kCtrDot = 5, kCtrCurv = 6, kCounterCount = 8 };

struct Region {
    double llx, lly, urx, ury;  // placement region (core)
    double bw, bh;              // bin width / height
    int nx, ny;                 // bins
};

// ---- wirelength ------------------------------------------------------------------------------
struct WireArgs {
    const double2* pos;
    const int* net_off;
    const double* net_w;
    const int* pin_part;
    const double2* pin_off;
    const int* small_nets;  // nets of degree 2..kSmallDegree, one thread each
    int n_small;
    const int* large_nets;  // nets of degree > kSmallDegree, one wavefront each
    int n_large;
    double gamma;
    int want_grad;
    float2* pin_grad;  // [I] per-pin d(WL)/d(pin), net order
    double* partials;
    unsigned* counter;
    double* scal;
};
constexpr int kSmallDegree = 8;
int wireBlocks(int n_small, int n_large);
void launchWire(const WireArgs& a, hipStream_t s);

// ---- density scatter + metrics ---------------------------------------------------------------
struct ScatterArgs {
    const double2* pos;
    const double2* size;   // (w, h)
    const int* layer;      // 0 std, 1 macro, 2 filler
    int n;
    Region r;
    double t;
    double* movable;       // [G] movable + t*macro (normalization layer)
    double* filler;        // [G] filler area
};
void launchScatter(const ScatterArgs& a, hipStream_t s);

struct MetricsArgs {
    double* movable;             // read then cleared
    double* filler;              // read then cleared
    const double* static_charge; // [G] t*(fixed + dark)
    Region r;
    double t;
    double dark_total, placeable_total;
    float* rho;                  // [nyp*nxp] density deviation, FP32, padded
    int ldrho;
    double* partials;
    unsigned* counter;
    double* scal;
};
int metricsBlocks(int G);
void launchMetrics(const MetricsArgs& a, hipStream_t s);

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
    double half_bin_area; // 0.5 * bin area (energy)
    double* partials;
    unsigned* counter;
    double* scal;
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
    const int* ppin_idx;
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
    unsigned* counter;
    double* scal;
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
// scal[slot] = sum(a . b) * sign
void launchDot(const double2* a, const double2* b, double sign, int slot, int n, double* partials,
               unsigned* counter, double* scal, hipStream_t s);
// scal[slot] = sum |b - pb|^2, scal[slot + 1] = sum |g - pg|^2
void launchCurvature(const double2* b, const double2* pb, const double2* g, const double2* pg, int n, int slot,
                     double* partials, unsigned* counter, double* scal, hipStream_t s);
// y = clamp(b + cap(-step * g, max_disp)); scal[kPredicted] = sum raw.(b - y); scal[kMaxDisp] = max |b - y|
void launchPropose(const double2* b, const double2* g, const double2* raw, double step, double max_disp,
                   double2* y, const ClampArgs& c, int n, double* partials, unsigned* counter, double* scal,
                   hipStream_t s);
// v = y - x
void launchDifference(const double2* y, const double2* x, double2* v, int n, hipStream_t s);
// fill with zero (used instead of memset so the sequence can be captured)
void launchZero(double* p, size_t count, hipStream_t s);

}  // namespace mpgp
