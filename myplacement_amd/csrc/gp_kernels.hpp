// gp_kernels.hpp — launch interface of the gfx950 kernels of the global-placement evaluate.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace mpgp {

// Slots of the scalar block (FP64), written by k_finalize after it reduces the per-workgroup
// partials in a fixed order.
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
    kAlignment = 25,  // -sum (pos - prev_pos) . pre of a gather with prev_pos
    kCurvDp = 26,     // sum |pos - px|^2 of a gather with px / pg
    kCurvDg = 27,     // sum |pre - pg|^2
    kDecidedLook = 28,  // the device's step decision (1: look-ahead base), step and displacement cap
    kDecidedStep = 29,
    kDecidedDisp = 30,
    kWireGammaMismatch = 31,  // 1: a wirelength evaluated ahead used another smoothing than the host's
    kScalarCount = 32
};
// The scalar block holds kScalarBanks banks of kScalarCount slots, so several evaluations can be
// enqueued back to back and read with one copy.
constexpr int kScalarBanks = 4;

// Reductions: every reducing kernel writes one partial row per workgroup (no fences, no atomics);
// one k_finalize launch before each read-back reduces all pending partial sets in index order.
enum RedKind : int { kRedWire = 0, kRedMetrics, kRedEnergy, kRedGather, kRedPropose, kRedKinds };
struct RedJob {
    const double* partials;
    int blocks;
    int kind;
    double* out;        // scalar bank (host-mapped)
    double* out_dev;    // its copy in device memory, or null
    double p0, p1, p2;  // energy: p0 half bin area; metrics: p0 bin count, p1 dark, p2 placeable
};
constexpr int kMaxRedJobs = 12;
struct RedList {
    RedJob job[kMaxRedJobs];
    int n;
    unsigned seq;              // completion value: workgroup b stores it to flag[b] after its results
    volatile unsigned* flag;   // host-mapped completion words (null: none)
};
int redWidth(int kind);
void launchFinalize(const RedList& l, hipStream_t s);

// Evaluations batched into one launch chain (the controller's current and look-ahead positions);
// per-evaluation pointers are indexed by blockIdx.y (blockIdx.z for the DCT products).
constexpr int kMaxEval = 2;

struct Region {
    double llx, lly, urx, ury;  // placement region (core)
    double bw, bh;              // bin width / height
    int nx, ny;                 // bins
};

// ---- wirelength ------------------------------------------------------------------------------
// Nets (degree >= 2 only) in the backend's internal order: n_small nets of degree 2..kSmallDegree
// grouped by degree, then n_medium of degree ..kMediumDegree, then n_large; pins follow the same
// order (CSR net_off), except inside a small-net degree group d, nets [small_net[d],
// small_net[d + 1]), which stores pin j of its i-th net at small_pin[d] + j * size + i.
constexpr int kSmallDegree = 8;
constexpr int kMediumDegree = 32;
struct WireArgs {
    const double2* pos[kMaxEval];
    const int* net_off;
    const double* net_w;
    const int* pin_part;
    const double2* pin_off;
    int n_small, n_medium, n_large;
    int small_net[kSmallDegree + 2];  // [d]: first net of degree d (d = 2..kSmallDegree), [kSmallDegree + 1] = n_small
    int small_pin[kSmallDegree + 2];  // [d]: first pin of that group
    const int* pin_dest;  // [pin] its slot in pin_grad, -1 for pins of fixed particles (null: pin order)
    double gamma;
    int want_grad;
    float2* pin_grad[kMaxEval];  // per-pin d(WL)/d(pin), at the slots pin_dest gives
    double* partials[kMaxEval];
    // Evaluated ahead (one evaluation): the smoothing is not `gamma` but the controller's law
    // (AdaptiveGlobalPlacer.cpp:456-462, exponent 1) at the optimizer overflow the device holds at
    // *ahead_overflow: (final + span * clamp(overflow / denominator, 0, 1)) * scale; workgroup 0
    // writes it after the partial rows, at partials[0][2 * ahead_rows].
    const double* ahead_overflow;
    double ahead_denominator, ahead_final, ahead_span, ahead_scale;
    int ahead_rows;
    // With ahead_metrics set, every workgroup forms that overflow itself from the metrics
    // reduction's partial rows (ahead_metric_rows of them), with k_finalize's sums in k_finalize's
    // order, instead of reading *ahead_overflow: the finalize then runs in the same launch.
    const double* ahead_metrics;
    int ahead_metric_rows;
};
int wireBlocks(int n_small, int n_medium, int n_large);
void launchWire(const WireArgs& a, int count, hipStream_t s);
// launchFinalize(l) and launchWire(w, 1) as one launch (k_finalize_wire): the finalize's
// workgroups first, then the wirelength evaluated ahead (w.ahead_metrics set)
void launchFinalizeWire(const RedList& l, const WireArgs& w, hipStream_t s);

// ---- density: row order, fixed-point deposit, bin metrics ---------------------------------------
constexpr int kSortThreads = 1024;    // particles per histogram / placement workgroup
constexpr int kDepositThreads = 1024;  // particles per deposit workgroup
constexpr int kWindowCells = 4096;     // LDS window of a deposit workgroup, cells per layer (64 KB)
// the same for the batch driver's jobs (small grids: more, smaller workgroups per job)
constexpr int kBatchDepositThreads = 512;
constexpr int kBatchWindowCells = 2048;
constexpr int kPermLayerShift = 30;   // row-order entries pack the particle layer above the id
constexpr unsigned kPermIdMask = (1U << kPermLayerShift) - 1U;
struct DensityArgs {
    const double2* pos[kMaxEval];  // the order is built from pos[0]
    const double2* size;   // (w, h)
    const int* layer;      // 0 std, 1 macro, 2 filler
    const unsigned char* tall;  // 1: taller than two bin rows (not row-sorted)
    const int* tall_ids;
    int n_tall;
    int n;
    Region r;
    double t;
    double fix, inv_fix;   // fixed-point scale: 2^40 per bin area
    int* key;              // [P]
    int* hist;             // [histogram blocks][ny]
    int* colprefix;        // [histogram blocks][ny]
    int* keytotal;         // [ny]
    int* keystart;         // [ny + 1]
    unsigned* perm;        // [P] particle id | layer << kPermLayerShift, in row order
    double2* sorted_size;  // [P] particle sizes in row order
    unsigned long long* grid[kMaxEval];  // [2][G] fixed point: movable + t*macro, filler; kept zero between evaluations
    const double* static_charge;   // [G] t*(fixed + dark)
    float* rho[kMaxEval];  // [nyp*nxp] density deviation, FP32, padded
    int ldrho;
    double* partials[kMaxEval];  // one metrics row per bin block
};
int densityHistBlocks(int n);
int densityBinBlocks(const Region& r);
void launchDensitySort(const DensityArgs& a, hipStream_t s);
// Deposit and bin metrics of `count` evaluations (partials: densityBinBlocks rows each).
void launchDensity(const DensityArgs& a, int count, hipStream_t s);  // launchDeposit, then the bin pass
void launchDeposit(const DensityArgs& a, int count, hipStream_t s);
// the bin pass of launchDensity(d, nd), after launchDeposit, and launchWire(w, nw) as one launch
// (k_wire_dbins); the same results
void launchWireBins(const WireArgs& w, int nw, const DensityArgs& d, int nd, hipStream_t s);

// ---- DCT products on MFMA (FP32) ---------------------------------------------------------------
// kEpiInterleave stores C[i][j] at C + 2 * (i * ldc + j) (two products into one interleaved grid)
enum Epilogue : int { kEpiStore = 0, kEpiSpectral = 1, kEpiInterleave = 2 };
constexpr int kMaxGemmBatch = 2 * kMaxEval;
struct GemmBatch {
    const float* A[kMaxGemmBatch];
    const float* B[kMaxGemmBatch];
    float* C[kMaxGemmBatch];
};
struct SpectralArgs {
    int nx, ny;
    double W, H;          // region width / height
    double normalization; // 1 / (4 nx ny)
    double* partials[kMaxGemmBatch];  // energy partials of the spectral epilogue, per product
};
// C[MxN] = A[MxK] * B[KxN], all row-major, M, N multiples of 32, K multiple of 16.
void launchGemm(const GemmBatch& g, int batch, int M, int N, int K, int lda, int ldb, int ldc, int epilogue,
                const SpectralArgs& sp, int precision, hipStream_t s);

// Half-length transforms.  Every factor of the path has the mirror symmetry
// F[n-1-i][k] = s_k F[i][k] (s_k = (-1)^k for the cosines, (-1)^(k+1) for the sines), so a
// length-n transform splits into two n/2 x n/2 products, one per parity of k, over the folded
// input x[i] +- x[n-1-i] (forward) or producing x[i] = E + O, x[n-1-i] = sigma (E - O)
// (inverse); half the multiply-adds of the dense product.  Operands (h = n / 2, row-major):
//   kFoldCols   out[m][2c+p] = sum_t (A[m][t] + (-1)^p A[m][n-1-t]) B_p[t][c]
//               A: M x n data, B: factor halves [2][h][h], out: M x n
//   kFoldRows   out[2r+p][k] = sum_t A_p[r][t] (B[t][k] + (-1)^p B[n-1-t][k]), spectral epilogue
//               A: factor halves [2][h][h], B: n x N data, out: n x N
//   kMirrorCols E/O[m][c] = sum_t A[m][2t+p] B_p[t][c]; columns c < h are the cosine half
//               (out[m][c], out[m][n-1-c], sigma +1), c >= h the sine half (written at column
//               offset n, sigma -1); A: M x n data, B: factor halves [2][h][2h], out: M x 2n
//   kMirrorRows E/O[r][k] = sum_t A_p[r][t] B[2t+p][k]; out rows r and n-1-r, sigma[z],
//               stored at out + step (row ldc + k) (step 2: one of two interleaved grids);
//               A: factor halves [2][h][h], B: n x N
enum DctMode : int { kFoldCols = 0, kFoldRows = 1, kMirrorCols = 2, kMirrorRows = 3 };
// Tile shapes of the half-length products: a workgroup holds w tiles of tr x tc half-size outputs
// (both parities) side by side, each tile's contraction split over s wavefronts; f = 1 rotates two
// operand buffers so that a chunk's loads fly while the previous chunk's products run (f = 0 loads
// two chunks, then runs their products).  Results depend on s only (the order of each element's
// sum), never on tr, tc, w or f.  Tile 0 is valid for every folded grid and is the default where the
// tuned table (dct_tiles.inc, written by myplacement_amd/autotune.py on an MI355X) names none.
struct DctTile {
    int tr, tc, w, s, f;
};
constexpr int kDctTileCount = 16;
constexpr DctTile kDctTiles[kDctTileCount] = {
    {32, 32, 1, 8, 0}, {32, 32, 1, 4, 0}, {16, 32, 1, 4, 0}, {32, 32, 1, 2, 0}, {32, 32, 1, 1, 0}, {32, 32, 2, 1, 0},
    {16, 32, 2, 1, 0}, {16, 16, 4, 1, 0}, {32, 32, 1, 8, 1}, {32, 32, 1, 4, 1}, {16, 32, 1, 4, 1}, {32, 32, 1, 2, 1},
    {32, 32, 1, 1, 1}, {32, 32, 2, 1, 1}, {16, 32, 2, 1, 1}, {16, 16, 4, 1, 1}};
bool dctTileValid(int tile, int mode, int rows, int cols);
int dctTileFor(int mode, int precision, int rows, int cols);  // the tuned tile of a product
int dctWorkgroups(int tile, int mode, int rows, int cols);    // per product (energy partial rows)
struct DctBatch {
    const float* A[kMaxGemmBatch];
    const float* B[kMaxGemmBatch];
    float* C[kMaxGemmBatch];
    float sigma[kMaxGemmBatch];  // kMirrorRows: sign of the mirrored half per product
};
// True when an n_x x n_y grid takes the half-length products (both sides multiples of 64).
bool foldedDct(int nx, int ny);
// `rows` x `cols` of the output grid (for kMirrorCols the transform side, cols = n); precision 0:
// FP32 products, 1: split-BF16 (bf16x3) products with FP32 accumulation.
void launchDct(int mode, const DctBatch& d, int batch, int rows, int cols, int lda, int ldb, int ldc, int step,
               int precision, const SpectralArgs& sp, hipStream_t s);
// launchDct with a given tile (the autotuner's timing runs)
void launchDctTile(int tile, int mode, const DctBatch& d, int batch, int rows, int cols, int lda, int ldb, int ldc,
                   int step, int precision, const SpectralArgs& sp, hipStream_t s);

// ---- field gather + gradient assembly ---------------------------------------------------------
struct GatherArgs {
    const double2* pos[kMaxEval];
    const double* charge;
    const float2* field[kMaxEval];  // interleaved (Ex, Ey), row stride ldf
    int ldf;
    Region r;
    const int* ppin_off;   // [P+1] particle -> pin CSR
    const int* ppin_idx;   // pin_grad slot of each CSR entry
    const float2* pin_grad[kMaxEval];
    int n;
    // backend mode (one evaluation): write the two gradients
    double2* wire_out;
    double2* density_out;
    // session mode: raw = wire + penalty * density, pre = raw * base / strength
    const double* precond;
    double penalty, inv_strength;
    double2* raw_out[kMaxEval];
    double2* pre_out[kMaxEval];
    // controller sums fused into the same pass (null pointers skip them)
    const double2* prev_pos[kMaxEval];  // momentum alignment with V = pos - prev_pos
    const double2* px;        // curvature: |pos - px|^2 and |pre - pg|^2
    const double2* pg;
    double2* ck_out[kMaxEval];          // checkpoint copy of the positions
    double* partials[kMaxEval];
};
int particleBlocks(int n);
constexpr int kGatherParticles = 256;  // particles per k_gather workgroup (one per thread)
// Regions of the per-pin gradient terms: region k holds the terms of the particles of the gather
// workgroups w = k mod kPinRegions (the backend's set-up assigns the slots, pin_dest).
constexpr int kPinRegions = 4;
void launchGather(const GatherArgs& a, int count, hipStream_t s);

// ---- controller vector passes (AdaptiveGlobalPlacer.cpp O(P) loops) ---------------------------
struct ClampArgs {
    const double2* size;
    Region r;
};
// out = clamp(x + beta * (x - x_prev))
// out = clamp(x + beta (x - x_prev))
struct LookaheadArgs {
    const double2* x;
    const double2* x_prev;
    double beta;
    double2* out;
    ClampArgs c;
    int n;
};
void launchLookahead(const LookaheadArgs& l, hipStream_t s);
// The first step of an iteration, decided on the device from the scalars of the current and
// look-ahead evaluations exactly as the host law decides it (AdaptiveGlobalPlacer.cpp:796-820 and
// estimateStep :470-492), by the propose kernel itself, so that the proposal and its evaluation
// follow without a read-back.  The decision (1 for the look-ahead base, step, displacement cap)
// is also stored to `mirror` for the host's check.
struct DecideArgs {
    const double* cur;   // device-memory copies of the current and look-ahead scalar banks
    const double* look;
    double penalty, growth_threshold, last_step, initial_overflow, move_cap;
    int try_momentum, restarted, has_previous;
    double* mirror;      // scalar-bank slots kDecidedLook.. (host-mapped)
};
// y = clamp(b + cap(-step * g, max_disp)); partials of sum raw.(b - y) and max |b - y|; optional
// staged copies b_copy = b, g_copy = g.  With decide_here set, the base (b or b_look, with its
// gradients), the step and the cap come from the decision taken from `decide` instead.
struct ProposeArgs {
    const double2* b;
    const double2* g;
    const double2* raw;
    double step, max_disp;
    double2* y;
    double2* b_copy;
    double2* g_copy;
    ClampArgs c;
    int n;
    double* partials;
    int decide_here = 0;
    DecideArgs decide{};
    const double2* b_look = nullptr;
    const double2* g_look = nullptr;
    const double2* raw_look = nullptr;
};
void launchPropose(const ProposeArgs& p, hipStream_t s);
// ---- launches batched over independent jobs -----------------------------------------------------
// One launch runs a stage for many jobs.  Entry e (an array in device memory) holds the arguments
// of one job's single launch and its grid extent gx x gy x gz; it owns the flattened workgroups
// [first[e], first[e + 1]), first[0] = 0.
struct WireEntry {
    WireArgs a;
    int large_blocks, medium_blocks, gx, gy;
};
struct SortEntry {  // histogram and placement: gx = histogram blocks; scan: gx = bin rows
    DensityArgs a;
    int rows;       // histogram blocks (the scan's row count)
    int gx;
};
struct DensityEntry {  // deposit (sorted_blocks set) and bin metrics
    DensityArgs a;
    int sorted_blocks, gx, gy;
};
struct GemmEntry {
    GemmBatch g;
    SpectralArgs sp;
    int K, lda, ldb, ldc, epi, gx, gy, gz;
};
struct DctEntry {
    DctBatch d;
    SpectralArgs sp;
    int mode, kpl, n, lda, ldb, ldc, step, gx, gy, gz, tile;
};
struct GatherEntry {
    GatherArgs a;
    int gx, gy;
};
struct LookaheadEntry {
    LookaheadArgs l;
    int gx;
};
struct ProposeEntry {
    ProposeArgs p;
    int gx;
};
struct CopyEntry {
    double2* dst;
    const double2* src;
    int n, gx;
};
struct ZeroEntry {
    double* p;
    int n, gx;
};
enum BatchKernel : int {
    kBatchWire = 0, kBatchSortHist, kBatchSortScan, kBatchSortPlace, kBatchDeposit, kBatchBins, kBatchGemmStore,
    kBatchGemmSpectral, kBatchGemmInterleave, kBatchGather, kBatchLookahead, kBatchPropose, kBatchCopy, kBatchZero,
    kBatchFinalize,  // entries: RedJob (one workgroup each, no completion word)
    kBatchDct  // + dctVariant(entry): (tile * 4 + mode) * 2 + (8 values per lane ? 1 : 0)
};
// Entry builders (the grid extent of the job's own launch; they validate as the launches do).
WireEntry wireEntry(const WireArgs& a, int count);
SortEntry sortEntry(const DensityArgs& a, int stage);  // 0 histogram, 1 scan, 2 placement
DensityEntry depositEntry(const DensityArgs& a, int count);
DensityEntry binsEntry(const DensityArgs& a, int count);
GemmEntry gemmEntry(const GemmBatch& g, int batch, int M, int N, int K, int lda, int ldb, int ldc, int epilogue,
                    const SpectralArgs& sp);
DctEntry dctEntry(int mode, const DctBatch& d, int batch, int rows, int cols, int lda, int ldb, int ldc, int step,
                  int precision, const SpectralArgs& sp);
int dctVariant(const DctEntry& e);
GatherEntry gatherEntry(const GatherArgs& a, int count);
LookaheadEntry lookaheadEntry(const LookaheadArgs& l);
ProposeEntry proposeEntry(const ProposeArgs& p);
CopyEntry copyEntry(double2* dst, const double2* src, int n);
ZeroEntry zeroEntry(double* p, int n);
// `entries` and `first` in device memory; total_blocks = first[n].
void launchBatched(BatchKernel k, const void* entries, const int* first, int n, int total_blocks, std::size_t dyn_lds,
                   hipStream_t s);
// The job's own (unbatched) launch described by an entry held in host memory.
void launchEntry(BatchKernel k, const void* entry, std::size_t dyn_lds, hipStream_t s);
// k_finalize over n reduction jobs; flags[b] receives seq after job b's results.
void launchFinalizeBatch(const RedJob* jobs, int n, volatile unsigned* flags, unsigned seq, hipStream_t s);

// Loads the library's device code on the stream's device (a service does this before its first job).
void loadDeviceCode(hipStream_t s);
// fill with zero (used instead of memset so the sequence can be captured)
void launchZero(double* p, size_t count, hipStream_t s);

}  // namespace mpgp
