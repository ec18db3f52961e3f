// gp_detailed.hip — detailed placement by four-cell window reordering on the GPU.
//
// The HIP counterpart of the reference's CUDA detailed-placement pass
// (src/placement/cuda/CudaDetailedPlacementBackend.cu:523-711, kernel :144-220), behind the same
// boundary (tryRunCudaDetailedPlacement, CudaDetailedPlacementBackend.hpp:14-17).  Each pass:
//   1. row segments of abutting standard cells at the current placement (buildRowSegments, :230-275);
//   2. windows of four consecutive cells from phase (pass % 2) * 2 in steps of four; a window
//      touching a net of more than maximum_net_degree pins, or shorter than four cells, goes to
//      the CPU list (makeWindow, :283-299);
//   3. one launch scores all 24 orderings of every device window against the placement at the
//      start of the pass and returns the best (lowest HPWL of the window's nets, lowest index on
//      ties);
//   4. on the host, in window order, each suggested ordering is re-scored on the current
//      placement (earlier windows of the pass may have moved shared nets) and applied when it
//      improves the window's HPWL by more than improvement_epsilon (:655-680);
//   5. the CPU windows are searched exhaustively in order (optimizeCpuWindow, :400-424).
// Orderings follow std::next_permutation from the identity; ordering p is decoded from its index
// in the factorial number system, which is that same lexicographic order.
//
// On MI355X a window takes half a wavefront (24 of 32 lanes score an ordering each), so one
// wavefront scores two windows and a 256-thread workgroup eight; the argmin is a 32-lane shuffle
// butterfly over (value, ordering) pairs, with no LDS and no barrier.  Every lane of a window reads
// the same net and pin entries, which the wavefront's loads broadcast.  The score of an ordering is
// computed with the same FP64 operations in the same order as the host's re-scoring
// (contraction off), so device and host agree bit for bit.

#include "mp_gp_backend.h"
#include "gp_host.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <climits>
#include <cmath>
#include <memory>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

namespace mpgp {

extern thread_local std::string g_last_error;
mp_gp_status translate(const std::exception_ptr& e);
std::size_t allocationLimit(int device, std::uint64_t maximum_bytes, std::string& reason);

namespace {

constexpr int kDpWindow = 4;
constexpr int kDpOrderings = 24;
constexpr int kDpLanes = 32;
constexpr int kDpThreads = 256;
constexpr int kDpWindowsPerBlock = kDpThreads / kDpLanes;
constexpr double kGeomEps = 1e-9;     // Geometry.hpp kEpsilon
constexpr double kSegmentEps = 1e-6;  // buildRowSegments tolerance of the pass

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Width rounded up to the row's site grid (alignUp(width, 0, spacing), DetailedPlacer.cpp:73-76).
__host__ __device__ inline double roundedWidth(double width, double spacing) {
    return spacing <= kGeomEps ? width : std::ceil((width - kGeomEps) / spacing) * spacing;
}

// Ordering `index` (0..23) of four items in lexicographic order: digits of the factorial number
// system pick from the items not yet used.
__host__ __device__ constexpr void decodeOrdering(int index, int (&order)[kDpWindow]) {
    int avail[kDpWindow] = {0, 1, 2, 3};
    int left = kDpWindow;
    const int radix[kDpWindow - 1] = {6, 2, 1};
    for (int k = 0; k < kDpWindow - 1; ++k) {
        const int d = index / radix[k];
        index -= d * radix[k];
        order[k] = avail[d];
        for (int j = d; j + 1 < left; ++j) avail[j] = avail[j + 1];
        --left;
    }
    order[kDpWindow - 1] = avail[0];
}

struct OrderingTable {
    int v[kDpOrderings][kDpWindow];
};
constexpr OrderingTable makeOrderings() {
    OrderingTable t{};
    for (int p = 0; p < kDpOrderings; ++p) decodeOrdering(p, t.v[p]);
    return t;
}
__constant__ OrderingTable kOrderings = makeOrderings();
static_assert(makeOrderings().v[1][2] == 3 && makeOrderings().v[23][0] == 3 && makeOrderings().v[9][3] == 0,
              "orderings follow std::next_permutation from the identity");

struct DpArgs {
    const double2* centers;   // [M] at the start of the pass
    const double* width;      // [M]
    const double* height;     // [M]
    const int* pin_module;    // [I] pins in net order
    const double2* pin_off;   // [I] offsets transformed by the module orientation at the start of the pass
    const int* net_start;     // [N+1]
    const double* net_w;      // [N]
    const int* win_modules;   // [4W]
    const double* win_x;      // [W] lower-left x of the window's first cell
    const double* win_bottom; // [W]
    const double* win_spacing;// [W]
    const int* win_net_start; // [W+1]
    const int* win_nets;      // ascending net ids per window
    int windows;
    int* best;                // [W] best ordering index
};

__global__ __launch_bounds__(kDpThreads) void k_dp_windows(DpArgs a) {
#pragma clang fp contract(off)
    const int window = blockIdx.x * kDpWindowsPerBlock + static_cast<int>(threadIdx.x) / kDpLanes;
    const int lane = threadIdx.x & (kDpLanes - 1);
    if (window >= a.windows) return;  // uniform over the window's 32 lanes
    const int ordering = lane < kDpOrderings ? lane : 0;
    int mod[kDpWindow];
#pragma unroll
    for (int k = 0; k < kDpWindow; ++k) mod[k] = a.win_modules[kDpWindow * window + k];
    // candidate centres of this ordering (centersForOrder, .cu:363-374)
    double wx[kDpWindow], wy[kDpWindow];
    double cursor = a.win_x[window];
    const double bottom = a.win_bottom[window], spacing = a.win_spacing[window];
#pragma unroll
    for (int k = 0; k < kDpWindow; ++k) {
        const int local = kOrderings.v[ordering][k];
        int m = mod[0];
#pragma unroll
        for (int j = 1; j < kDpWindow; ++j) m = local == j ? mod[j] : m;
        const double w = a.width[m], h = a.height[m];
        const double x = cursor + w * 0.5, y = bottom + h * 0.5;
#pragma unroll
        for (int j = 0; j < kDpWindow; ++j) {
            if (local == j) {
                wx[j] = x;
                wy[j] = y;
            }
        }
        cursor += roundedWidth(w, spacing);
    }
    // HPWL of the window's nets under this ordering (localHpwl, .cu:337-361)
    double value = 0.0;
    for (int e = a.win_net_start[window]; e < a.win_net_start[window + 1]; ++e) {
        const int net = a.win_nets[e];
        double lox = DBL_MAX, hix = -DBL_MAX, loy = DBL_MAX, hiy = -DBL_MAX;
        for (int p = a.net_start[net]; p < a.net_start[net + 1]; ++p) {
            const int m = a.pin_module[p];
            const double2 c = a.centers[m];
            double x = c.x, y = c.y;
#pragma unroll
            for (int j = 0; j < kDpWindow; ++j) {
                if (mod[j] == m) {
                    x = wx[j];
                    y = wy[j];
                }
            }
            const double2 o = a.pin_off[p];
            x += o.x;
            y += o.y;
            lox = fmin(lox, x);
            hix = fmax(hix, x);
            loy = fmin(loy, y);
            hiy = fmax(hiy, y);
        }
        value += a.net_w[net] * ((hix - lox) + (hiy - loy));
    }
    if (lane >= kDpOrderings) value = DBL_MAX;
    // lexicographic minimum of (value, ordering) over the 32 lanes
    int idx = lane;
#pragma unroll
    for (int o = kDpLanes / 2; o > 0; o >>= 1) {
        const double v2 = __shfl_xor(value, o, kDpLanes);
        const int i2 = __shfl_xor(idx, o, kDpLanes);
        if (v2 < value || (v2 == value && i2 < idx)) {
            value = v2;
            idx = i2;
        }
    }
    if (lane == 0) a.best[window] = idx;
}

// Device memory carved from a caller-provided block (the no-allocation entry point's scratch):
// a bump pointer over [base, base + size), 256-byte aligned pieces; mark / release rewind it.
class Arena {
public:
    Arena(unsigned char* base, std::size_t size) : base_(base), size_(size) {}
    void* take(std::size_t bytes, const char* what) {
        const std::size_t at = (used_ + 255) / 256 * 256;
        if (at > size_ || bytes > size_ - at)
            throw std::logic_error(std::string("detailed-placement scratch overrun while carving ") + what);
        used_ = at + bytes;
        peak_ = std::max(peak_, used_);
        return base_ + at;
    }
    std::size_t mark() const { return used_; }
    void release(std::size_t mark) { used_ = mark; }
    std::size_t peak() const { return peak_; }

private:
    unsigned char* base_;
    std::size_t size_, used_ = 0, peak_ = 0;
};

template <typename T>
class DeviceVec {
public:
    DeviceVec() = default;
    ~DeviceVec() {
        if (p_ && owned_) (void)hipFree(p_);
    }
    DeviceVec(const DeviceVec&) = delete;
    DeviceVec& operator=(const DeviceVec&) = delete;
    // (re)allocates for at least n elements, charging the budget with the growth; with an arena,
    // carves n elements from it instead (the arena's owner rewinds it between passes)
    void reserve(std::size_t n, std::size_t& used, std::size_t limit, const char* what, Arena* arena = nullptr) {
        if (n == 0) n = 1;
        if (arena) {
            p_ = static_cast<T*>(arena->take(n * sizeof(T), what));
            owned_ = false;
            n_ = n;
            return;
        }
        if (n <= n_) return;
        const std::size_t grow = (n - n_) * sizeof(T);
        if (grow > limit - used)
            throw std::runtime_error(std::string("HIP detailed-placement budget exceeded while reserving ") + what + ".");
        if (p_) check(hipFree(p_), what);
        p_ = nullptr;
        check(hipMalloc(reinterpret_cast<void**>(&p_), n * sizeof(T)), what);
        owned_ = true;
        used += grow;
        n_ = n;
    }
    void upload(const std::vector<T>& v, hipStream_t s, const char* what) {
        if (v.size() > n_) throw std::logic_error("device upload exceeds buffer");
        if (!v.empty()) check(hipMemcpyAsync(p_, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s), what);
    }
    T* get() const { return p_; }

private:
    T* p_ = nullptr;
    std::size_t n_ = 0;
    bool owned_ = false;
};

// The placement the pass works on: the PlacementDatabase fields window reordering reads and writes.
struct Placement {
    int64_t M = 0, N = 0, I = 0;
    std::vector<double> w, h, cx, cy;
    std::vector<int32_t> orientation;
    std::vector<uint8_t> standard;  // !is_fixed && !is_macro after refreshDerivedData
    std::vector<double> net_w;
    std::vector<int64_t> net_start;
    std::vector<int32_t> pin_module;
    std::vector<double> raw_x, raw_y;  // un-oriented pin offsets
    std::vector<int64_t> mod_pin_start, mod_pins, pin_net;
    struct Row {
        double bottom, height, x_start, spacing;
        int32_t orientation;
    };
    std::vector<Row> rows;
    std::vector<std::size_t> sorted_rows;

    double pinX(int64_t p, double cxm, double cym, double* y) const {
        const int64_t m = pin_module[p];
        double ox = raw_x[p], oy = raw_y[p];
        orient(ox, oy, orientation[m]);
        *y = cym + oy;
        return cxm + ox;
    }
    double lowerLeftX(int64_t m) const { return cx[m] - w[m] * 0.5; }
    double lowerLeftY(int64_t m) const { return cy[m] - h[m] * 0.5; }

    // calculateHpwl (Metrics.cpp:128-147)
    double hpwl() const {
        double total = 0.0;
        for (int64_t n = 0; n < N; ++n) {
            if (net_start[n + 1] - net_start[n] < 2) continue;
            double lox = INFINITY, hix = -INFINITY, loy = INFINITY, hiy = -INFINITY;
            for (int64_t p = net_start[n]; p < net_start[n + 1]; ++p) {
                double y;
                const double x = pinX(p, cx[pin_module[p]], cy[pin_module[p]], &y);
                lox = std::min(lox, x);
                hix = std::max(hix, x);
                loy = std::min(loy, y);
                hiy = std::max(hiy, y);
            }
            total += net_w[n] * ((hix - lox) + (hiy - loy));
        }
        return total;
    }

    // nearest row by row centre among rows sorted by bottom (nearestRow, DetailedPlacer.cpp:89-104)
    std::size_t nearestRow(double coordinate) const {
        const auto it = std::lower_bound(sorted_rows.begin(), sorted_rows.end(), coordinate,
                                         [&](std::size_t r, double v) { return rows[r].bottom + rows[r].height * 0.5 < v; });
        if (it == sorted_rows.begin()) return *it;
        if (it == sorted_rows.end()) return sorted_rows.back();
        const std::size_t right = *it, left = *(it - 1);
        const double dl = std::abs(rows[left].bottom + rows[left].height * 0.5 - coordinate);
        const double dr = std::abs(rows[right].bottom + rows[right].height * 0.5 - coordinate);
        return dl <= dr ? left : right;
    }
};

struct Segment {
    std::size_t row = 0;
    std::vector<int64_t> modules;
};

// buildRowSegments (.cu:230-275): standard cells on a row bottom, split where they do not abut
std::vector<Segment> rowSegments(const Placement& P, double eps) {
    std::vector<Segment> out;
    if (P.rows.empty()) return out;
    std::vector<std::vector<int64_t>> by_row(P.rows.size());
    for (int64_t m = 0; m < P.M; ++m) {
        if (!P.standard[m]) continue;
        const std::size_t r = P.nearestRow(P.cy[m]);
        if (std::abs(P.lowerLeftY(m) - P.rows[r].bottom) <= eps) by_row[r].push_back(m);
    }
    for (std::size_t r = 0; r < by_row.size(); ++r) {
        std::vector<int64_t>& cells = by_row[r];
        if (cells.size() < 2) continue;
        const double spacing = P.rows[r].spacing;
        std::sort(cells.begin(), cells.end(), [&](int64_t a, int64_t b) { return P.lowerLeftX(a) < P.lowerLeftX(b); });
        Segment s;
        s.row = r;
        for (const int64_t m : cells) {
            if (!s.modules.empty()) {
                const int64_t prev = s.modules.back();
                const double expected = P.lowerLeftX(prev) + roundedWidth(P.w[prev], spacing);
                if (std::abs(P.lowerLeftX(m) - expected) > eps) {
                    if (s.modules.size() >= 2) out.push_back(std::move(s));
                    s = Segment{};
                    s.row = r;
                }
            }
            s.modules.push_back(m);
        }
        if (s.modules.size() >= 2) out.push_back(std::move(s));
    }
    return out;
}

struct Window {
    std::size_t row = 0;
    std::vector<int64_t> modules;
    std::vector<int64_t> nets;  // ascending, unique
    bool device = false;
};

// makeWindow (.cu:283-299)
Window makeWindow(const Placement& P, std::size_t row, std::vector<int64_t> modules, uint64_t max_degree) {
    Window w;
    w.row = row;
    w.modules = std::move(modules);
    for (const int64_t m : w.modules)
        for (int64_t k = P.mod_pin_start[m]; k < P.mod_pin_start[m + 1]; ++k) w.nets.push_back(P.pin_net[P.mod_pins[k]]);
    std::sort(w.nets.begin(), w.nets.end());
    w.nets.erase(std::unique(w.nets.begin(), w.nets.end()), w.nets.end());
    w.device = w.modules.size() == static_cast<std::size_t>(kDpWindow) && !w.nets.empty();
    for (const int64_t n : w.nets)
        if (static_cast<uint64_t>(P.net_start[n + 1] - P.net_start[n]) > max_degree) w.device = false;
    return w;
}

// HPWL of the window's nets with the window's cells at `wx`, `wy` (localHpwl, .cu:337-361)
double windowHpwl(const Placement& P, const Window& win, const std::vector<double>& wx, const std::vector<double>& wy) {
    double total = 0.0;
    for (const int64_t n : win.nets) {
        double lox = INFINITY, hix = -INFINITY, loy = INFINITY, hiy = -INFINITY;
        for (int64_t p = P.net_start[n]; p < P.net_start[n + 1]; ++p) {
            const int64_t m = P.pin_module[p];
            double cxm = P.cx[m], cym = P.cy[m];
            for (std::size_t k = 0; k < win.modules.size(); ++k) {
                if (win.modules[k] == m) {
                    cxm = wx[k];
                    cym = wy[k];
                    break;
                }
            }
            double y;
            const double x = P.pinX(p, cxm, cym, &y);
            lox = std::min(lox, x);
            hix = std::max(hix, x);
            loy = std::min(loy, y);
            hiy = std::max(hiy, y);
        }
        total += P.net_w[n] * ((hix - lox) + (hiy - loy));
    }
    return total;
}

// centersForOrder (.cu:363-374)
void orderedCenters(const Placement& P, const Window& win, const std::vector<std::size_t>& order, std::vector<double>& wx,
                    std::vector<double>& wy) {
    const auto& row = P.rows[win.row];
    wx.assign(win.modules.size(), 0.0);
    wy.assign(win.modules.size(), 0.0);
    double cursor = P.lowerLeftX(win.modules.front());
    for (const std::size_t local : order) {
        const int64_t m = win.modules[local];
        wx[local] = cursor + P.w[m] * 0.5;
        wy[local] = row.bottom + P.h[m] * 0.5;
        cursor += roundedWidth(P.w[m], row.spacing);
    }
}

// applyOrder (.cu:376-386): Module::setLowerLeft, then the row's orientation unless it is Unknown
void applyOrdering(Placement& P, const Window& win, const std::vector<std::size_t>& order) {
    const auto& row = P.rows[win.row];
    double cursor = P.lowerLeftX(win.modules.front());
    for (const std::size_t local : order) {
        const int64_t m = win.modules[local];
        P.cx[m] = cursor + P.w[m] * 0.5;
        P.cy[m] = row.bottom + P.h[m] * 0.5;
        if (row.orientation != 8) P.orientation[m] = row.orientation;
        cursor += roundedWidth(P.w[m], row.spacing);
    }
}

struct CpuWindowResult {
    bool improved = false;
    std::vector<std::size_t> order;
    std::size_t orderings = 0;
};

// optimizeCpuWindow (.cu:400-424): every ordering, the first one better than the best so far by
// more than epsilon is kept
CpuWindowResult searchWindow(const Placement& P, const Window& win, const mp_dp_options& o) {
    CpuWindowResult r;
    if (win.nets.empty()) return r;
    for (const int64_t n : win.nets)
        if (static_cast<uint64_t>(P.net_start[n + 1] - P.net_start[n]) > o.maximum_net_degree) return r;
    std::vector<double> wx(win.modules.size()), wy(win.modules.size());
    for (std::size_t k = 0; k < win.modules.size(); ++k) {
        wx[k] = P.cx[win.modules[k]];
        wy[k] = P.cy[win.modules[k]];
    }
    double best = windowHpwl(P, win, wx, wy);
    std::vector<std::size_t> order(win.modules.size());
    std::iota(order.begin(), order.end(), std::size_t{0});
    do {
        orderedCenters(P, win, order, wx, wy);
        const double value = windowHpwl(P, win, wx, wy);
        ++r.orderings;
        if (value < best - o.improvement_epsilon) {
            best = value;
            r.order = order;
            r.improved = true;
        }
    } while (std::next_permutation(order.begin(), order.end()));
    return r;
}

void validate(const mp_dp_options& o) {  // validateDetailedPlacementOptions, DetailedPlacer.cpp:50-71
    if (o.method < MP_DP_NONE || o.method > MP_DP_WINDOW_REORDER || !std::isfinite(o.improvement_epsilon) ||
        o.improvement_epsilon < 0.0)
        throw std::invalid_argument("Detailed placement options are outside their valid range.");
    if (o.method == MP_DP_NONE) return;
    if (o.passes <= 0 || o.maximum_net_degree < 2)
        throw std::invalid_argument("Detailed placement options are outside their valid range.");
    const int window = o.method == MP_DP_ADJACENT_SWAP ? 2 : o.window_size;
    if (window < 2 || window > 6) throw std::invalid_argument("Detailed placement window size must be between 2 and 6.");
    if (o.maximum_device_bytes == 0) throw std::invalid_argument("Detailed placement needs a non-zero memory budget.");
}

struct Unavailable : std::runtime_error {
    using std::runtime_error::runtime_error;
};

Placement loadPlacement(const mp_design& d, const int32_t* row_orientation) {
    if (d.module_count < 0 || d.net_count < 0 || d.pin_count < 0 || d.row_count < 0)
        throw std::invalid_argument("negative counts in the design");
    if (d.module_count > INT_MAX || d.net_count >= INT_MAX || d.pin_count > INT_MAX)
        throw std::invalid_argument("the design is too large for 32-bit device indices");
    Netlist net(d);  // refreshDerivedData: terminals fixed, macro flags
    Placement P;
    P.M = d.module_count;
    P.N = d.net_count;
    P.I = d.pin_count;
    P.w.assign(d.width, d.width + P.M);
    P.h.assign(d.height, d.height + P.M);
    P.cx.resize(P.M);
    P.cy.resize(P.M);
    for (int64_t m = 0; m < P.M; ++m) {
        P.cx[m] = d.center[2 * m];
        P.cy[m] = d.center[2 * m + 1];
    }
    P.orientation.resize(P.M);
    for (int64_t m = 0; m < P.M; ++m)
        P.orientation[m] = d.orientation ? (d.orientation[m] < 0 || d.orientation[m] > 8 ? 8 : d.orientation[m]) : 0;
    P.standard.resize(P.M);
    for (int64_t m = 0; m < P.M; ++m) P.standard[m] = !net.fixed[m] && !net.macro[m];
    P.net_w.assign(d.net_weight, d.net_weight + P.N);
    P.net_start.assign(d.net_pin_start, d.net_pin_start + P.N + 1);
    if (P.N > 0 && (P.net_start[0] != 0 || P.net_start[P.N] != P.I))
        throw std::invalid_argument("net pin starts do not cover the pin array");
    P.pin_module.assign(d.pin_module, d.pin_module + P.I);
    P.raw_x.resize(P.I);
    P.raw_y.resize(P.I);
    P.pin_net.resize(P.I);
    for (int64_t n = 0; n < P.N; ++n) {
        if (P.net_start[n + 1] < P.net_start[n]) throw std::invalid_argument("net pin starts are not monotone");
        for (int64_t p = P.net_start[n]; p < P.net_start[n + 1]; ++p) P.pin_net[p] = n;
    }
    for (int64_t p = 0; p < P.I; ++p) {
        if (P.pin_module[p] < 0 || P.pin_module[p] >= P.M) throw std::invalid_argument("pin refers to an unknown module");
        P.raw_x[p] = d.pin_offset[2 * p];
        P.raw_y[p] = d.pin_offset[2 * p + 1];
    }
    // module -> pins in pin-id order (Module::pins as addPin appends them)
    P.mod_pin_start.assign(P.M + 1, 0);
    for (int64_t p = 0; p < P.I; ++p) ++P.mod_pin_start[P.pin_module[p] + 1];
    std::partial_sum(P.mod_pin_start.begin(), P.mod_pin_start.end(), P.mod_pin_start.begin());
    P.mod_pins.resize(P.I);
    {
        std::vector<int64_t> fill(P.mod_pin_start.begin(), P.mod_pin_start.end() - 1);
        for (int64_t p = 0; p < P.I; ++p) P.mod_pins[fill[P.pin_module[p]]++] = p;
    }
    P.rows.resize(d.row_count);
    for (int64_t r = 0; r < d.row_count; ++r) {
        const double* row = d.rows + 5 * r;
        P.rows[r] = {row[0], row[1], row[2], row[3], row_orientation ? row_orientation[r] : 0};
    }
    P.sorted_rows.resize(P.rows.size());
    std::iota(P.sorted_rows.begin(), P.sorted_rows.end(), std::size_t{0});
    std::sort(P.sorted_rows.begin(), P.sorted_rows.end(),
              [&](std::size_t a, std::size_t b) { return P.rows[a].bottom < P.rows[b].bottom; });
    return P;
}

// The device pass of one run: static netlist resident, per-pass placement and window lists.
class DeviceWindows {
public:
    // arena == nullptr: buffers from hipMalloc on a stream of its own; otherwise every buffer is
    // carved from the arena and the work runs on the null stream (no allocation of any kind)
    DeviceWindows(const Placement& P, int device, std::size_t limit, Arena* arena = nullptr)
        : limit_(limit), arena_(arena) {
        check(hipSetDevice(device), "select HIP device");
        if (!arena_) check(hipStreamCreate(&s_), "create detailed-placement stream");
        std::vector<int> net_start(P.N + 1);
        for (int64_t n = 0; n <= P.N; ++n) net_start[n] = static_cast<int>(P.net_start[n]);
        std::vector<int> pin_module(P.pin_module.begin(), P.pin_module.end());
        width_.reserve(P.M, used_, limit_, "module widths", arena_);
        height_.reserve(P.M, used_, limit_, "module heights", arena_);
        centers_.reserve(P.M, used_, limit_, "module centres", arena_);
        pin_module_.reserve(P.I, used_, limit_, "pin modules", arena_);
        pin_off_.reserve(P.I, used_, limit_, "pin offsets", arena_);
        net_start_.reserve(P.N + 1, used_, limit_, "net pin starts", arena_);
        net_w_.reserve(P.N, used_, limit_, "net weights", arena_);
        if (arena_) window_mark_ = arena_->mark();
        peak_ = arena_ ? arena_->peak() : used_;
        width_.upload(P.w, s_, "upload module widths");
        height_.upload(P.h, s_, "upload module heights");
        pin_module_.upload(pin_module, s_, "upload pin modules");
        net_start_.upload(net_start, s_, "upload net pin starts");
        net_w_.upload(P.net_w, s_, "upload net weights");
        check(hipStreamSynchronize(s_), "upload detailed-placement netlist");
    }
    ~DeviceWindows() {
        if (s_) {
            (void)hipStreamSynchronize(s_);
            (void)hipStreamDestroy(s_);
        } else if (arena_) {
            (void)hipStreamSynchronize(nullptr);
        }
    }
    DeviceWindows(const DeviceWindows&) = delete;
    DeviceWindows& operator=(const DeviceWindows&) = delete;

    // best ordering index of every window, scored at the placement P
    std::vector<int> score(const Placement& P, const std::vector<Window>& wins) {
        const std::size_t W = wins.size();
        if (W > static_cast<std::size_t>(INT_MAX / kDpWindow)) throw std::invalid_argument("too many windows");
        std::vector<double2> centers(P.M);
        for (int64_t m = 0; m < P.M; ++m) centers[m] = make_double2(P.cx[m], P.cy[m]);
        std::vector<double2> off(P.I);
        for (int64_t p = 0; p < P.I; ++p) {
            double ox = P.raw_x[p], oy = P.raw_y[p];
            orient(ox, oy, P.orientation[P.pin_module[p]]);
            off[p] = make_double2(ox, oy);
        }
        std::vector<int> mods(kDpWindow * W), net_start(W + 1, 0), nets;
        std::vector<double> x(W), bottom(W), spacing(W);
        for (std::size_t k = 0; k < W; ++k) {
            const Window& w = wins[k];
            for (int j = 0; j < kDpWindow; ++j) mods[kDpWindow * k + j] = static_cast<int>(w.modules[j]);
            x[k] = P.lowerLeftX(w.modules.front());
            bottom[k] = P.rows[w.row].bottom;
            spacing[k] = P.rows[w.row].spacing;
            for (const int64_t n : w.nets) nets.push_back(static_cast<int>(n));
            if (nets.size() > static_cast<std::size_t>(INT_MAX)) throw std::invalid_argument("too many window nets");
            net_start[k + 1] = static_cast<int>(nets.size());
        }
        if (arena_) arena_->release(window_mark_);  // the previous pass's window arrays are done
        win_mod_.reserve(mods.size(), used_, limit_, "window modules", arena_);
        win_x_.reserve(W, used_, limit_, "window starts", arena_);
        win_bottom_.reserve(W, used_, limit_, "window rows", arena_);
        win_spacing_.reserve(W, used_, limit_, "window site spacing", arena_);
        win_net_start_.reserve(W + 1, used_, limit_, "window net starts", arena_);
        win_nets_.reserve(nets.size(), used_, limit_, "window nets", arena_);
        best_.reserve(W, used_, limit_, "window orderings", arena_);
        peak_ = arena_ ? arena_->peak() : std::max(peak_, used_);
        centers_.upload(centers, s_, "upload module centres");
        pin_off_.upload(off, s_, "upload pin offsets");
        win_mod_.upload(mods, s_, "upload window modules");
        win_x_.upload(x, s_, "upload window starts");
        win_bottom_.upload(bottom, s_, "upload window rows");
        win_spacing_.upload(spacing, s_, "upload window site spacing");
        win_net_start_.upload(net_start, s_, "upload window net starts");
        win_nets_.upload(nets, s_, "upload window nets");
        DpArgs a{centers_.get(), width_.get(), height_.get(), pin_module_.get(), pin_off_.get(), net_start_.get(),
                 net_w_.get(), win_mod_.get(), win_x_.get(), win_bottom_.get(), win_spacing_.get(),
                 win_net_start_.get(), win_nets_.get(), static_cast<int>(W), best_.get()};
        const int blocks = static_cast<int>((W + kDpWindowsPerBlock - 1) / kDpWindowsPerBlock);
        hipLaunchKernelGGL(k_dp_windows, dim3(blocks), dim3(kDpThreads), 0, s_, a);
        check(hipGetLastError(), "launch window-ordering kernel");
        std::vector<int> best(W);
        check(hipMemcpyAsync(best.data(), best_.get(), W * sizeof(int), hipMemcpyDeviceToHost, s_), "download orderings");
        check(hipStreamSynchronize(s_), "score windows");
        return best;
    }
    std::size_t peakBytes() const { return peak_; }

private:
    std::size_t limit_ = 0, used_ = 0, peak_ = 0, window_mark_ = 0;
    Arena* arena_ = nullptr;
    hipStream_t s_ = nullptr;
    DeviceVec<double> width_, height_, net_w_, win_x_, win_bottom_, win_spacing_;
    DeviceVec<double2> centers_, pin_off_;
    DeviceVec<int> pin_module_, net_start_, win_mod_, win_net_start_, win_nets_, best_;
};

// Device bytes the no-allocation pass may carve, from the design's counts alone: the resident
// netlist (widths, heights, centres, pin modules and offsets, net starts and weights) and the
// window arrays of the largest possible pass (at most one window per two cells, and a window's
// nets are at most its cells' pins), each piece rounded up to 256 bytes.
std::size_t scratchBytesNeeded(int64_t M, int64_t N, int64_t I) {
    auto piece = [](std::size_t n, std::size_t elem) { return ((n == 0 ? 1 : n) * elem + 255) / 256 * 256; };
    const std::size_t m = static_cast<std::size_t>(M), n = static_cast<std::size_t>(N), i = static_cast<std::size_t>(I);
    const std::size_t w = m / 2 + 1;
    return piece(m, 8) * 2 + piece(m, 16) + piece(i, 4) + piece(i, 16) + piece(n + 1, 4) + piece(n, 8) +
           piece(kDpWindow * w, 4) + piece(w, 8) * 3 + piece(w + 1, 4) + piece(i, 4) + piece(w, 4);
}

// Scratch of the no-allocation entry point: one per GPU the library runs on, held until the
// process ends.  Host code never uses the symbol's name as an address: kScratch[device] holds the
// device address that hipGetSymbolAddress returned with that device current.
constexpr std::size_t kDpScratchBytes = 64ull << 20;
constexpr int kMaxDevices = 64;
__device__ __align__(256) unsigned char scratch[kDpScratchBytes];

struct DeviceScratch {
    std::once_flag once;
    unsigned char* ptr = nullptr;
    hipError_t status = hipSuccess;
    std::mutex lock;  // calls on one device take turns: they share its scratch
};
DeviceScratch g_scratch[kMaxDevices];

DeviceScratch& deviceScratch(int device) {
    DeviceScratch& d = g_scratch[device];
    std::call_once(d.once, [&] {
        d.status = hipSetDevice(device);
        if (d.status == hipSuccess) d.status = hipGetSymbolAddress(reinterpret_cast<void**>(&d.ptr), HIP_SYMBOL(scratch));
        if (d.status != hipSuccess) d.ptr = nullptr;
    });
    if (d.status != hipSuccess || d.ptr == nullptr)
        throw std::runtime_error(std::string("cannot locate the detailed-placement scratch on the device: ") +
                                 hipGetErrorString(d.status));
    return d;
}

void run(const mp_design& d, const int32_t* row_orientation, const mp_dp_options& o, double* centers,
         int32_t* orientation, mp_dp_result& r, bool no_alloc = false) {
    validate(o);
    if (no_alloc) {
        // size check first, from the counts alone: nothing runs for a design that does not fit
        const std::size_t need = scratchBytesNeeded(d.module_count, d.net_count, d.pin_count);
        if (d.module_count < 0 || d.net_count < 0 || d.pin_count < 0)
            throw std::invalid_argument("negative counts in the design");
        if (need > kDpScratchBytes)
            throw std::invalid_argument("design too large for the no-allocation detailed placement: it needs up to " +
                                        std::to_string(need) + " bytes of device scratch, the library holds " +
                                        std::to_string(kDpScratchBytes) + " per device");
    }
    Placement P = loadPlacement(d, row_orientation);
    r = {};
    r.device_used = -1;
    auto publish = [&] {
        for (int64_t m = 0; m < P.M; ++m) {
            centers[2 * m] = P.cx[m];
            centers[2 * m + 1] = P.cy[m];
            if (orientation) orientation[m] = P.orientation[m];
        }
    };
    r.hpwl_before = P.hpwl();
    if (o.method == MP_DP_NONE || P.rows.empty()) {
        r.hpwl_after = r.hpwl_before;
        publish();
        return;
    }
    if (o.method != MP_DP_WINDOW_REORDER || o.window_size != kDpWindow)
        throw Unavailable("HIP detailed placement supports window reordering with a four-cell window only.");
    std::unique_ptr<std::lock_guard<std::mutex>> turn;
    std::unique_ptr<Arena> arena;
    std::size_t limit = 0;
    if (no_alloc) {
        int count = 0;
        const hipError_t e = hipGetDeviceCount(&count);
        if (e != hipSuccess || count == 0)
            throw Unavailable(std::string("HIP detailed-placement backend is unavailable: HIP device discovery failed: ") +
                              (e == hipSuccess ? "no device" : hipGetErrorString(e)));
        if (o.device < 0 || o.device >= count || o.device >= kMaxDevices)
            throw Unavailable("HIP detailed-placement backend is unavailable: Requested HIP device does not exist on this host.");
        DeviceScratch& ds = deviceScratch(o.device);
        turn = std::make_unique<std::lock_guard<std::mutex>>(ds.lock);
        arena = std::make_unique<Arena>(ds.ptr, kDpScratchBytes);
        limit = kDpScratchBytes;
    } else {
        std::string reason;
        limit = allocationLimit(o.device, o.maximum_device_bytes, reason);
        if (limit == 0) throw Unavailable("HIP detailed-placement backend is unavailable: " + reason);
    }
    DeviceWindows dev(P, o.device, limit, arena.get());
    r.device_used = o.device;
    const auto started = std::chrono::steady_clock::now();
    std::vector<double> wx, wy;
    for (int pass = 0; pass < o.passes; ++pass) {
        const std::vector<Segment> segments = rowSegments(P, kSegmentEps);
        const std::size_t phase = static_cast<std::size_t>((pass % 2) * (kDpWindow / 2));
        std::vector<Window> dev_wins, cpu_wins;
        for (const Segment& s : segments) {
            for (std::size_t start = phase; start + 1 < s.modules.size(); start += kDpWindow) {
                const std::size_t count = std::min<std::size_t>(kDpWindow, s.modules.size() - start);
                if (count < 2) break;
                std::vector<int64_t> mods(s.modules.begin() + static_cast<std::ptrdiff_t>(start),
                                          s.modules.begin() + static_cast<std::ptrdiff_t>(start + count));
                Window w = makeWindow(P, s.row, std::move(mods), o.maximum_net_degree);
                (w.device ? dev_wins : cpu_wins).push_back(std::move(w));
            }
        }
        if (!dev_wins.empty()) {
            const std::vector<int> best = dev.score(P, dev_wins);
            for (std::size_t k = 0; k < dev_wins.size(); ++k) {
                ++r.evaluated_windows;
                r.evaluated_permutations += kDpOrderings;
                const int ordering = best[k];
                if (ordering <= 0 || ordering >= kDpOrderings) continue;
                int dec[kDpWindow];
                decodeOrdering(ordering, dec);
                const std::vector<std::size_t> order(dec, dec + kDpWindow);
                const Window& w = dev_wins[k];
                wx.resize(kDpWindow);
                wy.resize(kDpWindow);
                for (int j = 0; j < kDpWindow; ++j) {
                    wx[j] = P.cx[w.modules[j]];
                    wy[j] = P.cy[w.modules[j]];
                }
                const double before = windowHpwl(P, w, wx, wy);
                orderedCenters(P, w, order, wx, wy);
                const double after = windowHpwl(P, w, wx, wy);
                if (after < before - o.improvement_epsilon) {
                    applyOrdering(P, w, order);
                    ++r.accepted_operations;
                }
            }
        }
        for (const Window& w : cpu_wins) {
            const CpuWindowResult c = searchWindow(P, w, o);
            ++r.evaluated_windows;
            r.evaluated_permutations += c.orderings;
            if (c.improved) {
                applyOrdering(P, w, c.order);
                ++r.accepted_operations;
            }
        }
        r.completed_passes = pass + 1;
    }
    r.hpwl_after = P.hpwl();
    r.elapsed_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - started).count();
    r.reserved_device_bytes = dev.peakBytes();
    publish();
}

}  // namespace
}  // namespace mpgp

using namespace mpgp;

extern "C" {

void mp_dp_default_options(mp_dp_options* o) {
    if (!o) return;
    *o = {};
    o->method = MP_DP_WINDOW_REORDER;
    o->passes = 2;
    o->window_size = 4;
    o->maximum_net_degree = 64;
    o->improvement_epsilon = 1e-9;
    o->device = 0;
    o->maximum_device_bytes = 2ULL << 30;
}

mp_gp_status mp_dp_run(const mp_design* design, const int32_t* row_orientation, const mp_dp_options* options,
                       double* centers, int32_t* orientation, mp_dp_result* result) {
    g_last_error.clear();
    try {
        if (!design || !options || !centers || !result) throw std::invalid_argument("null design, options, centres or result");
        mp_dp_result r{};
        run(*design, row_orientation, *options, centers, orientation, r);
        *result = r;
        return MP_GP_OK;
    } catch (const Unavailable& e) {
        g_last_error = e.what();
        return MP_GP_UNAVAILABLE;
    } catch (...) {
        return translate(std::current_exception());
    }
}

mp_gp_status mp_dp_run_noalloc(const mp_design* design, const int32_t* row_orientation, const mp_dp_options* options,
                               double* centers, int32_t* orientation, mp_dp_result* result) {
    g_last_error.clear();
    try {
        if (!design || !options || !centers || !result) throw std::invalid_argument("null design, options, centres or result");
        mp_dp_result r{};
        run(*design, row_orientation, *options, centers, orientation, r, true);
        *result = r;
        return MP_GP_OK;
    } catch (const Unavailable& e) {
        g_last_error = e.what();
        return MP_GP_UNAVAILABLE;
    } catch (...) {
        return translate(std::current_exception());
    }
}

uint64_t mp_dp_noalloc_scratch_bytes(void) { return kDpScratchBytes; }

}  // extern "C"
