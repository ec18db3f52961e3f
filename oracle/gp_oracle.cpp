// gp_oracle.cpp — FP64 CPU restatement of MyPlacement's global-placement path.
//
// TEST INFRASTRUCTURE, NOT PRODUCT.  Only the tests, the smoke check and the CPU-baseline leg of
// the benchmark load this library, and only as the checker.  The HIP backend never links it.
//
// What it restates (paths relative to the MyPlacement tree), each cited at its function:
//   PlacementDatabase derived data     src/model/PlacementDatabase.cpp:55-68,138-142,162-213
//   DensityMap (course_eplace_v1)      src/metrics/Metrics.cpp:10-164, include/.../Metrics.hpp:17-47
//   static density, fillers, deviation src/placement/GlobalPlacementSupport.cpp:11-111
//   Neumann DCT field + parity sample  src/placement/ElectrostaticField.cpp:96-237,289-320
//   smooth (LSE) wirelength            src/placement/AdaptiveGlobalPlacer.cpp:29-82
//   session evaluate (CPU branch)      src/placement/AdaptiveGlobalPlacer.cpp:112-389
//   adaptive controller                src/placement/AdaptiveGlobalPlacer.cpp:413-953
//   HostLayout                         src/placement/cuda/CudaPlacementBackend.cu:784-839
//   random initial placement           src/placement/InitialPlacer.cpp:102-120
//   quadratic initial placement        src/placement/InitialPlacer.cpp:266-402 (BiCGSTAB restated)
//
// The reference CPU path uses FFTW r2r (REDFT10/REDFT01/RODFT01), which is not in this image.
// Its transforms are restated here from their definitions: a direct O(N^2) sum for any N and a
// Makhoul FFT form for power-of-two N (the CPU baseline).  Both are checked against the analytic
// Neumann modes of tests/DensityFieldTests.cpp and against scipy.fft golden vectors.
//
// Parity status of each function is listed in the header comment of tests/test_oracle.py.

#include "mp_gp_backend.h"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

namespace oracle {
namespace {

constexpr double kEpsilon = 1e-9;  // core/Geometry.hpp:8
constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr std::size_t kMaximumBinCount = 1U << 20U;  // core/ResourceLimits.hpp:10

thread_local std::string g_error;

struct V2 {
    double x = 0.0, y = 0.0;
};
inline V2 operator+(V2 a, V2 b) { return {a.x + b.x, a.y + b.y}; }
inline V2 operator-(V2 a, V2 b) { return {a.x - b.x, a.y - b.y}; }
inline V2 operator*(V2 a, double s) { return {a.x * s, a.y * s}; }
inline double sq(V2 a) { return a.x * a.x + a.y * a.y; }
inline double nrm(V2 a) { return std::sqrt(sq(a)); }
inline double dot(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }
inline double clampd(double v, double lo, double hi) { return std::max(lo, std::min(v, hi)); }

struct Rc {  // core/Geometry.hpp:30-86
    V2 ll, ur;
    double width() const { return ur.x - ll.x; }
    double height() const { return ur.y - ll.y; }
    double area() const { return std::max(0.0, width()) * std::max(0.0, height()); }
    V2 center() const { return {(ll.x + ur.x) * 0.5, (ll.y + ur.y) * 0.5}; }
    bool valid() const { return width() >= -kEpsilon && height() >= -kEpsilon; }
    Rc intersection(const Rc& o) const {
        return {{std::max(ll.x, o.ll.x), std::max(ll.y, o.ll.y)}, {std::min(ur.x, o.ur.x), std::min(ur.y, o.ur.y)}};
    }
    double overlapArea(const Rc& o) const { return intersection(o).area(); }
};

inline Rc rectOf(V2 c, double w, double h) {
    return {{c.x - w * 0.5, c.y - h * 0.5}, {c.x + w * 0.5, c.y + h * 0.5}};
}

// PlacementDatabase.cpp:55-68
V2 transformOffset(V2 o, int orientation) {
    switch (orientation) {
        case 0: return o;
        case 1: return {-o.x, -o.y};
        case 2: return {-o.y, o.x};
        case 3: return {o.y, -o.x};
        case 4: return {-o.x, o.y};
        case 5: return {o.x, -o.y};
        case 6: return {o.y, o.x};
        case 7: return {-o.y, -o.x};
        default: return o;
    }
}

// ---------------------------------------------------------------------------------------------
// Database (the fields of PlacementDatabase the path reads) and refreshDerivedData.
struct Db {
    std::vector<double> w, h;
    std::vector<V2> c;
    std::vector<int> orient;
    std::vector<uint8_t> terminal, fixed, macro;
    std::vector<double> net_weight;
    std::vector<int64_t> net_start;
    std::vector<int> pin_module;
    std::vector<V2> pin_offset;
    std::vector<int> module_pin_count;
    std::vector<Rc> rows;
    std::vector<double> row_heights;
    Rc core;
    double nominal_row_height = 0.0;
    std::vector<int64_t> movable;

    std::size_t modules() const { return w.size(); }
    Rc rect(std::size_t m) const { return rectOf(c[m], w[m], h[m]); }
    // PlacementDatabase.cpp:138-142
    V2 pinPosition(std::size_t p) const {
        const int m = pin_module[p];
        return c[m] + transformOffset(pin_offset[p], orient[m]);
    }

    // PlacementDatabase.cpp:162-200 (chip region is not read by the path)
    void refresh() {
        movable.clear();
        if (!rows.empty()) {
            std::vector<double> heights = row_heights;
            std::sort(heights.begin(), heights.end());
            nominal_row_height = heights[heights.size() / 2U];
            core = rows.front();
            for (const Rc& r : rows) {
                core.ll.x = std::min(core.ll.x, r.ll.x);
                core.ll.y = std::min(core.ll.y, r.ll.y);
                core.ur.x = std::max(core.ur.x, r.ur.x);
                core.ur.y = std::max(core.ur.y, r.ur.y);
            }
        } else {
            nominal_row_height = 0.0;
            core = {};
        }
        macro.assign(modules(), 0);
        for (std::size_t id = 0; id < modules(); ++id) {
            if (terminal[id]) fixed[id] = 1;
            macro[id] = !terminal[id] && nominal_row_height > kEpsilon &&
                        std::abs(h[id] - nominal_row_height) > kEpsilon;
            if (!fixed[id]) movable.push_back(static_cast<int64_t>(id));
        }
    }

    // PlacementDatabase.cpp:202-213
    void clampModuleToCore(std::size_t id) {
        if (core.area() <= kEpsilon) return;
        const double lx = core.ll.x + w[id] * 0.5, ux = core.ur.x - w[id] * 0.5;
        const double ly = core.ll.y + h[id] * 0.5, uy = core.ur.y - h[id] * 0.5;
        c[id].x = lx <= ux ? clampd(c[id].x, lx, ux) : core.center().x;
        c[id].y = ly <= uy ? clampd(c[id].y, ly, uy) : core.center().y;
    }
};

Db loadDesign(const mp_design* d) {
    if (d == nullptr) throw std::invalid_argument("design is null");
    Db db;
    const std::size_t M = static_cast<std::size_t>(d->module_count);
    db.w.assign(d->width, d->width + M);
    db.h.assign(d->height, d->height + M);
    db.c.resize(M);
    for (std::size_t i = 0; i < M; ++i) db.c[i] = {d->center[2 * i], d->center[2 * i + 1]};
    db.orient.assign(d->orientation, d->orientation + M);
    db.terminal.assign(d->is_terminal, d->is_terminal + M);
    db.fixed.assign(d->is_fixed, d->is_fixed + M);
    for (std::size_t i = 0; i < M; ++i)
        if (db.terminal[i]) db.fixed[i] = 1;  // BookshelfParser.cpp:92-95
    const std::size_t N = static_cast<std::size_t>(d->net_count);
    db.net_weight.assign(d->net_weight, d->net_weight + N);
    db.net_start.assign(d->net_pin_start, d->net_pin_start + N + 1);
    const std::size_t I = static_cast<std::size_t>(d->pin_count);
    db.pin_module.assign(d->pin_module, d->pin_module + I);
    db.pin_offset.resize(I);
    db.module_pin_count.assign(M, 0);
    for (std::size_t p = 0; p < I; ++p) {
        db.pin_offset[p] = {d->pin_offset[2 * p], d->pin_offset[2 * p + 1]};
        if (db.pin_module[p] < 0 || static_cast<std::size_t>(db.pin_module[p]) >= M)
            throw std::out_of_range("Pin refers to an unknown module or net.");
        db.module_pin_count[db.pin_module[p]] += 1;
    }
    if (db.net_start[0] != 0 || static_cast<std::size_t>(db.net_start[N]) != I)
        throw std::invalid_argument("net_pin_start does not cover the pin array");
    for (std::size_t r = 0; r < static_cast<std::size_t>(d->row_count); ++r) {
        const double* row = d->rows + 5 * r;
        const double bottom = row[0], height = row[1], x_start = row[2], spacing = row[3];
        const std::size_t count = static_cast<std::size_t>(row[4]);
        // SiteRow::xEnd / rect, PlacementDatabase.hpp:69-70
        db.rows.push_back({{x_start, bottom}, {x_start + static_cast<double>(count) * spacing, bottom + height}});
        db.row_heights.push_back(height);
    }
    db.refresh();
    return db;
}

// ---------------------------------------------------------------------------------------------
// DensityMap, Metrics.cpp:10-119.
struct Bin {
    Rc region;
    double fixed = 0, movable = 0, macro = 0, filler = 0, dark = 0;
    // Metrics.hpp:25-29
    double chargeArea(double t, bool include_fillers) const {
        const double scaled = fixed + macro + dark + (include_fillers ? filler : 0.0);
        return movable + t * scaled;
    }
    // Metrics.hpp:33-35
    double normalizationArea(double t) const { return movable + t * macro; }
};

enum Layer { kFixed, kMovable, kMacro, kFiller, kDark };

struct DensityMap {
    Rc region;
    int cols = 0, rows = 0;
    double t = 0, bw = 0, bh = 0;
    std::vector<Bin> bins;

    DensityMap(Rc r, int columns, int rws, double target) : region(r), cols(columns), rows(rws), t(target) {
        if (!region.valid() || region.area() <= kEpsilon || cols <= 0 || rows <= 0 || t <= 0.0)
            throw std::invalid_argument(
                "DensityMap requires a valid region, positive dimensions, and target density.");
        if (static_cast<std::size_t>(cols) > kMaximumBinCount / static_cast<std::size_t>(rows))
            throw std::invalid_argument("Density grid exceeds the safe one-million-bin workspace limit.");
        bw = region.width() / static_cast<double>(cols);
        bh = region.height() / static_cast<double>(rows);
        bins.resize(static_cast<std::size_t>(cols) * static_cast<std::size_t>(rows));
        for (int row = 0; row < rows; ++row)
            for (int col = 0; col < cols; ++col) {
                const double left = region.ll.x + static_cast<double>(col) * bw;
                const double bottom = region.ll.y + static_cast<double>(row) * bh;
                bins[static_cast<std::size_t>(row * cols + col)].region = {{left, bottom},
                                                                            {left + bw, bottom + bh}};
            }
    }

    void clearDynamic() {
        for (Bin& b : bins) b.movable = b.macro = b.filler = 0.0;
    }

    void binRange(const Rc& cl, int& c0, int& c1, int& r0, int& r1) const {
        c0 = std::max(0, std::min(cols - 1, static_cast<int>(std::floor((cl.ll.x - region.ll.x) / bw))));
        c1 = std::max(0, std::min(cols - 1, static_cast<int>(std::ceil((cl.ur.x - region.ll.x) / bw) - 1.0)));
        r0 = std::max(0, std::min(rows - 1, static_cast<int>(std::floor((cl.ll.y - region.ll.y) / bh))));
        r1 = std::max(0, std::min(rows - 1, static_cast<int>(std::ceil((cl.ur.y - region.ll.y) / bh) - 1.0)));
    }

    // Metrics.cpp:42-72
    void setPlaceableRegions(const std::vector<Rc>& regions) {
        for (Bin& b : bins) b.dark = b.region.area();
        for (const Rc& reg : regions) {
            const Rc cl = reg.intersection(region);
            if (cl.area() <= kEpsilon) continue;
            int c0, c1, r0, r1;
            binRange(cl, c0, c1, r0, r1);
            for (int row = r0; row <= r1; ++row)
                for (int col = c0; col <= c1; ++col) {
                    Bin& b = bins[static_cast<std::size_t>(row * cols + col)];
                    b.dark -= cl.overlapArea(b.region);
                }
        }
        for (Bin& b : bins) {
            const double tol = std::max(1.0, b.region.area()) * 1e-10;
            if (b.dark < -tol)
                throw std::invalid_argument("Overlapping SiteRows cannot define a density placeable-area mask.");
            b.dark = std::max(0.0, b.dark);
        }
    }

    // Metrics.cpp:74-99
    void addRectangle(const Rc& rect, Layer layer) {
        const Rc cl = rect.intersection(region);
        if (cl.area() <= kEpsilon) return;
        int c0, c1, r0, r1;
        binRange(cl, c0, c1, r0, r1);
        for (int row = r0; row <= r1; ++row)
            for (int col = c0; col <= c1; ++col) {
                Bin& b = bins[static_cast<std::size_t>(row * cols + col)];
                const double a = cl.overlapArea(b.region);
                if (a <= kEpsilon) continue;
                switch (layer) {
                    case kFixed: b.fixed += a; break;
                    case kMovable: b.movable += a; break;
                    case kMacro: b.macro += a; break;
                    case kFiller: b.filler += a; break;
                    case kDark: b.dark += a; break;
                }
            }
    }

    // Metrics.cpp:101-119
    mp_gp_density_metrics metrics(bool include_fillers) const {
        mp_gp_density_metrics r{};
        double total_density = 0.0;
        for (const Bin& b : bins) {
            const double ba = b.region.area();
            const double q = b.chargeArea(t, include_fillers);
            const double d = q / std::max(ba, kEpsilon);
            r.maximum_density = std::max(r.maximum_density, d);
            total_density += d;
            r.total_overflow_area += std::max(0.0, q - t * ba);
            r.total_charge_area += q;
            r.normalization_area += b.normalizationArea(t);
            r.dark_area += b.dark;
            r.placeable_area += ba - b.dark;
        }
        r.average_density = total_density / static_cast<double>(std::max<std::size_t>(1U, bins.size()));
        r.normalized_overflow = r.total_overflow_area / std::max(r.normalization_area, kEpsilon);
        return r;
    }
};

// ---------------------------------------------------------------------------------------------
// Fillers, GlobalPlacementSupport.cpp:52-103.
struct Filler {
    V2 c;
    double w = 0, h = 0;
    double area() const { return w * h; }
    Rc rect() const { return rectOf(c, w, h); }
};

// GlobalPlacementSupport.cpp:11-19
V2 clampCenterToRegion(V2 c, double w, double h, const Rc& region) {
    const double lx = region.ll.x + w * 0.5, ux = region.ur.x - w * 0.5;
    const double ly = region.ll.y + h * 0.5, uy = region.ur.y - h * 0.5;
    c.x = lx <= ux ? clampd(c.x, lx, ux) : region.center().x;
    c.y = ly <= uy ? clampd(c.y, ly, uy) : region.center().y;
    return c;
}

// GlobalPlacementSupport.cpp:21-30
void initializeStaticDensity(DensityMap& map, const Db& db) {
    if (!db.rows.empty()) map.setPlaceableRegions(db.rows);
    for (std::size_t m = 0; m < db.modules(); ++m)
        if (db.fixed[m]) map.addRectangle(db.rect(m), kFixed);
}

std::vector<Filler> createDensityFillers(const Db& db, const DensityMap& static_density, double t,
                                         std::size_t maximum_fillers, uint64_t seed) {
    double std_area = 0.0, macro_area = 0.0;
    std::vector<double> areas;
    for (const int64_t id : db.movable) {
        const double a = db.rect(id).intersection(db.core).area();
        if (db.macro[id]) macro_area += a; else std_area += a;
        if (a > kEpsilon) areas.push_back(a);
    }
    const double static_charge = static_density.metrics(false).total_charge_area;
    const double movable_charge = std_area + t * macro_area;
    const double required = std::max(0.0, t * db.core.area() - static_charge - movable_charge);
    const double filler_area = required / std::max(t, kEpsilon);
    if (filler_area <= kEpsilon || areas.empty() || maximum_fillers == 0U) return {};
    std::sort(areas.begin(), areas.end());
    const std::size_t lower = areas.size() / 10U;
    const std::size_t upper = std::max(lower + 1U, areas.size() - areas.size() / 10U);
    const double representative = std::accumulate(areas.begin() + static_cast<std::ptrdiff_t>(lower),
                                                  areas.begin() + static_cast<std::ptrdiff_t>(upper), 0.0) /
                                  static_cast<double>(upper - lower);
    const std::size_t count = std::min(
        maximum_fillers,
        std::max<std::size_t>(1U, static_cast<std::size_t>(std::ceil(filler_area / std::max(representative, kEpsilon)))));
    const double per = filler_area / static_cast<double>(count);
    const double aspect = std::max(db.core.width() / db.core.height(), 0.1);
    const double width = std::sqrt(per * aspect);
    const double height = per / std::max(width, kEpsilon);
    std::mt19937_64 gen(seed ^ 0x9e3779b97f4a7c15ULL);
    std::uniform_real_distribution<double> xd(db.core.ll.x + width * 0.5, db.core.ur.x - width * 0.5);
    std::uniform_real_distribution<double> yd(db.core.ll.y + height * 0.5, db.core.ur.y - height * 0.5);
    std::vector<Filler> out;
    out.reserve(count);
    for (std::size_t i = 0; i < count; ++i) {
        const double x = xd(gen);
        const double y = yd(gen);
        out.push_back({{x, y}, width, height});
    }
    return out;
}

// ---------------------------------------------------------------------------------------------
// 1-D real transforms with FFTW's unnormalised r2r definitions.
//   REDFT10 (DCT-II):  Y_k = 2 sum_n x_n cos(pi k (2n+1) / 2N)
//   REDFT01 (DCT-III): Y_k = x_0 + 2 sum_{n>=1} x_n cos(pi n (2k+1) / 2N)
//   RODFT01 (DST-III): Y_k = (-1)^k x_{N-1} + 2 sum_{n<N-1} x_n sin(pi (n+1)(2k+1) / 2N)
enum Kind { kRedft10 = 0, kRedft01 = 1, kRodft01 = 2 };
enum Method { kAuto = 0, kDirect = 1, kFft = 2 };
int g_method = kAuto;

struct Trig {  // cos(pi m / 2N) for m in [0, 4N)
    int n = 0;
    std::vector<double> cosv;
    explicit Trig(int N) : n(N), cosv(4 * static_cast<std::size_t>(N)) {
        for (int m = 0; m < 4 * N; ++m) cosv[m] = std::cos(kPi * static_cast<double>(m) / (2.0 * N));
    }
    double c(long long m) const { return cosv[static_cast<std::size_t>(((m % (4 * n)) + 4 * n) % (4 * n))]; }
    double s(long long m) const { return c(m - n); }  // sin(a) = cos(a - pi/2)
};

void directTransform(int kind, int N, const double* in, double* out, const Trig& tr) {
    for (int k = 0; k < N; ++k) {
        double acc = 0.0;
        if (kind == kRedft10) {
            for (int n = 0; n < N; ++n) acc += in[n] * tr.c(static_cast<long long>(k) * (2 * n + 1));
            acc *= 2.0;
        } else if (kind == kRedft01) {
            for (int n = 1; n < N; ++n) acc += in[n] * tr.c(static_cast<long long>(n) * (2 * k + 1));
            acc = in[0] + 2.0 * acc;
        } else {
            for (int n = 0; n < N - 1; ++n) acc += in[n] * tr.s(static_cast<long long>(n + 1) * (2 * k + 1));
            acc = ((k & 1) ? -in[N - 1] : in[N - 1]) + 2.0 * acc;
        }
        out[k] = acc;
    }
}

bool isPowerOfTwo(int n) { return n > 0 && (n & (n - 1)) == 0; }

struct Fft {  // iterative radix-2 complex FFT, twiddles computed directly per index
    int n = 0;
    std::vector<std::complex<double>> tw;
    std::vector<int> rev;
    explicit Fft(int N) : n(N), tw(N / 2), rev(N) {
        for (int k = 0; k < N / 2; ++k)
            tw[k] = {std::cos(2.0 * kPi * k / N), -std::sin(2.0 * kPi * k / N)};
        int bits = 0;
        while ((1 << bits) < N) ++bits;
        for (int i = 0; i < N; ++i) {
            int r = 0;
            for (int b = 0; b < bits; ++b)
                if (i & (1 << b)) r |= 1 << (bits - 1 - b);
            rev[i] = r;
        }
    }
    void run(std::complex<double>* a, bool inverse) const {
        for (int i = 0; i < n; ++i)
            if (i < rev[i]) std::swap(a[i], a[rev[i]]);
        for (int len = 2; len <= n; len <<= 1) {
            const int step = n / len;
            for (int i = 0; i < n; i += len)
                for (int j = 0; j < len / 2; ++j) {
                    std::complex<double> w = tw[j * step];
                    if (inverse) w = std::conj(w);
                    const std::complex<double> u = a[i + j], v = a[i + j + len / 2] * w;
                    a[i + j] = u + v;
                    a[i + j + len / 2] = u - v;
                }
        }
    }
};

// Makhoul's N-point forms of DCT-II and DCT-III; DST-III = (-1)^k DCT-III(reversed input).
void fftTransform(int kind, int N, const double* in, double* out, const Fft& fft, const Trig& tr,
                  std::vector<std::complex<double>>& work) {
    work.resize(N);
    if (kind == kRedft10) {
        for (int n = 0; n < N / 2; ++n) {
            work[n] = {in[2 * n], 0.0};
            work[N - 1 - n] = {in[2 * n + 1], 0.0};
        }
        fft.run(work.data(), false);
        for (int k = 0; k < N; ++k) {
            const std::complex<double> e(tr.c(k), -tr.s(k));  // exp(-i pi k / 2N)
            out[k] = 2.0 * (e * work[k]).real();
        }
        return;
    }
    const bool sine = kind == kRodft01;
    auto X = [&](int k) -> double {
        if (k >= N) return 0.0;
        return sine ? in[N - 1 - k] : in[k];
    };
    for (int k = 0; k < N; ++k) {
        const std::complex<double> e(tr.c(k), tr.s(k));  // exp(+i pi k / 2N)
        work[k] = e * std::complex<double>(X(k), k == 0 ? 0.0 : -X(N - k));
    }
    fft.run(work.data(), true);
    for (int m = 0; m < N / 2; ++m) {
        out[2 * m] = work[m].real();
        out[2 * m + 1] = work[N - 1 - m].real();
    }
    if (sine)
        for (int k = 1; k < N; k += 2) out[k] = -out[k];
}

struct Transformer {
    int N;
    Trig tr;
    std::unique_ptr<Fft> fft;
    std::vector<std::complex<double>> work;
    std::vector<double> a, b;
    explicit Transformer(int n) : N(n), tr(n), a(n), b(n) {
        if (isPowerOfTwo(n) && n >= 4) fft = std::make_unique<Fft>(n);
    }
    void run(int kind, const double* in, double* out) {
        const bool use_fft = fft && g_method != kDirect;
        if (use_fft) fftTransform(kind, N, in, out, *fft, tr, work);
        else directTransform(kind, N, in, out, tr);
    }
    // strided transform along one grid axis
    void runStrided(int kind, const double* in, double* out, std::size_t stride) {
        for (int i = 0; i < N; ++i) a[i] = in[i * stride];
        run(kind, a.data(), b.data());
        for (int i = 0; i < N; ++i) out[i * stride] = b[i];
    }
};

// 2-D r2r with FFTW's row-major (rows, cols) convention: kind_y along rows' index (dim 0),
// kind_x along columns (dim 1).  Matches fftw_plan_r2r_2d(rows, cols, ..., kind0, kind1).
void transform2d(int cols, int rows, int kind_y, int kind_x, const double* in, double* out, Transformer& tx,
                 Transformer& ty) {
    std::vector<double> tmp(static_cast<std::size_t>(cols) * rows);
    for (int r = 0; r < rows; ++r) tx.run(kind_x, in + static_cast<std::size_t>(r) * cols, tmp.data() + static_cast<std::size_t>(r) * cols);
    for (int c = 0; c < cols; ++c) ty.runStrided(kind_y, tmp.data() + c, out + c, static_cast<std::size_t>(cols));
}

// ---------------------------------------------------------------------------------------------
// ElectrostaticField (Neumann), ElectrostaticField.cpp:96-237,289-320.
struct Field {
    int cols, rows;
    Rc region;
    double bw, bh;
    std::size_t count;
    Transformer tx, ty;
    std::vector<double> source, coeff, pot_c, fx_c, fy_c, potential, fx, fy;

    Field(int c, int r, Rc reg)
        : cols(c), rows(r), region(reg), count(static_cast<std::size_t>(c) * r), tx(c), ty(r) {
        if (cols < 2 || rows < 2 || !region.valid() || region.area() <= kEpsilon)
            throw std::invalid_argument("ElectrostaticField requires a valid region and at least a 2 x 2 grid.");
        if (count > kMaximumBinCount)
            throw std::invalid_argument("Density grid exceeds the safe one-million-bin workspace limit.");
        bw = region.width() / static_cast<double>(cols);
        bh = region.height() / static_cast<double>(rows);
        for (auto* v : {&source, &coeff, &pot_c, &fx_c, &fy_c, &potential, &fx, &fy}) v->assign(count, 0.0);
    }

    // ElectrostaticField.cpp:119-133 then solveNeumann :205-237
    void solve(const std::vector<double>& dev) {
        if (dev.size() != count)
            throw std::invalid_argument("Density-field input does not match the configured grid dimensions.");
        const double mean = std::accumulate(dev.begin(), dev.end(), 0.0) / static_cast<double>(count);
        for (std::size_t i = 0; i < count; ++i) source[i] = dev[i] - mean;
        transform2d(cols, rows, kRedft10, kRedft10, source.data(), coeff.data(), tx, ty);
        std::fill(pot_c.begin(), pot_c.end(), 0.0);
        std::fill(fx_c.begin(), fx_c.end(), 0.0);
        std::fill(fy_c.begin(), fy_c.end(), 0.0);
        const double normalization = 1.0 / (4.0 * static_cast<double>(cols) * static_cast<double>(rows));
        for (int row = 0; row < rows; ++row) {
            const double wy = kPi * static_cast<double>(row) / region.height();
            for (int col = 0; col < cols; ++col) {
                const double wx = kPi * static_cast<double>(col) / region.width();
                const double den = wx * wx + wy * wy;
                if (den <= kEpsilon) continue;
                const std::size_t idx = static_cast<std::size_t>(row * cols + col);
                const double p = coeff[idx] * normalization / den;
                pot_c[idx] = p;
                if (col > 0) fx_c[static_cast<std::size_t>(row * cols + col - 1)] = p * wx;
                if (row > 0) fy_c[static_cast<std::size_t>((row - 1) * cols + col)] = p * wy;
            }
        }
        transform2d(cols, rows, kRedft01, kRedft01, pot_c.data(), potential.data(), tx, ty);
        transform2d(cols, rows, kRedft01, kRodft01, fx_c.data(), fx.data(), tx, ty);
        transform2d(cols, rows, kRodft01, kRedft01, fy_c.data(), fy.data(), tx, ty);
    }

    // ElectrostaticField.cpp:289-299
    static int reflectIndex(int index, int size, bool odd, bool& negate) {
        while (index < 0 || index >= size) {
            if (index < 0) index = -index - 1;
            else index = 2 * size - index - 1;
            if (odd) negate = !negate;
        }
        return index;
    }

    // ElectrostaticField.cpp:301-320
    double sampleWithParity(const double* values, V2 p, bool odd_x, bool odd_y) const {
        p.x = clampd(p.x, region.ll.x, region.ur.x);
        p.y = clampd(p.y, region.ll.y, region.ur.y);
        const double gx = (p.x - region.ll.x) / bw - 0.5;
        const double gy = (p.y - region.ll.y) / bh - 0.5;
        const int left = static_cast<int>(std::floor(gx));
        const int bottom = static_cast<int>(std::floor(gy));
        const double tx_ = gx - static_cast<double>(left);
        const double ty_ = gy - static_cast<double>(bottom);
        auto at = [&](int col, int row) {
            bool neg = false;
            const int rc = reflectIndex(col, cols, odd_x, neg);
            const int rr = reflectIndex(row, rows, odd_y, neg);
            const double v = values[static_cast<std::size_t>(rr * cols + rc)];
            return neg ? -v : v;
        };
        const double lower = at(left, bottom) * (1.0 - tx_) + at(left + 1, bottom) * tx_;
        const double upper = at(left, bottom + 1) * (1.0 - tx_) + at(left + 1, bottom + 1) * tx_;
        return lower * (1.0 - ty_) + upper * ty_;
    }
    V2 sampleField(V2 p) const {
        return {sampleWithParity(fx.data(), p, true, false), sampleWithParity(fy.data(), p, false, true)};
    }
};

// ---------------------------------------------------------------------------------------------
// Smooth wirelength, AdaptiveGlobalPlacer.cpp:29-82 (per module gradient).
struct Wire {
    double smooth = 0.0, hpwl = 0.0;
    std::vector<V2> grad;
};

Wire evaluateSmoothWirelength(const Db& db, double smoothing, bool want_grad) {
    Wire r;
    if (want_grad) r.grad.assign(db.modules(), V2{});
    const double gamma = std::max(smoothing, kEpsilon);
    const std::size_t N = db.net_weight.size();
    std::vector<V2> pos;
    for (std::size_t n = 0; n < N; ++n) {
        const int64_t b = db.net_start[n], e = db.net_start[n + 1];
        if (e - b < 2) continue;
        pos.resize(static_cast<std::size_t>(e - b));
        double mx = -std::numeric_limits<double>::infinity(), nx = std::numeric_limits<double>::infinity();
        double my = -std::numeric_limits<double>::infinity(), ny = std::numeric_limits<double>::infinity();
        for (int64_t p = b; p < e; ++p) {
            const V2 q = db.pinPosition(static_cast<std::size_t>(p));
            pos[p - b] = q;
            mx = std::max(mx, q.x); nx = std::min(nx, q.x);
            my = std::max(my, q.y); ny = std::min(ny, q.y);
        }
        double spx = 0, snx = 0, spy = 0, sny = 0;
        for (const V2 q : pos) {
            spx += std::exp((q.x - mx) / gamma);
            snx += std::exp((nx - q.x) / gamma);
            spy += std::exp((q.y - my) / gamma);
            sny += std::exp((ny - q.y) / gamma);
        }
        const double wgt = db.net_weight[n];
        r.smooth += wgt * (mx + gamma * std::log(spx) - nx + gamma * std::log(snx) + my + gamma * std::log(spy) -
                           ny + gamma * std::log(sny));
        r.hpwl += wgt * ((mx - nx) + (my - ny));
        if (!want_grad) continue;
        for (int64_t p = b; p < e; ++p) {
            const int m = db.pin_module[p];
            if (db.fixed[m]) continue;
            const V2 q = pos[p - b];
            const double gx = std::exp((q.x - mx) / gamma) / spx - std::exp((nx - q.x) / gamma) / snx;
            const double gy = std::exp((q.y - my) / gamma) / spy - std::exp((ny - q.y) / gamma) / sny;
            r.grad[m].x += wgt * gx;
            r.grad[m].y += wgt * gy;
        }
    }
    if (!std::isfinite(r.smooth) || !std::isfinite(r.hpwl))
        throw std::runtime_error("Smooth wirelength evaluation produced a non-finite value.");
    return r;
}

// ---------------------------------------------------------------------------------------------
// GlobalPlacementSession (CPU branch), AdaptiveGlobalPlacer.cpp:112-389.
struct Eval {
    mp_gp_scalars s{};
    double objective = 0.0;
    std::vector<V2> wire, density, raw, pre;
    double raw_norm = 0, pre_norm = 0, wire_l1 = 0, density_l1 = 0;
};

struct Session {
    Db db;
    double t;
    int bins_x, bins_y;
    DensityMap map;
    std::vector<Filler> fillers;
    std::vector<double> precond;
    std::unique_ptr<Field> field;
    std::vector<double> deviation;
    // HostLayout, .cu:764-839
    std::vector<double> pw, ph, pq, pos_static;
    std::vector<int32_t> players, pin_particles, net_offsets;
    std::vector<double> net_w;
    int64_t grad_evals = 0, value_evals = 0;

    Session(const mp_design* d, const mp_gp_options* o)
        : db(loadDesign(d)), t(o->target_density), bins_x(o->bins_x), bins_y(o->bins_y),
          map(db.core, o->bins_x, o->bins_y, o->target_density) {
        initializeStaticDensity(map, db);
        fillers = createDensityFillers(db, map, t, static_cast<std::size_t>(o->maximum_fillers), o->seed);
        initializePreconditioner();
        field = std::make_unique<Field>(bins_x, bins_y, db.core);
        buildLayout();
    }

    std::size_t particles() const { return db.movable.size() + fillers.size(); }
    double minBin() const { return std::min(map.bw, map.bh); }
    double maxBin() const { return std::max(map.bw, map.bh); }

    // .cu:784-839 (particle order: movable then fillers; pins in net order)
    void buildLayout() {
        std::vector<int> module_particle(db.modules(), -1);
        for (std::size_t i = 0; i < db.movable.size(); ++i) {
            const int64_t m = db.movable[i];
            module_particle[m] = static_cast<int>(i);
            pw.push_back(db.w[m]);
            ph.push_back(db.h[m]);
            pq.push_back(db.macro[m] ? t * db.w[m] * db.h[m] : db.w[m] * db.h[m]);  // :105-107
            players.push_back(db.macro[m] ? 1 : 0);
        }
        for (const Filler& f : fillers) {
            pw.push_back(f.w);
            ph.push_back(f.h);
            pq.push_back(t * f.area());  // :109-111
            players.push_back(2);
        }
        net_offsets.push_back(0);
        for (std::size_t n = 0; n < db.net_weight.size(); ++n) {
            net_w.push_back(db.net_weight[n]);
            for (int64_t p = db.net_start[n]; p < db.net_start[n + 1]; ++p) {
                const int m = db.pin_module[p];
                const int particle = module_particle[m];
                V2 v;
                if (particle >= 0) v = transformOffset(db.pin_offset[p], db.orient[m]);
                else v = db.pinPosition(static_cast<std::size_t>(p));
                pos_static.push_back(v.x);
                pos_static.push_back(v.y);
                pin_particles.push_back(particle);
            }
            net_offsets.push_back(static_cast<int32_t>(pin_particles.size()));
        }
    }

    // AdaptiveGlobalPlacer.cpp:358-375
    void initializePreconditioner() {
        double total = 0.0;
        for (const int64_t id : db.movable) total += db.w[id] * db.h[id];
        const double avg = total / static_cast<double>(std::max<std::size_t>(1U, db.movable.size()));
        auto add = [&](double area, std::size_t pins) {
            const double ratio = std::max(area / std::max(avg, kEpsilon), 0.1);
            const double pin_factor = std::sqrt(1.0 + static_cast<double>(pins));
            const double area_factor = 1.0 + std::sqrt(ratio);
            precond.push_back(1.0 / (pin_factor * area_factor));
        };
        for (const int64_t id : db.movable) add(db.w[id] * db.h[id], static_cast<std::size_t>(db.module_pin_count[id]));
        for (const Filler& f : fillers) add(f.area(), 0U);
    }

    std::vector<V2> capturePositions() const {  // :174-183
        std::vector<V2> p(particles());
        for (std::size_t i = 0; i < db.movable.size(); ++i) p[i] = db.c[db.movable[i]];
        for (std::size_t i = 0; i < fillers.size(); ++i) p[db.movable.size() + i] = fillers[i].c;
        return p;
    }

    void applyPositions(const std::vector<V2>& p) {  // :185-198
        if (p.size() != particles())
            throw std::invalid_argument("Adaptive placement state does not match the particle count.");
        for (std::size_t i = 0; i < db.movable.size(); ++i) {
            const int64_t m = db.movable[i];
            db.c[m] = clampCenterToRegion(p[i], db.w[m], db.h[m], db.core);
        }
        for (std::size_t i = 0; i < fillers.size(); ++i)
            fillers[i].c = clampCenterToRegion(p[db.movable.size() + i], fillers[i].w, fillers[i].h, db.core);
    }

    void setRaw(const std::vector<V2>& p) {  // positions as given, no clamp (backend evaluate contract)
        for (std::size_t i = 0; i < db.movable.size(); ++i) db.c[db.movable[i]] = p[i];
        for (std::size_t i = 0; i < fillers.size(); ++i) fillers[i].c = p[db.movable.size() + i];
    }

    // :333-356 with GlobalPlacementSupport.cpp:32-50
    void evaluateDensity(Eval& e) {
        map.clearDynamic();
        for (const int64_t m : db.movable) map.addRectangle(db.rect(m), db.macro[m] ? kMacro : kMovable);
        for (const Filler& f : fillers) map.addRectangle(f.rect(), kFiller);
        deviation.resize(map.bins.size());
        for (std::size_t i = 0; i < map.bins.size(); ++i)
            deviation[i] = map.bins[i].chargeArea(t, true) / std::max(map.bins[i].region.area(), kEpsilon) - t;
        field->solve(deviation);
        e.s.optimizer_density = map.metrics(true);
        e.s.design_density = map.metrics(false);
        const double mean = std::accumulate(deviation.begin(), deviation.end(), 0.0) /
                            static_cast<double>(std::max<std::size_t>(1U, deviation.size()));
        double energy = 0.0;
        for (std::size_t i = 0; i < map.bins.size(); ++i) {
            const double src = deviation[i] - mean;
            energy += 0.5 * src * field->potential[i] * map.bins[i].region.area();
        }
        if (!std::isfinite(energy)) throw std::runtime_error("Electrostatic density energy produced a non-finite value.");
        e.s.electrostatic_energy = std::max(0.0, energy);
    }

    // Backend-shaped evaluate: scalars + per-particle wire and density gradients (.cu:921-950
    // semantics, CPU arithmetic of :200-268).
    void evaluateRaw(double smoothing, bool want_grad, Eval& e) {
        evaluateDensity(e);
        Wire w = evaluateSmoothWirelength(db, smoothing, want_grad);
        e.s.hpwl = w.hpwl;
        e.s.smooth_wirelength = w.smooth;
        if (want_grad) ++grad_evals; else ++value_evals;
        if (!want_grad) return;
        const std::size_t P = particles();
        e.wire.assign(P, V2{});
        e.density.assign(P, V2{});
        for (std::size_t i = 0; i < db.movable.size(); ++i) {
            const int64_t m = db.movable[i];
            e.wire[i] = w.grad[m];
            e.density[i] = field->sampleField(db.c[m]) * (-pq[i]);
        }
        e.wire_l1 = 0;
        for (const V2 g : w.grad) e.wire_l1 += std::abs(g.x) + std::abs(g.y);
        for (std::size_t i = 0; i < fillers.size(); ++i) {
            const std::size_t k = db.movable.size() + i;
            e.density[k] = field->sampleField(fillers[i].c) * (-pq[k]);
        }
    }

    // :200-268
    Eval evaluate(double smoothing, double penalty, double reference_penalty, bool want_grad) {
        Eval e;
        evaluateRaw(smoothing, want_grad, e);
        e.objective = e.s.smooth_wirelength + penalty * e.s.electrostatic_energy;
        if (!std::isfinite(e.objective)) throw std::runtime_error("Adaptive placement objective produced a non-finite value.");
        if (!want_grad) return e;
        const std::size_t P = particles();
        e.raw.resize(P);
        e.pre.resize(P);
        e.density_l1 = 0;
        for (std::size_t i = 0; i < P; ++i) {
            e.density_l1 += std::abs(e.density[i].x) + std::abs(e.density[i].y);
            e.raw[i] = (i < db.movable.size() ? e.wire[i] : V2{}) + e.density[i] * penalty;
        }
        const double strength = std::sqrt(1.0 + std::max(0.0, penalty / std::max(reference_penalty, kEpsilon)));
        double rn = 0, pn = 0;
        for (std::size_t i = 0; i < P; ++i) {
            e.pre[i] = e.raw[i] * (precond[i] / strength);
            rn += sq(e.raw[i]);
            pn += sq(e.pre[i]);
        }
        e.raw_norm = std::sqrt(rn);
        e.pre_norm = std::sqrt(pn);
        return e;
    }
};

// ---------------------------------------------------------------------------------------------
// Adaptive controller, AdaptiveGlobalPlacer.cpp:413-953 (RUDY disabled).
void validateOptions(const Db& db, const mp_gp_options& o) {  // :413-454 (non-RUDY part)
    if (db.movable.empty()) return;
    if (o.iterations <= 0 || o.bins_x < 2 || o.bins_y < 2 || o.target_density <= 0.0 || o.target_density > 1.0 ||
        o.maximum_bin_overflow < 0.0 || o.initial_smoothing <= 0.0 || o.final_smoothing <= 0.0 ||
        o.maximum_movement_in_bins <= 0.0 || o.initial_movement_in_bins <= 0.0 || o.maximum_backtracks < 0 ||
        o.backtracking_ratio <= 0.0 || o.backtracking_ratio >= 1.0 || o.armijo_coefficient <= 0.0 ||
        o.armijo_coefficient >= 1.0 || o.maximum_momentum < 0.0 || o.maximum_momentum >= 1.0 ||
        o.density_penalty_increase <= 1.0 || o.density_penalty_decrease <= 0.0 || o.density_penalty_decrease >= 1.0 ||
        o.minimum_density_penalty <= 0.0 || o.maximum_density_penalty < o.minimum_density_penalty ||
        o.smoothing_overflow_exponent <= 0.0 || o.feasible_refinement_iterations <= 0)
        throw std::invalid_argument("Adaptive global placement options are outside their valid range.");
    if (static_cast<std::size_t>(o.bins_x) > kMaximumBinCount / static_cast<std::size_t>(o.bins_y))
        throw std::invalid_argument("Density grid exceeds the safe one-million-bin workspace limit.");
}

double smoothingFromOverflow(const mp_gp_options& o, double max_bin, double overflow, double initial) {  // :456-462
    const double ratio = clampd(overflow / std::max(initial, kEpsilon), 0.0, 1.0);
    const double annealing = std::pow(ratio, o.smoothing_overflow_exponent);
    return (o.final_smoothing + (o.initial_smoothing - o.final_smoothing) * annealing) * max_bin;
}

double initialStepSize(const Eval& e, double desired) {  // :464-468
    const double count = static_cast<double>(std::max<std::size_t>(1U, e.pre.size()));
    const double rms = e.pre_norm / std::sqrt(count);
    return desired / std::max(rms, kEpsilon);
}

struct StepEstimate {
    double step = 0, curvature = 0;
};
StepEstimate estimateStep(const std::vector<V2>& pos, const Eval& e, const std::vector<V2>& prev_pos,
                          const std::vector<V2>& prev_grad, bool has_prev, double fallback) {  // :470-492
    StepEstimate r;
    r.step = fallback;
    if (!has_prev || pos.size() != prev_pos.size() || e.pre.size() != prev_grad.size()) return r;
    double dp = 0, dg = 0;
    for (std::size_t i = 0; i < pos.size(); ++i) {
        dp += sq(pos[i] - prev_pos[i]);
        dg += sq(e.pre[i] - prev_grad[i]);
    }
    if (dp <= kEpsilon || dg <= kEpsilon) return r;
    r.curvature = std::sqrt(dg / dp);
    const double cs = 1.0 / std::max(r.curvature, kEpsilon);
    r.step = clampd(cs, fallback * 0.25, fallback * 4.0);
    return r;
}

struct Proposal {
    std::vector<V2> positions;
    double predicted = 0, max_disp = 0;
};
Proposal proposeStep(Session& s, const std::vector<V2>& base, const Eval& e, double step, double max_disp) {  // :494-519
    if (base.size() != e.pre.size()) throw std::invalid_argument("Adaptive gradient does not match the current placement state.");
    std::vector<V2> p = base;
    for (std::size_t i = 0; i < p.size(); ++i) {
        V2 d = e.pre[i] * (-step);
        const double dn = nrm(d);
        if (dn > max_disp) d = d * (max_disp / dn);
        p[i] = p[i] + d;
    }
    s.applyPositions(p);
    Proposal r;
    r.positions = s.capturePositions();
    for (std::size_t i = 0; i < r.positions.size(); ++i) {
        const V2 a = base[i] - r.positions[i];
        r.predicted += dot(e.raw[i], a);
        r.max_disp = std::max(r.max_disp, nrm(a));
    }
    return r;
}

bool acceptsCandidate(const Eval& base, const Eval& cand, double predicted, const mp_gp_options& o) {  // :521-534
    const double tol = std::max(1e-8, std::abs(base.objective) * 1e-10);
    const double armijo_target = base.objective - o.armijo_coefficient * std::max(0.0, predicted);
    const bool armijo = cand.objective <= armijo_target + tol;
    const bool stable = cand.objective <= base.objective + tol &&
                        cand.s.optimizer_density.normalized_overflow <= base.s.optimizer_density.normalized_overflow * 1.005;
    const double gain = std::max(1e-4, o.density_stall_threshold * 0.25);
    const bool trade = cand.s.optimizer_density.normalized_overflow <= base.s.optimizer_density.normalized_overflow - gain &&
                       cand.objective <= base.objective * (1.0 + o.objective_increase_for_density);
    return armijo || stable || trade;
}

double updateDensityPenalty(double penalty, double ref, const Eval& before, const Eval& after, const mp_gp_options& o) {  // :536-557
    const double growth = (after.s.hpwl - before.s.hpwl) / std::max(before.s.hpwl, 1.0);
    const double improvement = before.s.optimizer_density.normalized_overflow - after.s.optimizer_density.normalized_overflow;
    const bool critical = after.s.optimizer_density.normalized_overflow > std::max(2.0 * o.maximum_bin_overflow, 0.02);
    if (critical && improvement < o.density_stall_threshold) penalty *= o.density_penalty_increase;
    else if (!critical && growth > o.hpwl_growth_restart_threshold && improvement < o.density_stall_threshold)
        penalty *= o.density_penalty_decrease;
    else if (after.s.optimizer_density.normalized_overflow > o.maximum_bin_overflow && improvement < o.density_stall_threshold)
        penalty *= o.density_penalty_increase;
    else if (after.s.design_density.normalized_overflow <= o.maximum_bin_overflow)
        penalty *= o.density_penalty_decrease;
    const double lo = std::max(o.minimum_density_penalty, ref * 0.02);
    const double hi = std::max(o.maximum_density_penalty, lo * 10.0);
    return clampd(penalty, lo, hi);
}

struct Checkpoint {
    std::vector<V2> positions;
    int iteration = 0;
    double hpwl = std::numeric_limits<double>::infinity(), overflow = std::numeric_limits<double>::infinity();
    double score = std::numeric_limits<double>::infinity();
    bool found = false;
};

void considerCheckpoint(Checkpoint& feas, Checkpoint& fall, const std::vector<V2>& pos, const Eval& e, int it,
                        const mp_gp_options& o, bool& is_new) {  // :563-596 (routability weight 0)
    is_new = false;
    const double ovf = e.s.design_density.normalized_overflow;
    const bool feasible = ovf <= o.maximum_bin_overflow;
    const double score = e.s.hpwl;
    if (feasible && (!feas.found || score < feas.score - 1e-10 ||
                     (std::abs(score - feas.score) <= 1e-10 && e.s.hpwl < feas.hpwl))) {
        feas.positions = pos; feas.iteration = it; feas.hpwl = e.s.hpwl; feas.overflow = ovf; feas.score = score;
        feas.found = true; is_new = true;
    }
    if (!feas.found && (!fall.found || ovf < fall.overflow - 1e-12 ||
                        (std::abs(ovf - fall.overflow) <= 1e-12 && e.s.hpwl < fall.hpwl))) {
        fall.positions = pos; fall.iteration = it; fall.hpwl = e.s.hpwl; fall.overflow = ovf; fall.score = score;
        fall.found = true; is_new = true;
    }
}

mp_gp_history_row historyRow(int it, const Eval& e, double penalty, double smoothing, double step, double disp,
                             double gnorm, double curvature, int backtracks, bool restarted, bool accepted, bool best) {  // :619-648
    mp_gp_history_row r{};
    r.iteration = it;
    r.hpwl = e.s.hpwl;
    r.overflow = e.s.optimizer_density.normalized_overflow;
    r.design_overflow = e.s.design_density.normalized_overflow;
    r.smooth_wirelength = e.s.smooth_wirelength;
    r.density_energy = e.s.electrostatic_energy;
    r.objective = e.objective;
    r.penalty = penalty;
    r.smoothing = smoothing;
    r.step_size = step;
    r.maximum_displacement = disp;
    r.gradient_norm = gnorm;
    r.curvature = curvature;
    r.backtracks = backtracks;
    r.momentum_restarted = restarted;
    r.accepted = accepted;
    r.best_checkpoint = best;
    return r;
}

double designHpwl(const Db& db) {  // Metrics.cpp:128-147
    double h = 0;
    for (std::size_t n = 0; n < db.net_weight.size(); ++n) {
        const int64_t b = db.net_start[n], e = db.net_start[n + 1];
        if (e - b < 2) continue;
        double mnx = std::numeric_limits<double>::infinity(), mxx = -mnx, mny = mnx, mxy = -mnx;
        for (int64_t p = b; p < e; ++p) {
            const V2 q = db.pinPosition(static_cast<std::size_t>(p));
            mnx = std::min(mnx, q.x); mxx = std::max(mxx, q.x);
            mny = std::min(mny, q.y); mxy = std::max(mxy, q.y);
        }
        h += db.net_weight[n] * ((mxx - mnx) + (mxy - mny));
    }
    return h;
}

mp_gp_density_metrics designDensity(const Db& db, int cols, int rows, double t) {  // Metrics.cpp:149-164
    DensityMap map(db.core, cols, rows, t);
    if (!db.rows.empty()) map.setPlaceableRegions(db.rows);
    for (std::size_t m = 0; m < db.modules(); ++m) {
        if (db.fixed[m]) map.addRectangle(db.rect(m), kFixed);
        else map.addRectangle(db.rect(m), db.macro[m] ? kMacro : kMovable);
    }
    return map.metrics(true);
}

struct RunOutput {
    mp_gp_result result{};
    std::vector<mp_gp_history_row> history;
};

// :652-953
RunOutput runAdaptive(Session& s, const mp_gp_options& o, double t_started) {
    RunOutput out;
    mp_gp_result& res = out.result;
    validateOptions(s.db, o);
    using clock = std::chrono::steady_clock;
    const auto loop_start = clock::now();

    std::vector<V2> current = s.capturePositions();
    std::vector<V2> velocity(current.size());
    std::vector<V2> prev_grad, prev_pos;
    bool has_prev = false;

    double smoothing = o.initial_smoothing * s.maxBin();
    Eval init = s.evaluate(smoothing, 1.0, 1.0, true);
    const double initial_overflow = std::max(init.s.optimizer_density.normalized_overflow, kEpsilon);
    smoothing = smoothingFromOverflow(o, s.maxBin(), init.s.optimizer_density.normalized_overflow, initial_overflow);
    const double raw_penalty = init.density_l1 > kEpsilon ? init.wire_l1 / init.density_l1 : 1.0;
    double penalty = clampd(raw_penalty, o.minimum_density_penalty, o.maximum_density_penalty);
    const double ref_penalty = penalty;
    init = s.evaluate(smoothing, penalty, ref_penalty, true);

    double last_overflow = init.s.optimizer_density.normalized_overflow;
    double last_step = initialStepSize(init, o.initial_movement_in_bins * s.minBin());
    Checkpoint feas, fall;
    bool dummy = false;
    considerCheckpoint(feas, fall, current, init, 0, o, dummy);

    int momentum_age = 0, worse_streak = 0;
    int feasible_refinement = init.s.design_density.normalized_overflow <= o.maximum_bin_overflow ? 1 : 0;

    for (int it = 0; it < o.iterations; ++it) {
        Eval cur;
        if (it == 0) {
            cur = std::move(init);
        } else {
            smoothing = smoothingFromOverflow(o, s.maxBin(), last_overflow, initial_overflow);
            s.applyPositions(current);
            cur = s.evaluate(smoothing, penalty, ref_penalty, true);
        }
        bool restarted = worse_streak >= 2;
        bool momentum_used = false;
        std::vector<V2> base = current;
        const Eval* base_eval = &cur;
        Eval look;
        if (!restarted && momentum_age > 0) {
            const double beta = std::min(o.maximum_momentum, static_cast<double>(momentum_age) / static_cast<double>(momentum_age + 3));
            double align = 0.0;  // :610-617
            for (std::size_t i = 0; i < velocity.size(); ++i) align -= dot(velocity[i], cur.pre[i]);
            if (align <= 0.0) {
                restarted = true;
            } else {
                for (std::size_t i = 0; i < base.size(); ++i) base[i] = base[i] + velocity[i] * beta;
                s.applyPositions(base);
                base = s.capturePositions();
                look = s.evaluate(smoothing, penalty, ref_penalty, true);
                if (look.objective > cur.objective * (1.0 + o.hpwl_growth_restart_threshold)) {
                    restarted = true;
                } else {
                    base_eval = &look;
                    momentum_used = true;
                }
            }
        }
        if (restarted) {
            s.applyPositions(current);
            base = current;
            base_eval = &cur;
            momentum_age = 0;
            worse_streak = 0;
            ++res.momentum_restarts;
        }

        bool accepted = false, new_best = false;
        int backtracks = 0;
        double used_step = 0, used_curv = 0, used_disp = 0;
        Eval acc_eval;
        std::vector<V2> acc_pos;
        for (int pass = 0; pass < 2 && !accepted; ++pass) {
            const StepEstimate est = estimateStep(base, *base_eval, prev_pos, prev_grad, has_prev, std::max(last_step, kEpsilon));
            double step = est.step;
            const double ratio = clampd(base_eval->s.optimizer_density.normalized_overflow / std::max(initial_overflow, kEpsilon), 0.0, 1.0);
            const double max_disp = o.maximum_movement_in_bins * s.minBin() * (0.35 + 0.65 * ratio);
            for (int attempt = 0; attempt <= o.maximum_backtracks; ++attempt) {
                Proposal prop = proposeStep(s, base, *base_eval, step, max_disp);
                if (prop.predicted <= kEpsilon) {
                    s.applyPositions(base);
                    break;
                }
                Eval cand = s.evaluate(smoothing, penalty, ref_penalty, false);
                if (acceptsCandidate(*base_eval, cand, prop.predicted, o)) {
                    accepted = true;
                    used_step = step;
                    used_curv = est.curvature;
                    used_disp = prop.max_disp;
                    acc_pos = std::move(prop.positions);
                    acc_eval = std::move(cand);
                    break;
                }
                ++backtracks;
                ++res.rejected_candidates;
                step *= o.backtracking_ratio;
                s.applyPositions(base);
            }
            if (accepted || restarted || !momentum_used) break;
            restarted = true;
            ++res.momentum_restarts;
            momentum_age = 0;
            worse_streak = 0;
            s.applyPositions(current);
            base = current;
            base_eval = &cur;
            momentum_used = false;
        }

        const double penalty_used = penalty;
        if (accepted) {
            const std::vector<V2> old = current;
            current = std::move(acc_pos);
            velocity.resize(current.size());
            for (std::size_t i = 0; i < current.size(); ++i) velocity[i] = current[i] - old[i];
            last_step = used_step;
            if (restarted) momentum_age = 0; else ++momentum_age;
            const double growth = (acc_eval.s.hpwl - cur.s.hpwl) / std::max(cur.s.hpwl, 1.0);
            worse_streak = growth > o.hpwl_growth_restart_threshold ? worse_streak + 1 : 0;
            const double next = updateDensityPenalty(penalty, ref_penalty, cur, acc_eval, o);
            const bool changed = std::abs(std::log(next / std::max(penalty, kEpsilon))) > 0.01;
            penalty = next;
            if (changed) {
                prev_grad.clear(); prev_pos.clear(); has_prev = false;
            } else {
                prev_pos = base; prev_grad = base_eval->pre; has_prev = true;
            }
            last_overflow = acc_eval.s.optimizer_density.normalized_overflow;
            considerCheckpoint(feas, fall, current, acc_eval, it + 1, o, new_best);
            ++res.accepted_iterations;
            if (acc_eval.s.design_density.normalized_overflow <= o.maximum_bin_overflow) ++feasible_refinement;
            else feasible_refinement = 0;
            out.history.push_back(historyRow(it + 1, acc_eval, penalty_used, smoothing, used_step, used_disp,
                                             base_eval->pre_norm, used_curv, backtracks, restarted, true, new_best));
        } else {
            s.applyPositions(current);
            std::fill(velocity.begin(), velocity.end(), V2{});
            momentum_age = 0;
            worse_streak = 0;
            if (cur.s.optimizer_density.normalized_overflow > o.maximum_bin_overflow)
                penalty = clampd(penalty * o.density_penalty_increase, std::max(o.minimum_density_penalty, ref_penalty * 0.02),
                                 o.maximum_density_penalty);
            last_overflow = cur.s.optimizer_density.normalized_overflow;
            if (cur.s.design_density.normalized_overflow <= o.maximum_bin_overflow) ++feasible_refinement;
            else feasible_refinement = 0;
            out.history.push_back(historyRow(it + 1, cur, penalty_used, smoothing, 0.0, 0.0, cur.pre_norm, 0.0,
                                             backtracks, restarted, false, false));
        }
        res.completed_iterations = it + 1;
        if (it >= 20 && feasible_refinement >= o.feasible_refinement_iterations) break;
    }
    const Checkpoint& chosen = feas.found ? feas : fall;
    if (chosen.found) {
        s.applyPositions(chosen.positions);
        res.restored_best_checkpoint = 1;
        res.best_checkpoint_iteration = chosen.iteration;
        res.best_checkpoint_hpwl = chosen.hpwl;
        res.best_checkpoint_overflow = chosen.overflow;
    }
    const auto loop_end = clock::now();
    res.loop_seconds = std::chrono::duration<double>(loop_end - loop_start).count();
    res.hpwl_after = designHpwl(s.db);
    res.overflow_after = designDensity(s.db, o.bins_x, o.bins_y, o.target_density).normalized_overflow;
    res.elapsed_seconds = std::chrono::duration<double>(clock::now().time_since_epoch()).count() - t_started;
    res.gradient_evaluations = s.grad_evals;
    res.value_evaluations = s.value_evals;
    res.particle_count = static_cast<int64_t>(s.particles());
    res.filler_count = static_cast<int64_t>(s.fillers.size());
    res.device_used = -1;
    return out;
}

template <typename F>
mp_gp_status guarded(F&& body) {
    try {
        g_error.clear();
        body();
        return MP_GP_OK;
    } catch (const std::invalid_argument& e) {
        g_error = e.what();
        return MP_GP_INVALID_ARGUMENT;
    } catch (const std::out_of_range& e) {
        g_error = e.what();
        return MP_GP_INVALID_ARGUMENT;
    } catch (const std::exception& e) {
        g_error = e.what();
        return MP_GP_RUNTIME_ERROR;
    }
}

double nowSeconds() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace
}  // namespace oracle

namespace oracle {

// ---- quadratic initial placement (InitialPlacer.cpp:266-402) -------------------------------------

// Sparse symmetric system assembled from (row, col, value) entries; entries with equal (row, col)
// are summed in the order they were added, as a triplet assembly does.
struct Triplets {
    std::vector<int64_t> row, col;
    std::vector<double> val;
    void add(int64_t r, int64_t c, double v) {
        row.push_back(r);
        col.push_back(c);
        val.push_back(v);
    }
};

struct Csr {
    std::vector<int64_t> off;
    std::vector<int64_t> col;
    std::vector<double> val;
    std::vector<double> multiply(const std::vector<double>& x) const {
        std::vector<double> y(off.size() - 1, 0.0);
        for (std::size_t r = 0; r + 1 < off.size(); ++r) {
            double a = 0.0;
            for (int64_t k = off[r]; k < off[r + 1]; ++k) a += val[k] * x[col[k]];
            y[r] = a;
        }
        return y;
    }
};

Csr assemble(std::size_t n, const Triplets& t) {
    std::vector<std::size_t> order(t.val.size());
    std::iota(order.begin(), order.end(), 0U);
    std::stable_sort(order.begin(), order.end(), [&](std::size_t a, std::size_t b) {
        return t.row[a] != t.row[b] ? t.row[a] < t.row[b] : t.col[a] < t.col[b];
    });
    Csr m;
    m.off.assign(n + 1, 0);
    for (std::size_t k = 0; k < order.size(); ++k) {
        const std::size_t i = order[k];
        if (!m.col.empty() && k > 0 && t.row[order[k - 1]] == t.row[i] && t.col[order[k - 1]] == t.col[i]) {
            m.val.back() += t.val[i];
            continue;
        }
        m.col.push_back(t.col[i]);
        m.val.push_back(t.val[i]);
        ++m.off[static_cast<std::size_t>(t.row[i]) + 1];
    }
    std::partial_sum(m.off.begin(), m.off.end(), m.off.begin());
    return m;
}

inline double dotv(const std::vector<double>& a, const std::vector<double>& b) {
    double s = 0.0;
    for (std::size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
}

// Unpreconditioned BiCGSTAB (van der Vorst, SIAM J. Sci. Stat. Comput. 13(2), 1992) with the
// conventions of Eigen 3.4's BiCGSTAB, which the reference calls with an identity preconditioner
// (InitialPlacer.cpp:386-390): shadow residual r0 = b - A x0; stop when |r|^2 <= tol^2 |b|^2 or after
// max_iterations; a zero right-hand side gives x = 0; when |rho| < eps^2 |r0|^2 the shadow residual
// is reset to the current residual (and on the first such reset the iteration count restarts);
// omega = 0 when |t| = 0.  Returns the iterations run and the final relative residual.
int bicgstab(const Csr& A, const std::vector<double>& b, std::vector<double>& x, int max_iterations, double tol,
             double& relative_residual) {
    const std::size_t n = b.size();
    const double b2 = dotv(b, b);
    if (b2 == 0.0) {
        std::fill(x.begin(), x.end(), 0.0);
        relative_residual = 0.0;
        return 0;
    }
    std::vector<double> r = A.multiply(x);
    for (std::size_t i = 0; i < n; ++i) r[i] = b[i] - r[i];
    std::vector<double> r0 = r, p(n, 0.0), v(n, 0.0), s(n), t(n);
    double r0_2 = dotv(r0, r0);
    double rho = 1.0, alpha = 1.0, omega = 1.0;
    const double tol2 = tol * tol * b2;
    const double eps = std::numeric_limits<double>::epsilon();
    int i = 0, restarts = 0;
    while (dotv(r, r) > tol2 && i < max_iterations) {
        const double rho_old = rho;
        rho = dotv(r0, r);
        if (std::abs(rho) < eps * eps * r0_2) {
            r = A.multiply(x);
            for (std::size_t k = 0; k < n; ++k) r[k] = b[k] - r[k];
            r0 = r;
            rho = r0_2 = dotv(r, r);
            if (restarts++ == 0) i = 0;
        }
        const double beta = (rho / rho_old) * (alpha / omega);
        for (std::size_t k = 0; k < n; ++k) p[k] = r[k] + beta * (p[k] - omega * v[k]);
        v = A.multiply(p);
        alpha = rho / dotv(r0, v);
        for (std::size_t k = 0; k < n; ++k) s[k] = r[k] - alpha * v[k];
        t = A.multiply(s);
        const double t2 = dotv(t, t);
        omega = t2 > 0.0 ? dotv(t, s) / t2 : 0.0;
        for (std::size_t k = 0; k < n; ++k) {
            x[k] += alpha * p[k] + omega * s[k];
            r[k] = s[k] - omega * t[k];
        }
        ++i;
    }
    relative_residual = std::sqrt(dotv(r, r) / b2);
    return i;
}

struct QuadraticRun {
    int iterations = 0, solver_iterations = 0;
    double residual[2] = {0.0, 0.0};
};

// quadraticPlacement, InitialPlacer.cpp:315-402, with addPairToQuadraticSystem :266-313.
QuadraticRun quadraticPlacement(Db& db, const mp_qp_options& o) {
    QuadraticRun run;
    if (db.movable.empty()) return run;
    const V2 centre = db.core.center();
    for (const int64_t id : db.movable) {
        db.c[static_cast<std::size_t>(id)] = centre;
        db.clampModuleToCore(static_cast<std::size_t>(id));
    }
    const std::size_t n = db.movable.size();
    std::vector<int64_t> index(db.modules(), -1);
    for (std::size_t k = 0; k < n; ++k) index[static_cast<std::size_t>(db.movable[k])] = static_cast<int64_t>(k);
    std::vector<double> xs(n), ys(n);
    for (std::size_t k = 0; k < n; ++k) {
        xs[k] = db.c[static_cast<std::size_t>(db.movable[k])].x;
        ys[k] = db.c[static_cast<std::size_t>(db.movable[k])].y;
    }
    for (int outer = 0; outer < o.outer_iterations; ++outer) {
        Triplets tx, ty;
        std::vector<double> bx(n, 0.0), by(n, 0.0);
        auto addPair = [&](int64_t lp, int64_t rp, double base) {  // :266-313
            const int lm = db.pin_module[static_cast<std::size_t>(lp)], rm = db.pin_module[static_cast<std::size_t>(rp)];
            if (lm == rm) return;
            const int64_t li = index[static_cast<std::size_t>(lm)], ri = index[static_cast<std::size_t>(rm)];
            if (li < 0 && ri < 0) return;
            const V2 lo = transformOffset(db.pin_offset[static_cast<std::size_t>(lp)], db.orient[static_cast<std::size_t>(lm)]);
            const V2 ro = transformOffset(db.pin_offset[static_cast<std::size_t>(rp)], db.orient[static_cast<std::size_t>(rm)]);
            const V2 lpos = db.c[static_cast<std::size_t>(lm)] + lo, rpos = db.c[static_cast<std::size_t>(rm)] + ro;
            const double wx = base / std::max(std::abs(lpos.x - rpos.x), o.minimum_distance);
            const double wy = base / std::max(std::abs(lpos.y - rpos.y), o.minimum_distance);
            if (li >= 0 && ri >= 0) {
                tx.add(li, li, wx); ty.add(li, li, wy);
                tx.add(ri, ri, wx); ty.add(ri, ri, wy);
                tx.add(li, ri, -wx); tx.add(ri, li, -wx);
                ty.add(li, ri, -wy); ty.add(ri, li, -wy);
                const V2 d = lo - ro;
                bx[li] -= wx * d.x; bx[ri] += wx * d.x;
                by[li] -= wy * d.y; by[ri] += wy * d.y;
            } else if (li >= 0) {
                tx.add(li, li, wx); ty.add(li, li, wy);
                bx[li] += wx * (db.c[static_cast<std::size_t>(rm)].x + ro.x - lo.x);
                by[li] += wy * (db.c[static_cast<std::size_t>(rm)].y + ro.y - lo.y);
            } else {
                tx.add(ri, ri, wx); ty.add(ri, ri, wy);
                bx[ri] += wx * (db.c[static_cast<std::size_t>(lm)].x + lo.x - ro.x);
                by[ri] += wy * (db.c[static_cast<std::size_t>(lm)].y + lo.y - ro.y);
            }
        };
        for (std::size_t net = 0; net < db.net_weight.size(); ++net) {
            const int64_t b = db.net_start[net], e = db.net_start[net + 1];
            const int64_t degree = e - b;
            if (degree < 2) continue;
            const double w = db.net_weight[net];
            if (static_cast<uint64_t>(degree) <= o.clique_degree_limit) {
                const double cw = 2.0 * w / static_cast<double>(degree - 1);
                for (int64_t l = b; l < e; ++l)
                    for (int64_t r = l + 1; r < e; ++r) addPair(l, r, cw);
            } else {
                const double sw = w / static_cast<double>(degree - 1);
                for (int64_t k = b + 1; k < e; ++k) addPair(b, k, sw);
            }
        }
        constexpr double kAnchor = 1e-4;
        for (std::size_t k = 0; k < n; ++k) {
            const V2 c = db.c[static_cast<std::size_t>(db.movable[k])];
            tx.add(static_cast<int64_t>(k), static_cast<int64_t>(k), kAnchor);
            ty.add(static_cast<int64_t>(k), static_cast<int64_t>(k), kAnchor);
            bx[k] += kAnchor * c.x;
            by[k] += kAnchor * c.y;
        }
        const Csr ax = assemble(n, tx), ay = assemble(n, ty);
        run.solver_iterations += bicgstab(ax, bx, xs, o.solver_iterations, o.tolerance, run.residual[0]);
        run.solver_iterations += bicgstab(ay, by, ys, o.solver_iterations, o.tolerance, run.residual[1]);
        for (std::size_t k = 0; k < n; ++k)
            if (!std::isfinite(xs[k]) || !std::isfinite(ys[k]))
                throw std::runtime_error("Quadratic placement failed to produce finite coordinates.");
        for (std::size_t k = 0; k < n; ++k) {
            const std::size_t id = static_cast<std::size_t>(db.movable[k]);
            db.c[id] = {xs[k], ys[k]};
            db.clampModuleToCore(id);
            xs[k] = db.c[id].x;
            ys[k] = db.c[id].y;
        }
        run.iterations = outer + 1;
    }
    return run;
}

// ---------------------------------------------------------------------------------------------
// Detailed placement, window reordering as the reference's device pass runs it
// (src/placement/cuda/CudaDetailedPlacementBackend.cu:523-711): per pass, row segments and
// four-cell windows at the current placement; every eligible window's 24 orderings scored
// against the placement at the start of the pass (the device kernel :144-220, restated here on
// the CPU); then, in window order, the best ordering re-scored on the current placement and
// applied when it improves by more than epsilon; then the other windows searched exhaustively
// (optimizeCpuWindow :400-424, the same search as DetailedPlacer.cpp:186-236).
struct DpRow {
    double bottom, height, x_start, spacing;
    int orientation;
};

struct DpWork {
    Db db;
    std::vector<DpRow> rows;
    std::vector<std::size_t> by_bottom;               // sortedRows, DetailedPlacer.cpp:80-87
    std::vector<std::vector<std::size_t>> pins_of;    // Module::pins
    std::vector<std::size_t> net_of;                  // Pin::net

    static double alignUp(double value, double origin, double step) {  // DetailedPlacer.cpp:73-76
        if (step <= kEpsilon) return value;
        return origin + std::ceil((value - origin - kEpsilon) / step) * step;
    }
    double rowWidth(std::size_t m, std::size_t r) const { return alignUp(db.w[m], 0.0, rows[r].spacing); }
    double llx(std::size_t m) const { return db.c[m].x - db.w[m] * 0.5; }
    double lly(std::size_t m) const { return db.c[m].y - db.h[m] * 0.5; }

    std::size_t nearestRow(double y) const {  // DetailedPlacer.cpp:89-104
        auto it = std::lower_bound(by_bottom.begin(), by_bottom.end(), y, [&](std::size_t r, double v) {
            return rows[r].bottom + rows[r].height * 0.5 < v;
        });
        if (it == by_bottom.begin()) return *it;
        if (it == by_bottom.end()) return by_bottom.back();
        const std::size_t right = *it, left = *(it - 1);
        const double dl = std::abs(rows[left].bottom + rows[left].height * 0.5 - y);
        const double dr = std::abs(rows[right].bottom + rows[right].height * 0.5 - y);
        return dl <= dr ? left : right;
    }

    // buildRowSegments, DetailedPlacer.cpp:106-145
    std::vector<std::pair<std::size_t, std::vector<std::size_t>>> segments(double eps) const {
        std::vector<std::pair<std::size_t, std::vector<std::size_t>>> out;
        if (rows.empty()) return out;
        std::vector<std::vector<std::size_t>> cells(rows.size());
        for (std::size_t m = 0; m < db.modules(); ++m) {
            if (db.fixed[m] || db.macro[m]) continue;
            const std::size_t r = nearestRow(db.c[m].y);
            if (std::abs(lly(m) - rows[r].bottom) > eps) continue;
            cells[r].push_back(m);
        }
        for (std::size_t r = 0; r < cells.size(); ++r) {
            std::vector<std::size_t>& v = cells[r];
            if (v.size() < 2U) continue;
            std::sort(v.begin(), v.end(), [&](std::size_t a, std::size_t b) { return llx(a) < llx(b); });
            std::vector<std::size_t> cur;
            for (const std::size_t m : v) {
                if (!cur.empty() && std::abs(llx(m) - (llx(cur.back()) + rowWidth(cur.back(), r))) > eps) {
                    if (cur.size() >= 2U) out.emplace_back(r, cur);
                    cur.clear();
                }
                cur.push_back(m);
            }
            if (cur.size() >= 2U) out.emplace_back(r, cur);
        }
        return out;
    }

    std::vector<std::size_t> netsOf(const std::vector<std::size_t>& window) const {
        std::vector<std::size_t> nets;
        for (const std::size_t m : window)
            for (const std::size_t p : pins_of[m]) nets.push_back(net_of[p]);
        std::sort(nets.begin(), nets.end());
        nets.erase(std::unique(nets.begin(), nets.end()), nets.end());
        return nets;
    }
    std::size_t degree(std::size_t n) const { return static_cast<std::size_t>(db.net_start[n + 1] - db.net_start[n]); }

    // Window HPWL with the window's cells at `at` (centres) and pin offsets oriented by `orient`.
    double windowHpwl(const std::vector<std::size_t>& window, const std::vector<std::size_t>& nets,
                      const std::vector<V2>& at, const std::vector<V2>& centres, const std::vector<int>& orient) const {
        double total = 0.0;
        for (const std::size_t n : nets) {
            double lox = std::numeric_limits<double>::infinity(), hix = -lox, loy = lox, hiy = -lox;
            for (int64_t p = db.net_start[n]; p < db.net_start[n + 1]; ++p) {
                const std::size_t m = static_cast<std::size_t>(db.pin_module[static_cast<std::size_t>(p)]);
                V2 c = centres[m];
                for (std::size_t k = 0; k < window.size(); ++k)
                    if (window[k] == m) {
                        c = at[k];
                        break;
                    }
                const V2 q = c + transformOffset(db.pin_offset[static_cast<std::size_t>(p)], orient[m]);
                lox = std::min(lox, q.x);
                hix = std::max(hix, q.x);
                loy = std::min(loy, q.y);
                hiy = std::max(hiy, q.y);
            }
            total += db.net_weight[n] * ((hix - lox) + (hiy - loy));
        }
        return total;
    }

    // centersForOrder from the window's first lower-left x in `centres`
    std::vector<V2> packed(const std::vector<std::size_t>& window, std::size_t r, const std::vector<std::size_t>& order,
                           const std::vector<V2>& centres) const {
        std::vector<V2> at(window.size());
        double cursor = centres[window.front()].x - db.w[window.front()] * 0.5;
        for (const std::size_t local : order) {
            const std::size_t m = window[local];
            at[local] = {cursor + db.w[m] * 0.5, rows[r].bottom + db.h[m] * 0.5};
            cursor += rowWidth(m, r);
        }
        return at;
    }

    void apply(const std::vector<std::size_t>& window, std::size_t r, const std::vector<std::size_t>& order) {
        const std::vector<V2> at = packed(window, r, order, db.c);
        for (std::size_t k = 0; k < window.size(); ++k) {
            db.c[window[k]] = at[k];
            if (rows[r].orientation != 8) db.orient[window[k]] = rows[r].orientation;
        }
    }
};

mp_dp_result detailedPlace(DpWork& w, const mp_dp_options& o) {
    mp_dp_result r{};
    r.device_used = -1;
    r.hpwl_before = designHpwl(w.db);
    if (o.method == MP_DP_NONE || w.rows.empty()) {
        r.hpwl_after = r.hpwl_before;
        return r;
    }
    if (o.method != MP_DP_WINDOW_REORDER || o.window_size != 4)
        throw std::runtime_error("the device pass runs four-cell window reordering only");
    r.device_used = o.device;
    for (int pass = 0; pass < o.passes; ++pass) {
        const auto segs = w.segments(1e-6);
        const std::size_t phase = static_cast<std::size_t>((pass % 2) * 2);
        struct Win {
            std::size_t row;
            std::vector<std::size_t> cells, nets;
        };
        std::vector<Win> dev, cpu;
        for (const auto& seg : segs) {
            for (std::size_t start = phase; start + 1U < seg.second.size(); start += 4) {
                const std::size_t count = std::min<std::size_t>(4, seg.second.size() - start);
                if (count < 2U) break;
                Win win{seg.first, std::vector<std::size_t>(seg.second.begin() + static_cast<std::ptrdiff_t>(start),
                                                             seg.second.begin() + static_cast<std::ptrdiff_t>(start + count)),
                        {}};
                win.nets = w.netsOf(win.cells);
                bool eligible = win.cells.size() == 4U && !win.nets.empty();
                for (const std::size_t n : win.nets)
                    if (w.degree(n) > o.maximum_net_degree) eligible = false;
                (eligible ? dev : cpu).push_back(std::move(win));
            }
        }
        // scores of every ordering at the placement of the start of the pass
        const std::vector<V2> start_c = w.db.c;
        const std::vector<int> start_o = w.db.orient;
        std::vector<int> best(dev.size(), 0);
        for (std::size_t k = 0; k < dev.size(); ++k) {
            std::vector<std::size_t> order{0, 1, 2, 3};
            double best_value = std::numeric_limits<double>::infinity();
            int index = 0;
            do {
                const double v = w.windowHpwl(dev[k].cells, dev[k].nets, w.packed(dev[k].cells, dev[k].row, order, start_c),
                                              start_c, start_o);
                if (v < best_value) {  // strictly lower: ties keep the lower index
                    best_value = v;
                    best[k] = index;
                }
                ++index;
            } while (std::next_permutation(order.begin(), order.end()));
        }
        for (std::size_t k = 0; k < dev.size(); ++k) {
            ++r.evaluated_windows;
            r.evaluated_permutations += 24;
            if (best[k] <= 0) continue;
            std::vector<std::size_t> order{0, 1, 2, 3};
            for (int s = 0; s < best[k]; ++s) std::next_permutation(order.begin(), order.end());
            std::vector<V2> here(4);
            for (std::size_t j = 0; j < 4; ++j) here[j] = w.db.c[dev[k].cells[j]];
            const double before = w.windowHpwl(dev[k].cells, dev[k].nets, here, w.db.c, w.db.orient);
            const double after =
                w.windowHpwl(dev[k].cells, dev[k].nets, w.packed(dev[k].cells, dev[k].row, order, w.db.c), w.db.c, w.db.orient);
            if (after < before - o.improvement_epsilon) {
                w.apply(dev[k].cells, dev[k].row, order);
                ++r.accepted_operations;
            }
        }
        for (const Win& win : cpu) {
            ++r.evaluated_windows;
            bool skip = win.nets.empty();
            for (const std::size_t n : win.nets)
                if (w.degree(n) > o.maximum_net_degree) skip = true;
            if (skip) continue;
            std::vector<V2> here(win.cells.size());
            for (std::size_t j = 0; j < win.cells.size(); ++j) here[j] = w.db.c[win.cells[j]];
            double best_value = w.windowHpwl(win.cells, win.nets, here, w.db.c, w.db.orient);
            std::vector<std::size_t> order(win.cells.size()), best_order;
            std::iota(order.begin(), order.end(), std::size_t{0});
            do {
                const double v = w.windowHpwl(win.cells, win.nets, w.packed(win.cells, win.row, order, w.db.c), w.db.c,
                                              w.db.orient);
                ++r.evaluated_permutations;
                if (v < best_value - o.improvement_epsilon) {
                    best_value = v;
                    best_order = order;
                }
            } while (std::next_permutation(order.begin(), order.end()));
            if (!best_order.empty()) {
                w.apply(win.cells, win.row, best_order);
                ++r.accepted_operations;
            }
        }
        r.completed_passes = pass + 1;
    }
    r.hpwl_after = designHpwl(w.db);
    return r;
}

}  // namespace oracle

using namespace oracle;

struct orc_session {
    std::unique_ptr<Session> s;
    mp_gp_options options;
};

extern "C" {

const char* orc_last_error(void) { return g_error.c_str(); }

void orc_set_method(int method) { g_method = method; }

// One 1-D transform (kind: 0 REDFT10, 1 REDFT01, 2 RODFT01).
int orc_transform_1d(int kind, int n, const double* in, double* out) {
    return guarded([&] {
        if (n < 1 || kind < 0 || kind > 2) throw std::invalid_argument("bad transform");
        Transformer t(n);
        t.run(kind, in, out);
    });
}

// 2-D transform over a rows x cols row-major grid (FFTW r2r_2d convention).
int orc_transform_2d(int cols, int rows, int kind_y, int kind_x, const double* in, double* out) {
    return guarded([&] {
        Transformer tx(cols), ty(rows);
        transform2d(cols, rows, kind_y, kind_x, in, out, tx, ty);
    });
}

// ElectrostaticField::solve (Neumann) on a standalone grid.  region: llx, lly, urx, ury.
int orc_field_solve(int cols, int rows, const double* region, const double* source, double* potential,
                    double* field_x, double* field_y) {
    return guarded([&] {
        Field f(cols, rows, {{region[0], region[1]}, {region[2], region[3]}});
        std::vector<double> src(source, source + f.count);
        f.solve(src);
        std::copy(f.potential.begin(), f.potential.end(), potential);
        std::copy(f.fx.begin(), f.fx.end(), field_x);
        std::copy(f.fy.begin(), f.fy.end(), field_y);
    });
}

// ElectrostaticField::sampleField on given field grids (n points, xy interleaved).
int orc_sample_field(int cols, int rows, const double* region, const double* field_x, const double* field_y,
                     int64_t n, const double* xy, double* out) {
    return guarded([&] {
        Field f(cols, rows, {{region[0], region[1]}, {region[2], region[3]}});
        std::copy(field_x, field_x + f.count, f.fx.begin());
        std::copy(field_y, field_y + f.count, f.fy.begin());
        for (int64_t i = 0; i < n; ++i) {
            const V2 e = f.sampleField({xy[2 * i], xy[2 * i + 1]});
            out[2 * i] = e.x;
            out[2 * i + 1] = e.y;
        }
    });
}

// DensityMap known-answer driver: regions [4R], rects [4K] with layers [K] (0 fixed, 1 movable,
// 2 macro, 3 filler, 4 dark); outputs metrics with/without fillers and bins [5G]
// (fixed, movable, macro, filler, dark).
int orc_density_map(const double* region, int cols, int rows, double t, int64_t n_regions, const double* regions,
                    int64_t n_rects, const double* rects, const int32_t* layers, mp_gp_density_metrics* with_f,
                    mp_gp_density_metrics* without_f, double* bins) {
    return guarded([&] {
        DensityMap m({{region[0], region[1]}, {region[2], region[3]}}, cols, rows, t);
        std::vector<Rc> regs;
        for (int64_t i = 0; i < n_regions; ++i)
            regs.push_back({{regions[4 * i], regions[4 * i + 1]}, {regions[4 * i + 2], regions[4 * i + 3]}});
        if (n_regions > 0) m.setPlaceableRegions(regs);
        for (int64_t i = 0; i < n_rects; ++i)
            m.addRectangle({{rects[4 * i], rects[4 * i + 1]}, {rects[4 * i + 2], rects[4 * i + 3]}},
                           static_cast<Layer>(layers[i]));
        if (with_f) *with_f = m.metrics(true);
        if (without_f) *without_f = m.metrics(false);
        if (bins)
            for (std::size_t i = 0; i < m.bins.size(); ++i) {
                bins[5 * i] = m.bins[i].fixed;
                bins[5 * i + 1] = m.bins[i].movable;
                bins[5 * i + 2] = m.bins[i].macro;
                bins[5 * i + 3] = m.bins[i].filler;
                bins[5 * i + 4] = m.bins[i].dark;
            }
    });
}

int orc_random_placement(const mp_design* design, uint64_t seed, double* centers) {  // InitialPlacer.cpp:102-120
    return guarded([&] {
        Db db = loadDesign(design);
        std::mt19937_64 gen(seed);
        for (const int64_t id : db.movable) {
            const double lx = db.core.ll.x + db.w[id] * 0.5, ux = db.core.ur.x - db.w[id] * 0.5;
            const double ly = db.core.ll.y + db.h[id] * 0.5, uy = db.core.ur.y - db.h[id] * 0.5;
            if (lx <= ux) {
                std::uniform_real_distribution<double> d(lx, ux);
                db.c[id].x = d(gen);
            }
            if (ly <= uy) {
                std::uniform_real_distribution<double> d(ly, uy);
                db.c[id].y = d(gen);
            }
            db.clampModuleToCore(static_cast<std::size_t>(id));
        }
        for (std::size_t m = 0; m < db.modules(); ++m) {
            centers[2 * m] = db.c[m].x;
            centers[2 * m + 1] = db.c[m].y;
        }
    });
}

// InitialPlacer::run with the quadratic method (InitialPlacer.cpp:315-402, 423-436).
int orc_quadratic_placement(const mp_design* design, const mp_qp_options* options, double* centers,
                            mp_qp_result* result) {
    return guarded([&] {
        if (!options || !centers) throw std::invalid_argument("null options or centres");
        const mp_qp_options& o = *options;
        if (o.outer_iterations <= 0 || o.solver_iterations <= 0 || !std::isfinite(o.tolerance) || o.tolerance <= 0.0 ||
            !std::isfinite(o.minimum_distance) || o.minimum_distance <= 0.0)  // :52-57
            throw std::invalid_argument("Quadratic initial-placement options are outside their valid range.");
        const auto t0 = std::chrono::steady_clock::now();
        Db db = loadDesign(design);
        mp_qp_result r{};
        r.hpwl_before = designHpwl(db);
        const QuadraticRun run = quadraticPlacement(db, o);
        r.hpwl_after = designHpwl(db);
        r.iterations = run.iterations;
        r.solver_iterations = run.solver_iterations;
        r.relative_residual[0] = run.residual[0];
        r.relative_residual[1] = run.residual[1];
        r.elapsed_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        for (std::size_t m = 0; m < db.modules(); ++m) {
            centers[2 * m] = db.c[m].x;
            centers[2 * m + 1] = db.c[m].y;
        }
        if (result) *result = r;
    });
}

// Window-reordering detailed placement as the reference's device pass runs it (see detailedPlace).
int orc_detailed_place(const mp_design* design, const int32_t* row_orientation, const mp_dp_options* options,
                       double* centers, int32_t* orientation, mp_dp_result* result) {
    return guarded([&] {
        if (!options || !centers || !result) throw std::invalid_argument("null options, centres or result");
        const mp_dp_options& o = *options;
        if (o.method < MP_DP_NONE || o.method > MP_DP_WINDOW_REORDER || !std::isfinite(o.improvement_epsilon) ||
            o.improvement_epsilon < 0.0 ||
            (o.method != MP_DP_NONE && (o.passes <= 0 || o.maximum_net_degree < 2U)))  // DetailedPlacer.cpp:50-71
            throw std::invalid_argument("Detailed placement options are outside their valid range.");
        DpWork w;
        w.db = loadDesign(design);
        for (int64_t r = 0; r < design->row_count; ++r) {
            const double* row = design->rows + 5 * r;
            w.rows.push_back({row[0], row[1], row[2], row[3], row_orientation ? row_orientation[r] : 0});
        }
        w.by_bottom.resize(w.rows.size());
        std::iota(w.by_bottom.begin(), w.by_bottom.end(), std::size_t{0});
        std::sort(w.by_bottom.begin(), w.by_bottom.end(),
                  [&](std::size_t a, std::size_t b) { return w.rows[a].bottom < w.rows[b].bottom; });
        w.pins_of.assign(w.db.modules(), {});
        w.net_of.assign(w.db.pin_module.size(), 0);
        for (std::size_t n = 0; n + 1 < w.db.net_start.size(); ++n)
            for (int64_t p = w.db.net_start[n]; p < w.db.net_start[n + 1]; ++p) w.net_of[static_cast<std::size_t>(p)] = n;
        for (std::size_t p = 0; p < w.db.pin_module.size(); ++p) w.pins_of[static_cast<std::size_t>(w.db.pin_module[p])].push_back(p);
        const auto t0 = std::chrono::steady_clock::now();
        mp_dp_result r = detailedPlace(w, o);
        r.elapsed_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        for (std::size_t m = 0; m < w.db.modules(); ++m) {
            centers[2 * m] = w.db.c[m].x;
            centers[2 * m + 1] = w.db.c[m].y;
            if (orientation) orientation[m] = w.db.orient[m];
        }
        *result = r;
    });
}

int orc_design_metrics(const mp_design* design, int cols, int rows, double t, double* hpwl,
                       mp_gp_density_metrics* density) {
    return guarded([&] {
        Db db = loadDesign(design);
        if (hpwl) *hpwl = designHpwl(db);
        if (density) *density = designDensity(db, cols, rows, t);
    });
}

int orc_session_create(const mp_design* design, const mp_gp_options* options, orc_session** out) {
    return guarded([&] {
        auto h = std::make_unique<orc_session>();
        h->options = *options;
        h->s = std::make_unique<Session>(design, options);
        *out = h.release();
    });
}

void orc_session_destroy(orc_session* h) { delete h; }

int orc_session_info(const orc_session* h, mp_gp_prepared_info* info) {
    return guarded([&] {
        const Session& s = *h->s;
        *info = {};
        info->movable_count = static_cast<int64_t>(s.db.movable.size());
        info->filler_count = static_cast<int64_t>(s.fillers.size());
        info->particle_count = static_cast<int64_t>(s.particles());
        info->net_count = static_cast<int64_t>(s.net_w.size());
        info->pin_count = static_cast<int64_t>(s.pin_particles.size());
        info->bin_count = static_cast<int64_t>(s.map.bins.size());
        info->core_ll_x = s.db.core.ll.x; info->core_ll_y = s.db.core.ll.y;
        info->core_ur_x = s.db.core.ur.x; info->core_ur_y = s.db.core.ur.y;
        info->nominal_row_height = s.db.nominal_row_height;
        info->filler_width = s.fillers.empty() ? 0.0 : s.fillers[0].w;
        info->filler_height = s.fillers.empty() ? 0.0 : s.fillers[0].h;
    });
}

// Layout arrays (pointers into the session) and static grids.
int orc_session_layout(const orc_session* h, mp_gp_static* L) {
    return guarded([&] {
        const Session& s = *h->s;
        *L = {};
        L->particle_count = static_cast<int64_t>(s.particles());
        L->particle_width = s.pw.data();
        L->particle_height = s.ph.data();
        L->particle_charge_area = s.pq.data();
        L->particle_layer = s.players.data();
        L->net_count = static_cast<int64_t>(s.net_w.size());
        L->net_offsets = s.net_offsets.data();
        L->net_weights = s.net_w.data();
        L->pin_count = static_cast<int64_t>(s.pin_particles.size());
        L->pin_particles = s.pin_particles.data();
        L->pin_offsets_or_static = s.pos_static.data();
        L->bins_x = s.bins_x;
        L->bins_y = s.bins_y;
        L->region_ll_x = s.db.core.ll.x; L->region_ll_y = s.db.core.ll.y;
        L->region_ur_x = s.db.core.ur.x; L->region_ur_y = s.db.core.ur.y;
        L->target_density = s.t;
    });
}

int orc_session_static_grids(const orc_session* h, double* fixed, double* dark) {
    return guarded([&] {
        const Session& s = *h->s;
        for (std::size_t i = 0; i < s.map.bins.size(); ++i) {
            fixed[i] = s.map.bins[i].fixed;
            dark[i] = s.map.bins[i].dark;
        }
    });
}

int orc_session_arrays(const orc_session* h, int64_t* movable_ids, double* filler_centers, double* precond,
                       double* positions) {
    return guarded([&] {
        const Session& s = *h->s;
        if (movable_ids) std::copy(s.db.movable.begin(), s.db.movable.end(), movable_ids);
        if (filler_centers)
            for (std::size_t i = 0; i < s.fillers.size(); ++i) {
                filler_centers[2 * i] = s.fillers[i].c.x;
                filler_centers[2 * i + 1] = s.fillers[i].c.y;
            }
        if (precond) std::copy(s.precond.begin(), s.precond.end(), precond);
        if (positions) {
            const auto p = s.capturePositions();
            for (std::size_t i = 0; i < p.size(); ++i) {
                positions[2 * i] = p[i].x;
                positions[2 * i + 1] = p[i].y;
            }
        }
    });
}

// Backend-shaped evaluate at the given particle positions (not clamped).
int orc_session_evaluate(orc_session* h, const double* xy, double smoothing, int want_grad, mp_gp_scalars* out,
                         double* wire_grad, double* density_grad) {
    return guarded([&] {
        Session& s = *h->s;
        if (!std::isfinite(smoothing) || smoothing <= 0.0)
            throw std::invalid_argument("Global placement requires positive finite smoothing.");
        std::vector<V2> p(s.particles());
        for (std::size_t i = 0; i < p.size(); ++i) p[i] = {xy[2 * i], xy[2 * i + 1]};
        s.setRaw(p);
        Eval e;
        s.evaluateRaw(smoothing, want_grad != 0, e);
        *out = e.s;
        if (want_grad) {
            for (std::size_t i = 0; i < p.size(); ++i) {
                if (wire_grad) { wire_grad[2 * i] = e.wire[i].x; wire_grad[2 * i + 1] = e.wire[i].y; }
                if (density_grad) { density_grad[2 * i] = e.density[i].x; density_grad[2 * i + 1] = e.density[i].y; }
            }
        }
    });
}

// Grids of the last evaluate: centred source (deviation - mean), potential, field x/y.
int orc_session_grids(const orc_session* h, double* source, double* potential, double* fx, double* fy) {
    return guarded([&] {
        const Field& f = *h->s->field;
        if (source) std::copy(f.source.begin(), f.source.end(), source);
        if (potential) std::copy(f.potential.begin(), f.potential.end(), potential);
        if (fx) std::copy(f.fx.begin(), f.fx.end(), fx);
        if (fy) std::copy(f.fy.begin(), f.fy.end(), fy);
    });
}

// Whole adaptive run on the CPU.  centers [2M] in/out.
int orc_global_place(const mp_design* design, const mp_gp_options* options, double* centers, mp_gp_result* result,
                     mp_gp_history_row* history, int64_t capacity, int64_t* rows_written) {
    return guarded([&] {
        const double started = nowSeconds();
        std::vector<double> c(centers, centers + 2 * design->module_count);
        mp_design d = *design;
        d.center = c.data();
        {
            Db probe = loadDesign(&d);
            if (probe.movable.empty()) {
                *result = {};
                if (rows_written) *rows_written = 0;
                return;
            }
            validateOptions(probe, *options);
        }
        const double hpwl_before = designHpwl(loadDesign(&d));
        const double ovf_before = designDensity(loadDesign(&d), options->bins_x, options->bins_y, options->target_density).normalized_overflow;
        Session s(&d, options);
        RunOutput out = runAdaptive(s, *options, started);
        out.result.hpwl_before = hpwl_before;
        out.result.overflow_before = ovf_before;
        *result = out.result;
        for (std::size_t m = 0; m < s.db.modules(); ++m) {
            centers[2 * m] = s.db.c[m].x;
            centers[2 * m + 1] = s.db.c[m].y;
        }
        const int64_t n = std::min<int64_t>(capacity, static_cast<int64_t>(out.history.size()));
        if (history)
            for (int64_t i = 0; i < n; ++i) history[i] = out.history[static_cast<std::size_t>(i)];
        if (rows_written) *rows_written = static_cast<int64_t>(out.history.size());
    });
}

// The controller's scalar laws, one at a time, so that tests can hold them to objectives whose
// answers are known in closed form.  Vectors are interleaved (x, y) pairs.
double orc_estimate_step(const double* pos, const double* pre, const double* prev_pos, const double* prev_grad,
                         int64_t n, int has_prev, double fallback) {  // AdaptiveGlobalPlacer.cpp:470-492
    std::vector<V2> p(static_cast<std::size_t>(n)), pp(p.size()), pg(p.size());
    Eval e;
    e.pre.resize(p.size());
    for (std::size_t i = 0; i < p.size(); ++i) {
        p[i] = V2{pos[2 * i], pos[2 * i + 1]};
        e.pre[i] = V2{pre[2 * i], pre[2 * i + 1]};
        pp[i] = V2{prev_pos[2 * i], prev_pos[2 * i + 1]};
        pg[i] = V2{prev_grad[2 * i], prev_grad[2 * i + 1]};
    }
    return estimateStep(p, e, pp, pg, has_prev != 0, fallback).step;
}

double orc_smoothing(const mp_gp_options* o, double max_bin, double overflow, double initial) {  // :456-462
    return smoothingFromOverflow(*o, max_bin, overflow, initial);
}

double orc_update_penalty(const mp_gp_options* o, double penalty, double reference_penalty, double before_hpwl,
                          double before_overflow, double after_hpwl, double after_overflow,
                          double after_design_overflow) {  // :536-557
    Eval b, a;
    b.s.hpwl = before_hpwl;
    b.s.optimizer_density.normalized_overflow = before_overflow;
    a.s.hpwl = after_hpwl;
    a.s.optimizer_density.normalized_overflow = after_overflow;
    a.s.design_density.normalized_overflow = after_design_overflow;
    return updateDensityPenalty(penalty, reference_penalty, b, a, *o);
}

// considerCheckpoint (:563-596) over a sequence of (HPWL, design overflow) records, iterations 0..n-1:
// new_best[i] receives is_new_best of record i; returns the iteration runAdaptiveGlobalPlacement
// restores (:930-935), -1 when none; *feasible: whether it is the feasible checkpoint.
int32_t orc_checkpoint_sequence(const mp_gp_options* o, int64_t n, const double* hpwl, const double* design_overflow,
                                int32_t* new_best, int32_t* feasible) {
    Checkpoint feas, fall;
    const std::vector<V2> none;
    for (int64_t i = 0; i < n; ++i) {
        Eval e;
        e.s.hpwl = hpwl[i];
        e.s.design_density.normalized_overflow = design_overflow[i];
        bool is_new = false;
        considerCheckpoint(feas, fall, none, e, static_cast<int>(i), *o, is_new);
        if (new_best) new_best[i] = is_new ? 1 : 0;
    }
    const Checkpoint& chosen = feas.found ? feas : fall;
    if (feasible) *feasible = feas.found ? 1 : 0;
    return chosen.found ? chosen.iteration : -1;
}

int orc_accepts(const mp_gp_options* o, double base_objective, double base_overflow, double cand_objective,
                double cand_overflow, double predicted) {  // :521-534
    Eval b, c;
    b.objective = base_objective;
    b.s.optimizer_density.normalized_overflow = base_overflow;
    c.objective = cand_objective;
    c.s.optimizer_density.normalized_overflow = cand_overflow;
    return acceptsCandidate(b, c, predicted, *o) ? 1 : 0;
}

}  // extern "C"
