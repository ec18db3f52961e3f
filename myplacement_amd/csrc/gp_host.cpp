// gp_host.cpp — host-side set-up of a placement run: the parts of GlobalPlacementSession's
// constructor that feed the backend (static density, fillers, flattened layout, preconditioner),
// the random initial placement, and option defaults.
//
// Reference:
//   refreshDerivedData            src/model/PlacementDatabase.cpp:162-200
//   transformOffset / pinPosition src/model/PlacementDatabase.cpp:55-68, 138-142
//   DensityMap static layers      src/metrics/Metrics.cpp:10-99, GlobalPlacementSupport.cpp:21-30
//   createDensityFillers          src/placement/GlobalPlacementSupport.cpp:52-103
//   buildHostLayout               src/placement/cuda/CudaPlacementBackend.cu:784-839
//   initializePreconditioner      src/placement/AdaptiveGlobalPlacer.cpp:358-375
//   randomPlacement               src/placement/InitialPlacer.cpp:102-120
// The filler and random-placement centres use std::mt19937_64 and
// std::uniform_real_distribution<double> exactly as the reference does, so they are bit-identical
// to the reference's under the same standard library.

#include "gp_host.hpp"

#include <malloc.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <random>
#include <stdexcept>
#include <utility>

namespace mpgp {

namespace {
constexpr double kEps = 1e-9;  // core/Geometry.hpp:8
constexpr std::size_t kMaximumBinCount = 1U << 20U;

double clampd(double v, double lo, double hi) { return std::max(lo, std::min(v, hi)); }

struct Box {
    double lx, ly, ux, uy;
    double area() const { return std::max(0.0, ux - lx) * std::max(0.0, uy - ly); }
    Box clip(const Box& o) const {
        return {std::max(lx, o.lx), std::max(ly, o.ly), std::min(ux, o.ux), std::min(uy, o.uy)};
    }
};

Box boxAround(double cx, double cy, double w, double h) {
    return {cx - w * 0.5, cy - h * 0.5, cx + w * 0.5, cy + h * 0.5};
}

}  // namespace

void orient(double& x, double& y, int o) {  // PlacementDatabase.cpp:55-68
    const double ox = x, oy = y;
    switch (o) {
        case 1: x = -ox; y = -oy; break;
        case 2: x = -oy; y = ox; break;
        case 3: x = oy; y = -ox; break;
        case 4: x = -ox; y = oy; break;
        case 5: x = ox; y = -oy; break;
        case 6: x = oy; y = ox; break;
        case 7: x = -oy; y = -ox; break;
        default: break;
    }
}

namespace {

// Ascending sort of positive finite doubles: their bit patterns order as the values do, so an LSD
// radix sort on them (16 bits per pass) gives std::sort's sequence in linear time.
void sortPositive(std::vector<double>& v) {
    const std::size_t n = v.size();
    std::vector<std::uint64_t> a(n), b(n);
    std::memcpy(a.data(), v.data(), n * sizeof(double));
    std::vector<std::size_t> count(65537);
    for (int shift = 0; shift < 64; shift += 16) {
        std::fill(count.begin(), count.end(), std::size_t{0});
        for (std::size_t i = 0; i < n; ++i) ++count[((a[i] >> shift) & 0xffffU) + 1];
        std::partial_sum(count.begin(), count.end(), count.begin());
        for (std::size_t i = 0; i < n; ++i) b[count[(a[i] >> shift) & 0xffffU]++] = a[i];
        a.swap(b);
    }
    std::memcpy(v.data(), a.data(), n * sizeof(double));
}

// Bin span of a clipped rectangle (Metrics.cpp:78-85).
void binSpan(const Box& b, const Box& region, double bw, double bh, int nx, int ny, int& c0, int& c1, int& r0, int& r1) {
    c0 = std::max(0, std::min(nx - 1, static_cast<int>(std::floor((b.lx - region.lx) / bw))));
    c1 = std::max(0, std::min(nx - 1, static_cast<int>(std::ceil((b.ux - region.lx) / bw) - 1.0)));
    r0 = std::max(0, std::min(ny - 1, static_cast<int>(std::floor((b.ly - region.ly) / bh))));
    r1 = std::max(0, std::min(ny - 1, static_cast<int>(std::ceil((b.uy - region.ly) / bh) - 1.0)));
}

Box binBox(const Box& region, double bw, double bh, int col, int row) {
    const double left = region.lx + static_cast<double>(col) * bw;
    const double bottom = region.ly + static_cast<double>(row) * bh;
    return {left, bottom, left + bw, bottom + bh};
}

}  // namespace

Netlist::Netlist(const mp_design& d) {
    if (d.module_count < 0 || d.net_count < 0 || d.pin_count < 0 || d.row_count < 0)
        throw std::invalid_argument("negative counts in the design");
    M = static_cast<std::size_t>(d.module_count);
    width.assign(d.width, d.width + M);
    height.assign(d.height, d.height + M);
    cx.resize(M);
    cy.resize(M);
    for (std::size_t i = 0; i < M; ++i) {
        cx[i] = d.center[2 * i];
        cy[i] = d.center[2 * i + 1];
        if (width[i] < 0.0 || height[i] < 0.0) throw std::invalid_argument("A module cannot have a negative size.");
    }
    orientation.assign(d.orientation, d.orientation + M);
    terminal.assign(d.is_terminal, d.is_terminal + M);
    fixed.assign(d.is_fixed, d.is_fixed + M);
    const std::size_t N = static_cast<std::size_t>(d.net_count);
    net_weight.assign(d.net_weight, d.net_weight + N);
    net_start.assign(d.net_pin_start, d.net_pin_start + N + 1);
    const std::size_t I = static_cast<std::size_t>(d.pin_count);
    if (net_start.front() != 0 || static_cast<std::size_t>(net_start.back()) != I)
        throw std::invalid_argument("net pin ranges do not cover the pin array");
    pin_module.assign(d.pin_module, d.pin_module + I);
    pin_dx.resize(I);
    pin_dy.resize(I);
    module_pins.assign(M, 0);
    for (std::size_t p = 0; p < I; ++p) {
        if (pin_module[p] < 0 || static_cast<std::size_t>(pin_module[p]) >= M)
            throw std::out_of_range("Pin refers to an unknown module or net.");
        pin_dx[p] = d.pin_offset[2 * p];
        pin_dy[p] = d.pin_offset[2 * p + 1];
        ++module_pins[pin_module[p]];
    }
    for (std::size_t n = 0; n < N; ++n)
        if (net_start[n + 1] < net_start[n]) throw std::invalid_argument("net pin ranges are not monotone");
    // rows, core region, nominal height, macro flags, movable order (PlacementDatabase.cpp:162-200)
    std::vector<double> heights;
    for (std::size_t r = 0; r < static_cast<std::size_t>(d.row_count); ++r) {
        const double* row = d.rows + 5 * r;
        const double x_end = row[2] + static_cast<double>(static_cast<std::size_t>(row[4])) * row[3];
        rows.push_back({row[2], row[0], x_end, row[0] + row[1]});
        heights.push_back(row[1]);
    }
    if (!heights.empty()) {
        std::vector<double> sorted = heights;
        std::sort(sorted.begin(), sorted.end());
        nominal_row_height = sorted[sorted.size() / 2];
        core = rows.front();
        for (const auto& r : rows) {
            core[0] = std::min(core[0], r[0]);
            core[1] = std::min(core[1], r[1]);
            core[2] = std::max(core[2], r[2]);
            core[3] = std::max(core[3], r[3]);
        }
    }
    macro.assign(M, 0);
    movable.reserve(M);
    for (std::size_t i = 0; i < M; ++i) {
        if (terminal[i]) fixed[i] = 1;
        macro[i] = !terminal[i] && nominal_row_height > kEps && std::abs(height[i] - nominal_row_height) > kEps;
        if (!fixed[i]) movable.push_back(static_cast<int64_t>(i));
    }
}

Prepared::Prepared(const mp_design& d, const mp_gp_options& o) : Prepared(Netlist(d), o) {}

Prepared::Prepared(Netlist&& n, const mp_gp_options& o) : net(std::move(n)) {
    t = o.target_density;
    nx = o.bins_x;
    ny = o.bins_y;
    if (nx <= 0 || ny <= 0 || !(t > 0.0))
        throw std::invalid_argument("DensityMap requires a valid region, positive dimensions, and target density.");
    if (static_cast<std::size_t>(nx) > kMaximumBinCount / static_cast<std::size_t>(ny))
        throw std::invalid_argument("Density grid exceeds the safe one-million-bin workspace limit.");
    const Box region{net.core[0], net.core[1], net.core[2], net.core[3]};
    if (!(region.ux - region.lx >= -kEps && region.uy - region.ly >= -kEps) || region.area() <= kEps)
        throw std::invalid_argument("DensityMap requires a valid region, positive dimensions, and target density.");
    bw = (region.ux - region.lx) / nx;
    bh = (region.uy - region.ly) / ny;
    const std::size_t G = static_cast<std::size_t>(nx) * ny;

    // static layers: dark = bin area - rows, then fixed modules (GlobalPlacementSupport.cpp:21-30)
    fixed_area.assign(G, 0.0);
    dark_area.resize(G);
    for (int r = 0; r < ny; ++r)
        for (int c = 0; c < nx; ++c) dark_area[static_cast<std::size_t>(r) * nx + c] = binBox(region, bw, bh, c, r).area();
    if (!net.rows.empty()) {
        for (const auto& row : net.rows) {
            const Box cl = Box{row[0], row[1], row[2], row[3]}.clip(region);
            if (cl.area() <= kEps) continue;
            int c0, c1, r0, r1;
            binSpan(cl, region, bw, bh, nx, ny, c0, c1, r0, r1);
            for (int r = r0; r <= r1; ++r)
                for (int c = c0; c <= c1; ++c)
                    dark_area[static_cast<std::size_t>(r) * nx + c] -= cl.clip(binBox(region, bw, bh, c, r)).area();
        }
        for (int r = 0; r < ny; ++r)
            for (int c = 0; c < nx; ++c) {
                double& dk = dark_area[static_cast<std::size_t>(r) * nx + c];
                const double tol = std::max(1.0, binBox(region, bw, bh, c, r).area()) * 1e-10;
                if (dk < -tol) throw std::invalid_argument("Overlapping SiteRows cannot define a density placeable-area mask.");
                dk = std::max(0.0, dk);
            }
    }
    auto deposit = [&](const Box& rect, std::vector<double>& layer) {  // Metrics.cpp:74-99
        const Box cl = rect.clip(region);
        if (cl.area() <= kEps) return;
        int c0, c1, r0, r1;
        binSpan(cl, region, bw, bh, nx, ny, c0, c1, r0, r1);
        for (int r = r0; r <= r1; ++r)
            for (int c = c0; c <= c1; ++c) {
                const double a = cl.clip(binBox(region, bw, bh, c, r)).area();
                if (a > kEps) layer[static_cast<std::size_t>(r) * nx + c] += a;
            }
    };
    for (std::size_t m = 0; m < net.M; ++m)
        if (net.fixed[m]) deposit(boxAround(net.cx[m], net.cy[m], net.width[m], net.height[m]), fixed_area);

    // fillers (GlobalPlacementSupport.cpp:52-103)
    double std_area = 0.0, macro_area = 0.0;
    std::vector<double> areas;
    for (const int64_t id : net.movable) {
        const double a = boxAround(net.cx[id], net.cy[id], net.width[id], net.height[id]).clip(region).area();
        (net.macro[id] ? macro_area : std_area) += a;
        if (a > kEps) areas.push_back(a);
    }
    double static_charge = 0.0;  // DensityMap::metrics(false).total_charge_area of the static map
    for (std::size_t i = 0; i < G; ++i) static_charge += t * (fixed_area[i] + dark_area[i]);
    const double required = std::max(0.0, t * region.area() - static_charge - (std_area + t * macro_area));
    const double filler_area = required / std::max(t, kEps);
    if (!(filler_area <= kEps || areas.empty() || o.maximum_fillers == 0)) {
        sortPositive(areas);
        const std::size_t lower = areas.size() / 10U;
        const std::size_t upper = std::max(lower + 1U, areas.size() - areas.size() / 10U);
        const double representative = std::accumulate(areas.begin() + static_cast<std::ptrdiff_t>(lower),
                                                      areas.begin() + static_cast<std::ptrdiff_t>(upper), 0.0) /
                                      static_cast<double>(upper - lower);
        const std::size_t count = std::min<std::size_t>(
            static_cast<std::size_t>(o.maximum_fillers),
            std::max<std::size_t>(1U, static_cast<std::size_t>(std::ceil(filler_area / std::max(representative, kEps)))));
        const double per = filler_area / static_cast<double>(count);
        const double aspect = std::max((region.ux - region.lx) / (region.uy - region.ly), 0.1);
        filler_w = std::sqrt(per * aspect);
        filler_h = per / std::max(filler_w, kEps);
        std::mt19937_64 gen(o.seed ^ 0x9e3779b97f4a7c15ULL);
        std::uniform_real_distribution<double> xs(region.lx + filler_w * 0.5, region.ux - filler_w * 0.5);
        std::uniform_real_distribution<double> ys(region.ly + filler_h * 0.5, region.uy - filler_h * 0.5);
        filler_x.resize(count);
        filler_y.resize(count);
        for (std::size_t i = 0; i < count; ++i) {
            filler_x[i] = xs(gen);
            filler_y[i] = ys(gen);
        }
    }

    // flattened layout (.cu:784-839)
    const std::size_t Mv = net.movable.size(), F = filler_x.size(), P = Mv + F;
    if (P > static_cast<std::size_t>(std::numeric_limits<int>::max()) ||
        net.pin_module.size() > static_cast<std::size_t>(std::numeric_limits<int>::max()))
        throw std::invalid_argument("HIP backend cannot index more than INT_MAX particles or pins");
    std::vector<int> module_particle(net.M, -1);
    pw.reserve(P);
    ph.reserve(P);
    pq.reserve(P);
    player.reserve(P);
    precond.reserve(P);
    for (std::size_t i = 0; i < Mv; ++i) {
        const int64_t m = net.movable[i];
        module_particle[m] = static_cast<int>(i);
        pw.push_back(net.width[m]);
        ph.push_back(net.height[m]);
        pq.push_back(net.macro[m] ? t * net.width[m] * net.height[m] : net.width[m] * net.height[m]);
        player.push_back(net.macro[m] ? MP_LAYER_MACRO : MP_LAYER_STANDARD);
    }
    for (std::size_t i = 0; i < F; ++i) {
        pw.push_back(filler_w);
        ph.push_back(filler_h);
        pq.push_back(t * (filler_w * filler_h));
        player.push_back(MP_LAYER_FILLER);
    }
    const std::size_t N = net.net_weight.size();
    net_offsets.assign(N + 1, 0);
    pin_particles.resize(net.pin_module.size());
    pin_xy.resize(2 * net.pin_module.size());
    for (std::size_t n = 0; n < N; ++n) net_offsets[n + 1] = static_cast<int>(net.net_start[n + 1]);
    parallelFor(net.pin_module.size(), 131072, [&](std::size_t p0, std::size_t p1) {
        for (std::size_t p = p0; p < p1; ++p) {
            const int m = net.pin_module[p];
            const int part = module_particle[m];
            double dx = net.pin_dx[p], dy = net.pin_dy[p];
            orient(dx, dy, net.orientation[m]);
            if (part < 0) {  // fixed pin: absolute position
                dx = net.cx[m] + dx;
                dy = net.cy[m] + dy;
            }
            pin_particles[p] = part;
            pin_xy[2 * p] = dx;
            pin_xy[2 * p + 1] = dy;
        }
    });

    // preconditioner base (AdaptiveGlobalPlacer.cpp:358-375)
    double total = 0.0;
    for (const int64_t id : net.movable) total += net.width[id] * net.height[id];
    const double avg = total / static_cast<double>(std::max<std::size_t>(1U, Mv));
    auto scale = [&](double area, std::size_t pins) {
        const double ratio = std::max(area / std::max(avg, kEps), 0.1);
        return 1.0 / (std::sqrt(1.0 + static_cast<double>(pins)) * (1.0 + std::sqrt(ratio)));
    };
    for (const int64_t id : net.movable) precond.push_back(scale(net.width[id] * net.height[id], net.module_pins[id]));
    for (std::size_t i = 0; i < F; ++i) precond.push_back(scale(filler_w * filler_h, 0));
}

thread_local int t_parallel_limit = 8;

void keepFreedHostMemory() noexcept {
    static std::once_flag once;
    std::call_once(once, [] {
        (void)mallopt(M_TRIM_THRESHOLD, INT_MAX);
        (void)mallopt(M_TOP_PAD, 1 << 30);
    });
}

void stableSortByKey(std::uint32_t* idx, std::size_t n, const std::uint32_t* key, int key_bits) {
    constexpr int kBits = 11;
    constexpr std::uint32_t kRadix = 1U << kBits;
    std::vector<std::uint32_t> tmp(n);
    std::vector<std::size_t> count(kRadix + 1);
    std::uint32_t* src = idx;
    std::uint32_t* dst = tmp.data();
    for (int shift = 0; shift < key_bits; shift += kBits) {
        std::fill(count.begin(), count.end(), std::size_t{0});
        for (std::size_t i = 0; i < n; ++i) ++count[((key[src[i]] >> shift) & (kRadix - 1)) + 1];
        std::partial_sum(count.begin(), count.end(), count.begin());
        for (std::size_t i = 0; i < n; ++i) dst[count[(key[src[i]] >> shift) & (kRadix - 1)]++] = src[i];
        std::swap(src, dst);
    }
    if (src != idx) std::copy(src, src + n, idx);
}

mp_gp_static Prepared::layout(int device, int dct_precision, std::uint64_t maximum_bytes) const {
    mp_gp_static L{};
    L.particle_count = static_cast<int64_t>(pw.size());
    L.particle_width = pw.data();
    L.particle_height = ph.data();
    L.particle_charge_area = pq.data();
    L.particle_layer = player.data();
    L.net_count = static_cast<int64_t>(net_offsets.size()) - 1;
    L.net_offsets = net_offsets.data();
    L.net_weights = net.net_weight.data();
    L.pin_count = static_cast<int64_t>(pin_particles.size());
    L.pin_particles = pin_particles.data();
    L.pin_offsets_or_static = pin_xy.data();
    L.bins_x = nx;
    L.bins_y = ny;
    L.region_ll_x = net.core[0];
    L.region_ll_y = net.core[1];
    L.region_ur_x = net.core[2];
    L.region_ur_y = net.core[3];
    L.target_density = t;
    L.fixed_area = fixed_area.data();
    L.dark_area = dark_area.data();
    L.device = device;
    L.dct_precision = dct_precision;
    L.maximum_device_bytes = maximum_bytes;
    return L;
}

double Netlist::hpwl(const double* cxv, const double* cyv) const {  // Metrics.cpp:128-147
    // the nets' weighted extents in parallel parts, then their sum in net order
    const std::size_t N = net_start.empty() ? 0 : net_start.size() - 1;
    std::vector<double> term(N, 0.0);
    parallelFor(N, 32768, [&](std::size_t n0, std::size_t n1) {
        for (std::size_t n = n0; n < n1; ++n) {
            const int64_t b = net_start[n], e = net_start[n + 1];
            if (e - b < 2) continue;
            double lx = INFINITY, ux = -INFINITY, ly = INFINITY, uy = -INFINITY;
            for (int64_t p = b; p < e; ++p) {
                const int m = pin_module[p];
                double dx = pin_dx[p], dy = pin_dy[p];
                orient(dx, dy, orientation[m]);
                const double x = cxv[m] + dx, y = cyv[m] + dy;
                lx = std::min(lx, x); ux = std::max(ux, x);
                ly = std::min(ly, y); uy = std::max(uy, y);
            }
            term[n] = net_weight[n] * ((ux - lx) + (uy - ly));
        }
    });
    double total = 0.0;
    for (std::size_t n = 0; n < N; ++n)
        if (net_start[n + 1] - net_start[n] >= 2) total += term[n];
    return total;
}

double Netlist::designOverflow(const double* cxv, const double* cyv, int nx, int ny, double t) const {
    // calculateDensity(...).normalized_overflow, Metrics.cpp:149-164
    const Box region{core[0], core[1], core[2], core[3]};
    const double bw = (region.ux - region.lx) / nx, bh = (region.uy - region.ly) / ny;
    const std::size_t G = static_cast<std::size_t>(nx) * ny;
    std::vector<double> fixedl(G, 0.0), mov(G, 0.0), mac(G, 0.0), dark(G);
    for (int r = 0; r < ny; ++r)
        for (int c = 0; c < nx; ++c) dark[static_cast<std::size_t>(r) * nx + c] = binBox(region, bw, bh, c, r).area();
    for (const auto& row : rows) {
        const Box cl = Box{row[0], row[1], row[2], row[3]}.clip(region);
        if (cl.area() <= kEps) continue;
        int c0, c1, r0, r1;
        binSpan(cl, region, bw, bh, nx, ny, c0, c1, r0, r1);
        for (int r = r0; r <= r1; ++r)
            for (int c = c0; c <= c1; ++c)
                dark[static_cast<std::size_t>(r) * nx + c] -= cl.clip(binBox(region, bw, bh, c, r)).area();
    }
    for (auto& v : dark) v = std::max(0.0, v);
    for (std::size_t m = 0; m < M; ++m) {
        const Box cl = boxAround(cxv[m], cyv[m], width[m], height[m]).clip(region);
        if (cl.area() <= kEps) continue;
        std::vector<double>& layer = fixed[m] ? fixedl : (macro[m] ? mac : mov);
        int c0, c1, r0, r1;
        binSpan(cl, region, bw, bh, nx, ny, c0, c1, r0, r1);
        for (int r = r0; r <= r1; ++r)
            for (int c = c0; c <= c1; ++c) {
                const double a = cl.clip(binBox(region, bw, bh, c, r)).area();
                if (a > kEps) layer[static_cast<std::size_t>(r) * nx + c] += a;
            }
    }
    double overflow = 0.0, norm = 0.0;
    for (int r = 0; r < ny; ++r)
        for (int c = 0; c < nx; ++c) {
            const std::size_t i = static_cast<std::size_t>(r) * nx + c;
            const double area = binBox(region, bw, bh, c, r).area();
            const double q = mov[i] + t * (fixedl[i] + mac[i] + dark[i]);
            overflow += std::max(0.0, q - t * area);
            norm += mov[i] + t * mac[i];
        }
    return overflow / std::max(norm, kEps);
}

void randomPlacement(const mp_design& d, std::uint64_t seed, double* centers) {  // InitialPlacer.cpp:102-120
    Netlist net(d);
    std::mt19937_64 gen(seed);
    const double lx = net.core[0], ly = net.core[1], ux = net.core[2], uy = net.core[3];
    for (std::size_t m = 0; m < net.M; ++m) {
        centers[2 * m] = net.cx[m];
        centers[2 * m + 1] = net.cy[m];
    }
    const bool has_core = std::max(0.0, ux - lx) * std::max(0.0, uy - ly) > kEps;
    for (const int64_t id : net.movable) {
        const double w = net.width[id], h = net.height[id];
        double& x = centers[2 * id];
        double& y = centers[2 * id + 1];
        if (lx + w * 0.5 <= ux - w * 0.5) {
            std::uniform_real_distribution<double> dist(lx + w * 0.5, ux - w * 0.5);
            x = dist(gen);
        }
        if (ly + h * 0.5 <= uy - h * 0.5) {
            std::uniform_real_distribution<double> dist(ly + h * 0.5, uy - h * 0.5);
            y = dist(gen);
        }
        if (has_core) {  // clampModuleToCore, PlacementDatabase.cpp:202-213
            x = lx + w * 0.5 <= ux - w * 0.5 ? clampd(x, lx + w * 0.5, ux - w * 0.5) : (lx + ux) * 0.5;
            y = ly + h * 0.5 <= uy - h * 0.5 ? clampd(y, ly + h * 0.5, uy - h * 0.5) : (ly + uy) * 0.5;
        }
    }
}

}  // namespace mpgp
