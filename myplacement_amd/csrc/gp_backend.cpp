// gp_backend.cpp — Backend construction, evaluation orchestration and the backend half of the
// C ABI (mp_gp_create / mp_gp_evaluate / mp_gp_destroy ...).
//
// Reference: CudaPlacementBackendImpl and tryCreateCudaPlacementBackend,
// src/placement/cuda/CudaPlacementBackend.cu:841-1216.  Differences by design:
//   * no blocking scalar copy inside the evaluation (the reference does 14): one k_finalize
//     writes every scalar into host-mapped memory, and the host waits once;
//   * the controller's two gradient evaluations (current and look-ahead positions) run as one
//     chain of batched launches: twice the workgroups per launch, half the launches;
//   * the density is a fixed-point deposit through per-workgroup LDS windows over a particle
//     order sorted by bin row, reused across evaluations (the order affects speed, not results);
//   * the gradient scatter is atomic-free: per-pin terms, then a particle -> pin CSR sum;
//   * the DCT stages are dense bin-grid products on MFMA; the electrostatic energy is taken in
//     the cosine basis (sum_lk c_l c_k / 4 * chat * phihat), which equals the reference's grid sum
//     (.cu:569-573 / AdaptiveGlobalPlacer.cpp:341-354) by the orthogonality of the DCT-II basis,
//     so a value-only evaluation needs no inverse transform;
//   * the density mean is not subtracted before the forward transform: it only moves the DC
//     coefficient, which the solve drops (ElectrostaticField.cpp:228).

#include "gp_backend.hpp"
#include "gp_host.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <numeric>
#include <sstream>
#include <type_traits>
#include <vector>

namespace mpgp {

namespace {
constexpr double kPiD = 3.141592653589793238462643383279502884;
constexpr std::size_t kMaximumBinCount = 1U << 20U;  // core/ResourceLimits.hpp:10
constexpr std::size_t kDeviceReserveBytes = 4ULL << 30;


int roundUp32(int n) { return (n + 31) / 32 * 32; }

// Dense transform factors (FFTW r2r definitions, unnormalised), FP64 then rounded to FP32,
// zero in the padding.  `rows` x `cols` row-major.
std::vector<float> basis(int rows, int cols, int n, int kind, double wave_scale) {
    // kind 0: F[i][k] = 2 cos(pi k (2i+1) / 2n)        (REDFT10, x-side right factor, i = row)
    // kind 1: F[l][j] = 2 cos(pi l (2j+1) / 2n)        (REDFT10, y-side left factor)
    // kind 2: F[j][l] = c_l cos(pi l (2j+1) / 2n)      (REDFT01, left factor)
    // kind 3: F[j][l] = 2 sin(pi l (2j+1) / 2n) * w_l  (RODFT01 with the input shift folded in)
    std::vector<float> f(static_cast<std::size_t>(rows) * cols, 0.f);
    auto ang = [n](long long a, long long b) {  // cos(pi * a * b / 2n) with exact integer reduction
        const long long m = (a * b) % (4LL * n);
        return kPiD * static_cast<double>(m) / (2.0 * n);
    };
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c) {
            double v = 0.0;
            if (kind == 0) v = 2.0 * std::cos(ang(c, 2 * r + 1));
            else if (kind == 1) v = 2.0 * std::cos(ang(r, 2 * c + 1));
            else if (kind == 2) v = (c == 0 ? 1.0 : 2.0) * std::cos(ang(c, 2 * r + 1));
            else v = c == 0 ? 0.0 : 2.0 * std::sin(ang(c, 2 * r + 1)) * (kPiD * c / wave_scale);
            f[static_cast<std::size_t>(r) * cols + c] = static_cast<float>(v);
        }
    return f;
}
}  // namespace

void hipCheck(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

thread_local UploadStaging* t_upload_staging = nullptr;

void stagedUpload(void* dst, const void* src, std::size_t bytes, hipStream_t s, const char* what) {
    UploadStaging* st = t_upload_staging;
    const std::size_t at = st ? (st->used + 255) / 256 * 256 : 0;
    if (st && st->base && at <= st->size && bytes <= st->size - at) {
        std::memcpy(st->base + at, src, bytes);
        st->used = at + bytes;
        hipCheck(hipMemcpyAsync(dst, st->base + at, bytes, hipMemcpyHostToDevice, s), what);
    } else {
        hipCheck(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s), what);
    }
}

void DeviceBudget::reserve(std::size_t bytes, const char* what) {
    if (bytes > limit_ - reserved_) {
        std::ostringstream m;
        m << "HIP allocation budget exceeded while reserving " << what << " (requested " << bytes << " bytes, reserved "
          << reserved_ << " bytes, limit " << limit_ << " bytes).";
        throw std::runtime_error(m.str());
    }
    reserved_ += bytes;
}

Backend::Backend(const mp_gp_static& L, std::size_t allocation_limit, const BorrowedResources* borrowed)
    : device_(L.device), budget_(allocation_limit) {
    if (L.particle_count < 0 || L.net_count < 0 || L.pin_count < 0)
        throw InvalidArgument("negative counts in the placement layout");
    if (L.particle_count > std::numeric_limits<int>::max() || L.pin_count > std::numeric_limits<int>::max() ||
        L.net_count > std::numeric_limits<int>::max())
        throw InvalidArgument("HIP backend cannot index more than INT_MAX particles, pins or nets");
    if (L.bins_x < 2 || L.bins_y < 2) throw InvalidArgument("the density grid needs at least 2 x 2 bins");
    if (static_cast<std::size_t>(L.bins_x) > kMaximumBinCount / static_cast<std::size_t>(L.bins_y))
        throw InvalidArgument("Density grid exceeds the safe one-million-bin workspace limit.");
    const double W = L.region_ur_x - L.region_ll_x, H = L.region_ur_y - L.region_ll_y;
    if (!(W > 0.0 && H > 0.0 && W * H > 1e-9)) throw InvalidArgument("invalid placement region");
    if (!(L.target_density > 0.0)) throw InvalidArgument("target density must be positive");
    if (L.dct_precision != MP_DCT_FP32 && L.dct_precision != MP_DCT_BF16X3)
        throw InvalidArgument("unknown DCT precision");
    if (L.dct_precision == MP_DCT_BF16X3 && !foldedDct(L.bins_x, L.bins_y))
        throw InvalidArgument("split-BF16 DCT products need both bin counts to be multiples of 64");
    P_ = static_cast<int>(L.particle_count);
    N_ = static_cast<int>(L.net_count);
    I_ = static_cast<int>(L.pin_count);
    nx_ = L.bins_x;
    ny_ = L.bins_y;
    nxp_ = roundUp32(nx_);
    nyp_ = roundUp32(ny_);
    t_ = L.target_density;
    dct_precision_ = L.dct_precision;
    region_ = {L.region_ll_x, L.region_ll_y, L.region_ur_x, L.region_ur_y, W / nx_, H / ny_, nx_, ny_};
    const std::size_t G = static_cast<std::size_t>(nx_) * ny_;

    // ---- host-side derived arrays ----
    if (N_ > 0 && (L.net_offsets == nullptr || L.net_offsets[0] != 0 || L.net_offsets[N_] != I_))
        throw InvalidArgument("net offsets do not cover the pin array");
    // Internal net order for the wirelength kernel: degree classes 2..8, 9..32, >32 (nets with fewer
    // than two pins contribute nothing and are dropped, .cu:251-253, :324-326).  Small nets by degree,
    // then by the lowest particle index among their pins (the nets of one wavefront gather nearby
    // particles and share one degree); medium nets by lowest particle; large nets largest first
    // (their workgroups start first and take longest).  Pins follow the net order, except that each
    // small-net degree group stores its pins slot-major: pin j of the group's i-th net at
    // group base + j * group size + i, so one wavefront's loads of slot j are contiguous.
    auto degree = [&](int n) { return L.net_offsets[n + 1] - L.net_offsets[n]; };
    std::vector<uint32_t> lowest(static_cast<std::size_t>(N_), static_cast<uint32_t>(std::numeric_limits<int>::max()));
    std::vector<uint32_t> by_lowest;  // the nets of two pins or more, by lowest particle (ties: index)
    by_lowest.reserve(static_cast<std::size_t>(N_));
    for (int n = 0; n < N_; ++n)
        if (degree(n) < 0) throw InvalidArgument("net offsets are not monotone");
    parallelFor(static_cast<std::size_t>(N_), 32768, [&](std::size_t n0, std::size_t n1) {
        for (std::size_t n = n0; n < n1; ++n)
            for (int q = L.net_offsets[n]; q < L.net_offsets[n + 1]; ++q)
                if (L.pin_particles[q] >= 0) lowest[n] = std::min(lowest[n], static_cast<uint32_t>(L.pin_particles[q]));
    });
    for (int n = 0; n < N_; ++n)
        if (degree(n) >= 2) by_lowest.push_back(static_cast<uint32_t>(n));
    stableSortByKey(by_lowest.data(), by_lowest.size(), lowest.data(), 31);
    // then stably by bucket: small nets by degree (2 .. kSmallDegree), medium nets, large nets
    // (these last by degree descending)
    constexpr int kBuckets = kSmallDegree + 1;  // 0 .. kSmallDegree - 2 small, then medium, large
    auto bucket = [&](int n) {
        const int deg = degree(n);
        return deg <= kSmallDegree ? deg - 2 : (deg <= kMediumDegree ? kSmallDegree - 1 : kSmallDegree);
    };
    std::vector<int> start(kBuckets + 1, 0);
    for (uint32_t n : by_lowest) ++start[bucket(static_cast<int>(n)) + 1];
    std::partial_sum(start.begin(), start.end(), start.begin());
    std::vector<int> order(by_lowest.size());
    {
        std::vector<int> fill(start.begin(), start.end() - 1);
        for (uint32_t n : by_lowest) order[fill[bucket(static_cast<int>(n))]++] = static_cast<int>(n);
    }
    n_small_ = start[kSmallDegree - 1];
    n_medium_ = start[kSmallDegree] - start[kSmallDegree - 1];
    n_large_ = start[kBuckets] - start[kSmallDegree];
    std::stable_sort(order.begin() + start[kSmallDegree], order.end(), [&](int a, int b) { return degree(a) > degree(b); });
    std::vector<int> inet_off(order.size() + 1, 0);
    std::vector<double> inet_w(order.size());
    for (std::size_t k = 0; k < order.size(); ++k) {
        inet_w[k] = L.net_weights[order[k]];
        inet_off[k + 1] = inet_off[k] + degree(order[k]);
    }
    for (int d = 0; d <= kSmallDegree + 1; ++d) small_net_[d] = small_pin_[d] = 0;
    {
        int k = 0;
        for (int d = 2; d <= kSmallDegree; ++d) {
            small_net_[d] = k;
            while (k < n_small_ && degree(order[k]) == d) ++k;
        }
        small_net_[kSmallDegree + 1] = n_small_;
        for (int d = 2; d <= kSmallDegree + 1; ++d) small_pin_[d] = inet_off[small_net_[d]];
    }
    std::vector<int> ipin_part(static_cast<std::size_t>(inet_off.back()));
    std::vector<double2> ipin_off(ipin_part.size());
    // pins walked in their input order (sequential reads), each stored at its place in the net
    // order (scattered writes, which do not stall the walk), in parallel parts
    std::vector<int> net_rank(static_cast<std::size_t>(N_), -1);
    for (std::size_t k = 0; k < order.size(); ++k) net_rank[order[k]] = static_cast<int>(k);
    parallelFor(static_cast<std::size_t>(N_), 32768, [&](std::size_t n0, std::size_t n1) {
        for (std::size_t nn = n0; nn < n1; ++nn) {
            const int n = static_cast<int>(nn);
            const std::size_t k = static_cast<std::size_t>(net_rank[n]);
            if (net_rank[n] < 0) continue;  // fewer than two pins
            const int deg = degree(n);
            for (int j = 0; j < deg; ++j) {
                const int p = L.net_offsets[n] + j;
                const int part = L.pin_particles[p];
                if (part >= P_) throw InvalidArgument("pin refers to an unknown particle");
                std::size_t q = static_cast<std::size_t>(inet_off[k]) + j;
                if (static_cast<int>(k) < n_small_) {  // slot-major inside the degree group
                    const int count = small_net_[deg + 1] - small_net_[deg];
                    q = static_cast<std::size_t>(small_pin_[deg]) + static_cast<std::size_t>(j) * count +
                        (static_cast<int>(k) - small_net_[deg]);
                }
                ipin_part[q] = part;
                ipin_off[q] = make_double2(L.pin_offsets_or_static[2 * p], L.pin_offsets_or_static[2 * p + 1]);
            }
        }
    });
    const std::size_t Ii = ipin_part.size();

    // particle -> pin CSR over the movable pins, pins in internal order
    std::vector<int> count(static_cast<std::size_t>(P_) + 1, 0);
    for (std::size_t q = 0; q < Ii; ++q)
        if (ipin_part[q] >= 0) ++count[ipin_part[q] + 1];
    std::partial_sum(count.begin(), count.end(), count.begin());
    // Slots of the per-pin gradient terms: k_gather's workgroup w (kGatherParticles particles) runs
    // on XCD w mod 8, so the terms are stored in kPinRegions = 4 regions, region k holding, in pin
    // order, the pins of the particles of workgroups w = k mod 4.  The gather on XCDs k and k + 4
    // then reads one region, a quarter of the terms, which their L2s hold, and a wavefront's
    // stores of one pin slot fall in four runs.  A batch driver's jobs keep the terms in pin order
    // (no slot array): their launches give each job's workgroups to one XCD already.
    const bool regions = borrowed == nullptr;
    std::vector<int> pin_dest(Ii, -1);
    if (regions) {
        auto region = [](int part) { return (part / kGatherParticles) % kPinRegions; };
        std::array<int, kPinRegions + 1> base{};
        for (std::size_t q = 0; q < Ii; ++q)
            if (ipin_part[q] >= 0) ++base[region(ipin_part[q]) + 1];
        std::partial_sum(base.begin(), base.end(), base.begin());
        for (std::size_t q = 0; q < Ii; ++q)
            if (ipin_part[q] >= 0) pin_dest[q] = base[region(ipin_part[q])]++;
    }
    std::vector<int> ppin_idx(static_cast<std::size_t>(count[P_]));
    {
        std::vector<int> fill(count.begin(), count.end() - 1);
        for (std::size_t q = 0; q < Ii; ++q)
            if (ipin_part[q] >= 0) ppin_idx[fill[ipin_part[q]]++] = regions ? pin_dest[q] : static_cast<int>(q);
    }
    std::vector<double2> sizes(P_);
    for (int i = 0; i < P_; ++i) sizes[i] = make_double2(L.particle_width[i], L.particle_height[i]);
    std::vector<double> static_charge(G);
    for (std::size_t i = 0; i < G; ++i) static_charge[i] = t_ * (L.fixed_area[i] + L.dark_area[i]);
    // dark / placeable totals in bin order, as DensityMap::metrics sums them (Metrics.cpp:111-113)
    for (int row = 0; row < ny_; ++row)
        for (int col = 0; col < nx_; ++col) {
            const double left = region_.llx + static_cast<double>(col) * region_.bw;
            const double bottom = region_.lly + static_cast<double>(row) * region_.bh;
            const double area = std::max(0.0, (left + region_.bw) - left) * std::max(0.0, (bottom + region_.bh) - bottom);
            const double dark = L.dark_area[static_cast<std::size_t>(row) * nx_ + col];
            dark_total_ += dark;
            placeable_total_ += area - dark;
        }

    // ---- device ----
    setDevice();
    if (borrowed && borrowed->stream) {
        stream_ = borrowed->stream;
        owns_stream_ = false;
    } else {
        hipCheck(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "create HIP stream");
    }
    if (borrowed) budget_.setArena(borrowed->arena);
    own_launches_ = std::make_unique<StreamLaunches>(stream_);
    launches_ = own_launches_.get();
    hipStream_t stream = stream_;
    pos_.allocate(P_, budget_, "particle positions");
    size_.allocate(P_, budget_, "particle sizes");
    charge_.allocate(P_, budget_, "particle charge areas");
    layer_.allocate(P_, budget_, "particle layers");
    net_off_.allocate(order.size() + 1, budget_, "net offsets");
    net_w_.allocate(order.size(), budget_, "net weights");
    pin_part_.allocate(Ii, budget_, "pin-particle map");
    pin_off_.allocate(Ii, budget_, "pin offsets");
    ppin_idx_.allocate(ppin_idx.size(), budget_, "particle-pin index");
    if (regions) pin_dest_.allocate(Ii, budget_, "pin gradient slots");
    ppin_off_.allocate(static_cast<std::size_t>(P_) + 1, budget_, "particle-pin offsets");
    // density: particles taller than two bin rows are not row-sorted (their own deposit workgroups
    // add them to the global layers)
    if (ny_ > 8192) throw InvalidArgument("density grids taller than 8192 bins are not supported by the row sort");
    if (static_cast<unsigned>(P_) > kPermIdMask) throw InvalidArgument("HIP backend cannot order more than 2^30 - 1 particles");
    fix_ = std::ldexp(1.0, 40) / (region_.bw * region_.bh);
    std::vector<unsigned char> tall(P_);
    std::vector<int> tall_ids;
    for (int i = 0; i < P_; ++i) {
        tall[i] = L.particle_height[i] > 2.0 * region_.bh ? 1 : 0;
        if (tall[i]) tall_ids.push_back(i);
    }
    n_tall_ = static_cast<int>(tall_ids.size());
    const int hist_blocks = densityHistBlocks(P_);
    key_.allocate(P_, budget_, "density sort keys");
    hist_.allocate(static_cast<std::size_t>(hist_blocks) * ny_, budget_, "density row histograms");
    colprefix_.allocate(static_cast<std::size_t>(hist_blocks) * ny_, budget_, "density row prefixes");
    keytotal_.allocate(ny_, budget_, "density row totals");
    keystart_.allocate(static_cast<std::size_t>(ny_) + 1, budget_, "density row starts");
    perm_.allocate(P_, budget_, "particle row order");
    sorted_size_.allocate(P_, budget_, "row-ordered particle sizes");
    tall_.allocate(P_, budget_, "tall-particle flags");
    tall_ids_.allocate(tall_ids.size(), budget_, "tall-particle ids");
    static_charge_.allocate(G, budget_, "static density charge");
    const std::size_t Gp = static_cast<std::size_t>(nxp_) * nyp_;
    for (Slot& sl : slots_) {
        sl.pin_grad.allocate(Ii, budget_, "per-pin wirelength gradient");
        sl.rho.allocate(Gp, budget_, "density deviation grid");
        sl.tmp.allocate(Gp, budget_, "DCT intermediate");
        sl.phi.allocate(Gp, budget_, "potential coefficients");
        sl.t12.allocate(2 * Gp, budget_, "inverse DCT intermediate");
        sl.field.allocate(2 * Gp, budget_, "field grid");
        sl.grid.allocate(2 * G, budget_, "fixed-point density layers");
    }
    pot_.allocate(Gp, budget_, "potential grid");
    bxf_.allocate(static_cast<std::size_t>(nxp_) * nxp_, budget_, "forward x factor");
    ayf_.allocate(static_cast<std::size_t>(nyp_) * nyp_, budget_, "forward y factor");
    bxi_.allocate(static_cast<std::size_t>(nxp_) * 2 * nxp_, budget_, "inverse x factors");
    ayi_.allocate(static_cast<std::size_t>(nyp_) * nyp_, budget_, "inverse y cosine factor");
    syi_.allocate(static_cast<std::size_t>(nyp_) * nyp_, budget_, "inverse y sine factor");
    const int max_blocks = std::max({wireBlocks(n_small_, n_medium_, n_large_), densityBinBlocks(region_),
                                     particleBlocks(P_), (nxp_ / 32) * (nyp_ / 32), 1});
    for (int k = 0; k < kRedKinds * kScalarBanks; ++k)
        partials_[k].allocate(static_cast<std::size_t>(max_blocks + 1) * redWidth(k / kScalarBanks), budget_,
                              "reduction partials");  // + the row a wirelength evaluated ahead writes
    scal_copy_.allocate(static_cast<std::size_t>(kScalarCount) * kScalarBanks, budget_, "scalar block copy");
    wire_out_.allocate(P_, budget_, "wire gradient");
    density_out_.allocate(P_, budget_, "density gradient");
    // coherent host memory: k_finalize's stores go straight to the host, no copy after it
    if (borrowed && borrowed->scalars) {
        host_scal_ = borrowed->scalars;
        dev_scal_ = reinterpret_cast<double*>(borrowed->scalars_device);
        owns_scalars_ = false;
        std::memset(host_scal_, 0, sizeof(HostScalars));
    } else {
        hipCheck(hipHostMalloc(reinterpret_cast<void**>(&host_scal_), sizeof(HostScalars),
                               hipHostMallocMapped | hipHostMallocCoherent),
                 "allocate host-mapped scalars");
        std::memset(host_scal_, 0, sizeof(HostScalars));
        hipCheck(hipHostGetDevicePointer(reinterpret_cast<void**>(&dev_scal_), host_scal_, 0), "map scalar block");
    }
    dev_done_ = reinterpret_cast<unsigned*>(reinterpret_cast<char*>(dev_scal_) + offsetof(HostScalars, done));

    std::vector<int> layers(L.particle_layer, L.particle_layer + P_);
    for (int v : layers)
        if (v < 0 || v > 2) throw InvalidArgument("unknown particle layer");
    size_.upload(sizes.data(), P_, stream, "upload particle sizes");
    charge_.upload(L.particle_charge_area, P_, stream, "upload particle charges");
    layer_.upload(layers.data(), P_, stream, "upload particle layers");
    net_off_.upload(inet_off.data(), inet_off.size(), stream, "upload net offsets");
    net_w_.upload(inet_w.data(), inet_w.size(), stream, "upload net weights");
    pin_part_.upload(ipin_part.data(), Ii, stream, "upload pin-particle map");
    pin_off_.upload(ipin_off.data(), Ii, stream, "upload pin offsets");
    ppin_off_.upload(count.data(), count.size(), stream, "upload particle-pin offsets");
    ppin_idx_.upload(ppin_idx.data(), ppin_idx.size(), stream, "upload particle-pin index");
    if (regions) pin_dest_.upload(pin_dest.data(), Ii, stream, "upload pin gradient slots");
    static_charge_.upload(static_charge.data(), G, stream, "upload static density");
    const auto fx = basis(nxp_, nxp_, nx_, 0, W);
    const auto fy = basis(nyp_, nyp_, ny_, 1, H);
    const auto ix = basis(nxp_, nxp_, nx_, 2, W);
    const auto sx = basis(nxp_, nxp_, nx_, 3, W);
    const auto iy = basis(nyp_, nyp_, ny_, 2, H);
    const auto sy = basis(nyp_, nyp_, ny_, 3, H);
    folded_ = foldedDct(nx_, ny_);
    if (folded_) {
        // half-length factors (launchDct): parity p of the frequency index in the slowest dimension
        const std::size_t hx = nx_ / 2, hy = ny_ / 2, n = nx_;
        std::vector<float> fxh(2 * hx * hx), fyh(2 * hy * hy), bxh(2 * hx * n), iyh(2 * hy * hy), syh(2 * hy * hy);
        for (std::size_t p = 0; p < 2; ++p) {
            for (std::size_t t = 0; t < hx; ++t) {
                for (std::size_t c = 0; c < hx; ++c) fxh[(p * hx + t) * hx + c] = fx[t * nxp_ + 2 * c + p];
                for (std::size_t c = 0; c < n; ++c)
                    bxh[(p * hx + t) * n + c] = c < hx ? ix[c * nxp_ + 2 * t + p] : sx[(c - hx) * nxp_ + 2 * t + p];
            }
            for (std::size_t r = 0; r < hy; ++r)
                for (std::size_t t = 0; t < hy; ++t) {
                    fyh[(p * hy + r) * hy + t] = fy[(2 * r + p) * nyp_ + t];
                    iyh[(p * hy + r) * hy + t] = iy[r * nyp_ + 2 * t + p];
                    syh[(p * hy + r) * hy + t] = sy[r * nyp_ + 2 * t + p];
                }
        }
        bxf_.upload(fxh.data(), fxh.size(), stream, "upload forward x factor");
        ayf_.upload(fyh.data(), fyh.size(), stream, "upload forward y factor");
        bxi_.upload(bxh.data(), bxh.size(), stream, "upload inverse x factors");
        ayi_.upload(iyh.data(), iyh.size(), stream, "upload inverse y factor");
        syi_.upload(syh.data(), syh.size(), stream, "upload inverse y sine factor");
        hipCheck(hipStreamSynchronize(stream), "upload DCT factors");  // the host vectors go out of scope
    } else {
        // inverse x factors side by side: [k][i] = ix[i][k] (DCT-III), [k][nxp + i] = sx[i][k] (DST-III)
        std::vector<float> bx2(static_cast<std::size_t>(nxp_) * 2 * nxp_, 0.f);
        for (int k = 0; k < nxp_; ++k)
            for (int i = 0; i < nxp_; ++i) {
                bx2[static_cast<std::size_t>(k) * 2 * nxp_ + i] = ix[static_cast<std::size_t>(i) * nxp_ + k];
                bx2[static_cast<std::size_t>(k) * 2 * nxp_ + nxp_ + i] = sx[static_cast<std::size_t>(i) * nxp_ + k];
            }
        bxf_.upload(fx.data(), fx.size(), stream, "upload forward x factor");
        ayf_.upload(fy.data(), fy.size(), stream, "upload forward y factor");
        bxi_.upload(bx2.data(), bx2.size(), stream, "upload inverse x factors");
        ayi_.upload(iy.data(), iy.size(), stream, "upload inverse y factor");
        syi_.upload(sy.data(), sy.size(), stream, "upload inverse y sine factor");
        hipCheck(hipStreamSynchronize(stream), "upload DCT factors");
    }
    tall_.upload(tall.data(), tall.size(), stream, "upload tall-particle flags");
    tall_ids_.upload(tall_ids.data(), tall_ids.size(), stream, "upload tall-particle ids");
    // the padding of the deviation grids stays zero: the band kernel writes only the nx x ny part
    for (Slot& sl : slots_) {
        hipCheck(hipMemsetAsync(sl.rho.get(), 0, Gp * sizeof(float), stream), "clear density grid");
        // the bin kernel clears the layers after reading them, so they start every evaluation at zero
        hipCheck(hipMemsetAsync(sl.grid.get(), 0, 2 * G * sizeof(unsigned long long), stream), "clear density layers");
    }
    hipCheck(hipStreamSynchronize(stream), "finish backend upload");
}

Backend::~Backend() {
    if (stream_) {
        (void)hipSetDevice(device_);
        if (capture_shape_ >= 0) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(stream_, &g);
            if (g) (void)hipGraphDestroy(g);
        }
        (void)hipStreamSynchronize(stream_);
        for (auto& e : graphs_) (void)hipGraphExecDestroy(e.second);
        for (auto e : ev_pool_) (void)hipEventDestroy(e);
        if (owns_stream_) (void)hipStreamDestroy(stream_);
    }
    if (host_scal_ && owns_scalars_) (void)hipHostFree(host_scal_);
    if (host_stage_) (void)hipHostFree(host_stage_);
}

void Backend::setDevice() const { hipCheck(hipSetDevice(device_), "select HIP device"); }

hipEvent_t Backend::poolEvent() {
    if (ev_next_ == ev_pool_.size()) {
        // timing only: without the system-scope fence an event neither writes back nor invalidates
        // the caches, so the kernels after it run as they do untimed (the timestamps are read after
        // a stream synchronize)
        hipEvent_t e;
        hipCheck(hipEventCreateWithFlags(&e, hipEventDisableSystemFence), "create timing event");
        ev_pool_.push_back(e);
    }
    return ev_pool_[ev_next_++];
}

void Backend::timedBegin(int cls) {
    if (!((timing_mask_ >> cls) & 1U)) return;
    t_open_[cls] = t_calls_[cls]++ % static_cast<unsigned long long>(timing_every_) == 0;
    if (!t_open_[cls]) return;
    const hipEvent_t e = poolEvent();
    hipCheck(hipEventRecord(e, stream_), "record timing event");
    ev_begin_[cls].push_back(e);
}

void Backend::timedEnd(int cls, double work) {
    if (!((timing_mask_ >> cls) & 1U) || !t_open_[cls]) return;
    t_open_[cls] = false;
    pending_work_[cls].push_back(work);
    const hipEvent_t e = poolEvent();
    hipCheck(hipEventRecord(e, stream_), "record timing event");
    ev_end_[cls].push_back(e);
}

void Backend::enableTiming(unsigned class_mask, int every) {
    if (every < 1) throw InvalidArgument("the timing period must be at least 1");
    timing_mask_ = class_mask;
    timing_every_ = every;
    for (int c = 0; c < kTimingClasses; ++c) {
        t_calls_[c] = 0;
        t_open_[c] = false;
    }
    if (t_total_.empty()) {
        t_total_.assign(kTimingClasses, 0.0);
        t_work_.assign(kTimingClasses, 0.0);
        t_count_.assign(kTimingClasses, 0);
    }
}

std::vector<std::string> Backend::timingNames() const {
    return {"k_wire", "k_density", "k_density_sort", "k_gemm", "k_gather", "controller", "k_wire_bins"};
}

void Backend::timing(std::vector<double>& total_us, std::vector<long long>& launches, std::vector<double>& work) {
    if (t_total_.empty()) {
        t_total_.assign(kTimingClasses, 0.0);
        t_work_.assign(kTimingClasses, 0.0);
        t_count_.assign(kTimingClasses, 0);
    }
    hipCheck(hipStreamSynchronize(stream_), "synchronize for timing");
    for (int c = 0; c < kTimingClasses; ++c) {
        for (std::size_t i = 0; i < ev_begin_[c].size() && i < ev_end_[c].size(); ++i) {
            float ms = 0.f;
            hipCheck(hipEventElapsedTime(&ms, ev_begin_[c][i], ev_end_[c][i]), "read timing event");
            t_total_[c] += 1000.0 * ms;
            t_count_[c] += 1;
            if (i < pending_work_[c].size()) t_work_[c] += pending_work_[c][i];
        }
        pending_work_[c].clear();
        ev_begin_[c].clear();
        ev_end_[c].clear();
    }
    ev_next_ = 0;  // every pooled event has completed
    total_us = t_total_;
    launches = t_count_;
    work = t_work_;
}

DensityArgs Backend::densityArgs() const {
    DensityArgs da{};
    da.size = size_.get();
    da.layer = layer_.get();
    da.tall = tall_.get();
    da.tall_ids = tall_ids_.get();
    da.n_tall = n_tall_;
    da.n = P_;
    da.r = region_;
    da.t = t_;
    da.fix = fix_;
    da.inv_fix = 1.0 / fix_;
    da.key = key_.get();
    da.hist = hist_.get();
    da.colprefix = colprefix_.get();
    da.keytotal = keytotal_.get();
    da.keystart = keystart_.get();
    da.perm = perm_.get();
    da.sorted_size = sorted_size_.get();
    da.static_charge = static_charge_.get();
    da.ldrho = nxp_;
    return da;
}

void Backend::sortParticles(const double2* pos) {
    DensityArgs da = densityArgs();
    da.pos[0] = pos;
    timedBegin(kTSort);
    launches_->sort(da);
    // algorithmic bytes: positions and sizes read twice, keys written and read, order and sizes written
    timedEnd(kTSort, 64.0 * P_);
}

WireArgs Backend::wireArgs(double gamma, bool want_grad) const {
    WireArgs wa{};
    wa.net_off = net_off_.get();
    wa.net_w = net_w_.get();
    wa.pin_part = pin_part_.get();
    wa.pin_off = pin_off_.get();
    wa.pin_dest = pin_dest_.get();
    wa.n_small = n_small_;
    wa.n_medium = n_medium_;
    wa.n_large = n_large_;
    for (int d = 0; d <= kSmallDegree + 1; ++d) {
        wa.small_net[d] = small_net_[d];
        wa.small_pin[d] = small_pin_[d];
    }
    wa.gamma = gamma;
    wa.want_grad = want_grad ? 1 : 0;
    return wa;
}

void Backend::enqueueEvaluate(const EvalSpec* specs, int count, double gamma, bool want_grad) {
    if (count < 1 || count > kSlots) throw std::logic_error("one or two evaluations per batch");
    for (int e = 0; e < count; ++e) {
        if (!specs[e].session && count != 1) throw std::logic_error("backend-mode evaluations are not batched");
        if (specs[e].density_from >= 0 && (e != 0 || specs[e].density_from >= kScalarBanks))
            throw std::logic_error("only the first evaluation may take the density of slot 0");
    }
    if (specs[0].reorder) sortParticles(specs[0].pos);

    WireArgs wa = wireArgs(gamma, want_grad);
    const bool nets = n_small_ + n_medium_ + n_large_ > 0;
    const int wire_rows = wireBlocks(n_small_, n_medium_, n_large_);
    int nw = 0;  // evaluations whose wirelength runs here
    for (int e = 0; e < count; ++e) {
        if (specs[e].wire_ready) {
            if (e != 0 || !want_grad || !nets) throw std::logic_error("only a first gradient evaluation is evaluated ahead");
            addReduction(kRedWire, specs[e].bank, wire_rows, gamma, 1.0);
            continue;
        }
        wa.pos[nw] = specs[e].pos;
        wa.pin_grad[nw] = slots_[e].pin_grad.get();
        wa.partials[nw] = partials(kRedWire, specs[e].bank);
        ++nw;
    }
    // The wirelength goes with the density's bin pass (one launch, k_wire_dbins) when both run
    // here and the launcher fuses them; otherwise it is launched on its own, first.
    const bool wire_with_bins = nw > 0 && nets && launches_->fusesWireBins() &&
                                std::any_of(specs, specs + count, [](const EvalSpec& s) { return s.density_from < 0; });
    if (wire_with_bins) {
        for (int e = 0; e < count; ++e)
            if (!specs[e].wire_ready) addReduction(kRedWire, specs[e].bank, wire_rows);
    } else if (nw > 0) {
        timedBegin(kTWire);
        if (nets) {
            launches_->wire(wa, nw);
            for (int e = 0; e < count; ++e)
                if (!specs[e].wire_ready) addReduction(kRedWire, specs[e].bank, wire_rows);
        } else {
            for (int e = 0; e < count; ++e) launches_->zero(dev_scal_ + specs[e].bank * kScalarCount + kHpwl, 2);
        }
        // algorithmic bytes per evaluation, by the convention of one pass over each array at the
        // path's nominal widths (4-byte indices and scalars, 8-byte positions): 12 per pin, 8 per
        // net, 16 per particle with the gradient and 8 without
        timedEnd(kTWire, nw * (12.0 * I_ + 8.0 * N_ + (want_grad ? 16.0 : 8.0) * P_));
    }

    // The evaluations whose density is computed here (fwd[0..nf)); an evaluation at the positions
    // of the one before it takes that one's density, potential and partial rows instead.
    int fwd[kSlots], nf = 0;
    for (int e = 0; e < count; ++e)
        if (specs[e].density_from < 0) fwd[nf++] = e;
    const int energy_rows = folded_ ? dctWorkgroups(dctTileFor(kFoldRows, dct_precision_, ny_, nx_), kFoldRows, ny_, nx_)
                                    : (nxp_ / 32) * (nyp_ / 32);
    const double bins = static_cast<double>(nx_) * ny_;
    for (int e = 0; e < count; ++e)
        if (specs[e].density_from >= 0) {
            addReduction(kRedMetrics, specs[e].bank, densityBinBlocks(region_), bins, dark_total_, placeable_total_,
                         specs[e].density_from);
            addReduction(kRedEnergy, specs[e].bank, energy_rows, 0.5 * region_.bw * region_.bh,
                         0.0, 0.0, specs[e].density_from);
        }

    const double W = region_.urx - region_.llx, H = region_.ury - region_.lly;
    SpectralArgs sp{nx_, ny_, W, H, 1.0 / (4.0 * nx_ * static_cast<double>(ny_)), {}};
    if (nf > 0) {
        DensityArgs da = densityArgs();
        for (int k = 0; k < nf; ++k) {
            const int e = fwd[k];
            da.pos[k] = specs[e].pos;
            da.grid[k] = slots_[e].grid.get();
            da.rho[k] = slots_[e].rho.get();
            da.partials[k] = partials(kRedMetrics, specs[e].bank);
            sp.partials[k] = partials(kRedEnergy, specs[e].bank);
        }
        // algorithmic bytes per evaluation (same convention): scatter 20 per particle and 12 per bin,
        // metrics and deviation 24 per bin
        timedBegin(kTBand);
        if (wire_with_bins) {
            launches_->deposit(da, nf);
            timedEnd(kTBand, nf * (20.0 * P_ + 12.0 * bins));
            timedBegin(kTWireBins);
            launches_->wireBins(wa, nw, da, nf);
            timedEnd(kTWireBins, nw * (12.0 * I_ + 8.0 * N_ + (want_grad ? 16.0 : 8.0) * P_) + nf * 24.0 * bins);
        } else {
            launches_->density(da, nf);
            timedEnd(kTBand, nf * (20.0 * P_ + 36.0 * bins));
        }
        for (int k = 0; k < nf; ++k)
            addReduction(kRedMetrics, specs[fwd[k]].bank, densityBinBlocks(region_), bins, dark_total_,
                         placeable_total_);
    }
    auto gemm = [&](const GemmBatch& g, int batch, int M, int N, int K, int lda, int ldb, int ldc, int epi,
                    int product) {
        timedBegin(kTDct);
        launches_->gemm(g, batch, M, N, K, lda, ldb, ldc, epi, sp, dct_precision_, product);
        timedEnd(kTDct, 2.0 * M * static_cast<double>(N) * K * batch);
    };
    // half-length products: `rows` x `cols` outputs over a contraction of half the transform side
    auto dct = [&](int mode, const DctBatch& d, int batch, int lda, int ldb, int ldc, int step) {
        timedBegin(kTDct);
        launches_->dct(mode, d, batch, ny_, nx_, lda, ldb, ldc, step, dct_precision_, sp);
        const double half = (mode == kFoldCols || mode == kMirrorCols) ? 0.5 * nx_ : 0.5 * ny_;
        timedEnd(kTDct, 2.0 * ny_ * static_cast<double>(nx_) * half * batch * (mode == kMirrorCols ? 2 : 1));
    };
    if (folded_) {
        // forward: T = fold_x(rho) ; chat = fold_y(T) -> phihat (+ energy)
        if (nf > 0) {
            DctBatch f1{}, f2{};
            for (int k = 0; k < nf; ++k) {
                const Slot& sl = slots_[fwd[k]];
                f1.A[k] = sl.rho.get();
                f1.B[k] = bxf_.get();
                f1.C[k] = sl.tmp.get();
                f2.A[k] = ayf_.get();
                f2.B[k] = sl.tmp.get();
                f2.C[k] = sl.phi.get();
            }
            dct(kFoldCols, f1, nf, nxp_, 0, nxp_, 1);
            dct(kFoldRows, f2, nf, 0, nxp_, nxp_, 1);
            for (int k = 0; k < nf; ++k)
                addReduction(kRedEnergy, specs[fwd[k]].bank, energy_rows, 0.5 * region_.bw * region_.bh);
        }
        if (!want_grad) return;
        // [T1 | T2] = mirror_x(phihat) ; Ex = mirror_y(T2, cosine) ; Ey = mirror_y(T1, sine)
        DctBatch i1{}, i2{};
        for (int e = 0; e < count; ++e) {
            const Slot& sl = slots_[e];
            i1.A[e] = sl.phi.get();
            i1.B[e] = bxi_.get();
            i1.C[e] = sl.t12.get();
            i2.A[2 * e] = ayi_.get();
            i2.B[2 * e] = sl.t12.get() + nxp_;
            i2.C[2 * e] = sl.field.get();
            i2.sigma[2 * e] = 1.f;
            i2.A[2 * e + 1] = syi_.get();
            i2.B[2 * e + 1] = sl.t12.get();
            i2.C[2 * e + 1] = sl.field.get() + 1;
            i2.sigma[2 * e + 1] = -1.f;
        }
        dct(kMirrorCols, i1, count, nxp_, 0, 2 * nxp_, 1);
        dct(kMirrorRows, i2, 2 * count, 0, 2 * nxp_, nxp_, 2);
    } else {
        // forward: T = rho * Bx ; chat = Ay * T -> phihat (+ energy)
        if (nf > 0) {
            GemmBatch f1{}, f2{};
            for (int k = 0; k < nf; ++k) {
                const Slot& sl = slots_[fwd[k]];
                f1.A[k] = sl.rho.get();
                f1.B[k] = bxf_.get();
                f1.C[k] = sl.tmp.get();
                f2.A[k] = ayf_.get();
                f2.B[k] = sl.tmp.get();
                f2.C[k] = sl.phi.get();
            }
            gemm(f1, nf, nyp_, nxp_, nxp_, nxp_, nxp_, nxp_, kEpiStore, 0);
            gemm(f2, nf, nyp_, nxp_, nyp_, nyp_, nxp_, nxp_, kEpiSpectral, 1);
            for (int k = 0; k < nf; ++k)
                addReduction(kRedEnergy, specs[fwd[k]].bank, energy_rows, 0.5 * region_.bw * region_.bh);
        }
        if (!want_grad) return;
        // [T1 | T2] = phihat * [Cx_inv | Sx_inv] ; Ex = Ay_inv * T2 ; Ey = Sy_inv * T1
        GemmBatch i1{}, i2{};
        for (int e = 0; e < count; ++e) {
            const Slot& sl = slots_[e];
            i1.A[e] = sl.phi.get();
            i1.B[e] = bxi_.get();
            i1.C[e] = sl.t12.get();
            i2.A[2 * e] = ayi_.get();
            i2.B[2 * e] = sl.t12.get() + nxp_;
            i2.C[2 * e] = sl.field.get();
            i2.A[2 * e + 1] = syi_.get();
            i2.B[2 * e + 1] = sl.t12.get();
            i2.C[2 * e + 1] = sl.field.get() + 1;
        }
        gemm(i1, count, nyp_, 2 * nxp_, nxp_, nxp_, 2 * nxp_, 2 * nxp_, kEpiStore, 2);
        gemm(i2, 2 * count, nyp_, nxp_, nyp_, nyp_, 2 * nxp_, nxp_, kEpiInterleave, 3);
    }
    have_fields_ = true;

    GatherArgs ga{};
    ga.charge = charge_.get();
    ga.ldf = nxp_;
    ga.r = region_;
    ga.ppin_off = ppin_off_.get();
    ga.ppin_idx = ppin_idx_.get();
    ga.n = P_;
    for (int e = 0; e < count; ++e) {
        const Slot& sl = slots_[e];
        ga.pos[e] = specs[e].pos;
        ga.field[e] = reinterpret_cast<const float2*>(sl.field.get());
        ga.pin_grad[e] = sl.pin_grad.get();
        ga.partials[e] = partials(kRedGather, specs[e].bank);
        const SessionOut* so = specs[e].session;
        if (so) {
            ga.precond = so->precond;
            ga.penalty = so->penalty;
            ga.inv_strength = so->inv_strength;
            ga.raw_out[e] = so->raw;
            ga.pre_out[e] = so->pre;
            ga.prev_pos[e] = so->prev_pos;
            ga.px = so->px;
            ga.pg = so->pg;
            ga.ck_out[e] = so->ck_out;
        } else {
            ga.wire_out = wire_out_.get();
            ga.density_out = density_out_.get();
        }
    }
    timedBegin(kTGather);
    launches_->gather(ga, count);
    for (int e = 0; e < count; ++e) addReduction(kRedGather, specs[e].bank, particleBlocks(P_));
    // algorithmic bytes per evaluation (same convention): 20 per particle and 8 per bin
    timedEnd(kTGather, count * (20.0 * P_ + 8.0 * static_cast<double>(nx_) * ny_));
}

void StreamLaunches::sort(const DensityArgs& a) { launchDensitySort(a, s_); }
void StreamLaunches::wire(const WireArgs& a, int count) { launchWire(a, count, s_); }
void StreamLaunches::zero(double* p, int n) { launchZero(p, static_cast<size_t>(n), s_); }
void StreamLaunches::density(const DensityArgs& a, int count) { launchDensity(a, count, s_); }
void StreamLaunches::deposit(const DensityArgs& a, int count) { launchDeposit(a, count, s_); }
void StreamLaunches::wireBins(const WireArgs& w, int nw, const DensityArgs& d, int nd) {
    launchWireBins(w, nw, d, nd, s_);
}
void StreamLaunches::gemm(const GemmBatch& g, int batch, int M, int N, int K, int lda, int ldb, int ldc, int epilogue,
                          const SpectralArgs& sp, int precision, int) {
    launchGemm(g, batch, M, N, K, lda, ldb, ldc, epilogue, sp, precision, s_);
}
void StreamLaunches::dct(int mode, const DctBatch& d, int batch, int rows, int cols, int lda, int ldb, int ldc,
                         int step, int precision, const SpectralArgs& sp) {
    launchDct(mode, d, batch, rows, cols, lda, ldb, ldc, step, precision, sp, s_);
}
void StreamLaunches::gather(const GatherArgs& a, int count) { launchGather(a, count, s_); }
void StreamLaunches::lookahead(const LookaheadArgs& l) { launchLookahead(l, s_); }
void StreamLaunches::propose(const ProposeArgs& p) { launchPropose(p, s_); }
void StreamLaunches::finalize(const RedList& l) {
    RedList m = l;
    m.flag = nullptr;
    launchFinalize(m, s_);
}
void StreamLaunches::copy(double2* dst, const double2* src, int n) {
    if (n > 0)
        hipCheck(hipMemcpyAsync(dst, src, static_cast<size_t>(n) * sizeof(double2), hipMemcpyDeviceToDevice, s_),
                 "copy positions");
}

template <typename E>
void RecordLaunches::add(int stage, int kernel, const E& e, int blocks, std::size_t lds) {
    static_assert(sizeof(E) <= kMaxEntryBytes && alignof(E) <= 16, "batch entry does not fit a recorded launch");
    if (blocks <= 0) return;
    BatchOp& op = ops.emplace_back();
    op.stage = stage;
    op.kernel = kernel;
    op.blocks = blocks;
    op.bytes = static_cast<int>(sizeof(E));
    op.lds = lds;
    std::memcpy(op.entry, &e, sizeof(E));
}
void RecordLaunches::sort(const DensityArgs& a) {
    if (a.n == 0) return;
    const std::size_t ny = static_cast<std::size_t>(a.r.ny);
    const SortEntry h = sortEntry(a, 0), c = sortEntry(a, 1), p = sortEntry(a, 2);
    add(kStageSortHist, kBatchSortHist, h, h.gx, ny * sizeof(int));
    add(kStageSortScan, kBatchSortScan, c, c.gx);
    add(kStageSortPlace, kBatchSortPlace, p, p.gx, (2 * ny + 1) * sizeof(int));
}
void RecordLaunches::wire(const WireArgs& a, int count) {
    if (count < 1 || count > kMaxEval) throw std::invalid_argument("bad evaluation batch");
    const WireEntry e = wireEntry(a, count);
    add(kStageWire, kBatchWire, e, e.gx * e.gy);
}
void RecordLaunches::zero(double* p, int n) {
    const ZeroEntry e = zeroEntry(p, n);
    add(kStageZero, kBatchZero, e, e.gx);
}
void RecordLaunches::density(const DensityArgs& a, int count) {
    if (count < 1 || count > kMaxEval) throw std::invalid_argument("bad evaluation batch");
    const DensityEntry d = depositEntry(a, count), b = binsEntry(a, count);
    add(kStageDeposit, kBatchDeposit, d, d.gx * d.gy);
    add(kStageBins, kBatchBins, b, b.gx * b.gy);
}
void RecordLaunches::gemm(const GemmBatch& g, int batch, int M, int N, int K, int lda, int ldb, int ldc, int epilogue,
                          const SpectralArgs& sp, int precision, int product) {
    if (precision != 0) throw std::invalid_argument("only the FP32 DCT products are built");
    if (product < 0 || product >= 4) throw std::logic_error("an evaluation chain has four spectral products");
    const GemmEntry e = gemmEntry(g, batch, M, N, K, lda, ldb, ldc, epilogue, sp);
    const int k = epilogue == kEpiSpectral ? kBatchGemmSpectral
                  : epilogue == kEpiInterleave ? kBatchGemmInterleave : kBatchGemmStore;
    add(kStageSpectral + product, k, e, e.gx * e.gy * e.gz);
}
void RecordLaunches::dct(int mode, const DctBatch& d, int batch, int rows, int cols, int lda, int ldb, int ldc,
                         int step, int precision, const SpectralArgs& sp) {
    const DctEntry e = dctEntry(mode, d, batch, rows, cols, lda, ldb, ldc, step, precision, sp);
    // the modes are numbered in chain order: fold-cols, fold-rows, mirror-cols, mirror-rows
    add(kStageSpectral + mode, kBatchDct + dctVariant(e), e, e.gx * e.gy * e.gz);
}
void RecordLaunches::gather(const GatherArgs& a, int count) {
    if (count < 1 || count > kMaxEval) throw std::invalid_argument("bad evaluation batch");
    const GatherEntry e = gatherEntry(a, count);
    add(kStageGather, kBatchGather, e, e.gx * e.gy);
}
void RecordLaunches::lookahead(const LookaheadArgs& l) {
    const LookaheadEntry e = lookaheadEntry(l);
    add(kStageLookahead, kBatchLookahead, e, e.gx);
}
void RecordLaunches::propose(const ProposeArgs& p) {
    const ProposeEntry e = proposeEntry(p);
    add(kStagePropose, kBatchPropose, e, e.gx);
}
void RecordLaunches::finalize(const RedList& l) {
    for (int k = 0; k < l.n; ++k) add(kStageFinalize, kBatchFinalize, l.job[k], 1);
}
void RecordLaunches::copy(double2* dst, const double2* src, int n) {
    const CopyEntry e = copyEntry(dst, src, n);
    add(kStageCopy, kBatchCopy, e, e.gx);
}

void Backend::addReduction(int kind, int bank, int blocks, double p0, double p1, double p2, int from) {
    if (pending_.n == kMaxRedJobs) throw std::logic_error("too many pending reductions");
    pending_.job[pending_.n++] = RedJob{partials(kind, from >= 0 ? from : bank), blocks, kind,
                                        dev_scal_ + bank * kScalarCount, scal_copy_.get() + bank * kScalarCount, p0,
                                        p1, p2};
}

void Backend::finalizePending() {
    if (pending_.n == 0) return;
    RedList l = pending_;
    l.seq = 0;
    l.flag = nullptr;
    launches_->finalize(l);
    pending_.n = 0;
}

void Backend::beginCapture(int shape) {
    if (timing_mask_ != 0 || !dynamic_cast<StreamLaunches*>(launches_) || capture_shape_ >= 0) return;
    hipCheck(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "begin graph capture");
    capture_shape_ = shape;
}

const HostScalars& Backend::fetchScalars() {
    const int jobs = pending_.n;
    if (jobs == 0 && capture_shape_ < 0) {
        ahead_pending_ = false;  // nothing was evaluated to go ahead from
        hipCheck(hipStreamSynchronize(stream_), "synchronize placement evaluation");
        return *host_scal_;
    }
    pending_.seq = ++seq_;
    pending_.flag = dev_done_;
    const bool nets = n_small_ + n_medium_ + n_large_ > 0;
    auto aheadArgs = [&] {
        WireArgs wa = wireArgs(0.0, true);
        wa.pos[0] = ahead_.pos;
        wa.pin_grad[0] = slots_[0].pin_grad.get();
        wa.partials[0] = partials(kRedWire, ahead_.bank);
        wa.ahead_overflow = ahead_.overflow;
        wa.ahead_denominator = ahead_.denominator;
        wa.ahead_final = ahead_.final_smoothing;
        wa.ahead_span = ahead_.span;
        wa.ahead_scale = ahead_.scale;
        wa.ahead_rows = wireBlocks(n_small_, n_medium_, n_large_);
        return wa;
    };
    // The wirelength evaluated ahead goes into the finalize's launch (k_finalize_wire) when this
    // finalize forms the overflow its smoothing law reads (a pending metrics reduction whose
    // device-memory copy holds it) and no graph is being captured.
    const RedJob* metrics = nullptr;
    if (jobs > 0 && ahead_pending_ && nets && capture_shape_ < 0)
        for (int k = 0; k < jobs; ++k) {
            const RedJob& j = pending_.job[k];
            if (j.kind == kRedMetrics && j.out_dev && j.out_dev + kOptDensity + 1 == ahead_.overflow) metrics = &j;
        }
    if (metrics) {
        ahead_pending_ = false;
        WireArgs wa = aheadArgs();
        wa.ahead_metrics = metrics->partials;
        wa.ahead_metric_rows = metrics->blocks;
        timedBegin(kTWire);
        launchFinalizeWire(pending_, wa, stream_);
        timedEnd(kTWire, 12.0 * I_ + 8.0 * N_ + 16.0 * P_);
    } else if (jobs > 0) {
        launchFinalize(pending_, stream_);
    }
    pending_.n = 0;
    auto launchAhead = [&] {
        if (!ahead_pending_) return;
        ahead_pending_ = false;
        if (!nets) return;
        timedBegin(kTWire);
        launchWire(aheadArgs(), 1, stream_);
        timedEnd(kTWire, 12.0 * I_ + 8.0 * N_ + 16.0 * P_);
    };
    if (capture_shape_ >= 0) {
        const int shape = capture_shape_;
        capture_shape_ = -1;
        hipGraph_t g = nullptr;
        hipCheck(hipStreamEndCapture(stream_, &g), "end graph capture");
        std::unique_ptr<std::remove_pointer_t<hipGraph_t>, decltype(&hipGraphDestroy)> own(g, hipGraphDestroy);
        auto it = std::find_if(graphs_.begin(), graphs_.end(), [&](const auto& e) { return e.first == shape; });
        bool updated = false;
        if (it != graphs_.end()) {
            hipGraphNode_t bad = nullptr;
            hipGraphExecUpdateResult why{};
            updated = hipGraphExecUpdate(it->second, g, &bad, &why) == hipSuccess && why == hipGraphExecUpdateSuccess;
            if (!updated) {
                (void)hipGetLastError();
                (void)hipGraphExecDestroy(it->second);
                graphs_.erase(it);
            }
        }
        if (!updated) {
            hipGraphExec_t ex = nullptr;
            hipCheck(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0), "instantiate graph");
            graphs_.emplace_back(shape, ex);
            it = graphs_.end() - 1;
        }
        hipCheck(hipGraphLaunch(it->second, stream_), "launch graph");
        launchAhead();
        if (jobs == 0) {
            hipCheck(hipStreamSynchronize(stream_), "synchronize placement evaluation");
            return *host_scal_;
        }
    } else {
        launchAhead();
    }
    // Poll the completion words: the results are on the host as soon as the last one lands, some
    // microseconds before the stream reports the kernel retired.  The stream is queried now and
    // then so that a failed launch surfaces as an error instead of a hang.
    auto written = [&] {
        for (int b = 0; b < jobs; ++b)
            if (__atomic_load_n(&host_scal_->done[b], __ATOMIC_ACQUIRE) != seq_) return false;
        return true;
    };
    for (long spins = 0; !written(); ++spins) {
        if ((spins & 1023) != 1023) continue;
        const hipError_t st = hipStreamQuery(stream_);
        if (st == hipErrorNotReady) continue;
        hipCheck(st, "synchronize placement evaluation");
        if (!written()) throw std::runtime_error("reduction finalize retired without writing its completion words");
        break;
    }
    return *host_scal_;
}

void Backend::toScalars(const HostScalars& hs, mp_gp_scalars* out, int bank) const {
    const double* v = hs.bank(bank);
    out->hpwl = v[kHpwl];
    out->smooth_wirelength = v[kSmooth];
    out->electrostatic_energy = v[kEnergy];
    auto fill = [&](int base, mp_gp_density_metrics& m) {
        m.total_overflow_area = v[base + 0];
        m.normalized_overflow = v[base + 1];
        m.maximum_density = v[base + 2];
        m.average_density = v[base + 3];
        m.total_charge_area = v[base + 4];
        m.normalization_area = v[base + 5];
        m.placeable_area = v[base + 6];
        m.dark_area = v[base + 7];
    };
    fill(kOptDensity, out->optimizer_density);
    fill(kDesignDensity, out->design_density);
}

namespace {
constexpr std::size_t kStageMinBytes = 1u << 20;  // designs of 65,536 particles and more
constexpr std::size_t kStageGrain = 1u << 19;     // bytes per host copy thread
}  // namespace

void Backend::evaluateHost(const double* xy, double gamma, bool want_grad, mp_gp_scalars* out, double* wire_grad,
                           double* density_grad) {
    if (!std::isfinite(gamma) || gamma <= 0.0)
        throw InvalidArgument("HIP global placement requires positive finite smoothing.");
    if (P_ > 0 && xy == nullptr) throw InvalidArgument("HIP placement positions do not match the particle count.");
    if (want_grad && P_ > 0 && (wire_grad == nullptr || density_grad == nullptr))
        throw InvalidArgument("gradient outputs are required when want_grad is set");
    setDevice();
    hipStream_t s = stream_;
    const std::size_t bytes = static_cast<std::size_t>(P_) * sizeof(double2);
    // The positions go up through the runtime's pageable path (measured faster than staging them
    // for one 4 MB vector); a large design's two gradient vectors come down into pinned memory at the
    // link's pinned rate and are copied out on several host threads.
    pos_.upload(reinterpret_cast<const double2*>(xy), P_, s, "upload particle positions");
    const bool staged = want_grad && bytes >= kStageMinBytes;
    if (staged && host_stage_ == nullptr)
        hipCheck(hipHostMalloc(reinterpret_cast<void**>(&host_stage_), 2 * bytes, hipHostMallocDefault),
                 "allocate pinned staging");
    const EvalSpec spec{pos_.get(), nullptr, 0, true};
    enqueueEvaluate(&spec, 1, gamma, want_grad);
    if (want_grad && P_ > 0) {
        hipCheck(hipMemcpyAsync(staged ? static_cast<void*>(host_stage_) : wire_grad, wire_out_.get(), bytes,
                                hipMemcpyDeviceToHost, s),
                 "download wire gradients");
        hipCheck(hipMemcpyAsync(staged ? static_cast<void*>(host_stage_ + P_) : density_grad, density_out_.get(), bytes,
                                hipMemcpyDeviceToHost, s),
                 "download density gradients");
    }
    const HostScalars& h = fetchScalars();
    // the gradient copies follow the reductions on the stream: wait for them to finish
    if (want_grad && P_ > 0) {
        hipCheck(hipStreamSynchronize(s), "download gradients");
        if (staged) {
            parallelFor(2 * bytes, kStageGrain, [&](std::size_t a, std::size_t b) {
                // [0, bytes): wire gradient, [bytes, 2 bytes): density gradient
                const char* src = reinterpret_cast<const char*>(host_stage_);
                while (a < b) {
                    const bool wire = a < bytes;
                    const std::size_t end = wire ? std::min(b, bytes) : b;
                    char* dst = wire ? reinterpret_cast<char*>(wire_grad) + a : reinterpret_cast<char*>(density_grad) + (a - bytes);
                    std::memcpy(dst, src + a, end - a);
                    a = end;
                }
            });
        }
    }
    toScalars(h, out);
    if (!std::isfinite(out->smooth_wirelength) || !std::isfinite(out->hpwl) || !std::isfinite(out->electrostatic_energy))
        throw std::runtime_error("HIP global-placement evaluation produced a non-finite metric.");
}

void Backend::debugGrids(double* source, double* potential, double* field_x, double* field_y) {
    setDevice();
    if (!have_fields_) throw std::runtime_error("no gradient evaluation has produced field grids yet");
    const std::size_t Gp = static_cast<std::size_t>(nxp_) * nyp_;
    const Slot& sl = slots_[0];
    if (potential) {  // phi = Ay_inv * T1
        SpectralArgs sp{};
        if (folded_) {
            DctBatch d{{ayi_.get()}, {sl.t12.get()}, {pot_.get()}, {1.f}};
            launchDct(kMirrorRows, d, 1, ny_, nx_, 0, 2 * nxp_, nxp_, 1, dct_precision_, sp, stream_);
        } else {
            GemmBatch g{{ayi_.get()}, {sl.t12.get()}, {pot_.get()}};
            launchGemm(g, 1, nyp_, nxp_, nyp_, nyp_, 2 * nxp_, nxp_, kEpiStore, sp, dct_precision_, stream_);
        }
    }
    std::vector<float> buf(2 * Gp);
    auto pull = [&](const DeviceArray<float>& a, double* dst, int stride, int offset) {
        if (!dst) return;
        hipCheck(hipMemcpyAsync(buf.data(), a.get(), stride * Gp * sizeof(float), hipMemcpyDeviceToHost, stream_),
                 "read grid");
        hipCheck(hipStreamSynchronize(stream_), "read grid");
        for (int r = 0; r < ny_; ++r)
            for (int c = 0; c < nx_; ++c)
                dst[static_cast<std::size_t>(r) * nx_ + c] = buf[stride * (static_cast<std::size_t>(r) * nxp_ + c) + offset];
    };
    pull(sl.rho, source, 1, 0);
    if (source) {  // centre it as ElectrostaticField::solve does
        double mean = 0.0;
        for (std::size_t i = 0; i < static_cast<std::size_t>(nx_) * ny_; ++i) mean += source[i];
        mean /= static_cast<double>(nx_) * ny_;
        for (std::size_t i = 0; i < static_cast<std::size_t>(nx_) * ny_; ++i) source[i] -= mean;
    }
    pull(pot_, potential, 1, 0);
    pull(sl.field, field_x, 2, 0);
    pull(sl.field, field_y, 2, 1);
}

// ---------------------------------------------------------------------------------------------
// Shared C-ABI plumbing.
thread_local std::string g_last_error;

mp_gp_status translate(const std::exception_ptr& e) {
    try {
        std::rethrow_exception(e);
    } catch (const std::invalid_argument& x) {
        g_last_error = x.what();
        return MP_GP_INVALID_ARGUMENT;
    } catch (const std::out_of_range& x) {
        g_last_error = x.what();
        return MP_GP_INVALID_ARGUMENT;
    } catch (const std::exception& x) {
        g_last_error = x.what();
        return MP_GP_RUNTIME_ERROR;
    } catch (...) {
        g_last_error = "unknown error";
        return MP_GP_RUNTIME_ERROR;
    }
}

// Device discovery and allocation limit (.cu:1156-1207), HIP device policy: any present device.
std::size_t allocationLimit(int device, std::uint64_t maximum_bytes, std::string& reason) {
    int count = 0;
    const hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) {
        reason = std::string("HIP device discovery failed: ") + (e == hipSuccess ? "no device" : hipGetErrorString(e));
        return 0;
    }
    if (device < 0 || device >= count) {
        reason = "Requested HIP device does not exist on this host.";
        return 0;
    }
    if (hipSetDevice(device) != hipSuccess) {
        reason = "Unable to select the requested HIP device.";
        return 0;
    }
    std::size_t free_bytes = 0, total_bytes = 0;
    if (hipMemGetInfo(&free_bytes, &total_bytes) != hipSuccess) {
        reason = "Unable to query free HIP memory.";
        return 0;
    }
    if (free_bytes <= kDeviceReserveBytes) {
        reason = "Requested HIP device does not have the required 4 GiB memory reserve.";
        return 0;
    }
    const std::size_t limit = std::min<std::size_t>(maximum_bytes, free_bytes - kDeviceReserveBytes);
    if (limit == 0) reason = "HIP memory limit leaves no allocatable device memory.";
    return limit;
}

}  // namespace mpgp

using namespace mpgp;

extern "C" {

int mp_gp_compiled(void) { return 1; }

const char* mp_gp_last_error(void) { return g_last_error.c_str(); }

mp_gp_status mp_gp_create(const mp_gp_static* layout, mp_gp_backend** out) {
    g_last_error.clear();
    if (out == nullptr || layout == nullptr) {
        g_last_error = "null layout or output handle";
        return MP_GP_INVALID_ARGUMENT;
    }
    *out = nullptr;
    if (layout->maximum_device_bytes == 0) {
        g_last_error = "HIP global placement requires a non-zero memory budget.";
        return MP_GP_INVALID_ARGUMENT;
    }
    std::string reason;
    const std::size_t limit = allocationLimit(layout->device, layout->maximum_device_bytes, reason);
    if (limit == 0) {
        g_last_error = reason;
        return MP_GP_UNAVAILABLE;
    }
    try {
        auto h = std::make_unique<mp_gp_backend>();
        h->owned = std::make_unique<Backend>(*layout, limit);
        h->impl = h->owned.get();
        *out = h.release();
        return MP_GP_OK;
    } catch (const std::exception& e) {
        // the factory turns every construction failure into "unavailable + reason" (.cu:1209-1215)
        g_last_error = e.what();
        return MP_GP_UNAVAILABLE;
    }
}

mp_gp_status mp_gp_evaluate(mp_gp_backend* b, const double* xy, double smoothing, int want_grad, mp_gp_scalars* out,
                            double* wire_grad, double* density_grad) {
    g_last_error.clear();
    if (b == nullptr || out == nullptr) {
        g_last_error = "null backend or output";
        return MP_GP_INVALID_ARGUMENT;
    }
    try {
        b->impl->evaluateHost(xy, smoothing, want_grad != 0, out, wire_grad, density_grad);
        return MP_GP_OK;
    } catch (...) {
        return translate(std::current_exception());
    }
}

int mp_gp_device(const mp_gp_backend* b) { return b ? b->impl->device() : -1; }

uint64_t mp_gp_reserved_bytes(const mp_gp_backend* b) { return b ? b->impl->reservedBytes() : 0; }

void mp_gp_destroy(mp_gp_backend* b) {
    if (b && b->owned) delete b;  // session views are released with their session
}

mp_gp_status mp_gp_debug_grids(mp_gp_backend* b, double* source, double* potential, double* field_x, double* field_y) {
    g_last_error.clear();
    if (b == nullptr) {
        g_last_error = "null backend";
        return MP_GP_INVALID_ARGUMENT;
    }
    try {
        b->impl->debugGrids(source, potential, field_x, field_y);
        return MP_GP_OK;
    } catch (...) {
        return translate(std::current_exception());
    }
}

}  // extern "C"

// ---- autotuner timing of the half-length product tiles ---------------------------------------
namespace mpgp {
namespace {
struct DeviceScratchBuffer {
    void* p = nullptr;
    explicit DeviceScratchBuffer(std::size_t bytes) { hipCheck(hipMalloc(&p, bytes), "allocate tuning buffer"); }
    ~DeviceScratchBuffer() {
        if (p) (void)hipFree(p);
    }
    DeviceScratchBuffer(const DeviceScratchBuffer&) = delete;
    DeviceScratchBuffer& operator=(const DeviceScratchBuffer&) = delete;
};
}  // namespace
}  // namespace mpgp

namespace mpgp {
namespace {
// One product of a given tile on operands of the evaluation chain's shapes (gp_backend.cpp,
// enqueueEvaluate): data grids of ny x 2 nx floats per product (the widest, [T1 | T2] and the
// interleaved field) and factors of 2 n^2 floats, on device 0 and a stream of its own.
struct TileProduct {
    int nx, ny, precision, mode, tile, batch;
    std::size_t grid, factor;
    DeviceScratchBuffer data, fac, out, part;
    hipStream_t s = nullptr;
    DctBatch d{};
    SpectralArgs sp{};
    int lda = 0, ldb = 0, ldc = 0, step = 1;

    static std::size_t gridOf(int nx, int ny) { return static_cast<std::size_t>(ny) * 2 * nx; }
    static std::size_t factorOf(int nx, int ny) {
        return 2 * static_cast<std::size_t>(std::max(nx, ny)) * std::max(nx, ny);
    }
    static int batchOf(int mode, int evaluations) { return mode == kMirrorRows ? 2 * evaluations : evaluations; }
    static void check(int nx, int ny, int precision, int mode, int tile) {
        if (!foldedDct(nx, ny)) throw std::invalid_argument("half-length products need both bin counts multiples of 64");
        if (mode < kFoldCols || mode > kMirrorRows) throw std::invalid_argument("unknown DCT product mode");
        if (precision != MP_DCT_FP32 && precision != MP_DCT_BF16X3) throw std::invalid_argument("unknown DCT precision");
        if (!dctTileValid(tile, mode, ny, nx)) throw std::invalid_argument("DCT tile does not divide the product");
        std::string reason;
        if (allocationLimit(0, ~0ull, reason) == 0) throw std::runtime_error(reason);
    }
    TileProduct(int nx_, int ny_, int precision_, int mode_, int tile_, int evaluations)
        : nx(nx_), ny(ny_), precision(precision_), mode(mode_), tile(tile_), batch(batchOf(mode_, evaluations)),
          grid(gridOf(nx_, ny_)), factor(factorOf(nx_, ny_)), data(sizeof(float) * grid * batch),
          fac(sizeof(float) * factor), out(sizeof(float) * grid * batch),
          part(sizeof(double) * static_cast<std::size_t>(dctWorkgroups(tile_, mode_, ny_, nx_)) * batch) {
        hipCheck(hipStreamCreate(&s), "create tuning stream");
        hipCheck(hipMemsetAsync(out.p, 0, sizeof(float) * grid * batch, s), "clear tuning output");
        const int nxp = nx;  // multiples of 64 need no padding
        sp = SpectralArgs{nx, ny, static_cast<double>(nx), static_cast<double>(ny), 1.0 / (4.0 * nx * ny), {}};
        for (int e = 0; e < batch; ++e) {
            const bool data_left = mode == kFoldCols || mode == kMirrorCols;
            float* dg = static_cast<float*>(data.p) + grid * e;
            d.A[e] = data_left ? dg : static_cast<float*>(fac.p);
            d.B[e] = data_left ? static_cast<float*>(fac.p) : dg;
            d.C[e] = static_cast<float*>(out.p) + grid * e;
            d.sigma[e] = e % 2 ? -1.f : 1.f;
            sp.partials[e] = static_cast<double*>(part.p) + static_cast<std::size_t>(dctWorkgroups(tile, mode, ny, nx)) * e;
        }
        ldc = nxp;
        switch (mode) {
            case kFoldCols: lda = nxp; break;
            case kFoldRows: ldb = nxp; break;
            case kMirrorCols: lda = nxp; ldc = 2 * nxp; break;
            default: ldb = 2 * nxp; step = 2; break;
        }
    }
    ~TileProduct() {
        if (s) (void)hipStreamDestroy(s);
    }
    void fill(const float* host_data, const float* host_factors) {
        hipCheck(hipMemcpyAsync(data.p, host_data, sizeof(float) * grid * batch, hipMemcpyHostToDevice, s),
                 "upload tile operands");
        hipCheck(hipMemcpyAsync(fac.p, host_factors, sizeof(float) * factor, hipMemcpyHostToDevice, s),
                 "upload tile factors");
        hipCheck(hipStreamSynchronize(s), "upload tile operands");
    }
    void launch() { launchDctTile(tile, mode, d, batch, ny, nx, lda, ldb, ldc, step, precision, sp, s); }
};
}  // namespace
}  // namespace mpgp

extern "C" mp_gp_status mp_gp_dct_tile_time(int32_t nx, int32_t ny, int32_t precision, int32_t mode, int32_t tile,
                                            int32_t evaluations, int32_t reps, double* microseconds) {
    using namespace mpgp;
    g_last_error.clear();
    try {
        if (!microseconds || reps < 1 || evaluations < 1 || evaluations > kMaxEval)
            throw std::invalid_argument("bad tuning arguments");
        TileProduct::check(nx, ny, precision, mode, tile);
        TileProduct t(nx, ny, precision, mode, tile, evaluations);
        {
            std::vector<float> h(std::max(t.grid * t.batch, t.factor));
            for (std::size_t i = 0; i < h.size(); ++i) h[i] = static_cast<float>((i * 2654435761u % 1000) * 1e-3 - 0.5);
            t.fill(h.data(), h.data());
        }
        for (int r = 0; r < 3; ++r) t.launch();
        hipEvent_t a, b;
        hipCheck(hipEventCreate(&a), "create tuning event");
        hipCheck(hipEventCreate(&b), "create tuning event");
        hipCheck(hipEventRecord(a, t.s), "record tuning event");
        for (int r = 0; r < reps; ++r) t.launch();
        hipCheck(hipEventRecord(b, t.s), "record tuning event");
        hipCheck(hipEventSynchronize(b), "time tile");
        float ms = 0.f;
        hipCheck(hipEventElapsedTime(&ms, a, b), "read tuning events");
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        *microseconds = 1000.0 * ms / reps;
        return MP_GP_OK;
    } catch (...) {
        return translate(std::current_exception());
    }
}

extern "C" mp_gp_status mp_gp_dct_tile_run(int32_t nx, int32_t ny, int32_t precision, int32_t mode, int32_t tile,
                                           const float* data, int64_t data_count, const float* factors,
                                           int64_t factor_count, float* out) {
    using namespace mpgp;
    g_last_error.clear();
    try {
        if (!data || !factors || !out) throw std::invalid_argument("null tile operand");
        TileProduct::check(nx, ny, precision, mode, tile);
        const int batch = TileProduct::batchOf(mode, 1);
        if (data_count != static_cast<int64_t>(TileProduct::gridOf(nx, ny) * batch) ||
            factor_count != static_cast<int64_t>(TileProduct::factorOf(nx, ny)))
            throw std::invalid_argument("tile operand sizes do not match the product");
        TileProduct t(nx, ny, precision, mode, tile, 1);
        t.fill(data, factors);
        t.launch();
        hipCheck(hipMemcpyAsync(out, t.out.p, sizeof(float) * t.grid * t.batch, hipMemcpyDeviceToHost, t.s),
                 "download tile output");
        hipCheck(hipStreamSynchronize(t.s), "run tile");
        return MP_GP_OK;
    } catch (...) {
        return translate(std::current_exception());
    }
}

extern "C" int32_t mp_gp_dct_tile_for(int32_t nx, int32_t ny, int32_t precision, int32_t mode) {
    return mpgp::foldedDct(nx, ny) ? mpgp::dctTileFor(mode, precision, ny, nx) : -1;
}
