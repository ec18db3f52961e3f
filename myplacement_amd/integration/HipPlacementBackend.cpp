// HipPlacementBackend.cpp — the MyPlacement-side binding of the MI355X global-placement backend.
//
// The file a MyPlacement maintainer adds as src/placement/HipPlacementBackend.cpp: it implements the
// reference's abstract boundary myplacement::detail::CudaPlacementBackend
// (src/placement/CudaPlacementBackend.hpp:25-39) over the C ABI of include/mp_gp_backend.h, and the
// factory tryCreateHipPlacementBackend with the contract of tryCreateCudaPlacementBackend (:44-48,
// cuda/CudaPlacementBackend.cu:1156-1216): nullptr plus a reason for any recoverable failure.
// It compiles against the reference's own headers unchanged; GlobalPlacementSession selects it for
// ComputeBackend::Hip exactly where it selects the CUDA factory (AdaptiveGlobalPlacer.cpp:271-278).
//
// The flattening follows buildHostLayout (cuda/CudaPlacementBackend.cu:784-839): particles are the
// movable modules in movableModules() order then the fillers; charge areas from densityChargeArea
// (GlobalPlacementSupport.cpp:105-111); pins in net order, a movable module's pin as (particle,
// offset turned by the module orientation), a fixed module's pin as (-1, absolute pin position);
// the static layers from the DensityMap bins (.cu:907-917).

#include "CudaPlacementBackend.hpp"
#include "mp_gp_backend.h"

#include <cstdint>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace myplacement::detail {

static_assert(sizeof(Vec2) == 2 * sizeof(double), "Vec2 crosses the C boundary as two doubles");

namespace {

// The backend's flattened input, owning the arrays mp_gp_static points into.
struct HipHostLayout {
    std::vector<double> particle_width, particle_height, particle_charge_area;
    std::vector<int32_t> particle_layer;
    std::vector<int32_t> net_offsets;
    std::vector<double> net_weights;
    std::vector<int32_t> pin_particles;
    std::vector<double> pin_offsets_or_static;  // (x, y) per pin
    std::vector<double> fixed_area, dark_area;  // per bin, row-major (row = y)
    int bins_x = 0, bins_y = 0;
    Rect region;
    double target_density = 0.0;

    [[nodiscard]] mp_gp_static view() const {
        mp_gp_static L{};
        L.particle_count = static_cast<int64_t>(particle_width.size());
        L.particle_width = particle_width.data();
        L.particle_height = particle_height.data();
        L.particle_charge_area = particle_charge_area.data();
        L.particle_layer = particle_layer.data();
        L.net_count = static_cast<int64_t>(net_weights.size());
        L.net_offsets = net_offsets.data();
        L.net_weights = net_weights.data();
        L.pin_count = static_cast<int64_t>(pin_particles.size());
        L.pin_particles = pin_particles.data();
        L.pin_offsets_or_static = pin_offsets_or_static.data();
        L.bins_x = bins_x;
        L.bins_y = bins_y;
        L.region_ll_x = region.ll.x;
        L.region_ll_y = region.ll.y;
        L.region_ur_x = region.ur.x;
        L.region_ur_y = region.ur.y;
        L.target_density = target_density;
        L.fixed_area = fixed_area.data();
        L.dark_area = dark_area.data();
        return L;
    }
};

void requireIndexable(std::size_t value, const char* what) {
    if (value > static_cast<std::size_t>(std::numeric_limits<int32_t>::max()))
        throw std::invalid_argument(std::string("HIP backend cannot index ") + what + " larger than INT_MAX.");
}

HipHostLayout flattenForHip(const PlacementDatabase& database, const std::vector<ModuleId>& movable,
                            const std::vector<DensityFiller>& fillers, const DensityMap& static_density,
                            double target_density) {
    requireIndexable(database.modules.size(), "modules");
    requireIndexable(database.pins.size(), "pins");
    requireIndexable(database.nets.size(), "nets");
    requireIndexable(movable.size() + fillers.size(), "particles");
    HipHostLayout a;
    std::vector<int32_t> particle_of(database.modules.size(), -1);
    for (std::size_t k = 0; k < movable.size(); ++k) {
        const Module& m = database.modules[movable[k]];
        particle_of[movable[k]] = static_cast<int32_t>(k);
        a.particle_width.push_back(m.width);
        a.particle_height.push_back(m.height);
        a.particle_charge_area.push_back(densityChargeArea(m, target_density));
        a.particle_layer.push_back(m.is_macro ? MP_LAYER_MACRO : MP_LAYER_STANDARD);
    }
    for (const DensityFiller& f : fillers) {
        a.particle_width.push_back(f.width);
        a.particle_height.push_back(f.height);
        a.particle_charge_area.push_back(densityChargeArea(f, target_density));
        a.particle_layer.push_back(MP_LAYER_FILLER);
    }
    a.net_offsets.reserve(database.nets.size() + 1);
    a.net_offsets.push_back(0);
    a.net_weights.reserve(database.nets.size());
    a.pin_particles.reserve(database.pins.size());
    a.pin_offsets_or_static.reserve(2 * database.pins.size());
    for (const Net& net : database.nets) {
        a.net_weights.push_back(net.weight);
        for (const PinId id : net.pins) {
            const Pin& pin = database.pins[id];
            const int32_t particle = particle_of[pin.module];
            const Vec2 v = particle >= 0 ? transformOffset(pin.offset, database.modules[pin.module].orientation)
                                         : database.pinPosition(id);
            a.pin_particles.push_back(particle);
            a.pin_offsets_or_static.push_back(v.x);
            a.pin_offsets_or_static.push_back(v.y);
        }
        requireIndexable(a.pin_particles.size(), "flattened net pins");
        a.net_offsets.push_back(static_cast<int32_t>(a.pin_particles.size()));
    }
    const std::vector<DensityBin>& bins = static_density.bins();
    a.fixed_area.reserve(bins.size());
    a.dark_area.reserve(bins.size());
    for (const DensityBin& b : bins) {
        a.fixed_area.push_back(b.fixed_area);
        a.dark_area.push_back(b.dark_area);
    }
    a.bins_x = static_density.columns();
    a.bins_y = static_density.rows();
    a.region = static_density.region();
    a.target_density = target_density;
    return a;
}

DensityMetrics toMetrics(const mp_gp_density_metrics& m) {
    DensityMetrics r;
    r.total_overflow_area = m.total_overflow_area;
    r.normalized_overflow = m.normalized_overflow;
    r.maximum_density = m.maximum_density;
    r.average_density = m.average_density;
    r.total_charge_area = m.total_charge_area;
    r.normalization_area = m.normalization_area;
    r.placeable_area = m.placeable_area;
    r.dark_area = m.dark_area;
    return r;
}

[[noreturn]] void raise(mp_gp_status status) {
    if (status == MP_GP_INVALID_ARGUMENT) throw std::invalid_argument(mp_gp_last_error());
    throw std::runtime_error(mp_gp_last_error());
}

class HipPlacementBackend final : public CudaPlacementBackend {
public:
    HipPlacementBackend(mp_gp_backend* handle, std::size_t particles) : handle_(handle), particles_(particles) {}
    ~HipPlacementBackend() override { mp_gp_destroy(handle_); }

    CudaPlacementEvaluation evaluate(const std::vector<Vec2>& particle_positions, double smoothing,
                                     bool calculate_gradient) override {
        if (particle_positions.size() != particles_)
            throw std::invalid_argument("HIP placement positions do not match the particle count.");
        CudaPlacementEvaluation result;
        if (calculate_gradient) {
            result.wire_gradient.resize(particles_);
            result.density_gradient.resize(particles_);
        }
        mp_gp_scalars s{};
        const mp_gp_status status =
            mp_gp_evaluate(handle_, reinterpret_cast<const double*>(particle_positions.data()), smoothing,
                           calculate_gradient ? 1 : 0, &s,
                           calculate_gradient ? reinterpret_cast<double*>(result.wire_gradient.data()) : nullptr,
                           calculate_gradient ? reinterpret_cast<double*>(result.density_gradient.data()) : nullptr);
        if (status != MP_GP_OK) raise(status);
        result.hpwl = s.hpwl;
        result.smooth_wirelength = s.smooth_wirelength;
        result.electrostatic_energy = s.electrostatic_energy;
        result.optimizer_density = toMetrics(s.optimizer_density);
        result.design_density = toMetrics(s.design_density);
        return result;
    }
    [[nodiscard]] int device() const override { return mp_gp_device(handle_); }
    [[nodiscard]] std::size_t reservedBytes() const override {
        return static_cast<std::size_t>(mp_gp_reserved_bytes(handle_));
    }

private:
    mp_gp_backend* handle_;
    std::size_t particles_;
};

}  // namespace

// cudaPlacementBackendCompiled() for the HIP backend.
[[nodiscard]] bool hipPlacementBackendCompiled() { return mp_gp_compiled() != 0; }

// tryCreateCudaPlacementBackend for the HIP backend.  `device` is the HIP ordinal (a field the
// maintainer adds to GlobalPlacementOptions beside cuda_device; the shared-server CUDA allow-list
// of CudaDevicePolicy does not apply).  Invalid options throw, as in the CUDA factory's argument
// checks; every other failure returns nullptr with `reason` set.
std::unique_ptr<CudaPlacementBackend> tryCreateHipPlacementBackend(
    const PlacementDatabase& database, const std::vector<ModuleId>& movable, const std::vector<DensityFiller>& fillers,
    const DensityMap& static_density, const GlobalPlacementOptions& options, std::string& reason, int device = 0) {
    reason.clear();
    if (static_density.targetDensity() != options.target_density)
        throw std::invalid_argument("HIP density map target does not match global-placement options.");
    try {
        const HipHostLayout layout = flattenForHip(database, movable, fillers, static_density, options.target_density);
        mp_gp_static L = layout.view();
        L.device = device;
        L.dct_precision = MP_DCT_FP32;
        L.maximum_device_bytes = static_cast<uint64_t>(options.maximum_cuda_memory_bytes);
        mp_gp_backend* handle = nullptr;
        const mp_gp_status status = mp_gp_create(&L, &handle);
        if (status != MP_GP_OK) {
            reason = mp_gp_last_error();
            if (reason.empty()) reason = "HIP backend creation failed.";
            return nullptr;
        }
        return std::make_unique<HipPlacementBackend>(handle, movable.size() + fillers.size());
    } catch (const std::exception& e) {
        reason = e.what();
        return nullptr;
    }
}

}  // namespace myplacement::detail
