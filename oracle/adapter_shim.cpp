// adapter_shim.cpp — C entry points that drive the MyPlacement-side adapter
// (myplacement_amd/integration/HipPlacementBackend.cpp) the way GlobalPlacementSession does.
//
// TEST INFRASTRUCTURE.  Built by oracle/Makefile into oracle/_ref/libmpadapter.so together with the
// reference's own unmodified PlacementDatabase.cpp, Metrics.cpp, GlobalPlacementSupport.cpp and
// BookshelfParser.cpp (read in place from the reference tree), and linked against the product
// library libmp_gp_hip.so.  A design is read with the reference's BookshelfParser, then set up as
// GlobalPlacementSession's constructor does (AdaptiveGlobalPlacer.cpp:114-126): DensityMap over
// the core region, initializeStaticDensity, createDensityFillers, movableModules(); the backend
// comes from tryCreateHipPlacementBackend and is used only through the CudaPlacementBackend
// interface.

#include "../myplacement_amd/integration/HipPlacementBackend.cpp"

#include "myplacement/io/BookshelfParser.hpp"

#include <cstring>

using namespace myplacement;
using namespace myplacement::detail;

struct adp_run {
    PlacementDatabase db;
    GlobalPlacementOptions options;
    std::unique_ptr<DensityMap> density;
    std::vector<DensityFiller> fillers;
    std::vector<ModuleId> movable;
    HipHostLayout layout;
    std::unique_ptr<CudaPlacementBackend> backend;
};

namespace {
thread_local std::string g_adp_error;

void copyOut(const DensityMetrics& m, mp_gp_density_metrics* o) {
    o->total_overflow_area = m.total_overflow_area;
    o->normalized_overflow = m.normalized_overflow;
    o->maximum_density = m.maximum_density;
    o->average_density = m.average_density;
    o->total_charge_area = m.total_charge_area;
    o->normalization_area = m.normalization_area;
    o->placeable_area = m.placeable_area;
    o->dark_area = m.dark_area;
}

template <typename F>
int adpGuard(F&& body) {
    try {
        g_adp_error.clear();
        body();
        return 0;
    } catch (const std::invalid_argument& e) {
        g_adp_error = e.what();
        return 1;
    } catch (const std::exception& e) {
        g_adp_error = e.what();
        return 2;
    }
}
}  // namespace

extern "C" {

const char* adp_last_error(void) { return g_adp_error.c_str(); }

// Reads `aux`, sets the session up and asks the factory for a backend.  Returns 0 with *out set
// when the backend was created; 3 with *out set (no backend) and the factory's reason in `reason`
// when it returned nullptr; 1 / 2 when set-up threw invalid_argument / another error.
int adp_open(const char* aux, int bins_x, int bins_y, double target_density, uint64_t maximum_fillers, uint64_t seed,
             int device, uint64_t maximum_device_bytes, adp_run** out, char* reason, size_t reason_len) {
    *out = nullptr;
    const int status = adpGuard([&] {
        auto r = std::make_unique<adp_run>();
        r->db = BookshelfParser().parseAux(aux);
        r->options.bins_x = bins_x;
        r->options.bins_y = bins_y;
        r->options.target_density = target_density;
        r->options.maximum_fillers = maximum_fillers;
        r->options.seed = seed;
        r->options.maximum_cuda_memory_bytes = maximum_device_bytes;
        r->density = std::make_unique<DensityMap>(r->db.core_region, bins_x, bins_y, target_density);
        initializeStaticDensity(*r->density, r->db);
        r->fillers = createDensityFillers(r->db, *r->density, r->options);
        r->movable = r->db.movableModules();
        r->layout = flattenForHip(r->db, r->movable, r->fillers, *r->density, target_density);
        std::string why;
        r->backend = tryCreateHipPlacementBackend(r->db, r->movable, r->fillers, *r->density, r->options, why, device);
        if (reason && reason_len) {
            std::strncpy(reason, why.c_str(), reason_len - 1);
            reason[reason_len - 1] = '\0';
        }
        *out = r.release();
    });
    if (status != 0) return status;
    return (*out)->backend ? 0 : 3;
}

// The flattened layout the adapter hands to mp_gp_create (pointers valid while the run lives).
int adp_layout(const adp_run* r, mp_gp_static* layout) {
    *layout = r->layout.view();
    return 0;
}

// Particle positions as capturePositions gives them: movable centres, then filler centres [2P].
int adp_positions(const adp_run* r, double* xy) {
    std::size_t k = 0;
    for (const ModuleId m : r->movable) {
        xy[k++] = r->db.modules[m].center.x;
        xy[k++] = r->db.modules[m].center.y;
    }
    for (const DensityFiller& f : r->fillers) {
        xy[k++] = f.center.x;
        xy[k++] = f.center.y;
    }
    return 0;
}

int64_t adp_particle_count(const adp_run* r) { return static_cast<int64_t>(r->movable.size() + r->fillers.size()); }

// CudaPlacementBackend::evaluate through the adapter on `count` positions xy [2 count]; gradients
// [2 count] when want_grad (the adapter rejects a count other than the particle count before it
// writes anything).
int adp_evaluate(adp_run* r, const double* xy, int64_t count, double smoothing, int want_grad, mp_gp_scalars* out,
                 double* wire_grad, double* density_grad) {
    return adpGuard([&] {
        if (!r->backend) throw std::runtime_error("no backend");
        if (count < 0) throw std::invalid_argument("negative position count");
        std::vector<Vec2> positions(static_cast<std::size_t>(count));
        for (std::size_t i = 0; i < positions.size(); ++i) positions[i] = {xy[2 * i], xy[2 * i + 1]};
        const CudaPlacementEvaluation e = r->backend->evaluate(positions, smoothing, want_grad != 0);
        const std::size_t P = positions.size();
        out->hpwl = e.hpwl;
        out->smooth_wirelength = e.smooth_wirelength;
        out->electrostatic_energy = e.electrostatic_energy;
        copyOut(e.optimizer_density, &out->optimizer_density);
        copyOut(e.design_density, &out->design_density);
        if (want_grad) {
            if (e.wire_gradient.size() != P || e.density_gradient.size() != P)
                throw std::runtime_error("gradient size mismatch");
            for (std::size_t i = 0; i < P; ++i) {
                wire_grad[2 * i] = e.wire_gradient[i].x;
                wire_grad[2 * i + 1] = e.wire_gradient[i].y;
                density_grad[2 * i] = e.density_gradient[i].x;
                density_grad[2 * i + 1] = e.density_gradient[i].y;
            }
        } else if (!e.wire_gradient.empty() || !e.density_gradient.empty()) {
            throw std::runtime_error("value-only evaluation returned gradients");
        }
    });
}

int adp_device(const adp_run* r) { return r->backend ? r->backend->device() : -1; }
uint64_t adp_reserved_bytes(const adp_run* r) { return r->backend ? r->backend->reservedBytes() : 0; }
int adp_compiled(void) { return hipPlacementBackendCompiled() ? 1 : 0; }
void adp_close(adp_run* r) { delete r; }

}  // extern "C"
