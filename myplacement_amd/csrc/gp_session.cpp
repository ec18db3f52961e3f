// gp_session.cpp — the adaptive Nesterov / backtracking controller of runAdaptiveGlobalPlacement
// (src/placement/AdaptiveGlobalPlacer.cpp:413-953) driving the HIP backend with every O(P) vector
// resident on the device.  The scalar control law runs on the host in FP64 exactly as the reference
// writes it; each decision reads one small scalar block back from the device.
//
// Device vectors (double2[P]): X current and XP previous positions (the velocity is X - XP),
// B look-ahead base, Y proposal, PX / PG previous step base and its preconditioned gradient (PXn /
// PGn staged by the propose kernel), Rc/Gc and Rl/Gl raw and preconditioned gradients of the
// current and look-ahead evaluations, two checkpoint copies.  Accepting a step swaps pointers.
//
// One difference in sequencing, not in result: a proposal and its value-only evaluation are
// enqueued together and read back with one synchronisation; when the predicted reduction is
// below 1e-9 (AdaptiveGlobalPlacer.cpp:823-826) that evaluation is discarded unread.

#include "gp_session.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

namespace mpgp {

void validateOptions(const mp_gp_options& o) {  // AdaptiveGlobalPlacer.cpp:413-454 (non-RUDY part)
    if (o.iterations <= 0 || o.bins_x < 2 || o.bins_y < 2 || o.target_density <= 0.0 || o.target_density > 1.0 ||
        o.maximum_bin_overflow < 0.0 || o.initial_smoothing <= 0.0 || o.final_smoothing <= 0.0 ||
        o.maximum_movement_in_bins <= 0.0 || o.initial_movement_in_bins <= 0.0 || o.maximum_backtracks < 0 ||
        o.backtracking_ratio <= 0.0 || o.backtracking_ratio >= 1.0 || o.armijo_coefficient <= 0.0 ||
        o.armijo_coefficient >= 1.0 || o.maximum_momentum < 0.0 || o.maximum_momentum >= 1.0 ||
        o.density_penalty_increase <= 1.0 || o.density_penalty_decrease <= 0.0 || o.density_penalty_decrease >= 1.0 ||
        o.minimum_density_penalty <= 0.0 || o.maximum_density_penalty < o.minimum_density_penalty ||
        o.smoothing_overflow_exponent <= 0.0 || o.feasible_refinement_iterations <= 0 ||
        !std::isfinite(o.target_density) || !std::isfinite(o.maximum_bin_overflow) ||
        !std::isfinite(o.initial_smoothing) || !std::isfinite(o.final_smoothing) ||
        !std::isfinite(o.maximum_density_penalty) || !(o.density_order_reuse_rows >= 0.0) ||
        !std::isfinite(o.density_order_reuse_rows) || (o.capture_graphs != 0 && o.capture_graphs != 1) ||
        (o.recompute_density != 0 && o.recompute_density != 1))
        throw std::invalid_argument("Adaptive global placement options are outside their valid range.");
    if (static_cast<std::size_t>(o.bins_x) > (std::size_t{1} << 20) / static_cast<std::size_t>(o.bins_y))
        throw std::invalid_argument("Density grid exceeds the safe one-million-bin workspace limit.");
}


std::future<std::pair<double, double>> beforeMetrics(std::shared_ptr<const Prepared> prep, const mp_gp_options& o) {
    auto compute = [prep, nx = o.bins_x, ny = o.bins_y, t = o.target_density] {
        const Netlist& n = prep->net;
        return std::make_pair(n.hpwl(n.cx.data(), n.cy.data()), n.designOverflow(n.cx.data(), n.cy.data(), nx, ny, t));
    };
    constexpr std::size_t kLargePins = 200000;
    return std::async(prep->net.pin_module.size() > kLargePins ? std::launch::async : std::launch::deferred, compute);
}

}  // namespace mpgp

using namespace mpgp;

struct mp_gp_prepared {
    std::shared_ptr<const Prepared> p;
};

struct mp_gp_session {
    std::unique_ptr<Session> s;
    mp_gp_backend view;  // non-owning view of the session's backend
};

namespace {
template <typename F>
mp_gp_status guarded(F&& body) {
    g_last_error.clear();
    try {
        body();
        return MP_GP_OK;
    } catch (...) {
        return translate(std::current_exception());
    }
}
}  // namespace

extern "C" {

void mp_gp_default_options(mp_gp_options* o) {
    if (!o) return;
    *o = {};
    o->iterations = 280;
    o->bins_x = 64;
    o->bins_y = 64;
    o->target_density = 0.85;
    o->maximum_bin_overflow = 0.10;
    o->initial_smoothing = 0.75;
    o->final_smoothing = 0.12;
    o->maximum_movement_in_bins = 0.22;
    o->maximum_backtracks = 6;
    o->backtracking_ratio = 0.5;
    o->armijo_coefficient = 1e-4;
    o->maximum_momentum = 0.92;
    o->initial_movement_in_bins = 0.12;
    o->objective_increase_for_density = 0.002;
    o->hpwl_growth_restart_threshold = 0.02;
    o->density_stall_threshold = 0.002;
    o->density_penalty_increase = 1.05;
    o->density_penalty_decrease = 0.88;
    o->minimum_density_penalty = 1e-4;
    o->maximum_density_penalty = 1e6;
    o->smoothing_overflow_exponent = 1.0;
    o->feasible_refinement_iterations = 12;
    o->maximum_fillers = 50000;
    o->seed = 2026;
    o->device = 0;
    o->dct_precision = MP_DCT_FP32;
    o->maximum_device_bytes = 40ULL << 30;
    o->density_order_reuse_rows = 2.5;
}

mp_gp_status mp_gp_prepare(const mp_design* design, const mp_gp_options* options, mp_gp_prepared** out) {
    return guarded([&] {
        if (!design || !options || !out) throw std::invalid_argument("null design, options or output");
        *out = nullptr;
        auto h = std::make_unique<mp_gp_prepared>();
        h->p = std::make_shared<const Prepared>(*design, *options);
        *out = h.release();
    });
}

void mp_gp_prepared_destroy(mp_gp_prepared* p) { delete p; }

mp_gp_status mp_gp_prepared_query(const mp_gp_prepared* h, mp_gp_prepared_info* info) {
    return guarded([&] {
        if (!h || !info) throw std::invalid_argument("null prepared design or output");
        const Prepared& p = *h->p;
        *info = {};
        info->movable_count = static_cast<int64_t>(p.net.movable.size());
        info->filler_count = static_cast<int64_t>(p.filler_x.size());
        info->particle_count = static_cast<int64_t>(p.pw.size());
        info->net_count = static_cast<int64_t>(p.net_offsets.size()) - 1;
        info->pin_count = static_cast<int64_t>(p.pin_particles.size());
        info->bin_count = static_cast<int64_t>(p.nx) * p.ny;
        info->core_ll_x = p.net.core[0];
        info->core_ll_y = p.net.core[1];
        info->core_ur_x = p.net.core[2];
        info->core_ur_y = p.net.core[3];
        info->nominal_row_height = p.net.nominal_row_height;
        info->filler_width = p.filler_w;
        info->filler_height = p.filler_h;
    });
}

mp_gp_status mp_gp_prepared_layout(const mp_gp_prepared* h, mp_gp_static* layout) {
    return guarded([&] {
        if (!h || !layout) throw std::invalid_argument("null prepared design or output");
        const int device = layout->device;
        const int precision = layout->dct_precision;
        const std::uint64_t bytes = layout->maximum_device_bytes ? layout->maximum_device_bytes : (40ULL << 30);
        *layout = h->p->layout(device, precision, bytes);
    });
}

mp_gp_status mp_gp_prepared_arrays(const mp_gp_prepared* h, int64_t* movable_ids, double* filler_centers,
                                   double* preconditioner_base) {
    return guarded([&] {
        if (!h) throw std::invalid_argument("null prepared design");
        const Prepared& p = *h->p;
        if (movable_ids) std::copy(p.net.movable.begin(), p.net.movable.end(), movable_ids);
        if (filler_centers)
            for (std::size_t i = 0; i < p.filler_x.size(); ++i) {
                filler_centers[2 * i] = p.filler_x[i];
                filler_centers[2 * i + 1] = p.filler_y[i];
            }
        if (preconditioner_base) std::copy(p.precond.begin(), p.precond.end(), preconditioner_base);
    });
}

mp_gp_status mp_gp_random_placement(const mp_design* design, uint64_t seed, double* centers) {
    return guarded([&] {
        if (!design || !centers) throw std::invalid_argument("null design or output");
        randomPlacement(*design, seed, centers);
    });
}

mp_gp_status mp_gp_session_create(const mp_gp_prepared* prepared, const mp_gp_options* options, const double* centers,
                                  mp_gp_session** out) {
    return guarded([&] {
        if (!prepared || !options || !out) throw std::invalid_argument("null prepared design, options or output");
        *out = nullptr;
        if (options->maximum_device_bytes == 0)
            throw std::invalid_argument("HIP global placement requires a non-zero memory budget.");
        auto h = std::make_unique<mp_gp_session>();
        h->s = std::make_unique<Session>(prepared->p, *options, centers);
        h->view.impl = &h->s->backend();
        *out = h.release();
    });
}

mp_gp_status mp_gp_session_step(mp_gp_session* h, int32_t iterations, int32_t* completed, int32_t* stopped) {
    return guarded([&] {
        if (!h) throw std::invalid_argument("null session");
        if (iterations < 0) throw std::invalid_argument("negative iteration count");
        const int done = h->s->step(iterations);
        if (completed) *completed = done;
        if (stopped) *stopped = h->s->stopped() ? 1 : 0;
    });
}

mp_gp_status mp_gp_session_finish(mp_gp_session* h, double* centers, mp_gp_result* result) {
    return guarded([&] {
        if (!h) throw std::invalid_argument("null session");
        h->s->finish(centers, result);
    });
}

int64_t mp_gp_session_history(const mp_gp_session* h, mp_gp_history_row* rows, int64_t capacity) {
    if (!h) return 0;
    const auto& hist = h->s->history();
    if (rows == nullptr) return static_cast<int64_t>(hist.size());
    const int64_t n = std::min<int64_t>(capacity, static_cast<int64_t>(hist.size()));
    std::copy(hist.begin(), hist.begin() + n, rows);
    return n;
}

mp_gp_status mp_gp_session_positions(mp_gp_session* h, double* xy) {
    return guarded([&] {
        if (!h || !xy) throw std::invalid_argument("null session or output");
        h->s->positions(xy);
    });
}

mp_gp_status mp_gp_session_enable_kernel_timing(mp_gp_session* h, int enable) {
    return guarded([&] {
        if (!h) throw std::invalid_argument("null session");
        h->s->backend().enableTiming(static_cast<unsigned>(enable));
    });
}

mp_gp_status mp_gp_session_enable_kernel_timing_every(mp_gp_session* h, int enable, int every) {
    return guarded([&] {
        if (!h) throw std::invalid_argument("null session");
        h->s->backend().enableTiming(static_cast<unsigned>(enable), every);
    });
}

int32_t mp_gp_session_kernel_timing(mp_gp_session* h, char* names, size_t names_len, double* total_us,
                                    int64_t* launches, double* work, int32_t capacity) {
    if (!h) return 0;
    try {
        std::vector<double> t, wk;
        std::vector<long long> c;
        h->s->backend().timing(t, c, wk);
        const auto nm = h->s->backend().timingNames();
        std::string joined;
        for (std::size_t i = 0; i < nm.size(); ++i) joined += (i ? "," : "") + nm[i];
        if (names && names_len) {
            const std::size_t n = std::min(names_len - 1, joined.size());
            std::copy(joined.begin(), joined.begin() + n, names);
            names[n] = '\0';
        }
        const int32_t n = std::min<int32_t>(capacity, static_cast<int32_t>(t.size()));
        for (int32_t i = 0; i < n; ++i) {
            if (total_us) total_us[i] = t[i];
            if (launches) launches[i] = c[i];
            if (work) work[i] = wk[i];
        }
        return static_cast<int32_t>(t.size());
    } catch (...) {
        translate(std::current_exception());
        return -1;
    }
}

mp_gp_status mp_gp_session_particle_order(const mp_gp_session* h, int64_t* order) {
    return guarded([&] {
        if (!h || !order) throw std::invalid_argument("null session or output");
        const auto& o = h->s->order();
        std::copy(o.begin(), o.end(), order);
    });
}

mp_gp_backend* mp_gp_session_backend(mp_gp_session* h) { return h ? &h->view : nullptr; }

void mp_gp_session_destroy(mp_gp_session* h) { delete h; }

mp_gp_status mp_gp_global_place(const mp_design* design, const mp_gp_options* options, double* centers,
                                mp_gp_result* result, mp_gp_history_row* history, int64_t history_capacity) {
    return guarded([&] {
        if (!design || !options || !centers) throw std::invalid_argument("null design, options or centres");
        mp_design d = *design;
        d.center = centers;
        const auto t0 = std::chrono::steady_clock::now();
        Netlist probe(d);
        if (probe.movable.empty()) {
            if (result) *result = {};
            return;
        }
        validateOptions(*options);
        auto prep = std::make_shared<const Prepared>(std::move(probe), *options);
        auto before = beforeMetrics(prep, *options);
        Session s(prep, *options, nullptr);
        s.step(options->iterations);
        mp_gp_result r{};
        s.finish(centers, &r);
        std::tie(r.hpwl_before, r.overflow_before) = before.get();
        r.elapsed_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (result) *result = r;
        if (history) {
            const auto& hist = s.history();
            const int64_t n = std::min<int64_t>(history_capacity, static_cast<int64_t>(hist.size()));
            std::copy(hist.begin(), hist.begin() + n, history);
        }
    });
}

}  // extern "C"
