// ref_shim.cpp — C entry points over the reference's own density, filler and BookShelf code.
//
// TEST INFRASTRUCTURE.  Compiled together with these unmodified reference sources, read in place
// from the reference tree by oracle/Makefile (output: oracle/_ref/libmpref.so):
//   src/model/PlacementDatabase.cpp, src/metrics/Metrics.cpp,
//   src/placement/GlobalPlacementSupport.cpp, src/io/BookshelfParser.cpp,
//   src/placement/Legalizer.cpp, src/placement/DetailedPlacer.cpp and the reference's own
//   CPU-only stub src/placement/CudaDetailedPlacementBackendStub.cpp
// The rest of the reference GP path (ElectrostaticField.cpp needs FFTW3, InitialPlacer.cpp needs
// Eigen, AdaptiveGlobalPlacer.cpp links both) is not buildable in this image.

#include "mp_gp_backend.h"

#include "myplacement/io/BookshelfParser.hpp"
#include "myplacement/metrics/Metrics.hpp"
#include "myplacement/model/PlacementDatabase.hpp"
#include "myplacement/placement/DetailedPlacer.hpp"
#include "myplacement/placement/Legalizer.hpp"
#include "GlobalPlacementSupport.hpp"

#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

using namespace myplacement;

namespace {

thread_local std::string g_error;

const Orientation kOrientations[] = {Orientation::North,     Orientation::South,     Orientation::East,
                                     Orientation::West,      Orientation::FlipNorth, Orientation::FlipSouth,
                                     Orientation::FlipEast,  Orientation::FlipWest,  Orientation::Unknown};

int orientationIndex(Orientation o) {
    for (int i = 0; i < 9; ++i)
        if (kOrientations[i] == o) return i;
    return 8;
}

PlacementDatabase buildDatabase(const mp_design* d) {
    PlacementDatabase db;
    for (int64_t m = 0; m < d->module_count; ++m) {
        Module module;
        module.name = "m" + std::to_string(m);
        module.width = d->width[m];
        module.height = d->height[m];
        module.center = {d->center[2 * m], d->center[2 * m + 1]};
        module.orientation = kOrientations[d->orientation[m] < 0 || d->orientation[m] > 8 ? 8 : d->orientation[m]];
        module.is_terminal = d->is_terminal[m] != 0;
        module.is_fixed = d->is_fixed[m] != 0 || module.is_terminal;
        db.addModule(std::move(module));
    }
    for (int64_t n = 0; n < d->net_count; ++n) {
        Net net;
        net.name = "n" + std::to_string(n);
        net.weight = d->net_weight[n];
        const NetId id = db.addNet(std::move(net));
        for (int64_t p = d->net_pin_start[n]; p < d->net_pin_start[n + 1]; ++p) {
            Pin pin;
            pin.module = static_cast<ModuleId>(d->pin_module[p]);
            pin.net = id;
            pin.offset = {d->pin_offset[2 * p], d->pin_offset[2 * p + 1]};
            db.addPin(pin);
        }
    }
    for (int64_t r = 0; r < d->row_count; ++r) {
        const double* row = d->rows + 5 * r;
        SiteRow s;
        s.bottom = row[0];
        s.height = row[1];
        s.x_start = row[2];
        s.site_spacing = row[3];
        s.site_count = static_cast<std::size_t>(row[4]);
        db.rows.push_back(s);
    }
    db.refreshDerivedData();
    return db;
}

template <typename F>
int guarded(F&& body) {
    try {
        g_error.clear();
        body();
        return 0;
    } catch (const std::invalid_argument& e) {
        g_error = e.what();
        return 1;
    } catch (const std::exception& e) {
        g_error = e.what();
        return 2;
    }
}

void copyMetrics(const DensityMetrics& m, mp_gp_density_metrics* out) {
    out->total_overflow_area = m.total_overflow_area;
    out->normalized_overflow = m.normalized_overflow;
    out->maximum_density = m.maximum_density;
    out->average_density = m.average_density;
    out->total_charge_area = m.total_charge_area;
    out->normalization_area = m.normalization_area;
    out->placeable_area = m.placeable_area;
    out->dark_area = m.dark_area;
}

}  // namespace

struct ref_db {
    PlacementDatabase db;
};

extern "C" {

const char* ref_last_error(void) { return g_error.c_str(); }

// DensityMap driver with the same arguments as orc_density_map.
int ref_density_map(const double* region, int cols, int rows, double t, int64_t n_regions, const double* regions,
                    int64_t n_rects, const double* rects, const int32_t* layers, mp_gp_density_metrics* with_f,
                    mp_gp_density_metrics* without_f, double* bins) {
    return guarded([&] {
        DensityMap map({{region[0], region[1]}, {region[2], region[3]}}, cols, rows, t);
        std::vector<Rect> regs;
        for (int64_t i = 0; i < n_regions; ++i)
            regs.push_back({{regions[4 * i], regions[4 * i + 1]}, {regions[4 * i + 2], regions[4 * i + 3]}});
        if (n_regions > 0) map.setPlaceableRegions(regs);
        const DensityLayer kinds[] = {DensityLayer::Fixed, DensityLayer::Movable, DensityLayer::Macro,
                                      DensityLayer::Filler, DensityLayer::Dark};
        for (int64_t i = 0; i < n_rects; ++i)
            map.addRectangle({{rects[4 * i], rects[4 * i + 1]}, {rects[4 * i + 2], rects[4 * i + 3]}}, kinds[layers[i]]);
        if (with_f) copyMetrics(map.metrics(true), with_f);
        if (without_f) copyMetrics(map.metrics(false), without_f);
        if (bins)
            for (std::size_t i = 0; i < map.bins().size(); ++i) {
                const DensityBin& b = map.bins()[i];
                bins[5 * i] = b.fixed_area;
                bins[5 * i + 1] = b.movable_area;
                bins[5 * i + 2] = b.macro_area;
                bins[5 * i + 3] = b.filler_area;
                bins[5 * i + 4] = b.dark_area;
            }
    });
}

// initializeStaticDensity + createDensityFillers.  fillers_out: [4*capacity] cx, cy, w, h.
int ref_fillers(const mp_design* design, int cols, int rows, double t, uint64_t maximum_fillers, uint64_t seed,
                double* fixed, double* dark, int64_t* count, double* fillers_out, int64_t capacity) {
    return guarded([&] {
        PlacementDatabase db = buildDatabase(design);
        DensityMap map(db.core_region, cols, rows, t);
        detail::initializeStaticDensity(map, db);
        GlobalPlacementOptions options;
        options.target_density = t;
        options.bins_x = cols;
        options.bins_y = rows;
        options.maximum_fillers = static_cast<std::size_t>(maximum_fillers);
        options.seed = seed;
        const std::vector<detail::DensityFiller> fillers = detail::createDensityFillers(db, map, options);
        *count = static_cast<int64_t>(fillers.size());
        for (std::size_t i = 0; i < fillers.size() && static_cast<int64_t>(i) < capacity; ++i) {
            fillers_out[4 * i] = fillers[i].center.x;
            fillers_out[4 * i + 1] = fillers[i].center.y;
            fillers_out[4 * i + 2] = fillers[i].width;
            fillers_out[4 * i + 3] = fillers[i].height;
        }
        for (std::size_t i = 0; i < map.bins().size(); ++i) {
            if (fixed) fixed[i] = map.bins()[i].fixed_area;
            if (dark) dark[i] = map.bins()[i].dark_area;
        }
    });
}

// depositDynamicDensity + buildDensityDeviation + both metrics, for given movable centres [2M]
// (modules' centres in `design`) and fillers [4F] (cx, cy, w, h).
int ref_deposit(const mp_design* design, int cols, int rows, double t, int64_t n_fillers, const double* fillers_in,
                double* deviation, mp_gp_density_metrics* with_f, mp_gp_density_metrics* without_f) {
    return guarded([&] {
        PlacementDatabase db = buildDatabase(design);
        DensityMap map(db.core_region, cols, rows, t);
        detail::initializeStaticDensity(map, db);
        std::vector<detail::DensityFiller> fillers(static_cast<std::size_t>(n_fillers));
        for (int64_t i = 0; i < n_fillers; ++i)
            fillers[i] = {{fillers_in[4 * i], fillers_in[4 * i + 1]}, fillers_in[4 * i + 2], fillers_in[4 * i + 3]};
        detail::depositDynamicDensity(map, db, fillers);
        std::vector<double> dev;
        detail::buildDensityDeviation(map, dev);
        if (deviation) std::memcpy(deviation, dev.data(), dev.size() * sizeof(double));
        if (with_f) copyMetrics(map.metrics(true), with_f);
        if (without_f) copyMetrics(map.metrics(false), without_f);
    });
}

// calculateHpwl + calculateDensity (Metrics.cpp:128-164).
int ref_design_metrics(const mp_design* design, int cols, int rows, double t, double* hpwl,
                       mp_gp_density_metrics* density) {
    return guarded([&] {
        PlacementDatabase db = buildDatabase(design);
        if (hpwl) *hpwl = calculateHpwl(db).hpwl;
        if (density) copyMetrics(calculateDensity(db, cols, rows, t), density);
    });
}

// Movable ids, macro flags and core region after refreshDerivedData.
int ref_derived(const mp_design* design, int64_t* movable_ids, int64_t* movable_count, uint8_t* is_macro,
                double* core, double* nominal_row_height) {
    return guarded([&] {
        PlacementDatabase db = buildDatabase(design);
        *movable_count = static_cast<int64_t>(db.movableModules().size());
        if (movable_ids)
            for (std::size_t i = 0; i < db.movableModules().size(); ++i) movable_ids[i] = static_cast<int64_t>(db.movableModules()[i]);
        if (is_macro)
            for (std::size_t i = 0; i < db.modules.size(); ++i) is_macro[i] = db.modules[i].is_macro ? 1 : 0;
        core[0] = db.core_region.ll.x;
        core[1] = db.core_region.ll.y;
        core[2] = db.core_region.ur.x;
        core[3] = db.core_region.ur.y;
        *nominal_row_height = db.nominalRowHeight();
    });
}

// BookshelfParser::parseAux; the database is then exported through ref_db_* below.
int ref_parse_aux(const char* path, ref_db** out) {
    return guarded([&] {
        auto h = std::make_unique<ref_db>();
        h->db = BookshelfParser().parseAux(path);
        *out = h.release();
    });
}

void ref_db_destroy(ref_db* h) { delete h; }

// sizes: modules, nets, pins, rows, movable, fixed, macro
void ref_db_sizes(const ref_db* h, int64_t* sizes) {
    const DatabaseSummary s = h->db.summary();
    sizes[0] = static_cast<int64_t>(h->db.modules.size());
    sizes[1] = static_cast<int64_t>(h->db.nets.size());
    sizes[2] = static_cast<int64_t>(h->db.pins.size());
    sizes[3] = static_cast<int64_t>(h->db.rows.size());
    sizes[4] = static_cast<int64_t>(s.movable_count);
    sizes[5] = static_cast<int64_t>(s.fixed_count);
    sizes[6] = static_cast<int64_t>(s.macro_count);
}

// Exports the parsed database in the mp_design layout (pins in net order).
int ref_db_export(const ref_db* h, double* width, double* height, double* center, int32_t* orientation,
                  uint8_t* is_terminal, uint8_t* is_fixed, double* net_weight, int64_t* net_pin_start,
                  int32_t* pin_module, double* pin_offset, double* rows, double* core) {
    return guarded([&] {
        const PlacementDatabase& db = h->db;
        for (std::size_t m = 0; m < db.modules.size(); ++m) {
            width[m] = db.modules[m].width;
            height[m] = db.modules[m].height;
            center[2 * m] = db.modules[m].center.x;
            center[2 * m + 1] = db.modules[m].center.y;
            orientation[m] = orientationIndex(db.modules[m].orientation);
            is_terminal[m] = db.modules[m].is_terminal;
            is_fixed[m] = db.modules[m].is_fixed;
        }
        int64_t k = 0;
        net_pin_start[0] = 0;
        for (std::size_t n = 0; n < db.nets.size(); ++n) {
            net_weight[n] = db.nets[n].weight;
            for (const PinId p : db.nets[n].pins) {
                pin_module[k] = static_cast<int32_t>(db.pins[p].module);
                pin_offset[2 * k] = db.pins[p].offset.x;
                pin_offset[2 * k + 1] = db.pins[p].offset.y;
                ++k;
            }
            net_pin_start[n + 1] = k;
        }
        for (std::size_t r = 0; r < db.rows.size(); ++r) {
            rows[5 * r] = db.rows[r].bottom;
            rows[5 * r + 1] = db.rows[r].height;
            rows[5 * r + 2] = db.rows[r].x_start;
            rows[5 * r + 3] = db.rows[r].site_spacing;
            rows[5 * r + 4] = static_cast<double>(db.rows[r].site_count);
        }
        core[0] = db.core_region.ll.x;
        core[1] = db.core_region.ll.y;
        core[2] = db.core_region.ur.x;
        core[3] = db.core_region.ur.y;
    });
}

// Legalizer::legalize on the design (Legalizer.cpp:707-738).  report[12]: legal, placed,
// unplaced, overlap pairs, out-of-core, off-row, reverse pass selected, total / squared / weighted
// squared / maximum displacement, elapsed seconds.
int ref_legalize(const mp_design* design, int strategy, int candidate_rows, int macro_search_radius,
                 int bidirectional, double epsilon, double* centers, int32_t* orientation, double* report) {
    return guarded([&] {
        PlacementDatabase db = buildDatabase(design);
        LegalizationOptions o;
        o.strategy = strategy == 0 ? LegalizationStrategy::Abacus : LegalizationStrategy::GreedyIsotonic;
        o.candidate_rows = candidate_rows;
        o.macro_search_radius = macro_search_radius;
        o.abacus_bidirectional = bidirectional != 0;
        o.epsilon = epsilon;
        const LegalityReport r = Legalizer().legalize(db, o);
        const double v[12] = {r.legal ? 1.0 : 0.0, static_cast<double>(r.placed_standard_cells),
                              static_cast<double>(r.unplaced_standard_cells), static_cast<double>(r.overlap_pairs),
                              static_cast<double>(r.out_of_core_modules), static_cast<double>(r.off_row_modules),
                              r.abacus_reverse_pass_selected ? 1.0 : 0.0, r.standard_cell_total_displacement,
                              r.standard_cell_total_squared_displacement, r.standard_cell_weighted_squared_displacement,
                              r.standard_cell_maximum_displacement, r.elapsed_seconds};
        std::memcpy(report, v, sizeof v);
        for (std::size_t m = 0; m < db.modules.size(); ++m) {
            centers[2 * m] = db.modules[m].center.x;
            centers[2 * m + 1] = db.modules[m].center.y;
            if (orientation) orientation[m] = orientationIndex(db.modules[m].orientation);
        }
    });
}

// Legalizer::check (Legalizer.cpp:740-805); report[4]: legal, overlap pairs, out-of-core, off-row.
int ref_legality(const mp_design* design, double epsilon, double* report) {
    return guarded([&] {
        const PlacementDatabase db = buildDatabase(design);
        const LegalityReport r = Legalizer().check(db, epsilon);
        report[0] = r.legal ? 1.0 : 0.0;
        report[1] = static_cast<double>(r.overlap_pairs);
        report[2] = static_cast<double>(r.out_of_core_modules);
        report[3] = static_cast<double>(r.off_row_modules);
    });
}

// DetailedPlacer::run on the CPU (DetailedPlacer.cpp:289-335).  result[7]: hpwl before / after,
// evaluated windows, evaluated permutations, accepted operations, completed passes, elapsed.
int ref_detailed_place(const mp_design* design, int method, int passes, int window_size, uint64_t maximum_net_degree,
                       double improvement_epsilon, double* centers, int32_t* orientation, double* result) {
    return guarded([&] {
        PlacementDatabase db = buildDatabase(design);
        DetailedPlacementOptions o;
        o.method = method == 0 ? DetailedPlacementMethod::None
                               : (method == 1 ? DetailedPlacementMethod::AdjacentSwap : DetailedPlacementMethod::WindowReorder);
        o.passes = passes;
        o.window_size = window_size;
        o.maximum_net_degree = static_cast<std::size_t>(maximum_net_degree);
        o.improvement_epsilon = improvement_epsilon;
        o.compute_backend = DetailedPlacementBackend::Cpu;
        const DetailedPlacementResult r = DetailedPlacer().run(db, o);
        const double v[7] = {r.hpwl_before, r.hpwl_after, static_cast<double>(r.evaluated_windows),
                             static_cast<double>(r.evaluated_permutations), static_cast<double>(r.accepted_operations),
                             static_cast<double>(r.completed_passes), r.elapsed_seconds};
        std::memcpy(result, v, sizeof v);
        for (std::size_t m = 0; m < db.modules.size(); ++m) {
            centers[2 * m] = db.modules[m].center.x;
            centers[2 * m + 1] = db.modules[m].center.y;
            if (orientation) orientation[m] = orientationIndex(db.modules[m].orientation);
        }
    });
}

}  // extern "C"
