// gp_host.hpp — host-side netlist view and run set-up (see gp_host.cpp).
#pragma once

#include "mp_gp_backend.h"

#include <algorithm>
#include <array>
#include <cstdint>
#include <exception>
#include <limits>
#include <thread>
#include <vector>

namespace mpgp {

// The PlacementDatabase fields the GP path reads, after refreshDerivedData.
struct Netlist {
    explicit Netlist(const mp_design& d);
    std::size_t M = 0;
    std::vector<double> width, height, cx, cy;
    std::vector<int32_t> orientation;
    std::vector<uint8_t> terminal, fixed, macro;
    std::vector<double> net_weight;
    std::vector<int64_t> net_start;
    std::vector<int32_t> pin_module;
    std::vector<double> pin_dx, pin_dy;
    std::vector<int> module_pins;
    std::vector<std::array<double, 4>> rows;  // lx, ly, ux, uy
    std::array<double, 4> core{{0, 0, 0, 0}};
    double nominal_row_height = 0.0;
    std::vector<int64_t> movable;

    double hpwl(const double* cx, const double* cy) const;
    double designOverflow(const double* cx, const double* cy, int nx, int ny, double t) const;
};

// Static density, fillers, flattened layout and preconditioner of one run.
struct Prepared {
    Prepared(const mp_design& d, const mp_gp_options& o);
    // from a netlist already read from the design (taken over)
    Prepared(Netlist&& n, const mp_gp_options& o);
    Netlist net;
    double t = 0.0;
    int nx = 0, ny = 0;
    double bw = 0.0, bh = 0.0;
    std::vector<double> fixed_area, dark_area;
    std::vector<double> filler_x, filler_y;
    double filler_w = 0.0, filler_h = 0.0;
    std::vector<double> pw, ph, pq;
    std::vector<int32_t> player;
    std::vector<int32_t> net_offsets, pin_particles;
    std::vector<double> pin_xy;
    std::vector<double> precond;

    mp_gp_static layout(int device, int dct_precision, std::uint64_t maximum_bytes) const;
};

void randomPlacement(const mp_design& d, std::uint64_t seed, double* centers);

// From the first call on (once per process), freed host memory stays in the process heap: glibc's
// trim threshold is raised past any set-up's size and its heap grows in 1 GiB steps.  A job's
// set-up arrays then reuse pages an earlier job already touched instead of faulting in fresh ones
// (a fifth of a 150k-cell job's set-up time).  The placement service and the batch driver call it;
// the process keeps its peak heap for the rest of its life.
void keepFreedHostMemory() noexcept;

// Most host threads parallelFor may use from this thread (8; a caller that already runs many
// preparations side by side lowers it for its small jobs).
extern thread_local int t_parallel_limit;

// Runs body(begin, end) over [0, n) in contiguous parts on up to t_parallel_limit host threads (the
// calling thread takes the first part), or in one part below `grain` items; rethrows a part's
// exception.
template <typename F>
void parallelFor(std::size_t n, std::size_t grain, F&& body) {
    const unsigned hw = std::max(1U, std::thread::hardware_concurrency());
    const std::size_t limit = static_cast<std::size_t>(std::max(1, t_parallel_limit));
    const std::size_t parts =
        std::min<std::size_t>({limit, hw, std::max<std::size_t>(1, n / std::max<std::size_t>(grain, 1))});
    if (parts <= 1) {
        body(std::size_t{0}, n);
        return;
    }
    std::vector<std::exception_ptr> errors(parts);
    std::vector<std::thread> threads;
    auto run = [&](std::size_t k) {
        try {
            body(n * k / parts, n * (k + 1) / parts);
        } catch (...) {
            errors[k] = std::current_exception();
        }
    };
    for (std::size_t k = 1; k < parts; ++k) threads.emplace_back(run, k);
    run(0);
    for (auto& t : threads) t.join();
    for (auto& e : errors)
        if (e) std::rethrow_exception(e);
}

// Sorts the n indices idx[] stably by key[idx[i]] (the low key_bits bits): an LSD radix sort, 11
// bits per pass, the same order as std::stable_sort by key in linear time.
void stableSortByKey(std::uint32_t* idx, std::size_t n, const std::uint32_t* key, int key_bits);

// transformOffset, PlacementDatabase.cpp:55-68 (orientation codes N,S,E,W,FN,FS,FE,FW = 0..7)
void orient(double& x, double& y, int o);

}  // namespace mpgp
