// gp_host.hpp — host-side netlist view and run set-up (see gp_host.cpp).
#pragma once

#include "mp_gp_backend.h"

#include <array>
#include <cstdint>
#include <limits>
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

// transformOffset, PlacementDatabase.cpp:55-68 (orientation codes N,S,E,W,FN,FS,FE,FW = 0..7)
void orient(double& x, double& y, int o);

}  // namespace mpgp
