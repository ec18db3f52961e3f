// gp_quadratic.hip — quadratic initial placement (InitialPlacer quadraticPlacement,
// src/placement/InitialPlacer.cpp:266-402) with its linear solves on the device.
//
// Each outer iteration builds, from the current pin positions, the bound-to-bound weighted system
// of the clique (degree <= limit) or star net model plus a 1e-4 anchor at the current position,
// for x and for y, and solves both with unpreconditioned BiCGSTAB from the current positions.
// The sparsity pattern depends only on the netlist and is built once; the host fills the values in
// the order a triplet assembly sums them.  The two solves run in lockstep as one batch of launches
// (blockIdx.y selects the axis); the host applies BiCGSTAB's scalar recurrences between launches
// from deterministic per-block partial sums that one finishing launch writes into host-mapped
// memory.
//
// BiCGSTAB follows van der Vorst (SIAM J. Sci. Stat. Comput. 13(2), 1992) with the conventions of
// Eigen 3.4's BiCGSTAB, which the reference calls with an identity preconditioner
// (InitialPlacer.cpp:386-390): shadow residual r0 = b - A x0; stop when |r|^2 <= tol^2 |b|^2 or after
// the iteration limit; a zero right-hand side gives x = 0; when |rho| < eps^2 |r0|^2 the shadow
// residual is reset to the current residual (the first reset restarts the iteration count);
// omega = 0 when |t| = 0.

#include "gp_backend.hpp"
#include "gp_host.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

namespace mpgp {

extern thread_local std::string g_last_error;
mp_gp_status translate(const std::exception_ptr& e);
std::size_t allocationLimit(int device, std::uint64_t maximum_bytes, std::string& reason);

namespace {

constexpr int kQpThreads = 256;
constexpr int kQpWaves = kQpThreads / 64;
constexpr int kRowLanes = 16;                       // lanes per sparse row
constexpr int kRowsPerBlock = kQpThreads / kRowLanes;
constexpr int kMaxPartials = 2;                     // values per partial row

// Per-axis pointers of one launch (blockIdx.y = axis).
struct Axes {
    double* v[2];
};
struct CAxes {
    const double* v[2];
};

struct Matrix {
    const int* off;   // [n+1], shared pattern
    const int* col;
    CAxes val;        // values per axis
    int n;
};

__device__ inline double waveSumD(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Block sums of K values, written as one partial row of this block (fixed order).
template <int K>
__device__ void blockPartials(double (&v)[K], double* partials) {
    __shared__ double lds[kQpWaves * kMaxPartials];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = waveSumD(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) lds[wave * K + k] = v[k];
    __syncthreads();
    if (threadIdx.x < K) {
        double a = 0.0;
        for (int w = 0; w < kQpWaves; ++w) a += lds[w * K + threadIdx.x];
        partials[static_cast<size_t>(blockIdx.x) * K + threadIdx.x] = a;
    }
}

// Row r of A times x by a group of kRowLanes lanes; the result is in every lane of the group.
__device__ inline double rowProduct(const Matrix& A, const double* val, const double* x, int r, int sub) {
    double a = 0.0;
    for (int k = A.off[r] + sub; k < A.off[r + 1]; k += kRowLanes) a += val[k] * x[A.col[k]];
#pragma unroll
    for (int o = kRowLanes / 2; o > 0; o >>= 1) a += __shfl_xor(a, o, kRowLanes);
    return a;
}

// r = b - A x, r0 = r; partials (r.r, b.b)
__global__ __launch_bounds__(kQpThreads) void k_qp_residual(Matrix A, CAxes b, CAxes x, Axes r, Axes r0, Axes partials,
                                                            unsigned mask) {
    const int axis = blockIdx.y;
    if (!((mask >> axis) & 1U)) return;
    const int row = blockIdx.x * kRowsPerBlock + threadIdx.x / kRowLanes, sub = threadIdx.x % kRowLanes;
    double v[2] = {0.0, 0.0};
    if (row < A.n) {
        const double ax = rowProduct(A, A.val.v[axis], x.v[axis], row, sub);
        if (sub == 0) {
            const double bb = b.v[axis][row], rr = bb - ax;
            r.v[axis][row] = rr;
            r0.v[axis][row] = rr;
            v[0] = rr * rr;
            v[1] = bb * bb;
        }
    }
    blockPartials<2>(v, partials.v[axis]);
}

// y = A u; partials (w.y) or, with SELF, (w.y, y.y)
template <bool SELF>
__global__ __launch_bounds__(kQpThreads) void k_qp_spmv(Matrix A, CAxes u, Axes y, CAxes w, Axes partials,
                                                        unsigned mask) {
    const int axis = blockIdx.y;
    if (!((mask >> axis) & 1U)) return;
    const int row = blockIdx.x * kRowsPerBlock + threadIdx.x / kRowLanes, sub = threadIdx.x % kRowLanes;
    double v[SELF ? 2 : 1] = {};
    if (row < A.n) {
        const double a = rowProduct(A, A.val.v[axis], u.v[axis], row, sub);
        if (sub == 0) {
            y.v[axis][row] = a;
            v[0] = w.v[axis][row] * a;
            if constexpr (SELF) v[SELF ? 1 : 0] = a * a;
        }
    }
    blockPartials<SELF ? 2 : 1>(v, partials.v[axis]);
}

struct Coef {
    double c[2];
};

// p = r + beta (p - omega v)
__global__ __launch_bounds__(kQpThreads) void k_qp_direction(Axes p, CAxes r, CAxes v, Coef beta, Coef omega, int n,
                                                             unsigned mask) {
    const int axis = blockIdx.y;
    if (!((mask >> axis) & 1U)) return;
    const int i = blockIdx.x * kQpThreads + threadIdx.x;
    if (i >= n) return;
    p.v[axis][i] = r.v[axis][i] + beta.c[axis] * (p.v[axis][i] - omega.c[axis] * v.v[axis][i]);
}

// s = r - alpha v
__global__ __launch_bounds__(kQpThreads) void k_qp_half(Axes s, CAxes r, CAxes v, Coef alpha, int n, unsigned mask) {
    const int axis = blockIdx.y;
    if (!((mask >> axis) & 1U)) return;
    const int i = blockIdx.x * kQpThreads + threadIdx.x;
    if (i >= n) return;
    s.v[axis][i] = r.v[axis][i] - alpha.c[axis] * v.v[axis][i];
}

// x += alpha p + omega s; r = s - omega t; partials (r0.r, r.r)
__global__ __launch_bounds__(kQpThreads) void k_qp_update(Axes x, Axes r, CAxes p, CAxes s, CAxes t, CAxes r0,
                                                          Coef alpha, Coef omega, int n, Axes partials,
                                                          unsigned mask) {
    const int axis = blockIdx.y;
    if (!((mask >> axis) & 1U)) return;
    const int i = blockIdx.x * kQpThreads + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (i < n) {
        const double si = s.v[axis][i];
        x.v[axis][i] += alpha.c[axis] * p.v[axis][i] + omega.c[axis] * si;
        const double ri = si - omega.c[axis] * t.v[axis][i];
        r.v[axis][i] = ri;
        v[0] = r0.v[axis][i] * ri;
        v[1] = ri * ri;
    }
    blockPartials<2>(v, partials.v[axis]);
}

// out[axis * kMaxPartials + k] = sum over blocks of partial k, in block order per thread then a
// fixed combination (one workgroup per value and axis).
__global__ __launch_bounds__(kQpThreads) void k_qp_finish(CAxes partials, int blocks, int K, double* out,
                                                          unsigned mask) {
    const int k = blockIdx.x, axis = blockIdx.y;
    if (!((mask >> axis) & 1U)) return;
    double a = 0.0;
    for (int b = threadIdx.x; b < blocks; b += kQpThreads) a += partials.v[axis][static_cast<size_t>(b) * K + k];
    __shared__ double lds[kQpWaves];
    a = waveSumD(a);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < kQpWaves; ++w) s += lds[w];
        out[axis * kMaxPartials + k] = s;
    }
}

void check(hipError_t e, const char* what) { hipCheck(e, what); }

// ---- host side -----------------------------------------------------------------------------------

// One pair term of the system (addPairToQuadraticSystem, InitialPlacer.cpp:266-313).
struct Pair {
    int lp, rp;        // pins
    int li, ri;        // variable indices, -1 for a fixed module
    double base;       // clique or star weight
};

class QuadraticPlacer {
public:
    QuadraticPlacer(const mp_design& d, const mp_qp_options& o, double* centers) : net_(d), o_(o), centers_(centers) {}

    mp_qp_result run() {
        const auto t0 = std::chrono::steady_clock::now();
        mp_qp_result res{};
        res.hpwl_before = net_.hpwl(net_.cx.data(), net_.cy.data());
        cx_ = net_.cx;
        cy_ = net_.cy;
        if (!net_.movable.empty()) solve(res);
        res.hpwl_after = net_.hpwl(cx_.data(), cy_.data());
        for (std::size_t m = 0; m < net_.M; ++m) {
            centers_[2 * m] = cx_[m];
            centers_[2 * m + 1] = cy_[m];
        }
        res.elapsed_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return res;
    }

private:
    void clampToCore(std::size_t m) {  // PlacementDatabase.cpp:202-213
        const auto& c = net_.core;
        if (std::max(0.0, c[2] - c[0]) * std::max(0.0, c[3] - c[1]) <= 1e-9) return;
        const double lx = c[0] + net_.width[m] * 0.5, ux = c[2] - net_.width[m] * 0.5;
        const double ly = c[1] + net_.height[m] * 0.5, uy = c[3] - net_.height[m] * 0.5;
        cx_[m] = lx <= ux ? std::max(lx, std::min(cx_[m], ux)) : (c[0] + c[2]) * 0.5;
        cy_[m] = ly <= uy ? std::max(ly, std::min(cy_[m], uy)) : (c[1] + c[3]) * 0.5;
    }

    void buildPattern() {
        const std::size_t M = net_.M;
        index_.assign(M, -1);
        for (std::size_t k = 0; k < n_; ++k) index_[static_cast<std::size_t>(net_.movable[k])] = static_cast<int>(k);
        auto addPair = [&](int64_t lp, int64_t rp, double base) {
            const int lm = net_.pin_module[lp], rm = net_.pin_module[rp];
            if (lm == rm) return;
            const int li = index_[lm], ri = index_[rm];
            if (li < 0 && ri < 0) return;
            pairs_.push_back(Pair{static_cast<int>(lp), static_cast<int>(rp), li, ri, base});
        };
        for (std::size_t net = 0; net + 1 < net_.net_start.size(); ++net) {
            const int64_t b = net_.net_start[net], e = net_.net_start[net + 1];
            const int64_t degree = e - b;
            if (degree < 2) continue;
            const double w = net_.net_weight[net];
            if (static_cast<uint64_t>(degree) <= o_.clique_degree_limit) {
                const double cw = 2.0 * w / static_cast<double>(degree - 1);
                for (int64_t l = b; l < e; ++l)
                    for (int64_t r = l + 1; r < e; ++r) addPair(l, r, cw);
            } else {
                const double sw = w / static_cast<double>(degree - 1);
                for (int64_t k = b + 1; k < e; ++k) addPair(b, k, sw);
            }
        }
        // entries in the order the triplets are added: per pair (li,li) (ri,ri) (li,ri) (ri,li), then
        // the anchors; a stable bucket sort by row, then by column within the row, gives the slots
        std::vector<int> erow, ecol;
        erow.reserve(pairs_.size() * 4 + n_);
        ecol.reserve(pairs_.size() * 4 + n_);
        for (const Pair& p : pairs_) {
            if (p.li >= 0 && p.ri >= 0) {
                erow.insert(erow.end(), {p.li, p.ri, p.li, p.ri});
                ecol.insert(ecol.end(), {p.li, p.ri, p.ri, p.li});
            } else {
                const int i = p.li >= 0 ? p.li : p.ri;
                erow.push_back(i);
                ecol.push_back(i);
            }
        }
        for (std::size_t k = 0; k < n_; ++k) {
            erow.push_back(static_cast<int>(k));
            ecol.push_back(static_cast<int>(k));
        }
        const std::size_t E = erow.size();
        if (E > static_cast<std::size_t>(std::numeric_limits<int>::max()))
            throw std::runtime_error("quadratic system too large for 32-bit indices");
        std::vector<int> start(n_ + 1, 0);
        for (std::size_t e = 0; e < E; ++e) ++start[erow[e] + 1];
        std::partial_sum(start.begin(), start.end(), start.begin());
        std::vector<int> byrow(E);
        {
            std::vector<int> fill(start.begin(), start.end() - 1);
            for (std::size_t e = 0; e < E; ++e) byrow[fill[erow[e]]++] = static_cast<int>(e);
        }
        entry_slot_.assign(E, -1);
        off_.assign(n_ + 1, 0);
        col_.clear();
        for (std::size_t r = 0; r < n_; ++r) {
            auto first = byrow.begin() + start[r], last = byrow.begin() + start[r + 1];
            std::stable_sort(first, last, [&](int a, int b) { return ecol[a] < ecol[b]; });
            for (auto it = first; it != last; ++it) {
                if (it == first || ecol[*it] != ecol[*(it - 1)]) col_.push_back(ecol[*it]);
                entry_slot_[*it] = static_cast<int>(col_.size()) - 1;
            }
            off_[r + 1] = static_cast<int>(col_.size());
        }
    }

    // Fills the values and right-hand sides from the current positions (:266-313, :369-376).
    void assemble(std::vector<double>& vx, std::vector<double>& vy, std::vector<double>& bx, std::vector<double>& by) {
        vx.assign(col_.size(), 0.0);
        vy.assign(col_.size(), 0.0);
        bx.assign(n_, 0.0);
        by.assign(n_, 0.0);
        std::size_t e = 0;
        for (const Pair& p : pairs_) {
            const int lm = net_.pin_module[p.lp], rm = net_.pin_module[p.rp];
            double lox = net_.pin_dx[p.lp], loy = net_.pin_dy[p.lp], rox = net_.pin_dx[p.rp], roy = net_.pin_dy[p.rp];
            orient(lox, loy, net_.orientation[lm]);
            orient(rox, roy, net_.orientation[rm]);
            const double lpx = cx_[lm] + lox, lpy = cy_[lm] + loy, rpx = cx_[rm] + rox, rpy = cy_[rm] + roy;
            const double wx = p.base / std::max(std::abs(lpx - rpx), o_.minimum_distance);
            const double wy = p.base / std::max(std::abs(lpy - rpy), o_.minimum_distance);
            if (p.li >= 0 && p.ri >= 0) {
                vx[entry_slot_[e]] += wx; vy[entry_slot_[e]] += wy;
                vx[entry_slot_[e + 1]] += wx; vy[entry_slot_[e + 1]] += wy;
                vx[entry_slot_[e + 2]] += -wx; vy[entry_slot_[e + 2]] += -wy;
                vx[entry_slot_[e + 3]] += -wx; vy[entry_slot_[e + 3]] += -wy;
                e += 4;
                const double dx = lox - rox, dy = loy - roy;
                bx[p.li] -= wx * dx; bx[p.ri] += wx * dx;
                by[p.li] -= wy * dy; by[p.ri] += wy * dy;
            } else if (p.li >= 0) {
                vx[entry_slot_[e]] += wx; vy[entry_slot_[e]] += wy;
                e += 1;
                bx[p.li] += wx * (cx_[rm] + rox - lox);
                by[p.li] += wy * (cy_[rm] + roy - loy);
            } else {
                vx[entry_slot_[e]] += wx; vy[entry_slot_[e]] += wy;
                e += 1;
                bx[p.ri] += wx * (cx_[lm] + lox - rox);
                by[p.ri] += wy * (cy_[lm] + loy - roy);
            }
        }
        constexpr double kAnchor = 1e-4;
        for (std::size_t k = 0; k < n_; ++k) {
            const std::size_t m = static_cast<std::size_t>(net_.movable[k]);
            vx[entry_slot_[e]] += kAnchor;
            vy[entry_slot_[e]] += kAnchor;
            ++e;
            bx[k] += kAnchor * cx_[m];
            by[k] += kAnchor * cy_[m];
        }
    }

    void solve(mp_qp_result& res);

    Netlist net_;
    mp_qp_options o_;
    double* centers_;
    std::vector<double> cx_, cy_;
    std::size_t n_ = 0;
    std::vector<int> index_;
    std::vector<Pair> pairs_;
    std::vector<int> entry_slot_;
    std::vector<int> off_, col_;
};

void QuadraticPlacer::solve(mp_qp_result& res) {
    n_ = net_.movable.size();
    if (n_ > static_cast<std::size_t>(std::numeric_limits<int>::max() / kQpThreads))
        throw InvalidArgument("too many movable modules for the quadratic solve");
    const Netlist& N = net_;
    const auto& core = N.core;
    for (const int64_t id : N.movable) {  // :318-322
        cx_[id] = (core[0] + core[2]) * 0.5;
        cy_[id] = (core[1] + core[3]) * 0.5;
        clampToCore(static_cast<std::size_t>(id));
    }
    buildPattern();
    const int n = static_cast<int>(n_);
    const std::size_t nnz = col_.size();

    std::string reason;
    const std::size_t limit = allocationLimit(o_.device, 64ULL << 30, reason);
    if (limit == 0) throw std::runtime_error("HIP quadratic placement is unavailable: " + reason);
    hipCheck(hipSetDevice(o_.device), "select HIP device");
    DeviceBudget budget(limit);
    hipStream_t s = nullptr;
    hipCheck(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "create HIP stream");
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    } guard{s};
    DeviceArray<int> d_off, d_col;
    d_off.allocate(n_ + 1, budget, "quadratic system offsets");
    d_col.allocate(nnz, budget, "quadratic system columns");
    d_off.upload(off_.data(), off_.size(), s, "upload quadratic pattern");
    d_col.upload(col_.data(), nnz, s, "upload quadratic pattern");
    enum Vec { kX, kB, kR, kR0, kP, kV, kS, kT, kVecs };
    DeviceArray<double> val[2], vec[2][kVecs], part[2];
    const int row_blocks = (n + kRowsPerBlock - 1) / kRowsPerBlock;
    const int el_blocks = (n + kQpThreads - 1) / kQpThreads;
    const int part_rows = std::max(row_blocks, el_blocks);
    for (int a = 0; a < 2; ++a) {
        val[a].allocate(nnz, budget, "quadratic system values");
        for (int v = 0; v < kVecs; ++v) vec[a][v].allocate(n_, budget, "BiCGSTAB vector");
        part[a].allocate(static_cast<std::size_t>(part_rows) * kMaxPartials, budget, "BiCGSTAB partials");
    }
    double* host = nullptr;
    hipCheck(hipHostMalloc(reinterpret_cast<void**>(&host), 4 * sizeof(double), hipHostMallocMapped | hipHostMallocCoherent),
             "allocate host-mapped sums");
    struct HostGuard {
        double* p;
        ~HostGuard() { (void)hipHostFree(p); }
    } hguard{host};
    double* dev_sums = nullptr;
    hipCheck(hipHostGetDevicePointer(reinterpret_cast<void**>(&dev_sums), host, 0), "map host sums");

    auto ax = [&](Vec v) { return Axes{{vec[0][v].get(), vec[1][v].get()}}; };
    auto cax = [&](Vec v) { return CAxes{{vec[0][v].get(), vec[1][v].get()}}; };
    const Matrix A{d_off.get(), d_col.get(), CAxes{{val[0].get(), val[1].get()}}, n};
    const Axes P{{part[0].get(), part[1].get()}};
    const CAxes CP{{part[0].get(), part[1].get()}};
    // finish K partial values of `blocks` rows for the axes in `mask`, wait, return them
    auto finish = [&](int blocks, int K, unsigned mask) {
        hipLaunchKernelGGL(k_qp_finish, dim3(K, 2), dim3(kQpThreads), 0, s, CP, blocks, K, dev_sums, mask);
        check(hipGetLastError(), "launch BiCGSTAB finish kernel");
        hipCheck(hipStreamSynchronize(s), "BiCGSTAB reduction");
    };
    auto residual = [&](unsigned mask) {
        hipLaunchKernelGGL(k_qp_residual, dim3(row_blocks, 2), dim3(kQpThreads), 0, s, A, cax(kB), cax(kX), ax(kR),
                           ax(kR0), P, mask);
        check(hipGetLastError(), "launch BiCGSTAB residual kernel");
        finish(row_blocks, 2, mask);
    };

    std::vector<double> vx, vy, bx, by, xs(n_), ys(n_);
    const double eps = std::numeric_limits<double>::epsilon();
    for (int outer = 0; outer < o_.outer_iterations; ++outer) {
        assemble(vx, vy, bx, by);
        for (std::size_t k = 0; k < n_; ++k) {
            xs[k] = cx_[N.movable[k]];
            ys[k] = cy_[N.movable[k]];
        }
        val[0].upload(vx.data(), nnz, s, "upload quadratic values");
        val[1].upload(vy.data(), nnz, s, "upload quadratic values");
        vec[0][kB].upload(bx.data(), n_, s, "upload right-hand side");
        vec[1][kB].upload(by.data(), n_, s, "upload right-hand side");
        vec[0][kX].upload(xs.data(), n_, s, "upload positions");
        vec[1][kX].upload(ys.data(), n_, s, "upload positions");
        for (int a = 0; a < 2; ++a)
            for (Vec v : {kP, kV}) hipCheck(hipMemsetAsync(vec[a][v].get(), 0, n_ * sizeof(double), s), "clear vector");

        residual(3U);
        double r2[2] = {host[0], host[2]}, b2[2] = {host[1], host[3]};
        double r0_2[2] = {r2[0], r2[1]}, rho[2] = {1.0, 1.0}, alpha[2] = {1.0, 1.0}, omega[2] = {1.0, 1.0};
        double next_rho[2] = {r2[0], r2[1]};  // r0 = r at the start: r0.r = r.r
        double tol2[2];
        int iters[2] = {0, 0}, restarts[2] = {0, 0};
        unsigned live = 0;
        for (int a = 0; a < 2; ++a) {
            tol2[a] = o_.tolerance * o_.tolerance * b2[a];
            if (b2[a] == 0.0) hipCheck(hipMemsetAsync(vec[a][kX].get(), 0, n_ * sizeof(double), s), "zero solution");
            else live |= 1U << a;
        }
        for (;;) {
            unsigned mask = 0;
            for (int a = 0; a < 2; ++a)
                if (((live >> a) & 1U) && r2[a] > tol2[a] && iters[a] < o_.solver_iterations) mask |= 1U << a;
            if (!mask) break;
            Coef beta{}, om{}, al{};
            unsigned restart = 0;
            for (int a = 0; a < 2; ++a) {
                if (!((mask >> a) & 1U)) continue;
                if (std::abs(next_rho[a]) < eps * eps * r0_2[a]) restart |= 1U << a;
            }
            if (restart) {
                residual(restart);
                for (int a = 0; a < 2; ++a) {
                    if (!((restart >> a) & 1U)) continue;
                    r2[a] = host[2 * a];
                    r0_2[a] = r2[a];
                    next_rho[a] = r2[a];
                    if (restarts[a]++ == 0) iters[a] = 0;
                }
            }
            for (int a = 0; a < 2; ++a) {
                if (!((mask >> a) & 1U)) continue;
                const double rho_old = rho[a];
                rho[a] = next_rho[a];
                beta.c[a] = (rho[a] / rho_old) * (alpha[a] / omega[a]);
                om.c[a] = omega[a];
            }
            hipLaunchKernelGGL(k_qp_direction, dim3(el_blocks, 2), dim3(kQpThreads), 0, s, ax(kP), cax(kR), cax(kV),
                               beta, om, n, mask);
            check(hipGetLastError(), "launch BiCGSTAB direction kernel");
            hipLaunchKernelGGL(k_qp_spmv<false>, dim3(row_blocks, 2), dim3(kQpThreads), 0, s, A, cax(kP), ax(kV),
                               cax(kR0), P, mask);
            check(hipGetLastError(), "launch BiCGSTAB product kernel");
            finish(row_blocks, 1, mask);
            for (int a = 0; a < 2; ++a)
                if ((mask >> a) & 1U) al.c[a] = alpha[a] = rho[a] / host[2 * a];
            hipLaunchKernelGGL(k_qp_half, dim3(el_blocks, 2), dim3(kQpThreads), 0, s, ax(kS), cax(kR), cax(kV), al, n,
                               mask);
            check(hipGetLastError(), "launch BiCGSTAB half-step kernel");
            hipLaunchKernelGGL(k_qp_spmv<true>, dim3(row_blocks, 2), dim3(kQpThreads), 0, s, A, cax(kS), ax(kT),
                               cax(kS), P, mask);
            check(hipGetLastError(), "launch BiCGSTAB product kernel");
            finish(row_blocks, 2, mask);
            for (int a = 0; a < 2; ++a) {
                if (!((mask >> a) & 1U)) continue;
                const double ts = host[2 * a], tt = host[2 * a + 1];
                om.c[a] = omega[a] = tt > 0.0 ? ts / tt : 0.0;
            }
            hipLaunchKernelGGL(k_qp_update, dim3(el_blocks, 2), dim3(kQpThreads), 0, s, ax(kX), ax(kR), cax(kP),
                               cax(kS), cax(kT), cax(kR0), al, om, n, P, mask);
            check(hipGetLastError(), "launch BiCGSTAB update kernel");
            finish(el_blocks, 2, mask);
            for (int a = 0; a < 2; ++a) {
                if (!((mask >> a) & 1U)) continue;
                next_rho[a] = host[2 * a];
                r2[a] = host[2 * a + 1];
                ++iters[a];
            }
        }
        res.solver_iterations += iters[0] + iters[1];
        for (int a = 0; a < 2; ++a) res.relative_residual[a] = b2[a] > 0.0 ? std::sqrt(r2[a] / b2[a]) : 0.0;
        hipCheck(hipMemcpyAsync(xs.data(), vec[0][kX].get(), n_ * sizeof(double), hipMemcpyDeviceToHost, s),
                 "download solution");
        hipCheck(hipMemcpyAsync(ys.data(), vec[1][kX].get(), n_ * sizeof(double), hipMemcpyDeviceToHost, s),
                 "download solution");
        hipCheck(hipStreamSynchronize(s), "download solution");
        for (std::size_t k = 0; k < n_; ++k)
            if (!std::isfinite(xs[k]) || !std::isfinite(ys[k]))
                throw std::runtime_error("Quadratic placement failed to produce finite coordinates.");
        for (std::size_t k = 0; k < n_; ++k) {  // :392-399
            const std::size_t m = static_cast<std::size_t>(N.movable[k]);
            cx_[m] = xs[k];
            cy_[m] = ys[k];
            clampToCore(m);
        }
        res.iterations = outer + 1;
    }
}

}  // namespace
}  // namespace mpgp

using namespace mpgp;

extern "C" {

void mp_gp_default_qp_options(mp_qp_options* o) {  // InitialPlacer.hpp:21-25
    if (!o) return;
    *o = {};
    o->outer_iterations = 5;
    o->solver_iterations = 100;
    o->tolerance = 1e-5;
    o->minimum_distance = 1.0;
    o->clique_degree_limit = 96;
    o->device = 0;
}

mp_gp_status mp_gp_quadratic_placement(const mp_design* design, const mp_qp_options* options, double* centers,
                                       mp_qp_result* result) {
    g_last_error.clear();
    try {
        if (!design || !options || !centers) throw std::invalid_argument("null design, options or centres");
        const mp_qp_options& o = *options;
        if (o.outer_iterations <= 0 || o.solver_iterations <= 0 || !std::isfinite(o.tolerance) || o.tolerance <= 0.0 ||
            !std::isfinite(o.minimum_distance) || o.minimum_distance <= 0.0)  // InitialPlacer.cpp:52-57
            throw std::invalid_argument("Quadratic initial-placement options are outside their valid range.");
        QuadraticPlacer qp(*design, o, centers);
        const mp_qp_result r = qp.run();
        if (result) *result = r;
        return MP_GP_OK;
    } catch (...) {
        return translate(std::current_exception());
    }
}

}  // extern "C"
