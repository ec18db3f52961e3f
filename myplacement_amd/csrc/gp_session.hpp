// gp_session.hpp — the adaptive Nesterov / backtracking controller of runAdaptiveGlobalPlacement
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
//
// The single-run entry points (gp_session.cpp) and the batch driver (gp_batch.cpp) share it.

#pragma once

#include "gp_backend.hpp"
#include "gp_host.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <coroutine>
#include <exception>
#include <future>
#include <limits>
#include <memory>
#include <numeric>
#include <string>
#include <utility>
#include <vector>

namespace mpgp {

extern thread_local std::string g_last_error;
mp_gp_status translate(const std::exception_ptr& e);
std::size_t allocationLimit(int device, std::uint64_t maximum_bytes, std::string& reason);
// AdaptiveGlobalPlacer.cpp:413-454 (non-RUDY part)
void validateOptions(const mp_gp_options& o);
// HPWL and design overflow of the starting placement, the "before" figures of the result
// (GlobalPlacer.cpp:134-146): for a large design on a thread of its own beside the run (they read
// only the prepared netlist), else at once.
std::future<std::pair<double, double>> beforeMetrics(std::shared_ptr<const Prepared> prep, const mp_gp_options& o);


namespace session_detail {
constexpr double kEps = 1e-9;
inline double clampd(double v, double lo, double hi) { return std::max(lo, std::min(v, hi)); }

struct EvalRec {
    mp_gp_scalars s{};
    double objective = 0.0;
    double pre_norm = 0.0;
    double wire_l1 = 0.0, density_l1 = 0.0;
};

struct Checkpoint {
    int iteration = 0;
    double hpwl = std::numeric_limits<double>::infinity();
    double overflow = std::numeric_limits<double>::infinity();
    double score = std::numeric_limits<double>::infinity();
    bool found = false;
};
}  // namespace session_detail

// A step of the controller that runs until it needs the device's scalars.  Started eagerly; when
// the session launches straight onto its stream the wait is served in place and the step runs to
// the end inside the call.  Under a batch driver the step suspends at each wait, after recording
// its launches, and the driver resumes it once the batched launches have delivered the scalars.
class Step {
public:
    struct promise_type {
        std::exception_ptr error;
        Step get_return_object() { return Step(std::coroutine_handle<promise_type>::from_promise(*this)); }
        std::suspend_never initial_suspend() noexcept { return {}; }
        std::suspend_always final_suspend() noexcept { return {}; }
        void return_void() {}
        void unhandled_exception() { error = std::current_exception(); }
    };
    Step() = default;
    explicit Step(std::coroutine_handle<promise_type> h) : h_(h) {}
    Step(Step&& o) noexcept : h_(std::exchange(o.h_, {})) {}
    Step& operator=(Step&& o) noexcept {
        if (this != &o) {
            if (h_) h_.destroy();
            h_ = std::exchange(o.h_, {});
        }
        return *this;
    }
    ~Step() {
        if (h_) h_.destroy();
    }
    explicit operator bool() const { return static_cast<bool>(h_); }
    bool done() const { return h_.done(); }
    // rethrows what the finished step threw
    void result() const {
        if (h_.promise().error) std::rethrow_exception(h_.promise().error);
    }

private:
    std::coroutine_handle<promise_type> h_;
};

class Session {
    using EvalRec = session_detail::EvalRec;
    using Checkpoint = session_detail::Checkpoint;
    static constexpr double kEps = session_detail::kEps;
    static double clampd(double v, double lo, double hi) { return session_detail::clampd(v, lo, hi); }

public:
    // batched: the session's launches are recorded for a batch driver, which also starts and
    // resumes its steps (initialize, then iterate until stopped)
    Session(std::shared_ptr<const Prepared> prep, const mp_gp_options& o, const double* centers, bool batched = false,
            const BorrowedResources* borrowed = nullptr)
        : prep_(std::move(prep)), o_(o) {
        validateOptions(o_);
        const Prepared& p = *prep_;
        std::string reason;
        const std::size_t limit = allocationLimit(o_.device, o_.maximum_device_bytes, reason);
        if (limit == 0) throw std::runtime_error("HIP global-placement backend is unavailable: " + reason);
        const std::size_t Mv = p.net.movable.size();
        const std::size_t P = p.pw.size();
        std::vector<double2> x0(P);  // starting positions in the prepared particle order
        for (std::size_t i = 0; i < Mv; ++i) {
            const int64_t m = p.net.movable[i];
            x0[i] = centers ? make_double2(centers[2 * m], centers[2 * m + 1]) : make_double2(p.net.cx[m], p.net.cy[m]);
        }
        for (std::size_t i = 0; i < p.filler_x.size(); ++i) x0[Mv + i] = make_double2(p.filler_x[i], p.filler_y[i]);
        // The device keeps the particles renumbered along a Z-order curve of their starting
        // positions (movable modules first, then fillers): particles close on the chip get close
        // indices, so the kernels' gathers by particle and by pin hit fewer cache lines.  Results are
        // mapped back to the prepared order on the way out.
        order_.resize(P);
        std::iota(order_.begin(), order_.end(), 0U);
        auto zkey = [&](std::size_t i) {
            const double fx = (x0[i].x - p.net.core[0]) / std::max(p.net.core[2] - p.net.core[0], 1e-9);
            const double fy = (x0[i].y - p.net.core[1]) / std::max(p.net.core[3] - p.net.core[1], 1e-9);
            const uint32_t ix = static_cast<uint32_t>(clampd(fx, 0.0, 1.0) * 1023.0);
            const uint32_t iy = static_cast<uint32_t>(clampd(fy, 0.0, 1.0) * 1023.0);
            uint32_t k = 0;
            for (int b = 9; b >= 0; --b) k = (k << 2) | (((iy >> b) & 1U) << 1) | ((ix >> b) & 1U);
            return k;
        };
        std::vector<uint32_t> keys(P);
        parallelFor(P, 65536, [&](std::size_t i0, std::size_t i1) {
            for (std::size_t i = i0; i < i1; ++i) keys[i] = zkey(i);
        });
        stableSortByKey(order_.data(), Mv, keys.data(), 20);
        stableSortByKey(order_.data() + Mv, P - Mv, keys.data(), 20);
        std::vector<int32_t> rank(P);
        for (std::size_t k = 0; k < P; ++k) rank[order_[k]] = static_cast<int32_t>(k);
        std::vector<double> pw(P), ph(P), pq(P), precond(P);
        std::vector<int32_t> player(P), pin_particles(p.pin_particles.size());
        std::vector<double2> x(P);
        parallelFor(P, 65536, [&](std::size_t k0, std::size_t k1) {
            for (std::size_t k = k0; k < k1; ++k) {
                const uint32_t i = order_[k];
                pw[k] = p.pw[i];
                ph[k] = p.ph[i];
                pq[k] = p.pq[i];
                player[k] = p.player[i];
                precond[k] = p.precond[i];
                x[k] = x0[i];
            }
        });
        parallelFor(pin_particles.size(), 131072, [&](std::size_t q0, std::size_t q1) {
            for (std::size_t q = q0; q < q1; ++q)
                pin_particles[q] = p.pin_particles[q] < 0 ? -1 : rank[p.pin_particles[q]];
        });
        mp_gp_static L = p.layout(o_.device, o_.dct_precision, o_.maximum_device_bytes);
        L.particle_width = pw.data();
        L.particle_height = ph.data();
        L.particle_charge_area = pq.data();
        L.particle_layer = player.data();
        L.pin_particles = pin_particles.data();
        be_ = std::make_unique<Backend>(L, limit, borrowed);
        P_ = be_->particles();
        DeviceBudget& b = be_->budget();
        for (DeviceArray<double2>* a : {&X_, &XP_, &B_, &Y_, &PX_, &PG_, &PXn_, &PGn_, &Rc_, &Gc_, &Rl_, &Gl_, &CKf_, &CKb_})
            a->allocate(P_, b, "controller vector");
        precond_.allocate(P_, b, "preconditioner");
        hipStream_t s = be_->stream();
        X_.upload(x.data(), P_, s, "upload positions");
        precond_.upload(precond.data(), P_, s, "upload preconditioner");
        hipCheck(hipStreamSynchronize(s), "upload session vectors");  // the host vectors above go out of scope
        min_bin_ = std::min(p.bw, p.bh);
        max_bin_ = std::max(p.bw, p.bh);
        clamp_ = ClampArgs{be_->sizes(), be_->region()};
        bin_height_ = be_->region().bh;
        if (batched) {
            rec_ = std::make_unique<RecordLaunches>();
            be_->setLaunches(rec_.get());
            return;
        }
        const auto t0 = std::chrono::steady_clock::now();
        runInPlace(initialize());
        setup_seconds_ += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    Backend& backend() { return *be_; }
    bool stopped() const { return stopped_; }
    const std::vector<mp_gp_history_row>& history() const { return history_; }

    int step(int iterations) {
        if (rec_) throw std::logic_error("a batched session is stepped by its batch driver");
        const auto t0 = std::chrono::steady_clock::now();
        int done = 0;
        while (done < iterations && !stopped_) {
            runInPlace(iterate());
            ++done;
        }
        loop_seconds_ += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return done;
    }

    // positions in the prepared particle order
    void positions(double* xy) {
        hipCheck(hipSetDevice(o_.device), "select HIP device");
        std::vector<double2> pos(P_);
        hipCheck(hipMemcpyAsync(pos.data(), X_.get(), P_ * sizeof(double2), hipMemcpyDeviceToHost, be_->stream()),
                 "download positions");
        hipCheck(hipStreamSynchronize(be_->stream()), "download positions");
        for (int k = 0; k < P_; ++k) {
            xy[2 * order_[k]] = pos[k].x;
            xy[2 * order_[k] + 1] = pos[k].y;
        }
    }

    // ---- batch driver interface ----------------------------------------------------------------
    RecordLaunches* recorder() { return rec_.get(); }
    // the suspended step's continuation (null when none waits)
    std::coroutine_handle<> takeWaiting() { return std::exchange(waiting_, {}); }
    bool waiting() const { return static_cast<bool>(waiting_); }
    // hands the scalars the batched launches delivered to the suspended step
    void deliver(const HostScalars* h) { scalars_ = h; }
    Step iteration() { return iterate(); }
    int particles() const { return P_; }
    void addTime(double setup, double loop) {
        setup_seconds_ += setup;
        loop_seconds_ += loop;
    }
    // back to launching onto the session's own stream (after the driver's last round for it)
    void detach() {
        be_->setLaunches(nullptr);
        rec_.reset();
    }

    // :672-733
    Step initialize() {
        smoothing_ = o_.initial_smoothing * max_bin_;
        penalty_ = 1.0;
        reference_penalty_ = 1.0;
        enqueueGradEval();
        EvalRec init = record(co_await ScalarWait{this}, true, kBankCurrent);
        initial_overflow_ = std::max(init.s.optimizer_density.normalized_overflow, kEps);
        smoothing_ = smoothingFromOverflow(init.s.optimizer_density.normalized_overflow);
        const double raw_penalty = init.density_l1 > kEps ? init.wire_l1 / init.density_l1 : 1.0;
        penalty_ = clampd(raw_penalty, o_.minimum_density_penalty, o_.maximum_density_penalty);
        reference_penalty_ = penalty_;
        enqueueGradEval();
        init = record(co_await ScalarWait{this}, true, kBankCurrent);
        last_overflow_ = init.s.optimizer_density.normalized_overflow;
        const double rms = init.pre_norm / std::sqrt(static_cast<double>(std::max(1, P_)));
        last_step_ = o_.initial_movement_in_bins * min_bin_ / std::max(rms, kEps);
        bool dummy = false;
        considerCheckpoint(init, 0, dummy);
        feasible_refinement_ = init.s.design_density.normalized_overflow <= o_.maximum_bin_overflow ? 1 : 0;
        current_ = init;
    }

    void finish(double* centers, mp_gp_result* r) {
        hipCheck(hipSetDevice(o_.device), "select HIP device");
        flushCheckpoint();
        const Prepared& p = *prep_;
        const Checkpoint& chosen = feasible_.found ? feasible_ : fallback_;
        std::vector<double2> pos(P_);
        const DeviceArray<double2>& src = feasible_.found ? CKf_ : (fallback_.found ? CKb_ : X_);
        hipCheck(hipMemcpyAsync(pos.data(), src.get(), P_ * sizeof(double2), hipMemcpyDeviceToHost, be_->stream()),
                 "download checkpoint");
        hipCheck(hipStreamSynchronize(be_->stream()), "download checkpoint");
        std::vector<double> cx = p.net.cx, cy = p.net.cy;
        for (int k = 0; k < P_; ++k) {
            const std::size_t i = order_[k];
            if (i >= p.net.movable.size()) continue;  // filler
            const int64_t m = p.net.movable[i];
            cx[m] = pos[k].x;
            cy[m] = pos[k].y;
        }
        if (centers)
            for (std::size_t m = 0; m < p.net.M; ++m) {
                centers[2 * m] = cx[m];
                centers[2 * m + 1] = cy[m];
            }
        result_.restored_best_checkpoint = chosen.found ? 1 : 0;
        if (chosen.found) {
            result_.best_checkpoint_iteration = chosen.iteration;
            result_.best_checkpoint_hpwl = chosen.hpwl;
            result_.best_checkpoint_overflow = chosen.overflow;
        }
        result_.hpwl_after = p.net.hpwl(cx.data(), cy.data());
        result_.overflow_after = p.net.designOverflow(cx.data(), cy.data(), p.nx, p.ny, p.t);
        result_.loop_seconds = loop_seconds_;
        result_.elapsed_seconds = setup_seconds_ + loop_seconds_;
        result_.particle_count = P_;
        result_.filler_count = static_cast<int64_t>(p.filler_x.size());
        result_.device_used = be_->device();
        result_.reserved_device_bytes = be_->reservedBytes();
        if (r) *r = result_;
    }

    mp_gp_result& result() { return result_; }
    const std::vector<uint32_t>& order() const { return order_; }

private:
    // Awaiting the device's scalars: served in place (fetchScalars) when the session launches onto
    // its own stream; otherwise the step suspends until the batch driver delivers them.
    struct ScalarWait {
        Session* s;
        bool await_ready() {
            if (s->rec_) return false;
            s->scalars_ = &s->be_->fetchScalars();
            return true;
        }
        void await_suspend(std::coroutine_handle<> h) { s->waiting_ = h; }
        const HostScalars& await_resume() const { return *s->scalars_; }
    };
    // runs a step that launches onto the session's own stream: it never suspends
    static void runInPlace(const Step& st) {
        if (!st.done()) throw std::logic_error("a session step suspended outside a batch driver");
        st.result();
    }

    // ---- evaluation helpers -----------------------------------------------------------------
    double strength() const { return std::sqrt(1.0 + std::max(0.0, penalty_ / std::max(reference_penalty_, kEps))); }

    // Scalar banks: 0 current-position gradient evaluation, 1 look-ahead gradient evaluation,
    // 2 proposal + candidate value evaluation.
    enum Bank : int { kBankCurrent = 0, kBankLook = 1, kBankCandidate = 2 };

    Backend::SessionOut sessionOut(double2* raw, double2* pre, const double2* prev, double2* ck) const {
        Backend::SessionOut so{precond_.get(), penalty_, 1.0 / strength(), raw, pre};
        so.prev_pos = prev;
        if (has_previous_) {
            so.px = PX_.get();
            so.pg = PG_.get();
        }
        so.ck_out = ck;
        return so;
    }
    EvalRec record(const HostScalars& h, bool grad, int bank) {
        EvalRec e;
        be_->toScalars(h, &e.s, bank);
        const double* v = h.bank(bank);
        e.objective = e.s.smooth_wirelength + penalty_ * e.s.electrostatic_energy;
        if (!std::isfinite(e.objective) || !std::isfinite(e.s.hpwl))
            throw std::runtime_error("HIP adaptive placement objective produced a non-finite value.");
        if (grad) {
            e.pre_norm = std::sqrt(v[kPreNorm2]);
            e.wire_l1 = v[kWireL1];
            e.density_l1 = v[kDensityL1];
            ++result_.gradient_evaluations;
        } else {
            ++result_.value_evaluations;
        }
        return e;
    }
    // gradient evaluation at X into bank 0 (read back by the caller's wait)
    void enqueueGradEval() {
        const Backend::SessionOut so = sessionOut(Rc_.get(), Gc_.get(), nullptr, nullptr);
        const Backend::EvalSpec spec{X_.get(), &so, kBankCurrent, true};
        be_->enqueueEvaluate(&spec, 1, smoothing_, true);
        dX_ = 0.0;
    }
    // After a candidate's evaluation at Y: its wirelength gradient at the smoothing the next
    // iteration will use if Y is accepted (the law at Y's optimizer overflow) is enqueued right after
    // the read-back's finalize, so it runs while the host decides.  Only where the device evaluates
    // the law as the host does (exponent 1) and the launches are the session's own (a batch driver
    // runs its own finalize).
    void requestWireAhead() {
        wire_ahead_ = !rec_ && o_.recompute_density == 0 && o_.smoothing_overflow_exponent == 1.0;
        if (wire_ahead_)
            be_->wireAhead(Backend::WireAhead{Y_.get(), kBankCurrent,
                                              be_->deviceBank(kBankCandidate) + kOptDensity + 1,
                                              std::max(initial_overflow_, kEps), o_.final_smoothing,
                                              o_.initial_smoothing - o_.final_smoothing, max_bin_});
    }
    // A new best position is copied by the next gradient evaluation at X (its gather writes the
    // copy); a copy still pending when that evaluation will not come is made here.
    void flushCheckpoint() {
        if (!pending_ck_) return;
        be_->launches().copy(pending_ck_->get(), X_.get(), P_);
        pending_ck_ = nullptr;
    }

    // ---- control law (AdaptiveGlobalPlacer.cpp:456-596) ------------------------------------------
    double smoothingFromOverflow(double overflow) const {
        const double ratio = clampd(overflow / std::max(initial_overflow_, kEps), 0.0, 1.0);
        const double annealing = std::pow(ratio, o_.smoothing_overflow_exponent);
        return (o_.final_smoothing + (o_.initial_smoothing - o_.final_smoothing) * annealing) * max_bin_;
    }
    bool accepts(const EvalRec& base, const EvalRec& cand, double predicted) const {
        const double tol = std::max(1e-8, std::abs(base.objective) * 1e-10);
        const double armijo_target = base.objective - o_.armijo_coefficient * std::max(0.0, predicted);
        const bool armijo = cand.objective <= armijo_target + tol;
        const double bo = base.s.optimizer_density.normalized_overflow, co = cand.s.optimizer_density.normalized_overflow;
        const bool stable = cand.objective <= base.objective + tol && co <= bo * 1.005;
        const double gain = std::max(1e-4, o_.density_stall_threshold * 0.25);
        const bool trade = co <= bo - gain && cand.objective <= base.objective * (1.0 + o_.objective_increase_for_density);
        return armijo || stable || trade;
    }
    double updatePenalty(double penalty, const EvalRec& before, const EvalRec& after) const {
        const double growth = (after.s.hpwl - before.s.hpwl) / std::max(before.s.hpwl, 1.0);
        const double ao = after.s.optimizer_density.normalized_overflow;
        const double improvement = before.s.optimizer_density.normalized_overflow - ao;
        const bool critical = ao > std::max(2.0 * o_.maximum_bin_overflow, 0.02);
        if (critical && improvement < o_.density_stall_threshold) penalty *= o_.density_penalty_increase;
        else if (!critical && growth > o_.hpwl_growth_restart_threshold && improvement < o_.density_stall_threshold)
            penalty *= o_.density_penalty_decrease;
        else if (ao > o_.maximum_bin_overflow && improvement < o_.density_stall_threshold)
            penalty *= o_.density_penalty_increase;
        else if (after.s.design_density.normalized_overflow <= o_.maximum_bin_overflow)
            penalty *= o_.density_penalty_decrease;
        const double lo = std::max(o_.minimum_density_penalty, reference_penalty_ * 0.02);
        const double hi = std::max(o_.maximum_density_penalty, lo * 10.0);
        return clampd(penalty, lo, hi);
    }
    void considerCheckpoint(const EvalRec& e, int iteration, bool& is_new) {
        is_new = false;
        const double ovf = e.s.design_density.normalized_overflow;
        const double score = e.s.hpwl;
        if (ovf <= o_.maximum_bin_overflow &&
            (!feasible_.found || score < feasible_.score - 1e-10 ||
             (std::abs(score - feasible_.score) <= 1e-10 && e.s.hpwl < feasible_.hpwl))) {
            pending_ck_ = &CKf_;
            feasible_ = {iteration, e.s.hpwl, ovf, score, true};
            is_new = true;
        }
        if (!feasible_.found && (!fallback_.found || ovf < fallback_.overflow - 1e-12 ||
                                 (std::abs(ovf - fallback_.overflow) <= 1e-12 && e.s.hpwl < fallback_.hpwl))) {
            pending_ck_ = &CKb_;
            fallback_ = {iteration, e.s.hpwl, ovf, score, true};
            is_new = true;
        }
    }
    mp_gp_history_row row(int it, const EvalRec& e, double penalty, double step, double disp, double gnorm,
                          double curvature, int backtracks, bool restarted, bool accepted, bool best) const {
        mp_gp_history_row r{};
        r.iteration = it;
        r.hpwl = e.s.hpwl;
        r.overflow = e.s.optimizer_density.normalized_overflow;
        r.design_overflow = e.s.design_density.normalized_overflow;
        r.smooth_wirelength = e.s.smooth_wirelength;
        r.density_energy = e.s.electrostatic_energy;
        r.objective = e.objective;
        r.penalty = penalty;
        r.smoothing = smoothing_;
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

    // One pass of the loop body :735-928.  Everything the first decisions need is enqueued before
    // one read-back: the gradient evaluation at X (slot 0) with the momentum alignment and the
    // curvature sums fused into its gather, and -- when momentum may be used -- the look-ahead
    // positions and their gradient evaluation (slot 1, concurrently).  Work whose result the
    // control law then does not use is discarded and not counted as an evaluation.
    //
    // The density deposit visits particles in a row order built at earlier positions; it stays
    // good (its workgroups' particles cover few bin rows) while positions stay close to where it was
    // built.  The movement is bounded from the accepted displacements: |X - S| <= dX_, |B - X| <=
    // beta * mX_ (mX_ bounds |X - X_prev|), |Y - base| <= max_disp, and the order is rebuilt when
    // the bound passes the reuse limit.  The bound only steers speed: results do not depend on the
    // order.
    Step iterate() {
        const int it = result_.completed_iterations;
        const bool try_momentum = worse_streak_ < 2 && momentum_age_ > 0;
        const double beta = std::min(o_.maximum_momentum, static_cast<double>(momentum_age_) / (momentum_age_ + 3));
        const double look_drift = try_momentum ? beta * mX_ : 0.0;
        if (it == 0) flushCheckpoint();
        if (it > 0) smoothing_ = smoothingFromOverflow(last_overflow_);
        const double cap = o_.maximum_movement_in_bins * min_bin_;
        // rebuild the row order when the positions about to be evaluated may have moved more than
        // the reuse limit from where it was built (a limit of 0 rebuilds it for every evaluation,
        // which then run one after another)
        const bool fresh = o_.density_order_reuse_rows <= 0.0;
        const bool reorder = fresh || dX_ + look_drift + cap > o_.density_order_reuse_rows * bin_height_;
        if (it > 0 && reorder) dX_ = 0.0;
        // try_momentum implies it > 0: momentum_age_ only grows on an accepted step
        const Backend::SessionOut so_cur =
            sessionOut(Rc_.get(), Gc_.get(), try_momentum ? XP_.get() : nullptr, pending_ck_ ? pending_ck_->get() : nullptr);
        const Backend::SessionOut so_look = sessionOut(Rl_.get(), Gl_.get(), nullptr, nullptr);
        // X is the candidate accepted last iteration, and that candidate's evaluation was the last one
        // enqueued: its density, potential and energy are X's (smoothing and penalty, which changed
        // since, enter neither), so X's evaluation takes them instead of computing them again
        const bool reuse = it > 0 && density_at_x_ && o_.recompute_density == 0;
        density_at_x_ = false;
        // and its wirelength gradient at this iteration's smoothing went ahead after that candidate's
        // read-back (requestWireAhead)
        const bool wire_ready = it > 0 && wire_at_x_;
        wire_at_x_ = false;
        Backend::EvalSpec specs[2] = {
            {X_.get(), &so_cur, kBankCurrent, reorder, reuse ? int{kBankCandidate} : -1, wire_ready},
            {B_.get(), &so_look, kBankLook, fresh}};
        // the chain up to the first read-back, as one graph when asked (its shape: the momentum
        // look-ahead, the row-order rebuild, the evaluations one after another, the density taken
        // over, the wirelength gone ahead)
        if (it > 0 && o_.capture_graphs)
            be_->beginCapture((try_momentum ? 1 : 0) | (reorder ? 2 : 0) | (fresh ? 4 : 0) | (reuse ? 8 : 0) |
                              (wire_ready ? 16 : 0));
        if (try_momentum) {
            be_->timedBegin(kTController);
            be_->launches().lookahead(LookaheadArgs{X_.get(), XP_.get(), beta, B_.get(), clamp_, P_});
            be_->timedEnd(kTController);
        }
        if (it > 0) {
            pending_ck_ = nullptr;
            if (try_momentum && !fresh) {
                be_->enqueueEvaluate(specs, 2, smoothing_, true);
            } else {
                be_->enqueueEvaluate(&specs[0], 1, smoothing_, true);
                if (try_momentum) be_->enqueueEvaluate(&specs[1], 1, smoothing_, true);
            }
        }
        // The first step is decided on the device (by k_propose, the law below restated) and proposed
        // and evaluated before the one read-back of the iteration; the host then replays the
        // decision from the same scalars and checks it.
        const bool early = it > 0;
        const double move_cap = o_.maximum_movement_in_bins * min_bin_;
        if (early) {
            be_->finalizePending();
            ProposeArgs pa{X_.get(), Gc_.get(), Rc_.get(), 0.0, 0.0, Y_.get(), PXn_.get(), PGn_.get(), clamp_, P_,
                           be_->partials(kRedPropose, kBankCandidate)};
            pa.decide_here = 1;
            pa.decide = DecideArgs{be_->deviceBank(kBankCurrent), be_->deviceBank(kBankLook), penalty_,
                                   o_.hpwl_growth_restart_threshold, last_step_, initial_overflow_, move_cap,
                                   try_momentum ? 1 : 0, worse_streak_ >= 2 ? 1 : 0, has_previous_ ? 1 : 0,
                                   be_->mappedBank(kBankCandidate) + kDecidedLook};
            pa.b_look = B_.get();
            pa.g_look = Gl_.get();
            pa.raw_look = Rl_.get();
            be_->timedBegin(kTController);
            be_->launches().propose(pa);
            be_->addReduction(kRedPropose, kBankCandidate, particleBlocks(P_));
            be_->timedEnd(kTController);
            const Backend::SessionOut so_cand = sessionOut(nullptr, nullptr, nullptr, nullptr);
            const Backend::EvalSpec cand_spec{Y_.get(), &so_cand, kBankCandidate, fresh};
            be_->enqueueEvaluate(&cand_spec, 1, smoothing_, false);
        }
        const bool need_read = it > 0;
        HostScalars first{};
        if (need_read) {
            requestWireAhead();
            first = co_await ScalarWait{this};
        }
        if (it > 0) current_ = record(first, true, kBankCurrent);
        if (wire_ready && first.bank(kBankCurrent)[kWireGammaMismatch] != 0.0)
            throw std::logic_error("the wirelength evaluated ahead used another smoothing than the host's");
        const EvalRec& cur = current_;

        bool restarted = worse_streak_ >= 2;
        bool momentum_used = false;
        const double2* base = X_.get();
        const double2* base_raw = Rc_.get();
        const double2* base_pre = Gc_.get();
        const EvalRec* base_eval = &cur;
        int base_bank = kBankCurrent;
        double base_drift = dX_;
        EvalRec look;
        if (try_momentum) {
            if (first.bank(kBankCurrent)[kAlignment] <= 0.0) {
                restarted = true;
            } else {
                look = record(first, true, kBankLook);
                if (look.objective > cur.objective * (1.0 + o_.hpwl_growth_restart_threshold)) {
                    restarted = true;
                } else {
                    base = B_.get();
                    base_raw = Rl_.get();
                    base_pre = Gl_.get();
                    base_eval = &look;
                    base_bank = kBankLook;
                    base_drift = dX_ + look_drift;
                    momentum_used = true;
                }
            }
        }
        if (restarted) {
            base = X_.get();
            base_raw = Rc_.get();
            base_pre = Gc_.get();
            base_eval = &cur;
            base_bank = kBankCurrent;
            base_drift = dX_;
            momentum_age_ = 0;
            worse_streak_ = 0;
            ++result_.momentum_restarts;
        }

        bool accepted = false, new_best = false;
        int backtracks = 0;
        double used_step = 0.0, used_curvature = 0.0, used_disp = 0.0;
        EvalRec acc;
        for (int pass = 0; pass < 2 && !accepted; ++pass) {
            // estimateStep :470-492 from the sums the base's gather took
            double step = std::max(last_step_, kEps), curvature = 0.0;
            if (has_previous_) {
                const double dp = first.bank(base_bank)[kCurvDp], dg = first.bank(base_bank)[kCurvDg];
                if (!(dp <= kEps || dg <= kEps)) {
                    curvature = std::sqrt(dg / dp);
                    step = clampd(1.0 / std::max(curvature, kEps), step * 0.25, step * 4.0);
                }
            }
            const double ratio = clampd(base_eval->s.optimizer_density.normalized_overflow / std::max(initial_overflow_, kEps), 0.0, 1.0);
            const double max_disp = move_cap * (0.35 + 0.65 * ratio);
            for (int attempt = 0; attempt <= o_.maximum_backtracks; ++attempt) {
                HostScalars fetched{};
                if (early && pass == 0 && attempt == 0) {
                    // proposed and evaluated already, from the device's decision: the same one
                    const double* dv = first.bank(kBankCandidate) + kDecidedLook;
                    if ((dv[0] != 0.0) != momentum_used || dv[1] != step || dv[2] != max_disp)
                        throw std::logic_error("device step decision differs from the host's");
                    fetched = first;
                } else {
                    be_->timedBegin(kTController);
                    be_->launches().propose(ProposeArgs{base, base_pre, base_raw, step, max_disp, Y_.get(), PXn_.get(),
                                                        PGn_.get(), clamp_, P_, be_->partials(kRedPropose, kBankCandidate)});
                    be_->addReduction(kRedPropose, kBankCandidate, particleBlocks(P_));
                    be_->timedEnd(kTController);
                    const Backend::SessionOut so_cand = sessionOut(nullptr, nullptr, nullptr, nullptr);
                    const Backend::EvalSpec cand_spec{Y_.get(), &so_cand, kBankCandidate, fresh};
                    be_->enqueueEvaluate(&cand_spec, 1, smoothing_, false);
                    requestWireAhead();
                    fetched = co_await ScalarWait{this};
                }
                const HostScalars& h = fetched;
                const double predicted = h.bank(kBankCandidate)[kPredicted];
                const double disp = h.bank(kBankCandidate)[kMaxDisp];
                if (predicted <= kEps) break;
                EvalRec cand = record(h, false, kBankCandidate);
                if (accepts(*base_eval, cand, predicted)) {
                    accepted = true;
                    used_step = step;
                    used_curvature = curvature;
                    used_disp = disp;
                    acc = cand;
                    break;
                }
                ++backtracks;
                ++result_.rejected_candidates;
                step *= o_.backtracking_ratio;
            }
            if (accepted || restarted || !momentum_used) break;
            restarted = true;
            ++result_.momentum_restarts;
            momentum_age_ = 0;
            worse_streak_ = 0;
            base = X_.get();
            base_raw = Rc_.get();
            base_pre = Gc_.get();
            base_eval = &cur;
            base_bank = kBankCurrent;
            base_drift = dX_;
            momentum_used = false;
        }

        const double penalty_used = penalty_;
        if (accepted) {
            last_step_ = used_step;
            momentum_age_ = restarted ? 0 : momentum_age_ + 1;
            const double growth = (acc.s.hpwl - cur.s.hpwl) / std::max(cur.s.hpwl, 1.0);
            worse_streak_ = growth > o_.hpwl_growth_restart_threshold ? worse_streak_ + 1 : 0;
            const double next = updatePenalty(penalty_, cur, acc);
            const bool changed = std::abs(std::log(next / std::max(penalty_, kEps))) > 0.01;
            penalty_ = next;
            if (changed) {
                has_previous_ = false;
            } else {  // the propose kernel staged the accepted step's base and gradient
                PX_.swapWith(PXn_);
                PG_.swapWith(PGn_);
                has_previous_ = true;
            }
            const double gnorm = base_eval->pre_norm;
            // X_prev <- X, X <- Y (velocity X - X_prev, :861-863)
            X_.swapWith(Y_);
            XP_.swapWith(Y_);
            density_at_x_ = true;  // the accepted candidate's evaluation was the last one enqueued
            wire_at_x_ = wire_ahead_;
            mX_ = used_disp + (momentum_used ? look_drift : 0.0);
            dX_ = base_drift + used_disp;
            last_overflow_ = acc.s.optimizer_density.normalized_overflow;
            considerCheckpoint(acc, it + 1, new_best);
            ++result_.accepted_iterations;
            if (acc.s.design_density.normalized_overflow <= o_.maximum_bin_overflow) ++feasible_refinement_;
            else feasible_refinement_ = 0;
            history_.push_back(row(it + 1, acc, penalty_used, used_step, used_disp, gnorm, used_curvature, backtracks,
                                   restarted, true, new_best));
        } else {
            momentum_age_ = 0;
            worse_streak_ = 0;
            if (cur.s.optimizer_density.normalized_overflow > o_.maximum_bin_overflow)
                penalty_ = clampd(penalty_ * o_.density_penalty_increase,
                                  std::max(o_.minimum_density_penalty, reference_penalty_ * 0.02), o_.maximum_density_penalty);
            last_overflow_ = cur.s.optimizer_density.normalized_overflow;
            if (cur.s.design_density.normalized_overflow <= o_.maximum_bin_overflow) ++feasible_refinement_;
            else feasible_refinement_ = 0;
            history_.push_back(row(it + 1, cur, penalty_used, 0.0, 0.0, cur.pre_norm, 0.0, backtracks, restarted, false, false));
        }
        result_.completed_iterations = it + 1;
        if ((it >= 20 && feasible_refinement_ >= o_.feasible_refinement_iterations) ||
            result_.completed_iterations >= o_.iterations)
            stopped_ = true;
    }

    std::shared_ptr<const Prepared> prep_;
    mp_gp_options o_;
    std::unique_ptr<Backend> be_;
    int P_ = 0;
    ClampArgs clamp_{};
    DeviceArray<double2> X_, XP_, B_, Y_, PX_, PG_, PXn_, PGn_, Rc_, Gc_, Rl_, Gl_, CKf_, CKb_;
    DeviceArray<double> precond_;
    std::vector<uint32_t> order_;  // device particle k is prepared particle order_[k]
    DeviceArray<double2>* pending_ck_ = nullptr;
    double min_bin_ = 0, max_bin_ = 0, bin_height_ = 1;
    double smoothing_ = 0, penalty_ = 1, reference_penalty_ = 1, initial_overflow_ = 0, last_overflow_ = 0, last_step_ = 0;
    double dX_ = 0.0, mX_ = 0.0;  // movement bounds: X from the row order's positions, X from X_prev
    int momentum_age_ = 0, worse_streak_ = 0, feasible_refinement_ = 0;
    bool has_previous_ = false, stopped_ = false;
    bool density_at_x_ = false;  // slot 0 and the candidate bank's partial rows hold X's density
    bool wire_ahead_ = false;    // the last read-back ran the wirelength ahead at Y
    bool wire_at_x_ = false;     // slot 0 and the current bank's wire rows hold X's at this smoothing
    EvalRec current_;
    Checkpoint feasible_, fallback_;
    mp_gp_result result_{};
    std::vector<mp_gp_history_row> history_;
    double setup_seconds_ = 0.0, loop_seconds_ = 0.0;
    std::unique_ptr<RecordLaunches> rec_;  // batched: the launches go to the batch driver
    std::coroutine_handle<> waiting_;      // batched: the step suspended on a wait
    const HostScalars* scalars_ = nullptr; // the scalars a wait resumes with
};


}  // namespace mpgp
