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

#include "gp_backend.hpp"
#include "gp_batch.hpp"
#include "gp_host.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <coroutine>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <exception>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <thread>
#include <tuple>
#include <utility>
#include <vector>

namespace mpgp {

extern thread_local std::string g_last_error;
mp_gp_status translate(const std::exception_ptr& e);
std::size_t allocationLimit(int device, std::uint64_t maximum_bytes, std::string& reason);

namespace {

constexpr double kEps = 1e-9;
double clampd(double v, double lo, double hi) { return std::max(lo, std::min(v, hi)); }

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
        !std::isfinite(o.density_order_reuse_rows))
        throw std::invalid_argument("Adaptive global placement options are outside their valid range.");
    if (static_cast<std::size_t>(o.bins_x) > (std::size_t{1} << 20) / static_cast<std::size_t>(o.bins_y))
        throw std::invalid_argument("Density grid exceeds the safe one-million-bin workspace limit.");
}

}  // namespace

class Session {
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
        for (std::size_t i = 0; i < P; ++i) keys[i] = zkey(i);
        stableSortByKey(order_.data(), Mv, keys.data(), 20);
        stableSortByKey(order_.data() + Mv, P - Mv, keys.data(), 20);
        std::vector<int32_t> rank(P);
        for (std::size_t k = 0; k < P; ++k) rank[order_[k]] = static_cast<int32_t>(k);
        std::vector<double> pw(P), ph(P), pq(P), precond(P);
        std::vector<int32_t> player(P), pin_particles(p.pin_particles.size());
        std::vector<double2> x(P);
        for (std::size_t k = 0; k < P; ++k) {
            const uint32_t i = order_[k];
            pw[k] = p.pw[i];
            ph[k] = p.ph[i];
            pq[k] = p.pq[i];
            player[k] = p.player[i];
            precond[k] = p.precond[i];
            x[k] = x0[i];
        }
        for (std::size_t q = 0; q < pin_particles.size(); ++q)
            pin_particles[q] = p.pin_particles[q] < 0 ? -1 : rank[p.pin_particles[q]];
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
        Backend::EvalSpec specs[2] = {{X_.get(), &so_cur, kBankCurrent, reorder},
                                      {B_.get(), &so_look, kBankLook, fresh}};
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
        const bool need_read = it > 0;
        HostScalars first{};
        if (need_read) first = co_await ScalarWait{this};
        if (it > 0) current_ = record(first, true, kBankCurrent);
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
            const double max_disp = o_.maximum_movement_in_bins * min_bin_ * (0.35 + 0.65 * ratio);
            for (int attempt = 0; attempt <= o_.maximum_backtracks; ++attempt) {
                be_->timedBegin(kTController);
                be_->launches().propose(ProposeArgs{base, base_pre, base_raw, step, max_disp, Y_.get(), PXn_.get(),
                                                    PGn_.get(), clamp_, P_, be_->partials(kRedPropose, kBankCandidate)});
                be_->addReduction(kRedPropose, kBankCandidate, particleBlocks(P_));
                be_->timedEnd(kTController);
                const Backend::SessionOut so_cand = sessionOut(nullptr, nullptr, nullptr, nullptr);
                const Backend::EvalSpec cand_spec{Y_.get(), &so_cand, kBankCandidate, fresh};
                be_->enqueueEvaluate(&cand_spec, 1, smoothing_, false);
                const HostScalars& h = co_await ScalarWait{this};
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
    EvalRec current_;
    Checkpoint feasible_, fallback_;
    mp_gp_result result_{};
    std::vector<mp_gp_history_row> history_;
    double setup_seconds_ = 0.0, loop_seconds_ = 0.0;
    std::unique_ptr<RecordLaunches> rec_;  // batched: the launches go to the batch driver
    std::coroutine_handle<> waiting_;      // batched: the step suspended on a wait
    const HostScalars* scalars_ = nullptr; // the scalars a wait resumes with
};

// ---- batch driver ----------------------------------------------------------------------------------
// Jobs are prepared (netlist preparation, device allocation and upload) and finished (download,
// result metrics, release) by a pool of host threads, so that this host work overlaps the rounds
// of the running jobs; the driver thread only runs rounds.
struct BatchDriver::Impl {
    struct Job {
        mp_design design{};
        mp_gp_options options{};
        // borrowed by the session's backend; released after the session
        DeviceArena arena;
        HostScalars* scalars = nullptr;
        HostScalars* scalars_device = nullptr;
        std::unique_ptr<Session> s;
        Step step;
        bool initialized = false, finished = false;
        std::exception_ptr error;
        double* centers = nullptr;
        double hpwl_before = 0.0, overflow_before = 0.0;
        std::chrono::steady_clock::time_point t0, t_loop;
        Callback done;
    };
    // one launch of the round: a run of entries of one kernel in the table, or, for a single
    // entry, the job's own launch from its recorded entry
    struct Launch {
        int kernel, n, blocks;
        std::size_t lds, entries, first;
        const BatchOp* single;
    };

    // A cohort of running jobs with its own table and completion words.  Two cohorts take turns on
    // the stream: while the device runs one cohort's round, the host resumes the other cohort's jobs
    // and builds and issues its next round.
    struct Cohort {
        std::vector<std::unique_ptr<Job>> jobs;
        std::vector<std::unique_ptr<Job>> joining;  // prepared, joining at the cohort's next turn
        long long particles = 0;                    // of its jobs and joining jobs
        unsigned char* host_table = nullptr;  // pinned staging of the round's table
        unsigned char* dev_table = nullptr;
        std::size_t table_bytes = 0;
        unsigned* host_flags = nullptr;  // host-mapped completion words, one per reduction job
        unsigned* dev_flags = nullptr;
        int flag_count = 0;
        unsigned seq = 0;
        bool issued = false;  // a round is issued and not yet collected
        int reductions = 0;   // of the issued round
    };

    int device;
    hipStream_t stream = nullptr;
    Cohort cohorts[2];
    int turn = 0;
    // MP_GP_BATCH_STATS set: a summary of the rounds on stderr when the driver ends
    bool stats = std::getenv("MP_GP_BATCH_STATS") != nullptr;
    long long st_rounds = 0, st_job_rounds = 0, st_launches = 0, st_reductions = 0, st_table_bytes = 0;
    double st_build = 0.0, st_issue = 0.0, st_wait = 0.0, st_advance = 0.0, st_upload = 0.0;
    long long st_prepared = 0, st_finished = 0;
    double st_prepare = 0.0, st_finish = 0.0, st_last_ready = 0.0, st_last_done = 0.0;  // guarded by mu
    const std::chrono::steady_clock::time_point t_created = std::chrono::steady_clock::now();

    // host work pool
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::unique_ptr<Job>> to_prepare, ready, to_finish;
    int in_pool = 0;  // jobs queued for or inside preparation / finishing
    bool stop = false;
    std::vector<std::thread> pool;
    std::vector<hipStream_t> pool_streams;  // one per pool thread, lent to the jobs it prepares
    // pinned upload staging of a pool thread, made when it first prepares a large job
    static constexpr std::size_t kStagingBytes = std::size_t{64} << 20;
    static constexpr int64_t kStagingPins = 200000;
    std::vector<char*> staging_blocks;
    // Device blocks the jobs' arrays are carved from, and host-mapped scalar blocks: kept and reused
    // for later jobs, so that starting and ending a job allocates and frees nothing on the device
    // (a free waits for the whole device).  Guarded by mu.
    std::multimap<std::size_t, char*> free_slabs;
    std::map<char*, std::size_t> slabs;  // every block held, in use or kept
    std::size_t slab_bytes = 0;          // their total
    std::size_t slab_cap = 0;            // kept blocks are freed while a new one would pass this
    std::vector<HostScalars*> free_scalars;
    std::vector<HostScalars*> scalar_chunks;
    std::map<HostScalars*, HostScalars*> scalar_device;  // chunk host address -> device address

    Impl(int d, std::uint64_t memory_cap) : device(d) {
        hipCheck(hipSetDevice(device), "select HIP device");
        hipCheck(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "create batch stream");
        std::size_t free_bytes = 0, total_bytes = 0;
        hipCheck(hipMemGetInfo(&free_bytes, &total_bytes), "query device memory");
        slab_cap = memory_cap ? static_cast<std::size_t>(memory_cap) : total_bytes / 5 * 4;
        const unsigned hw = std::thread::hardware_concurrency();
        const int threads = static_cast<int>(std::min(16U, std::max(2U, hw)));
        for (int k = 0; k < threads; ++k) pool.emplace_back([this] { work(); });
    }
    ~Impl() {
        {
            std::lock_guard<std::mutex> g(mu);
            stop = true;
        }
        cv.notify_all();
        for (auto& t : pool) t.join();
        if (stats && st_rounds > 0)
            std::fprintf(stderr,
                         "batch driver: %lld cohort rounds, %.1f jobs/round, %.1f launches/round, %.1f reductions/round, "
                         "%.0f table bytes/round; per round: build %.1f us, issue %.1f us, wait %.1f us, "
                         "resume %.1f us (upload call %.1f us); %lld jobs prepared (%.2f ms each, last ready at %.3f s), %lld finished "
                         "(%.2f ms each, last at %.3f s); %.3f s in rounds\n",
                         st_rounds, double(st_job_rounds) / st_rounds, double(st_launches) / st_rounds,
                         double(st_reductions) / st_rounds, double(st_table_bytes) / st_rounds,
                         1e6 * st_build / st_rounds, 1e6 * st_issue / st_rounds, 1e6 * st_wait / st_rounds,
                         1e6 * st_advance / st_rounds, 1e6 * st_upload / st_rounds, st_prepared, 1e3 * st_prepare / std::max(1LL, st_prepared),
                         st_last_ready, st_finished, 1e3 * st_finish / std::max(1LL, st_finished), st_last_done,
                         st_build + st_issue + st_wait + st_advance);
        (void)hipSetDevice(device);
        if (stream) (void)hipStreamSynchronize(stream);
        ready.clear();
        to_prepare.clear();
        for (Cohort& c : cohorts) {
            c.jobs.clear();
            c.joining.clear();
            if (c.host_table) (void)hipHostFree(c.host_table);
            if (c.dev_table) (void)hipFree(c.dev_table);
            if (c.host_flags) (void)hipHostFree(c.host_flags);
        }
        for (auto& kv : slabs) (void)hipFree(kv.first);
        for (HostScalars* p : scalar_chunks) (void)hipHostFree(p);
        for (hipStream_t t : pool_streams) (void)hipStreamDestroy(t);
        for (char* p : staging_blocks) (void)hipHostFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }

    // Block sizes come in classes 2 MiB x ceil(2^(k/4)), so that later jobs of similar size reuse them.
    static std::size_t slabClass(std::size_t bytes) {
        const std::size_t unit = std::size_t{2} << 20;
        double c = 1.0;
        while (static_cast<double>(unit) * std::ceil(c) < static_cast<double>(bytes)) c *= 1.189207115002721;  // 2^(1/4)
        return unit * static_cast<std::size_t>(std::ceil(c));
    }
    // a device block of at least `bytes`: a kept one up to twice that size, else a new one (kept
    // blocks, largest first, are freed while the total would pass the cap)
    DeviceArena takeSlab(std::size_t bytes) {
        const std::size_t size = slabClass(bytes);
        std::vector<char*> drop;
        {
            std::lock_guard<std::mutex> g(mu);
            auto it = free_slabs.lower_bound(size);
            if (it != free_slabs.end() && it->first <= 2 * size) {
                DeviceArena a{it->second, it->first, 0};
                free_slabs.erase(it);
                return a;
            }
            while (!free_slabs.empty() && slab_bytes + size > slab_cap) {
                auto last = std::prev(free_slabs.end());
                drop.push_back(last->second);
                slab_bytes -= last->first;
                slabs.erase(last->second);
                free_slabs.erase(last);
            }
            slab_bytes += size;  // counted now, so that concurrent takers see it
        }
        for (char* q : drop) (void)hipFree(q);
        char* p = nullptr;
        const hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), size);
        std::lock_guard<std::mutex> g(mu);
        if (e != hipSuccess) {
            slab_bytes -= size;
            hipCheck(e, "allocate job memory");
        }
        slabs.emplace(p, size);
        return DeviceArena{p, size, 0};
    }
    // a host-mapped scalar block: host and device addresses
    std::pair<HostScalars*, HostScalars*> takeScalars() {
        std::lock_guard<std::mutex> g(mu);
        if (free_scalars.empty()) {
            constexpr int kChunk = 256;
            HostScalars* h = nullptr;
            HostScalars* d = nullptr;
            hipCheck(hipHostMalloc(reinterpret_cast<void**>(&h), kChunk * sizeof(HostScalars),
                                   hipHostMallocMapped | hipHostMallocCoherent),
                     "allocate host-mapped scalars");
            scalar_chunks.push_back(h);
            hipCheck(hipHostGetDevicePointer(reinterpret_cast<void**>(&d), h, 0), "map scalar blocks");
            for (int k = kChunk - 1; k >= 0; --k) {
                free_scalars.push_back(h + k);
                scalar_device[h + k] = d + k;
            }
        }
        HostScalars* h = free_scalars.back();
        free_scalars.pop_back();
        return {h, scalar_device[h]};
    }
    // ends the job's session and keeps what it borrowed for later jobs
    void release(Job& j) {
        j.step = Step();
        j.s.reset();
        std::lock_guard<std::mutex> g(mu);
        if (j.arena.base) free_slabs.emplace(j.arena.size, j.arena.base);
        if (j.scalars) free_scalars.push_back(j.scalars);
        j.arena = DeviceArena{};
        j.scalars = nullptr;
    }

    static double since(std::chrono::steady_clock::time_point t) {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
    }

    // Runs the job's steps until one waits for the device (its launches are then recorded) or
    // the run has stopped.
    static void advance(Job& j) {
        try {
            if (j.s->waiting()) j.s->takeWaiting().resume();
            for (;;) {
                if (j.s->waiting()) return;
                if (j.step) {  // the step ran to its end
                    j.step.result();
                    j.step = Step();
                    if (!j.initialized) {
                        j.initialized = true;
                        j.s->addTime(since(j.t0), 0.0);
                        j.t_loop = std::chrono::steady_clock::now();
                    }
                }
                if (!j.initialized) j.step = j.s->initialize();
                else if (!j.s->stopped()) j.step = j.s->iteration();
                else {
                    j.finished = true;
                    return;
                }
            }
        } catch (...) {
            j.error = std::current_exception();
            j.finished = true;
        }
    }

    void work() {
        bool selected = hipSetDevice(device) == hipSuccess;
        hipStream_t lent = nullptr;
        if (selected && hipStreamCreateWithFlags(&lent, hipStreamNonBlocking) == hipSuccess) {
            std::lock_guard<std::mutex> g(mu);
            pool_streams.push_back(lent);
        } else {
            lent = nullptr;  // the jobs then make their own
        }
        UploadStaging staging;
        for (;;) {
            std::unique_ptr<Job> j;
            bool finishing = false;
            {
                std::unique_lock<std::mutex> g(mu);
                cv.wait(g, [&] { return stop || !to_finish.empty() || !to_prepare.empty(); });
                if (!to_finish.empty()) {
                    j = std::move(to_finish.front());
                    to_finish.pop_front();
                    finishing = true;
                } else if (!to_prepare.empty()) {
                    j = std::move(to_prepare.front());
                    to_prepare.pop_front();
                } else {
                    return;  // stopping with nothing queued
                }
            }
            if (!selected) selected = hipSetDevice(device) == hipSuccess;
            bool to_ready = false;
            const auto t0 = std::chrono::steady_clock::now();
            if (finishing) {
                complete(*j);
            } else {
                if (!staging.base && selected && j->design.pin_count > kStagingPins &&
                    hipHostMalloc(reinterpret_cast<void**>(&staging.base), kStagingBytes, hipHostMallocDefault) ==
                        hipSuccess) {
                    staging.size = kStagingBytes;
                    std::lock_guard<std::mutex> g(mu);
                    staging_blocks.push_back(staging.base);
                }
                staging.used = 0;
                t_upload_staging = staging.base ? &staging : nullptr;  // every preparation ends synchronized
                prepare(*j, lent);
                t_upload_staging = nullptr;
                if (j->s) to_ready = true;
            }
            if (!to_ready) {  // done: end it here, off the driver thread
                release(*j);
                j.reset();
            }
            const double took = since(t0), at = since(t_created);
            std::lock_guard<std::mutex> g(mu);
            if (finishing) {
                ++st_finished;
                st_finish += took;
                st_last_done = at;
            } else {
                ++st_prepared;
                st_prepare += took;
                st_last_ready = at;
            }
            if (to_ready) ready.push_back(std::move(j));
            --in_pool;
            cv.notify_all();
        }
    }

    // the host side of starting a job; a failure completes it at once
    void prepare(Job& j, hipStream_t lent) {
        try {
            if (!j.centers) throw std::invalid_argument("null centres");
            if (j.options.device != device) throw std::invalid_argument("a batch runs on one device");
            hipCheck(hipSetDevice(device), "select HIP device");
            mp_design d = j.design;
            d.center = j.centers;
            Netlist probe(d);
            if (probe.movable.empty()) {  // nothing to place
                j.done(MP_GP_OK, "", mp_gp_result{});
                return;
            }
            validateOptions(j.options);
            j.hpwl_before = probe.hpwl(probe.cx.data(), probe.cy.data());
            j.overflow_before = probe.designOverflow(probe.cx.data(), probe.cy.data(), j.options.bins_x,
                                                     j.options.bins_y, j.options.target_density);
            auto prep = std::make_shared<const Prepared>(std::move(probe), j.options);
            j.arena = takeSlab(estimateBytes(d, j.options));
            std::tie(j.scalars, j.scalars_device) = takeScalars();
            const BorrowedResources borrowed{lent, &j.arena, j.scalars, j.scalars_device};
            j.s = std::make_unique<Session>(prep, j.options, nullptr, true, &borrowed);
        } catch (...) {
            // copies a failed set-up left in flight finish before its block and staging are reused
            if (lent) (void)hipStreamSynchronize(lent);
            release(j);
            const mp_gp_status st = translate(std::current_exception());
            const std::string msg = g_last_error;
            j.done(st, msg.c_str(), mp_gp_result{});
        }
    }

    int running() {
        std::lock_guard<std::mutex> g(mu);
        std::size_t n = ready.size();
        for (const Cohort& c : cohorts) n += c.jobs.size() + c.joining.size();
        return static_cast<int>(n) + in_pool;
    }

    void complete(Job& j) {
        mp_gp_result r{};
        mp_gp_status st = MP_GP_OK;
        std::string msg;
        if (!j.error) {
            try {
                j.s->addTime(0.0, since(j.t_loop));
                j.s->detach();
                j.s->finish(j.centers, &r);
                r.hpwl_before = j.hpwl_before;
                r.overflow_before = j.overflow_before;
                r.elapsed_seconds = since(j.t0);
            } catch (...) {
                j.error = std::current_exception();
            }
        }
        if (j.error) {
            st = translate(j.error);
            msg = g_last_error;
            r = mp_gp_result{};
        }
        j.done(st, msg.c_str(), r);
    }

    // queue finished jobs for the pool (their callbacks and release)
    void handOver(std::vector<std::unique_ptr<Job>>& done) {
        if (done.empty()) return;
        {
            std::lock_guard<std::mutex> g(mu);
            for (auto& j : done) {
                to_finish.push_back(std::move(j));
                ++in_pool;
            }
        }
        done.clear();
        cv.notify_all();
    }
    // a round of the cohort failed: every job of it fails with that error
    void failCohort(Cohort& c, const std::exception_ptr& e) {
        (void)hipStreamSynchronize(stream);
        c.issued = false;
        std::vector<std::unique_ptr<Job>> gone;
        gone.swap(c.jobs);
        c.particles = 0;
        for (auto& j : c.joining) c.particles += j->s->particles();
        for (auto& j : gone) {
            j->error = e;
            j->finished = true;
        }
        handOver(gone);
    }

    static void reserve(Cohort& c, std::size_t bytes, int flags) {
        if (bytes > c.table_bytes) {
            if (c.host_table) hipCheck(hipHostFree(c.host_table), "free batch table");
            if (c.dev_table) hipCheck(hipFree(c.dev_table), "free batch table");
            c.host_table = nullptr;
            c.dev_table = nullptr;
            c.table_bytes = 0;
            const std::size_t cap = std::max<std::size_t>(bytes + bytes / 2, std::size_t{1} << 20);
            hipCheck(hipHostMalloc(reinterpret_cast<void**>(&c.host_table), cap, hipHostMallocDefault),
                     "allocate batch table");
            hipCheck(hipMalloc(reinterpret_cast<void**>(&c.dev_table), cap), "allocate batch table");
            c.table_bytes = cap;
        }
        if (flags > c.flag_count) {
            if (c.host_flags) hipCheck(hipHostFree(c.host_flags), "free completion words");
            c.host_flags = nullptr;
            c.flag_count = 0;
            const int cap = std::max(flags + flags / 2, 1024);
            hipCheck(hipHostMalloc(reinterpret_cast<void**>(&c.host_flags), cap * sizeof(unsigned),
                                   hipHostMallocMapped | hipHostMallocCoherent),
                     "allocate completion words");
            std::memset(c.host_flags, 0, cap * sizeof(unsigned));
            hipCheck(hipHostGetDevicePointer(reinterpret_cast<void**>(&c.dev_flags), c.host_flags, 0),
                     "map completion words");
            c.flag_count = cap;
        }
    }

    // One turn: collect the cohort's issued round (wait, resume its jobs), let prepared jobs join it,
    // retire its finished jobs, and issue its next round without waiting for it.
    void round() {
        try {
            turnOnce();
        } catch (...) {
            // the device or the host failed outside any one cohort's round: every running job fails
            const std::exception_ptr e = std::current_exception();
            for (Cohort& c : cohorts) {
                for (auto& j : c.joining) c.jobs.push_back(std::move(j));
                c.joining.clear();
                failCohort(c, e);
            }
        }
    }

    void turnOnce() {
        hipCheck(hipSetDevice(device), "select HIP device");
        Cohort& c = cohorts[turn];
        Cohort& other = cohorts[turn ^ 1];
        turn ^= 1;
        // prepared jobs go to the cohort with fewer particles, joining it at its turn
        {
            std::lock_guard<std::mutex> g(mu);
            for (; !ready.empty(); ready.pop_front()) {
                Cohort& to = c.particles <= other.particles ? c : other;
                to.particles += ready.front()->s->particles();
                to.joining.push_back(std::move(ready.front()));
            }
        }
        if (c.issued) {
            try {
                collect(c);
            } catch (...) {
                failCohort(c, std::current_exception());
            }
        }
        for (auto& j : c.joining) {
            advance(*j);  // its first step records its first launches
            c.jobs.push_back(std::move(j));
        }
        c.joining.clear();
        // Retire finished jobs.  Every launch they recorded ran in a round that has been collected.
        std::vector<std::unique_ptr<Job>> done;
        for (std::size_t k = 0; k < c.jobs.size();) {
            Job& j = *c.jobs[k];
            if (j.finished && (j.error || j.s->recorder()->ops.empty())) {
                c.particles -= j.s->particles();
                done.push_back(std::move(c.jobs[k]));
                c.jobs.erase(c.jobs.begin() + static_cast<std::ptrdiff_t>(k));
            } else {
                ++k;
            }
        }
        handOver(done);
        if (!c.jobs.empty()) {
            try {
                issue(c);
            } catch (...) {
                failCohort(c, std::current_exception());
            }
            return;
        }
        if (other.jobs.empty() && other.joining.empty()) {
            // nothing runs: wait until the pool has a job ready or has nothing left
            std::unique_lock<std::mutex> g(mu);
            cv.wait(g, [&] { return !ready.empty() || in_pool == 0; });
        }
    }

    // builds the cohort's table from its jobs' recorded launches and issues the round
    void issue(Cohort& c) {
        const auto t_build = std::chrono::steady_clock::now();
        // every recorded launch, keyed (pass, stage, kernel, job, position); a job's launches that go
        // back to an earlier stage start its next pass
        struct Ref {
            int pass, stage, kernel, job, op;
        };
        std::vector<Ref> refs;
        std::vector<RedJob> reds;
        for (int ji = 0; ji < static_cast<int>(c.jobs.size()); ++ji) {
            Job& j = *c.jobs[ji];
            if (j.error) continue;
            const std::vector<BatchOp>& ops = j.s->recorder()->ops;
            int pass = 0, prev = -1;
            for (int k = 0; k < static_cast<int>(ops.size()); ++k) {
                if (ops[k].stage < prev) ++pass;
                prev = ops[k].stage;
                refs.push_back(Ref{pass, ops[k].stage, ops[k].kernel, ji, k});
            }
            const RedList l = j.s->backend().takePending();
            reds.insert(reds.end(), l.job, l.job + l.n);
        }
        std::sort(refs.begin(), refs.end(), [](const Ref& a, const Ref& b) {
            if (a.pass != b.pass) return a.pass < b.pass;
            if (a.stage != b.stage) return a.stage < b.stage;
            if (a.kernel != b.kernel) return a.kernel < b.kernel;
            if (a.job != b.job) return a.job < b.job;
            return a.op < b.op;
        });
        auto align = [](std::size_t o, std::size_t a) { return (o + a - 1) / a * a; };
        std::vector<Launch> launches;
        std::size_t off = 0;
        for (std::size_t a = 0; a < refs.size();) {
            std::size_t b = a;
            while (b < refs.size() && refs[b].pass == refs[a].pass && refs[b].stage == refs[a].stage &&
                   refs[b].kernel == refs[a].kernel)
                ++b;
            const BatchOp& op0 = c.jobs[refs[a].job]->s->recorder()->ops[refs[a].op];
            Launch L{refs[a].kernel, static_cast<int>(b - a), 0, 0, 0, 0, nullptr};
            if (L.n == 1) {
                L.single = &op0;
            } else {
                off = align(off, 256);
                L.entries = off;
                off += static_cast<std::size_t>(L.n) * op0.bytes;
                off = align(off, 16);
                L.first = off;
                off += (static_cast<std::size_t>(L.n) + 1) * sizeof(int);
            }
            launches.push_back(L);
            a = b;
        }
        // up to kMaxRedJobs reductions go to the finalize as a kernel argument
        const bool red_in_table = reds.size() > static_cast<std::size_t>(kMaxRedJobs);
        off = align(off, 256);
        const std::size_t red_off = off;
        if (red_in_table) off += reds.size() * sizeof(RedJob);
        reserve(c, off, static_cast<int>(reds.size()));
        // fill the staging table (the cohort's previous round has been collected: its copy is done)
        std::size_t r = 0;
        for (Launch& L : launches) {
            if (L.single) {
                L.blocks = L.single->blocks;
                L.lds = L.single->lds;
                ++r;
                continue;
            }
            int* first = reinterpret_cast<int*>(c.host_table + L.first);
            long long total = 0;
            first[0] = 0;
            for (int e = 0; e < L.n; ++e, ++r) {
                const BatchOp& op = c.jobs[refs[r].job]->s->recorder()->ops[refs[r].op];
                std::memcpy(c.host_table + L.entries + static_cast<std::size_t>(e) * op.bytes, op.entry, op.bytes);
                total += op.blocks;
                if (total > (1LL << 30)) throw std::runtime_error("batched launch exceeds the grid limit");
                first[e + 1] = static_cast<int>(total);
                L.lds = std::max(L.lds, op.lds);
            }
            L.blocks = static_cast<int>(total);
        }
        if (red_in_table) std::memcpy(c.host_table + red_off, reds.data(), reds.size() * sizeof(RedJob));
        const auto t_issue = std::chrono::steady_clock::now();
        c.issued = true;
        c.reductions = static_cast<int>(reds.size());
        if (off > 0)
            hipCheck(hipMemcpyAsync(c.dev_table, c.host_table, off, hipMemcpyHostToDevice, stream), "upload batch table");
        const auto t_launch = std::chrono::steady_clock::now();
        for (const Launch& L : launches) {
            if (L.single)
                launchEntry(static_cast<BatchKernel>(L.kernel), L.single->entry, L.lds, stream);
            else
                launchBatched(static_cast<BatchKernel>(L.kernel), c.dev_table + L.entries,
                              reinterpret_cast<const int*>(c.dev_table + L.first), L.n, L.blocks, L.lds, stream);
        }
        for (auto& j : c.jobs) j->s->recorder()->ops.clear();
        if (c.reductions > 0) {
            if (++c.seq == 0) c.seq = 1;
            if (red_in_table) {
                launchFinalizeBatch(reinterpret_cast<const RedJob*>(c.dev_table + red_off), c.reductions, c.dev_flags,
                                    c.seq, stream);
            } else {
                RedList l{};
                std::copy(reds.begin(), reds.end(), l.job);
                l.n = c.reductions;
                l.seq = c.seq;
                l.flag = c.dev_flags;
                launchFinalize(l, stream);
            }
        }
        if (stats) {
            using D = std::chrono::duration<double>;
            ++st_rounds;
            st_job_rounds += static_cast<long long>(c.jobs.size());
            st_launches += static_cast<long long>(launches.size());
            st_reductions += c.reductions;
            st_table_bytes += static_cast<long long>(off);
            st_build += D(t_issue - t_build).count();
            st_issue += D(std::chrono::steady_clock::now() - t_issue).count();
            st_upload += D(t_launch - t_issue).count();
        }
    }

    // waits for the cohort's issued round and resumes its jobs with their scalars
    void collect(Cohort& c) {
        const auto t_wait = std::chrono::steady_clock::now();
        c.issued = false;
        const int nred = c.reductions;
        if (nred > 0) {
            auto written = [&] {
                for (int b = 0; b < nred; ++b)
                    if (__atomic_load_n(&c.host_flags[b], __ATOMIC_ACQUIRE) != c.seq) return false;
                return true;
            };
            for (long spins = 0; !written(); ++spins) {
                if ((spins & 1023) != 1023) continue;
                const hipError_t st = hipStreamQuery(stream);
                if (st == hipErrorNotReady) continue;
                hipCheck(st, "synchronize batched evaluation");
                if (!written()) throw std::runtime_error("batched finalize retired without writing its completion words");
                break;
            }
        } else {
            hipCheck(hipStreamSynchronize(stream), "synchronize batched launches");
        }
        const auto t_advance = std::chrono::steady_clock::now();
        for (auto& jp : c.jobs) {
            Job& j = *jp;
            if (j.finished) continue;
            if (j.s->waiting()) j.s->deliver(&j.s->backend().hostScalars());
            advance(j);
        }
        if (stats) {
            using D = std::chrono::duration<double>;
            st_wait += D(t_advance - t_wait).count();
            st_advance += D(std::chrono::steady_clock::now() - t_advance).count();
        }
    }
};

BatchDriver::BatchDriver(int device, std::uint64_t memory_cap) : impl_(std::make_unique<Impl>(device, memory_cap)) {}
BatchDriver::~BatchDriver() = default;
int BatchDriver::running() const { return impl_->running(); }
void BatchDriver::round() { impl_->round(); }

void BatchDriver::add(const mp_design& design, const mp_gp_options& options, double* centers, Callback done) {
    auto j = std::make_unique<Impl::Job>();
    j->t0 = std::chrono::steady_clock::now();
    j->design = design;
    j->options = options;
    j->centers = centers;
    j->done = std::move(done);
    {
        std::lock_guard<std::mutex> g(impl_->mu);
        impl_->to_prepare.push_back(std::move(j));
        ++impl_->in_pool;
    }
    impl_->cv.notify_all();
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
        const double hpwl_before = probe.hpwl(probe.cx.data(), probe.cy.data());
        const double ovf_before = probe.designOverflow(probe.cx.data(), probe.cy.data(), options->bins_x, options->bins_y,
                                                       options->target_density);
        auto prep = std::make_shared<const Prepared>(std::move(probe), *options);
        Session s(prep, *options, nullptr);
        s.step(options->iterations);
        mp_gp_result r{};
        s.finish(centers, &r);
        r.hpwl_before = hpwl_before;
        r.overflow_before = ovf_before;
        r.elapsed_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (result) *result = r;
        if (history) {
            const auto& hist = s.history();
            const int64_t n = std::min<int64_t>(history_capacity, static_cast<int64_t>(hist.size()));
            std::copy(hist.begin(), hist.begin() + n, history);
        }
    });
}

mp_gp_status mp_gp_global_place_batch(int32_t count, const mp_design* designs, const mp_gp_options* options,
                                       double* const* centers, mp_gp_result* results, mp_gp_status* statuses) {
    std::vector<mp_gp_status> st(count > 0 ? static_cast<std::size_t>(count) : 0, MP_GP_OK);
    std::vector<std::string> err(st.size());
    const mp_gp_status first = guarded([&] {
        if (count < 0 || (count > 0 && (!designs || !options || !centers)))
            throw std::invalid_argument("bad job count or null designs, options or centres");
        if (count == 0) return;
        BatchDriver driver(options[0].device);
        for (int32_t i = 0; i < count; ++i) {
            mp_gp_result* r = results ? results + i : nullptr;
            driver.add(designs[i], options[i], centers[i], [&st, &err, r, i](mp_gp_status s, const char* e, const mp_gp_result& res) {
                st[i] = s;
                err[i] = e ? e : "";
                if (r) *r = res;
            });
        }
        while (driver.running() > 0) driver.round();
    });
    if (statuses)
        for (std::size_t i = 0; i < st.size(); ++i) statuses[i] = st[i];
    if (first != MP_GP_OK) return first;
    for (std::size_t i = 0; i < st.size(); ++i)
        if (st[i] != MP_GP_OK) {
            g_last_error = "job " + std::to_string(i) + ": " + err[i];
            return st[i];
        }
    return MP_GP_OK;
}

}  // extern "C"
