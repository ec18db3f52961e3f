// gp_serve.cpp — a placement service over the device-resident adaptive run.
//
// Jobs of mixed sizes are submitted with their netlist and options; a job is admitted when one of
// `max_active_jobs` slots is free and its estimated device footprint fits the memory budget next to
// the jobs already running (first come first served, a later job that fits may pass a head job that
// does not).  In stream mode each admitted job runs to completion on its own stream, driven by its
// own worker thread, so the jobs' kernels overlap on the device.  In batched mode one driver thread
// runs all admitted jobs in lockstep rounds (one batched launch per stage over every job, see
// gp_batch.hpp) and admits waiting jobs between rounds.  Either way a job's result is the same as
// when it runs alone (every evaluation is deterministic).  The reference places one design per
// process and has no counterpart; the run itself is mp_gp_global_place (runAdaptiveGlobalPlacement,
// AdaptiveGlobalPlacer.cpp:652-953).

#include "mp_gp_backend.h"
#include "gp_batch.hpp"
#include "gp_host.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace mpgp {

extern thread_local std::string g_last_error;
mp_gp_status translate(const std::exception_ptr& e);

namespace {
int roundUp32(long long n) { return static_cast<int>((n + 31) / 32 * 32); }
}  // namespace

// Upper bound of the device bytes one run reserves (the backend's and the session's allocations,
// gp_backend.cpp / gp_session.cpp), from the design's counts.
uint64_t estimateBytes(const mp_design& d, const mp_gp_options& o) {
    long long movable = 0;
    for (int64_t m = 0; m < d.module_count; ++m) movable += (d.is_fixed[m] || d.is_terminal[m]) ? 0 : 1;
    const long long P = movable + static_cast<long long>(o.maximum_fillers);
    const long long I = d.pin_count, N = d.net_count;
    const long long G = static_cast<long long>(o.bins_x) * o.bins_y;
    const long long nxp = roundUp32(o.bins_x), nyp = roundUp32(o.bins_y), Gp = nxp * nyp;
    const long long sort_rows = (P + 1023) / 1024 * static_cast<long long>(o.bins_y);
    long long bytes = 0;
    bytes += P * (16 + 16 + 8 + 4 + 4 + 4 + 4 + 16 + 1 + 4 + 32);      // backend particle arrays
    bytes += P * (14 * 16 + 8);                                        // session vectors, preconditioner
    bytes += I * (4 + 16 + 4) + N * (4 + 8) + 8 * sort_rows;           // pins, nets, row histograms
    bytes += 2 * (I * 8 + Gp * (4 + 4 + 4 + 8 + 8) + G * 16);          // two evaluation slots
    bytes += G * 8 + Gp * 4 + nxp * nxp * 12 + nyp * nyp * 12;         // static charge, potential, factors
    bytes += 64LL << 20;                                               // partials, scalars, alignment slack
    return static_cast<uint64_t>(bytes);
}

namespace {

struct Job {
    int64_t id = 0;
    // owned copy of the design
    std::vector<double> width, height, center, net_weight, pin_offset, rows;
    std::vector<int32_t> orientation, pin_module;
    std::vector<uint8_t> is_terminal, is_fixed;
    std::vector<int64_t> net_pin_start;
    mp_design design{};
    mp_gp_options options{};
    uint64_t estimate = 0;
    double submit = 0.0, start = 0.0, finish = 0.0;
    bool done = false;
    mp_gp_status status = MP_GP_OK;
    std::string error;
    mp_gp_result result{};
};

}  // namespace

class Server {
public:
    explicit Server(const mp_gp_server_options& o) : o_(o), t0_(std::chrono::steady_clock::now()) {
        if (o_.max_active_jobs < 1 || o_.memory_budget_bytes == 0)
            throw std::invalid_argument("the placement service needs at least one slot and a memory budget");
        if (o_.mode != MP_GP_SERVE_BATCHED && o_.mode != MP_GP_SERVE_STREAMS)
            throw std::invalid_argument("unknown placement service mode");
        keepFreedHostMemory();  // jobs' set-up arrays reuse the pages earlier jobs touched
        if (o_.mode == MP_GP_SERVE_BATCHED) {
            workers_.emplace_back([this] { drive(); });  // its driver loads the device code
        } else {
            loadDeviceCodeOn(o_.device);
            for (int k = 0; k < o_.max_active_jobs; ++k) workers_.emplace_back([this] { work(); });
        }
    }
    ~Server() {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : workers_) t.join();
    }

    int64_t submit(const mp_design& d, const mp_gp_options& o) {
        auto j = std::make_unique<Job>();
        const std::size_t M = static_cast<std::size_t>(d.module_count), N = static_cast<std::size_t>(d.net_count);
        const std::size_t I = static_cast<std::size_t>(d.pin_count), R = static_cast<std::size_t>(d.row_count);
        j->width.assign(d.width, d.width + M);
        j->height.assign(d.height, d.height + M);
        j->center.assign(d.center, d.center + 2 * M);
        j->orientation.assign(d.orientation, d.orientation + M);
        j->is_terminal.assign(d.is_terminal, d.is_terminal + M);
        j->is_fixed.assign(d.is_fixed, d.is_fixed + M);
        j->net_weight.assign(d.net_weight, d.net_weight + N);
        j->net_pin_start.assign(d.net_pin_start, d.net_pin_start + N + 1);
        j->pin_module.assign(d.pin_module, d.pin_module + I);
        j->pin_offset.assign(d.pin_offset, d.pin_offset + 2 * I);
        j->rows.assign(d.rows, d.rows + 5 * R);
        j->design = mp_design{d.module_count, j->width.data(), j->height.data(), j->center.data(), j->orientation.data(),
                              j->is_terminal.data(), j->is_fixed.data(), d.net_count, j->net_weight.data(),
                              j->net_pin_start.data(), d.pin_count, j->pin_module.data(), j->pin_offset.data(),
                              d.row_count, j->rows.data()};
        j->options = o;
        j->options.device = o_.device;
        j->estimate = estimateBytes(j->design, j->options);
        if (j->estimate > o_.memory_budget_bytes)
            throw std::invalid_argument("the job's device footprint exceeds the service's whole memory budget");
        j->options.maximum_device_bytes = std::min<uint64_t>(o.maximum_device_bytes, j->estimate + (256ULL << 20));
        std::lock_guard<std::mutex> g(m_);
        if (!broken_.empty()) throw std::runtime_error("the placement service is down: " + broken_);
        j->id = next_id_++;
        j->submit = now();
        const int64_t id = j->id;
        pending_.push_back(j.get());
        jobs_[id] = std::move(j);
        cv_.notify_all();
        return id;
    }

    void wait(int64_t id, double* centers, mp_gp_job_report* report) {
        std::unique_lock<std::mutex> g(m_);
        auto it = jobs_.find(id);
        if (it == jobs_.end()) throw std::invalid_argument("unknown or already collected placement job");
        Job* j = it->second.get();
        cv_.wait(g, [&] { return j->done; });
        if (j->status != MP_GP_OK) {
            const std::string msg = j->error;
            const mp_gp_status st = j->status;
            jobs_.erase(it);
            if (st == MP_GP_INVALID_ARGUMENT) throw std::invalid_argument(msg);
            throw std::runtime_error(msg);
        }
        if (centers) std::copy(j->center.begin(), j->center.end(), centers);
        if (report) {
            *report = {};
            report->submit_seconds = j->submit;
            report->start_seconds = j->start;
            report->finish_seconds = j->finish;
            report->queue_seconds = j->start - j->submit;
            report->latency_seconds = j->finish - j->submit;
            report->estimated_device_bytes = j->estimate;
            report->result = j->result;
        }
        jobs_.erase(it);
    }

private:
    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count(); }

    // admission: first pending job (in submission order) whose footprint fits next to the running ones
    Job* admit() {
        for (auto it = pending_.begin(); it != pending_.end(); ++it) {
            if (used_ + (*it)->estimate <= o_.memory_budget_bytes) {
                Job* j = *it;
                pending_.erase(it);
                used_ += j->estimate;
                return j;
            }
        }
        return nullptr;
    }

    void work() {
        for (;;) {
            Job* j = nullptr;
            {
                std::unique_lock<std::mutex> g(m_);
                cv_.wait(g, [&] { return stop_ || (j = admit()) != nullptr; });
                if (!j) return;
                j->start = now();
            }
            mp_gp_result r{};
            const mp_gp_status st =
                mp_gp_global_place(&j->design, &j->options, j->center.data(), &r, nullptr, 0);
            const std::string err = st == MP_GP_OK ? std::string() : std::string(mp_gp_last_error());
            {
                std::lock_guard<std::mutex> g(m_);
                j->result = r;
                j->status = st;
                j->error = err;
                j->finish = now();
                j->done = true;
                used_ -= j->estimate;
            }
            cv_.notify_all();
        }
    }

    void finished(Job* j, mp_gp_status st, const char* err, const mp_gp_result& r) {
        {
            std::lock_guard<std::mutex> g(m_);
            j->result = r;
            j->status = st;
            j->error = err ? err : "";
            j->finish = now();
            j->done = true;
            used_ -= j->estimate;
            --active_;
        }
        cv_.notify_all();
    }

    // batched mode: one thread admits jobs between rounds and runs the rounds
    void drive() {
        try {
            BatchDriver driver(o_.device, o_.memory_budget_bytes);
            for (;;) {
                std::vector<Job*> start;
                {
                    std::unique_lock<std::mutex> g(m_);
                    auto take = [&] {
                        Job* j = nullptr;
                        while (!stop_ && active_ < o_.max_active_jobs && (j = admit()) != nullptr) {
                            j->start = now();
                            ++active_;
                            start.push_back(j);
                        }
                    };
                    take();
                    if (start.empty() && driver.running() == 0) {
                        cv_.wait(g, [&] {
                            take();
                            return stop_ || !start.empty();
                        });
                        if (start.empty()) return;  // stopping with nothing running
                    }
                }
                for (Job* j : start)
                    driver.add(j->design, j->options, j->center.data(),
                               [this, j](mp_gp_status st, const char* err, const mp_gp_result& r) { finished(j, st, err, r); });
                if (driver.running() > 0) driver.round();
            }
        } catch (const std::exception& e) {
            // the driver could not be set up: fail what waits and what is admitted from now on
            std::lock_guard<std::mutex> g(m_);
            broken_ = e.what();
            for (Job* j : pending_) {
                j->status = MP_GP_RUNTIME_ERROR;
                j->error = broken_;
                j->done = true;
            }
            pending_.clear();
            cv_.notify_all();
        }
    }

    mp_gp_server_options o_;
    std::chrono::steady_clock::time_point t0_;
    int active_ = 0;           // batched mode: admitted jobs not finished
    std::string broken_;       // batched mode: why the driver stopped
    std::mutex m_;
    std::condition_variable cv_;
    std::deque<Job*> pending_;
    std::map<int64_t, std::unique_ptr<Job>> jobs_;
    uint64_t used_ = 0;
    int64_t next_id_ = 1;
    bool stop_ = false;
    std::vector<std::thread> workers_;
};

}  // namespace mpgp

using namespace mpgp;

struct mp_gp_server {
    std::unique_ptr<Server> s;
};

extern "C" {

void mp_gp_default_server_options(mp_gp_server_options* o) {
    if (!o) return;
    *o = {};
    o->device = 0;
    o->max_active_jobs = 8;
    o->memory_budget_bytes = 200ULL << 30;
    o->mode = MP_GP_SERVE_STREAMS;
}

uint64_t mp_gp_estimate_device_bytes(const mp_design* design, const mp_gp_options* options) {
    if (!design || !options) return 0;
    return estimateBytes(*design, *options);
}

mp_gp_status mp_gp_server_create(const mp_gp_server_options* options, mp_gp_server** out) {
    g_last_error.clear();
    try {
        if (!options || !out) throw std::invalid_argument("null options or output");
        *out = nullptr;
        auto h = std::make_unique<mp_gp_server>();
        h->s = std::make_unique<Server>(*options);
        *out = h.release();
        return MP_GP_OK;
    } catch (...) {
        return translate(std::current_exception());
    }
}

mp_gp_status mp_gp_server_submit(mp_gp_server* server, const mp_design* design, const mp_gp_options* options,
                                 int64_t* job) {
    g_last_error.clear();
    try {
        if (!server || !design || !options || !job) throw std::invalid_argument("null server, design, options or job");
        *job = server->s->submit(*design, *options);
        return MP_GP_OK;
    } catch (...) {
        return translate(std::current_exception());
    }
}

mp_gp_status mp_gp_server_wait(mp_gp_server* server, int64_t job, double* centers, mp_gp_job_report* report) {
    g_last_error.clear();
    try {
        if (!server) throw std::invalid_argument("null server");
        server->s->wait(job, centers, report);
        return MP_GP_OK;
    } catch (...) {
        return translate(std::current_exception());
    }
}

void mp_gp_server_destroy(mp_gp_server* server) { delete server; }

}  // extern "C"
