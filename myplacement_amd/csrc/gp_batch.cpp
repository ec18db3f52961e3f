// gp_batch.cpp — many placements in lockstep rounds on one device (gp_batch.hpp): the batch
// driver and mp_gp_global_place_batch.

#include "gp_batch.hpp"
#include "gp_session.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <coroutine>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <exception>
#include <future>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <utility>
#include <vector>

namespace mpgp {

void loadDeviceCodeOn(int device) noexcept {
    try {
        hipCheck(hipSetDevice(device), "select HIP device");
        hipStream_t s = nullptr;
        hipCheck(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "create stream");
        try {
            loadDeviceCode(s);
        } catch (...) {
        }
        (void)hipStreamDestroy(s);
    } catch (...) {
    }
}

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
        std::future<std::pair<double, double>> before;  // HPWL and design overflow at the start
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
        keepFreedHostMemory();  // jobs' set-up arrays reuse the pages earlier jobs touched
        hipCheck(hipSetDevice(device), "select HIP device");
        hipCheck(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "create batch stream");
        loadDeviceCode(stream);
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
                // the pool already prepares jobs side by side: only a large one splits its own passes
                t_parallel_limit = j->design.pin_count > kStagingPins ? 8 : 1;
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
            auto prep = std::make_shared<const Prepared>(std::move(probe), j.options);
            j.before = beforeMetrics(prep, j.options);
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
                std::tie(r.hpwl_before, r.overflow_before) = j.before.get();
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
