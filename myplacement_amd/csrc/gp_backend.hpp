// gp_backend.hpp — the device-resident numeric backend behind the C ABI.
//
// One Backend owns every device buffer of one placement problem (the role of
// CudaPlacementBackendImpl, src/placement/cuda/CudaPlacementBackend.cu:841-1148).  Static data is
// uploaded once.  One or two evaluations (the controller's current and look-ahead positions) are
// enqueued as one chain of batched launches on one stream, each with its own work buffers; every
// reduction result lands in a small host-mapped scalar block that k_finalize writes directly.
#pragma once

#include "mp_gp_backend.h"
#include "gp_kernels.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <stdexcept>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace mpgp {

class InvalidArgument : public std::invalid_argument {
public:
    using std::invalid_argument::invalid_argument;
};

void hipCheck(hipError_t e, const char* what);

// Device memory that a job's arrays are carved from (256-byte aligned) instead of being allocated
// one by one; the owner of the block keeps it alive past every array carved from it.
struct DeviceArena {
    char* base = nullptr;
    std::size_t size = 0, used = 0;
    void* carve(std::size_t bytes) {
        const std::size_t at = (used + 255) / 256 * 256;
        if (!base || at > size || bytes > size - at) return nullptr;
        used = at + bytes;
        return base + at;
    }
};

// Allocation accounting against a byte limit (DeviceMemoryBudget, .cu:60-79); with an arena the
// arrays are carved from it while it has room.
class DeviceBudget {
public:
    explicit DeviceBudget(std::size_t limit) : limit_(limit) {}
    void reserve(std::size_t bytes, const char* what);
    std::size_t reserved() const { return reserved_; }
    void setArena(DeviceArena* a) { arena_ = a; }
    void* carve(std::size_t bytes) { return arena_ ? arena_->carve(bytes) : nullptr; }

private:
    std::size_t limit_;
    std::size_t reserved_ = 0;
    DeviceArena* arena_ = nullptr;
};

// Pinned host memory that uploads made on this thread are staged through (set by a batch driver's
// preparing thread for large jobs): the copies then run as asynchronous DMA from pinned memory
// instead of each blocking on the runtime's staging of pageable memory.  The owner resets it only
// after the stream has been synchronized past every copy from it.
struct UploadStaging {
    char* base = nullptr;
    std::size_t size = 0, used = 0;
};
extern thread_local UploadStaging* t_upload_staging;
void stagedUpload(void* dst, const void* src, std::size_t bytes, hipStream_t s, const char* what);

template <typename T>
class DeviceArray {
public:
    DeviceArray() = default;
    ~DeviceArray() {
        if (p_ && owned_) (void)hipFree(p_);
    }
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
    void allocate(std::size_t n, DeviceBudget& budget, const char* what) {
        if (n == 0) n = 1;  // keep a valid pointer for empty sets
        budget.reserve(n * sizeof(T), what);
        if (void* q = budget.carve(n * sizeof(T))) {
            p_ = static_cast<T*>(q);
            owned_ = false;
        } else {
            hipCheck(hipMalloc(reinterpret_cast<void**>(&p_), n * sizeof(T)), what);
            owned_ = true;
        }
        n_ = n;
    }
    void upload(const T* host, std::size_t n, hipStream_t s, const char* what) {
        if (n > n_) throw std::logic_error("device upload exceeds buffer");
        if (n) stagedUpload(p_, host, n * sizeof(T), s, what);
    }
    T* get() const { return p_; }
    std::size_t size() const { return n_; }
    void swapWith(DeviceArray& o) {
        std::swap(p_, o.p_);
        std::swap(n_, o.n_);
        std::swap(owned_, o.owned_);
    }

private:
    T* p_ = nullptr;
    std::size_t n_ = 0;
    bool owned_ = false;
};

// Where a backend's kernels go: launched at once on its stream (StreamLaunches), or recorded for
// launches batched over many jobs.
class Launches {
public:
    virtual ~Launches() = default;
    virtual void sort(const DensityArgs& a) = 0;
    virtual void wire(const WireArgs& a, int count) = 0;
    virtual void zero(double* p, int n) = 0;
    virtual void density(const DensityArgs& a, int count) = 0;
    // Launchers that run the density's bin pass and the wirelength as one launch: density(d, nd)
    // and wire(w, nw) become deposit(d, nd) followed by wireBins(w, nw, d, nd).
    virtual bool fusesWireBins() const { return false; }
    virtual void deposit(const DensityArgs&, int) { throw std::logic_error("this launcher does not split the density"); }
    virtual void wireBins(const WireArgs&, int, const DensityArgs&, int) {
        throw std::logic_error("this launcher does not fuse the wirelength and the bin pass");
    }
    // `product`: the product's place 0..3 in the evaluation chain (two forward, two inverse)
    virtual void gemm(const GemmBatch& g, int batch, int M, int N, int K, int lda, int ldb, int ldc, int epilogue,
                      const SpectralArgs& sp, int precision, int product) = 0;
    virtual void dct(int mode, const DctBatch& d, int batch, int rows, int cols, int lda, int ldb, int ldc, int step,
                     int precision, const SpectralArgs& sp) = 0;
    virtual void gather(const GatherArgs& a, int count) = 0;
    virtual void lookahead(const LookaheadArgs& l) = 0;
    virtual void propose(const ProposeArgs& p) = 0;
    virtual void copy(double2* dst, const double2* src, int n) = 0;
    // reductions finalized into the scalar block for a later kernel (no completion words)
    virtual void finalize(const RedList& l) = 0;
};

class StreamLaunches final : public Launches {
public:
    explicit StreamLaunches(hipStream_t s) : s_(s) {}
    void sort(const DensityArgs& a) override;
    void wire(const WireArgs& a, int count) override;
    void zero(double* p, int n) override;
    void density(const DensityArgs& a, int count) override;
    bool fusesWireBins() const override { return true; }
    void deposit(const DensityArgs& a, int count) override;
    void wireBins(const WireArgs& w, int nw, const DensityArgs& d, int nd) override;
    void gemm(const GemmBatch& g, int batch, int M, int N, int K, int lda, int ldb, int ldc, int epilogue,
              const SpectralArgs& sp, int precision, int product) override;
    void dct(int mode, const DctBatch& d, int batch, int rows, int cols, int lda, int ldb, int ldc, int step,
             int precision, const SpectralArgs& sp) override;
    void gather(const GatherArgs& a, int count) override;
    void lookahead(const LookaheadArgs& l) override;
    void propose(const ProposeArgs& p) override;
    void copy(double2* dst, const double2* src, int n) override;
    void finalize(const RedList& l) override;

private:
    hipStream_t s_;
};

// Stages of a job's launches in the order one evaluation chain issues them (sorting, wirelength,
// density, up to four spectral products, gather; the controller passes before).  A batch driver
// runs every job's launches of one stage in one batched launch; a job whose recorded launches go
// back to an earlier stage continues in a second pass, so each job's own order is kept.
enum BatchStage : int {
    kStageCopy = 0, kStageLookahead, kStagePropose, kStageSortHist, kStageSortScan, kStageSortPlace, kStageWire,
    kStageZero, kStageDeposit, kStageBins, kStageSpectral,  // + product index 0..3
    kStageGather = kStageSpectral + 4,
    kStageFinalize,  // reductions a later kernel of the round reads (the step decision's inputs)
    kStageCount
};
constexpr std::size_t kMaxEntryBytes =
    std::max({sizeof(WireEntry), sizeof(SortEntry), sizeof(DensityEntry), sizeof(GemmEntry), sizeof(DctEntry),
              sizeof(GatherEntry), sizeof(LookaheadEntry), sizeof(ProposeEntry), sizeof(CopyEntry), sizeof(ZeroEntry),
              sizeof(RedJob)});

// One recorded launch: the job's entry of a batched kernel.
struct BatchOp {
    int stage;        // BatchStage
    int kernel;       // BatchKernel
    int blocks;       // workgroups of the job's own launch
    int bytes;        // sizeof the entry type (the table's stride for this kernel)
    std::size_t lds;  // dynamic LDS the job's launch needs
    alignas(16) unsigned char entry[kMaxEntryBytes];
};

// Records a job's launches as batch entries instead of launching them.
class RecordLaunches final : public Launches {
public:
    void sort(const DensityArgs& a) override;
    void wire(const WireArgs& a, int count) override;
    void zero(double* p, int n) override;
    void density(const DensityArgs& a, int count) override;
    void gemm(const GemmBatch& g, int batch, int M, int N, int K, int lda, int ldb, int ldc, int epilogue,
              const SpectralArgs& sp, int precision, int product) override;
    void dct(int mode, const DctBatch& d, int batch, int rows, int cols, int lda, int ldb, int ldc, int step,
             int precision, const SpectralArgs& sp) override;
    void gather(const GatherArgs& a, int count) override;
    void lookahead(const LookaheadArgs& l) override;
    void propose(const ProposeArgs& p) override;
    void copy(double2* dst, const double2* src, int n) override;
    void finalize(const RedList& l) override;

    std::vector<BatchOp> ops;

private:
    template <typename E>
    void add(int stage, int kernel, const E& e, int blocks, std::size_t lds = 0);
};

struct HostScalars {
    double v[kScalarCount * kScalarBanks];
    unsigned done[kMaxRedJobs];  // k_finalize's completion words, one per reduction job
    const double* bank(int b) const { return v + b * kScalarCount; }
};

// What a backend borrows instead of creating its own (the jobs of a batch driver): the stream of
// its uploads and downloads, device memory its arrays are carved from, and a host-mapped scalar
// block (host and device addresses).  Each may be null.
struct BorrowedResources {
    hipStream_t stream = nullptr;
    DeviceArena* arena = nullptr;
    HostScalars* scalars = nullptr;
    HostScalars* scalars_device = nullptr;
};

class Backend {
public:
    Backend(const mp_gp_static& layout, std::size_t allocation_limit, const BorrowedResources* borrowed = nullptr);
    ~Backend();
    Backend(const Backend&) = delete;
    Backend& operator=(const Backend&) = delete;

    // Session-mode outputs of a gradient evaluation: the gather writes the combined raw /
    // preconditioned gradient and, where the pointers are set, the controller sums and a
    // checkpoint copy of the positions.
    struct SessionOut {
        const double* precond;
        double penalty, inv_strength;
        double2* raw;
        double2* pre;
        const double2* prev_pos = nullptr;
        const double2* px = nullptr;
        const double2* pg = nullptr;
        double2* ck_out = nullptr;
    };
    struct EvalSpec {
        const double2* pos;          // device positions
        const SessionOut* session;   // null: write the separate wire / density gradients
        int bank;                    // scalar bank of the results
        bool reorder;                // first rebuild the particle row order at these positions (first spec only)
        // >= 0 (first spec only): the last evaluation enqueued, with scalar bank `density_from`, was
        // at these same positions, so slot 0 still holds their density and potential and that
        // bank's metric and energy partial rows are still current: the deposit, the bin metrics and
        // the forward products are not run again (their results would be the same bits)
        int density_from = -1;
        // (first spec only) the wirelength at these positions was evaluated ahead (wireAhead) with
        // smoothing `gamma`, into slot 0 and this bank's partial rows: it is not run again, and the
        // finalize checks that the smoothing it used is `gamma`
        bool wire_ready = false;
    };
    // Enqueues `count` (1 or 2) evaluations with smoothing `gamma` as one chain of batched
    // launches.  The density deposit visits particles in the shared row order, which only has to
    // be roughly current for speed: results do not depend on it.
    void enqueueEvaluate(const EvalSpec* specs, int count, double gamma, bool want_grad);
    // Rebuilds the particle row order at `pos`.
    void sortParticles(const double2* pos);
    // Asks the next fetchScalars() to enqueue, right after its finalize and before it waits, the
    // wirelength gradient at `pos` into slot 0 and bank `bank`'s partial rows, with the smoothing
    // law of WireArgs at the optimizer overflow the finalize leaves at `overflow` (device memory):
    // the controller's next evaluation at `pos`, run while the host decides.
    struct WireAhead {
        const double2* pos;
        int bank;
        const double* overflow;
        double denominator, final_smoothing, span, scale;
    };
    void wireAhead(const WireAhead& w) { ahead_ = w; ahead_pending_ = true; }
    // Reduces the pending partials into the scalar block and waits until k_finalize has written
    // it (by polling its completion words; the rest of the kernel's retirement is not waited for).
    const HostScalars& fetchScalars();
    // Captures the launches from here to the next fetchScalars() (its finalize included) into a HIP
    // graph and launches that graph instead: one executable graph per `shape`, updated in place with
    // the new launch arguments while the chain's shape repeats.  No effect while kernel timing is on
    // or the launches are recorded for a batch.
    void beginCapture(int shape);
    // Finalizes the queued reductions into the scalar block without waiting for them (a later
    // kernel on the stream reads them: deviceBank).
    void finalizePending();
    // a scalar bank: device address of the host-mapped block, and its copy in device memory
    double* mappedBank(int bank) const { return dev_scal_ + bank * kScalarCount; }
    const double* deviceBank(int bank) const { return scal_copy_.get() + bank * kScalarCount; }
    void toScalars(const HostScalars& h, mp_gp_scalars* out, int bank = 0) const;

    // Reference-shaped blocking evaluate (CudaPlacementBackend::evaluate).
    void evaluateHost(const double* xy, double gamma, bool want_grad, mp_gp_scalars* out, double* wire_grad,
                      double* density_grad);
    void debugGrids(double* source, double* potential, double* field_x, double* field_y);

    int device() const { return device_; }
    std::size_t reservedBytes() const { return budget_.reserved(); }
    hipStream_t stream() const { return stream_; }
    int particles() const { return P_; }
    const Region& region() const { return region_; }
    DeviceBudget& budget() { return budget_; }
    const double2* sizes() const { return size_.get(); }
    // Partial rows of reduction `kind` destined for scalar bank `bank`.
    double* partials(int kind, int bank) const { return partials_[kind * kScalarBanks + bank].get(); }
    // Queues the final reduction of `blocks` partial rows; done by the next fetchScalars().  The
    // rows are those of bank `from` when it is given (>= 0), else those of `bank`.
    void addReduction(int kind, int bank, int blocks, double p0 = 0, double p1 = 0, double p2 = 0, int from = -1);

    // The kernels' destination: the backend's stream by default; a batch driver substitutes a
    // recorder (and then finalizes the pending reductions itself, see takePending).
    Launches& launches() { return *launches_; }
    void setLaunches(Launches* l) { launches_ = l ? l : own_launches_.get(); }
    // Hands over the queued reductions (for a batched finalize) and clears the queue.
    RedList takePending() {
        RedList l = pending_;
        pending_.n = 0;
        return l;
    }
    const HostScalars& hostScalars() const { return *host_scal_; }

    // Optional per-kernel-class timing with HIP events on the backend's stream; bit i of the mask
    // times class i of timingNames() (0: off).
    // every: time only every every-th launch of each enabled class (fewer events in the stream)
    void enableTiming(unsigned class_mask, int every = 1);
    std::vector<std::string> timingNames() const;
    void timing(std::vector<double>& total_us, std::vector<long long>& launches, std::vector<double>& work);
    void timedBegin(int cls);
    void timedEnd(int cls, double work = 0.0);

private:
    // work buffers of one evaluation of a batch
    struct Slot {
        DeviceArray<float2> pin_grad;
        DeviceArray<float> rho, tmp, phi, t12;
        DeviceArray<float> field;  // interleaved (Ex, Ey)
        DeviceArray<unsigned long long> grid;  // fixed-point density layers, zero between evaluations
    };
    static constexpr int kSlots = kMaxEval;
    void setDevice() const;
    DensityArgs densityArgs() const;

    int device_ = 0;
    hipStream_t stream_ = nullptr;
    std::unique_ptr<StreamLaunches> own_launches_;
    Launches* launches_ = nullptr;
    int P_ = 0, N_ = 0, I_ = 0;
    int nx_ = 0, ny_ = 0, nxp_ = 0, nyp_ = 0;
    Region region_{};
    double t_ = 0.0;
    double dark_total_ = 0.0, placeable_total_ = 0.0;
    int dct_precision_ = 0;
    bool folded_ = false;  // half-length DCT products (both grid sides multiples of 64)
    bool have_fields_ = false;
    DeviceBudget budget_;
    Slot slots_[kSlots];

    DeviceArray<double2> pos_, size_;
    DeviceArray<double> charge_;
    DeviceArray<int> layer_;
    DeviceArray<int> net_off_, pin_part_, ppin_off_, ppin_idx_, pin_dest_;
    int n_small_ = 0, n_medium_ = 0, n_large_ = 0;
    int small_net_[kSmallDegree + 2] = {}, small_pin_[kSmallDegree + 2] = {};  // small-net degree groups
    DeviceArray<double> net_w_;
    DeviceArray<double2> pin_off_;
    DeviceArray<double> static_charge_;
    DeviceArray<int> key_, hist_, colprefix_, keytotal_, keystart_, tall_ids_;
    DeviceArray<unsigned> perm_;
    DeviceArray<double2> sorted_size_;
    DeviceArray<unsigned char> tall_;
    int n_tall_ = 0;
    double fix_ = 1.0;
    DeviceArray<float> pot_;
    DeviceArray<float> bxf_, ayf_, bxi_, ayi_, syi_;
    DeviceArray<double> partials_[kRedKinds * kScalarBanks];
    RedList pending_{};
    DeviceArray<double2> wire_out_, density_out_;
    bool owns_stream_ = true, owns_scalars_ = true;
    HostScalars* host_scal_ = nullptr;  // host-mapped, written by k_finalize
    // pinned staging of evaluateHost's gradients (2 x P double2), allocated on the first gradient
    // evaluateHost call of a large design
    double2* host_stage_ = nullptr;
    double* dev_scal_ = nullptr;        // device address of host_scal_
    DeviceArray<double> scal_copy_;     // the banks' copy in device memory (k_finalize writes both)
    unsigned* dev_done_ = nullptr;      // device address of host_scal_->done
    unsigned seq_ = 0;                  // completion value of the last finalize

    // timing (events come from a pool that is reused once they have been read)
    static constexpr int kTimingClasses = 7;
    hipEvent_t poolEvent();
    unsigned timing_mask_ = 0;
    int timing_every_ = 1;
    unsigned long long t_calls_[kTimingClasses] = {};  // launches of each class since timing was enabled
    bool t_open_[kTimingClasses] = {};                 // the class's current launch is being timed
    int capture_shape_ = -1;                              // the chain being captured, or -1
    WireAhead ahead_{};
    bool ahead_pending_ = false;
    WireArgs wireArgs(double gamma, bool want_grad) const;
    std::vector<std::pair<int, hipGraphExec_t>> graphs_;  // executable graph per chain shape
    std::vector<hipEvent_t> ev_pool_;
    std::size_t ev_next_ = 0;
    std::vector<hipEvent_t> ev_begin_[kTimingClasses], ev_end_[kTimingClasses];
    std::vector<double> t_total_, t_work_;
    std::vector<double> pending_work_[kTimingClasses];
    std::vector<long long> t_count_;
};

enum TimingClass : int { kTWire = 0, kTBand = 1, kTSort = 2, kTDct = 3, kTGather = 4, kTController = 5, kTWireBins = 6 };

}  // namespace mpgp

// C handle of a backend: owns it when made by mp_gp_create, views it when taken from a session.
struct mp_gp_backend {
    std::unique_ptr<mpgp::Backend> owned;
    mpgp::Backend* impl = nullptr;
};
