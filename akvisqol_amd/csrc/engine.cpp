// Batched device scoring pipeline (see engine.h).
//
// Phase A (device): envelope sums, packed envelopes, FFT cross-correlation,
//   first-max lag; SPL energies of the reference and of the lag-aligned
//   degraded signal; speech VAD chunk activity.  One readback.
// Host: lag application, SPL factor, frame counts, patch indices, search
//   windows (visqol.zig:489-555).
// Phase B (device): aligned+scaled degraded signal, gammatone spectrograms,
//   dB prep, deg planes, similarity table, DP.  Readback of the DP choice,
//   then per-patch NSIM statistics (dp.zig:313-352).
// Phase C (device, unless disabled): per-match slices, envelope lags,
//   trimmed spectrograms (grid-aligned frames copied from the full
//   spectrograms), dB prep and statistics; the host keeps a realigned
//   patch unless it is strictly worse (realign.zig:294-460).
// Host: aggregation and MOS mapping (visqol.zig:595-709).
#include "engine.h"
#include "engine_state.h"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <condition_variable>
#include <map>
#include <mutex>
#include <system_error>
#include <thread>

namespace vq {

// ---------------------------------------------------------------- arenas
DeviceArena::~DeviceArena() {
    for (auto& c : chunks_) (void)hipFree(c.base);
}

void* DeviceArena::alloc(size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    if (bytes == 0) bytes = 256;
    for (auto& c : chunks_)
        if (c.size - c.used >= bytes) {
            void* p = c.base + c.used;
            c.used += bytes;
            total_used_ += bytes;
            high_water_ = std::max(high_water_, total_used_);
            return p;
        }
    size_t sz = std::max(bytes, (size_t)256 << 20);
    if (!chunks_.empty()) sz = std::max(sz, chunks_.back().size);
    char* base = nullptr;
    if (hipMalloc(&base, sz) != hipSuccess) {
        (void)hipGetLastError();
        throw BackendError{-2, "device out of memory"};
    }
    chunks_.push_back({base, sz, bytes});
    total_used_ += bytes;
    high_water_ = std::max(high_water_, total_used_);
    return base;
}

// Between batches: one chunk of the high-water size, so steady-state
// batches allocate nothing.
void DeviceArena::reset() {
    if (chunks_.size() > 1) {
        for (auto& c : chunks_) (void)hipFree(c.base);
        chunks_.clear();
        char* base = nullptr;
        const size_t sz = (high_water_ + ((size_t)1 << 20)) & ~(((size_t)1 << 20) - 1);
        if (hipMalloc(&base, sz) == hipSuccess) chunks_.push_back({base, sz, 0});
        else (void)hipGetLastError();
    }
    for (auto& c : chunks_) c.used = 0;
    total_used_ = 0;
}

PinnedArena::~PinnedArena() {
    for (auto& c : chunks_) (void)hipHostFree(c.base);
}

void* PinnedArena::alloc(size_t bytes) {
    bytes = (bytes + 63) & ~(size_t)63;
    if (bytes == 0) bytes = 64;
    for (auto& c : chunks_)
        if (c.size - c.used >= bytes) {
            void* p = c.base + c.used;
            c.used += bytes;
            return p;
        }
    size_t sz = std::max(bytes, (size_t)16 << 20);
    char* base = nullptr;
    if (hipHostMalloc(&base, sz, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        throw BackendError{-2, "pinned host memory exhausted"};
    }
    chunks_.push_back({base, sz, bytes});
    return base;
}

void PinnedArena::reset() {
    for (auto& c : chunks_) c.used = 0;
}

// Host threads for per-pair host work over n pairs, one per `per` pairs,
// at most kMaxHostThreads (the host cores a process on one GPU has).
int host_threads(size_t n, size_t per) {
    constexpr unsigned kMaxHostThreads = 16;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    if (n < 2 * per) return 1;
    return (int)std::min<size_t>({(size_t)kMaxHostThreads, (size_t)hw, n / per});
}

// ---------------------------------------------------------------- engine
Engine::Engine(const EngineOptions& o, std::unique_ptr<SvrModel> svr,
               std::unique_ptr<LatticeModel, LatticeDeleter> lattice)
    : opts_(o), svr_(std::move(svr)), lattice_(lattice.release(), LatticeDeleter()) {
    check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    if (const char* e = getenv("VISQOL_HIP_LAG_CAP")) {
        char* end = nullptr;
        const long v = strtol(e, &end, 10);
        if (end != e && *end == 0 && v >= 0 && v <= kMaxLagCands) opts_.lag_cap = (int)v;
    }
    if (const char* e = getenv("VISQOL_HIP_SCORERS")) {
        char* end = nullptr;
        const long v = strtol(e, &end, 10);
        if (end != e && *end == 0 && v >= 2 && v <= 4) opts_.scorers = (int)v;
    }
    if (const char* e = getenv("VISQOL_HIP_SCORER_STREAMS")) opts_.scorer_streams = strcmp(e, "0") != 0;
    if (const char* e = getenv("VISQOL_HIP_GT_EXCLUSIVE")) opts_.gt_exclusive = strcmp(e, "0") != 0;
    if (const char* e = getenv("VISQOL_HIP_SPEC_SUMS")) opts_.spec_sums = strcmp(e, "0") != 0;
    fence_ = std::make_shared<StreamFence>();
    fence_->streams.push_back(stream_);
    if (const char* e = getenv("VISQOL_HIP_GAMMATONE_SHAPE")) {
        char* end = nullptr;
        const long v = strtol(e, &end, 10);
        if (end != e && *end == 0 && v >= 0 && v < gammatone_shape_count()) gt_force_ = (int)v;
    }
}

Engine::Engine(HelperTag, Engine& parent)
    : gt_force_(parent.gt_force_), opts_(parent.opts_), svr_(parent.svr_), lattice_(parent.lattice_) {
    opts_.interleave_chunks = false;
    fence_ = parent.fence_;
    if (parent.opts_.scorer_streams) {
        check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
        std::lock_guard<std::mutex> lk(fence_->mu);
        fence_->streams.push_back(stream_);
    } else {
        stream_ = parent.stream_;
        owns_stream_ = false;
    }
    enqueue_mu_ = parent.enqueue_mu_;
    profiling_ = parent.profiling_;
}

void Engine::merge_profile(const Profile& p) {
    prof_.gammatone_ms += p.gammatone_ms;
    prof_.gammatone_flop += p.gammatone_flop;
    prof_.gammatone_bytes += p.gammatone_bytes;
    prof_.launches += p.launches;
    prof_.sim_ms += p.sim_ms;
    prof_.sim_cells += p.sim_cells;
    prof_.sim_launches += p.sim_launches;
    for (int i = 0; i < Profile::kPhases; i++) prof_.phase_ms[i] += p.phase_ms[i];
}

Engine::~Engine() {
    helpers_.clear();  // before the stream they share
    uploader_.reset();
    for (double* p : slots_)
        if (p) (void)hipFree(p);
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (auto e : ev_)
        if (e) (void)hipEventDestroy(e);
    for (auto& kv : rates_) {
        (void)hipFree(kv.second->coef);
        (void)hipFree(kv.second->fir);
        (void)hipFree(kv.second->hann);
    }
    for (hipStream_t cs : copy_streams_)
        if (cs) (void)hipStreamDestroy(cs);
    if (fence_) {
        std::lock_guard<std::mutex> lk(fence_->mu);
        auto& ss = fence_->streams;
        for (size_t i = 0; i < ss.size(); i++)
            if (ss[i] == stream_ && owns_stream_) {
                if (i < fence_->events.size()) {
                    if (fence_->events[i]) (void)hipEventDestroy(fence_->events[i]);
                    fence_->events.erase(fence_->events.begin() + i);
                }
                ss.erase(ss.begin() + i);
                break;
            }
        if (ss.empty())
            for (hipEvent_t e : fence_->events)
                if (e) (void)hipEventDestroy(e);
    }
    if (stream_ && owns_stream_) (void)hipStreamDestroy(stream_);
}

template <typename F>
void Engine::exclusive(F&& enqueue) {
    if (!opts_.scorer_streams || !opts_.gt_exclusive) return enqueue();
    std::lock_guard<std::mutex> lk(fence_->mu);
    auto& ss = fence_->streams;
    auto& ev = fence_->events;
    while (ev.size() < ss.size()) {
        hipEvent_t e = nullptr;
        check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
        ev.push_back(e);
    }
    size_t self = ss.size();
    for (size_t i = 0; i < ss.size(); i++) {
        if (ss[i] == stream_) {
            self = i;
            continue;
        }
        check(hipEventRecord(ev[i], ss[i]), "hipEventRecord");
        check(hipStreamWaitEvent(stream_, ev[i], 0), "hipStreamWaitEvent");
    }
    enqueue();
    if (self == ss.size()) return;
    check(hipEventRecord(ev[self], stream_), "hipEventRecord");
    for (size_t i = 0; i < ss.size(); i++)
        if (i != self) check(hipStreamWaitEvent(ss[i], ev[self], 0), "hipStreamWaitEvent");
}

void Engine::check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw BackendError{-6, std::string(what) + ": " + hipGetErrorString(e)};
    }
}

void Engine::sync() {
    check(hipGetLastError(), "kernel launch");
    check(hipStreamSynchronize(stream_), "device execution");
}

template <typename T>
T* Engine::upload(const T* v, size_t count) {
    if (count == 0) return nullptr;
    const size_t bytes = sizeof(T) * count;
    void* h = pinned_.alloc(bytes);
    memcpy(h, v, bytes);
    T* d = (T*)arena_.alloc(bytes);
    check(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, stream_), "upload");
    return d;
}

template <typename T>
T* Engine::upload(const std::vector<T>& v) {
    return upload(v.data(), v.size());
}

// the gradient pass (engine_grad.cpp) uploads these
template double* Engine::upload(const std::vector<double>&);
template int* Engine::upload(const std::vector<int>&);
template FrameGradJob* Engine::upload(const std::vector<FrameGradJob>&);
template AccRun* Engine::upload(const std::vector<AccRun>&);

void Engine::readback(void* host_dst, const void* dev_src, size_t bytes) {
    if (bytes == 0) return;
    check(hipMemcpyAsync(host_dst, dev_src, bytes, hipMemcpyDeviceToHost, stream_), "readback");
}

// Pageable host memory reaches the device at a fraction of the link's rate
// through the runtime's own staging, with the host thread blocked
// throughout.  Here up to eight host threads each copy a contiguous stripe
// of a pinned staging block in 16 MiB pieces and send every piece by DMA as
// soon as it is in, pieces alternating between this engine's two copy
// streams (two DMA engines: 56 GB/s of pinned H2D against 45 on one), so
// host copies, transfers and (on the interleaved path) the other chunk's
// kernels on the compute stream all overlap.  When the destinations are laid
// out as the staging block is (signal i at dst[0] + the sum of the padded
// lengths before it), a piece goes as one DMA; otherwise one per signal
// segment.  Signals go in groups of at most 2^28 samples (one chunk's
// worth), the staging block reused from group to group.
size_t upload_span(long long len) { return ((size_t)std::max(1LL, len) + 31) & ~(size_t)31; }

void Engine::upload_signals(int n, const double* const* src, const long long* len, double* const* dst) {
    constexpr size_t kGroup = (size_t)1 << 28;       // samples staged at once
    constexpr size_t kMaxPiece = (size_t)2 << 20;    // samples per piece: 16 MiB
    constexpr size_t kMinPiece = (size_t)128 << 10;  // 1 MiB
    std::vector<int> groups{0};
    size_t stage_len = 0, acc = 0;
    for (int i = 0; i < n; i++) {
        if (acc > 0 && acc + upload_span(len[i]) > kGroup) {
            groups.push_back(i);
            acc = 0;
        }
        acc += upload_span(len[i]);
        stage_len = std::max(stage_len, acc);
    }
    groups.push_back(n);
    if (stage_len == 0) return;
    for (hipStream_t& cs : copy_streams_)
        if (!cs) check(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "hipStreamCreate");
    double* stage = (double*)pinned_.alloc(sizeof(double) * stage_len);
    int dev = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    for (size_t g = 0; g + 1 < groups.size(); g++) {
        const int g0 = groups[g], m = groups[g + 1] - g0;
        std::vector<size_t> off(m + 1, 0);  // staging offsets in samples
        for (int i = 0; i < m; i++) off[i + 1] = off[i] + upload_span(len[g0 + i]);
        const size_t total = off[m];
        bool mirrored = true;  // dst[g0 + i] == dst[g0] + off[i]
        for (int i = 1; i < m && mirrored; i++) mirrored = dst[g0 + i] == dst[g0] + off[i];
        // pieces of about an eighth of the group (1-16 MiB, whole 256-byte
        // runs): a small upload (a pair or two) is then copied by several
        // threads, its DMAs overlapping the copies, instead of one thread's
        // copy followed by one DMA
        const size_t kPiece = std::min(kMaxPiece, std::max(kMinPiece, (total / 8 + 31) & ~(size_t)31));
        const int T = std::min(8, host_threads((total + kPiece - 1) / kPiece, 2));
        std::atomic<int> failed{(int)hipSuccess};
        auto send = [&](double* to, const double* from, size_t count, size_t piece) {
            const hipError_t r = hipMemcpyAsync(to, from, sizeof(double) * count, hipMemcpyHostToDevice,
                                                copy_streams_[piece & 1]);
            if (r != hipSuccess) failed.store((int)r);
        };
        auto stripe = [&](int t) {
            (void)hipSetDevice(dev);
            const size_t a0 = total * t / T, a1 = total * (t + 1) / T;
            if (a0 >= a1) return;
            // the signal whose padded span holds a0
            size_t i = (size_t)(std::upper_bound(off.begin(), off.end(), a0) - off.begin()) - 1;
            for (size_t a = a0; a < a1 && failed.load(std::memory_order_relaxed) == hipSuccess; a += kPiece) {
                const size_t b = std::min(a1, a + kPiece), piece = a / kPiece;
                while (off[i + 1] <= a) i++;
                size_t lo = b, hi = a;  // staged samples of this piece
                for (size_t k = i; k < (size_t)m && off[k] < b; k++) {
                    const size_t s = std::max(a, off[k]), e = std::min(b, off[k] + (size_t)std::max(0LL, len[g0 + k]));
                    if (s >= e) continue;
                    memcpy(stage + s, src[g0 + k] + (s - off[k]), sizeof(double) * (e - s));
                    if (mirrored) {
                        lo = std::min(lo, s);
                        hi = std::max(hi, e);
                    } else {
                        send(dst[g0 + k] + (s - off[k]), stage + s, e - s, piece);
                    }
                }
                // one DMA from the first staged sample to the last (padding
                // between signals goes along)
                if (mirrored && lo < hi) send(dst[g0] + lo, stage + lo, hi - lo, piece);
            }
        };
        run_striped(T, stripe);
        // drain what was enqueued before any failure (the staging block is
        // reused), then report it
        hipError_t drained = hipSuccess;
        for (hipStream_t cs : copy_streams_) {
            const hipError_t r = hipStreamSynchronize(cs);
            if (drained == hipSuccess) drained = r;
        }
        check(failed.load() != hipSuccess ? (hipError_t)failed.load() : drained, "upload");
    }
}

const RateParams& Engine::rate(uint32_t sr) { return dev_rate(sr).p; }

Engine::DevRate& Engine::dev_rate(uint32_t sr) {
    auto it = rates_.find(sr);
    if (it != rates_.end()) return *it->second;
    auto d = std::make_unique<DevRate>();
    d->p = make_rate_params(opts_.mode, sr);
    check(hipMalloc(&d->coef, sizeof(BandCoef) * d->p.nb), "hipMalloc");
    check(hipMalloc(&d->fir, sizeof(FirCoef) * d->p.nb), "hipMalloc");
    check(hipMemcpy(d->fir, d->p.fir.data(), sizeof(FirCoef) * d->p.nb, hipMemcpyHostToDevice), "upload");
    check(hipMalloc(&d->hann, sizeof(double) * std::max(1, d->p.window)), "hipMalloc");
    check(hipMemcpy(d->coef, d->p.coef.data(), sizeof(BandCoef) * d->p.nb, hipMemcpyHostToDevice), "upload");
    check(hipMemcpy(d->hann, d->p.hann.data(), sizeof(double) * d->p.window, hipMemcpyHostToDevice), "upload");
    auto& ref = *d;
    rates_[sr] = std::move(d);
    return ref;
}

namespace {

// Level twiddle tables shared by every engine of the process on one device,
// one per precision, each grown to the largest transform asked for: the
// table for n begins with the table of every smaller power of two, bit for
// bit, so one table serves every size.  A table replaced by a larger one
// stays allocated, since kernels of other engines may still read it.  The
// first transform of a new largest size pays the host generation (16 host
// threads: ~30 ms for 2^24 points) and one upload; later ones, in any
// engine, pay nothing.
class TwiddleStore {
   public:
    static TwiddleStore& of_current_device() {
        static std::mutex mu;
        static std::map<int, TwiddleStore*>* stores = new std::map<int, TwiddleStore*>();
        int dev = 0;
        (void)hipGetDevice(&dev);
        std::lock_guard<std::mutex> lk(mu);
        TwiddleStore*& s = (*stores)[dev];
        if (!s) s = new TwiddleStore();  // lives until the process ends
        return *s;
    }
    const double2* f64(long long n) {
        std::lock_guard<std::mutex> lk(mu_);
        if (n > cap64_) {
            grow_host(n);
            double2* d = nullptr;
            put(d, host_.data(), host_.size());
            d64_ = d;
            cap64_ = host_cap_;
        }
        return d64_;
    }
    const float2* f32(long long n) {
        std::lock_guard<std::mutex> lk(mu_);
        if (n > cap32_) {
            grow_host(n);
            std::vector<float> t(host_.begin(), host_.end());
            float2* d = nullptr;
            put(d, t.data(), t.size());
            d32_ = d;
            cap32_ = host_cap_;
        }
        return d32_;
    }

   private:
    static constexpr long long kMinCap = 1LL << 20;
    void grow_host(long long n) {
        if (n <= host_cap_) return;
        long long cap = kMinCap;
        while (cap < n) cap <<= 1;
        host_ = make_level_twiddles(cap, 16);
        host_cap_ = cap;
    }
    template <typename T, typename V>
    static void put(T*& d, const V* src, size_t count) {
        if (hipMalloc(&d, sizeof(V) * count) != hipSuccess) {
            (void)hipGetLastError();
            throw BackendError{-2, "device out of memory (twiddle table)"};
        }
        if (hipMemcpy(d, src, sizeof(V) * count, hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipGetLastError();
            throw BackendError{-6, "twiddle table upload failed"};
        }
    }
    std::mutex mu_;
    std::vector<double> host_;
    long long host_cap_ = 0, cap64_ = 0, cap32_ = 0;
    double2* d64_ = nullptr;
    float2* d32_ = nullptr;
};

}  // namespace

const double2* Engine::twiddles(long long n) { return TwiddleStore::of_current_device().f64(n); }

const float2* Engine::twiddles32(long long n) { return TwiddleStore::of_current_device().f32(n); }

void Engine::lap(int phase) {
    const auto now = std::chrono::steady_clock::now();
    if (profiling_ && phase >= 0)
        prof_.phase_ms[phase] += std::chrono::duration<double, std::milli>(now - lap_t_).count();
    lap_t_ = now;
}

// Reference-order sums of few long signals take the speculative form: as
// one serial chain per signal they would leave all but a handful of CUs
// idle, while a batch's many chains already fill the device.  The choice
// follows the form's time model (below two thirds of the serial chain's).
constexpr size_t kSpecMaxJobs = 64;
constexpr long long kSpecMinLen = 1 << 14;

void Engine::sums(const std::vector<SumJob>& jobs, double* d_out, bool serial) {
    if (jobs.empty()) return;
    long long max_len = 0, total_len = 0;
    for (const SumJob& j : jobs) {
        max_len = std::max(max_len, j.len);
        total_len += j.len;
    }
    if (serial && opts_.spec_sums && jobs.size() <= kSpecMaxJobs && max_len >= kSpecMinLen && spec_sums_fit(max_len) &&
        spec_sums_cost_ratio((int)jobs.size(), max_len, total_len) < 2.0 / 3.0) {
        void* scratch = arena_.alloc(spec_sums_scratch_bytes((int)jobs.size(), max_len));
        launch_sums_spec(upload(jobs), (int)jobs.size(), max_len, d_out, scratch, stream_);
        return;
    }
    double* parts =
        serial ? nullptr : (double*)arena_.alloc(sizeof(double) * kSumOutputs * kSumParts * jobs.size());
    launch_sums(upload(jobs), (int)jobs.size(), max_len, d_out, serial, parts, stream_);
}

// Lag searches grouped by transform size; pos[i] is job i's slot in the
// returned device array of lags.
long long* Engine::lag_search(const std::vector<LagJob>& jobs, const double* d_sums, std::vector<int>& pos,
                              int** d_fallback, bool exact) {
    if (d_fallback) *d_fallback = nullptr;
    pos.assign(jobs.size(), 0);
    long long* d_lags = (long long*)arena_.alloc(sizeof(long long) * std::max<size_t>(1, jobs.size()));
    if (jobs.empty()) return d_lags;
    std::vector<int> idx(jobs.size());
    for (size_t i = 0; i < idx.size(); i++) idx[i] = (int)i;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return jobs[a].n < jobs[b].n; });
    std::vector<LagJob> sorted;
    sorted.reserve(jobs.size());
    for (size_t k = 0; k < idx.size(); k++) {
        pos[idx[k]] = (int)k;
        sorted.push_back(jobs[idx[k]]);
    }
    LagJob* d_jobs = upload(sorted);
    size_t parts = 0;
    for (const LagJob& j : sorted) parts += 2 * (size_t)lag_partials_per_job(j.n);
    auto* d_parts = (unsigned long long*)arena_.alloc(sizeof(unsigned long long) * parts);
    const bool nominate = opts_.fast && !exact && d_fallback;
    int* d_fb = nullptr;
    size_t start = 0, poff = 0;
    while (start < sorted.size()) {
        size_t end = start;
        while (end < sorted.size() && sorted[end].n == sorted[start].n) end++;
        const long long n = sorted[start].n;
        const int rows = (int)(end - start);
        if (nominate && lag32_handles(n)) {
            if (!d_fb) {
                d_fb = (int*)arena_.alloc(sizeof(int) * sorted.size());
                check(hipMemsetAsync(d_fb, 0, sizeof(int) * sorted.size(), stream_), "memset");
            }
            void* scratch = arena_.alloc(lag32_scratch_bytes(rows, n));
            launch_lag_search32(d_jobs + start, rows, n, twiddles32(n), d_sums, scratch, opts_.lag_cap, d_lags + start,
                                d_fb + start, stream_);
        } else {
            launch_lag_search(d_jobs + start, rows, n, twiddles(n), d_sums, d_parts + poff, d_lags + start, stream_);
        }
        poff += 2 * (size_t)lag_partials_per_job(n) * (end - start);
        start = end;
    }
    if (d_fallback) *d_fallback = d_fb;
    return d_lags;
}

namespace {

void fail(PairResult& r, PairState& s, int code, const char* msg) {
    r.status = code;
    r.error = msg;
    s.live = false;
}

}  // namespace

void Engine::run(const PairInput* pairs, int n, uint32_t sr, PairResult* out, bool want_stages,
                 double* const* grads) {
    // Chunks bounded by samples so a batch of any size fits in HBM.  Host
    // signals scored on the interleaved path go up a chunk ahead of scoring
    // (below): there chunks of 2^26 samples start the device sooner and keep
    // the link and the device busier than 2^28 (5600 against 4200 ten-second
    // pairs/s on MI355X), while signals already on the device keep 2^28,
    // whose larger launches fill the chip better.
    bool host_side = false;
    for (int i = 0; i < n && !host_side; i++) host_side = !pairs[i].on_device;
    const bool prefetch = host_side && opts_.interleave_chunks && !want_stages && !grads;
    const long long kChunkSamples = prefetch ? 1LL << 26 : 1LL << 28;
    constexpr int kChunkPairs = 512;
    // Greedy chunks at the cap give their count; a device-resident batch
    // is then cut into that many chunks of about equal samples (a multiple
    // of the scorer count), so the scorers finish together instead of one
    // scoring a short last chunk alone (1024 ten-second pairs: four chunks
    // of 256 instead of 3 x 279 + 187, 9150-9170 against 8990-9000
    // pairs/s).  Host-array batches keep the greedy cut (measured no
    // better balanced).
    auto cut = [&](long long cap) {
        std::vector<std::pair<int, int>> c;
        int start = 0;
        while (start < n) {
            long long samples = 0;
            int end = start;
            while (end < n && (end == start || (samples + pairs[end].ref_len + pairs[end].deg_len <= cap &&
                                                end - start < kChunkPairs))) {
                samples += pairs[end].ref_len + pairs[end].deg_len;
                end++;
            }
            c.emplace_back(start, end);
            start = end;
        }
        return c;
    };
    std::vector<std::pair<int, int>> chunks = cut(kChunkSamples);
    const bool interleave = chunks.size() >= 2 && opts_.interleave_chunks && !want_stages && !grads;
    if (interleave && !prefetch) {
        const size_t nscore = (size_t)std::max(2, std::min(4, opts_.scorers));
        const size_t want = (chunks.size() + nscore - 1) / nscore * nscore;
        long long total = 0;
        for (int i = 0; i < n; i++) total += pairs[i].ref_len + pairs[i].deg_len;
        // the smallest cap (at most the chunk cap) that cuts at most `want` chunks
        long long lo = (total + (long long)want - 1) / (long long)want, hi = kChunkSamples;
        if (lo < hi && cut(hi).size() <= want) {
            while (lo < hi) {
                const long long mid = lo + (hi - lo) / 2;
                if (cut(mid).size() <= want) hi = mid;
                else lo = mid + 1;
            }
            chunks = cut(hi);
        }
    }
    if (!interleave) {
        for (const auto& c : chunks)
            run_chunk(pairs + c.first, c.second - c.first, sr, out + c.first, want_stages,
                      grads ? grads + c.first : nullptr);
        return;
    }
    // Chunk k on scorer k mod S (scorer 0 this thread, the others helper
    // engines on threads of their own), each on its own stream (or all on
    // the one in-order stream without scorer streams), so one chunk's
    // kernels run while the others' host work proceeds.
    const size_t nscore = (size_t)std::max(2, std::min(4, opts_.scorers));
    while (helpers_.size() + 1 < nscore) helpers_.emplace_back(new Engine(HelperTag{}, *this));
    for (auto& h : helpers_) h->profiling_ = profiling_;
    int dev = 0;  // the calling thread's device, made current on every thread here
    check(hipGetDevice(&dev), "hipGetDevice");

    // Host-resident signals: one more thread uploads chunk after chunk into a
    // ring of kSlots device buffers ahead of the scoring threads, so the
    // link carries the next chunks while earlier ones are scored.  Slot
    // k % kSlots holds chunk k from its upload until its scoring returns.
    const size_t kSlots = nscore + 1;
    std::vector<std::vector<PairInput>> staged(host_side ? chunks.size() : 0);
    std::mutex mu;
    std::condition_variable cv;
    std::vector<char> ready(chunks.size(), 0), busy(kSlots, 0);
    bool abort = false;
    std::exception_ptr upload_error;
    std::thread up;
    if (host_side) {
        size_t need = 0;
        for (const auto& c : chunks) {
            size_t sz = 0;
            for (int i = c.first; i < c.second; i++) {
                if (!pairs[i].on_device && !pairs[i].prep) sz += upload_span(pairs[i].ref_len);
                if (!pairs[i].on_device) sz += upload_span(pairs[i].deg_len);
            }
            need = std::max(need, sz);
        }
        if (need > slot_len_ || slots_.size() != kSlots) {
            for (double*& p : slots_) {
                if (p) (void)hipFree(p);
                p = nullptr;
            }
            slot_len_ = 0;
            slots_.assign(kSlots, nullptr);
            for (double*& p : slots_)
                if (hipMalloc(&p, sizeof(double) * need) != hipSuccess) {
                    (void)hipGetLastError();
                    throw BackendError{-2, "device out of memory"};
                }
            slot_len_ = need;
        }
        if (!uploader_) uploader_.reset(new Engine(HelperTag{}, *this));
        up = std::thread([&] {
            (void)hipSetDevice(dev);
            std::vector<const double*> src;
            std::vector<long long> len;
            std::vector<double*> dst;
            for (size_t k = 0; k < chunks.size(); k++) {
                const size_t slot = k % kSlots;
                {
                    std::unique_lock<std::mutex> lk(mu);
                    cv.wait(lk, [&] { return !busy[slot] || abort; });
                    if (abort) return;
                    busy[slot] = 1;
                }
                try {
                    src.clear();
                    len.clear();
                    dst.clear();
                    double* at = slots_[slot];
                    std::vector<PairInput>& sp = staged[k];
                    sp.assign(pairs + chunks[k].first, pairs + chunks[k].second);
                    for (PairInput& p : sp) {
                        if (p.on_device) continue;
                        if (!p.prep) {
                            src.push_back(p.ref);
                            len.push_back(p.ref_len);
                            dst.push_back(at);
                            p.ref = at;
                            at += upload_span(p.ref_len);
                        }
                        src.push_back(p.deg);
                        len.push_back(p.deg_len);
                        dst.push_back(at);
                        p.deg = at;
                        at += upload_span(p.deg_len);
                        p.on_device = true;
                    }
                    uploader_->pinned_.reset();
                    uploader_->upload_signals((int)src.size(), src.data(), len.data(), dst.data());
                } catch (...) {
                    std::lock_guard<std::mutex> lk(mu);
                    upload_error = std::current_exception();
                    abort = true;
                    cv.notify_all();
                    return;
                }
                std::lock_guard<std::mutex> lk(mu);
                ready[k] = 1;
                cv.notify_all();
            }
        });
    }
    auto score = [&](Engine& e, size_t k) {
        const int first = chunks[k].first, count = chunks[k].second - chunks[k].first;
        if (!host_side) {
            e.run_chunk(pairs + first, count, sr, out + first, false, nullptr);
            return;
        }
        {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] { return ready[k] || abort; });
            if (!ready[k]) return;  // the uploader or the other side failed and reports it
        }
        try {
            e.run_chunk(staged[k].data(), count, sr, out + first, false, nullptr);
        } catch (...) {
            std::lock_guard<std::mutex> lk(mu);
            busy[k % kSlots] = 0;
            abort = true;
            cv.notify_all();
            throw;
        }
        std::lock_guard<std::mutex> lk(mu);
        busy[k % kSlots] = 0;
        cv.notify_all();
    };
    std::vector<std::exception_ptr> helper_error(nscore - 1);
    auto helper = [&](size_t h) {
        (void)hipSetDevice(dev);
        try {
            for (size_t k = h; k < chunks.size(); k += nscore) score(*helpers_[h - 1], k);
        } catch (...) {
            helper_error[h - 1] = std::current_exception();
        }
    };
    std::vector<std::thread> ts;
    ts.reserve(nscore);  // only thread creation can throw below
    std::vector<size_t> inline_helpers;  // scorers no thread could be made for
    inline_helpers.reserve(nscore);
    for (size_t h = 1; h < nscore; h++) {
        try {
            ts.emplace_back(helper, h);
        } catch (const std::system_error&) {
            if (host_side) {
                // the upload ring needs every scorer draining its slots:
                // stop the uploader and the started scorers, fail the batch
                {
                    std::lock_guard<std::mutex> lk(mu);
                    abort = true;
                    cv.notify_all();
                }
                for (auto& t : ts) t.join();
                if (up.joinable()) up.join();
                throw;
            }
            inline_helpers.push_back(h);
        }
    }
    std::exception_ptr own_error;
    try {
        for (size_t k = 0; k < chunks.size(); k += nscore) score(*this, k);
    } catch (...) {
        own_error = std::current_exception();
    }
    for (size_t h : inline_helpers) helper(h);
    for (auto& t : ts) t.join();
    if (up.joinable()) up.join();
    if (profiling_)
        for (auto& h : helpers_) {
            merge_profile(h->prof_);
            h->prof_ = Profile();
        }
    if (own_error) std::rethrow_exception(own_error);
    for (auto& e : helper_error)
        if (e) std::rethrow_exception(e);
    if (upload_error) std::rethrow_exception(upload_error);
}

void Engine::run_chunk(const PairInput* pairs, int n, uint32_t sr, PairResult* out, bool want_stages,
                       double* const* grads) {
    arena_.reset();
    pinned_.reset();
    DevRate& dr = dev_rate(sr);
    const RateParams& rp = dr.p;
    const int nb = rp.nb, W = rp.window, hop = rp.hop, S = rp.patch;
    const bool serial = !opts_.fast;
    std::vector<PairState> st(n);
    lap(-1);

    // ------------------------------------------------------------ phase A
    std::vector<const double*> up_src;  // host signals to upload
    std::vector<long long> up_len;
    std::vector<const double**> up_to;  // where each one's device address goes
    for (int i = 0; i < n; i++) {
        out[i] = PairResult();
        PairState& s = st[i];
        s.prep = pairs[i].prep;
        s.rl = s.prep ? s.prep->len : pairs[i].ref_len;
        s.dl = pairs[i].deg_len;
        if (s.prep) {
            s.dref = s.prep->d_signal;
        } else if (pairs[i].on_device) {
            s.dref = pairs[i].ref;
        } else {
            up_src.push_back(pairs[i].ref);
            up_len.push_back(s.rl);
            up_to.push_back(&s.dref);
        }
        if (pairs[i].on_device) {
            s.ddeg = pairs[i].deg;
        } else {
            up_src.push_back(pairs[i].deg);
            up_len.push_back(s.dl);
            up_to.push_back(&s.ddeg);
        }
    }
    if (!up_src.empty()) {
        // host signals side by side in one block, laid out as their staging
        size_t span = 0;
        for (long long l : up_len) span += upload_span(l);
        double* at = (double*)arena_.alloc(sizeof(double) * span);
        std::vector<double*> up_dst(up_src.size());
        for (size_t k = 0; k < up_src.size(); k++) {
            up_dst[k] = at;
            *up_to[k] = at;
            at += upload_span(up_len[k]);
        }
        upload_signals((int)up_src.size(), up_src.data(), up_len.data(), up_dst.data());
    }
    // Slots per pair: Σref, Σdeg, Nyquist ref, Nyquist deg, ref energy,
    // aligned deg energy; a prepared reference's three come precomputed.
    bool any_prep = false;
    for (int i = 0; i < n; i++) any_prep = any_prep || st[i].prep;
    double* d_sums = nullptr;
    if (any_prep) {
        std::vector<double> init(6 * n, 0.0);
        for (int i = 0; i < n; i++)
            if (st[i].prep) {
                init[6 * i] = st[i].prep->sum;
                init[6 * i + 2] = st[i].prep->nyq;
                init[6 * i + 4] = st[i].prep->energy;
            }
        d_sums = upload(init);
    } else {
        d_sums = (double*)arena_.alloc(sizeof(double) * 6 * n);
        check(hipMemsetAsync(d_sums, 0, sizeof(double) * 6 * n, stream_), "memset");
    }
    std::vector<SumJob> j1, j2, j3;
    std::vector<LagJob> lj;
    for (int i = 0; i < n; i++) {
        PairState& s = st[i];
        const bool global = !opts_.no_global && s.rl > 0 && s.dl > 0;
        const bool pr = s.prep != nullptr;  // reference slots already filled
        if (global && opts_.fast) {
            // one pass per signal: sum, Nyquist sum, and the reference energy
            if (!pr) j1.push_back(SumJob{s.dref, s.rl, kMoments, 6 * i, 6 * i + 2, 6 * i + 4, nullptr, 0});
            j1.push_back(SumJob{s.ddeg, s.dl, kMoments, 6 * i + 1, 6 * i + 3, -1, nullptr, 0});
        } else if (global) {
            if (!pr) j1.push_back(SumJob{s.dref, s.rl, kSum, 6 * i, 0, 0, nullptr, 0});
            j1.push_back(SumJob{s.ddeg, s.dl, kSum, 6 * i + 1, 0, 0, nullptr, 0});
            if (!pr) j2.push_back(SumJob{s.dref, s.rl, kNyquist, 6 * i + 2, 6 * i, 0, nullptr, 0});
            j2.push_back(SumJob{s.ddeg, s.dl, kNyquist, 6 * i + 3, 6 * i + 1, 0, nullptr, 0});
        }
        if (global) {
            const long long nfft = fft_size_for(std::max(s.rl, s.dl));
            double2* z = (double2*)arena_.alloc(sizeof(double2) * nfft);
            s.lagjob = (int)lj.size();
            lj.push_back(LagJob{s.dref, s.ddeg, s.rl, s.dl, z, nfft, 6 * i, 6 * i + 1, 6 * i + 2, 6 * i + 3});
        }
        if (!(global && opts_.fast) && !pr) j1.push_back(SumJob{s.dref, s.rl, kSumSq, 6 * i + 4, 0, 0, nullptr, 0});
    }
    sums(j1, d_sums, serial);
    sums(j2, d_sums, serial);
    long long* d_zero = (long long*)arena_.alloc(sizeof(long long));
    check(hipMemsetAsync(d_zero, 0, sizeof(long long), stream_), "memset");
    std::vector<int> lag_pos;
    int* d_fb = nullptr;
    long long* d_lags = lag_search(lj, d_sums, lag_pos, &d_fb);
    for (int i = 0; i < n; i++) {
        PairState& s = st[i];
        const long long* lp = s.lagjob >= 0 ? d_lags + lag_pos[s.lagjob] : d_zero;
        j3.push_back(SumJob{s.ddeg, s.dl, kSumSqFrom, 6 * i + 5, 0, 0, lp, s.rl / 2});
    }
    sums(j3, d_sums, serial);
    std::vector<VadJob> vj;
    size_t vad_bytes = 0;
    if (opts_.mode == 1) {
        for (int i = 0; i < n; i++) {
            PairState& s = st[i];
            if (s.rl <= W || s.prep) continue;
            s.vp = vad_plan(num_frames(s.rl, W, hop), S, W);
            if (s.vp.n_chunks <= 0) continue;
            s.vadjob = (int)vj.size();
            s.vad_off = vad_bytes;
            vad_bytes += (size_t)s.vp.n_chunks;
            vj.push_back(VadJob{s.dref, s.rl, nullptr, s.vp.n_chunks, s.vp.frame_len, s.vp.first, 0});
        }
    }
    unsigned char* d_active = vad_bytes ? (unsigned char*)arena_.alloc(vad_bytes) : nullptr;
    if (!vj.empty()) {
        int max_chunks = 0;
        for (auto& v : vj) max_chunks = std::max(max_chunks, v.n_chunks);
        for (int i = 0; i < n; i++)
            if (st[i].vadjob >= 0) vj[st[i].vadjob].active = d_active + st[i].vad_off;
        auto* scratch = (unsigned long long*)arena_.alloc(sizeof(unsigned long long) * vj.size());
        launch_vad(upload(vj), (int)vj.size(), max_chunks, scratch, stream_);
    }
    std::vector<long long> h_lags(lj.size());
    std::vector<double> h_sums(6 * n);
    std::vector<unsigned char> h_active(vad_bytes);
    std::vector<int> h_fb(d_fb ? lj.size() : 0);
    {
        long long* pl = (long long*)pinned_.alloc(sizeof(long long) * (lj.size() + 1));
        double* ps = (double*)pinned_.alloc(sizeof(double) * 6 * n);
        unsigned char* pa = (unsigned char*)pinned_.alloc(vad_bytes + 1);
        int* pf = (int*)pinned_.alloc(sizeof(int) * (h_fb.size() + 1));
        readback(pl, d_lags, sizeof(long long) * lj.size());
        readback(ps, d_sums, sizeof(double) * 6 * n);
        readback(pa, d_active, vad_bytes);
        readback(pf, d_fb, sizeof(int) * h_fb.size());
        sync();
        memcpy(h_lags.data(), pl, sizeof(long long) * lj.size());
        memcpy(h_sums.data(), ps, sizeof(double) * 6 * n);
        memcpy(h_active.data(), pa, vad_bytes);
        memcpy(h_fb.data(), pf, sizeof(int) * h_fb.size());
    }
    // Rows the nominating search could not settle (more near-maximal lags
    // than the cap, xcorr.zig:292-307): the f64 search, then the SPL energy
    // of those pairs' aligned degraded signals again.
    if (!h_fb.empty()) {
        std::vector<LagJob> redo;
        std::vector<int> redo_pair;
        for (int i = 0; i < n; i++)
            if (st[i].lagjob >= 0 && h_fb[lag_pos[st[i].lagjob]]) {
                redo.push_back(lj[st[i].lagjob]);
                redo_pair.push_back(i);
            }
        if (!redo.empty()) {
            std::vector<int> rpos;
            long long* d_l2 = lag_search(redo, d_sums, rpos, nullptr, true);
            std::vector<SumJob> j4;
            for (size_t k = 0; k < redo.size(); k++) {
                const PairState& s = st[redo_pair[k]];
                j4.push_back(SumJob{s.ddeg, s.dl, kSumSqFrom, 6 * redo_pair[k] + 5, 0, 0, d_l2 + rpos[k], s.rl / 2});
            }
            sums(j4, d_sums, serial);
            long long* pl = (long long*)pinned_.alloc(sizeof(long long) * redo.size());
            double* ps = (double*)pinned_.alloc(sizeof(double) * 6 * n);
            readback(pl, d_l2, sizeof(long long) * redo.size());
            readback(ps, d_sums, sizeof(double) * 6 * n);
            sync();
            for (size_t k = 0; k < redo.size(); k++) {
                const int i = redo_pair[k];
                h_lags[lag_pos[st[i].lagjob]] = pl[rpos[k]];
                h_sums[6 * i + 5] = ps[6 * i + 5];
            }
        }
    }
    lap(0);

    // ------------------------------------------------------------ host plan
    for (int i = 0; i < n; i++) {
        PairState& s = st[i];
        PairResult& r = out[i];
        s.lag = s.lagjob >= 0 ? h_lags[lag_pos[s.lagjob]] : 0;
        r.lag = s.lag;
        // realign.zig:91-110 (applyLag)
        s.dl2 = s.dl;
        if (!opts_.no_global) {
            const long long a = s.lag < 0 ? -s.lag : s.lag;
            if (!(s.lag == 0 || a > s.rl / 2)) {
                s.lag_s = (double)s.lag / (double)sr;
                if (s.lag < 0) {
                    const long long k = std::min(-s.lag, s.dl);
                    s.shift = -k;
                    s.dl2 = s.dl - k;
                } else {
                    s.shift = s.lag;
                    s.dl2 = s.dl + s.lag;
                }
            }
        }
        // audio.zig:132-159 (scaleToMatchSpl)
        s.e_ref = h_sums[6 * i + 4];
        s.e_deg = h_sums[6 * i + 5];
        s.factor = spl_factor(spl_of_energy(s.e_ref, s.rl), spl_of_energy(s.e_deg, s.dl2));
        r.spl_factor = s.factor;
        r.lag_s = s.lag_s;
        if (s.rl <= W || s.dl2 <= W) { fail(r, s, -5, "SpectrogramTooShort"); continue; }
        s.nf_ref = num_frames(s.rl, W, hop);
        s.nf_deg = num_frames(s.dl2, W, hop);
        r.nf_ref = s.nf_ref;
        r.nf_deg = s.nf_deg;
        if (opts_.mode == 0) {
            if (!ref_patch_indices(s.nf_ref, S, s.idx)) { fail(r, s, -5, "ReferenceSpectrogramTooSmall"); continue; }
        } else {
            s.idx.clear();
            if (s.prep) s.idx = s.prep->vad_idx;
            else if (s.vadjob >= 0) vad_patches(s.vp, h_active.data() + s.vad_off, S, s.idx);
        }
        if (s.idx.empty()) { fail(r, s, -5, "NoPatches"); continue; }
        if (!plan_search(s.idx, s.nf_ref, s.nf_deg, S, opts_.radius, s.np, s.lo, s.hi)) {
            fail(r, s, -5, "SignalsTooDifferent");
            continue;
        }
        r.np = s.np;
        r.ref_idx.assign(s.idx.begin(), s.idx.begin() + s.np);
        s.width = s.nf_deg;
        for (int p = 0; p < s.np; p++) s.width = std::max(s.width, s.hi[p] + S);
    }

    lap(1);
    // ------------------------------------------------------------ phase B
    std::vector<ScaleJob> scj;
    std::vector<GtJob> gj;
    std::vector<FrameCopyJob> fcj;  // prepared references' raw spectrograms
    std::vector<PrepJob> pj;
    std::vector<PlaneJob> plj;
    std::vector<SimJob> sj;
    std::vector<DpJob> dj;
    int max_nf = 0, max_width = 0, max_np = 0, max_win = 0, max_span = 0;
    long long max_len = 0;
    // Chunk-wide contiguous blocks: one allocation, upload, memset and
    // readback each instead of one per pair.
    size_t spec_total = 0, cells_total = 0, np_total = 0, plane_total = 0, degs_total = 0;
    std::vector<size_t> spec_off(n), cell_off(n), np_off(n), plane_off(n), degs_off(n);
    for (int i = 0; i < n; i++) {
        PairState& s = st[i];
        if (!s.live) continue;
        spec_off[i] = spec_total;
        spec_total += (size_t)nb * (s.nf_ref + s.nf_deg);
        cell_off[i] = cells_total;
        cells_total += (size_t)s.np * s.nf_deg;
        np_off[i] = np_total;
        np_total += (size_t)s.np;
        plane_off[i] = plane_total;
        plane_total += (size_t)kPlanes * nb * s.width;
        degs_off[i] = degs_total;
        degs_total += (size_t)s.dl2;
    }
    double* d_spec = (double*)arena_.alloc(sizeof(double) * std::max<size_t>(1, spec_total));
    double* d_raw = nullptr;
    if (!opts_.no_realign || want_stages || grads)
        d_raw = (double*)arena_.alloc(sizeof(double) * std::max<size_t>(1, spec_total));
    double* d_degs = (double*)arena_.alloc(sizeof(double) * std::max<size_t>(1, degs_total));
    double* d_planes = (double*)arena_.alloc(sizeof(double) * std::max<size_t>(1, plane_total));
    double* d_sims = (double*)arena_.alloc(sizeof(double) * std::max<size_t>(1, cells_total));
    double* d_dp = (double*)arena_.alloc(sizeof(double) * std::max<size_t>(1, cells_total));
    int* d_bt = (int*)arena_.alloc(sizeof(int) * std::max<size_t>(1, cells_total));
    int* d_offs = (int*)arena_.alloc(sizeof(int) * std::max<size_t>(1, np_total));
    int* d_states = (int*)arena_.alloc(sizeof(int) * std::max<size_t>(1, np_total));
    std::vector<int> h_ilh(3 * np_total);  // idx | lo | hi
    for (int i = 0; i < n; i++) {
        PairState& s = st[i];
        if (!s.live) continue;
        for (int p = 0; p < s.np; p++) {
            h_ilh[np_off[i] + p] = s.idx[p];
            h_ilh[np_total + np_off[i] + p] = s.lo[p];
            h_ilh[2 * np_total + np_off[i] + p] = s.hi[p];
        }
    }
    int* d_ilh = upload(h_ilh);
    if (cells_total) {
        check(hipMemsetAsync(d_sims, 0, sizeof(double) * cells_total, stream_), "memset");
        check(hipMemsetAsync(d_dp, 0, sizeof(double) * cells_total, stream_), "memset");
        check(hipMemsetAsync(d_bt, 0, sizeof(int) * cells_total, stream_), "memset");
    }
    for (int i = 0; i < n; i++) {
        PairState& s = st[i];
        if (!s.live) continue;
        s.degs = d_degs + degs_off[i];
        scj.push_back(ScaleJob{s.ddeg, s.degs, s.dl, s.dl2, s.shift, s.factor});
        max_len = std::max(max_len, s.dl2);
        s.spec_ref = d_spec + spec_off[i];
        s.spec_deg = s.spec_ref + (size_t)nb * s.nf_ref;
        if (s.prep) fcj.push_back(FrameCopyJob{s.spec_ref, s.prep->d_spec, s.nf_ref, s.nf_ref});
        else gj.push_back(GtJob{s.dref, s.spec_ref, s.nf_ref, 0});
        gj.push_back(GtJob{s.degs, s.spec_deg, s.nf_deg, 0});
        max_nf = std::max(max_nf, std::max(s.nf_ref, s.nf_deg));
        if (d_raw) {
            s.raw_ref = d_raw + spec_off[i];
            s.raw_deg = s.raw_ref + (size_t)nb * s.nf_ref;
        }
        pj.push_back(PrepJob{s.spec_ref, s.spec_deg, s.nf_ref, s.nf_deg});
        s.plane = d_planes + plane_off[i];
        plj.push_back(PlaneJob{s.spec_deg, s.plane, s.nf_deg, s.width});
        max_width = std::max(max_width, s.width);
        s.sims = d_sims + cell_off[i];
        s.dp = d_dp + cell_off[i];
        s.bt = d_bt + cell_off[i];
        s.d_idx = d_ilh + np_off[i];
        s.d_lo = d_ilh + np_total + np_off[i];
        s.d_hi = d_ilh + 2 * np_total + np_off[i];
        s.d_off = d_offs + np_off[i];
        s.d_state = d_states + np_off[i];
        sj.push_back(SimJob{s.spec_ref, s.plane, s.sims, s.d_idx, s.d_lo, s.d_hi, s.np, s.nf_ref, s.nf_deg, s.width});
        dj.push_back(DpJob{s.sims, s.dp, s.bt, s.d_lo, s.d_hi, s.d_off, s.d_state, s.np, s.nf_deg});
        max_np = std::max(max_np, s.np);
        for (int p = 0; p < s.np; p++)  // the DP's scan range of row p
            max_span = std::max(max_span, s.hi[p] - (p == 0 ? s.lo[0] : std::min(s.lo[p - 1], s.lo[p])) + 1);
        // widest union of the search windows of a similarity-table patch group
        for (int p0 = 0; p0 < s.np; p0 += kSimPatchGroup) {
            int lo = s.lo[p0], hi = s.hi[p0];
            for (int p = p0 + 1; p < std::min(s.np, p0 + kSimPatchGroup); p++) {
                lo = std::min(lo, s.lo[p]);
                hi = std::max(hi, s.hi[p]);
            }
            max_win = std::max(max_win, hi - lo + 1);
        }
    }
    bool timed = false;
    if (!scj.empty()) {
        launch_scale(upload(scj), (int)scj.size(), max_len, stream_);
        GtJob* d_gj = upload(gj);
        if (profiling_ && !ev_[0])
            for (auto& e : ev_) check(hipEventCreate(&e), "hipEventCreate");
        exclusive([&] {
            // the events bracket this launch alone on a shared stream
            std::unique_lock<std::recursive_mutex> g(*enqueue_mu_, std::defer_lock);
            if (profiling_) {
                g.lock();
                check(hipEventRecord(ev_[0], stream_), "hipEventRecord");
            }
            gammatone(d_gj, (int)gj.size(), max_nf, dr, 0, &gj);
            if (profiling_) check(hipEventRecord(ev_[1], stream_), "hipEventRecord");
        });
        if (profiling_) {
            timed = true;
            for (const GtJob& g : gj) {
                prof_.gammatone_flop += 30.0 * W * nb * (double)g.nf;
                prof_.gammatone_bytes += 8.0 * ((double)(g.nf - 1) * hop + W) + 8.0 * nb * g.nf;
            }
            prof_.launches++;
        }
        if (!fcj.empty()) {
            int fmax = 0;
            for (const auto& f : fcj) fmax = std::max(fmax, f.nf);
            launch_frame_copies(upload(fcj), (int)fcj.size(), nb, fmax, stream_);
        }
        if (d_raw && spec_total)
            check(hipMemcpyAsync(d_raw, d_spec, sizeof(double) * spec_total, hipMemcpyDeviceToDevice, stream_), "copy");
        auto* prep_scratch = (unsigned long long*)arena_.alloc(sizeof(unsigned long long) * pj.size());
        launch_prep(upload(pj), (int)pj.size(), nb, max_nf, prep_scratch, stream_);
        launch_planes(upload(plj), (int)plj.size(), nb, S, max_width, opts_.fast, stream_);
        double* terms = (double*)arena_.alloc(sizeof(double) * 4 * (size_t)nb * S * max_np * sj.size());
        const SimJob* d_sj = upload(sj);
        {
            std::unique_lock<std::recursive_mutex> g(*enqueue_mu_, std::defer_lock);
            if (timed) {
                g.lock();
                check(hipEventRecord(ev_[2], stream_), "hipEventRecord");
            }
            launch_sim_table(d_sj, (int)sj.size(), nb, S, max_np, max_win, terms, opts_.fast, stream_);
            if (timed) check(hipEventRecord(ev_[3], stream_), "hipEventRecord");
        }
        if (timed) {
            for (int i = 0; i < n; i++)
                if (st[i].live)
                    for (int p = 0; p < st[i].np; p++)
                        prof_.sim_cells += (double)(st[i].hi[p] - st[i].lo[p] + 1) * nb * S;
            prof_.sim_launches++;
        }
        launch_dp(upload(dj), (int)dj.size(), max_span, stream_);
    }
    // Readback of the DP choice (and stage outputs on request).
    std::vector<std::vector<int>> h_off(n), h_state(n);
    {
        std::vector<double*> pr(n, nullptr), pd(n, nullptr), pss(n, nullptr);
        int* p_offs = (int*)pinned_.alloc(sizeof(int) * (np_total + 1));
        int* p_states = (int*)pinned_.alloc(sizeof(int) * (np_total + 1));
        readback(p_offs, d_offs, sizeof(int) * np_total);
        readback(p_states, d_states, sizeof(int) * np_total);
        for (int i = 0; i < n; i++) {
            PairState& s = st[i];
            if (!s.live) continue;
            if (want_stages) {
                pr[i] = (double*)pinned_.alloc(sizeof(double) * nb * s.nf_ref);
                pd[i] = (double*)pinned_.alloc(sizeof(double) * nb * s.nf_deg);
                pss[i] = (double*)pinned_.alloc(sizeof(double) * s.np * s.nf_deg);
                readback(pr[i], s.raw_ref, sizeof(double) * nb * s.nf_ref);
                readback(pd[i], s.raw_deg, sizeof(double) * nb * s.nf_deg);
                readback(pss[i], s.sims, sizeof(double) * s.np * s.nf_deg);
            }
        }
        sync();
        lap(2);
        if (timed) {
            float ms = 0.f;
            check(hipEventElapsedTime(&ms, ev_[0], ev_[1]), "hipEventElapsedTime");
            prof_.gammatone_ms += ms;
            check(hipEventElapsedTime(&ms, ev_[2], ev_[3]), "hipEventElapsedTime");
            prof_.sim_ms += ms;
        }
        for (int i = 0; i < n; i++) {
            if (!st[i].live) continue;
            h_off[i].assign(p_offs + np_off[i], p_offs + np_off[i] + st[i].np);
            h_state[i].assign(p_states + np_off[i], p_states + np_off[i] + st[i].np);
            st[i].match_state = h_state[i];
            if (want_stages) {
                out[i].ref_spec.assign(pr[i], pr[i] + (size_t)nb * st[i].nf_ref);
                out[i].deg_spec.assign(pd[i], pd[i] + (size_t)nb * st[i].nf_deg);
                out[i].sims.assign(pss[i], pss[i] + (size_t)st[i].np * st[i].nf_deg);
            }
        }
    }

    // Per-patch NSIM statistics of the chosen patches (dp.zig:313-352).
    std::vector<StatJob> stj;
    std::vector<std::pair<int, int>> stat_of;  // (pair, patch)
    double* d_stats = nullptr;
    const int stat_len = 3 * nb + 1;
    {
        int total = 0;
        for (int i = 0; i < n; i++)
            if (st[i].live)
                for (int p = 0; p < st[i].np; p++)
                    if (h_state[i][p] != 0) total++;
        d_stats = (double*)arena_.alloc(sizeof(double) * stat_len * std::max(1, total));
        for (int i = 0; i < n; i++) {
            PairState& s = st[i];
            if (!s.live) continue;
            for (int p = 0; p < s.np; p++) {
                if (h_state[i][p] == 0) continue;
                double* o = d_stats + (size_t)stat_len * stj.size();
                stj.push_back(StatJob{s.spec_ref, s.spec_deg, s.nf_ref, s.nf_deg, s.idx[p], h_off[i][p], s.nf_deg, S, o});
                stat_of.emplace_back(i, p);
            }
        }
        launch_stats(upload(stj), (int)stj.size(), nb, S, stream_);
        std::vector<double> h(stat_len * stj.size());
        double* ph = (double*)pinned_.alloc(sizeof(double) * (h.size() + 1));
        readback(ph, d_stats, sizeof(double) * h.size());
        sync();
        lap(3);
        memcpy(h.data(), ph, sizeof(double) * h.size());
        for (int i = 0; i < n; i++)
            if (st[i].live) {
                out[i].matches.assign(st[i].np, Match());
                out[i].dp_offset.assign(st[i].np, 0);
                out[i].match_valid.assign(st[i].np, 1);
                out[i].realign_lag.assign(st[i].np, INT64_MIN);
                out[i].realign_frames.assign(st[i].np, 0);
                out[i].realign_accepted.assign(st[i].np, 0);
            }
        const double pdur = rp.frame_duration * (double)S;
        for (size_t k = 0; k < stj.size(); k++) {
            const int i = stat_of[k].first, p = stat_of[k].second;
            Match& m = out[i].matches[p];
            const double* v = &h[k * stat_len];
            for (int b = 0; b < nb; b++) {
                m.means[b] = v[b];
                m.stddevs[b] = v[nb + b];
                m.energy[b] = v[2 * nb + b];
            }
            m.similarity = v[3 * nb];
            const int off = h_off[i][p];
            out[i].dp_offset[p] = off;
            if (h_state[i][p] == 2) {
                m.deg_start = 0.0;
                m.deg_end = 0.0;
                m.similarity = 0.0;
                for (int b = 0; b < nb; b++) m.means[b] = 0.0;
                m.valid = false;
            } else {
                m.deg_start = (double)off * rp.frame_duration;
                m.deg_end = m.deg_start + pdur;
                m.valid = true;
            }
            m.ref_start = (double)st[i].idx[p] * rp.frame_duration;
            m.ref_end = m.ref_start + pdur;
            out[i].match_valid[p] = m.valid ? 1 : 0;
        }
    }

    lap(4);
    // ------------------------------------------------------------ phase C
    if (!opts_.no_realign) {
        struct RJ {
            int pair, match;
            Layout rl, dl;
            long long rlen, dlen;
            const double *sref, *sdeg;
            int lagjob;
            Geometry geo;
            int nf = 0;
            double *spec_r = nullptr, *spec_d = nullptr;
        };
        std::vector<RJ> rjs;
        std::vector<SliceJob> slj;
        long long max_slice = 0;
        for (int i = 0; i < n; i++) {
            PairState& s = st[i];
            if (!s.live) continue;
            for (int p = 0; p < s.np; p++) {
                const Match& m = out[i].matches[p];
                if (!m.valid) continue;
                RJ rj;
                rj.pair = i;
                rj.match = p;
                rj.rl = slice_layout(s.rl, sr, m.ref_start, m.ref_end);
                rj.dl = slice_layout(s.dl2, sr, m.deg_start, m.deg_end);
                rj.rlen = rj.rl.pad + rj.rl.len;
                rj.dlen = rj.dl.pad + rj.dl.len;
                if (rj.rlen < W || rj.dlen < W) continue;
                // A slice with no padding is a window of its signal, read in
                // place; a padded one is copied with its zeros.
                auto slice = [&](const double* src, const Layout& l, long long total) -> const double* {
                    if (l.pad == 0) return src + l.base;
                    double* dst = (double*)arena_.alloc(sizeof(double) * total);
                    slj.push_back(SliceJob{src, dst, l.pad, l.base, l.len, total});
                    max_slice = std::max(max_slice, total);
                    return dst;
                };
                rj.sref = slice(s.dref, rj.rl, rj.rlen);
                rj.sdeg = slice(s.degs, rj.dl, rj.dlen);
                rjs.push_back(rj);
            }
        }
        if (!rjs.empty()) {
            launch_slices(upload(slj), (int)slj.size(), max_slice, stream_);
            double* d_rs = (double*)arena_.alloc(sizeof(double) * 4 * rjs.size());
            std::vector<SumJob> a1, a2;
            std::vector<LagJob> rlj;
            for (size_t k = 0; k < rjs.size(); k++) {
                RJ& rj = rjs[k];
                const int b = 4 * (int)k;
                if (opts_.fast) {
                    a1.push_back(SumJob{rj.sref, rj.rlen, kMoments, b, b + 2, -1, nullptr, 0});
                    a1.push_back(SumJob{rj.sdeg, rj.dlen, kMoments, b + 1, b + 3, -1, nullptr, 0});
                } else {
                    a1.push_back(SumJob{rj.sref, rj.rlen, kSum, b, 0, 0, nullptr, 0});
                    a1.push_back(SumJob{rj.sdeg, rj.dlen, kSum, b + 1, 0, 0, nullptr, 0});
                    a2.push_back(SumJob{rj.sref, rj.rlen, kNyquist, b + 2, b, 0, nullptr, 0});
                    a2.push_back(SumJob{rj.sdeg, rj.dlen, kNyquist, b + 3, b + 1, 0, nullptr, 0});
                }
                const long long nfft = fft_size_for(std::max(rj.rlen, rj.dlen));
                double2* z = (double2*)arena_.alloc(sizeof(double2) * nfft);
                rj.lagjob = (int)rlj.size();
                rlj.push_back(LagJob{rj.sref, rj.sdeg, rj.rlen, rj.dlen, z, nfft, b, b + 1, b + 2, b + 3});
            }
            sums(a1, d_rs, serial);
            sums(a2, d_rs, serial);
            std::vector<int> rpos;
            int* d_rfb = nullptr;
            long long* d_rlags = lag_search(rlj, d_rs, rpos, &d_rfb);
            std::vector<long long> h_rl(rlj.size());
            long long* pl = (long long*)pinned_.alloc(sizeof(long long) * rlj.size());
            int* pf = (int*)pinned_.alloc(sizeof(int) * (rlj.size() + 1));
            readback(pl, d_rlags, sizeof(long long) * rlj.size());
            if (d_rfb) readback(pf, d_rfb, sizeof(int) * rlj.size());
            sync();
            for (size_t q = 0; q < rlj.size(); q++) h_rl[q] = pl[rpos[q]];
            if (d_rfb) {
                // rows the nominating search could not settle: the f64 search
                std::vector<LagJob> redo;
                std::vector<int> redo_row;
                for (size_t q = 0; q < rlj.size(); q++)
                    if (pf[rpos[q]]) {
                        redo.push_back(rlj[q]);
                        redo_row.push_back((int)q);
                    }
                if (!redo.empty()) {
                    std::vector<int> p2;
                    long long* d_l2 = lag_search(redo, d_rs, p2, nullptr, true);
                    long long* pl2 = (long long*)pinned_.alloc(sizeof(long long) * redo.size());
                    readback(pl2, d_l2, sizeof(long long) * redo.size());
                    sync();
                    for (size_t k = 0; k < redo.size(); k++) h_rl[redo_row[k]] = pl2[p2[k]];
                }
            }
            lap(5);

            // Trim geometry, frame reuse, spectrograms of what is kept.
            std::vector<GtJob> rg;
            std::vector<SliceJob> tslj;
            std::vector<PrepJob> rpj;
            std::vector<StatJob> rstj;
            int rmax_nf = 0;
            long long max_t = 0;
            std::vector<FrameCopyJob> copies, raw_copies;
            for (size_t k = 0; k < rjs.size(); k++) {
                RJ& rj = rjs[k];
                PairState& s = st[rj.pair];
                const long long lag = h_rl[rj.lagjob];
                out[rj.pair].realign_lag[rj.match] = lag;
                rj.geo = trim_geometry(rj.rlen, rj.dlen, lag, sr);
                rj.nf = num_frames(rj.geo.len, W, hop);
                out[rj.pair].realign_frames[rj.match] = rj.nf;
                if (rj.nf == 0) continue;
                rj.spec_r = (double*)arena_.alloc(sizeof(double) * nb * rj.nf);
                rj.spec_d = (double*)arena_.alloc(sizeof(double) * nb * rj.nf);
                rmax_nf = std::max(rmax_nf, rj.nf);
                // realign.zig:262-281 (RawSpectrograms.reuse)
                auto reuse = [&](const Layout& l, long long off, int nf_global, long long& g0) {
                    if (off < l.pad) return false;
                    const long long start = l.base + off - l.pad;
                    if (start % hop != 0) return false;
                    g0 = start / hop;
                    return g0 + rj.nf <= nf_global;
                };
                long long g0 = 0;
                RowTrace tr;
                if (grads) {
                    tr.match = rj.match;
                    tr.nf = rj.nf;
                    tr.spec_r = rj.spec_r;
                    tr.spec_d = rj.spec_d;
                    double* raw = (double*)arena_.alloc(sizeof(double) * 2 * nb * rj.nf);
                    tr.raw_r = raw;
                    tr.raw_d = raw + (size_t)nb * rj.nf;
                    raw_copies.push_back(FrameCopyJob{raw, rj.spec_r, rj.nf, rj.nf});
                    raw_copies.push_back(FrameCopyJob{raw + (size_t)nb * rj.nf, rj.spec_d, rj.nf, rj.nf});
                    tr.first = rj.dl.base - rj.dl.pad + rj.geo.deg_off;
                    tr.valid_lo = rj.dl.base;
                    tr.valid_hi = rj.dl.base + rj.dl.len;
                }
                if (reuse(rj.rl, rj.geo.ref_off, s.nf_ref, g0)) {
                    copies.push_back(FrameCopyJob{rj.spec_r, s.raw_ref + g0, s.nf_ref, rj.nf});
                } else {
                    rg.push_back(GtJob{rj.sref + rj.geo.ref_off, rj.spec_r, rj.nf, 0});
                }
                if (reuse(rj.dl, rj.geo.deg_off, s.nf_deg, g0)) {
                    copies.push_back(FrameCopyJob{rj.spec_d, s.raw_deg + g0, s.nf_deg, rj.nf});
                    tr.deg_reused = true;
                    tr.g0 = g0;
                } else {
                    const long long off = rj.geo.deg_off;
                    if (off >= 0 && off + rj.geo.len <= rj.dlen) {
                        rg.push_back(GtJob{rj.sdeg + off, rj.spec_d, rj.nf, 0});
                    } else {
                        double* t = (double*)arena_.alloc(sizeof(double) * rj.geo.len);
                        const long long pad = off < 0 ? std::min(-off, rj.geo.len) : 0;
                        const long long base = off < 0 ? 0 : off;
                        const long long avail = std::max(0LL, std::min(rj.geo.len - pad, rj.dlen - base));
                        tslj.push_back(SliceJob{rj.sdeg, t, pad, base, avail, rj.geo.len});
                        max_t = std::max(max_t, rj.geo.len);
                        rg.push_back(GtJob{t, rj.spec_d, rj.nf, 0});
                    }
                }
                rpj.push_back(PrepJob{rj.spec_r, rj.spec_d, rj.nf, rj.nf});
                if (grads) s.rows.push_back(tr);
            }
            launch_slices(upload(tslj), (int)tslj.size(), max_t, stream_);
            gammatone(upload(rg), (int)rg.size(), rmax_nf, dr, 1);
            launch_frame_copies(upload(copies), (int)copies.size(), nb, rmax_nf, stream_);
            // the rows' raw spectrograms, kept for the gradient before the dB preparation
            launch_frame_copies(upload(raw_copies), (int)raw_copies.size(), nb, rmax_nf, stream_);
            auto* rscratch = (unsigned long long*)arena_.alloc(sizeof(unsigned long long) * std::max<size_t>(1, rpj.size()));
            launch_prep(upload(rpj), (int)rpj.size(), nb, rmax_nf, rscratch, stream_);
            double* d_rstats = (double*)arena_.alloc(sizeof(double) * stat_len * std::max<size_t>(1, rjs.size()));
            std::vector<int> rstat_of;
            for (size_t k = 0; k < rjs.size(); k++) {
                RJ& rj = rjs[k];
                if (rj.nf == 0) continue;
                rstj.push_back(StatJob{rj.spec_r, rj.spec_d, rj.nf, rj.nf, 0, 0, rj.nf, rj.nf,
                                       d_rstats + (size_t)stat_len * rstj.size()});
                rstat_of.push_back((int)k);
            }
            launch_stats(upload(rstj), (int)rstj.size(), nb, rmax_nf, stream_);
            std::vector<double> h(stat_len * rstj.size());
            double* ph = (double*)pinned_.alloc(sizeof(double) * (h.size() + 1));
            readback(ph, d_rstats, sizeof(double) * h.size());
            sync();
            lap(6);
            memcpy(h.data(), ph, sizeof(double) * h.size());
            // realign.zig:398-460 (finishRealign)
            for (size_t q = 0; q < rstj.size(); q++) {
                const RJ& rj = rjs[rstat_of[q]];
                Match& m = out[rj.pair].matches[rj.match];
                const double* v = &h[q * stat_len];
                const double mean = v[3 * nb];
                if (!(mean < m.similarity)) {
                    for (int b = 0; b < nb; b++) {
                        m.means[b] = v[b];
                        m.stddevs[b] = v[nb + b];
                        m.energy[b] = v[2 * nb + b];
                    }
                    m.similarity = mean;
                    const double new_dur = (double)rj.geo.len / (double)sr;
                    if (rj.geo.lag_seconds > 0.0) m.ref_start += rj.geo.lag_seconds;
                    else m.deg_start -= rj.geo.lag_seconds;
                    m.ref_end = m.ref_start + new_dur;
                    m.deg_end = m.deg_start + new_dur;
                    out[rj.pair].realign_accepted[rj.match] = 1;
                }
            }
        }
    }

    // Aggregation and MOS mapping per pair.  Every map is a pure function of
    // the pair's result (the lattice interpreter only reads the model's
    // constant tensors and writes its own per-call values), so larger
    // batches spread them over a few host threads.
    const int per = opts_.mode == 1 && opts_.lattice ? 1 : 8;  // pairs per thread worth a thread
    const int nthreads = host_threads(n, per);
    if (nthreads > 1) {
        run_striped(nthreads, [&](int t) {
            for (int i = t; i < n; i += nthreads)
                if (st[i].live) finish(out[i], rp);
        });
    } else {
        for (int i = 0; i < n; i++)
            if (st[i].live) finish(out[i], rp);
    }
    lap(7);
    if (grads) backward(pairs, n, sr, out, st, grads);
}

// visqol.zig:595-709 (buildResult) and visqol.zig:81-98 (mapping).
void Engine::finish(PairResult& r, const RateParams& rp) {
    const int nb = rp.nb;
    const size_t n = r.matches.size();
    const double n_f = (double)n;
    r.cf = rp.cf;
    r.fvnsim.assign(nb, 0.0);
    r.fvdegenergy.assign(nb, 0.0);
    r.fvnsim10.assign(nb, 0.0);
    r.fstdnsim.assign(nb, 0.0);
    for (const Match& m : r.matches)
        for (int b = 0; b < nb; b++) {
            r.fvnsim[b] += m.means[b];
            r.fvdegenergy[b] += m.energy[b];
        }
    for (int b = 0; b < nb; b++) {
        r.fvnsim[b] /= n_f;
        r.fvdegenergy[b] /= n_f;
    }
    const size_t k = std::max<size_t>(1, (size_t)trunc(n_f * 0.10));
    std::vector<double> col(n);
    for (int b = 0; b < nb; b++) {
        for (size_t i = 0; i < n; i++) col[i] = r.matches[i].means[b];
        std::sort(col.begin(), col.end());
        double acc = 0.0;
        for (size_t i = 0; i < k && i < n; i++) acc += col[i];
        r.fvnsim10[b] = acc / (double)k;
    }
    std::vector<double> contrib(nb, 0.0);
    long long total = 0;
    for (const Match& m : r.matches) {
        const long long fc_i = (long long)ceil((m.ref_end - m.ref_start) / rp.frame_duration);
        total += fc_i;
        const double fc = (double)fc_i;
        for (int b = 0; b < nb; b++) {
            contrib[b] += (fc - 1.0) * m.stddevs[b] * m.stddevs[b];
            contrib[b] += fc * m.means[b] * m.means[b];
        }
    }
    const double total_f = (double)total;
    for (int b = 0; b < nb; b++) {
        const double v = (contrib[b] - r.fvnsim[b] * r.fvnsim[b] * total_f) / (total_f - 1.0);
        r.fstdnsim[b] = v < 0.0 ? 0.0 : sqrt(v);
    }
    double sum = 0.0;
    for (int b = 0; b < nb; b++) sum += r.fvnsim[b];
    r.vnsim = sum / (double)nb;
    double mos = 1.0;
    if (opts_.mode == 0) {
        mos = fmax(1.0, fmin(svr_->predict(r.fvnsim.data()), 5.0));
    } else if (opts_.lattice) {
        if (!lattice_predict(*lattice_, r.fvnsim.data(), r.fvnsim10.data(), r.fstdnsim.data(), r.fvdegenergy.data(), mos)) {
            r.status = -4;
            r.error = "lattice model evaluation failed";
            return;
        }
    } else {
        double s2 = 0.0;
        for (int b = 0; b < nb; b++) s2 += r.fvnsim[b];
        const double mean = s2 / (double)nb;
        const double m = -262.847869 + exp(0.0154302525 * (mean - -361.063949));
        const double scale = opts_.unscaled_speech ? 1.0 : (double)1.245063f;
        mos = fmax(1.0, fmin(m * scale, 5.0));
    }
    if (r.vnsim < 0.15) mos = 1.0;
    r.moslqo = mos;
}

// ---------------------------------------------------------------- prepared references
PreparedSet::~PreparedSet() {
    if (d_block) (void)hipFree(d_block);
}

// Per reference, the values an unprepared pair computes for it, by the same
// kernels with the same per-signal work split: Σx, the Nyquist sum and the
// energy (exact: serial; fast: the one-pass moments when global alignment
// runs, else the Σx² tree), the raw spectrogram, and in speech mode the VAD
// patch indices.
std::unique_ptr<PreparedSet> Engine::prepare(const double* const* refs, const long long* lens, int n,
                                             uint32_t sr) {
    arena_.reset();
    pinned_.reset();
    DevRate& dr = dev_rate(sr);
    const RateParams& rp = dr.p;
    const int nb = rp.nb, W = rp.window, hop = rp.hop, S = rp.patch;
    auto set = std::make_unique<PreparedSet>();
    set->sample_rate = sr;
    set->mode = opts_.mode;
    set->fast = opts_.fast;
    set->refs.resize(n);
    std::vector<size_t> sig_off(n), spec_off(n);
    size_t total = 0;
    for (int i = 0; i < n; i++) {
        PreparedRef& r = set->refs[i];
        r.len = lens[i];
        r.nf = r.len > W ? num_frames(r.len, W, hop) : 0;
        sig_off[i] = total;
        total += ((size_t)std::max(1LL, r.len) * sizeof(double) + 255) & ~(size_t)255;
        spec_off[i] = total;
        total += ((size_t)nb * r.nf * sizeof(double) + 255) & ~(size_t)255;
    }
    if (hipMalloc(&set->d_block, std::max<size_t>(total, 256)) != hipSuccess) {
        (void)hipGetLastError();
        throw BackendError{-2, "device out of memory"};
    }
    char* base = (char*)set->d_block;
    std::vector<SumJob> s1, s2;
    std::vector<GtJob> gj;
    std::vector<VadJob> vj;
    std::vector<int> vad_of(n, -1);
    std::vector<const double*> up_src(n);
    std::vector<long long> up_len(n);
    std::vector<double*> up_dst(n);
    std::vector<VadPlan> vps(n);
    std::vector<size_t> vad_off(n, 0);
    size_t vad_bytes = 0;
    int max_nf = 0, max_chunks = 0;
    for (int i = 0; i < n; i++) {
        PreparedRef& r = set->refs[i];
        double* sig = (double*)(base + sig_off[i]);
        r.d_signal = sig;
        up_src[i] = refs[i];
        up_len[i] = r.len;
        up_dst[i] = sig;
        if (r.len > 0) {
            if (opts_.fast && !opts_.no_global) {
                s1.push_back(SumJob{sig, r.len, kMoments, 3 * i, 3 * i + 1, 3 * i + 2, nullptr, 0});
            } else {
                s1.push_back(SumJob{sig, r.len, kSum, 3 * i, 0, 0, nullptr, 0});
                s2.push_back(SumJob{sig, r.len, kNyquist, 3 * i + 1, 3 * i, 0, nullptr, 0});
                s1.push_back(SumJob{sig, r.len, kSumSq, 3 * i + 2, 0, 0, nullptr, 0});
            }
        }
        if (r.nf > 0) {
            r.d_spec = (const double*)(base + spec_off[i]);
            gj.push_back(GtJob{sig, (double*)(base + spec_off[i]), r.nf, 0});
            max_nf = std::max(max_nf, r.nf);
            if (opts_.mode == 1) {
                vps[i] = vad_plan(r.nf, S, W);
                if (vps[i].n_chunks > 0) {
                    vad_of[i] = (int)vj.size();
                    vad_off[i] = vad_bytes;
                    vad_bytes += (size_t)vps[i].n_chunks;
                    max_chunks = std::max(max_chunks, vps[i].n_chunks);
                    vj.push_back(VadJob{sig, r.len, nullptr, vps[i].n_chunks, vps[i].frame_len, vps[i].first, 0});
                }
            }
        }
    }
    upload_signals(n, up_src.data(), up_len.data(), up_dst.data());
    double* d_out = (double*)arena_.alloc(sizeof(double) * 3 * std::max(1, n));
    check(hipMemsetAsync(d_out, 0, sizeof(double) * 3 * std::max(1, n), stream_), "memset");
    sums(s1, d_out, !opts_.fast);
    sums(s2, d_out, !opts_.fast);
    if (!gj.empty())
        gammatone(upload(gj), (int)gj.size(), max_nf, dr, 0, &gj);
    unsigned char* d_active = vad_bytes ? (unsigned char*)arena_.alloc(vad_bytes) : nullptr;
    if (!vj.empty()) {
        for (int i = 0; i < n; i++)
            if (vad_of[i] >= 0) vj[vad_of[i]].active = d_active + vad_off[i];
        auto* scratch = (unsigned long long*)arena_.alloc(sizeof(unsigned long long) * vj.size());
        launch_vad(upload(vj), (int)vj.size(), max_chunks, scratch, stream_);
    }
    double* ps = (double*)pinned_.alloc(sizeof(double) * 3 * std::max(1, n));
    unsigned char* pa = (unsigned char*)pinned_.alloc(vad_bytes + 1);
    readback(ps, d_out, sizeof(double) * 3 * n);
    readback(pa, d_active, vad_bytes);
    sync();
    for (int i = 0; i < n; i++) {
        PreparedRef& r = set->refs[i];
        r.sum = ps[3 * i];
        r.nyq = ps[3 * i + 1];
        r.energy = ps[3 * i + 2];
        if (opts_.mode == 1 && vad_of[i] >= 0) vad_patches(vps[i], pa + vad_off[i], S, r.vad_idx);
        r.vad_done = true;
    }
    return set;
}

// ---------------------------------------------------------------- spectrogram shape tuning
namespace {
// Process-wide tuning table: (device, fast, bands, window, kind) -> state.
struct GtTuneEntry {
    int chosen = -1;
    bool busy = false;  // an engine is timing the shapes now
    std::vector<float> ms;
};
std::mutex g_gt_mu;
std::map<std::array<int, 5>, GtTuneEntry> g_gt_tune;
// Launches of less work than this, counted in (signal, 64 frames, 16 bands)
// units, run the first shape untuned: too small to fill 256 CUs several
// times over.
constexpr long long kTuneUnits = 4096;
// Launches of at most this many units (one block of the first shape each,
// at most one per CU) run the wide shape.
constexpr long long kWideUnits = 256;
}  // namespace

void Engine::gammatone(const GtJob* d_jobs, int njobs, int max_nf, const DevRate& dr, int kind,
                       const std::vector<GtJob>* host_jobs) {
    const int nb = dr.p.nb, W = dr.p.window, hop = dr.p.hop;
    // the unit list of each shape's band-row count, made on first use
    std::map<int, std::pair<const GtUnit*, int>> lists;
    auto launch = [&](int shape) {
        if (host_jobs && max_nf > kFramesPerUnit) {
            const GtShape sh = gammatone_shape(shape);
            const int rows = (nb + sh.waves * sh.bands_per_lane - 1) / (sh.waves * sh.bands_per_lane);
            long long active = 0;
            for (const GtJob& j : *host_jobs) active += (j.nf + kFramesPerUnit - 1) / kFramesPerUnit;
            const long long grid = (long long)xcd_spread((max_nf + kFramesPerUnit - 1) / kFramesPerUnit) * njobs;
            // a quarter or more of the grid's blocks would exit at once
            if (4 * active <= 3 * grid) {
                auto it = lists.find(rows);
                if (it == lists.end()) {
                    std::vector<GtUnit> units;
                    units.reserve((size_t)active * rows);
                    for (int jb = 0; jb < njobs; jb++) {
                        const int blocks = ((*host_jobs)[jb].nf + kFramesPerUnit - 1) / kFramesPerUnit;
                        for (int b = 0; b < blocks; b++)
                            for (int r = 0; r < rows; r++) units.push_back(GtUnit{jb, b, r, 0});
                    }
                    it = lists.emplace(rows, std::make_pair((const GtUnit*)upload(units), (int)units.size())).first;
                }
                launch_gammatone_units(d_jobs, it->second.first, it->second.second, dr.coef, dr.fir, dr.hann, nb, W,
                                       hop, opts_.fast, shape, stream_);
                return;
            }
        }
        launch_gammatone(d_jobs, njobs, max_nf, dr.coef, dr.fir, dr.hann, nb, W, hop, opts_.fast, shape, stream_);
    };
    if (njobs == 0 || max_nf == 0) return;
    if (gt_force_ >= 0) return launch(gt_force_);
    int dev = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    const std::array<int, 5> key{dev, opts_.fast ? 1 : 0, nb, W, kind};
    const long long units = (long long)njobs * ((max_nf + 63) / 64) * ((nb + 15) / 16);
    {
        std::lock_guard<std::mutex> g(g_gt_mu);
        GtTuneEntry& t = g_gt_tune[key];
        // Launches too small to give every CU a block (a few served pairs,
        // their realignment patches) run the wide shape: their time is one
        // block's serial work, which the most blocks split finest.
        if (units <= kWideUnits) return launch(gammatone_wide_shape());
        if (t.chosen >= 0) return launch(t.chosen);
        // Realignment launches (kind 1) keep the first shape: their sizes
        // vary from batch to batch, and a timing run in the middle of a
        // served batch would stall it for the little they could gain.
        if (t.busy || kind != 0 || units < kTuneUnits) return launch(0);
        t.busy = true;
    }
    // Every shape writes the same output, so this launch's spectrograms
    // come out of whichever runs last.  Two rounds; each shape keeps its
    // faster time (the first round also absorbs one-time launch costs).
    const int count = gammatone_shape_count();
    std::vector<hipEvent_t> ev(count + 1, nullptr);
    std::vector<float> best(count, 0.f);
    int chosen = 0;
    try {
        for (auto& e : ev) check(hipEventCreate(&e), "hipEventCreate");
        // nothing else enters the stream between the events
        std::lock_guard<std::recursive_mutex> eg(*enqueue_mu_);
        for (int round = 0; round < 2; round++) {
            check(hipEventRecord(ev[0], stream_), "hipEventRecord");
            for (int v = 0; v < count; v++) {
                launch(v);
                check(hipEventRecord(ev[v + 1], stream_), "hipEventRecord");
            }
            check(hipEventSynchronize(ev[count]), "hipEventSynchronize");
            for (int v = 0; v < count; v++) {
                float ms = 0.f;
                check(hipEventElapsedTime(&ms, ev[v], ev[v + 1]), "hipEventElapsedTime");
                best[v] = round == 0 ? ms : std::min(best[v], ms);
            }
        }
        // The fastest shape; within 2% of it, the one with the most bands
        // per block (it reads each frame's samples from HBM fewer times,
        // leaving bandwidth to work running beside it).
        int fastest = 0;
        for (int v = 1; v < count; v++)
            if (best[v] < best[fastest]) fastest = v;
        chosen = fastest;
        auto bands = [](int v) { const GtShape g = gammatone_shape(v); return g.waves * g.bands_per_lane; };
        for (int v = 0; v < count; v++)
            if (best[v] <= 1.02f * best[fastest] &&
                (bands(v) > bands(chosen) || (bands(v) == bands(chosen) && best[v] < best[chosen])))
                chosen = v;
    } catch (...) {
        for (auto e : ev)
            if (e) (void)hipEventDestroy(e);
        std::lock_guard<std::mutex> g(g_gt_mu);
        g_gt_tune[key].busy = false;
        throw;
    }
    for (auto e : ev) (void)hipEventDestroy(e);
    std::lock_guard<std::mutex> g(g_gt_mu);
    GtTuneEntry& t = g_gt_tune[key];
    t.busy = false;
    t.ms = best;
    t.chosen = chosen;
}

Engine::GtTuning Engine::gammatone_tuning(uint32_t sample_rate, int kind) {
    const RateParams& rp = rate(sample_rate);
    int dev = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> g(g_gt_mu);
    auto it = g_gt_tune.find({dev, opts_.fast ? 1 : 0, rp.nb, rp.window, kind});
    GtTuning out;
    if (it != g_gt_tune.end()) {
        out.chosen = it->second.chosen;
        out.ms = it->second.ms;
    }
    return out;
}

// ---------------------------------------------------------------- stage access
void Engine::spectrogram(uint32_t sr, const double* host_sig, long long len, std::vector<double>& out, int& nf) {
    arena_.reset();
    pinned_.reset();
    DevRate& dr = dev_rate(sr);
    nf = num_frames(len, dr.p.window, dr.p.hop);
    out.assign((size_t)dr.p.nb * nf, 0.0);
    if (nf == 0) return;
    double* d = (double*)arena_.alloc(sizeof(double) * len);
    check(hipMemcpyAsync(d, host_sig, sizeof(double) * len, hipMemcpyHostToDevice, stream_), "upload");
    double* o = (double*)arena_.alloc(sizeof(double) * out.size());
    std::vector<GtJob> j{GtJob{d, o, nf, 0}};
    gammatone(upload(j), 1, nf, dr, 0);
    double* p = (double*)pinned_.alloc(sizeof(double) * out.size());
    readback(p, o, sizeof(double) * out.size());
    sync();
    memcpy(out.data(), p, sizeof(double) * out.size());
}

long long Engine::best_lag(const double* a, long long la, const double* b, long long lb) {
    arena_.reset();
    pinned_.reset();
    if (std::max(la, lb) == 0) return 0;
    double* da = (double*)arena_.alloc(sizeof(double) * std::max(1LL, la));
    double* db = (double*)arena_.alloc(sizeof(double) * std::max(1LL, lb));
    if (la) check(hipMemcpyAsync(da, a, sizeof(double) * la, hipMemcpyHostToDevice, stream_), "upload");
    if (lb) check(hipMemcpyAsync(db, b, sizeof(double) * lb, hipMemcpyHostToDevice, stream_), "upload");
    double* d_s = (double*)arena_.alloc(sizeof(double) * 4);
    check(hipMemsetAsync(d_s, 0, sizeof(double) * 4, stream_), "memset");
    std::vector<SumJob> s1{SumJob{da, la, kSum, 0, 0, 0, nullptr, 0}, SumJob{db, lb, kSum, 1, 0, 0, nullptr, 0}};
    std::vector<SumJob> s2{SumJob{da, la, kNyquist, 2, 0, 0, nullptr, 0}, SumJob{db, lb, kNyquist, 3, 1, 0, nullptr, 0}};
    sums(s1, d_s, !opts_.fast);
    sums(s2, d_s, !opts_.fast);
    const long long nfft = fft_size_for(std::max(la, lb));
    double2* z = (double2*)arena_.alloc(sizeof(double2) * nfft);
    std::vector<LagJob> lj{LagJob{da, db, la, lb, z, nfft, 0, 1, 2, 3}};
    std::vector<int> pos;
    long long* d_lag = lag_search(lj, d_s, pos);
    long long* p = (long long*)pinned_.alloc(sizeof(long long));
    readback(p, d_lag, sizeof(long long));
    sync();
    return *p;
}

double Engine::reference_sum(int kind, const double* x, long long len, double plain_sum, long long lag, long long half,
                             int form) {
    arena_.reset();
    pinned_.reset();
    double* dx = (double*)arena_.alloc(sizeof(double) * std::max(1LL, len));
    if (len) check(hipMemcpyAsync(dx, x, sizeof(double) * len, hipMemcpyHostToDevice, stream_), "upload");
    // slot 0: the plain sum (kNyquist's mean), slot 1: the result, then the lag
    double* d_s = (double*)arena_.alloc(sizeof(double) * 2 + sizeof(long long));
    double* h = (double*)pinned_.alloc(sizeof(double) * 2 + sizeof(long long));
    h[0] = plain_sum;
    h[1] = 0.0;
    memcpy(h + 2, &lag, sizeof(long long));
    check(hipMemcpyAsync(d_s, h, sizeof(double) * 2 + sizeof(long long), hipMemcpyHostToDevice, stream_), "upload");
    const long long* d_lag = reinterpret_cast<const long long*>(d_s + 2);
    std::vector<SumJob> jobs{SumJob{dx, len, kind, 1, 0, 0, kind == kSumSqFrom ? d_lag : nullptr, half}};
    if (form == 1 && spec_sums_fit(len)) {
        void* scratch = arena_.alloc(spec_sums_scratch_bytes(1, len));
        launch_sums_spec(upload(jobs), 1, len, d_s, scratch, stream_);
    } else if (form == 2) {
        sums(jobs, d_s, true);
    } else {
        launch_sums(upload(jobs), 1, len, d_s, true, nullptr, stream_);
    }
    double* r = (double*)pinned_.alloc(sizeof(double));
    readback(r, d_s + 1, sizeof(double));
    sync();
    return *r;
}

void Engine::xcorr(const double* a, long long la, const double* b, long long lb, std::vector<double>& out) {
    arena_.reset();
    pinned_.reset();
    out.clear();
    if (std::max(la, lb) == 0) return;
    double* da = (double*)arena_.alloc(sizeof(double) * std::max(1LL, la));
    double* db = (double*)arena_.alloc(sizeof(double) * std::max(1LL, lb));
    if (la) check(hipMemcpyAsync(da, a, sizeof(double) * la, hipMemcpyHostToDevice, stream_), "upload");
    if (lb) check(hipMemcpyAsync(db, b, sizeof(double) * lb, hipMemcpyHostToDevice, stream_), "upload");
    double* d_s = (double*)arena_.alloc(sizeof(double) * 4);
    check(hipMemsetAsync(d_s, 0, sizeof(double) * 4, stream_), "memset");
    std::vector<SumJob> s1{SumJob{da, la, kSum, 0, 0, 0, nullptr, 0}, SumJob{db, lb, kSum, 1, 0, 0, nullptr, 0}};
    std::vector<SumJob> s2{SumJob{da, la, kNyquist, 2, 0, 0, nullptr, 0}, SumJob{db, lb, kNyquist, 3, 1, 0, nullptr, 0}};
    sums(s1, d_s, !opts_.fast);
    sums(s2, d_s, !opts_.fast);
    const long long nfft = fft_size_for(std::max(la, lb));
    double2* z = (double2*)arena_.alloc(sizeof(double2) * nfft);
    std::vector<LagJob> lj{LagJob{da, db, la, lb, z, nfft, 0, 1, 2, 3}};
    LagJob* d_lj = upload(lj);
    launch_lag_search(d_lj, 1, nfft, twiddles(nfft), d_s, nullptr, nullptr, stream_, true);
    double* p = (double*)pinned_.alloc(sizeof(double2) * nfft);
    readback(p, z, sizeof(double2) * nfft);
    sync();
    out.resize(nfft);
    for (long long i = 0; i < nfft; i++) out[i] = p[2 * i];
}

// Both probes form the sums (Σx, Nyquist sum) with the throughput mode's
// one-pass reduction, as the pipeline's phase A does.
long long Engine::best_lag_fast(const double* a, long long la, const double* b, long long lb, bool* nominated) {
    arena_.reset();
    pinned_.reset();
    if (nominated) *nominated = false;
    if (std::max(la, lb) == 0) return 0;
    double* da = (double*)arena_.alloc(sizeof(double) * std::max(1LL, la));
    double* db = (double*)arena_.alloc(sizeof(double) * std::max(1LL, lb));
    if (la) check(hipMemcpyAsync(da, a, sizeof(double) * la, hipMemcpyHostToDevice, stream_), "upload");
    if (lb) check(hipMemcpyAsync(db, b, sizeof(double) * lb, hipMemcpyHostToDevice, stream_), "upload");
    double* d_s = (double*)arena_.alloc(sizeof(double) * 4);
    check(hipMemsetAsync(d_s, 0, sizeof(double) * 4, stream_), "memset");
    sums({SumJob{da, la, kMoments, 0, 2, -1, nullptr, 0}, SumJob{db, lb, kMoments, 1, 3, -1, nullptr, 0}}, d_s, false);
    const long long nfft = fft_size_for(std::max(la, lb));
    double2* z = (double2*)arena_.alloc(sizeof(double2) * nfft);
    std::vector<LagJob> lj{LagJob{da, db, la, lb, z, nfft, 0, 1, 2, 3}};
    std::vector<int> pos;
    int* d_fb = nullptr;
    long long* d_lag = lag_search(lj, d_s, pos, &d_fb);
    long long* p = (long long*)pinned_.alloc(sizeof(long long));
    int* pf = (int*)pinned_.alloc(sizeof(int));
    *pf = 1;
    readback(p, d_lag, sizeof(long long));
    if (d_fb) readback(pf, d_fb, sizeof(int));
    sync();
    if (!*pf) {
        if (nominated) *nominated = true;
        return *p;
    }
    d_lag = lag_search(lj, d_s, pos, nullptr, true);
    readback(p, d_lag, sizeof(long long));
    sync();
    return *p;
}

void Engine::xcorr32(const double* a, long long la, const double* b, long long lb, std::vector<float>& out) {
    arena_.reset();
    pinned_.reset();
    out.clear();
    const long long nfft = fft_size_for(std::max(la, lb));
    if (std::max(la, lb) == 0 || !lag32_handles(nfft)) return;
    double* da = (double*)arena_.alloc(sizeof(double) * std::max(1LL, la));
    double* db = (double*)arena_.alloc(sizeof(double) * std::max(1LL, lb));
    if (la) check(hipMemcpyAsync(da, a, sizeof(double) * la, hipMemcpyHostToDevice, stream_), "upload");
    if (lb) check(hipMemcpyAsync(db, b, sizeof(double) * lb, hipMemcpyHostToDevice, stream_), "upload");
    double* d_s = (double*)arena_.alloc(sizeof(double) * 4);
    check(hipMemsetAsync(d_s, 0, sizeof(double) * 4, stream_), "memset");
    sums({SumJob{da, la, kMoments, 0, 2, -1, nullptr, 0}, SumJob{db, lb, kMoments, 1, 3, -1, nullptr, 0}}, d_s, false);
    double2* z = (double2*)arena_.alloc(sizeof(double2) * nfft);
    std::vector<LagJob> lj{LagJob{da, db, la, lb, z, nfft, 0, 1, 2, 3}};
    void* scratch = arena_.alloc(lag32_scratch_bytes(1, nfft));
    launch_lag_search32(upload(lj), 1, nfft, twiddles32(nfft), d_s, scratch, opts_.lag_cap, nullptr, nullptr, stream_,
                        true);
    float* p = (float*)pinned_.alloc(sizeof(float2) * nfft);
    readback(p, z, sizeof(float2) * nfft);
    sync();
    out.resize(nfft);
    for (long long h = 0; h < nfft / 2; h++) {
        out[2 * h] = p[2 * lag32_xcorr_slot(h)];
        out[2 * h + 1] = p[2 * lag32_xcorr_slot(h) + 1];
    }
}

void Engine::fft_probe(double* z, long long n, bool dit) {
    arena_.reset();
    pinned_.reset();
    if (n < 2) return;
    double2* d = (double2*)arena_.alloc(sizeof(double2) * n);
    check(hipMemcpyAsync(d, z, sizeof(double2) * n, hipMemcpyHostToDevice, stream_), "upload");
    std::vector<LagJob> rows{LagJob{nullptr, nullptr, 0, 0, d, n, 0, 0, 0, 0}};
    LagJob* d_rows = upload(rows);
    if (dit) launch_fft_dit(d_rows, 1, n, twiddles(n), stream_);
    else launch_fft_dif(d_rows, 1, n, twiddles(n), stream_);
    double* p = (double*)pinned_.alloc(sizeof(double2) * n);
    readback(p, d, sizeof(double2) * n);
    sync();
    memcpy(z, p, sizeof(double2) * n);
}

void Engine::math_probe(int op, const double* a, const double* b, double* out, long long n) {
    arena_.reset();
    pinned_.reset();
    if (n == 0) return;
    double* da = (double*)arena_.alloc(sizeof(double) * n);
    double* db = (double*)arena_.alloc(sizeof(double) * n);
    double* dout = (double*)arena_.alloc(sizeof(double) * n);
    check(hipMemcpyAsync(da, a, sizeof(double) * n, hipMemcpyHostToDevice, stream_), "upload");
    if (b) check(hipMemcpyAsync(db, b, sizeof(double) * n, hipMemcpyHostToDevice, stream_), "upload");
    launch_math_probe(op, da, db, dout, n, stream_);
    double* p = (double*)pinned_.alloc(sizeof(double) * n);
    readback(p, dout, sizeof(double) * n);
    sync();
    memcpy(out, p, sizeof(double) * n);
}

}  // namespace vq
