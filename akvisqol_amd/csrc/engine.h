// The device scoring pipeline: batches of pairs at one sample rate run
// stage by stage on one gfx950 device (visqol.zig:394-487's eight stages),
// with host planning between the phases whose shapes depend on device
// results (the global lag, the DP choice, the realignment lags).
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "host.h"
#include "kernels.h"

namespace vq {

struct EngineOptions {
    int mode = 0;
    int radius = 60;
    bool unscaled_speech = false;
    bool no_global = false;
    bool no_realign = false;
    bool lattice = true;
    bool fast = false;
    // Batches of several chunks: `scorers` engines, each on its own host
    // thread, take the chunks in turn, so one chunk's host work (planning,
    // decoding, aggregation) overlaps the others' kernels
    // (VISQOL_HIP_SCORERS, 2-4, read when an engine is created).  Two
    // measured best on 1024 ten-second pairs (on one shared stream 8300
    // pairs/s against 8100 with three and 7840 with four; on streams of
    // their own, three no faster than two).
    bool interleave_chunks = true;
    int scorers = 2;
    // Each scorer on a stream of its own, so one chunk's lag searches,
    // similarity tables and small kernels run beside another's and in the
    // gaps of its host work (VISQOL_HIP_SCORER_STREAMS=0: one shared
    // in-order stream instead).  1024 ten-second pairs: 9000-9100 against
    // 8600-8670 pairs/s on one shared stream.
    bool scorer_streams = true;
    // With scorer streams: whole-signal spectrogram launches run with
    // nothing from the other scorers' streams beside them, so their own
    // duration is measurable (VISQOL_HIP_GT_EXCLUSIVE=0 lets other kernels
    // share the device with them: ~2% more pairs/s, each spectrogram launch
    // then stretched from ~16 to ~23 ms).
    bool gt_exclusive = true;
    // Exact mode: launches of reference-order sums over few long signals
    // take the speculative form (same bits; a ten-second signal's sum in
    // ~0.2 ms instead of ~2 ms as one serial chain); VISQOL_HIP_SPEC_SUMS=0
    // keeps one serial chain per signal.
    bool spec_sums = true;
    // Throughput mode: rows of the lag search with more nominated lags than
    // this (at most kMaxLagCands) take the f64 search instead; 0 sends every
    // row there (VISQOL_HIP_LAG_CAP, read when an engine is created).
    int lag_cap = kMaxLagCands;
};

// A reference prepared for scoring many degraded signals against it
// (backend.zig:84-99, visqol.zig:194-244): what depends on the reference
// alone, computed once by the same kernels as an unprepared pair.
struct PreparedRef {
    const double* d_signal = nullptr;  // device samples
    long long len = 0;
    const double* d_spec = nullptr;    // raw spectrogram [bands][nf] (before dB prep), or null if too short
    int nf = 0;
    double sum = 0, nyq = 0, energy = 0;  // Σx, Nyquist sum, Σx² (the pair path's slots 0, 2, 4)
    bool vad_done = false;               // speech: patch indices from the VAD
    std::vector<int> vad_idx;
};
struct PreparedSet {
    uint32_t sample_rate = 0;
    int mode = 0;
    bool fast = false;
    void* d_block = nullptr;  // one device allocation for every reference's samples and spectrogram
    std::vector<PreparedRef> refs;
    ~PreparedSet();
};

struct PairInput {
    const double* ref;
    long long ref_len;
    const double* deg;
    long long deg_len;
    bool on_device;                      // pointers are device memory
    const PreparedRef* prep = nullptr;   // ref/ref_len are then the prepared reference's
};

struct Match {
    double similarity = 0, ref_start = 0, ref_end = 0, deg_start = 0, deg_end = 0;
    double means[kMaxBands] = {0}, stddevs[kMaxBands] = {0}, energy[kMaxBands] = {0};
    bool valid = true;
};

struct PairResult {
    int status = 0;  // visqol.h code
    std::string error;
    double moslqo = 0, vnsim = 0, lag_s = 0;
    std::vector<double> fvnsim, fvnsim10, fstdnsim, fvdegenergy, cf;
    std::vector<Match> matches;
    // Trace of the discrete decisions.
    long long lag = 0;
    double spl_factor = 1.0;
    int nf_ref = 0, nf_deg = 0, np = 0;
    std::vector<int> ref_idx, dp_offset;
    std::vector<unsigned char> match_valid, realign_accepted;
    std::vector<long long> realign_lag;
    std::vector<int> realign_frames;
    // Optional stage outputs.
    std::vector<double> ref_spec, deg_spec, sims;
};

// Device memory in chunks; allocations stay put until reset().
class DeviceArena {
   public:
    ~DeviceArena();
    void* alloc(size_t bytes);
    void reset();

   private:
    struct Chunk {
        char* base;
        size_t size, used;
    };
    std::vector<Chunk> chunks_;
    size_t total_used_ = 0, high_water_ = 0;
};

// Pinned host staging for job tables and readbacks.
class PinnedArena {
   public:
    ~PinnedArena();
    void* alloc(size_t bytes);
    void reset();

   private:
    struct Chunk {
        char* base;
        size_t size, used;
    };
    std::vector<Chunk> chunks_;
};

struct PairState;  // engine_state.h

// Host threads for per-pair host work over n pairs: one per `per` pairs, at
// most 16 and the machine's hardware threads, 1 below 2·per pairs.
int host_threads(size_t n, size_t per);

// Samples a signal of len samples takes in upload staging: at least one,
// rounded up to 256 bytes.
size_t upload_span(long long len);

struct BackendError {
    int code;
    std::string msg;
};

class Engine {
   public:
    Engine(const EngineOptions& o, std::unique_ptr<SvrModel> svr,
           std::unique_ptr<LatticeModel, LatticeDeleter> lattice);
    ~Engine();
    // Scores n pairs; per-pair outcome in out[i].status.  Throws
    // BackendError for device failures that poison the whole batch.  With
    // grads, grads[i] (host, deg_len values, or NULL to skip the pair)
    // receives ∂moslqo/∂deg of each pair scored (backend.zig:101-110), the
    // discrete choices of the forward pass held fixed; zero for a pair that
    // fails.
    void run(const PairInput* pairs, int n, uint32_t sample_rate, PairResult* out, bool want_stages,
             double* const* grads = nullptr);
    const EngineOptions& options() const { return opts_; }
    void set_lag_cap(int cap) { opts_.lag_cap = cap; }

    // References of one sample rate prepared for run() (PairInput::prep);
    // throws BackendError on device failure.
    std::unique_ptr<PreparedSet> prepare(const double* const* refs, const long long* lens, int n, uint32_t sample_rate);

    // Stage access for checks.
    void spectrogram(uint32_t sample_rate, const double* host_sig, long long len, std::vector<double>& out, int& nf);
    long long best_lag(const double* a, long long la, const double* b, long long lb);
    // One reference-order sum over host samples x[0..len): kind a SumKind
    // (kNyquist's mean from plain_sum, kSumSqFrom's start from lag and
    // half); form 0 the serial chain, 1 the speculative form, 2 as the
    // engine's own launches choose.
    double reference_sum(int kind, const double* x, long long len, double plain_sum, long long lag, long long half,
                         int form);
    // The throughput mode's lag (nominated in f32, rescored in f64) with
    // the engine's cap; *nominated is false when the row fell back to the
    // f64 search (or its size is outside the nominating search's range).
    long long best_lag_fast(const double* a, long long la, const double* b, long long lb, bool* nominated);
    // The nominating search's f32 correlation of the two envelopes (Re,
    // times n), fft_size_for(max(la, lb)) values; empty when that size is
    // outside its range.
    void xcorr32(const double* a, long long la, const double* b, long long lb, std::vector<float>& out);
    // Re of the circular cross-correlation of the two signals' envelopes
    // (the values best_lag maximises), fft_size_for(max(la, lb)) of them.
    void xcorr(const double* a, long long la, const double* b, long long lb, std::vector<double>& out);
    void math_probe(int op, const double* a, const double* b, double* out, long long n);
    // The spectrogram adjoint on one frame (window samples, nb band slopes
    // g): out = ∂(Σ_b g_b · rms_b)/∂frame, samples outside [lo, hi) read as
    // zero and given zero slope.
    void frame_grad(uint32_t sample_rate, const double* frame, const double* g, int lo, int hi, double* out);
    void fft_probe(double* z, long long n, bool dit);
    const RateParams& rate(uint32_t sample_rate);

    // Device time of the full-signal gammatone launches, measured with HIP
    // events on the engine's stream, and their algorithmic work.
    struct Profile {
        double gammatone_ms = 0;
        double gammatone_flop = 0;  // 30 per (sample, band) of every frame
        double gammatone_bytes = 0; // signal samples read + spectrogram written
        long long launches = 0;
        // Similarity-table launches (reference-term prep + table), same way:
        // device time and cells evaluated (patch offsets × bands × patch
        // frames).
        double sim_ms = 0;
        double sim_cells = 0;
        long long sim_launches = 0;
        // Host wall time of each stretch of a batch call: phase A up to its
        // readback, host planning, phase B up to the DP readback, the match
        // statistics, their host decoding, phase C's lag searches, phase C's
        // spectrograms and statistics, and the final aggregation.
        static constexpr int kPhases = 8;
        double phase_ms[kPhases] = {0};
    };
    // Spectrogram block shape for launches of one kind (0: whole signals,
    // 1: realignment slices, which always run the first shape) at one
    // precision, band count and window: the shape chosen by timing every
    // shape on the first whole-signal launch large enough to fill the
    // device (shared by the process's engines on the device), or -1 while
    // untuned; ms holds the timed launch of each shape.
    struct GtTuning {
        int chosen = -1;
        std::vector<float> ms;
    };
    GtTuning gammatone_tuning(uint32_t sample_rate, int kind);

    void set_profiling(bool on) {
        profiling_ = on;
        for (auto& h : helpers_) h->set_profiling(on);
    }
    Profile profile() const { return prof_; }
    void reset_profile() {
        prof_ = Profile();
        for (auto& h : helpers_) h->reset_profile();
    }

   private:
    // A helper for interleaved chunks: shares the parent's options, models,
    // stream fence and enqueue lock (and stream, without scorer streams);
    // owns its device arena, staging and tables.
    struct HelperTag {};
    Engine(HelperTag, Engine& parent);
    std::vector<std::unique_ptr<Engine>> helpers_;  // scorers - 1 of them, made on first use
    // Uploads host-resident chunks ahead of scoring into slots_ (device
    // buffers of slot_len_ samples each, kept between calls).
    std::unique_ptr<Engine> uploader_;
    std::vector<double*> slots_;
    size_t slot_len_ = 0;
    // The streams of an engine and its helpers, for launches that must run
    // with nothing beside them on the device: such a launch waits for the
    // work already enqueued on the other streams, and their later work
    // waits for it.
    struct StreamFence {
        std::mutex mu;
        std::vector<hipStream_t> streams;
        std::vector<hipEvent_t> events;  // one per stream, created on first use
    };
    std::shared_ptr<StreamFence> fence_;
    template <typename F>
    void exclusive(F&& enqueue);
    std::recursive_mutex own_enqueue_mu_;
    std::recursive_mutex* enqueue_mu_ = &own_enqueue_mu_;  // held around timed launches on a shared stream
    bool owns_stream_ = true;
    void merge_profile(const Profile& p);
    bool profiling_ = false;
    std::chrono::steady_clock::time_point lap_t_;
    void lap(int phase);
    Profile prof_;
    hipEvent_t ev_[4] = {nullptr, nullptr, nullptr, nullptr};
    struct DevRate {
        RateParams p;
        BandCoef* coef = nullptr;
        FirCoef* fir = nullptr;
        double* hann = nullptr;
    };
    DevRate& dev_rate(uint32_t sample_rate);
    // Level twiddles for transforms of up to n points, shared by the
    // process's engines on the current device.
    const double2* twiddles(long long n);
    const float2* twiddles32(long long n);  // the f64 table rounded to f32
    template <typename T>
    T* upload(const std::vector<T>& v);
    template <typename T>
    T* upload(const T* v, size_t count);
    void readback(void* host_dst, const void* dev_src, size_t bytes);
    // Host signals src[i] (len[i] samples each) to device memory dst[i]
    // through pinned staging, on this engine's copy stream; returns once
    // every sample is on the device.
    // A destination laid out as upload_span of each length in turn from
    // dst[0] gets one DMA per staged piece.
    void upload_signals(int n, const double* const* src, const long long* len, double* const* dst);
    hipStream_t copy_streams_[2] = {nullptr, nullptr};  // created on the first upload_signals
    void sync();
    void check(hipError_t e, const char* what);
    // One chunk of pairs (bounded HBM footprint).
    void run_chunk(const PairInput* pairs, int n, uint32_t sr, PairResult* out, bool want_stages,
                   double* const* grads);
    // The gradient pass of a chunk whose forward pass left its state in st.
    void backward(const PairInput* pairs, int n, uint32_t sr, PairResult* out, std::vector<PairState>& st,
                  double* const* grads);
    // ∂moslqo/∂(fvnsim, fvnsim10, fstdnsim, fvdegenergy) of this engine's
    // quality mapping at the given features (visqol.zig:100-149).
    // The lattice's slopes use up to `threads` host threads.
    void mapper_grad(int nb, const double* f0, const double* f1, const double* f2, const double* f3, double* g0,
                     double* g1, double* g2, double* g3, int threads = 1) const;
    // Reductions of `jobs` into d_out (serial: the reference's order).
    void sums(const std::vector<SumJob>& jobs, double* d_out, bool serial);
    // FFT cross-correlation lags of `jobs` (envelope sums already in
    // d_sums); pos[i] is job i's slot in the returned device lag array.  In
    // the throughput mode rows of a size the nominating search handles take
    // it (unless `exact`), and *d_fallback (one int per slot; nullptr when
    // no row took it) flags the rows it could not nominate, whose lags are
    // left for lag_search(..., exact = true) on those rows.
    long long* lag_search(const std::vector<LagJob>& jobs, const double* d_sums, std::vector<int>& pos,
                          int** d_fallback = nullptr, bool exact = false);
    void finish(PairResult& r, const RateParams& rp);
    // Spectrograms of d_jobs with the tuned block shape (tuning on the
    // first large launch of this kind); kind as in gammatone_tuning.
    // With the host copy of the jobs, whole-signal launches whose grid
    // would be mostly blocks past a shorter job's end take the persistent
    // unit-list form instead.
    void gammatone(const GtJob* d_jobs, int njobs, int max_nf, const DevRate& dr, int kind,
                   const std::vector<GtJob>* host_jobs = nullptr);
    int gt_force_ = -1;  // block shape forced by VISQOL_HIP_GAMMATONE_SHAPE, else -1

    EngineOptions opts_;
    std::shared_ptr<const SvrModel> svr_;
    std::shared_ptr<const LatticeModel> lattice_;
    hipStream_t stream_ = nullptr;
    DeviceArena arena_;
    PinnedArena pinned_;
    std::map<uint32_t, std::unique_ptr<DevRate>> rates_;
};

}  // namespace vq
