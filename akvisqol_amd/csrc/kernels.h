// Launch interface of the gfx950 kernels (host side).  Every launcher takes
// device pointers and a stream; job tables live in device memory.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "dev_common.h"

namespace vq {

// One signal's spectrogram: out[b * nf + f] for frames f < nf.
struct GtJob {
    const double* sig;
    double* out;
    int nf;
    int pad_;
};

// Block shape of the spectrogram kernel: waves per block, bands per lane,
// samples per staged chunk (64 frames per block, one per lane).  Every shape
// gives the same output bits.
struct GtShape {
    int waves, bands_per_lane, chunk;
};
int gammatone_shape_count();
GtShape gammatone_shape(int v);
// The shape with the most blocks per launch (one band per lane), for
// launches too small to fill the device.  Short jobs (two per block) run
// shape 0 or this one.
int gammatone_wide_shape();
void launch_gammatone(const GtJob* d_jobs, int njobs, int max_nf, const BandCoef* d_coef, const FirCoef* d_fir,
                      const double* d_hann, int nb, int window, int hop, bool fast, int shape, hipStream_t stream);

// The persistent form over jobs of different lengths: units lists the
// (job, 64-frame block, band row of WAVES·BPL bands of `shape`) blocks that
// have frames; a grid of at most the blocks the device holds at once walks
// them.  Same bits as launch_gammatone with the same shape.
struct GtUnit {
    int job, block, row, pad_;
};
constexpr int kFramesPerUnit = 64;  // frames of one block (one per lane)
void launch_gammatone_units(const GtJob* d_jobs, const GtUnit* d_units, int nunits, const BandCoef* d_coef,
                            const FirCoef* d_fir, const double* d_hann, int nb, int window, int hop, bool fast,
                            int shape, hipStream_t stream);

// Grid extent for a dimension whose trailing blocks exit at once when a job
// is shorter than the longest (jobs of uneven size).  Workgroups go to the
// eight XCDs round-robin by linear block index, so with an even extent the
// busy leading blocks of every row land on the same few XCDs (a 48-wide
// spectrogram grid over 1-60 s signals put 1.35x the mean load on XCD 0);
// an odd extent rotates them from row to row.
inline unsigned xcd_spread(unsigned n) { return n > 1 && n % 2 == 0 ? n + 1 : n; }

// ---------------------------------------------------------------- reductions
// Serial (reference-order) or tree reductions of one signal each.
enum SumKind : int {
    kSum = 0,        // Σ x
    kSumSq = 1,      // Σ x·x
    kNyquist = 2,    // Σ (-1)^i (x - mean), mean = out[aux] / len
    kSumSqFrom = 3,  // Σ x·x over x[start..len), start from a lag slot (see SumJob)
    kMoments = 4     // tree only, one pass: out[slot] = Σ x, out[aux] = the
                     // Nyquist sum as Σ (-1)^i x - mean·(len mod 2), and
                     // out[slot2] = Σ x·x when slot2 >= 0
};
struct SumJob {
    const double* x;
    long long len;
    int kind;
    int slot;              // result goes to out[slot]
    int aux;               // kNyquist: slot holding the plain sum; kMoments: Nyquist output slot
    int slot2;             // kMoments: Σ x·x output slot, or -1
    const long long* lag;  // kSumSqFrom: global lag; start = -lag if 0 < -lag <= half, else 0
    long long half;        // kSumSqFrom: ref_len / 2
};
// out[job.slot] = reduction of each job.  `serial` keeps the reference's
// left-to-right order (bit-exact); otherwise a deterministic tree order over
// up to kSumParts shares per job (max_len: longest job), d_partials holding
// njobs * kSumOutputs * kSumParts doubles.
constexpr int kSumParts = 64;
constexpr int kSumOutputs = 3;
void launch_sums(const SumJob* d_jobs, int njobs, long long max_len, double* d_out, bool serial, double* d_partials,
                 hipStream_t stream);
// The serial sums' speculative form (the same bits): each job's terms in
// segments (8 K terms from 2048 candidate starts each when the longest
// job has kSpecLongLen terms or more, else 4 K from 512) run at once,
// then resolved segment by segment (a segment whose true start falls
// outside its candidates is run again serially).  For launches of few
// jobs, where the serial form would run a handful of chains on an
// otherwise idle device.  d_scratch holds spec_sums_scratch_bytes(njobs,
// max_len) bytes; spec_sums_fit(max_len): at most kSpecMaxSegs segments;
// spec_sums_cost_ratio: the form's modelled time over the serial form's
// (total_len: the jobs' summed lengths).
constexpr int kSpecMaxSegs = 2048;
constexpr long long kSpecLongLen = 1 << 16;
size_t spec_sums_scratch_bytes(int njobs, long long max_len);
bool spec_sums_fit(long long max_len);
double spec_sums_cost_ratio(int njobs, long long max_len, long long total_len);
void launch_sums_spec(const SumJob* d_jobs, int njobs, long long max_len, double* d_out, void* d_scratch,
                      hipStream_t stream);

// ---------------------------------------------------------------- envelope + FFT xcorr
// One lag search: z[i] = (env(a)[i], env(b)[i]) zero padded to n (a power
// of two), z being the job's n-element work row.
struct LagJob {
    const double* a;
    const double* b;
    long long la, lb;
    double2* z;
    long long n;
    int sum_a, sum_b;  // slots of the plain sums in the sums array
    int nyq_a, nyq_b;  // slots of the alternating (Nyquist) sums
};
// Lag search for njobs jobs of one size n (contiguous in d_jobs): lags[j] =
// first argmax of |Re xcorr| in the order -max_lag..-1, 0..max_lag.
// d_partials holds 2 * njobs * lag_partials_per_job(n) words.  keep_xcorr:
// leave the circular cross-correlation in each job's z instead (no lags).
long long lag_partials_per_job(long long n);
void launch_lag_search(const LagJob* d_jobs, int njobs, long long n, const double2* d_tw, const double* d_sums,
                       unsigned long long* d_partials, long long* d_lags, hipStream_t stream, bool keep_xcorr = false);
// Twiddle tables (d_tw) hold W_S^j for every span S = 2..n at complex index
// S/2 + j.
// Radix-2 DIF forward / DIT inverse of each job's z.
void launch_fft_dif(const LagJob* d_jobs, int njobs, long long n, const double2* d_tw, hipStream_t stream);
void launch_fft_dit(const LagJob* d_jobs, int njobs, long long n, const double2* d_tw, hipStream_t stream);

// Throughput-mode lag search (xcorr.zig:269-340 refineLag / rescore,
// engine.zig:800-822): an f32 correlation of the same packed envelopes
// nominates the lags within its error bound of the row's maximum, and f64
// dot products of the f64 envelopes pick the first maximum among them.  A
// row is nominated when at most `cap` (<= kMaxLagCands,
// MAX_REFINE_CANDIDATES) lags qualify; lags[j] is then written and
// fallback[j] = 0, else fallback[j] = 1 and lags[j] is left for the f64
// search.  d_tw32: the f32 twiddle table (W_S^j at S/2 + j, rounded from
// the f64 one); each job's z holds at least n float2.  keep_xcorr: leave
// the f32 circular correlation c (times n) in each job's z instead, for
// checks: c[2h] and c[2h + 1] are the two floats of the float2 at
// lag32_xcorr_slot(h), h < n/2.
constexpr int kMaxLagCands = 64;
__host__ __device__ inline long long lag32_xcorr_slot(long long h) { return h + (h & ~2047LL); }
bool lag32_handles(long long n);  // n a power of two in [2^14, 2^30]
size_t lag32_scratch_bytes(int njobs, long long n);
void launch_lag_search32(const LagJob* d_jobs, int njobs, long long n, const float2* d_tw32, const double* d_sums,
                         void* d_scratch, int cap, long long* d_lags, int* d_fallback, hipStream_t stream,
                         bool keep_xcorr = false);

// ---------------------------------------------------------------- signal prep
// dst[i] = src-aligned sample i times factor: negative lag drops -lag
// leading samples, positive lag prepends zeros (realign.zig:91-110).
struct ScaleJob {
    const double* src;
    double* dst;
    long long src_len, dst_len;
    long long shift;  // dst[i] = src[i - shift]
    double factor;
};
void launch_scale(const ScaleJob* d_jobs, int njobs, long long max_len, hipStream_t stream);

// A zero-padded window of a signal, `total` long: dst[i] = src[base + i - pad]
// for pad <= i < pad + len, else 0.
struct SliceJob {
    const double* src;
    double* dst;
    long long pad, base, len, total;
};
void launch_slices(const SliceJob* d_jobs, int njobs, long long max_len, hipStream_t stream);

// Speech-mode VAD chunk activity (vad.zig:31-48): per chunk c,
// RMS of trunc(clamp(x / max(ref) * 32768)) over frame_len samples starting
// at first + c·frame_len, compared with 5000.
struct VadJob {
    const double* ref;
    long long len;
    unsigned char* active;
    int n_chunks, frame_len, first, pad_;
};
void launch_vad(const VadJob* d_jobs, int njobs, int max_chunks, unsigned long long* d_scratch, hipStream_t stream);

// Spectrogram frames on the full-signal grid reused by a realigned patch
// (realign.zig:262-281): dst[b * nf + f] = src[b * src_stride + f].
struct FrameCopyJob {
    double* dst;
    const double* src;
    int src_stride, nf;
};
void launch_frame_copies(const FrameCopyJob* d_jobs, int njobs, int nb, int max_nf, hipStream_t stream);

// ---------------------------------------------------------------- spectrogram prep
struct PrepJob {
    double* ref;
    double* deg;
    int nfr, nfd;
};
// dB, per-frame floor, joint-minimum subtract (spectrogram.zig:69-80).
void launch_prep(const PrepJob* d_jobs, int njobs, int nb, int max_nf, unsigned long long* d_scratch,
                 hipStream_t stream);

// ---------------------------------------------------------------- NSIM
// Planes of [nb][width] doubles per degraded spectrogram: the zero-padded
// values and their squares, then the 3x3 window means and variances for the
// interior column set and for a patch's first and last columns (variances
// as signed square roots in the fast mode).
enum : int {
    kPlanePad = 0,
    kPlaneSq,
    kPlaneMu,
    kPlaneMuFirst,
    kPlaneMuLast,
    kPlaneVar,
    kPlaneVarFirst,
    kPlaneVarLast,
    kPlanes
};
struct PlaneJob {
    const double* spec;  // [nb][nf]
    double* plane;       // kPlanes planes of [nb][width]
    int nf, width;
};
void launch_planes(const PlaneJob* d_jobs, int njobs, int nb, int patch, int max_width, bool fast, hipStream_t stream);

struct SimJob {
    const double* ref_spec;  // [nb][nf_ref], comparison-ready dB
    const double* plane;     // from launch_planes
    double* sims;            // [np][nf_deg]
    const int* ref_idx;      // [np]
    const int* lo;           // [np]
    const int* hi;           // [np]
    int np, nf_ref, nf_deg, width;
};
// Patches of one pair scored together by a similarity-table block: they
// share the degraded signal's loads at each offset.
constexpr int kSimPatchGroup = 2;
// sims[p][off] for off in [lo[p], hi[p]] (accel.zig:126-146); d_terms holds
// 4 * nb * patch * max_np * njobs doubles of ref-patch terms; max_window is
// the widest union of the search windows of patches p..p+kSimPatchGroup-1
// (p a multiple of kSimPatchGroup).  `fast`: fused multiply-adds and one
// division per cell instead of the reference's order.
void launch_sim_table(const SimJob* d_jobs, int njobs, int nb, int patch, int max_np, int max_window, double* d_terms,
                      bool fast, hipStream_t stream);

struct DpJob {
    const double* sims;
    double* dp;     // [np][nf_deg], zeroed
    int* bt;        // [np][nf_deg], zeroed
    const int* lo;
    const int* hi;
    int* offsets;   // [np] chosen deg start frame
    int* state;     // [np] 0 untouched, 1 valid, 2 no real match
    int np, nf_deg;
};
// max_span: the longest scan range of any row (frames from the previous
// row's lo to the row's hi); from kDpWideSpan up a pair gets 1024 lanes.
constexpr int kDpWideSpan = 2048;
void launch_dp(const DpJob* d_jobs, int njobs, int max_span, hipStream_t stream);

// Per-band NSIM statistics of one patch pair (nsim.zig:427-510).
struct StatJob {
    const double* ref;  // ref spectrogram base
    const double* deg;  // deg spectrogram base
    int ref_stride, deg_stride;  // row strides (frames per band)
    int ref_col, deg_col;        // first column; deg columns >= deg_valid read 0
    int deg_valid;
    int ncols;
    double* out;  // [3*nb + 1]: means, stddevs, energy, mean of means
};
void launch_stats(const StatJob* d_jobs, int njobs, int nb, int max_cols, hipStream_t stream);

// ---------------------------------------------------------------- gradients
// One frame's slope (gammatone_bwd.slang:84-150): frame sample t is
// sig[start + t] for t in [lo, hi) and zero elsewhere (padding the forward
// added); band b's spectrogram slope is g_col[b * stride]; out receives the
// window's W sample slopes (zero outside [lo, hi)).  The host keeps
// start + lo >= 0 and start + hi within the signal.
struct FrameGradJob {
    const double* sig;
    long long start;
    const double* g_col;
    double* out;
    int stride, lo, hi, pad_;
};
// Samples per checkpoint of the adjoint's forward pass.
constexpr int kAdjSpan = 64;
// Doubles of checkpoint scratch a launch of njobs frames needs.
size_t gammatone_adjoint_scratch(int njobs, int window);
void launch_gammatone_adjoint(const FrameGradJob* d_jobs, int njobs, const BandCoef* d_coef, const double* d_hann,
                              int nb, int window, double* d_ckpt, hipStream_t stream);

// dst[i] = Σ over frames list[first .. first+count) (in that order) of
// frame.out[pos + i - frame.start] where that index is in the frame's
// [lo, hi), for i < len (gammatone_bwd.slang:153-183's overlap sum, in a
// fixed order).
struct AccRun {
    double* dst;
    long long pos;
    int first, count, len, pad_;
};
void launch_frame_accumulate(const AccRun* d_runs, int nruns, const int* d_list, const FrameGradJob* d_frames,
                             hipStream_t stream);

// Device math checks (log10, sqrt, division) against the host.
void launch_math_probe(int op, const double* a, const double* b, double* out, long long n, hipStream_t stream);

}  // namespace vq
