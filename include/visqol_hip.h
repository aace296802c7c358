/* visqol_hip.h — additive entry points of the MI355X ViSQOL library.
 *
 * None of these has a counterpart in jdek/akvisqol's visqol.h; they are
 * for callers that keep signals in device memory (benchmarks, serving
 * loops) and for stage-by-stage parity checks of the device path.  All
 * take plain pointers and sizes; results use visqol.h's types.
 */
#ifndef VISQOL_HIP_H
#define VISQOL_HIP_H

#include "visqol.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Discrete decisions and shapes of one scored pair: the global lag, the
 * SPL factor applied to the degraded signal, frame counts, reference patch
 * starts, the DP's chosen degraded starts and validity, and per match the
 * realignment lag (INT64_MIN when not realigned), realigned frame count
 * and whether the realigned score replaced the DP score. */
typedef struct {
    int64_t lag;
    double spl_factor;
    size_t nf_ref, nf_deg, num_patches;
    size_t ref_indices[512];
    int64_t realign_lag[512];
    size_t realign_frames[512];
    uint8_t realign_accepted[512];
    uint8_t match_valid[512];
    size_t dp_offset[512];
} visqol_hip_trace;

/* visqol_measure plus the trace and, when non-NULL, the linear gammatone
 * spectrograms ([bands][frames], nf_ref / nf_deg frames) and the DP
 * similarity table ([patches][nf_deg]).  The output buffers must hold
 * bands * (ref_len / hop + 1), bands * ((deg_len + ref_len / 2) / hop + 1)
 * and (ref_len / hop / patch + 2) * ((deg_len + ref_len / 2) / hop + 1)
 * doubles. */
int visqol_hip_measure_trace(visqol_ctx *ctx, const double *ref, size_t ref_len, const double *deg, size_t deg_len,
                             uint32_t sample_rate, visqol_result **out, visqol_hip_trace *trace,
                             double *ref_spec_out, double *deg_spec_out, double *sims_out);

/* visqol_measure_batch over signals already resident in device memory
 * (hipMalloc'd on the context's device): no host-to-device copy of the
 * samples.  Per-pair failures land in status[i]; returns the first. */
int visqol_hip_measure_batch_device(visqol_ctx *ctx, size_t n, const double *const *dev_refs, const size_t *ref_lens,
                                    const double *const *dev_degs, const size_t *deg_lens, uint32_t sample_rate,
                                    visqol_result **out, int32_t *status);

/* One batch of pairs of several contexts (modes, options) and sample
 * rates, signals resident in device memory: pair i is scored by context
 * ctxs[ctx_index[i]] at sample_rates[i].  Each context scores its pairs
 * (one batch per sample rate) on a host thread and device streams of its
 * own, so the contexts' batches run on the device at once: 48 kHz audio
 * and 16 kHz speech pairs of one uneven mix in one call.  Results equal
 * each context's visqol_hip_measure_batch_device on its own pairs.  The
 * contexts must be distinct and not in use elsewhere during the call.
 * Per-pair failures land in status[i]; returns the first failure in pair
 * order. */
int visqol_hip_measure_mixed_batch_device(visqol_ctx *const *ctxs, size_t num_ctxs, size_t n,
                                          const uint32_t *ctx_index, const double *const *dev_refs,
                                          const size_t *ref_lens, const double *const *dev_degs,
                                          const size_t *deg_lens, const uint32_t *sample_rates,
                                          visqol_result **out, int32_t *status);

/* Per-context kernel timing: when enabled, HIP events on the context's
 * stream bracket every full-signal gammatone launch; *ms is their summed
 * device time, *flop the algorithmic work (30 flop per sample and band of
 * every frame), *bytes the signal read plus spectrogram written, *launches
 * the launch count since the last reset. */
int visqol_hip_profile(visqol_ctx *ctx, int enable, int reset);
int visqol_hip_profile_read(visqol_ctx *ctx, double *ms, double *flop, double *bytes, int64_t *launches);
/* The same for the similarity-table launches: *ms device time, *cells the
 * cells evaluated (search offsets x bands x patch frames), *launches. */
int visqol_hip_profile_sim(visqol_ctx *ctx, double *ms, double *cells, int64_t *launches);
/* Host wall time (ms) of the stretches of batch calls since the last reset,
 * writing min(n, 8) values: phase A to its readback, host planning, phase B
 * to the DP readback, match statistics, their decoding, realignment lag
 * searches, realignment spectrograms and statistics, final aggregation.
 * Returns the number of stretches available (8). */
int visqol_hip_profile_phases(visqol_ctx *ctx, double *ms, int n);

/* References prepared once for scoring many degraded signals against them
 * (backend.zig:84-99 prepare / measurePrepared; visqol.zig:194-244
 * PreparedRef): each reference's samples and raw spectrogram stay in device
 * memory with its sums, energy and (speech) VAD patch indices, computed by
 * the same kernels as an unprepared pair, so scores are identical to
 * visqol_measure_batch on the same pairs while the reference's spectrogram
 * is not recomputed.  visqol_hip_measure_prepared scores degs[i] (host
 * arrays) against reference ref_index[i] (ref_index NULL: reference i) at
 * the prepared sample rate, with visqol_measure_batch's per-pair status and
 * return contract.  The context must have the preparing context's mode and
 * precision. */
typedef struct visqol_hip_prepared visqol_hip_prepared;
int visqol_hip_prepare(visqol_ctx *ctx, size_t n, const double *const *refs, const size_t *ref_lens,
                       uint32_t sample_rate, visqol_hip_prepared **out);
int visqol_hip_measure_prepared(visqol_ctx *ctx, const visqol_hip_prepared *prep, size_t n,
                                const size_t *ref_index, const double *const *degs, const size_t *deg_lens,
                                visqol_result **out, int32_t *status);
void visqol_hip_prepared_free(visqol_hip_prepared *prep);

/* Gradients (backend.zig:101-110 measureBatchGrad, :93-99 measurePrepared
 * with grads): visqol_measure_batch's scores, and into grads[i] (deg_lens[i]
 * values, caller-allocated host memory; NULL skips the pair) ∂moslqo/∂deg[i]
 * with every discrete choice of the forward pass held fixed (the global and
 * realignment lags, the matched patches, which realignments are accepted,
 * the cells the dB floors pin).  A pair that fails, or whose score is
 * clamped, gets zeros.  Both precisions compute in f64. */
int visqol_hip_measure_batch_grad(visqol_ctx *ctx, size_t n, const double *const *refs, const size_t *ref_lens,
                                  const double *const *degs, const size_t *deg_lens, uint32_t sample_rate,
                                  visqol_result **out, int32_t *status, double *const *grads);
int visqol_hip_measure_prepared_grad(visqol_ctx *ctx, const visqol_hip_prepared *prep, size_t n,
                                     const size_t *ref_index, const double *const *degs, const size_t *deg_lens,
                                     visqol_result **out, int32_t *status, double *const *grads);

/* Continuous-batching server (no counterpart in the reference's C ABI; a
 * serving front end for callers with a stream of requests, in the manner of
 * the reference's batch lanes, backend.zig:182-221).
 *
 * visqol_hip_server_submit copies the request's signals to device memory
 * and queues it; it blocks while admitting it would exceed
 * max_inflight_bytes of estimated device footprint (queued plus running
 * requests) or the slab_bytes signal slab is full.  `workers` threads, each
 * with its own engines and HIP stream, take the queued requests of one
 * (mode, sample rate) at a time, up to max_batch, and score them as a batch
 * while other workers' batches run.  visqol_hip_server_wait blocks until the
 * ticket's result is ready, returns it as visqol_measure would (the caller
 * frees it with visqol_result_free), and gives the time from the submit
 * call to completion.  Each ticket is waited for exactly once.  submit and
 * wait may be called from any threads; destroy must not overlap them.
 * Options as for visqol_create, backend HIP; mode comes with each request;
 * model_path must be NULL (the embedded models serve both modes). */
typedef struct visqol_hip_server visqol_hip_server;
typedef struct {
    uint32_t workers;            /* concurrent batches, default 3 (1..16) */
    uint32_t max_batch;          /* pairs per batch, default 256 */
    uint64_t slab_bytes;         /* device memory for queued signals, default 16 GiB */
    uint64_t max_inflight_bytes; /* admission budget, default 96 GiB */
} visqol_hip_server_config;
void visqol_hip_server_config_default(visqol_hip_server_config *cfg);
int visqol_hip_server_create(const visqol_options *opts, const visqol_hip_server_config *cfg,
                             visqol_hip_server **out);
int visqol_hip_server_submit(visqol_hip_server *s, uint32_t mode, const double *ref, size_t ref_len,
                             const double *deg, size_t deg_len, uint32_t sample_rate, uint64_t *ticket);
int visqol_hip_server_wait(visqol_hip_server *s, uint64_t ticket, visqol_result **out, double *latency_ms);
/* visqol_hip_server_wait with where the request's time went, each on the
 * steady clock in ms: waiting for admission (footprint budget and slab
 * space), uploading its signals, queued until a worker took it into a
 * batch, that batch's run, and the total (submit call to completion), plus
 * the size of the batch it ran in. */
typedef struct {
    double admit_ms, upload_ms, queue_ms, run_ms, total_ms;
    uint32_t batch_size;
} visqol_hip_request_timing;
int visqol_hip_server_wait_timed(visqol_hip_server *s, uint64_t ticket, visqol_result **out,
                                 visqol_hip_request_timing *timing);
/* Batches run, pairs scored, and the footprint now admitted. */
int visqol_hip_server_stats(visqol_hip_server *s, uint64_t *batches, uint64_t *pairs, uint64_t *inflight_bytes);
void visqol_hip_server_destroy(visqol_hip_server *s);

/* Device-resident buffers for the batch entry point above. */
/* Host-side pieces of the gradient, for checks: measurePatchSimilarity's
 * per-band (mean, stddev, deg energy) slopes (any of the three may be NULL)
 * taken back to the comparison-ready ref and deg patch cells (row-major
 * [nrows][ncols]), and prepareForComparison's prepared-spectrogram slopes
 * taken back to the raw spectrograms ([nb][nfr] and [nb][nfd]); the outputs
 * are overwritten. */
int visqol_hip_patch_similarity_grad(const double *ref, const double *deg, size_t nrows, size_t ncols,
                                     const double *g_means, const double *g_stddevs, const double *g_energy,
                                     double *g_ref, double *g_deg);
int visqol_hip_prepare_grad(const double *raw_ref, const double *raw_deg, size_t nb, size_t nfr, size_t nfd,
                            const double *g_ref_prep, const double *g_deg_prep, double *g_raw_ref, double *g_raw_deg);

/* The TF-Lattice quality mapping on the host (lattice.zig:49-71): the
 * model's flatbuffer bytes, then `count` feature vectors of nb bands each
 * (row k of each array at k·nb), one MOS per vector into y. */
int visqol_hip_lattice_predict(const uint8_t *model, size_t len, size_t nb, size_t count, const double *fvnsim,
                               const double *fvnsim10, const double *fstdnsim, const double *fvdegenergy, double *y);

/* Spectrogram kernel block shapes.  visqol_hip_gammatone_shape describes
 * shape `shape` (waves per block, bands per lane, samples per staged chunk;
 * 64 frames per block) and returns the number of shapes.  Every shape gives
 * bit-identical spectrograms; an engine times them all on its first
 * whole-signal launch large enough to fill the device and keeps the fastest
 * for the process (per device, precision, band count and window; kind 0:
 * whole signals; kind 1, realignment slices, always runs shape 0).  visqol_hip_gammatone_tuning reports that choice
 * (*chosen = -1 while untuned) and, if `ms` holds `cap` floats, the timed
 * launch of each shape; it returns the number of shapes.  The environment
 * variable VISQOL_HIP_GAMMATONE_SHAPE=k, read when a context is created,
 * forces shape k for that context. */
int visqol_hip_gammatone_shape(int shape, int *waves, int *bands_per_lane, int *chunk);
int visqol_hip_gammatone_tuning(visqol_ctx *ctx, uint32_t sample_rate, int kind, int *chosen, float *ms, int cap);

/* WAV decoding as the file entry points do it (audio.zig:48-130: RIFF/WAVE,
 * PCM 8/16/24/32-bit, float 32/64-bit, WAVE_FORMAT_EXTENSIBLE, channel
 * mean), host only.  Sets *n_samples (frames) and *sample_rate; copies the
 * mono samples when `samples` is non-NULL and holds `cap` >= frames.  Errors
 * are the file entry points' codes (E_IO for format errors, E_SIGNAL for
 * empty audio or zero channels) with the reference's error name as the
 * last-error message. */
int visqol_hip_decode_wav(const uint8_t *data, size_t len, double *samples, size_t cap, size_t *n_samples,
                          uint32_t *sample_rate);

int visqol_hip_alloc(size_t bytes, void **dev_ptr);
int visqol_hip_free(void *dev_ptr);
int visqol_hip_upload(void *dev_dst, const void *host_src, size_t bytes);

/* Stage checks, each on host arrays. */
/* Linear gammatone spectrogram of one signal ([bands][frames]); precision
 * as visqol.h; returns the frame count through *num_frames. */
int visqol_hip_spectrogram(uint32_t mode, uint32_t sample_rate, uint32_t precision, const double *signal, size_t len,
                           double *out, size_t *num_frames);
/* The spectrogram's adjoint on one frame (gammatone_bwd.slang:84-150):
 * out[t] = ∂(Σ_b g[b] · rms_b)/∂frame[t] for the `window` samples of one
 * frame (window must be the sample rate's frame length) with the
 * filterbank's band slopes g; samples outside [lo, hi) read as zero and get
 * zero slope. */
int visqol_hip_frame_grad(uint32_t mode, uint32_t sample_rate, const double *frame, size_t window, const double *g,
                          size_t lo, size_t hi, double *out);
/* Filter-bank design shared with the device: 8 doubles per band
 * (a0 a11 a12 a13 a14 b1 b2 gain), bands ascending, plus centre
 * frequencies. */
int visqol_hip_filters(uint32_t mode, uint32_t sample_rate, double *bands8, double *center_freqs);
/* FFT cross-correlation lag of the upper envelopes of a and b. */
int visqol_hip_best_lag(const double *a, size_t la, const double *b, size_t lb, int64_t *lag);
/* One reference-order (left-to-right, f64) sum over x[0..len) on the
 * device, as the exact path forms it (envelope.zig:23-53,
 * realign.zig:21-51): kind 0 sum of x, 1 sum of x*x, 2 the alternating
 * sum of +-(x[i] - plain_sum/len) (+ for even i), 3 sum of x*x from
 * sample -lag when -half <= lag < 0, else from 0.  form 0: one serial
 * chain; 1: the speculative segmented form; 2: as the scoring path chooses.
 * Every form gives the same bits. */
int visqol_hip_sum(int kind, const double *x, size_t len, double plain_sum, int64_t lag, size_t half, int form,
                   double *out);
/* visqol_hip_best_lag on signals in device memory (dev_a, dev_b on the
 * current device), allocating nothing.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under graph
 * capture, serving loops that forbid hipMalloc).  Its work buffers are a
 * static device array of 16 MiB + 20 KiB per device (plus a 16 MiB twiddle
 * table), held until the process ends; signals of more than 2^19 samples
 * (a transform beyond 2^20 points) are rejected with
 * VISQOL_E_INVALID_ARGUMENT.  Calls that overlap on one device race (they
 * share the array); the lag equals visqol_hip_best_lag's. */
int visqol_hip_best_lag_noalloc(const double *dev_a, size_t la, const double *dev_b, size_t lb, int64_t *lag);
/* The values visqol_hip_best_lag maximises: Re of the circular
 * cross-correlation of the two upper envelopes, n = the power of two
 * >= 2·max(la, lb) - 1 (max(la, lb) >= 2).  Writes min(n, cap) values to
 * out and stores n in *n_out. */
int visqol_hip_xcorr(const double *a, size_t la, const double *b, size_t lb, double *out, size_t cap, size_t *n_out);
/* The throughput mode's lag search on one pair (xcorr.zig:269-340
 * refineLag): the lags whose f32 correlation lies within the f32 error
 * bound of the maximum are rescored by f64 dot products of the envelopes,
 * in scan order, first maximum.  cap (0..64; negative: the default 64,
 * MAX_REFINE_CANDIDATES) bounds the lags rescored: a pair with more, or of
 * a transform size outside 2^14..2^30, takes the f64 search instead
 * (*nominated = 0). */
int visqol_hip_best_lag_fast(const double *a, size_t la, const double *b, size_t lb, int cap, int64_t *lag,
                             int *nominated);
/* The f32 correlation that search nominates from: Re of the circular
 * cross-correlation of the two envelopes times n (n as visqol_hip_xcorr);
 * *n_out = 0 when n is outside 2^14..2^30. */
int visqol_hip_xcorr32(const double *a, size_t la, const double *b, size_t lb, float *out, size_t cap,
                       size_t *n_out);
/* The device radix-2 transform in place on n interleaved complex values
 * (n a power of two): dit = 0 runs the DIF forward transform (natural in,
 * bit-reversed out), dit = 1 the DIT inverse (bit-reversed in, natural out),
 * both unscaled. */
int visqol_hip_fft(double *z, size_t n, int dit);
/* Device arithmetic probes: op 0 log10(a), 1 sqrt(a), 2 a / b. */
int visqol_hip_math_probe(int op, const double *a, const double *b, double *out, size_t n);

#ifdef __cplusplus
}
#endif

#endif /* VISQOL_HIP_H */
