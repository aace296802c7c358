/* ViSQOL v3.3.3 scoring path: CPU f64 oracle (test infrastructure only).
 *
 * This is a plain-C restatement of jdek/akvisqol's CPU path, operation for
 * operation, used ONLY as the checker by tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg.  The shipped library (akvisqol_amd/)
 * never links, loads or calls anything here.
 *
 * Reference files restated (paths relative to the reference checkout):
 *   src/visqol.zig, src/accel.zig, src/backend.zig (batch threading),
 *   src/gammatone.zig, src/dsp/erb.zig, src/spectrogram.zig,
 *   src/envelope.zig, src/xcorr.zig, src/dsp/fft.zig, src/realign.zig,
 *   src/dp.zig, src/vad.zig, src/nsim.zig, src/audio.zig (SPL),
 *   src/svr.zig, src/lattice.zig, src/tflite.zig.
 * Each function cites the file:line it follows.
 *
 * Build: gcc -O2 -ffp-contract=off (no FMA contraction, IEEE double).
 */
#ifndef VO_H
#define VO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VO_MAX_BANDS 32

enum { VO_OK = 0, VO_E_INVALID_ARGUMENT = -1, VO_E_OUT_OF_MEMORY = -2, VO_E_IO = -3,
       VO_E_MODEL = -4, VO_E_SIGNAL = -5, VO_E_BACKEND = -6, VO_E_INTERNAL = -7 };

/* gammatone.zig:17-30 */
typedef struct { double a0, a11, a12, a13, a14, b1, b2, gain; } vo_band;

/* Same layout as visqol.h's visqol_options (capi.zig:28-38). */
typedef struct {
    uint32_t mode;
    int32_t search_window_radius;
    uint8_t use_unscaled_speech_mos_mapping;
    uint8_t disable_global_alignment;
    uint8_t disable_realignment;
    uint8_t use_lattice_model;
    const char *model_path;
    uint32_t backend;
    uint32_t precision;
} vo_options;

/* Same layouts as visqol.h's visqol_patch / visqol_result (capi.zig:40-63). */
typedef struct {
    double similarity, ref_patch_start_time, ref_patch_end_time, deg_patch_start_time, deg_patch_end_time;
    const double *freq_band_means, *freq_band_stddevs, *freq_band_deg_energy;
} vo_patch;

typedef struct {
    double moslqo, vnsim, alignment_lag_s;
    size_t num_bands;
    const double *fvnsim, *fvnsim10, *fstdnsim, *fvdegenergy, *center_freq_bands;
    size_t num_patches;
    const vo_patch *patch_sims;
} vo_result;

/* Per-pair intermediates the parity tests compare stage by stage. */
typedef struct {
    int64_t lag;             /* global lag in samples (0 if alignment disabled) */
    double spl_factor;       /* deg scale factor applied (1 if skipped) */
    size_t nf_ref, nf_deg, num_patches;
    size_t ref_indices[512];
    int64_t realign_lag[512];     /* per match; INT64_MIN if not realigned */
    size_t realign_frames[512];
    uint8_t realign_accepted[512];
    uint8_t match_valid[512];
    size_t dp_offset[512];        /* chosen deg start frame per patch */
} vo_trace;

typedef struct vo_ctx vo_ctx;

const char *vo_last_error(void);
void vo_options_default(vo_options *o);
int vo_create(const vo_options *o, vo_ctx **out);
void vo_destroy(vo_ctx *c);
int vo_measure(vo_ctx *c, const double *ref, size_t ref_len, const double *deg, size_t deg_len,
               uint32_t sample_rate, vo_result **out);
int vo_measure_trace(vo_ctx *c, const double *ref, size_t ref_len, const double *deg, size_t deg_len,
                     uint32_t sample_rate, vo_result **out, vo_trace *trace,
                     double *ref_spec_out, double *deg_spec_out, double *sims_out);
int vo_measure_batch(vo_ctx *c, size_t n, const double *const *refs, const size_t *ref_lens,
                     const double *const *degs, const size_t *deg_lens, uint32_t sample_rate,
                     uint32_t num_threads, vo_result **out, int32_t *status);
void vo_result_free(vo_result *r);
/* buildResult and the mapping over given matches ([n][nb] per-band means,
 * stddevs, deg energies; reference patch times): aggregation checks. */
int vo_build_result_from(vo_ctx *c, size_t n, const double *means, const double *stddevs, const double *energy,
                         const double *ref_start, const double *ref_end, double frame_duration, vo_result **out);

/* ---- stage entry points (unit tests / per-stage parity) ---- */
double vo_log10(double x);
void vo_center_freqs_desc(size_t n, double low, double high, double *out);
void vo_filters(uint32_t sample_rate, size_t nbands, double low, double high, vo_band *bands, double *cf_asc);
void vo_filter_band(const vo_band *band, const double *in, size_t n, double *out);
double vo_band_rms_fresh(const vo_band *band, const double *frame, size_t n);
void vo_hann(size_t size, double *out);
size_t vo_num_frames(size_t len, size_t window, size_t hop);
int vo_spectrogram(const vo_band *bands, size_t nbands, const double *signal, size_t len,
                   size_t window, size_t hop, const double *hann, double *out);
void vo_db_floor(double *data, size_t n, double floor_db);
void vo_raise_floor_per_frame(double *ref, double *deg, size_t nb, size_t nfr, size_t nfd, double rel);
void vo_subtract_joint_min(double *ref, size_t nr, double *deg, size_t nd);
void vo_prepare_for_comparison(double *ref, double *deg, size_t nb, size_t nfr, size_t nfd);
void vo_upper_envelope(const double *s, size_t n, double *out);
size_t vo_fft_size_for(size_t biggest);
void vo_fft_dif_forward(double *z_interleaved, size_t n);
void vo_fft_dit_inverse(double *z_interleaved, size_t n);
void vo_fft_natural(double *z_interleaved, size_t n, int inverse);
int64_t vo_best_lag(const double *s1, size_t l1, const double *s2, size_t l2);
/* Re of the circular cross-correlation behind vo_best_lag: n =
 * vo_fft_size_for(max(l1, l2)) values written to out. */
int vo_xcorr_re(const double *s1, size_t l1, const double *s2, size_t l2, double *out);
int64_t vo_best_lag_reference(const double *s1, size_t l1, const double *s2, size_t l2);
int64_t vo_best_lag_direct(const double *s1, size_t l1, const double *s2, size_t l2);
double vo_spl(const double *x, size_t n);
void vo_scale_to_match_spl(const double *ref, size_t nr, double *deg, size_t nd);
void vo_conv2d3x3(const double *in, double *out, size_t nrows, size_t ncols);
void vo_patch_similarity(const double *ref, const double *deg, size_t nrows, size_t ncols,
                         double *means, double *stddevs, double *energy, double *mean_of_means);
double vo_similarity_scalar(const double *ref, const double *deg, size_t nrows, size_t ncols);
void vo_similarity_table(const double *ref_spec, size_t nf_ref, const double *deg_spec, size_t nf_deg,
                         size_t nbands, size_t patch, const size_t *ref_idx, const size_t *lo,
                         const size_t *hi, size_t np, double *sims);
size_t vo_ref_patch_indices(size_t nf, size_t patch, size_t *out, size_t cap);
size_t vo_calc_max_num_patches(const size_t *idx, size_t n, size_t nf_deg, size_t patch);
size_t vo_vad_patch_indices(const double *ref, size_t len, size_t nf, size_t patch, size_t window,
                            double overlap, size_t *out, size_t cap);
/* DP over a precomputed similarity table; writes the chosen deg offsets and
 * validity per patch (dp.zig:205-355, the discrete part). */
double vo_dp_search(const double *sims, size_t np, size_t nf_deg, const size_t *lo, const size_t *hi,
                  size_t *offsets, uint8_t *valid);
/* Whole findOptimalDegPatches on explicit patch indices (dp.zig:358-387). */
int vo_find_optimal_patches(const double *ref_spec, const size_t *ref_idx, size_t n_idx,
                            const double *deg_spec, size_t nb, size_t nf_ref, size_t nf_deg,
                            size_t patch, double frame_duration, int32_t radius, double *deg_start_times);
void vo_slice_layout(size_t len, uint32_t sr, double start, double end, size_t *pad, size_t *base, size_t *n);
int64_t vo_lag_samples_back(int64_t lag, uint32_t sr);
void vo_trim_geometry(size_t ref_len, size_t deg_len, int64_t lag, uint32_t sr, size_t *len,
                      size_t *ref_off, int64_t *deg_off, double *lag_seconds);
/* SVR over the model text in `model` (svr.zig:37-165). */
int vo_svr_predict(const char *model, size_t model_len, const double *obs, size_t n, double *y);
/* Lattice over the tflite bytes (lattice.zig:49-71). */
int vo_lattice_predict(const uint8_t *bytes, size_t len, const double *fvnsim, const double *fvnsim10,
                       const double *fstdnsim, const double *fvdegenergy, size_t nb, double *y);
int vo_lattice_op_count(const uint8_t *bytes, size_t len);

#ifdef __cplusplus
}
#endif
#endif
