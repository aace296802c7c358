/* visqol.h — C ABI of the MI355X ViSQOL v3.3.3 scoring library.
 *
 * Drop-in for jdek/akvisqol's src/include/visqol.h (ABI version 1): the same
 * twelve functions, the same three structs with the same field order and
 * x86-64 layout (visqol_options 32 B, visqol_patch 64 B, visqol_result 88 B),
 * the same error codes and ownership rules.  Each declaration cites the
 * reference line it replaces (src/include/visqol.h and its implementation in
 * src/capi.zig).
 *
 * One additive value: VISQOL_BACKEND_HIP (4) names this library's backend.
 * The reference's four backends have no AMD slot, and reusing "cuda" would
 * make this a CUDA shim; ABI version 1 is kept because no existing value,
 * layout or signature changes.  visqol_options_default() keeps the
 * reference's default, VISQOL_BACKEND_CPU, which this library serves with
 * its exact f64 device path: that path's scores are the reference CPU
 * path's bit for bit, so callers that never name a backend (the reference's
 * Python and C++ wrappers with VISQOL_BACKEND unset) run unchanged.
 *
 * Threading: a visqol_ctx is used by one thread at a time;
 * visqol_measure_batch packs its pairs onto the GPU itself.
 */
#ifndef VISQOL_H
#define VISQOL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VISQOL_ABI_VERSION 1

/* Return codes (visqol.h:20-29; capi.zig:17-24). */
enum {
    VISQOL_OK = 0,
    VISQOL_E_INVALID_ARGUMENT = -1,
    VISQOL_E_OUT_OF_MEMORY = -2,
    VISQOL_E_IO = -3,
    VISQOL_E_MODEL = -4,
    VISQOL_E_SIGNAL = -5,
    VISQOL_E_BACKEND = -6,
    VISQOL_E_INTERNAL = -7
};

/* visqol.h:31 */
enum { VISQOL_MODE_AUDIO = 0, VISQOL_MODE_SPEECH = 1 };

/* visqol.h:33-38, plus the additive HIP slot. */
enum {
    VISQOL_BACKEND_CPU = 0,
    VISQOL_BACKEND_METAL = 1,
    VISQOL_BACKEND_CUDA = 2,
    VISQOL_BACKEND_VULKAN = 3,
    VISQOL_BACKEND_HIP = 4
};

/* visqol.h:43.  On this library: EXACT = f64 with the reference CPU path's
 * operation order (bit-compatible scores); FAST and BALANCED = the
 * throughput mode (f64 with fused multiply-add and parallel reductions,
 * lags nominated from an f32 cross-correlation and rescored in f64,
 * |dMOS| <= 1e-6); AUTO = EXACT.  VISQOL_BACKEND_CPU always runs EXACT,
 * as the reference's CPU backend computes in f64 whatever the precision. */
enum {
    VISQOL_PRECISION_EXACT = 0,
    VISQOL_PRECISION_FAST = 1,
    VISQOL_PRECISION_BALANCED = 2,
    VISQOL_PRECISION_AUTO = 3
};

/* visqol.h:45-57; capi.zig:28-38. */
typedef struct {
    uint32_t mode;
    int32_t search_window_radius; /* <= 0 means 60 */
    uint8_t use_unscaled_speech_mos_mapping;
    uint8_t disable_global_alignment;
    uint8_t disable_realignment;
    uint8_t use_lattice_model;
    const char *model_path;       /* NULL or "" selects the bundled model */
    uint32_t backend;
    uint32_t precision;
} visqol_options;

/* visqol.h:59-68; capi.zig:40-49. */
typedef struct {
    double similarity;
    double ref_patch_start_time;
    double ref_patch_end_time;
    double deg_patch_start_time;
    double deg_patch_end_time;
    const double *freq_band_means;
    const double *freq_band_stddevs;
    const double *freq_band_deg_energy;
} visqol_patch;

/* visqol.h:70-82; capi.zig:51-63. */
typedef struct {
    double moslqo;
    double vnsim;
    double alignment_lag_s;
    size_t num_bands;
    const double *fvnsim;
    const double *fvnsim10;
    const double *fstdnsim;
    const double *fvdegenergy;
    const double *center_freq_bands;
    size_t num_patches;
    const visqol_patch *patch_sims;
} visqol_result;

typedef struct visqol_ctx visqol_ctx;

/* visqol.h:86; capi.zig:124-126 */
uint32_t visqol_abi_version(void);
/* visqol.h:87; capi.zig:128-130 */
const char *visqol_version(void);
/* visqol.h:88; capi.zig:119-122: thread-local, <= 511 bytes. */
const char *visqol_last_error(void);
/* visqol.h:90; capi.zig:132-134: (1 << VISQOL_BACKEND_CPU) |
 * (1 << VISQOL_BACKEND_HIP) when a gfx950 device is present (both run on
 * it), else 0. */
uint32_t visqol_available_backends(void);
/* visqol.h:92; capi.zig:136-149: the reference's defaults, backend CPU. */
void visqol_options_default(visqol_options *opts);
/* visqol.h:94; capi.zig:155-223 */
int visqol_create(const visqol_options *opts, visqol_ctx **out);
/* visqol.h:95; capi.zig:225-230 */
void visqol_destroy(visqol_ctx *ctx);
/* visqol.h:98-102; capi.zig:280-298 */
int visqol_measure(visqol_ctx *ctx, const double *ref, size_t ref_len, const double *deg, size_t deg_len,
                   uint32_t sample_rate, visqol_result **out);
/* visqol.h:105-106; capi.zig:300-318 */
int visqol_measure_files(visqol_ctx *ctx, const char *ref_path, const char *deg_path, visqol_result **out);
/* visqol.h:110-114; capi.zig:331-372 */
int visqol_measure_batch(visqol_ctx *ctx, size_t n, const double *const *refs, const size_t *ref_lens,
                         const double *const *degs, const size_t *deg_lens, uint32_t sample_rate,
                         uint32_t num_threads, visqol_result **out, int32_t *status);
/* visqol.h:117-121; capi.zig:375-456 */
int visqol_measure_files_batch(visqol_ctx *ctx, size_t n, const char *const *ref_paths,
                               const char *const *deg_paths, uint32_t num_threads, visqol_result **out,
                               int32_t *status);
/* visqol.h:123; capi.zig:272-278 */
void visqol_result_free(visqol_result *result);

#ifdef __cplusplus
}
#endif

#endif /* VISQOL_H */
