/* Oracle pipeline: per-pair scoring (visqol.zig:394-709), the batch driver
 * with one worker per thread and an atomic work index (backend.zig:174-221),
 * and the context / result plumbing of the C ABI (capi.zig).
 * Test infrastructure only (see vo.h). */
#include "vo_internal.h"

#include <dlfcn.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdatomic.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

static __thread char g_err[512];

void vo_set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}
const char *vo_last_error(void) { return g_err; }

/* visqol.zig:18-30 */
#define NUM_BANDS_AUDIO 32
#define NUM_BANDS_SPEECH 21
#define PATCH_AUDIO 30
#define PATCH_SPEECH 20

enum { MAP_SVR, MAP_EXP, MAP_LATTICE };

struct vo_ctx {
    vo_options opts;
    int mapper;
    vo_svr *svr;
    vo_lattice *lat;
    size_t nb, patch;
    double max_freq; /* < 0: Nyquist */
};

/* Per-rate cache (visqol.zig:346-359 ensureCache). */
typedef struct {
    uint32_t sr;
    size_t window, hop;
    double frame_duration;
    vo_band bands[VO_MAX_BANDS];
    double cf[VO_MAX_BANDS];
    double *hann;
} rate_cache;

static int rate_cache_init(rate_cache *rc, const struct vo_ctx *c, uint32_t sr) {
    const double sr_f = (double)sr;
    rc->sr = sr;
    vo_filters(sr, c->nb, 50.0, c->max_freq < 0 ? sr_f / 2.0 : c->max_freq, rc->bands, rc->cf);
    rc->window = (size_t)round(sr_f * 0.08);
    rc->hop = (size_t)((double)rc->window * 0.25);
    if (rc->hop == 0) rc->hop = 1;
    rc->hann = (double *)malloc(sizeof(double) * (rc->window ? rc->window : 1));
    if (!rc->hann) return VO_E_OUT_OF_MEMORY;
    vo_hann(rc->window, rc->hann);
    rc->frame_duration = (double)rc->hop / sr_f;
    return VO_OK;
}

static char *read_file(const char *path, size_t *len) {
    FILE *f = fopen(path, "rb");
    if (!f) return NULL;
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    char *buf = (char *)malloc((size_t)n + 1);
    if (fread(buf, 1, (size_t)n, f) != (size_t)n) { fclose(f); free(buf); return NULL; }
    fclose(f);
    buf[n] = 0;
    *len = (size_t)n;
    return buf;
}

/* The default models live beside the shipped package; this library sits in
 * <repo>/oracle/build/, the models in <repo>/akvisqol_amd/models/. */
static void default_model_path(int lattice, char *out, size_t cap) {
    Dl_info info;
    char dir[1024] = ".";
    if (dladdr((void *)default_model_path, &info) && info.dli_fname) {
        snprintf(dir, sizeof dir, "%s", info.dli_fname);
        char *s = strrchr(dir, '/');
        if (s) *s = 0; else snprintf(dir, sizeof dir, ".");
    }
    snprintf(out, cap, "%s/../../akvisqol_amd/models/%s", dir,
             lattice ? "lattice_speech.tflite" : "libsvm_nu_svr_model.txt");
}

void vo_options_default(vo_options *o) { /* capi.zig:136-149 */
    if (!o) return;
    memset(o, 0, sizeof *o);
    o->search_window_radius = 60;
    o->use_lattice_model = 1;
    o->precision = 3;
}

int vo_create(const vo_options *in, vo_ctx **out) {
    if (!out) { vo_set_error("out is NULL"); return VO_E_INVALID_ARGUMENT; }
    *out = NULL;
    vo_options o;
    vo_options_default(&o);
    if (in) o = *in;
    if (o.mode > 1) { vo_set_error("invalid mode %u", o.mode); return VO_E_INVALID_ARGUMENT; }
    if (o.search_window_radius <= 0) o.search_window_radius = 60;
    struct vo_ctx *c = (struct vo_ctx *)calloc(1, sizeof *c);
    c->opts = o;
    c->nb = o.mode == 0 ? NUM_BANDS_AUDIO : NUM_BANDS_SPEECH;
    c->patch = o.mode == 0 ? PATCH_AUDIO : PATCH_SPEECH;
    c->max_freq = o.mode == 0 ? -1.0 : 8000.0;
    int need_model = o.mode == 0 || o.use_lattice_model;
    c->mapper = o.mode == 0 ? MAP_SVR : (o.use_lattice_model ? MAP_LATTICE : MAP_EXP);
    if (need_model) {
        char path[1200];
        if (o.model_path && o.model_path[0]) snprintf(path, sizeof path, "%s", o.model_path);
        else default_model_path(c->mapper == MAP_LATTICE, path, sizeof path);
        size_t len = 0;
        char *text = read_file(path, &len);
        if (!text) { vo_set_error("cannot read model '%s'", path); free(c); return VO_E_IO; }
        int err = VO_OK;
        if (c->mapper == MAP_SVR) c->svr = vo_svr_parse(text, len, &err);
        else c->lat = vo_lattice_load((const uint8_t *)text, len, NUM_BANDS_SPEECH, &err);
        free(text);
        if (err != VO_OK) { vo_set_error("invalid model '%s'", path); free(c); return err; }
    }
    *out = c;
    return VO_OK;
}

void vo_destroy(vo_ctx *c) {
    if (!c) return;
    vo_svr_free(c->svr);
    vo_lattice_free(c->lat);
    free(c);
}

/* visqol.zig:81-98 (QualityMapper.predict) */
static int map_predict(const struct vo_ctx *c, const double *fvnsim, const double *fvnsim10, const double *fstdnsim,
                       const double *fvdegenergy, size_t nb, double *mos) {
    if (c->mapper == MAP_SVR) {
        double y = vo_svr_eval(c->svr, fvnsim, nb);
        *mos = fmax(1.0, fmin(y, 5.0)); /* std.math.clamp */
        return VO_OK;
    }
    if (c->mapper == MAP_LATTICE) return vo_lattice_eval(c->lat, fvnsim, fvnsim10, fstdnsim, fvdegenergy, mos);
    const double a = -262.847869, b = 0.0154302525, x0 = -361.063949;
    double sum = 0.0;
    for (size_t i = 0; i < nb; i++) sum += fvnsim[i];
    double nsim_mean = sum / (double)nb;
    double m = a + exp(b * (nsim_mean - x0));
    double scale = c->opts.use_unscaled_speech_mos_mapping ? 1.0 : (double)1.245063f;
    double v = m * scale;
    *mos = fmax(1.0, fmin(v, 5.0));
    return VO_OK;
}

static int cmp_dbl(const void *a, const void *b) {
    double x = *(const double *)a, y = *(const double *)b;
    return x < y ? -1 : x > y ? 1 : 0;
}

/* visqol.zig:595-709 (buildResult) */
static int build_result(const struct vo_ctx *c, const vo_match *items, size_t n, double frame_duration,
                        const double *cf, double lag_s, vo_result **out) {
    const size_t nb = c->nb;
    const double n_f = (double)n;
    size_t sz = sizeof(vo_result) + sizeof(double) * 5 * nb + sizeof(vo_patch) * n + sizeof(double) * n * 3 * nb;
    char *blk = (char *)calloc(1, sz);
    if (!blk) return VO_E_OUT_OF_MEMORY;
    vo_result *r = (vo_result *)blk;
    double *fvnsim = (double *)(blk + sizeof(vo_result));
    double *fvnsim10 = fvnsim + nb, *fstdnsim = fvnsim10 + nb, *fvdeg = fstdnsim + nb, *cfb = fvdeg + nb;
    vo_patch *patches = (vo_patch *)(cfb + nb);
    double *band_data = (double *)(patches + n);
    memcpy(cfb, cf, sizeof(double) * nb);
    for (size_t i = 0; i < n; i++)
        for (size_t b = 0; b < nb; b++) {
            fvnsim[b] += items[i].means[b];
            fvdeg[b] += items[i].energy[b];
        }
    for (size_t b = 0; b < nb; b++) { fvnsim[b] /= n_f; fvdeg[b] /= n_f; }
    {
        double *sorted = (double *)malloc(sizeof(double) * (n ? n : 1));
        size_t k = (size_t)trunc(n_f * 0.10);
        if (k < 1) k = 1;
        for (size_t b = 0; b < nb; b++) {
            for (size_t i = 0; i < n; i++) sorted[i] = items[i].means[b];
            qsort(sorted, n, sizeof(double), cmp_dbl);
            double acc = 0.0;
            for (size_t i = 0; i < k; i++) acc += sorted[i];
            fvnsim10[b] = acc / (double)k;
        }
        free(sorted);
    }
    {
        double contrib[VO_MAX_BANDS] = {0};
        int64_t total = 0;
        for (size_t i = 0; i < n; i++) {
            double secs = items[i].ref_end - items[i].ref_start;
            int64_t fc_i = (int64_t)ceil(secs / frame_duration);
            total += fc_i;
            double fc = (double)fc_i;
            for (size_t b = 0; b < nb; b++) {
                double dev = items[i].stddevs[b], mean = items[i].means[b];
                contrib[b] += (fc - 1.0) * dev * dev;
                contrib[b] += fc * mean * mean;
            }
        }
        double total_f = (double)total;
        for (size_t b = 0; b < nb; b++) {
            double v = (contrib[b] - fvnsim[b] * fvnsim[b] * total_f) / (total_f - 1.0);
            fstdnsim[b] = v < 0.0 ? 0.0 : sqrt(v);
        }
    }
    double sum = 0.0;
    for (size_t b = 0; b < nb; b++) sum += fvnsim[b];
    double vnsim = sum / (double)nb;
    double mos;
    int rc = map_predict(c, fvnsim, fvnsim10, fstdnsim, fvdeg, nb, &mos);
    if (rc) { free(blk); vo_set_error("mapping model failed"); return rc; }
    if (vnsim < 0.15) mos = 1.0;
    for (size_t i = 0; i < n; i++) {
        double *dst = band_data + i * nb * 3;
        memcpy(dst, items[i].means, sizeof(double) * nb);
        memcpy(dst + nb, items[i].stddevs, sizeof(double) * nb);
        memcpy(dst + 2 * nb, items[i].energy, sizeof(double) * nb);
        patches[i] = (vo_patch){items[i].similarity, items[i].ref_start, items[i].ref_end, items[i].deg_start,
                                items[i].deg_end, dst, dst + nb, dst + 2 * nb};
    }
    r->moslqo = mos; r->vnsim = vnsim; r->alignment_lag_s = lag_s; r->num_bands = nb;
    r->fvnsim = fvnsim; r->fvnsim10 = fvnsim10; r->fstdnsim = fstdnsim; r->fvdegenergy = fvdeg;
    r->center_freq_bands = cfb; r->num_patches = n; r->patch_sims = patches;
    *out = r;
    return VO_OK;
}

void vo_result_free(vo_result *r) { free(r); }

/* buildResult (visqol.zig:595-709) over caller-given matches, for checks of
 * the aggregation and the mappings: n matches of the context's band count,
 * per-band means / stddevs / deg energies row-major [n][nb], each match's
 * reference start and end times; centre frequencies and lag read zero. */
int vo_build_result_from(vo_ctx *c, size_t n, const double *means, const double *stddevs, const double *energy,
                         const double *ref_start, const double *ref_end, double frame_duration, vo_result **out) {
    if (!c || !out || n == 0 || !means || !stddevs || !energy || !ref_start || !ref_end) return VO_E_INVALID_ARGUMENT;
    const size_t nb = c->nb;
    vo_match *items = (vo_match *)calloc(n, sizeof(vo_match));
    if (!items) return VO_E_OUT_OF_MEMORY;
    for (size_t i = 0; i < n; i++) {
        memcpy(items[i].means, means + i * nb, sizeof(double) * nb);
        memcpy(items[i].stddevs, stddevs + i * nb, sizeof(double) * nb);
        memcpy(items[i].energy, energy + i * nb, sizeof(double) * nb);
        items[i].ref_start = ref_start[i];
        items[i].ref_end = ref_end[i];
        items[i].valid = 1;
    }
    double cf[VO_MAX_BANDS] = {0};
    const int rc = build_result(c, items, n, frame_duration, cf, 0.0, out);
    free(items);
    return rc;
}

/* realign.zig:262-281 (RawSpectrograms.reuse) */
static int reuse_frames(size_t nb, const double *global, size_t nf_global, vo_layout l, int64_t off, size_t hop,
                        size_t nf_local, double *out) {
    if (off < (int64_t)l.pad) return 0;
    size_t start = l.base + (size_t)off - l.pad;
    if (start % hop != 0) return 0;
    size_t g0 = start / hop;
    if (g0 + nf_local > nf_global) return 0;
    for (size_t b = 0; b < nb; b++) memcpy(out + b * nf_local, global + b * nf_global + g0, sizeof(double) * nf_local);
    return 1;
}

size_t vo_truncate_aligned(const double *ref, size_t ref_len, const double *deg, size_t deg_len, int64_t lag,
                           uint32_t sr, double *ref_out, double *deg_out, size_t *ref_off, int64_t *deg_off);
int vo_find_matches(const double *sims, size_t np, size_t nf_deg, const size_t *lo, const size_t *hi,
                    const double *ref_spec, const size_t *ref_idx, const double *deg_spec, size_t nb, size_t nf_ref,
                    size_t patch, double frame_duration, vo_match *items, size_t *dp_offsets);
int vo_plan_search(const size_t *ref_idx, size_t n_idx, size_t nf_ref, size_t nf_deg, size_t patch, int32_t radius,
                   size_t *lo, size_t *hi, size_t *np_out);

#define FAIL(code, ...) do { vo_set_error(__VA_ARGS__); rc = (code); goto done; } while (0)

/* One pair through visqol.zig:394-487's eight stages. */
static int measure_one(const struct vo_ctx *c, const rate_cache *rcache, const double *ref, size_t ref_len,
                       const double *deg_in, size_t deg_len, uint32_t sr, vo_result **out, vo_trace *tr,
                       double *ref_spec_out, double *deg_spec_out, double *sims_out) {
    const size_t nb = c->nb, patch = c->patch, W = rcache->window, hop = rcache->hop;
    int rc = VO_OK;
    double *deg = NULL, *ref_spec = NULL, *deg_spec = NULL, *ref_raw = NULL, *deg_raw = NULL, *sims = NULL;
    size_t *idx = NULL, *lo = NULL, *hi = NULL;
    vo_match *items = NULL;
    if (tr) {
        memset(tr, 0, sizeof *tr);
        tr->spl_factor = 1.0;
        for (size_t i = 0; i < 512; i++) tr->realign_lag[i] = INT64_MIN;
    }
    /* 1: global lag (visqol.zig:425-434; accel.zig:100-107) */
    int64_t lag = 0;
    if (!c->opts.disable_global_alignment && ref_len > 0 && deg_len > 0) {
        double *er = (double *)malloc(sizeof(double) * ref_len), *ed = (double *)malloc(sizeof(double) * deg_len);
        vo_upper_envelope(ref, ref_len, er);
        vo_upper_envelope(deg_in, deg_len, ed);
        lag = vo_best_lag(er, ref_len, ed, deg_len);
        free(er);
        free(ed);
    }
    if (tr) tr->lag = lag;
    /* 2: apply lag, SPL match (visqol.zig:489-517; realign.zig:91-110) */
    double lag_s = 0.0;
    size_t dl = deg_len;
    if (c->opts.disable_global_alignment) {
        deg = (double *)malloc(sizeof(double) * (deg_len ? deg_len : 1));
        memcpy(deg, deg_in, sizeof(double) * deg_len);
    } else {
        int64_t half = (int64_t)(ref_len / 2), al = lag < 0 ? -lag : lag;
        if (lag == 0 || al > half) {
            deg = (double *)malloc(sizeof(double) * (deg_len ? deg_len : 1));
            memcpy(deg, deg_in, sizeof(double) * deg_len);
        } else {
            lag_s = (double)lag / (double)sr;
            if (lag < 0) {
                size_t k = (size_t)(-lag) < deg_len ? (size_t)(-lag) : deg_len;
                dl = deg_len - k;
                deg = (double *)malloc(sizeof(double) * (dl ? dl : 1));
                memcpy(deg, deg_in + k, sizeof(double) * dl);
            } else {
                size_t k = (size_t)lag;
                dl = k + deg_len;
                deg = (double *)calloc(dl, sizeof(double));
                memcpy(deg + k, deg_in, sizeof(double) * deg_len);
            }
        }
    }
    double factor = vo_spl_factor_for(ref, ref_len, deg, dl);
    {
        double rs = vo_spl(ref, ref_len), ds = vo_spl(deg, dl);
        if (isfinite(rs) && isfinite(ds)) {
            for (size_t i = 0; i < dl; i++) deg[i] *= factor;
        } else {
            factor = 1.0;
        }
    }
    if (tr) tr->spl_factor = factor;
    if (ref_len <= W || dl <= W) FAIL(VO_E_SIGNAL, "SpectrogramTooShort");
    const size_t nf_ref = vo_num_frames(ref_len, W, hop), nf_deg = vo_num_frames(dl, W, hop);
    if (tr) { tr->nf_ref = nf_ref; tr->nf_deg = nf_deg; }
    /* 3: spectrograms */
    ref_spec = (double *)malloc(sizeof(double) * nb * nf_ref);
    deg_spec = (double *)malloc(sizeof(double) * nb * nf_deg);
    if (!ref_spec || !deg_spec) FAIL(VO_E_OUT_OF_MEMORY, "OutOfMemory");
    vo_spectrogram(rcache->bands, nb, ref, ref_len, W, hop, rcache->hann, ref_spec);
    vo_spectrogram(rcache->bands, nb, deg, dl, W, hop, rcache->hann, deg_spec);
    if (ref_spec_out) memcpy(ref_spec_out, ref_spec, sizeof(double) * nb * nf_ref);
    if (deg_spec_out) memcpy(deg_spec_out, deg_spec, sizeof(double) * nb * nf_deg);
    /* 4-5: prep, patch indices, plan (visqol.zig:519-555) */
    if (!c->opts.disable_realignment) {
        ref_raw = (double *)malloc(sizeof(double) * nb * nf_ref);
        deg_raw = (double *)malloc(sizeof(double) * nb * nf_deg);
        memcpy(ref_raw, ref_spec, sizeof(double) * nb * nf_ref);
        memcpy(deg_raw, deg_spec, sizeof(double) * nb * nf_deg);
    }
    vo_prepare_for_comparison(ref_spec, deg_spec, nb, nf_ref, nf_deg);
    size_t cap = nf_ref / (patch ? patch : 1) + 4, n_idx;
    idx = (size_t *)malloc(sizeof(size_t) * cap);
    if (c->opts.mode == 0) {
        n_idx = vo_ref_patch_indices(nf_ref, patch, idx, cap);
        if (n_idx == (size_t)-1) FAIL(VO_E_SIGNAL, "ReferenceSpectrogramTooSmall");
    } else {
        n_idx = vo_vad_patch_indices(ref, ref_len, nf_ref, patch, W, 0.25, idx, cap);
    }
    if (n_idx == 0) FAIL(VO_E_SIGNAL, "NoPatches");
    lo = (size_t *)malloc(sizeof(size_t) * n_idx);
    hi = (size_t *)malloc(sizeof(size_t) * n_idx);
    size_t np = 0;
    if (vo_plan_search(idx, n_idx, nf_ref, nf_deg, patch, c->opts.search_window_radius, lo, hi, &np))
        FAIL(VO_E_SIGNAL, "SignalsTooDifferent");
    if (tr) {
        tr->num_patches = np;
        for (size_t i = 0; i < np && i < 512; i++) tr->ref_indices[i] = idx[i];
    }
    /* 6a: similarity table; 6b: DP + per-patch NSIM */
    sims = (double *)calloc(np * nf_deg, sizeof(double));
    vo_similarity_table(ref_spec, nf_ref, deg_spec, nf_deg, nb, patch, idx, lo, hi, np, sims);
    if (sims_out) memcpy(sims_out, sims, sizeof(double) * np * nf_deg);
    items = (vo_match *)malloc(sizeof(vo_match) * np);
    size_t *offs = (size_t *)malloc(sizeof(size_t) * np);
    vo_find_matches(sims, np, nf_deg, lo, hi, ref_spec, idx, deg_spec, nb, nf_ref, patch, rcache->frame_duration,
                    items, offs);
    if (tr)
        for (size_t i = 0; i < np && i < 512; i++) { tr->dp_offset[i] = offs[i]; tr->match_valid[i] = (uint8_t)items[i].valid; }
    free(offs);
    /* 7: realignment (realign.zig:294-460) */
    if (!c->opts.disable_realignment) {
        for (size_t mi = 0; mi < np; mi++) {
            vo_match *m = &items[mi];
            if (!m->valid) continue;
            vo_layout rl = vo_layout_of(ref_len, sr, m->ref_start, m->ref_end);
            vo_layout dlay = vo_layout_of(dl, sr, m->deg_start, m->deg_end);
            size_t rlen = rl.pad + rl.len, dlen = dlay.pad + dlay.len;
            if (rlen < W || dlen < W) continue;
            double *ra = (double *)calloc(rlen ? rlen : 1, sizeof(double));
            double *da = (double *)calloc(dlen ? dlen : 1, sizeof(double));
            memcpy(ra + rl.pad, ref + rl.base, sizeof(double) * rl.len);
            memcpy(da + dlay.pad, deg + dlay.base, sizeof(double) * dlay.len);
            double *er = (double *)malloc(sizeof(double) * (rlen ? rlen : 1));
            double *ed = (double *)malloc(sizeof(double) * (dlen ? dlen : 1));
            vo_upper_envelope(ra, rlen, er);
            vo_upper_envelope(da, dlen, ed);
            int64_t plag = vo_best_lag(er, rlen, ed, dlen);
            free(er);
            free(ed);
            double *tref = (double *)malloc(sizeof(double) * (rlen ? rlen : 1));
            double *tdeg = (double *)malloc(sizeof(double) * (rlen ? rlen : 1));
            size_t roff;
            int64_t doff;
            size_t tlen = vo_truncate_aligned(ra, rlen, da, dlen, plag, sr, tref, tdeg, &roff, &doff);
            size_t nf = vo_num_frames(tlen, W, hop);
            if (tr && mi < 512) { tr->realign_lag[mi] = plag; tr->realign_frames[mi] = nf; }
            if (nf > 0) {
                double *rs = (double *)malloc(sizeof(double) * nb * nf), *ds = (double *)malloc(sizeof(double) * nb * nf);
                if (!reuse_frames(nb, ref_raw, nf_ref, rl, (int64_t)roff, hop, nf, rs))
                    vo_spectrogram(rcache->bands, nb, tref, tlen, W, hop, rcache->hann, rs);
                if (!reuse_frames(nb, deg_raw, nf_deg, dlay, doff, hop, nf, ds))
                    vo_spectrogram(rcache->bands, nb, tdeg, tlen, W, hop, rcache->hann, ds);
                /* finishRealign (realign.zig:398-460) */
                size_t glen, gro;
                int64_t gdo;
                double glag_s;
                vo_trim_geometry(rlen, dlen, plag, sr, &glen, &gro, &gdo, &glag_s);
                vo_prepare_for_comparison(rs, ds, nb, nf, nf);
                double nm[VO_MAX_BANDS], nsd[VO_MAX_BANDS], ne[VO_MAX_BANDS];
                vo_psim ps = {nm, nsd, ne, 0.0};
                vo_patch_similarity_into(rs, ds, nb, nf, &ps);
                if (!(ps.mean_of_means < m->similarity)) {
                    memcpy(m->means, nm, sizeof(double) * nb);
                    memcpy(m->stddevs, nsd, sizeof(double) * nb);
                    memcpy(m->energy, ne, sizeof(double) * nb);
                    m->similarity = ps.mean_of_means;
                    double new_dur = (double)glen / (double)sr;
                    if (glag_s > 0.0) m->ref_start += glag_s;
                    else m->deg_start -= glag_s;
                    m->ref_end = m->ref_start + new_dur;
                    m->deg_end = m->deg_start + new_dur;
                    if (tr && mi < 512) tr->realign_accepted[mi] = 1;
                }
                free(rs);
                free(ds);
            }
            free(tref); free(tdeg); free(ra); free(da);
        }
    }
    /* 8: aggregate, map */
    rc = build_result(c, items, np, rcache->frame_duration, rcache->cf, lag_s, out);
done:
    free(deg); free(ref_spec); free(deg_spec); free(ref_raw); free(deg_raw); free(sims);
    free(idx); free(lo); free(hi); free(items);
    return rc;
}

int vo_measure_trace(vo_ctx *c, const double *ref, size_t ref_len, const double *deg, size_t deg_len,
                     uint32_t sample_rate, vo_result **out, vo_trace *trace, double *ref_spec_out,
                     double *deg_spec_out, double *sims_out) {
    if (!c) { vo_set_error("ctx is NULL"); return VO_E_INVALID_ARGUMENT; }
    if (!out) { vo_set_error("out is NULL"); return VO_E_INVALID_ARGUMENT; }
    *out = NULL;
    if ((!ref && ref_len > 0) || (!deg && deg_len > 0)) { vo_set_error("signal pointer is NULL"); return VO_E_INVALID_ARGUMENT; }
    if (sample_rate == 0) { vo_set_error("InvalidInput"); return VO_E_INVALID_ARGUMENT; }
    rate_cache rcache;
    int rc = rate_cache_init(&rcache, c, sample_rate);
    if (rc) return rc;
    rc = measure_one(c, &rcache, ref, ref_len, deg, deg_len, sample_rate, out, trace, ref_spec_out, deg_spec_out,
                     sims_out);
    free(rcache.hann);
    return rc;
}

int vo_measure(vo_ctx *c, const double *ref, size_t ref_len, const double *deg, size_t deg_len, uint32_t sample_rate,
               vo_result **out) {
    return vo_measure_trace(c, ref, ref_len, deg, deg_len, sample_rate, out, NULL, NULL, NULL, NULL);
}

/* backend.zig:174-221: workers pull pair indices from an atomic counter. */
typedef struct {
    const struct vo_ctx *c;
    const rate_cache *rcache;
    size_t n;
    const double *const *refs;
    const size_t *ref_lens;
    const double *const *degs;
    const size_t *deg_lens;
    uint32_t sr;
    vo_result **out;
    int32_t *codes;
    atomic_size_t next;
} batch_shared;

static void *batch_worker(void *arg) {
    batch_shared *s = (batch_shared *)arg;
    for (;;) {
        size_t i = atomic_fetch_add(&s->next, 1);
        if (i >= s->n) return NULL;
        s->out[i] = NULL;
        s->codes[i] = measure_one(s->c, s->rcache, s->refs[i], s->ref_lens[i], s->degs[i], s->deg_lens[i], s->sr,
                                  &s->out[i], NULL, NULL, NULL, NULL);
    }
}

int vo_measure_batch(vo_ctx *c, size_t n, const double *const *refs, const size_t *ref_lens,
                     const double *const *degs, const size_t *deg_lens, uint32_t sr, uint32_t num_threads,
                     vo_result **out, int32_t *status) {
    if (!c) { vo_set_error("ctx is NULL"); return VO_E_INVALID_ARGUMENT; }
    if (n == 0) return VO_OK;
    if (!out || !refs || !ref_lens || !degs || !deg_lens) { vo_set_error("NULL argument"); return VO_E_INVALID_ARGUMENT; }
    if (sr == 0) { vo_set_error("InvalidInput"); return VO_E_INVALID_ARGUMENT; }
    rate_cache rcache;
    int rc = rate_cache_init(&rcache, c, sr);
    if (rc) return rc;
    int32_t *codes = (int32_t *)calloc(n, sizeof(int32_t));
    batch_shared s = {c, &rcache, n, refs, ref_lens, degs, deg_lens, sr, out, codes, 0};
    long cpus = sysconf(_SC_NPROCESSORS_ONLN);
    size_t want = num_threads == 0 ? (size_t)(cpus > 0 ? cpus : 1) : num_threads;
    size_t nt = want < n ? want : n;
    if (nt < 1) nt = 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nt);
    size_t spawned = 0;
    for (size_t t = 1; t < nt; t++)
        if (pthread_create(&th[spawned], NULL, batch_worker, &s) == 0) spawned++;
    batch_worker(&s);
    for (size_t t = 0; t < spawned; t++) pthread_join(th[t], NULL);
    free(th);
    int first = VO_OK;
    for (size_t i = 0; i < n; i++) {
        if (status) status[i] = codes[i];
        if (codes[i] != VO_OK && first == VO_OK) first = codes[i];
    }
    free(codes);
    free(rcache.hann);
    return first;
}
