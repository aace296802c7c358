/* Oracle NSIM, patch creation, DP search, VAD and realignment geometry.
 * Test infrastructure only (see vo.h). */
#include "vo_internal.h"

#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------
 * NSIM (nsim.zig).
 * ---------------------------------------------------------------------- */
/* nsim.zig:111-115; flipped taps as Taps.init (nsim.zig:130-142). */
static const double WIN[3][3] = {
    {0.0113033910173052, 0.0838251475442633, 0.0113033910173052},
    {0.0838251475442633, 0.619485845753726, 0.0838251475442633},
    {0.0113033910173052, 0.0838251475442633, 0.0113033910173052},
};
#define W00 WIN[2][2]
#define W01 WIN[2][1]
#define W02 WIN[2][0]
#define W10 WIN[1][2]
#define W11 WIN[1][1]
#define W12 WIN[1][0]
#define W20 WIN[0][2]
#define W21 WIN[0][1]
#define W22 WIN[0][0]

/* nsim.zig:145-155 (Taps.at): the single tap summation order. */
static inline double tap_at(const double *in, size_t rm, size_t r0, size_t rp, size_t cm, size_t c, size_t cp) {
    return W00 * in[rm + cm] + W01 * in[rm + c] + W02 * in[rm + cp] + W10 * in[r0 + cm] + W11 * in[r0 + c] +
           W12 * in[r0 + cp] + W20 * in[rp + cm] + W21 * in[rp + c] + W22 * in[rp + cp];
}

/* nsim.zig:182-192 (Taps.row) */
static void tap_row(const double *in, size_t rm, size_t r0, size_t rp, size_t ncols, double *out) {
    if (ncols == 1) { out[0] = tap_at(in, rm, r0, rp, 0, 0, 0); return; }
    out[0] = tap_at(in, rm, r0, rp, 0, 0, 1);
    for (size_t c = 1; c + 1 < ncols; c++) out[c] = tap_at(in, rm, r0, rp, c - 1, c, c + 1);
    out[ncols - 1] = tap_at(in, rm, r0, rp, ncols - 2, ncols - 1, ncols - 1);
}

static inline size_t row_above(size_t r) { return r == 0 ? 0 : r - 1; }
static inline size_t row_below(size_t r, size_t n) { return r + 1 >= n ? n - 1 : r + 1; }

/* nsim.zig:204-218 */
void vo_conv2d3x3(const double *in, double *out, size_t nrows, size_t ncols) {
    for (size_t r = 0; r < nrows; r++)
        tap_row(in, row_above(r) * ncols, r * ncols, row_below(r, nrows) * ncols, ncols, out + r * ncols);
}

static const double INTENSITY_RANGE = 1.0;
#define NSIM_C1 ((0.01 * INTENSITY_RANGE) * (0.01 * INTENSITY_RANGE))
#define NSIM_C3 ((0.03 * INTENSITY_RANGE) * (0.03 * INTENSITY_RANGE) / 2.0)

/* nsim.zig:293-308 */
static inline double sim_cell(double mr, double rms, double srs, double md, double dms, double sds, double crd,
                              double c1, double c3) {
    double mrd = mr * md;
    double srd = crd - mrd;
    double intensity = (2.0 * mrd + c1) / (rms + dms + c1);
    double prod = srs * sds;
    double den = prod < 0.0 ? c3 : sqrt(prod) + c3;
    double structure = (srd + c3) / den;
    return intensity * structure;
}

/* nsim.zig:332-360 (measureSimilarityScalar, ref cache computed inline). */
double vo_similarity_scalar(const double *ref, const double *deg, size_t nr, size_t nc) {
    const double c1 = NSIM_C1, c3 = NSIM_C3;
    size_t n = nr * nc;
    double *buf = (double *)malloc(sizeof(double) * n * 8);
    double *mu_ref = buf, *ref_sq = buf + n, *conv_ref_sq = buf + 2 * n, *mu_deg = buf + 3 * n;
    double *deg_sq = buf + 4 * n, *ref_deg = buf + 5 * n, *conv_deg_sq = buf + 6 * n, *conv_rd = buf + 7 * n;
    vo_conv2d3x3(ref, mu_ref, nr, nc);
    for (size_t i = 0; i < n; i++) ref_sq[i] = ref[i] * ref[i];
    vo_conv2d3x3(ref_sq, conv_ref_sq, nr, nc);
    vo_conv2d3x3(deg, mu_deg, nr, nc);
    for (size_t i = 0; i < n; i++) deg_sq[i] = deg[i] * deg[i];
    for (size_t i = 0; i < n; i++) ref_deg[i] = ref[i] * deg[i];
    vo_conv2d3x3(deg_sq, conv_deg_sq, nr, nc);
    vo_conv2d3x3(ref_deg, conv_rd, nr, nc);
    double sum = 0.0;
    for (size_t i = 0; i < n; i++) {
        double mr = mu_ref[i], md = mu_deg[i];
        double rms = mr * mr;
        double srs = conv_ref_sq[i] - rms;
        double dms = md * md;
        sum += sim_cell(mr, mr * mr, srs, md, dms, conv_deg_sq[i] - dms, conv_rd[i], c1, c3);
    }
    free(buf);
    return sum / (double)n;
}

/* nsim.zig:224-263 (DegPlane) */
typedef struct { size_t nrows, width; double *pad, *sq, *mu, *dms, *sds; } deg_plane;

static int plane_init(deg_plane *p, const double *spec, size_t nrows, size_t nf, size_t width) {
    size_t n = nrows * width;
    double *slab = (double *)malloc(sizeof(double) * 5 * n);
    if (!slab) return -1;
    p->nrows = nrows; p->width = width;
    p->pad = slab; p->sq = slab + n; p->mu = slab + 2 * n; p->dms = slab + 3 * n; p->sds = slab + 4 * n;
    for (size_t r = 0; r < nrows; r++) {
        memcpy(p->pad + r * width, spec + r * nf, sizeof(double) * nf);
        memset(p->pad + r * width + nf, 0, sizeof(double) * (width - nf));
    }
    for (size_t i = 0; i < n; i++) p->sq[i] = p->pad[i] * p->pad[i];
    vo_conv2d3x3(p->pad, p->mu, nrows, width);
    vo_conv2d3x3(p->sq, p->sds, nrows, width);
    for (size_t i = 0; i < n; i++) {
        p->dms[i] = p->mu[i] * p->mu[i];
        p->sds[i] = p->sds[i] - p->dms[i];
    }
    return 0;
}

/* nsim.zig:318-329 (precomputeRef) + 364-424 (measureSimilarityAt). */
typedef struct { size_t nr, nc; double *mu_ref, *rms, *srs, *ref_deg, *row_crd, *row_sim, *tmp; } ref_cache;

static void ref_cache_init(ref_cache *s, size_t nr, size_t nc) {
    size_t n = nr * nc;
    s->nr = nr; s->nc = nc;
    s->mu_ref = (double *)malloc(sizeof(double) * (5 * n + 2 * nc));
    s->rms = s->mu_ref + n; s->srs = s->rms + n; s->ref_deg = s->srs + n; s->tmp = s->ref_deg + n;
    s->row_crd = s->tmp + n; s->row_sim = s->row_crd + nc;
}
static void ref_cache_free(ref_cache *s) { free(s->mu_ref); }

static void precompute_ref(ref_cache *s, const double *ref) {
    size_t n = s->nr * s->nc;
    vo_conv2d3x3(ref, s->mu_ref, s->nr, s->nc);
    for (size_t i = 0; i < n; i++) s->tmp[i] = ref[i] * ref[i];
    vo_conv2d3x3(s->tmp, s->srs, s->nr, s->nc); /* conv_ref_sq, finished below */
    for (size_t i = 0; i < n; i++) {
        s->rms[i] = s->mu_ref[i] * s->mu_ref[i];
        s->srs[i] = s->srs[i] - s->rms[i];
    }
}

static double similarity_at(ref_cache *s, const double *ref, const deg_plane *pl, size_t off) {
    const size_t nr = s->nr, nc = s->nc, w = pl->width;
    const double c1 = NSIM_C1, c3 = NSIM_C3;
    for (size_t r = 0; r < nr; r++)
        for (size_t c = 0; c < nc; c++) s->ref_deg[r * nc + c] = ref[r * nc + c] * pl->pad[r * w + off + c];
    double sum = 0.0;
    for (size_t r = 0; r < nr; r++) {
        size_t rm = row_above(r), rp = row_below(r, nr);
        tap_row(s->ref_deg, rm * nc, r * nc, rp * nc, nc, s->row_crd);
        size_t pr = r * w + off, ri = r * nc;
        size_t n_edges = nc < 2 ? nc : 2;
        for (size_t e = 0; e < n_edges; e++) {
            size_t c = e == 0 ? 0 : nc - 1;
            size_t cm = off + (c == 0 ? 0 : c - 1);
            size_t cp = off + (c + 1 >= nc ? nc - 1 : c + 1);
            double md = tap_at(pl->pad, rm * w, r * w, rp * w, cm, off + c, cp);
            double cdsq = tap_at(pl->sq, rm * w, r * w, rp * w, cm, off + c, cp);
            double dms = md * md;
            s->row_sim[c] = sim_cell(s->mu_ref[ri + c], s->rms[ri + c], s->srs[ri + c], md, dms, cdsq - dms,
                                     s->row_crd[c], c1, c3);
        }
        for (size_t c = 1; c + 1 < nc; c++)
            s->row_sim[c] = sim_cell(s->mu_ref[ri + c], s->rms[ri + c], s->srs[ri + c], pl->mu[pr + c], pl->dms[pr + c],
                                     pl->sds[pr + c], s->row_crd[c], c1, c3);
        for (size_t c = 0; c < nc; c++) sum += s->row_sim[c];
    }
    return sum / (double)(nr * nc);
}

/* accel.zig:126-146 (Cpu.patchSimilarities) for one pair; sims[p*nf_deg+off]. */
void vo_similarity_table(const double *ref_spec, size_t nf_ref, const double *deg_spec, size_t nf_deg, size_t nb,
                         size_t patch, const size_t *ref_idx, const size_t *lo, const size_t *hi, size_t np,
                         double *sims) {
    size_t width = nf_deg;
    for (size_t p = 0; p < np; p++) if (hi[p] + patch > width) width = hi[p] + patch;
    deg_plane pl = {0};
    if (plane_init(&pl, deg_spec, nb, nf_deg, width)) return;
    ref_cache s;
    ref_cache_init(&s, nb, patch);
    double *ref_patch = (double *)malloc(sizeof(double) * nb * patch);
    for (size_t p = 0; p < np; p++) {
        for (size_t b = 0; b < nb; b++)
            for (size_t c = 0; c < patch; c++) ref_patch[b * patch + c] = ref_spec[b * nf_ref + ref_idx[p] + c];
        precompute_ref(&s, ref_patch);
        for (size_t off = lo[p]; off <= hi[p]; off++) sims[p * nf_deg + off] = similarity_at(&s, ref_patch, &pl, off);
    }
    free(ref_patch);
    ref_cache_free(&s);
    free(pl.pad);
}

/* nsim.zig:427-510 (measurePatchSimilarity) */
void vo_patch_similarity_into(const double *ref, const double *deg, size_t nr, size_t nc, vo_psim *out) {
    const double c1 = NSIM_C1, c3 = NSIM_C3;
    size_t n = nr * nc;
    double *buf = (double *)malloc(sizeof(double) * n * 8);
    double *mu_ref = buf, *mu_deg = buf + n, *ref_sq = buf + 2 * n, *deg_sq = buf + 3 * n;
    double *ref_deg = buf + 4 * n, *conv_ref_sq = buf + 5 * n, *conv_deg_sq = buf + 6 * n, *conv_rd = buf + 7 * n;
    vo_conv2d3x3(ref, mu_ref, nr, nc);
    vo_conv2d3x3(deg, mu_deg, nr, nc);
    for (size_t i = 0; i < n; i++) ref_sq[i] = ref[i] * ref[i];
    for (size_t i = 0; i < n; i++) deg_sq[i] = deg[i] * deg[i];
    for (size_t i = 0; i < n; i++) ref_deg[i] = ref[i] * deg[i];
    vo_conv2d3x3(ref_sq, conv_ref_sq, nr, nc);
    vo_conv2d3x3(deg_sq, conv_deg_sq, nr, nc);
    vo_conv2d3x3(ref_deg, conv_rd, nr, nc);
    const double inv_ncols = 1.0 / (double)nc;
    for (size_t r = 0; r < nr; r++) {
        double deg_sum = 0.0;
        for (size_t c = 0; c < nc; c++) deg_sum += deg[r * nc + c];
        out->energy[r] = deg_sum * inv_ncols;
    }
    for (size_t r = 0; r < nr; r++) {
        double sim_sum = 0.0;
        for (size_t c = 0; c < nc; c++) {
            size_t idx = r * nc + c;
            double mr = mu_ref[idx], md = mu_deg[idx];
            double rms = mr * mr, dms = md * md, mrd = mr * md;
            double srs = conv_ref_sq[idx] - rms, sds = conv_deg_sq[idx] - dms, srd = conv_rd[idx] - mrd;
            double intensity = (2.0 * mrd + c1) / (rms + dms + c1);
            double prod = srs * sds;
            double den = prod < 0.0 ? c3 : sqrt(prod) + c3;
            double structure = (srd + c3) / den;
            double sim = intensity * structure;
            sim_sum += sim;
            conv_rd[idx] = sim;
        }
        out->means[r] = sim_sum * inv_ncols;
    }
    for (size_t r = 0; r < nr; r++) {
        if (nc <= 1) { out->stddevs[r] = 0.0; continue; }
        double mean = out->means[r], ss = 0.0;
        for (size_t c = 0; c < nc; c++) {
            double diff = conv_rd[r * nc + c] - mean;
            ss += diff * diff;
        }
        out->stddevs[r] = sqrt(ss / ((double)nc - 1.0));
    }
    double acc = 0.0;
    for (size_t r = 0; r < nr; r++) acc += out->means[r];
    out->mean_of_means = acc / (double)nr;
    free(buf);
}

void vo_patch_similarity(const double *ref, const double *deg, size_t nr, size_t nc, double *means,
                         double *stddevs, double *energy, double *mean_of_means) {
    vo_psim ps = {means, stddevs, energy, 0.0};
    vo_patch_similarity_into(ref, deg, nr, nc, &ps);
    *mean_of_means = ps.mean_of_means;
}

/* ------------------------------------------------------------------------
 * Patch indices (dp.zig:47-87) and VAD patches (vad.zig:15-81).
 * ---------------------------------------------------------------------- */
size_t vo_ref_patch_indices(size_t nf, size_t patch, size_t *out, size_t cap) {
    size_t init = patch / 2;
    if (nf < patch + init) return (size_t)-1; /* ReferenceSpectrogramTooSmall */
    size_t max_index = init < nf - patch ? nf - patch : init + 1;
    size_t k = 0;
    for (size_t i = init; i < max_index; i += patch) {
        if (k < cap) out[k] = i - 1;
        k++;
    }
    return k;
}

size_t vo_calc_max_num_patches(const size_t *idx, size_t n, size_t nf_deg, size_t patch) {
    if (n == 0) return 0;
    while (n > 0) {
        size_t last = idx[n - 1];
        if (last < patch / 2) break;
        if (last - patch / 2 <= nf_deg) break;
        n--;
    }
    return n;
}

size_t vo_vad_patch_indices(const double *ref, size_t len, size_t nf, size_t patch, size_t window, double overlap,
                            size_t *out, size_t cap) {
    const double frame_size = (double)window * overlap;
    const size_t frame_len = (size_t)frame_size;
    const size_t patch_sample_len = (size_t)((double)patch * frame_size);
    const size_t first = patch / 2 - 1;
    if (nf <= first || frame_len == 0) return 0;
    const size_t patch_count = (nf - first) / patch;
    const size_t total = patch_count * patch_sample_len;
    double max_val = -INFINITY;
    for (size_t i = 0; i < len; i++) max_val = ref[i] > max_val ? ref[i] : max_val;
    const size_t n_chunks = total / frame_len;
    unsigned char *active = (unsigned char *)calloc(n_chunks ? n_chunks : 1, 1);
    for (size_t c = 0; c < n_chunks; c++) {
        double square = 0.0;
        for (size_t j = 0; j < frame_len; j++) {
            size_t idx = first + c * frame_len + j;
            double x = idx < len ? ref[idx] / max_val : 0.0;
            /* std.math.clamp = @max(lo, @min(v, hi)); Zig's float min/max
             * return the non-NaN operand, as fmin/fmax do. */
            double scaled = fmax(-32768.0, fmin(x * 32768.0, 32767.0));
            double s = trunc(scaled);
            square += s * s;
        }
        double rms = sqrt(square / (double)frame_len);
        active[c] = rms >= 5000.0;
    }
    const size_t vad_len = n_chunks > 2 ? n_chunks : 2;
    double *vad = (double *)malloc(sizeof(double) * vad_len);
    vad[0] = 1.0;
    vad[1] = 1.0;
    for (size_t i = 2; i < n_chunks; i++) {
        int prev_silent = !active[i - 1] && !active[i - 2];
        vad[i] = (!active[i] && prev_silent) ? 0.0 : 1.0;
    }
    size_t k = 0, patch_idx = first;
    for (size_t p = 0; p < patch_count; p++) {
        double fva = 0.0;
        size_t lo = p * patch;
        for (size_t q = lo; q < lo + patch; q++) if (q < vad_len) fva += vad[q];
        if (fva >= 1.0) {
            if (k < cap) out[k] = patch_idx;
            k++;
        }
        patch_idx += patch;
    }
    free(active);
    free(vad);
    return k;
}

/* ------------------------------------------------------------------------
 * DP (dp.zig:205-355).
 * ---------------------------------------------------------------------- */
#define NO_BACK ((size_t)-1)

/* Fills dp/bt rows and returns the last row's first argmax (dp.zig:227-287). */
static size_t dp_fill(const double *sims, size_t np, size_t cols, const size_t *lo, const size_t *hi, double *dp,
                      size_t *bt, double *best) {
    memset(dp, 0, sizeof(double) * np * cols);
    memset(bt, 0, sizeof(size_t) * np * cols);
    for (size_t pi = 0; pi < np; pi++) {
        size_t start = lo[pi], end = hi[pi];
        size_t prev_row = 0, prev_lower = 0;
        if (pi > 0) { prev_row = (pi - 1) * cols; prev_lower = lo[pi - 1]; }
        double rmax = -INFINITY;
        int64_t rmax_idx = -1;
        if (pi > 0)
            for (size_t i = prev_lower; i < start; i++) {
                double v = dp[prev_row + i];
                if (v >= rmax) { rmax = v; rmax_idx = (int64_t)i; }
            }
        for (size_t slide = start; slide <= end; slide++) {
            double sim = sims[pi * cols + slide];
            int64_t past = -1;
            if (pi > 0) {
                sim += rmax;
                past = rmax_idx;
                double no_move = dp[prev_row + slide];
                if (no_move > sim) { sim = no_move; past = (int64_t)slide; }
                if (no_move >= rmax) { rmax = no_move; rmax_idx = (int64_t)slide; }
            }
            dp[pi * cols + slide] = sim;
            bt[pi * cols + slide] = past < 0 ? NO_BACK : (size_t)past;
        }
    }
    size_t last = np - 1, last_off = 0;
    double max_score = -INFINITY;
    for (size_t off = lo[last]; off <= hi[last]; off++) {
        double v = dp[last * cols + off];
        if (v > max_score) { max_score = v; last_off = off; }
    }
    if (best) *best = max_score;
    return last_off;
}

double vo_dp_search(const double *sims, size_t np, size_t nf_deg, const size_t *lo, const size_t *hi, size_t *offsets,
                    uint8_t *valid) {
    double *dp = (double *)malloc(sizeof(double) * np * nf_deg);
    size_t *bt = (size_t *)malloc(sizeof(size_t) * np * nf_deg);
    double best = 0.0;
    size_t cur = dp_fill(sims, np, nf_deg, lo, hi, dp, bt, &best);
    for (size_t i = 0; i < np; i++) { offsets[i] = 0; valid[i] = 1; }
    for (int64_t p = (int64_t)np - 1; p >= 0; p--) {
        size_t back = bt[(size_t)p * nf_deg + cur];
        offsets[p] = cur;
        valid[p] = back != cur;
        if (back == NO_BACK) break;
        cur = back;
    }
    free(dp);
    free(bt);
    return best;
}

/* dp.zig:90-132 */
static void build_patch(const double *spec, size_t nb, size_t nf, size_t start, size_t cols, double *out) {
    size_t real_end = start + cols < nf ? start + cols : nf;
    size_t real_cols = real_end > start ? real_end - start : 0;
    for (size_t b = 0; b < nb; b++) {
        size_t c = 0;
        for (; c < real_cols; c++) out[b * cols + c] = spec[b * nf + start + c];
        for (; c < cols; c++) out[b * cols + c] = 0.0;
    }
}

/* dp.zig:205-355 (findOptimalDegPatchesFromSims): DP, backtrack, and full
 * per-band NSIM of each chosen patch.  `items` has np entries. */
int vo_find_matches(const double *sims, size_t np, size_t nf_deg, const size_t *lo, const size_t *hi,
                    const double *ref_spec, const size_t *ref_idx, const double *deg_spec, size_t nb, size_t nf_ref,
                    size_t patch, double frame_duration, vo_match *items, size_t *dp_offsets) {
    double *dp = (double *)malloc(sizeof(double) * np * nf_deg);
    size_t *bt = (size_t *)malloc(sizeof(size_t) * np * nf_deg);
    if (!dp || !bt) { free(dp); free(bt); return VO_E_OUT_OF_MEMORY; }
    size_t cur = dp_fill(sims, np, nf_deg, lo, hi, dp, bt, NULL);
    for (size_t i = 0; i < np; i++) {
        memset(&items[i], 0, sizeof(vo_match));
        items[i].valid = 1;
        if (dp_offsets) dp_offsets[i] = 0;
    }
    double *rp = (double *)malloc(sizeof(double) * nb * patch);
    double *dpb = (double *)malloc(sizeof(double) * nb * patch);
    const double patch_duration = frame_duration * (double)patch;
    for (int64_t p = (int64_t)np - 1; p >= 0; p--) {
        size_t i = (size_t)p;
        size_t ridx = ref_idx[i];
        build_patch(deg_spec, nb, nf_deg, cur, patch, dpb);
        build_patch(ref_spec, nb, nf_ref, ridx, patch, rp);
        vo_psim ps = {items[i].means, items[i].stddevs, items[i].energy, 0.0};
        vo_patch_similarity_into(rp, dpb, nb, patch, &ps);
        items[i].similarity = ps.mean_of_means;
        if (dp_offsets) dp_offsets[i] = cur;
        size_t back = bt[i * nf_deg + cur];
        if (back == cur) {
            items[i].deg_start = 0.0;
            items[i].deg_end = 0.0;
            items[i].similarity = 0.0;
            memset(items[i].means, 0, sizeof(items[i].means));
            items[i].valid = 0;
        } else {
            items[i].deg_start = (double)cur * frame_duration;
            items[i].deg_end = items[i].deg_start + patch_duration;
            items[i].valid = 1;
        }
        items[i].ref_start = (double)ridx * frame_duration;
        items[i].ref_end = items[i].ref_start + patch_duration;
        if (back == NO_BACK) break;
        cur = back;
    }
    free(rp);
    free(dpb);
    free(dp);
    free(bt);
    return VO_OK;
}

/* dp.zig:153-201 (searchPatches / searchWindow / planSearch) */
int vo_plan_search(const size_t *ref_idx, size_t n_idx, size_t nf_ref, size_t nf_deg, size_t patch, int32_t radius,
                   size_t *lo, size_t *hi, size_t *np_out) {
    size_t np = vo_calc_max_num_patches(ref_idx, n_idx, nf_deg, patch);
    if (np == 0 || nf_deg == 0) return VO_E_SIGNAL;
    for (size_t i = 0; i < np; i++) if (ref_idx[i] + patch > nf_ref) return VO_E_SIGNAL;
    int64_t sw = (int64_t)radius * (int64_t)patch;
    int64_t last = (int64_t)nf_deg - 1;
    for (size_t i = 0; i < np; i++) {
        int64_t r = (int64_t)ref_idx[i];
        int64_t l = r - sw > 0 ? r - sw : 0;
        int64_t h = r + sw < last ? r + sw : last;
        lo[i] = (size_t)l;
        hi[i] = (size_t)(h > l ? h : l);
    }
    *np_out = np;
    return VO_OK;
}

int vo_find_optimal_patches(const double *ref_spec, const size_t *ref_idx, size_t n_idx, const double *deg_spec,
                            size_t nb, size_t nf_ref, size_t nf_deg, size_t patch, double frame_duration,
                            int32_t radius, double *deg_start_times) {
    size_t *lo = (size_t *)malloc(sizeof(size_t) * n_idx), *hi = (size_t *)malloc(sizeof(size_t) * n_idx), np = 0;
    int rc = vo_plan_search(ref_idx, n_idx, nf_ref, nf_deg, patch, radius, lo, hi, &np);
    if (rc) { free(lo); free(hi); return rc; }
    double *sims = (double *)calloc(np * nf_deg, sizeof(double));
    vo_similarity_table(ref_spec, nf_ref, deg_spec, nf_deg, nb, patch, ref_idx, lo, hi, np, sims);
    vo_match *items = (vo_match *)malloc(sizeof(vo_match) * np);
    vo_find_matches(sims, np, nf_deg, lo, hi, ref_spec, ref_idx, deg_spec, nb, nf_ref, patch, frame_duration, items,
                    NULL);
    for (size_t i = 0; i < np; i++) deg_start_times[i] = items[i].deg_start;
    free(items); free(sims); free(lo); free(hi);
    return (int)np;
}

/* ------------------------------------------------------------------------
 * Realignment geometry (realign.zig:21-51, 147-250).
 * ---------------------------------------------------------------------- */
vo_layout vo_layout_of(size_t signal_len, uint32_t sr, double start, double end) {
    const double sr_f = (double)sr;
    const int64_t n = (int64_t)signal_len;
    int64_t start_idx = (int64_t)(start * sr_f);
    if (start_idx < 0) start_idx = 0;
    int64_t end_idx = (int64_t)(end * sr_f);
    if (end_idx > n - 1) end_idx = n - 1;
    size_t real_len = end_idx > start_idx ? (size_t)(end_idx - start_idx) : 0;
    double end_diff = end * sr_f - (double)signal_len;
    size_t post = end_diff > 0.0 ? (size_t)end_diff : 0;
    size_t pre = start < 0.0 ? (size_t)(-start * sr_f) : 0;
    vo_layout l = {pre + post, real_len > 0 ? (size_t)start_idx : 0, real_len};
    return l;
}

void vo_slice_layout(size_t len, uint32_t sr, double start, double end, size_t *pad, size_t *base, size_t *n) {
    vo_layout l = vo_layout_of(len, sr, start, end);
    *pad = l.pad; *base = l.base; *n = l.len;
}

int64_t vo_lag_samples_back(int64_t lag, uint32_t sr) {
    double sr_f = (double)sr;
    double seconds = (double)lag / sr_f;
    return (int64_t)(seconds * sr_f);
}

/* realign.zig:229-250 */
void vo_trim_geometry(size_t ref_len, size_t deg_len, int64_t lag, uint32_t sr, size_t *len, size_t *ref_off,
                      int64_t *deg_off, double *lag_seconds) {
    int64_t r = (int64_t)ref_len, d = (int64_t)deg_len;
    int64_t alag = lag < 0 ? -lag : lag;
    int64_t shift = (lag == 0 || alag > (int64_t)(ref_len / 2)) ? 0 : lag;
    int64_t dl = shift < 0 ? d - (-shift < d ? -shift : d) : d + shift;
    int64_t head = 0, ln = r;
    if (r > dl) {
        ln = dl;
    } else if (r < dl && shift > 0) {
        head = vo_lag_samples_back(shift, sr);
        if (head < 0) head = 0;
        ln = r - head;
    }
    *len = (size_t)ln;
    *ref_off = (size_t)head;
    *deg_off = head - shift;
    *lag_seconds = shift == 0 ? 0.0 : (double)shift / (double)sr;
}

/* realign.zig:147-210 (truncateAligned) with real buffers: the trimmed
 * copies go to ref_out / deg_out (each at least ref_len long); returns the
 * shared length and the offsets used for spectrogram frame reuse. */
size_t vo_truncate_aligned(const double *ref, size_t ref_len, const double *deg, size_t deg_len, int64_t lag,
                           uint32_t sr, double *ref_out, double *deg_out, size_t *ref_off, int64_t *deg_off) {
    /* applyLag (realign.zig:91-110) */
    int64_t half_ref = (int64_t)(ref_len / 2);
    int64_t alag = lag < 0 ? -lag : lag;
    double lag_seconds = 0.0;
    int64_t lag_samples = 0;
    size_t al = deg_len;
    double *aligned;
    if (lag == 0 || alag > half_ref) {
        aligned = (double *)malloc(sizeof(double) * (deg_len ? deg_len : 1));
        memcpy(aligned, deg, sizeof(double) * deg_len);
    } else {
        lag_seconds = (double)lag / (double)sr;
        lag_samples = lag;
        if (lag < 0) {
            size_t k = (size_t)(-lag) < deg_len ? (size_t)(-lag) : deg_len;
            al = deg_len - k;
            aligned = (double *)malloc(sizeof(double) * (al ? al : 1));
            memcpy(aligned, deg + k, sizeof(double) * al);
        } else {
            size_t k = (size_t)lag;
            al = k + deg_len;
            aligned = (double *)malloc(sizeof(double) * al);
            memset(aligned, 0, sizeof(double) * k);
            memcpy(aligned + k, deg, sizeof(double) * deg_len);
        }
    }
    size_t head = 0, rl, dl;
    const double *rsrc, *dsrc;
    if (ref_len > al) {
        rsrc = ref; rl = al;
        dsrc = aligned; dl = al;
    } else if (ref_len < al) {
        int64_t ls_i = (int64_t)(lag_seconds * (double)sr);
        if (ls_i > 0) {
            size_t ls = (size_t)ls_i;
            head = ls;
            rsrc = ref + ls; rl = ref_len - ls;
            size_t d_take = ls + rl < al ? ls + rl : al;
            dsrc = aligned + ls; dl = d_take - ls;
        } else {
            dsrc = aligned; dl = ref_len;
            rsrc = ref; rl = ref_len;
        }
    } else {
        rsrc = ref; rl = ref_len;
        dsrc = aligned; dl = al;
    }
    size_t min_len = rl < dl ? rl : dl;
    memcpy(ref_out, rsrc, sizeof(double) * min_len);
    memcpy(deg_out, dsrc, sizeof(double) * min_len);
    free(aligned);
    *ref_off = head;
    *deg_off = (int64_t)head - lag_samples;
    return min_len;
}
