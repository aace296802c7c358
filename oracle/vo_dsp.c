/* Oracle DSP stages: log10, ERB filterbank, gammatone spectrogram, dB prep,
 * upper envelope, radix-2 FFT and the FFT cross-correlation lag search, SPL.
 * Test infrastructure only (see vo.h). */
#include "vo_internal.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------
 * log10: the fdlibm/FreeBSD e_log10.c algorithm as published in musl
 * (src/math/log10.c, MIT).  The reference calls std.math.log10, which Zig's
 * compiler-rt implements with this same published algorithm; the shipped
 * device code carries its own copy so CPU and GPU agree bit for bit.
 * Constants are given by their IEEE-754 bit patterns.
 * ---------------------------------------------------------------------- */
static double bits_to_f64(uint64_t b) { double d; memcpy(&d, &b, 8); return d; }
static uint64_t f64_to_bits(double d) { uint64_t b; memcpy(&b, &d, 8); return b; }

double vo_log10(double x) {
    const double ivln10hi = bits_to_f64(0x3fdbcb7b15200000ULL);
    const double ivln10lo = bits_to_f64(0x3dbb9438ca9aadd5ULL);
    const double log10_2hi = bits_to_f64(0x3fd34413509f6000ULL);
    const double log10_2lo = bits_to_f64(0x3d59fef311f12b36ULL);
    const double Lg1 = bits_to_f64(0x3fe5555555555593ULL);
    const double Lg2 = bits_to_f64(0x3fd999999997fa04ULL);
    const double Lg3 = bits_to_f64(0x3fd2492494229359ULL);
    const double Lg4 = bits_to_f64(0x3fcc71c51d8e78afULL);
    const double Lg5 = bits_to_f64(0x3fc7466496cb03deULL);
    const double Lg6 = bits_to_f64(0x3fc39a09d078c69fULL);
    const double Lg7 = bits_to_f64(0x3fc2f112df3e5244ULL);
    uint64_t ui = f64_to_bits(x);
    uint32_t hx = (uint32_t)(ui >> 32);
    int k = 0;
    if (hx < 0x00100000u || (hx >> 31)) {
        if ((ui << 1) == 0) return -1.0 / (x * x);
        if (hx >> 31) return (x - x) / 0.0;
        k -= 54;
        x *= bits_to_f64(0x4350000000000000ULL); /* 2^54 */
        ui = f64_to_bits(x);
        hx = (uint32_t)(ui >> 32);
    } else if (hx >= 0x7ff00000u) {
        return x;
    } else if (hx == 0x3ff00000u && (ui << 32) == 0) {
        return 0.0;
    }
    hx += 0x3ff00000u - 0x3fe6a09eu;
    k += (int)(hx >> 20) - 0x3ff;
    hx = (hx & 0x000fffffu) + 0x3fe6a09eu;
    ui = ((uint64_t)hx << 32) | (ui & 0xffffffffULL);
    x = bits_to_f64(ui);

    double f = x - 1.0;
    double hfsq = 0.5 * f * f;
    double s = f / (2.0 + f);
    double z = s * s;
    double w = z * z;
    double t1 = w * (Lg2 + w * (Lg4 + w * Lg6));
    double t2 = z * (Lg1 + w * (Lg3 + w * (Lg5 + w * Lg7)));
    double R = t2 + t1;
    double hi = f - hfsq;
    hi = bits_to_f64(f64_to_bits(hi) & 0xffffffff00000000ULL);
    double lo = f - hi - hfsq + s * (hfsq + R);
    double val_hi = hi * ivln10hi;
    double dk = (double)k;
    double y = dk * log10_2hi;
    double val_lo = dk * log10_2lo + (lo + hi) * ivln10lo + lo * ivln10hi;
    w = y + val_hi;
    val_lo += (y - w) + val_hi;
    val_hi = w;
    return val_lo + val_hi;
}

/* ------------------------------------------------------------------------
 * ERB centre frequencies and gammatone coefficients.
 * ---------------------------------------------------------------------- */
#define EAR_Q 9.26449
#define MIN_BW 24.7

/* dsp/erb.zig:30-52 (uniformCenterFreqs), descending. */
void vo_center_freqs_desc(size_t n, double low, double high, double *out) {
    const double q = EAR_Q, m = MIN_BW;
    const double a = -(q * m);
    const double b = -log(high + q * m);
    const double c = log(low + q * m);
    const double d = high + q * m;
    const double e = (b + c) / (double)n;
    for (size_t i = 0; i < n; i++) {
        double idx = (double)i + 1.0;
        double f = exp(idx * e) * d;
        out[i] = a + f;
    }
}

typedef struct { double re, im; } cpx;
static inline cpx c_add(cpx a, cpx b) { return (cpx){a.re + b.re, a.im + b.im}; }
static inline cpx c_sub(cpx a, cpx b) { return (cpx){a.re - b.re, a.im - b.im}; }
static inline cpx c_mul(cpx a, cpx b) { return (cpx){a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
static inline cpx c_muls(cpx a, double s) { return (cpx){a.re * s, a.im * s}; }
static inline cpx c_div(cpx a, cpx b) {
    double den = b.re * b.re + b.im * b.im;
    return (cpx){(a.re * b.re + a.im * b.im) / den, (a.im * b.re - a.re * b.im) / den};
}
static inline cpx c_exp(cpx z) { double e = exp(z.re); return (cpx){e * cos(z.im), e * sin(z.im)}; }

/* gammatone.zig:76-153 (computeBand). Operation order kept left to right. */
static vo_band compute_band(double cf, double t, double pi) {
    const double erb = cf / EAR_Q + MIN_BW; /* erb.zig:20-23, order 1 */
    const double big_b = 1.019 * 2.0 * pi * erb;
    const double exp_bt = exp(big_b * t);
    const double inv_exp_bt = 1.0 / exp_bt;
    const double exp_neg_2bt = exp(-2.0 * big_b * t);
    const double omega = 2.0 * cf * pi * t;
    const double cos_w = cos(omega), sin_w = sin(omega);
    const double b1_re = -2.0 * cos_w * inv_exp_bt;
    const double b2_re = exp_neg_2bt;
    const double sin_t = sin_w * t;
    const double sqrt_3p_p1 = sqrt(3.0 + pow(2.0, 1.5));
    const double sqrt_3m_p1 = sqrt(3.0 - pow(2.0, 1.5));
    const double b_pos = sin_t * 2.0 * sqrt_3p_p1;
    const double b_neg = sin_t * 2.0 * sqrt_3m_p1;
    const double a_term = cos_w * 2.0 * t;
    const double a11_re = -(a_term * inv_exp_bt + b_pos * inv_exp_bt) / 2.0;
    const double a12_re = -(a_term * inv_exp_bt - b_pos * inv_exp_bt) / 2.0;
    const double a13_re = -(a_term * inv_exp_bt + b_neg * inv_exp_bt) / 2.0;
    const double a14_re = -(a_term * inv_exp_bt - b_neg * inv_exp_bt) / 2.0;

    const cpx i_unit = {0.0, 1.0};
    const cpx four_i_cf_pi_t = c_muls(i_unit, 4.0 * cf * pi * t);
    const cpx x_exp = c_exp(four_i_cf_pi_t);
    const cpx x01 = c_muls(x_exp, -2.0 * t);
    const cpx inner = {-big_b * t, 2.0 * cf * pi * t};
    const cpx x02 = c_muls(c_exp(inner), 2.0 * t);
    const cpx xcos = {cos_w, 0.0}, xsin = {sin_w, 0.0};
    const double s1 = sqrt(3.0 - pow(2.0, 1.5));
    const double s2 = sqrt(3.0 + pow(2.0, 1.5));
    const cpx x12 = c_sub(xcos, c_muls(xsin, s1));
    const cpx x22 = c_add(xcos, c_muls(xsin, s1));
    const cpx x32 = c_sub(xcos, c_muls(xsin, s2));
    const cpx x42 = c_add(xcos, c_muls(xsin, s2));
    const cpx x1 = c_add(x01, c_mul(x02, x12));
    const cpx x2 = c_add(x01, c_mul(x02, x22));
    const cpx x3 = c_add(x01, c_mul(x02, x32));
    const cpx x4 = c_add(x01, c_mul(x02, x42));
    const cpx exp_2bt = {exp(2.0 * big_b * t), 0.0};
    const cpx exp_bt_c = {exp_bt, 0.0};
    const cpx term1 = c_div((cpx){-2.0, 0.0}, exp_2bt);
    const cpx term2 = c_muls(x_exp, -2.0);
    const cpx one_p_xexp = c_add((cpx){1.0, 0.0}, x_exp);
    const cpx term3 = c_muls(c_div(one_p_xexp, exp_bt_c), 2.0);
    const cpx x5 = c_add(c_add(term1, term2), term3);
    const cpx x5_p4 = c_mul(c_mul(x5, x5), c_mul(x5, x5));
    const cpx num = c_mul(c_mul(x1, x2), c_mul(x3, x4));
    const cpx ratio = c_div(num, x5_p4);
    const double gain = sqrt(ratio.re * ratio.re + ratio.im * ratio.im);
    return (vo_band){t, a11_re, a12_re, a13_re, a14_re, b1_re, b2_re, gain};
}

/* gammatone.zig:41-70 (Filters.init): bands ascending by centre frequency. */
void vo_filters(uint32_t sample_rate, size_t nbands, double low, double high_in, vo_band *bands, double *cf_asc) {
    const double sr_f = (double)sample_rate;
    const double high = fmin(high_in, sr_f / 2.0);
    double cf_desc[VO_MAX_BANDS];
    vo_center_freqs_desc(nbands, low, high, cf_desc);
    const double t = 1.0 / sr_f;
    const double pi = M_PI;
    for (size_t i = 0; i < nbands; i++) {
        bands[nbands - 1 - i] = compute_band(cf_desc[i], t, pi);
        if (cf_asc) cf_asc[nbands - 1 - i] = cf_desc[i];
    }
}

/* gammatone.zig:193-265 (applyStage / filterBand), fresh state. */
void vo_filter_band(const vo_band *band, const double *in, size_t n, double *out) {
    const double gi = 1.0 / band->gain;
    const double n0s[4] = {band->a0 * gi, band->a0, band->a0, band->a0};
    const double n1s[4] = {band->a11 * gi, band->a12, band->a13, band->a14};
    double *buf = (double *)malloc(sizeof(double) * (n ? n : 1));
    memcpy(buf, in, sizeof(double) * n);
    for (int st = 0; st < 4; st++) {
        double c0 = 0.0, c1 = 0.0;
        for (size_t i = 0; i < n; i++) {
            double s = buf[i];
            double filtered = n0s[st] * s + c0;
            c0 = n1s[st] * s + c1 - band->b1 * filtered;
            c1 = -(band->b2 * filtered);
            buf[i] = filtered;
        }
    }
    memcpy(out, buf, sizeof(double) * n);
    free(buf);
}

/* gammatone.zig:268-306 (filterBandRmsFresh). */
double vo_band_rms_fresh(const vo_band *band, const double *frame, size_t n) {
    const double gi = 1.0 / band->gain;
    const double s1n0 = band->a0 * gi, s1n1 = band->a11 * gi;
    const double s2n1 = band->a12, s3n1 = band->a13, s4n1 = band->a14;
    const double ca0 = band->a0, d1 = band->b1, d2 = band->b2;
    double s1c0 = 0, s1c1 = 0, s2c0 = 0, s2c1 = 0, s3c0 = 0, s3c1 = 0, s4c0 = 0, s4c1 = 0, sum_sq = 0;
    for (size_t i = 0; i < n; i++) {
        const double s = frame[i];
        const double f1 = s1n0 * s + s1c0;
        s1c0 = s1n1 * s + s1c1 - d1 * f1;
        s1c1 = -(d2 * f1);
        const double f2 = ca0 * f1 + s2c0;
        s2c0 = s2n1 * f1 + s2c1 - d1 * f2;
        s2c1 = -(d2 * f2);
        const double f3 = ca0 * f2 + s3c0;
        s3c0 = s3n1 * f2 + s3c1 - d1 * f3;
        s3c1 = -(d2 * f3);
        const double f4 = ca0 * f3 + s4c0;
        s4c0 = s4n1 * f3 + s4c1 - d1 * f4;
        s4c1 = -(d2 * f4);
        sum_sq += f4 * f4;
    }
    const double inv_n = 1.0 / (double)n;
    return sqrt(sum_sq * inv_n);
}

/* gammatone.zig:309-316 */
void vo_hann(size_t size, double *out) {
    double denom = fmax(1.0, (double)size - 1.0);
    for (size_t i = 0; i < size; i++) out[i] = 0.5 - 0.5 * cos(2.0 * M_PI * (double)i / denom);
}

/* gammatone.zig:320-323 */
size_t vo_num_frames(size_t len, size_t window, size_t hop) {
    if (len < window) return 0;
    return 1 + (len - window) / hop;
}

/* gammatone.zig:432-467 (buildSpectrogramBands); out[b*nf + f].  Each
 * frame's windowed samples x[i]*hann[i] run the fresh recurrence; the
 * reference's SIMD lanes are bit-identical to this scalar form
 * (gammatone.zig:579-610). */
int vo_spectrogram(const vo_band *bands, size_t nbands, const double *signal, size_t len,
                   size_t window, size_t hop, const double *hann, double *out) {
    if (len < window || hop == 0) return VO_E_SIGNAL;
    const size_t nf = vo_num_frames(len, window, hop);
    double *w = (double *)malloc(sizeof(double) * window);
    if (!w) return VO_E_OUT_OF_MEMORY;
    for (size_t f = 0; f < nf; f++) {
        for (size_t i = 0; i < window; i++) w[i] = signal[f * hop + i] * hann[i];
        for (size_t b = 0; b < nbands; b++) out[b * nf + f] = vo_band_rms_fresh(&bands[b], w, window);
    }
    free(w);
    return VO_OK;
}

/* ------------------------------------------------------------------------
 * Spectrogram preparation (spectrogram.zig).
 * ---------------------------------------------------------------------- */
/* spectrogram.zig:14-20 */
void vo_db_floor(double *data, size_t n, double floor_db) {
    for (size_t i = 0; i < n; i++) {
        double sample = data[i] == 0.0 ? DBL_EPSILON : fabs(data[i]);
        double db = 10.0 * vo_log10(sample);
        data[i] = db > floor_db ? db : floor_db;
    }
}

/* spectrogram.zig:24-54 */
void vo_raise_floor_per_frame(double *ref, double *deg, size_t nb, size_t nfr, size_t nfd, double rel) {
    size_t mf = nfr < nfd ? nfr : nfd;
    for (size_t c = 0; c < mf; c++) {
        double col_max = -INFINITY;
        for (size_t r = 0; r < nb; r++) {
            double sv = ref[r * nfr + c], ov = deg[r * nfd + c];
            double m = fmax(sv, ov); /* Zig @max: the non-NaN operand */
            if (m > col_max) col_max = m;
        }
        double floor_db = col_max - rel;
        for (size_t r = 0; r < nb; r++) {
            if (ref[r * nfr + c] < floor_db) ref[r * nfr + c] = floor_db;
            if (deg[r * nfd + c] < floor_db) deg[r * nfd + c] = floor_db;
        }
    }
}

/* spectrogram.zig:56-66 */
void vo_subtract_joint_min(double *ref, size_t nr, double *deg, size_t nd) {
    double lo = INFINITY;
    for (size_t i = 0; i < nr; i++) if (ref[i] < lo) lo = ref[i];
    for (size_t i = 0; i < nd; i++) if (deg[i] < lo) lo = deg[i];
    for (size_t i = 0; i < nr; i++) ref[i] -= lo;
    for (size_t i = 0; i < nd; i++) deg[i] -= lo;
}

/* spectrogram.zig:69-80 */
void vo_prepare_for_comparison(double *ref, double *deg, size_t nb, size_t nfr, size_t nfd) {
    vo_db_floor(ref, nb * nfr, -45.0);
    vo_db_floor(deg, nb * nfd, -45.0);
    vo_raise_floor_per_frame(ref, deg, nb, nfr, nfd, 45.0);
    vo_subtract_joint_min(ref, nb * nfr, deg, nb * nfd);
}

/* ------------------------------------------------------------------------
 * Upper envelope (envelope.zig:23-53).
 * ---------------------------------------------------------------------- */
static size_t next_pow_two(size_t x) {
    if (x <= 1) return 1;
    size_t n = 1;
    while (n < x) n <<= 1;
    return n;
}

void vo_upper_envelope(const double *s, size_t l, double *out) {
    if (l == 0) return;
    double sum = 0.0;
    for (size_t i = 0; i < l; i++) sum += s[i];
    const double mean = sum / (double)l;
    size_t n = next_pow_two(l);
    if (n < 32) n = 32;
    double nyquist = 0.0, sign = 1.0;
    for (size_t i = 0; i < l; i++) {
        nyquist += sign * (s[i] - mean);
        sign = -sign;
    }
    const double nyq_over_n = nyquist / (double)n;
    const double eps = 0.000001;
    int pos = 1;
    for (size_t i = 0; i < l; i++) {
        double pc = s[i] - mean;
        double nyq_term = pos ? nyq_over_n : -nyq_over_n;
        double re = 2.0 * (pc - nyq_term) - eps;
        out[i] = fabs(re) + mean;
        pos = !pos;
    }
}

/* ------------------------------------------------------------------------
 * FFT (dsp/fft.zig).  Twiddles W_n^k = (cos(-2πk/n), sin(-2πk/n)) for
 * k < n/2; span-S stages read W_n^{j·n/S} (fft.zig:45-66).  Fusing stages
 * into radix-4 groups (fft.zig:127-166, 285-356) does not change any
 * butterfly's arithmetic, so the plain radix-2 DIF / DIT loops below give
 * the same bits.
 * ---------------------------------------------------------------------- */
typedef struct { size_t n; cpx *tw; } fft_plan;

static int plan_init(fft_plan *p, size_t n) {
    p->n = n;
    size_t half = n / 2 ? n / 2 : 1;
    p->tw = (cpx *)malloc(sizeof(cpx) * half);
    if (!p->tw) return -1;
    const double two_pi = 2.0 * M_PI;
    const double n_f = (double)n;
    for (size_t k = 0; k < n / 2; k++) {
        double angle = -two_pi * (double)k / n_f;
        p->tw[k] = (cpx){cos(angle), sin(angle)};
    }
    return 0;
}
static void plan_free(fft_plan *p) { free(p->tw); p->tw = NULL; }

/* DIF: natural order in, bit-reversed out, unscaled (fft.zig:127-145). */
static void dif_forward(const fft_plan *p, cpx *buf) {
    const size_t n = p->n;
    for (size_t S = n; S >= 2; S >>= 1) {
        const size_t half = S / 2, stride = n / S;
        for (size_t k = 0; k < n; k += S) {
            for (size_t j = 0; j < half; j++) {
                cpx a = buf[k + j], b = buf[k + j + half];
                cpx w = p->tw[j * stride];
                buf[k + j] = c_add(a, b);
                buf[k + j + half] = c_mul(c_sub(a, b), w);
            }
        }
    }
}

/* cmul(x, conj(w)) with the reference's operation order (fft.zig:185-188). */
static inline cpx c_mul_conj(cpx x, cpx w) {
    double wi = -w.im;
    return (cpx){x.re * w.re - x.im * wi, x.re * wi + x.im * w.re};
}

/* DIT: bit-reversed in, natural order out, unscaled (fft.zig:148-166). */
static void dit_inverse(const fft_plan *p, cpx *buf) {
    const size_t n = p->n;
    for (size_t S = 2; S <= n; S <<= 1) {
        const size_t half = S / 2, stride = n / S;
        for (size_t k = 0; k < n; k += S) {
            for (size_t j = 0; j < half; j++) {
                cpx a = buf[k + j];
                cpx b = c_mul_conj(buf[k + j + half], p->tw[j * stride]);
                buf[k + j] = c_add(a, b);
                buf[k + j + half] = c_sub(a, b);
            }
        }
    }
}

void vo_fft_dif_forward(double *z, size_t n) {
    fft_plan p;
    plan_init(&p, n);
    dif_forward(&p, (cpx *)z);
    plan_free(&p);
}
void vo_fft_dit_inverse(double *z, size_t n) {
    fft_plan p;
    plan_init(&p, n);
    dit_inverse(&p, (cpx *)z);
    plan_free(&p);
}

static size_t bit_reverse(size_t x, unsigned log2n) {
    size_t r = 0;
    for (unsigned i = 0; i < log2n; i++) { r = (r << 1) | (x & 1); x >>= 1; }
    return r;
}

/* fft.zig:103-122 (run): permute then DIT-style stages; used by the
 * three-FFT reference lag search (xcorr.zig:71-136, 260-267). */
void vo_fft_natural(double *zi, size_t n, int inverse) {
    cpx *buf = (cpx *)zi;
    fft_plan p;
    plan_init(&p, n);
    unsigned lg = 0;
    while (((size_t)1 << lg) < n) lg++;
    for (size_t i = 0; i < n; i++) {
        size_t j = bit_reverse(i, lg);
        if (j > i) { cpx t = buf[i]; buf[i] = buf[j]; buf[j] = t; }
    }
    for (size_t S = 2; S <= n; S <<= 1) {
        size_t half = S / 2, stride = n / S;
        for (size_t k = 0; k < n; k += S)
            for (size_t j = 0; j < half; j++) {
                cpx a = buf[k + j];
                cpx b = inverse ? c_mul_conj(buf[k + j + half], p.tw[j * stride]) : c_mul(buf[k + j + half], p.tw[j * stride]);
                buf[k + j] = c_add(a, b);
                buf[k + j + half] = c_sub(a, b);
            }
    }
    if (inverse) {
        double inv_n = 1.0 / (double)n;
        for (size_t i = 0; i < n; i++) { buf[i].re *= inv_n; buf[i].im *= inv_n; }
    }
    plan_free(&p);
}

/* xcorr.zig:64-69 */
size_t vo_fft_size_for(size_t biggest) {
    size_t needed = biggest == 0 ? 1 : 2 * biggest - 1;
    size_t n = 1;
    while (n < needed) n <<= 1;
    return n;
}

/* xcorr.zig:182-192 */
static inline cpx cross_term(cpx x, cpx zy) {
    cpx y = {zy.re, -zy.im};
    cpx sum = {x.re + y.re, x.im + y.im};
    cpx dif_c = {x.re - y.re, -(x.im - y.im)};
    cpx idc = {-dif_c.im, dif_c.re};
    return (cpx){0.25 * (sum.re * idc.re - sum.im * idc.im), 0.25 * (sum.re * idc.im + sum.im * idc.re)};
}

/* xcorr.zig:197-229: the circular cross-correlation of s1 and s2 as the
 * real part of z after the DIF forward transform of s1 + i·s2 (zero padded
 * to n), the cross terms in bit-reversed order and the DIT inverse.
 * Returns z (n elements, caller frees) or NULL. */
static cpx *xcorr_z(const double *s1, size_t l1, const double *s2, size_t l2, size_t n) {
    fft_plan p;
    if (plan_init(&p, n)) return NULL;
    cpx *z = (cpx *)calloc(n, sizeof(cpx));
    if (!z) {
        plan_free(&p);
        return NULL;
    }
    size_t m = l1 < l2 ? l1 : l2;
    for (size_t i = 0; i < m; i++) z[i] = (cpx){s1[i], s2[i]};
    for (size_t i = m; i < l1; i++) z[i] = (cpx){s1[i], 0.0};
    for (size_t i = m; i < l2; i++) z[i] = (cpx){0.0, s2[i]};
    dif_forward(&p, z);
    z[0] = cross_term(z[0], z[0]);
    if (n > 1) z[1] = cross_term(z[1], z[1]);
    for (size_t h = 2; h < n; h <<= 1) {
        size_t a = h, b = 2 * h - 1;
        while (a < b) {
            cpx za = z[a], zb = z[b];
            z[a] = cross_term(za, zb);
            z[b] = cross_term(zb, za);
            a++;
            b--;
        }
    }
    dit_inverse(&p, z);
    plan_free(&p);
    return z;
}

int vo_xcorr_re(const double *s1, size_t l1, const double *s2, size_t l2, double *out) {
    size_t biggest = l1 > l2 ? l1 : l2;
    if (biggest == 0) return 0;
    size_t n = vo_fft_size_for(biggest);
    cpx *z = xcorr_z(s1, l1, s2, l2, n);
    if (!z) return -1;
    for (size_t i = 0; i < n; i++) out[i] = z[i].re;
    free(z);
    return 0;
}

/* xcorr.zig:230-251 (bestLagCached): argmax of |Σ_t s1[t+lag]·s2[t]|,
 * first maximum scanning lags -max..-1 then 0..max. */
int64_t vo_best_lag(const double *s1, size_t l1, const double *s2, size_t l2) {
    size_t biggest = l1 > l2 ? l1 : l2;
    if (biggest == 0) return 0;
    size_t n = vo_fft_size_for(biggest);
    cpx *z = xcorr_z(s1, l1, s2, l2, n);
    if (!z) return 0;
    size_t max_lag = biggest - 1;
    double best_abs = -INFINITY;
    int64_t best_lag = 0;
    for (size_t i = n - max_lag; i < n; i++) {
        double a = fabs(z[i].re);
        if (a > best_abs) { best_abs = a; best_lag = (int64_t)i - (int64_t)n; }
    }
    for (size_t k = 0; k <= max_lag; k++) {
        double a = fabs(z[k].re);
        if (a > best_abs) { best_abs = a; best_lag = (int64_t)k; }
    }
    free(z);
    return best_lag;
}

/* xcorr.zig:71-136 (xcorrCircular + bestLagWithPlan): three full FFTs. */
int64_t vo_best_lag_reference(const double *s1, size_t l1, const double *s2, size_t l2) {
    size_t biggest = l1 > l2 ? l1 : l2;
    if (biggest == 0) return 0;
    size_t n = vo_fft_size_for(biggest);
    cpx *a = (cpx *)calloc(n, sizeof(cpx)), *b = (cpx *)calloc(n, sizeof(cpx)), *s = (cpx *)calloc(n, sizeof(cpx));
    for (size_t i = 0; i < l1; i++) a[i].re = s1[i];
    for (size_t i = 0; i < l2; i++) b[i].re = s2[i];
    vo_fft_natural((double *)a, n, 0);
    vo_fft_natural((double *)b, n, 0);
    for (size_t i = 0; i <= n / 2; i++) {
        double ar = a[i].re, ai = a[i].im, br = b[i].re, bi = b[i].im;
        s[i] = (cpx){ar * br + ai * bi, ai * br - ar * bi};
    }
    for (size_t i = 1; i < n / 2; i++) s[n - i] = (cpx){s[i].re, -s[i].im};
    vo_fft_natural((double *)s, n, 1);
    size_t max_lag = biggest - 1;
    double best_abs = -INFINITY;
    int64_t best_lag = 0;
    for (size_t i = n - max_lag; i < n; i++) {
        double v = fabs(s[i].re);
        if (v > best_abs) { best_abs = v; best_lag = (int64_t)i - (int64_t)n; }
    }
    for (size_t k = 0; k <= max_lag; k++) {
        double v = fabs(s[k].re);
        if (v > best_abs) { best_abs = v; best_lag = (int64_t)k; }
    }
    free(a); free(b); free(s);
    return best_lag;
}

/* Brute-force O(L^2) direct correlation, same scan order (test pin only). */
int64_t vo_best_lag_direct(const double *s1, size_t l1, const double *s2, size_t l2) {
    size_t biggest = l1 > l2 ? l1 : l2;
    if (biggest == 0) return 0;
    int64_t max_lag = (int64_t)biggest - 1;
    double best_abs = -INFINITY;
    int64_t best = 0;
    for (int64_t lag = -max_lag; lag <= max_lag; lag++) {
        double acc = 0.0;
        for (size_t t = 0; t < l2; t++) {
            int64_t i = (int64_t)t + lag;
            if (i >= 0 && i < (int64_t)l1) acc += s1[i] * s2[t];
        }
        if (fabs(acc) > best_abs) { best_abs = fabs(acc); best = lag; }
    }
    return best;
}

/* ------------------------------------------------------------------------
 * SPL (audio.zig:132-159).
 * ---------------------------------------------------------------------- */
static double spl_of_energy(double energy, size_t len) {
    if (len == 0) return -INFINITY;
    double rms = sqrt(energy / (double)len);
    return 20.0 * vo_log10(rms / 0.00002);
}

double vo_spl(const double *x, size_t n) {
    double e = 0.0;
    for (size_t i = 0; i < n; i++) e += x[i] * x[i];
    return spl_of_energy(e, n);
}

double vo_spl_factor_for(const double *ref, size_t nr, const double *deg, size_t nd) {
    double ref_spl = vo_spl(ref, nr), deg_spl = vo_spl(deg, nd);
    if (!isfinite(ref_spl) || !isfinite(deg_spl)) return 1.0;
    return pow(10.0, (ref_spl - deg_spl) / 20.0);
}

void vo_scale_to_match_spl(const double *ref, size_t nr, double *deg, size_t nd) {
    double ref_spl = vo_spl(ref, nr), deg_spl = vo_spl(deg, nd);
    if (!isfinite(ref_spl) || !isfinite(deg_spl)) return;
    double factor = pow(10.0, (ref_spl - deg_spl) / 20.0);
    for (size_t i = 0; i < nd; i++) deg[i] *= factor;
}
