// Host-side planning for the gfx950 pipeline: gammatone filter design,
// framing, FFT twiddles, SPL, patch indices, VAD decisions, search windows
// and realignment geometry.  Operation orders follow the cited reference
// lines so the device sees the same coefficients and decisions as the
// reference CPU path.
#include <math.h>

#include <algorithm>
#include <thread>

#include "host.h"

namespace vq {

namespace {

struct Cx {
    double re, im;
};
inline Cx add(Cx a, Cx b) { return {a.re + b.re, a.im + b.im}; }
inline Cx sub(Cx a, Cx b) { return {a.re - b.re, a.im - b.im}; }
inline Cx mul(Cx a, Cx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline Cx scale(Cx a, double s) { return {a.re * s, a.im * s}; }
inline Cx quot(Cx a, Cx b) {
    const double den = b.re * b.re + b.im * b.im;
    return {(a.re * b.re + a.im * b.im) / den, (a.im * b.re - a.re * b.im) / den};
}
inline Cx expc(Cx z) {
    const double e = exp(z.re);
    return {e * cos(z.im), e * sin(z.im)};
}

constexpr double kEarQ = 9.26449;  // erb.zig:8
constexpr double kMinBw = 24.7;    // erb.zig:10

// gammatone.zig:76-153: 4 biquads sharing the denominator; the gain is the
// magnitude of a complex product, evaluated left to right.
void design_band(double cf, double t, double pi, double* o) {
    const double erb = cf / kEarQ + kMinBw;
    const double big_b = 1.019 * 2.0 * pi * erb;
    const double exp_bt = exp(big_b * t);
    const double inv_exp_bt = 1.0 / exp_bt;
    const double exp_neg_2bt = exp(-2.0 * big_b * t);
    const double omega = 2.0 * cf * pi * t;
    const double cos_w = cos(omega), sin_w = sin(omega);
    const double sin_t = sin_w * t;
    const double root_p = sqrt(3.0 + pow(2.0, 1.5));
    const double root_m = sqrt(3.0 - pow(2.0, 1.5));
    const double b_pos = sin_t * 2.0 * root_p;
    const double b_neg = sin_t * 2.0 * root_m;
    const double a_term = cos_w * 2.0 * t;
    o[0] = t;
    o[1] = -(a_term * inv_exp_bt + b_pos * inv_exp_bt) / 2.0;
    o[2] = -(a_term * inv_exp_bt - b_pos * inv_exp_bt) / 2.0;
    o[3] = -(a_term * inv_exp_bt + b_neg * inv_exp_bt) / 2.0;
    o[4] = -(a_term * inv_exp_bt - b_neg * inv_exp_bt) / 2.0;
    o[5] = -2.0 * cos_w * inv_exp_bt;
    o[6] = exp_neg_2bt;

    const Cx x_exp = expc(scale(Cx{0.0, 1.0}, 4.0 * cf * pi * t));
    const Cx x01 = scale(x_exp, -2.0 * t);
    const Cx x02 = scale(expc(Cx{-big_b * t, 2.0 * cf * pi * t}), 2.0 * t);
    const Cx xc{cos_w, 0.0}, xs{sin_w, 0.0};
    const double s1 = sqrt(3.0 - pow(2.0, 1.5));
    const double s2 = sqrt(3.0 + pow(2.0, 1.5));
    const Cx x1 = add(x01, mul(x02, sub(xc, scale(xs, s1))));
    const Cx x2 = add(x01, mul(x02, add(xc, scale(xs, s1))));
    const Cx x3 = add(x01, mul(x02, sub(xc, scale(xs, s2))));
    const Cx x4 = add(x01, mul(x02, add(xc, scale(xs, s2))));
    const Cx term1 = quot(Cx{-2.0, 0.0}, Cx{exp(2.0 * big_b * t), 0.0});
    const Cx term2 = scale(x_exp, -2.0);
    const Cx term3 = scale(quot(add(Cx{1.0, 0.0}, x_exp), Cx{exp_bt, 0.0}), 2.0);
    const Cx x5 = add(add(term1, term2), term3);
    const Cx ratio = quot(mul(mul(x1, x2), mul(x3, x4)), mul(mul(x5, x5), mul(x5, x5)));
    o[7] = sqrt(ratio.re * ratio.re + ratio.im * ratio.im);
}

}  // namespace

// visqol.zig:346-359 with gammatone.zig:41-70 and dsp/erb.zig:30-52.
RateParams make_rate_params(int mode, uint32_t sample_rate) {
    RateParams p;
    p.sample_rate = sample_rate;
    p.nb = mode == 0 ? kBandsAudio : kBandsSpeech;
    p.patch = mode == 0 ? kPatchAudio : kPatchSpeech;
    const double sr = (double)sample_rate;
    const double high = std::min(mode == 0 ? sr / 2.0 : 8000.0, sr / 2.0);
    const double low = 50.0;
    // ERB-uniform centre frequencies, descending.
    std::vector<double> desc(p.nb);
    {
        const double q = kEarQ, m = kMinBw;
        const double a = -(q * m), b = -log(high + q * m), c = log(low + q * m), d = high + q * m;
        const double e = (b + c) / (double)p.nb;
        for (int i = 0; i < p.nb; i++) desc[i] = a + exp(((double)i + 1.0) * e) * d;
    }
    const double t = 1.0 / sr;
    p.cf.resize(p.nb);
    p.coef.resize(p.nb);
    p.raw_band.resize(8 * p.nb);
    for (int i = 0; i < p.nb; i++) {
        const int asc = p.nb - 1 - i;
        double* o = &p.raw_band[8 * asc];
        design_band(desc[i], t, M_PI, o);
        p.cf[asc] = desc[i];
        // gammatone.zig:268-278
        const double gi = 1.0 / o[7];
        p.coef[asc] = BandCoef{o[0] * gi, o[1] * gi, o[2], o[3], o[4], o[0], o[5], o[6]};
    }
    // FIR taps: (s1n0 + s1n1 z^-1)(ca0 + s2n1 z^-1)(ca0 + s3n1 z^-1)(ca0 + s4n1 z^-1),
    // multiplied out in extended precision and rounded once.
    p.fir.resize(p.nb);
    for (int b = 0; b < p.nb; b++) {
        const BandCoef& k = p.coef[b];
        long double t[5] = {(long double)k.s1n0, (long double)k.s1n1, 0, 0, 0};
        const long double n1[3] = {(long double)k.s2n1, (long double)k.s3n1, (long double)k.s4n1};
        for (int st = 0; st < 3; st++) {
            long double q[5] = {0, 0, 0, 0, 0};
            for (int i = 0; i <= st + 1; i++) {
                q[i] += t[i] * (long double)k.ca0;
                q[i + 1] += t[i] * n1[st];
            }
            for (int i = 0; i < 5; i++) t[i] = q[i];
        }
        p.fir[b] = FirCoef{(double)t[0], (double)t[1], (double)t[2], (double)t[3], (double)t[4], k.d1, k.d2, 0.0};
    }
    p.window = (int)round(sr * 0.08);
    p.hop = (int)((double)p.window * 0.25);
    if (p.hop == 0) p.hop = 1;
    p.hann.resize(p.window);
    const double denom = std::max(1.0, (double)p.window - 1.0);
    for (int i = 0; i < p.window; i++) p.hann[i] = 0.5 - 0.5 * cos(2.0 * M_PI * (double)i / denom);
    p.frame_duration = (double)p.hop / sr;
    return p;
}

// dsp/fft.zig:55-61
std::vector<double> make_twiddles(long long n) {
    std::vector<double> tw(2 * std::max(1LL, n / 2));
    const double two_pi = 2.0 * M_PI;
    const double n_f = (double)n;
    for (long long k = 0; k < n / 2; k++) {
        const double angle = -two_pi * (double)k / n_f;
        tw[2 * k] = cos(angle);
        tw[2 * k + 1] = sin(angle);
    }
    return tw;
}

// Twiddles of every radix-2 span S = 2, 4, ..., n in one table: W_S^j at
// complex index S/2 + j (j < S/2), so that a stage's twiddles for
// consecutive j are consecutive in memory.  Index 0 is unused.  The top
// span holds W_n^k as make_twiddles computes it (on up to `threads` host
// threads); every smaller span copies W_S^j = W_n^{j·n/S} from it.  The
// angle -2π·(j·n/S)/n differs from -2π·j/S only by exact power-of-two
// scalings, so the table for n starts with the table of every smaller
// power of two, bit for bit.
std::vector<double> make_level_twiddles(long long n, int threads) {
    std::vector<double> t(2 * std::max(2LL, n), 0.0);
    const double two_pi = 2.0 * M_PI;
    const double n_f = (double)n;
    double* top = t.data() + n;  // complex index n/2
    const long long half = n / 2;
    auto fill = [&](long long k0, long long k1) {
        for (long long k = k0; k < k1; k++) {
            const double angle = -two_pi * (double)k / n_f;
            top[2 * k] = cos(angle);
            top[2 * k + 1] = sin(angle);
        }
    };
    const long long nt = std::max(1LL, std::min<long long>(threads, half >> 16));
    run_striped((int)nt, [&](int i) { fill(half * i / nt, half * (i + 1) / nt); });
    for (long long S = 2; S < n; S <<= 1)
        for (long long j = 0; j < S / 2; j++) {
            t[2 * (S / 2 + j)] = top[2 * (j * (n / S))];
            t[2 * (S / 2 + j) + 1] = top[2 * (j * (n / S)) + 1];
        }
    return t;
}

long long fft_size_for(long long biggest) {
    const long long need = biggest == 0 ? 1 : 2 * biggest - 1;
    long long n = 1;
    while (n < need) n <<= 1;
    return n;
}

int num_frames(long long len, int window, int hop) {
    if (len < window) return 0;
    return (int)(1 + (len - window) / hop);
}

double spl_of_energy(double energy, long long len) {
    if (len == 0) return -INFINITY;
    const double rms = sqrt(energy / (double)len);
    return 20.0 * fdlibm_log10(rms / 0.00002);
}

double spl_factor(double ref_spl, double deg_spl) {
    if (!isfinite(ref_spl) || !isfinite(deg_spl)) return 1.0;
    return pow(10.0, (ref_spl - deg_spl) / 20.0);
}

bool ref_patch_indices(int nf, int patch, std::vector<int>& out) {
    out.clear();
    const int init = patch / 2;
    if (nf < patch + init) return false;
    const int max_index = init < nf - patch ? nf - patch : init + 1;
    for (int i = init; i < max_index; i += patch) out.push_back(i - 1);
    return true;
}

VadPlan vad_plan(int nf, int patch, int window) {
    VadPlan v;
    const double frame_size = (double)window * 0.25;
    v.frame_len = (int)frame_size;
    const long long patch_sample_len = (long long)((double)patch * frame_size);
    v.first = patch / 2 - 1;
    if (nf <= v.first || v.frame_len == 0) return v;
    v.patch_count = (nf - v.first) / patch;
    const long long total = (long long)v.patch_count * patch_sample_len;
    v.n_chunks = (int)(total / v.frame_len);
    return v;
}

// vad.zig:50-80 from the per-chunk RMS decisions.
void vad_patches(const VadPlan& vp, const uint8_t* active, int patch, std::vector<int>& out) {
    out.clear();
    if (vp.patch_count == 0) return;
    const int vad_len = std::max(vp.n_chunks, 2);
    std::vector<double> vad(vad_len, 1.0);
    for (int i = 2; i < vp.n_chunks; i++) {
        const bool prev_silent = !active[i - 1] && !active[i - 2];
        vad[i] = (!active[i] && prev_silent) ? 0.0 : 1.0;
    }
    int idx = vp.first;
    for (int p = 0; p < vp.patch_count; p++) {
        double fva = 0.0;
        for (int k = p * patch; k < p * patch + patch; k++)
            if (k < vad_len) fva += vad[k];
        if (fva >= 1.0) out.push_back(idx);
        idx += patch;
    }
}

bool plan_search(const std::vector<int>& idx, int nf_ref, int nf_deg, int patch, int radius, int& np,
                 std::vector<int>& lo, std::vector<int>& hi) {
    int n = (int)idx.size();
    while (n > 0) {  // dp.zig:73-87
        const int last = idx[n - 1];
        if (last < patch / 2) break;
        if (last - patch / 2 <= nf_deg) break;
        n--;
    }
    if (n == 0 || nf_deg == 0) return false;
    for (int i = 0; i < n; i++)
        if (idx[i] + patch > nf_ref) return false;
    const long long sw = (long long)radius * patch, last = (long long)nf_deg - 1;
    lo.resize(n);
    hi.resize(n);
    for (int i = 0; i < n; i++) {
        const long long r = idx[i];
        const long long l = std::max(0LL, r - sw);
        lo[i] = (int)l;
        hi[i] = (int)std::max(l, std::min(r + sw, last));
    }
    np = n;
    return true;
}

Layout slice_layout(long long n, uint32_t sr, double start, double end) {
    const double sr_f = (double)sr;
    const long long start_idx = std::max(0LL, (long long)(start * sr_f));
    const long long end_idx = std::min(n - 1, (long long)(end * sr_f));
    const long long real_len = end_idx > start_idx ? end_idx - start_idx : 0;
    const double end_diff = end * sr_f - (double)n;
    const long long post = end_diff > 0.0 ? (long long)end_diff : 0;
    const long long pre = start < 0.0 ? (long long)(-start * sr_f) : 0;
    Layout l;
    l.pad = pre + post;
    l.base = real_len > 0 ? start_idx : 0;
    l.len = real_len;
    return l;
}

Geometry trim_geometry(long long r, long long d, long long lag, uint32_t sr) {
    const long long alag = lag < 0 ? -lag : lag;
    const long long shift = (lag == 0 || alag > r / 2) ? 0 : lag;
    const long long dl = shift < 0 ? d - std::min(-shift, d) : d + shift;
    long long head = 0, len = r;
    if (r > dl) {
        len = dl;
    } else if (r < dl && shift > 0) {
        const double sr_f = (double)sr;
        head = std::max((long long)(((double)shift / sr_f) * sr_f), 0LL);
        len = r - head;
    }
    Geometry g;
    g.len = len;
    g.ref_off = head;
    g.deg_off = head - shift;
    g.lag_seconds = shift == 0 ? 0.0 : (double)shift / (double)sr;
    return g;
}

}  // namespace vq
