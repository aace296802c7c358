// Host-side pieces of ∂moslqo/∂deg (backend.zig:101-110, visqol.zig:
// 714-800): the gradient of the aggregation and the quality mapping with
// respect to each match's per-band statistics, of the per-patch NSIM
// statistics with respect to the comparison-ready spectrogram cells
// (nsim.zig:427-510), of the dB preparation with respect to the raw
// spectrograms (spectrogram.zig:14-80), and of the aligned, SPL-scaled
// degraded signal with respect to the input (audio.zig:132-159,
// realign.zig:91-110).  Every discrete choice of the forward pass (lags,
// matches, realignment, the argmin/argmax cells of the floors) is held
// fixed; f64 throughout.
#include "grad.h"

#include <float.h>
#include <math.h>

#include <algorithm>
#include <numeric>

namespace vq {

namespace {

// the forward's dB preparation constants (spectrogram.zig:14-80, as k_db and
// k_frame_floor apply them)
constexpr double kF64Eps = DBL_EPSILON;
constexpr double kNoiseFloorAbsDb = -45.0;
constexpr double kNoiseFloorRelDb = 45.0;
inline double vq_log10(double x) { return fdlibm_log10(x); }

// The 3x3 window of the NSIM statistics (nsim.zig:19-23) and its
// replicate-padded convolution (conv2d3x3Replicate), forward and adjoint.
constexpr double kW[3][3] = {{VQ_WC, VQ_WE, VQ_WC}, {VQ_WE, VQ_WM, VQ_WE}, {VQ_WC, VQ_WE, VQ_WC}};

void conv3x3(const double* in, double* out, int nrows, int ncols) {
    for (int r = 0; r < nrows; r++)
        for (int c = 0; c < ncols; c++) {
            double acc = 0.0;
            for (int dr = -1; dr <= 1; dr++)
                for (int dc = -1; dc <= 1; dc++) {
                    const int rr = std::min(std::max(r + dr, 0), nrows - 1);
                    const int cc = std::min(std::max(c + dc, 0), ncols - 1);
                    acc += kW[dr + 1][dc + 1] * in[rr * ncols + cc];
                }
            out[r * ncols + c] = acc;
        }
}

// g_in[·] += Σ over outputs of w · g_out (the transpose of conv3x3).
void conv3x3_adjoint(const double* g_out, double* g_in, int nrows, int ncols) {
    for (int r = 0; r < nrows; r++)
        for (int c = 0; c < ncols; c++) {
            const double g = g_out[r * ncols + c];
            if (g == 0.0) continue;
            for (int dr = -1; dr <= 1; dr++)
                for (int dc = -1; dc <= 1; dc++) {
                    const int rr = std::min(std::max(r + dr, 0), nrows - 1);
                    const int cc = std::min(std::max(c + dc, 0), ncols - 1);
                    g_in[rr * ncols + cc] += kW[dr + 1][dc + 1] * g;
                }
        }
}

}  // namespace

void patch_similarity_grad(const double* ref, const double* deg, int nrows, int ncols, const double* g_means,
                           const double* g_stddevs, const double* g_energy, double* g_ref, double* g_deg) {
    const int n = nrows * ncols;
    const double c1 = nsim_c1(), c3 = nsim_c3();
    std::vector<double> rr(n), dd(n), rd(n), mu_r(n), mu_d(n), c_rr(n), c_dd(n), c_rd(n);
    for (int e = 0; e < n; e++) {
        rr[e] = ref[e] * ref[e];
        dd[e] = deg[e] * deg[e];
        rd[e] = ref[e] * deg[e];
    }
    conv3x3(ref, mu_r.data(), nrows, ncols);
    conv3x3(deg, mu_d.data(), nrows, ncols);
    conv3x3(rr.data(), c_rr.data(), nrows, ncols);
    conv3x3(dd.data(), c_dd.data(), nrows, ncols);
    conv3x3(rd.data(), c_rd.data(), nrows, ncols);

    // forward cells, the band means and stddevs (nsim.zig:465-500)
    std::vector<double> sim(n);
    std::vector<double> mean(nrows, 0.0), sd(nrows, 0.0);
    for (int r = 0; r < nrows; r++) {
        double s = 0.0;
        for (int c = 0; c < ncols; c++) {
            const int e = r * ncols + c;
            const double mr = mu_r[e], md = mu_d[e];
            const double rms = mr * mr, dms = md * md, mrd = mr * md;
            const double srs = c_rr[e] - rms, sds = c_dd[e] - dms, srd = c_rd[e] - mrd;
            const double intensity = (2.0 * mrd + c1) / (rms + dms + c1);
            const double prod = srs * sds;
            const double den = prod < 0.0 ? c3 : sqrt(prod) + c3;
            sim[e] = intensity * ((srd + c3) / den);
            s += sim[e];
        }
        mean[r] = s / (double)ncols;
        if (ncols > 1) {
            double ss = 0.0;
            for (int c = 0; c < ncols; c++) {
                const double d = sim[r * ncols + c] - mean[r];
                ss += d * d;
            }
            sd[r] = sqrt(ss / ((double)ncols - 1.0));
        }
    }

    // ∂/∂sim: the mean's and the sample stddev's slopes
    std::vector<double> g_mr(n, 0.0), g_md(n, 0.0), g_crr(n, 0.0), g_cdd(n, 0.0), g_crd(n, 0.0);
    for (int r = 0; r < nrows; r++) {
        const double gm = g_means ? g_means[r] : 0.0;
        const double gs = (g_stddevs && ncols > 1 && sd[r] > 0.0) ? g_stddevs[r] / (((double)ncols - 1.0) * sd[r]) : 0.0;
        const double ge = g_energy ? g_energy[r] / (double)ncols : 0.0;
        for (int c = 0; c < ncols; c++) {
            const int e = r * ncols + c;
            g_deg[e] += ge;  // energy: the band mean of the degraded patch
            const double g = gm / (double)ncols + gs * (sim[e] - mean[r]);
            if (g == 0.0) continue;
            const double mr = mu_r[e], md = mu_d[e];
            const double rms = mr * mr, dms = md * md, mrd = mr * md;
            const double srs = c_rr[e] - rms, sds = c_dd[e] - dms, srd = c_rd[e] - mrd;
            const double a = rms + dms + c1;
            const double intensity = (2.0 * mrd + c1) / a;
            const double prod = srs * sds;
            const double den = prod < 0.0 ? c3 : sqrt(prod) + c3;
            const double structure = (srd + c3) / den;
            const double g_int = g * structure, g_str = g * intensity;
            double g_mrd = g_int * 2.0 / a;
            double g_rms = -g_int * (2.0 * mrd + c1) / (a * a);
            double g_dms = g_rms;
            const double g_srd = g_str / den;
            const double g_den = -g_str * (srd + c3) / (den * den);
            const double g_prod = prod > 0.0 ? g_den / (2.0 * sqrt(prod)) : 0.0;
            const double g_srs = g_prod * sds, g_sds = g_prod * srs;
            g_crr[e] += g_srs;
            g_rms -= g_srs;
            g_cdd[e] += g_sds;
            g_dms -= g_sds;
            g_crd[e] += g_srd;
            g_mrd -= g_srd;
            g_mr[e] += 2.0 * mr * g_rms + md * g_mrd;
            g_md[e] += 2.0 * md * g_dms + mr * g_mrd;
        }
    }
    // back through the window convolutions
    std::vector<double> t(n);
    std::fill(t.begin(), t.end(), 0.0);
    conv3x3_adjoint(g_mr.data(), t.data(), nrows, ncols);
    for (int e = 0; e < n; e++) g_ref[e] += t[e];
    std::fill(t.begin(), t.end(), 0.0);
    conv3x3_adjoint(g_md.data(), t.data(), nrows, ncols);
    for (int e = 0; e < n; e++) g_deg[e] += t[e];
    std::fill(t.begin(), t.end(), 0.0);
    conv3x3_adjoint(g_crr.data(), t.data(), nrows, ncols);
    for (int e = 0; e < n; e++) g_ref[e] += 2.0 * ref[e] * t[e];
    std::fill(t.begin(), t.end(), 0.0);
    conv3x3_adjoint(g_cdd.data(), t.data(), nrows, ncols);
    for (int e = 0; e < n; e++) g_deg[e] += 2.0 * deg[e] * t[e];
    std::fill(t.begin(), t.end(), 0.0);
    conv3x3_adjoint(g_crd.data(), t.data(), nrows, ncols);
    for (int e = 0; e < n; e++) {
        g_ref[e] += deg[e] * t[e];
        g_deg[e] += ref[e] * t[e];
    }
}

void prepare_grad(const double* raw_ref, const double* raw_deg, int nb, int nfr, int nfd, const double* g_ref_prep,
                  const double* g_deg_prep, double* g_raw_ref, double* g_raw_deg) {
    const int nr = nb * nfr, nd = nb * nfd;
    // forward: dB with the absolute floor, then the per-frame floor
    std::vector<double> v_ref(nr), v_deg(nd);
    std::vector<char> abs_ref(nr), abs_deg(nd);  // clamped at the absolute floor
    auto db = [](double x, double& v, char& clamped) {
        const double s = x == 0.0 ? kF64Eps : fabs(x);
        const double d = 10.0 * vq_log10(s);
        clamped = d < kNoiseFloorAbsDb;
        v = clamped ? kNoiseFloorAbsDb : d;
    };
    for (int e = 0; e < nr; e++) db(raw_ref[e], v_ref[e], abs_ref[e]);
    for (int e = 0; e < nd; e++) db(raw_deg[e], v_deg[e], abs_deg[e]);
    const int shared = std::min(nfr, nfd);
    std::vector<double> v2_ref(v_ref), v2_deg(v_deg);
    std::vector<int> col_arg(shared);    // argmax cell of each shared frame: band, +nb for the degraded side
    std::vector<char> fl_ref(nr, 0), fl_deg(nd, 0);  // raised to the frame floor
    for (int c = 0; c < shared; c++) {
        double col_max = -INFINITY;
        int arg = 0;
        for (int r = 0; r < nb; r++) {
            const double sv = v_ref[r * nfr + c], ov = v_deg[r * nfd + c];
            const double m = sv >= ov ? sv : ov;
            if (m > col_max) {
                col_max = m;
                arg = sv >= ov ? r : nb + r;
            }
        }
        col_arg[c] = arg;
        const double floor_db = col_max - kNoiseFloorRelDb;
        for (int r = 0; r < nb; r++) {
            if (v_ref[r * nfr + c] < floor_db) {
                v2_ref[r * nfr + c] = floor_db;
                fl_ref[r * nfr + c] = 1;
            }
            if (v_deg[r * nfd + c] < floor_db) {
                v2_deg[r * nfd + c] = floor_db;
                fl_deg[r * nfd + c] = 1;
            }
        }
    }
    // joint minimum: the first cell holding it, reference cells first
    int min_side = 0, min_e = 0;
    double lo = INFINITY;
    for (int e = 0; e < nr; e++)
        if (v2_ref[e] < lo) {
            lo = v2_ref[e];
            min_side = 0;
            min_e = e;
        }
    for (int e = 0; e < nd; e++)
        if (v2_deg[e] < lo) {
            lo = v2_deg[e];
            min_side = 1;
            min_e = e;
        }

    // backward
    std::vector<double> g2_ref(g_ref_prep, g_ref_prep + nr), g2_deg(g_deg_prep, g_deg_prep + nd);
    double total = 0.0;
    for (int e = 0; e < nr; e++) total += g_ref_prep[e];
    for (int e = 0; e < nd; e++) total += g_deg_prep[e];
    if (nr + nd > 0) (min_side == 0 ? g2_ref[min_e] : g2_deg[min_e]) -= total;
    std::vector<double> g1_ref(g2_ref), g1_deg(g2_deg);
    for (int c = 0; c < shared; c++) {
        double g_floor = 0.0;
        for (int r = 0; r < nb; r++) {
            if (fl_ref[r * nfr + c]) {
                g_floor += g1_ref[r * nfr + c];
                g1_ref[r * nfr + c] = 0.0;
            }
            if (fl_deg[r * nfd + c]) {
                g_floor += g1_deg[r * nfd + c];
                g1_deg[r * nfd + c] = 0.0;
            }
        }
        const int a = col_arg[c];
        if (a < nb) g1_ref[a * nfr + c] += g_floor;
        else g1_deg[(a - nb) * nfd + c] += g_floor;
    }
    // dB: 10·log10(x) has slope 10 / (x ln 10); a clamped or zero cell none
    const double k = 10.0 / log(10.0);
    for (int e = 0; e < nr; e++)
        g_raw_ref[e] = (abs_ref[e] || raw_ref[e] == 0.0) ? 0.0 : g1_ref[e] * k / raw_ref[e];
    for (int e = 0; e < nd; e++)
        g_raw_deg[e] = (abs_deg[e] || raw_deg[e] == 0.0) ? 0.0 : g1_deg[e] * k / raw_deg[e];
}

void aligned_to_input_grad(const double* deg, long long deg_len, long long shift, long long aligned_len,
                           double factor, double deg_energy, bool scaled, const double* g_aligned, double* g_deg) {
    for (long long j = 0; j < deg_len; j++) g_deg[j] = 0.0;
    // aligned[t] = factor · deg[t - shift] where 0 <= t - shift < deg_len
    double dot = 0.0;  // Σ_t g_aligned[t] · deg[t - shift]
    for (long long t = 0; t < aligned_len; t++) {
        const long long j = t - shift;
        if (j < 0 || j >= deg_len) continue;
        g_deg[j] = factor * g_aligned[t];
        dot += g_aligned[t] * deg[j];
    }
    // factor ∝ (Σ deg² over the aligned samples)^(-1/2)
    if (!scaled || !(deg_energy > 0.0)) return;
    const double c = factor * dot / deg_energy;
    for (long long t = 0; t < aligned_len; t++) {
        const long long j = t - shift;
        if (j < 0 || j >= deg_len) continue;
        g_deg[j] -= c * deg[j];
    }
}

void result_grad(const std::vector<Match>& matches, int nb, double frame_duration, const MapperGrad& mapper,
                 std::vector<double>& out) {
    const size_t n = matches.size();
    out.assign(n * nb * 3, 0.0);
    if (n == 0) return;
    const double n_f = (double)n;
    std::vector<double> fvnsim(nb, 0.0), fvdeg(nb, 0.0), fvnsim10(nb, 0.0), fstd(nb, 0.0), variance(nb, 0.0);
    for (const Match& m : matches)
        for (int b = 0; b < nb; b++) {
            fvnsim[b] += m.means[b];
            fvdeg[b] += m.energy[b];
        }
    for (int b = 0; b < nb; b++) {
        fvnsim[b] /= n_f;
        fvdeg[b] /= n_f;
    }
    double sum = 0.0;
    for (int b = 0; b < nb; b++) sum += fvnsim[b];
    if (sum / (double)nb < 0.15) return;  // the score is pinned to 1
    const size_t k = std::max<size_t>(1, (size_t)trunc(n_f * 0.10));
    std::vector<std::vector<size_t>> lowest(nb);
    for (int b = 0; b < nb; b++) {
        std::vector<size_t> idx(n);
        std::iota(idx.begin(), idx.end(), 0);
        std::stable_sort(idx.begin(), idx.end(),
                         [&](size_t x, size_t y) { return matches[x].means[b] < matches[y].means[b]; });
        idx.resize(std::min(k, n));
        double acc = 0.0;
        for (size_t i : idx) acc += matches[i].means[b];
        fvnsim10[b] = acc / (double)k;
        lowest[b] = idx;
    }
    std::vector<double> counts(n);
    double total = 0.0;
    for (size_t i = 0; i < n; i++) {
        counts[i] = (double)(long long)ceil((matches[i].ref_end - matches[i].ref_start) / frame_duration);
        total += counts[i];
    }
    for (int b = 0; b < nb; b++) {
        double c = 0.0;
        for (size_t i = 0; i < n; i++) {
            c += (counts[i] - 1.0) * matches[i].stddevs[b] * matches[i].stddevs[b];
            c += counts[i] * matches[i].means[b] * matches[i].means[b];
        }
        variance[b] = (c - fvnsim[b] * fvnsim[b] * total) / (total - 1.0);
        fstd[b] = variance[b] < 0.0 ? 0.0 : sqrt(variance[b]);
    }
    std::vector<double> g0(nb, 0.0), g1(nb, 0.0), g2(nb, 0.0), g3(nb, 0.0);
    mapper(fvnsim.data(), fvnsim10.data(), fstd.data(), fvdeg.data(), g0.data(), g1.data(), g2.data(), g3.data());
    for (int b = 0; b < nb; b++) {
        const double g_v = variance[b] > 0.0 ? g2[b] / (2.0 * fstd[b]) : 0.0;
        const double g_fv = g0[b] - g_v * 2.0 * fvnsim[b] * total / (total - 1.0);
        for (size_t i = 0; i < n; i++) {
            double* o = &out[(i * nb + b) * 3];
            o[0] += g_fv / n_f + g_v * 2.0 * counts[i] * matches[i].means[b] / (total - 1.0);
            o[1] += g_v * 2.0 * (counts[i] - 1.0) * matches[i].stddevs[b] / (total - 1.0);
            o[2] += g3[b] / n_f;
        }
        for (size_t i : lowest[b]) out[(i * nb + b) * 3] += g1[b] / (double)k;
    }
}

void SvrModel::gradient(const double* x, double* g) const {
    for (int k = 0; k < nf_; k++) g[k] = 0.0;
    for (size_t i = 0; i < coef_.size(); i++) {
        const double* sv = &sv_[i * nf_];
        double dist = 0.0;
        for (int k = 0; k < nf_; k++) {
            const double d = sv[k] - x[k];
            dist += d * d;
        }
        const double w = coef_[i] * exp(-gamma_ * dist) * -2.0 * gamma_;
        for (int k = 0; k < nf_; k++) g[k] += w * (x[k] - sv[k]);
    }
}

}  // namespace vq
