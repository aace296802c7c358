// The gradient pass of a chunk (backend.zig:101-110, pipeline.zig:842-1063):
// ∂moslqo/∂deg with every discrete choice of the forward pass held fixed
// (the lags, the DP's matches, the realignment decisions, the floors'
// argmax and argmin cells).
//
//   host    the aggregation and quality mapping, back to each match's
//           per-band statistics (visqol.zig:714-800);
//   host    the NSIM statistics back to the comparison-ready patch cells
//           (nsim.zig:427-510), per full-signal patch or accepted
//           realignment row;
//   host    the dB preparation back to the raw spectrograms
//           (spectrogram.zig:14-80), per pair and per row; rows whose
//           degraded frames were copied from the full-signal grid add into
//           its frames;
//   device  every degraded frame with a nonzero slope through the
//           gammatone adjoint, then the frames' overlap summed into the
//           aligned degraded signal's slope in a fixed order;
//   host    the lag shift and the SPL scale back to the input samples
//           (audio.zig:132-159, realign.zig:91-110).
//
// The host stages are small (tens of patches of 32 x 20 cells per pair);
// the device stage is the one that scales with the signal.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <thread>

#include "engine.h"
#include "engine_state.h"
#include "grad.h"

namespace vq {

namespace {

// Frames per adjoint launch: bounds the checkpoint scratch.
constexpr int kAdjBatch = 2048;
// Output samples per accumulation block.
constexpr int kAccLen = 1024;

bool column_nonzero(const double* g, int nb, int stride) {
    for (int b = 0; b < nb; b++)
        if (g[(size_t)b * stride] != 0.0) return true;
    return false;
}

}  // namespace

void Engine::mapper_grad(int nb, const double* f0, const double* f1, const double* f2, const double* f3, double* g0,
                         double* g1, double* g2, double* g3, int threads) const {
    double* g[4] = {g0, g1, g2, g3};
    for (double* x : g) std::fill(x, x + nb, 0.0);
    if (opts_.mode == 0) {
        // svr.zig kernel exp(-γ|sv - x|²), zero where the score is clamped
        const double y = svr_->predict(f0);
        if (y <= 1.0 || y >= 5.0) return;
        svr_->gradient(f0, g0);
    } else if (opts_.lattice) {
        // piecewise linear and evaluated in f32: central differences with
        // steps of many f32 ulps (visqol.zig:132-146), one feature at a
        // time, spread over `threads` host threads
        const double* f[4] = {f0, f1, f2, f3};
        auto slope = [&](int idx) {
            double x[4][kMaxBands];
            for (int k = 0; k < 4; k++) std::copy(f[k], f[k] + nb, x[k]);
            const int k = idx / nb, b = idx % nb;
            const double v = x[k][b];
            const double h = 1e-3 * std::max(1.0, fabs(v));
            double up = 0.0, down = 0.0;
            x[k][b] = v + h;
            const bool ok_up = lattice_predict(*lattice_, x[0], x[1], x[2], x[3], up);
            x[k][b] = v - h;
            const bool ok_down = lattice_predict(*lattice_, x[0], x[1], x[2], x[3], down);
            g[k][b] = ok_up && ok_down ? (up - down) / (2.0 * h) : 0.0;
        };
        const int total = 4 * nb;
        if (threads > 1) {
            run_striped(threads, [&](int t) {
                for (int idx = t; idx < total; idx += threads) slope(idx);
            });
        } else {
            for (int idx = 0; idx < total; idx++) slope(idx);
        }
    } else {
        // exponential speech mapping (visqol.zig:117-131)
        double sum = 0.0;
        for (int b = 0; b < nb; b++) sum += f0[b];
        const double mean = sum / (double)nb;
        const double bexp = 0.0154302525, x0 = -361.063949;
        const double scale = opts_.unscaled_speech ? 1.0 : (double)1.245063f;
        const double mos = fmax(1.0, fmin((-262.847869 + exp(bexp * (mean - x0))) * scale, 5.0));
        if (mos <= 1.0 || mos >= 5.0) return;
        const double slope = scale * bexp * exp(bexp * (mean - x0)) / (double)nb;
        std::fill(g0, g0 + nb, slope);
    }
}

void Engine::frame_grad(uint32_t sr, const double* frame, const double* g, int lo, int hi, double* out) {
    arena_.reset();
    pinned_.reset();
    DevRate& dr = dev_rate(sr);
    const int nb = dr.p.nb, W = dr.p.window;
    const double* d_sig = upload(std::vector<double>(frame, frame + W));
    const double* d_g = upload(std::vector<double>(g, g + nb));
    double* d_out = (double*)arena_.alloc(sizeof(double) * W);
    double* d_ck = (double*)arena_.alloc(sizeof(double) * gammatone_adjoint_scratch(1, W));
    const FrameGradJob* d_job = upload(std::vector<FrameGradJob>{FrameGradJob{d_sig, 0, d_g, d_out, 1, lo, hi, 0}});
    launch_gammatone_adjoint(d_job, 1, dr.coef, dr.hann, nb, W, d_ck, stream_);
    double* p = (double*)pinned_.alloc(sizeof(double) * W);
    readback(p, d_out, sizeof(double) * W);
    sync();
    memcpy(out, p, sizeof(double) * W);
}

void Engine::backward(const PairInput* pairs, int n, uint32_t sr, PairResult* out, std::vector<PairState>& st,
                      double* const* grads) {
    DevRate& dr = dev_rate(sr);
    const RateParams& rp = dr.p;
    const int nb = rp.nb, W = rp.window, hop = rp.hop, S = rp.patch;
    std::vector<int> todo;
    for (int i = 0; i < n; i++) {
        if (!grads[i]) continue;
        std::fill(grads[i], grads[i] + std::max(0LL, pairs[i].deg_len), 0.0);
        if (st[i].live && out[i].status == 0) todo.push_back(i);
    }
    if (todo.empty()) return;

    // ---------------------------------------------------------------- readback
    struct RowHost {
        std::vector<double> prep_r, prep_d, raw_r, raw_d;
    };
    struct Host {
        std::vector<double> raw_ref, raw_deg, prep_ref, prep_deg, deg;
        std::vector<RowHost> rows;  // parallel to PairState::rows
        std::vector<double> g_raw_deg;
        std::vector<std::vector<double>> g_row_deg;  // raw-level slopes of rows computed on their own slices
        const double* g_al = nullptr;                // the aligned signal's slopes (pinned readback)
    };
    std::vector<Host> hs(todo.size());
    std::vector<std::pair<const double*, std::vector<double>*>> staged;
    auto fetch = [&](const double* d, size_t count, std::vector<double>& dst) {
        dst.resize(count);
        if (count == 0) return;
        double* p = (double*)pinned_.alloc(sizeof(double) * count);
        readback(p, d, sizeof(double) * count);
        staged.emplace_back(p, &dst);
    };
    for (size_t q = 0; q < todo.size(); q++) {
        const int i = todo[q];
        PairState& s = st[i];
        Host& h = hs[q];
        fetch(s.raw_ref, (size_t)nb * s.nf_ref, h.raw_ref);
        fetch(s.raw_deg, (size_t)nb * s.nf_deg, h.raw_deg);
        fetch(s.spec_ref, (size_t)nb * s.nf_ref, h.prep_ref);
        fetch(s.spec_deg, (size_t)nb * s.nf_deg, h.prep_deg);
        if (pairs[i].on_device) fetch(pairs[i].deg, (size_t)pairs[i].deg_len, h.deg);
        h.rows.resize(s.rows.size());
        for (size_t k = 0; k < s.rows.size(); k++) {
            const RowTrace& tr = s.rows[k];
            if (!out[i].realign_accepted[tr.match]) continue;
            const size_t cells = (size_t)nb * tr.nf;
            fetch(tr.spec_r, cells, h.rows[k].prep_r);
            fetch(tr.spec_d, cells, h.rows[k].prep_d);
            fetch(tr.raw_r, cells, h.rows[k].raw_r);
            fetch(tr.raw_d, cells, h.rows[k].raw_d);
        }
    }
    sync();
    for (auto& sv : staged) memcpy(sv.second->data(), sv.first, sizeof(double) * sv.second->size());

    // ---------------------------------------------------------------- host: stats → raw spectrograms
    // Pairs spread over host threads; with fewer pairs than threads, each
    // pair's lattice slopes (4·nb pairs of predictions) take the rest.
    const int nthreads = host_threads(todo.size(), 1);
    const int inner_threads = opts_.mode == 1 && opts_.lattice
                                  ? std::max(1, host_threads(16, 1) / std::max<int>(1, (int)todo.size()))
                                  : 1;
    auto host_pass = [&](size_t q) {
        const int i = todo[q];
        PairState& s = st[i];
        PairResult& r = out[i];
        Host& h = hs[q];
        std::vector<double> gst;
        result_grad(r.matches, nb, rp.frame_duration,
                    [&](const double* f0, const double* f1, const double* f2, const double* f3, double* g0,
                        double* g1, double* g2, double* g3) {
                        mapper_grad(nb, f0, f1, f2, f3, g0, g1, g2, g3, inner_threads);
                    },
                    gst);
        std::vector<int> row_of(s.np, -1);
        for (size_t k = 0; k < s.rows.size(); k++)
            if (r.realign_accepted[s.rows[k].match]) row_of[s.rows[k].match] = (int)k;
        std::vector<double> g_prep_ref((size_t)nb * s.nf_ref, 0.0), g_prep_deg((size_t)nb * s.nf_deg, 0.0);
        std::vector<double> gm(nb), gs(nb), ge(nb);
        std::vector<double> ref_patch((size_t)nb * S), deg_patch((size_t)nb * S), gr((size_t)nb * S), gd((size_t)nb * S);
        std::vector<std::pair<int, std::vector<double>>> reused;  // (row, raw-level slopes) onto full-signal frames
        h.g_row_deg.assign(s.rows.size(), {});
        for (int p = 0; p < s.np; p++) {
            const double* g = &gst[(size_t)p * nb * 3];
            for (int b = 0; b < nb; b++) {
                gm[b] = g[b * 3];
                gs[b] = g[b * 3 + 1];
                ge[b] = g[b * 3 + 2];
            }
            if (row_of[p] >= 0) {
                // the accepted realignment row's statistics replaced the match's
                const int k = row_of[p];
                const RowTrace& tr = s.rows[k];
                const RowHost& rh = h.rows[k];
                const size_t cells = (size_t)nb * tr.nf;
                std::vector<double> g_pr(cells, 0.0), g_pd(cells, 0.0), g_rr(cells), g_rd(cells);
                patch_similarity_grad(rh.prep_r.data(), rh.prep_d.data(), nb, tr.nf, gm.data(), gs.data(), ge.data(),
                                      g_pr.data(), g_pd.data());
                prepare_grad(rh.raw_r.data(), rh.raw_d.data(), nb, tr.nf, tr.nf, g_pr.data(), g_pd.data(), g_rr.data(),
                             g_rd.data());
                if (tr.deg_reused) reused.emplace_back(k, std::move(g_rd));
                else h.g_row_deg[k] = std::move(g_rd);
                continue;
            }
            const int state = s.match_state[p];
            if (state == 0) continue;  // not reached by the backtrack: constant defaults
            if (state == 2) std::fill(gm.begin(), gm.end(), 0.0);  // no real match: means pinned to 0
            const int idx = s.idx[p], off = r.dp_offset[p];
            for (int b = 0; b < nb; b++)
                for (int c = 0; c < S; c++) {
                    ref_patch[(size_t)b * S + c] = h.prep_ref[(size_t)b * s.nf_ref + idx + c];
                    deg_patch[(size_t)b * S + c] = off + c < s.nf_deg ? h.prep_deg[(size_t)b * s.nf_deg + off + c] : 0.0;
                }
            std::fill(gr.begin(), gr.end(), 0.0);
            std::fill(gd.begin(), gd.end(), 0.0);
            patch_similarity_grad(ref_patch.data(), deg_patch.data(), nb, S, gm.data(), gs.data(), ge.data(), gr.data(),
                                  gd.data());
            for (int b = 0; b < nb; b++)
                for (int c = 0; c < S; c++) {
                    g_prep_ref[(size_t)b * s.nf_ref + idx + c] += gr[(size_t)b * S + c];
                    if (off + c < s.nf_deg) g_prep_deg[(size_t)b * s.nf_deg + off + c] += gd[(size_t)b * S + c];
                }
        }
        std::vector<double> g_raw_ref((size_t)nb * s.nf_ref);
        h.g_raw_deg.assign((size_t)nb * s.nf_deg, 0.0);
        prepare_grad(h.raw_ref.data(), h.raw_deg.data(), nb, s.nf_ref, s.nf_deg, g_prep_ref.data(), g_prep_deg.data(),
                     g_raw_ref.data(), h.g_raw_deg.data());
        for (auto& rg : reused) {
            const RowTrace& tr = s.rows[rg.first];
            for (int b = 0; b < nb; b++)
                for (int f = 0; f < tr.nf; f++)
                    h.g_raw_deg[(size_t)b * s.nf_deg + tr.g0 + f] += rg.second[(size_t)b * tr.nf + f];
        }
    };
    if (nthreads > 1) {
        run_striped(nthreads, [&](int t) {
            for (size_t q = t; q < todo.size(); q += nthreads) host_pass(q);
        });
    } else {
        for (size_t q = 0; q < todo.size(); q++) host_pass(q);
    }

    // ---------------------------------------------------------------- device: frames → aligned signals
    std::vector<double> gcols;
    std::vector<FrameGradJob> jobs;
    std::vector<size_t> gcol_off;  // per job: offset of its column in gcols
    std::vector<size_t> first_job(todo.size() + 1, 0);
    for (size_t q = 0; q < todo.size(); q++) {
        const int i = todo[q];
        const PairState& s = st[i];
        Host& h = hs[q];
        first_job[q] = jobs.size();
        const size_t base = gcols.size();
        gcols.insert(gcols.end(), h.g_raw_deg.begin(), h.g_raw_deg.end());
        for (int f = 0; f < s.nf_deg; f++) {
            if (!column_nonzero(&h.g_raw_deg[f], nb, s.nf_deg)) continue;
            jobs.push_back(FrameGradJob{s.degs, (long long)f * hop, nullptr, nullptr, s.nf_deg, 0, W, 0});
            gcol_off.push_back(base + f);
        }
        for (size_t k = 0; k < s.rows.size(); k++) {
            const std::vector<double>& g = h.g_row_deg[k];
            if (g.empty()) continue;
            const RowTrace& tr = s.rows[k];
            const size_t rbase = gcols.size();
            gcols.insert(gcols.end(), g.begin(), g.end());
            for (int f = 0; f < tr.nf; f++) {
                if (!column_nonzero(&g[f], nb, tr.nf)) continue;
                const long long start = tr.first + (long long)f * hop;
                const long long lo = std::min<long long>(std::max(0LL, tr.valid_lo - start), W);
                const long long hi = std::min<long long>(std::max(0LL, tr.valid_hi - start), W);
                if (lo >= hi) continue;
                if (start + lo < 0 || start + hi > s.dl2)
                    throw BackendError{-7, "gradient: realignment frame outside the aligned signal"};
                jobs.push_back(FrameGradJob{s.degs, start, nullptr, nullptr, tr.nf, (int)lo, (int)hi, 0});
                gcol_off.push_back(rbase + f);
            }
        }
    }
    first_job[todo.size()] = jobs.size();
    if (!jobs.empty()) {
        const double* d_gcols = upload(gcols);
        double* d_fout = (double*)arena_.alloc(sizeof(double) * W * jobs.size());
        for (size_t j = 0; j < jobs.size(); j++) {
            jobs[j].g_col = d_gcols + gcol_off[j];
            jobs[j].out = d_fout + j * W;
        }
        FrameGradJob* d_jobs = upload(jobs);
        const int batch = std::min<int>(kAdjBatch, (int)jobs.size());
        double* d_ck = (double*)arena_.alloc(sizeof(double) * gammatone_adjoint_scratch(batch, W));
        for (size_t j0 = 0; j0 < jobs.size(); j0 += batch)
            launch_gammatone_adjoint(d_jobs + j0, (int)std::min<size_t>(batch, jobs.size() - j0), dr.coef, dr.hann, nb,
                                     W, d_ck, stream_);

        // overlap sum: each block of kAccLen samples lists the frames over it, in job order
        std::vector<AccRun> runs;
        std::vector<int> list;
        std::vector<size_t> gal_off(todo.size());
        size_t gal_total = 0;
        for (size_t q = 0; q < todo.size(); q++) {
            gal_off[q] = gal_total;
            gal_total += (size_t)st[todo[q]].dl2;
        }
        double* d_gal = (double*)arena_.alloc(sizeof(double) * std::max<size_t>(1, gal_total));
        for (size_t q = 0; q < todo.size(); q++) {
            const long long len = st[todo[q]].dl2;
            const long long nruns = (len + kAccLen - 1) / kAccLen;
            std::vector<std::vector<int>> per(nruns);
            for (size_t j = first_job[q]; j < first_job[q + 1]; j++) {
                const long long a = jobs[j].start + jobs[j].lo, b = jobs[j].start + jobs[j].hi;  // within [0, len)
                for (long long rr = a / kAccLen; rr <= (b - 1) / kAccLen; rr++) per[rr].push_back((int)j);
            }
            for (long long rr = 0; rr < nruns; rr++) {
                runs.push_back(AccRun{d_gal + gal_off[q] + rr * kAccLen, rr * kAccLen, (int)list.size(),
                                      (int)per[rr].size(), (int)std::min<long long>(kAccLen, len - rr * kAccLen), 0});
                list.insert(list.end(), per[rr].begin(), per[rr].end());
            }
        }
        launch_frame_accumulate(upload(runs), (int)runs.size(), upload(list), d_jobs, stream_);
        double* p = (double*)pinned_.alloc(sizeof(double) * std::max<size_t>(1, gal_total));
        readback(p, d_gal, sizeof(double) * gal_total);
        sync();
        for (size_t q = 0; q < todo.size(); q++) hs[q].g_al = p + gal_off[q];
    }

    // ---------------------------------------------------------------- host: aligned → input
    // (a pair whose frames all have zero slope keeps the zeros written above)
    auto input_pass = [&](size_t q) {
        if (!hs[q].g_al) return;
        const int i = todo[q];
        const PairState& s = st[i];
        const double* deg = pairs[i].on_device ? hs[q].deg.data() : pairs[i].deg;
        const bool scaled = isfinite(spl_of_energy(s.e_ref, s.rl)) && isfinite(spl_of_energy(s.e_deg, s.dl2));
        aligned_to_input_grad(deg, s.dl, s.shift, s.dl2, s.factor, s.e_deg, scaled, hs[q].g_al, grads[i]);
    };
    if (nthreads > 1) {
        run_striped(nthreads, [&](int t) {
            for (size_t q = t; q < todo.size(); q += nthreads) input_pass(q);
        });
    } else {
        for (size_t q = 0; q < todo.size(); q++) input_pass(q);
    }
}

}  // namespace vq
