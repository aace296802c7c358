// `visqol`: the command-line tool over libvisqol.so's C ABI (main.zig:35-257,
// flags after google/visqol's CLI).  One pair (--reference_file,
// --degraded_file) prints its MOS-LQO, with --verbose the per-band features
// and the patches; a batch CSV (--batch_input_csv) decodes every WAV on host
// threads, requires one sample rate, scores all pairs in one
// visqol_measure_batch call, prints a line per pair and optionally writes
// reference,degraded,moslqo rows (--results_csv).  Errors print
// "error: ..." and exit 1 (usage errors on stdout, failed calls on stderr);
// an unknown flag prints the usage and exits 2.
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <atomic>
#include <fstream>
#include <sstream>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "../../include/visqol.h"
#include "../../include/visqol_hip.h"
#include "cli_args.h"

using namespace vq_cli;

namespace {

// A usage error, on stdout (main.zig:259-263).
[[noreturn]] void fail(const std::string& msg) {
    printf("error: %s\n", msg.c_str());
    fflush(stdout);
    exit(1);
}

// A failed call (a file, the model, the device), on stderr as an error
// leaving main does.
[[noreturn]] void fail_call(const std::string& msg) {
    fflush(stdout);
    fprintf(stderr, "error: %s\n", msg.c_str());
    exit(1);
}

bool read_file(const std::string& path, std::string& out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::ostringstream ss;
    ss << f.rdbuf();
    out = ss.str();
    return true;
}

struct Wav {
    std::vector<double> x;
    uint32_t sr = 0;
    int rc = VISQOL_OK;
    std::string err;
};

Wav load(const std::string& path) {
    Wav w;
    std::string bytes;
    if (!read_file(path, bytes)) {
        w.rc = VISQOL_E_IO;
        w.err = "FileNotFound";
        return w;
    }
    size_t n = 0;
    w.rc = visqol_hip_decode_wav((const uint8_t*)bytes.data(), bytes.size(), nullptr, 0, &n, &w.sr);
    if (w.rc == VISQOL_OK) {
        w.x.resize(n);
        w.rc = visqol_hip_decode_wav((const uint8_t*)bytes.data(), bytes.size(), w.x.data(), n, &n, &w.sr);
    }
    if (w.rc != VISQOL_OK) w.err = visqol_last_error();
    return w;
}

void print_result(const std::string& rp, const std::string& dp, const visqol_result* r, bool verbose) {
    printf("Reference Filepath:\t%s\nDegraded Filepath:\t%s\nMOS-LQO:\t\t%.6f\n", rp.c_str(), dp.c_str(), r->moslqo);
    if (!verbose) return;
    printf("VNSIM:\t\t\t%.6f\nAlignment lag (s):\t%.6f\n", r->vnsim, r->alignment_lag_s);
    printf("\n  band    freq (Hz)    fvnsim  fvnsim10  fstdnsim  fvdegenergy\n");
    for (size_t b = 0; b < r->num_bands; b++)
        printf("  %4zu %12.2f %9.6f %9.6f %9.6f %12.6f\n", b, r->center_freq_bands[b], r->fvnsim[b], r->fvnsim10[b],
               r->fstdnsim[b], r->fvdegenergy[b]);
    printf("\n  patch  ref_start  ref_end  deg_start  deg_end  similarity\n");
    for (size_t k = 0; k < r->num_patches; k++) {
        const visqol_patch& p = r->patch_sims[k];
        printf("  %5zu %10.4f %8.4f %10.4f %8.4f %11.6f\n", k, p.ref_patch_start_time, p.ref_patch_end_time,
               p.deg_patch_start_time, p.deg_patch_end_time, p.similarity);
    }
}

int run_batch(visqol_ctx* ctx, const Args& a) {
    std::string text;
    if (!read_file(a.batch_csv, text)) fail_call("FileNotFound");
    const auto names = parse_csv(text);
    // decode every file on host threads (threads 0: all cores)
    std::vector<Wav> refs(names.size()), degs(names.size());
    std::atomic<size_t> next{0};
    auto worker = [&] {
        for (size_t i; (i = next.fetch_add(1)) < names.size();) {
            refs[i] = load(names[i].first);
            if (refs[i].rc == VISQOL_OK) degs[i] = load(names[i].second);
        }
    };
    const size_t want = a.threads ? a.threads : std::max(1u, std::thread::hardware_concurrency());
    std::vector<std::thread> pool;
    pool.reserve(std::min(want, names.size()));
    for (size_t t = 1; t < std::min(want, names.size()); t++) {
        try {
            pool.emplace_back(worker);
        } catch (const std::system_error&) {
            break;  // the queue is shared: the threads there are (and this one) take the rest
        }
    }
    worker();
    for (auto& t : pool) t.join();
    // the first failure in CSV order ends the run
    for (size_t i = 0; i < names.size(); i++) {
        if (refs[i].rc != VISQOL_OK) fail_call(refs[i].err);
        if (degs[i].rc != VISQOL_OK) fail_call(degs[i].err);
    }
    uint32_t sr = 0;
    for (size_t i = 0; i < names.size(); i++) {
        if (sr == 0) sr = refs[i].sr;
        if (refs[i].sr != sr || degs[i].sr != sr) fail("batch mode requires every file to share one sample rate");
    }
    const size_t n = names.size();
    std::vector<const double*> rp(n), dp(n);
    std::vector<size_t> rl(n), dl(n);
    for (size_t i = 0; i < n; i++) {
        rp[i] = refs[i].x.data();
        dp[i] = degs[i].x.data();
        rl[i] = refs[i].x.size();
        dl[i] = degs[i].x.size();
    }
    std::vector<visqol_result*> out(n, nullptr);
    std::vector<int32_t> st(n, 0);
    if (n) {
        const int rc = visqol_measure_batch(ctx, n, rp.data(), rl.data(), dp.data(), dl.data(), sr, a.threads, out.data(),
                                            st.data());
        // a code that poisons the whole batch (every status equal and no result) is a device failure
        bool any = false;
        for (auto* r : out) any = any || r;
        if (rc != VISQOL_OK && !any) fail_call(visqol_last_error());
    }
    std::string csv = "reference,degraded,moslqo\n";
    for (size_t i = 0; i < n; i++) {
        const auto& nm = names[i];
        if (out[i]) {
            printf("%s %s MOS-LQO: %.6f\n", nm.first.c_str(), nm.second.c_str(), out[i]->moslqo);
            csv += nm.first + "," + nm.second + "," + shortest_decimal(out[i]->moslqo) + "\n";
            visqol_result_free(out[i]);
        } else {
            // the pair's own error name: score it alone once more
            visqol_result* r = nullptr;
            const int rc = visqol_measure(ctx, rp[i], rl[i], dp[i], dl[i], sr, &r);
            if (r) visqol_result_free(r);
            printf("%s %s error: %s\n", nm.first.c_str(), nm.second.c_str(),
                   rc != VISQOL_OK ? visqol_last_error() : "unknown");
        }
    }
    if (!a.results_csv.empty()) {
        std::ofstream f(a.results_csv, std::ios::binary);
        if (!f || !(f << csv)) fail("cannot write " + a.results_csv);
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    Args a;
    std::string err;
    const int pr = parse_args(argc, argv, a, err);
    if (pr == 2) {
        printf("unknown flag: %s\n\n%s", err.c_str(), usage());
        return 2;
    }
    if (pr == 3) fail(err);  // main.zig:89-93: usage errors on stdout
    if (pr != 0) fail_call(err);
    if (a.help) {
        printf("%s", usage());
        return 0;
    }
    visqol_options o;
    visqol_options_default(&o);
    o.mode = a.speech ? VISQOL_MODE_SPEECH : VISQOL_MODE_AUDIO;
    o.search_window_radius = a.radius;
    o.use_unscaled_speech_mos_mapping = a.unscaled;
    o.disable_global_alignment = a.no_global;
    o.disable_realignment = a.no_realign;
    o.use_lattice_model = a.lattice;
    o.model_path = a.model.empty() ? nullptr : a.model.c_str();
    o.backend = a.backend;
    o.precision = a.precision;
    visqol_ctx* ctx = nullptr;
    if (visqol_create(&o, &ctx) != VISQOL_OK) fail_call(visqol_last_error());
    int rc = 0;
    if (!a.batch_csv.empty()) {
        rc = run_batch(ctx, a);
    } else {
        if (a.ref.empty()) fail("--reference_file is required");
        if (a.deg.empty()) fail("--degraded_file is required");
        const Wav r = load(a.ref), d = load(a.deg);
        if (r.rc != VISQOL_OK) fail_call(r.err);
        if (d.rc != VISQOL_OK) fail_call(d.err);
        if (r.sr != d.sr) fail("reference and degraded sample rates differ");
        visqol_result* res = nullptr;
        if (visqol_measure(ctx, r.x.data(), r.x.size(), d.x.data(), d.x.size(), r.sr, &res) != VISQOL_OK)
            fail_call(visqol_last_error());
        print_result(a.ref, a.deg, res, a.verbose);
        visqol_result_free(res);
    }
    visqol_destroy(ctx);
    fflush(stdout);
    return rc;
}
