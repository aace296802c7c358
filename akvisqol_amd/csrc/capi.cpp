// C ABI of the gfx950 ViSQOL library (include/visqol.h, include/visqol_hip.h).
// Mirrors the reference's src/capi.zig: contexts, thread-local last error,
// heap results owning every pointer they expose, per-pair batch status.
#include <dlfcn.h>
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <system_error>
#include <thread>

#include "../../include/visqol.h"
#include "../../include/visqol_hip.h"
#include "capi_internal.h"
#include "engine.h"
#include "grad.h"
#include "wav.h"

using namespace vq;

struct visqol_ctx {
    EngineOptions opts;
    std::unique_ptr<Engine> engine;
};

struct visqol_hip_prepared {
    std::unique_ptr<PreparedSet> set;
};

namespace {
thread_local char g_error[512] = {0};
}  // namespace

namespace vq_capi {

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof g_error, fmt, ap);
    va_end(ap);
}

int invalid(const char* msg) {
    set_error("%s", msg);
    return VISQOL_E_INVALID_ARGUMENT;
}

int exception_code() {
    try {
        throw;
    } catch (const BackendError& e) {
        set_error("%s", e.msg.c_str());
        return e.code;
    } catch (const std::bad_alloc&) {
        set_error("OutOfMemory");
        return VISQOL_E_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        // std::system_error from thread creation, length errors, ...
        set_error("%s", e.what());
        return VISQOL_E_BACKEND;
    } catch (...) {
        set_error("unknown exception");
        return VISQOL_E_INTERNAL;
    }
}

// One allocation holding the visqol_result, its arrays and its patches
// (capi.zig:65-70, 236-278).
visqol_result* box(const PairResult& r) {
    const size_t nb = r.fvnsim.size(), np = r.matches.size();
    const size_t bytes = sizeof(visqol_result) + sizeof(double) * 5 * nb + sizeof(visqol_patch) * np +
                         sizeof(double) * 3 * nb * np;
    char* blk = (char*)calloc(1, bytes);
    if (!blk) return nullptr;
    visqol_result* out = (visqol_result*)blk;
    double* arr = (double*)(blk + sizeof(visqol_result));
    visqol_patch* patches = (visqol_patch*)(arr + 5 * nb);
    double* band = (double*)(patches + np);
    memcpy(arr, r.fvnsim.data(), sizeof(double) * nb);
    memcpy(arr + nb, r.fvnsim10.data(), sizeof(double) * nb);
    memcpy(arr + 2 * nb, r.fstdnsim.data(), sizeof(double) * nb);
    memcpy(arr + 3 * nb, r.fvdegenergy.data(), sizeof(double) * nb);
    memcpy(arr + 4 * nb, r.cf.data(), sizeof(double) * nb);
    for (size_t i = 0; i < np; i++) {
        const Match& m = r.matches[i];
        double* d = band + i * 3 * nb;
        memcpy(d, m.means, sizeof(double) * nb);
        memcpy(d + nb, m.stddevs, sizeof(double) * nb);
        memcpy(d + 2 * nb, m.energy, sizeof(double) * nb);
        patches[i] = visqol_patch{m.similarity, m.ref_start, m.ref_end, m.deg_start, m.deg_end, d, d + nb, d + 2 * nb};
    }
    out->moslqo = r.moslqo;
    out->vnsim = r.vnsim;
    out->alignment_lag_s = r.lag_s;
    out->num_bands = nb;
    out->fvnsim = arr;
    out->fvnsim10 = arr + nb;
    out->fstdnsim = arr + 2 * nb;
    out->fvdegenergy = arr + 3 * nb;
    out->center_freq_bands = arr + 4 * nb;
    out->num_patches = np;
    out->patch_sims = patches;
    return out;
}

bool device_ok(std::string& why) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        (void)hipGetLastError();
        why = "no HIP device";
        return false;
    }
    int dev = 0;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        why = "cannot query the HIP device";
        return false;
    }
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        why = std::string("device is ") + prop.gcnArchName + ", this library is built for gfx950";
        return false;
    }
    return true;
}

}  // namespace vq_capi

namespace {
using namespace vq_capi;

std::string library_dir() {
    Dl_info info;
    if (dladdr((void*)&library_dir, &info) && info.dli_fname) {
        std::string p = info.dli_fname;
        const size_t s = p.rfind('/');
        if (s != std::string::npos) return p.substr(0, s);
    }
    return ".";
}

bool read_file(const std::string& path, std::string& out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::ostringstream ss;
    ss << f.rdbuf();
    out = ss.str();
    return true;
}

int run_pairs(visqol_ctx* c, const PairInput* pairs, size_t n, uint32_t sr, std::vector<PairResult>& res,
              double* const* grads = nullptr) {
    res.assign(n, PairResult());
    if (sr == 0) {
        for (auto& r : res) { r.status = VISQOL_E_INVALID_ARGUMENT; r.error = "InvalidInput"; }
        for (size_t i = 0; grads && i < n; i++)
            if (grads[i]) std::fill(grads[i], grads[i] + pairs[i].deg_len, 0.0);
        return VISQOL_OK;
    }
    try {
        c->engine->run(pairs, (int)n, sr, res.data(), false, grads);
    } catch (...) {
        return exception_code();
    }
    return VISQOL_OK;
}

// Boxes per-pair results; status[i] per pair; the first failure is returned.
int finish_batch(std::vector<PairResult>& res, visqol_result** out, int32_t* status) {
    int first = VISQOL_OK;
    for (size_t i = 0; i < res.size(); i++) {
        out[i] = nullptr;
        int rc = res[i].status;
        if (rc == VISQOL_OK) {
            out[i] = box(res[i]);
            if (!out[i]) rc = VISQOL_E_OUT_OF_MEMORY;
        }
        if (rc != VISQOL_OK) set_error("%s", res[i].error.empty() ? "OutOfMemory" : res[i].error.c_str());
        if (status) status[i] = rc;
        if (rc != VISQOL_OK && first == VISQOL_OK) first = rc;
    }
    return first;
}

std::mutex g_stage_mu;
std::unique_ptr<Engine> g_stage[2][2];  // [mode][fast]

Engine* stage_engine(uint32_t mode, bool fast) {
    auto& e = g_stage[mode ? 1 : 0][fast ? 1 : 0];
    if (!e) {
        EngineOptions o;
        o.mode = mode ? 1 : 0;
        o.fast = fast;
        e.reset(new Engine(o, nullptr, nullptr));
    }
    return e.get();
}

}  // namespace

namespace vq_capi {

// Engine options and models for validated options (visqol.zig:346-359 model
// selection; capi.zig:188-203): the embedded model unless model_path names
// one, read only in audio mode or in speech mode with the lattice.
int build_engine(const visqol_options& o, EngineOptions& eo, std::unique_ptr<Engine>& engine, bool interleave) {
    eo = EngineOptions();
    eo.interleave_chunks = interleave;
    eo.mode = (int)o.mode;
    eo.radius = o.search_window_radius > 0 ? o.search_window_radius : 60;
    eo.unscaled_speech = o.use_unscaled_speech_mos_mapping != 0;
    eo.no_global = o.disable_global_alignment != 0;
    eo.no_realign = o.disable_realignment != 0;
    eo.lattice = o.use_lattice_model != 0;
    eo.fast = o.precision == VISQOL_PRECISION_FAST || o.precision == VISQOL_PRECISION_BALANCED;
    std::unique_ptr<SvrModel> svr;
    std::unique_ptr<LatticeModel, LatticeDeleter> lat;
    if (eo.mode == 0 || eo.lattice) {
        std::string path = o.model_path && o.model_path[0]
                               ? std::string(o.model_path)
                               : library_dir() + "/models/" + (eo.mode == 0 ? "libsvm_nu_svr_model.txt" : "lattice_speech.tflite");
        std::string text, err;
        if (!read_file(path, text)) {
            set_error("cannot read model '%s'", path.c_str());
            return VISQOL_E_IO;
        }
        if (eo.mode == 0) svr = SvrModel::parse(text, err);
        else lat = load_lattice(text, kBandsSpeech, err);
        if (!svr && !lat) {
            set_error("invalid model '%s': %s", path.c_str(), err.c_str());
            return VISQOL_E_MODEL;
        }
    }
    try {
        engine.reset(new Engine(eo, std::move(svr), std::move(lat)));
    } catch (const BackendError& e) {
        set_error("%s", e.msg.c_str());
        return e.code;
    }
    return VISQOL_OK;
}

}  // namespace vq_capi

extern "C" {

uint32_t visqol_abi_version(void) { return VISQOL_ABI_VERSION; }

const char* visqol_version(void) { return "0.1.0"; }

const char* visqol_last_error(void) { return g_error; }

// Both slots run on the device: VISQOL_BACKEND_CPU as the exact f64 path.
uint32_t visqol_available_backends(void) {
    std::string why;
    return device_ok(why) ? (1u << VISQOL_BACKEND_CPU) | (1u << VISQOL_BACKEND_HIP) : 0u;
}

void visqol_options_default(visqol_options* o) {
    if (!o) return;
    o->mode = VISQOL_MODE_AUDIO;
    o->search_window_radius = 60;
    o->use_unscaled_speech_mos_mapping = 0;
    o->disable_global_alignment = 0;
    o->disable_realignment = 0;
    o->use_lattice_model = 1;
    o->model_path = nullptr;
    o->backend = VISQOL_BACKEND_CPU;
    o->precision = VISQOL_PRECISION_AUTO;
}

int visqol_create(const visqol_options* in, visqol_ctx** out) try {
    if (!out) return invalid("out is NULL");
    *out = nullptr;
    visqol_options o;
    visqol_options_default(&o);
    if (in) o = *in;
    if (o.mode > 1) {
        set_error("invalid mode %u", o.mode);
        return VISQOL_E_INVALID_ARGUMENT;
    }
    if (o.backend > VISQOL_BACKEND_HIP) {
        set_error("invalid backend %u", o.backend);
        return VISQOL_E_INVALID_ARGUMENT;
    }
    if (o.precision > VISQOL_PRECISION_AUTO) {
        set_error("invalid precision %u", o.precision);
        return VISQOL_E_INVALID_ARGUMENT;
    }
    if (o.backend != VISQOL_BACKEND_HIP && o.backend != VISQOL_BACKEND_CPU) {
        set_error("backend %u is not provided by this library; it implements VISQOL_BACKEND_HIP (4) on gfx950 "
                  "and serves VISQOL_BACKEND_CPU (0) with its exact f64 device path",
                  o.backend);
        return VISQOL_E_BACKEND;
    }
    // The reference's CPU backend computes in f64 whatever the precision
    // (backend.zig:163 reads it only for GPU prepared references); the exact
    // device path gives that backend's scores bit for bit.
    if (o.backend == VISQOL_BACKEND_CPU) o.precision = VISQOL_PRECISION_EXACT;
    std::string why;
    if (!device_ok(why)) {
        set_error("%s", why.c_str());
        return VISQOL_E_BACKEND;
    }
    std::unique_ptr<visqol_ctx> c(new visqol_ctx());
    const int rc = build_engine(o, c->opts, c->engine);
    if (rc != VISQOL_OK) return rc;
    *out = c.release();
    return VISQOL_OK;
} catch (...) {
    return exception_code();
}

void visqol_destroy(visqol_ctx* c) { delete c; }

void visqol_result_free(visqol_result* r) { free(r); }

int visqol_measure(visqol_ctx* c, const double* ref, size_t ref_len, const double* deg, size_t deg_len,
                   uint32_t sample_rate, visqol_result** out) try {
    if (!c) return invalid("ctx is NULL");
    if (!out) return invalid("out is NULL");
    *out = nullptr;
    if ((!ref && ref_len) || (!deg && deg_len)) return invalid("signal pointer is NULL");
    PairInput p{ref, (long long)ref_len, deg, (long long)deg_len, false};
    std::vector<PairResult> res;
    const int rc = run_pairs(c, &p, 1, sample_rate, res);
    if (rc != VISQOL_OK) return rc;
    return finish_batch(res, out, nullptr);
} catch (...) {
    return exception_code();
}

int visqol_measure_batch(visqol_ctx* c, size_t n, const double* const* refs, const size_t* ref_lens,
                         const double* const* degs, const size_t* deg_lens, uint32_t sample_rate, uint32_t num_threads,
                         visqol_result** out, int32_t* status) try {
    (void)num_threads;  // the device packs the whole batch
    if (!c) return invalid("ctx is NULL");
    if (n == 0) return VISQOL_OK;
    if (!out) return invalid("out is NULL");
    if (!refs) return invalid("refs is NULL");
    if (!ref_lens) return invalid("ref_lens is NULL");
    if (!degs) return invalid("degs is NULL");
    if (!deg_lens) return invalid("deg_lens is NULL");
    std::vector<PairInput> pairs(n);
    for (size_t i = 0; i < n; i++) pairs[i] = PairInput{refs[i], (long long)ref_lens[i], degs[i], (long long)deg_lens[i], false};
    std::vector<PairResult> res;
    const int rc = run_pairs(c, pairs.data(), n, sample_rate, res);
    if (rc != VISQOL_OK) {
        for (size_t i = 0; i < n; i++) {
            out[i] = nullptr;
            if (status) status[i] = rc;
        }
        return rc;
    }
    return finish_batch(res, out, status);
} catch (...) {
    return exception_code();
}

int visqol_hip_measure_batch_grad(visqol_ctx* c, size_t n, const double* const* refs, const size_t* ref_lens,
                                  const double* const* degs, const size_t* deg_lens, uint32_t sample_rate,
                                  visqol_result** out, int32_t* status, double* const* grads) try {
    if (!c) return invalid("ctx is NULL");
    if (n == 0) return VISQOL_OK;
    if (!out || !refs || !ref_lens || !degs || !deg_lens || !grads) return invalid("NULL argument");
    std::vector<PairInput> pairs(n);
    for (size_t i = 0; i < n; i++) {
        if ((!refs[i] && ref_lens[i]) || (!degs[i] && deg_lens[i])) return invalid("NULL signal with nonzero length");
        pairs[i] = PairInput{refs[i], (long long)ref_lens[i], degs[i], (long long)deg_lens[i], false};
    }
    std::vector<PairResult> res;
    const int rc = run_pairs(c, pairs.data(), n, sample_rate, res, grads);
    if (rc != VISQOL_OK) {
        for (size_t i = 0; i < n; i++) {
            out[i] = nullptr;
            if (status) status[i] = rc;
        }
        return rc;
    }
    return finish_batch(res, out, status);
} catch (...) {
    return exception_code();
}

int visqol_measure_files(visqol_ctx* c, const char* ref_path, const char* deg_path, visqol_result** out) try {
    if (!c) return invalid("ctx is NULL");
    if (!out) return invalid("out is NULL");
    *out = nullptr;
    if (!ref_path) return invalid("ref_path is NULL");
    if (!deg_path) return invalid("deg_path is NULL");
    WavAudio ra, da;
    std::string err;
    if (!load_wav_mono(ref_path, ra, err)) { set_error("%s: %s", ref_path, err.c_str()); return wav_error_code(err); }
    if (!load_wav_mono(deg_path, da, err)) { set_error("%s: %s", deg_path, err.c_str()); return wav_error_code(err); }
    if (ra.sample_rate != da.sample_rate) {
        set_error("Input audio signals have different sample rates! Reference audio sample rate: %u. Degraded audio sample rate: %u",
                  ra.sample_rate, da.sample_rate);
        return VISQOL_E_SIGNAL;
    }
    return visqol_measure(c, ra.samples.data(), ra.samples.size(), da.samples.data(), da.samples.size(), ra.sample_rate, out);
} catch (...) {
    return exception_code();
}

int visqol_hip_decode_wav(const uint8_t* data, size_t len, double* samples, size_t cap, size_t* n_samples,
                          uint32_t* sample_rate) try {
    if (!data && len) return invalid("data is NULL");
    if (!n_samples) return invalid("n_samples is NULL");
    *n_samples = 0;
    WavAudio a;
    std::string err;
    if (!decode_wav_mono(data, len, a, err)) {
        set_error("%s", err.c_str());
        return wav_error_code(err);
    }
    *n_samples = a.samples.size();
    if (sample_rate) *sample_rate = a.sample_rate;
    if (samples) {
        if (cap < a.samples.size()) return invalid("samples holds fewer than the decoded frames");
        memcpy(samples, a.samples.data(), a.samples.size() * sizeof(double));
    }
    return VISQOL_OK;
} catch (...) {
    return exception_code();
}

int visqol_measure_files_batch(visqol_ctx* c, size_t n, const char* const* ref_paths, const char* const* deg_paths,
                               uint32_t num_threads, visqol_result** out, int32_t* status) try {
    (void)num_threads;
    if (!c) return invalid("ctx is NULL");
    if (n == 0) return VISQOL_OK;
    if (!out) return invalid("out is NULL");
    if (!ref_paths) return invalid("ref_paths is NULL");
    if (!deg_paths) return invalid("deg_paths is NULL");
    std::vector<WavAudio> refs(n), degs(n);
    std::vector<int> codes(n, VISQOL_OK);
    for (size_t i = 0; i < n; i++) {
        out[i] = nullptr;
        std::string err;
        if (!load_wav_mono(ref_paths[i], refs[i], err)) {
            set_error("%s: %s", ref_paths[i], err.c_str());
            codes[i] = wav_error_code(err);
            continue;
        }
        if (!load_wav_mono(deg_paths[i], degs[i], err)) {
            set_error("%s: %s", deg_paths[i], err.c_str());
            codes[i] = wav_error_code(err);
            continue;
        }
        if (refs[i].sample_rate != degs[i].sample_rate) {
            set_error("pair %zu: reference and degraded sample rates differ (%u vs %u)", i, refs[i].sample_rate,
                      degs[i].sample_rate);
            codes[i] = VISQOL_E_SIGNAL;
        }
    }
    // One device batch per sample rate (capi.zig:429-448).
    std::vector<bool> done(n);
    for (size_t i = 0; i < n; i++) done[i] = codes[i] != VISQOL_OK;
    for (size_t first = 0; first < n; first++) {
        if (done[first]) continue;
        const uint32_t rate = refs[first].sample_rate;
        std::vector<size_t> idx;
        std::vector<PairInput> pairs;
        for (size_t i = first; i < n; i++) {
            if (done[i] || refs[i].sample_rate != rate) continue;
            idx.push_back(i);
            pairs.push_back(PairInput{refs[i].samples.data(), (long long)refs[i].samples.size(), degs[i].samples.data(),
                                      (long long)degs[i].samples.size(), false});
            done[i] = true;
        }
        std::vector<PairResult> res;
        const int rc = run_pairs(c, pairs.data(), pairs.size(), rate, res);
        if (rc != VISQOL_OK) return rc;
        for (size_t k = 0; k < idx.size(); k++) {
            if (res[k].status == VISQOL_OK) {
                out[idx[k]] = box(res[k]);
                if (!out[idx[k]]) codes[idx[k]] = VISQOL_E_OUT_OF_MEMORY;
            } else {
                set_error("%s", res[k].error.c_str());
                codes[idx[k]] = res[k].status;
            }
        }
    }
    int first_err = VISQOL_OK;
    for (size_t i = 0; i < n; i++) {
        if (status) status[i] = codes[i];
        if (codes[i] != VISQOL_OK && first_err == VISQOL_OK) first_err = codes[i];
    }
    return first_err;
} catch (...) {
    return exception_code();
}

// ---------------------------------------------------------------- extensions
int visqol_hip_measure_trace(visqol_ctx* c, const double* ref, size_t ref_len, const double* deg, size_t deg_len,
                             uint32_t sample_rate, visqol_result** out, visqol_hip_trace* tr, double* ref_spec_out,
                             double* deg_spec_out, double* sims_out) try {
    if (!c) return invalid("ctx is NULL");
    if (!out) return invalid("out is NULL");
    *out = nullptr;
    if ((!ref && ref_len) || (!deg && deg_len)) return invalid("signal pointer is NULL");
    if (sample_rate == 0) return invalid("InvalidInput");
    PairInput p{ref, (long long)ref_len, deg, (long long)deg_len, false};
    std::vector<PairResult> res(1);
    try {
        c->engine->run(&p, 1, sample_rate, res.data(), ref_spec_out || deg_spec_out || sims_out);
    } catch (const BackendError& e) {
        set_error("%s", e.msg.c_str());
        return e.code;
    }
    const PairResult& r = res[0];
    if (tr) {
        memset(tr, 0, sizeof *tr);
        tr->lag = r.lag;
        tr->spl_factor = r.spl_factor;
        tr->nf_ref = r.nf_ref;
        tr->nf_deg = r.nf_deg;
        tr->num_patches = r.np;
        for (int i = 0; i < 512; i++) tr->realign_lag[i] = INT64_MIN;
        for (int i = 0; i < r.np && i < 512; i++) {
            tr->ref_indices[i] = r.ref_idx[i];
            if (i < (int)r.dp_offset.size()) tr->dp_offset[i] = r.dp_offset[i];
            if (i < (int)r.match_valid.size()) tr->match_valid[i] = r.match_valid[i];
            if (i < (int)r.realign_lag.size()) tr->realign_lag[i] = r.realign_lag[i];
            if (i < (int)r.realign_frames.size()) tr->realign_frames[i] = r.realign_frames[i];
            if (i < (int)r.realign_accepted.size()) tr->realign_accepted[i] = r.realign_accepted[i];
        }
    }
    if (ref_spec_out && !r.ref_spec.empty()) memcpy(ref_spec_out, r.ref_spec.data(), sizeof(double) * r.ref_spec.size());
    if (deg_spec_out && !r.deg_spec.empty()) memcpy(deg_spec_out, r.deg_spec.data(), sizeof(double) * r.deg_spec.size());
    if (sims_out && !r.sims.empty()) memcpy(sims_out, r.sims.data(), sizeof(double) * r.sims.size());
    return finish_batch(res, out, nullptr);
} catch (...) {
    return exception_code();
}

int visqol_hip_measure_batch_device(visqol_ctx* c, size_t n, const double* const* dev_refs, const size_t* ref_lens,
                                    const double* const* dev_degs, const size_t* deg_lens, uint32_t sample_rate,
                                    visqol_result** out, int32_t* status) try {
    if (!c) return invalid("ctx is NULL");
    if (n == 0) return VISQOL_OK;
    if (!out || !dev_refs || !ref_lens || !dev_degs || !deg_lens) return invalid("NULL argument");
    std::vector<PairInput> pairs(n);
    for (size_t i = 0; i < n; i++)
        pairs[i] = PairInput{dev_refs[i], (long long)ref_lens[i], dev_degs[i], (long long)deg_lens[i], true};
    std::vector<PairResult> res;
    const int rc = run_pairs(c, pairs.data(), n, sample_rate, res);
    if (rc != VISQOL_OK) {
        for (size_t i = 0; i < n; i++) {
            out[i] = nullptr;
            if (status) status[i] = rc;
        }
        return rc;
    }
    return finish_batch(res, out, status);
} catch (...) {
    return exception_code();
}

int visqol_hip_measure_mixed_batch_device(visqol_ctx* const* ctxs, size_t num_ctxs, size_t n,
                                          const uint32_t* ctx_index, const double* const* dev_refs,
                                          const size_t* ref_lens, const double* const* dev_degs,
                                          const size_t* deg_lens, const uint32_t* sample_rates,
                                          visqol_result** out, int32_t* status) try {
    if (n == 0) return VISQOL_OK;
    if (!ctxs || num_ctxs == 0 || !ctx_index || !out || !dev_refs || !ref_lens || !dev_degs || !deg_lens ||
        !sample_rates)
        return invalid("NULL argument");
    for (size_t k = 0; k < num_ctxs; k++) {
        if (!ctxs[k]) return invalid("ctx is NULL");
        for (size_t j = 0; j < k; j++)
            if (ctxs[j] == ctxs[k]) return invalid("contexts must be distinct");
    }
    for (size_t i = 0; i < n; i++)
        if (ctx_index[i] >= num_ctxs) return invalid("ctx_index out of range");
    // per context, its pairs of each sample rate in pair order
    std::vector<std::map<uint32_t, std::vector<size_t>>> groups(num_ctxs);
    for (size_t i = 0; i < n; i++) groups[ctx_index[i]][sample_rates[i]].push_back(i);
    std::vector<PairResult> res(n);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        (void)hipGetLastError();
        set_error("hipGetDevice failed");
        return VISQOL_E_BACKEND;
    }
    // A batch that fails as a whole marks its pairs with its code and the
    // message its thread recorded.
    auto fail = [&](const std::vector<size_t>& idx, int rc) {
        const std::string msg = visqol_last_error();
        for (size_t i : idx) {
            res[i].status = rc;
            res[i].error = msg;
        }
    };
    auto work = [&](size_t k) {
        (void)hipSetDevice(dev);
        for (const auto& g : groups[k]) try {
            const std::vector<size_t>& idx = g.second;
            std::vector<PairInput> pairs(idx.size());
            for (size_t j = 0; j < idx.size(); j++) {
                const size_t i = idx[j];
                pairs[j] = PairInput{dev_refs[i], (long long)ref_lens[i], dev_degs[i], (long long)deg_lens[i], true};
            }
            std::vector<PairResult> r;
            const int rc = run_pairs(ctxs[k], pairs.data(), pairs.size(), g.first, r);
            if (rc != VISQOL_OK) {
                fail(idx, rc);
                continue;
            }
            for (size_t j = 0; j < idx.size(); j++) res[idx[j]] = std::move(r[j]);
        } catch (...) {
            fail(g.second, exception_code());
        }
    };
    std::vector<std::thread> ts;
    ts.reserve(num_ctxs);  // only thread creation can throw below
    for (size_t k = 1; k < num_ctxs; k++) {
        if (groups[k].empty()) continue;
        try {
            ts.emplace_back(work, k);
        } catch (const std::system_error&) {
            work(k);  // no thread to be had: this context's batches run here
        }
    }
    work(0);
    for (auto& t : ts) t.join();
    return finish_batch(res, out, status);
} catch (...) {
    return exception_code();
}

int visqol_hip_prepare(visqol_ctx* c, size_t n, const double* const* refs, const size_t* ref_lens,
                       uint32_t sample_rate, visqol_hip_prepared** out) try {
    if (!c || !out) return invalid("NULL argument");
    *out = nullptr;
    if (n == 0 || !refs || !ref_lens) return invalid("no references");
    if (sample_rate == 0) return invalid("InvalidInput");
    for (size_t i = 0; i < n; i++)
        if (!refs[i] && ref_lens[i]) return invalid("NULL reference with nonzero length");
    auto* p = new (std::nothrow) visqol_hip_prepared();
    if (!p) {
        set_error("OutOfMemory");
        return VISQOL_E_OUT_OF_MEMORY;
    }
    std::vector<long long> lens(n);
    for (size_t i = 0; i < n; i++) lens[i] = (long long)ref_lens[i];
    try {
        p->set = c->engine->prepare(refs, lens.data(), (int)n, sample_rate);
    } catch (const BackendError& e) {
        delete p;
        set_error("%s", e.msg.c_str());
        return e.code;
    } catch (const std::bad_alloc&) {
        delete p;
        set_error("OutOfMemory");
        return VISQOL_E_OUT_OF_MEMORY;
    }
    *out = p;
    return VISQOL_OK;
} catch (...) {
    return exception_code();
}

int visqol_hip_measure_prepared(visqol_ctx* c, const visqol_hip_prepared* prep, size_t n, const size_t* ref_index,
                                const double* const* degs, const size_t* deg_lens, visqol_result** out,
                                int32_t* status) try {
    return visqol_hip_measure_prepared_grad(c, prep, n, ref_index, degs, deg_lens, out, status, nullptr);
} catch (...) {
    return exception_code();
}

int visqol_hip_measure_prepared_grad(visqol_ctx* c, const visqol_hip_prepared* prep, size_t n,
                                     const size_t* ref_index, const double* const* degs, const size_t* deg_lens,
                                     visqol_result** out, int32_t* status, double* const* grads) try {
    if (!c || !prep || !prep->set) return invalid("NULL argument");
    if (n == 0) return VISQOL_OK;
    if (!out || !degs || !deg_lens) return invalid("NULL argument");
    const PreparedSet& ps = *prep->set;
    if (ps.mode != c->opts.mode || ps.fast != c->opts.fast)
        return invalid("prepared references were made by a context of another mode or precision");
    std::vector<PairInput> pairs(n);
    for (size_t i = 0; i < n; i++) {
        const size_t k = ref_index ? ref_index[i] : i;
        if (k >= ps.refs.size()) return invalid("reference index out of range");
        if (!degs[i] && deg_lens[i]) return invalid("NULL signal with nonzero length");
        pairs[i] = PairInput{nullptr, 0, degs[i], (long long)deg_lens[i], false, &ps.refs[k]};
    }
    std::vector<PairResult> res;
    const int rc = run_pairs(c, pairs.data(), n, ps.sample_rate, res, grads);
    if (rc != VISQOL_OK) {
        for (size_t i = 0; i < n; i++) {
            out[i] = nullptr;
            if (status) status[i] = rc;
        }
        return rc;
    }
    return finish_batch(res, out, status);
} catch (...) {
    return exception_code();
}

void visqol_hip_prepared_free(visqol_hip_prepared* p) { delete p; }

int visqol_hip_profile(visqol_ctx* c, int enable, int reset) try {
    if (!c) return invalid("ctx is NULL");
    c->engine->set_profiling(enable != 0);
    if (reset) c->engine->reset_profile();
    return VISQOL_OK;
} catch (...) {
    return exception_code();
}

int visqol_hip_profile_read(visqol_ctx* c, double* ms, double* flop, double* bytes, int64_t* launches) try {
    if (!c) return invalid("ctx is NULL");
    const Engine::Profile p = c->engine->profile();
    if (ms) *ms = p.gammatone_ms;
    if (flop) *flop = p.gammatone_flop;
    if (bytes) *bytes = p.gammatone_bytes;
    if (launches) *launches = p.launches;
    return VISQOL_OK;
} catch (...) {
    return exception_code();
}

int visqol_hip_profile_sim(visqol_ctx* c, double* ms, double* cells, int64_t* launches) try {
    if (!c) return invalid("ctx is NULL");
    const Engine::Profile p = c->engine->profile();
    if (ms) *ms = p.sim_ms;
    if (cells) *cells = p.sim_cells;
    if (launches) *launches = p.sim_launches;
    return VISQOL_OK;
} catch (...) {
    return exception_code();
}

int visqol_hip_profile_phases(visqol_ctx* c, double* ms, int n) try {
    if (!c) return invalid("ctx is NULL");
    if (n > 0 && !ms) return invalid("ms is NULL");
    const Engine::Profile p = c->engine->profile();
    for (int i = 0; i < n && i < Engine::Profile::kPhases; i++) ms[i] = p.phase_ms[i];
    return Engine::Profile::kPhases;
} catch (...) {
    return exception_code();
}

int visqol_hip_patch_similarity_grad(const double* ref, const double* deg, size_t nrows, size_t ncols,
                                     const double* g_means, const double* g_stddevs, const double* g_energy,
                                     double* g_ref, double* g_deg) try {
    if (!ref || !deg || !g_ref || !g_deg || nrows == 0 || ncols == 0) return invalid("invalid argument");
    memset(g_ref, 0, sizeof(double) * nrows * ncols);
    memset(g_deg, 0, sizeof(double) * nrows * ncols);
    patch_similarity_grad(ref, deg, (int)nrows, (int)ncols, g_means, g_stddevs, g_energy, g_ref, g_deg);
    return VISQOL_OK;
} catch (...) {
    return exception_code();
}

int visqol_hip_prepare_grad(const double* raw_ref, const double* raw_deg, size_t nb, size_t nfr, size_t nfd,
                            const double* g_ref_prep, const double* g_deg_prep, double* g_raw_ref, double* g_raw_deg) try {
    if (!raw_ref || !raw_deg || !g_ref_prep || !g_deg_prep || !g_raw_ref || !g_raw_deg || nb == 0)
        return invalid("invalid argument");
    prepare_grad(raw_ref, raw_deg, (int)nb, (int)nfr, (int)nfd, g_ref_prep, g_deg_prep, g_raw_ref, g_raw_deg);
    return VISQOL_OK;
} catch (...) {
    return exception_code();
}

int visqol_hip_lattice_predict(const uint8_t* model, size_t len, size_t nb, size_t count, const double* fvnsim,
                               const double* fvnsim10, const double* fstdnsim, const double* fvdegenergy, double* y) try {
    if (!model || !y || nb == 0 || nb > kMaxBands || (count && (!fvnsim || !fvnsim10 || !fstdnsim || !fvdegenergy)))
        return invalid("invalid argument");
    std::string err;
    auto m = load_lattice(std::string((const char*)model, len), (int)nb, err);
    if (!m) {
        set_error("lattice model: %s", err.c_str());
        return VISQOL_E_MODEL;
    }
    for (size_t k = 0; k < count; k++)
        if (!lattice_predict(*m, fvnsim + k * nb, fvnsim10 + k * nb, fstdnsim + k * nb, fvdegenergy + k * nb, y[k])) {
            set_error("lattice model evaluation failed");
            return VISQOL_E_MODEL;
        }
    return VISQOL_OK;
} catch (...) {
    return exception_code();
}

int visqol_hip_gammatone_shape(int shape, int* waves, int* bands_per_lane, int* chunk) try {
    if (shape < 0 || shape >= gammatone_shape_count()) return invalid("no such spectrogram block shape");
    const GtShape s = gammatone_shape(shape);
    if (waves) *waves = s.waves;
    if (bands_per_lane) *bands_per_lane = s.bands_per_lane;
    if (chunk) *chunk = s.chunk;
    return gammatone_shape_count();
} catch (...) {
    return exception_code();
}

int visqol_hip_gammatone_tuning(visqol_ctx* c, uint32_t sample_rate, int kind, int* chosen, float* ms, int cap) try {
    if (!c) return invalid("ctx is NULL");
    if (!chosen) return invalid("chosen is NULL");
    if (kind < 0 || kind > 1 || sample_rate == 0) return invalid("invalid argument");
    if (cap > 0 && !ms) return invalid("ms is NULL with cap > 0");
    try {
        const Engine::GtTuning t = c->engine->gammatone_tuning(sample_rate, kind);
        *chosen = t.chosen;
        for (int v = 0; v < cap && v < (int)t.ms.size(); v++) ms[v] = t.ms[v];
    } catch (const BackendError& e) {
        set_error("%s", e.msg.c_str());
        return e.code;
    }
    return gammatone_shape_count();
} catch (...) {
    return exception_code();
}

int visqol_hip_alloc(size_t bytes, void** p) try {
    if (!p) return invalid("NULL argument");
    if (hipMalloc(p, bytes ? bytes : 1) != hipSuccess) {
        (void)hipGetLastError();
        set_error("hipMalloc of %zu bytes failed", bytes);
        return VISQOL_E_OUT_OF_MEMORY;
    }
    return VISQOL_OK;
} catch (...) {
    return exception_code();
}

int visqol_hip_free(void* p) try {
    if (p && hipFree(p) != hipSuccess) {
        set_error("hipFree failed");
        return VISQOL_E_BACKEND;
    }
    return VISQOL_OK;
} catch (...) {
    return exception_code();
}

int visqol_hip_upload(void* dst, const void* src, size_t bytes) try {
    if (!dst || (!src && bytes)) return invalid("NULL argument");
    if (hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) != hipSuccess) {
        set_error("hipMemcpy failed");
        return VISQOL_E_BACKEND;
    }
    return VISQOL_OK;
} catch (...) {
    return exception_code();
}

int visqol_hip_spectrogram(uint32_t mode, uint32_t sample_rate, uint32_t precision, const double* signal, size_t len,
                           double* out, size_t* num_frames_out) try {
    if (!signal || !out || mode > 1 || sample_rate == 0) return invalid("invalid argument");
    std::string why;
    if (!device_ok(why)) { set_error("%s", why.c_str()); return VISQOL_E_BACKEND; }
    std::lock_guard<std::mutex> g(g_stage_mu);
    try {
        std::vector<double> spec;
        int nf = 0;
        stage_engine(mode, precision == VISQOL_PRECISION_FAST || precision == VISQOL_PRECISION_BALANCED)
            ->spectrogram(sample_rate, signal, (long long)len, spec, nf);
        memcpy(out, spec.data(), sizeof(double) * spec.size());
        if (num_frames_out) *num_frames_out = (size_t)nf;
    } catch (const BackendError& e) {
        set_error("%s", e.msg.c_str());
        return e.code;
    }
    return VISQOL_OK;
} catch (...) {
    return exception_code();
}

int visqol_hip_frame_grad(uint32_t mode, uint32_t sample_rate, const double* frame, size_t window, const double* g,
                          size_t lo, size_t hi, double* out) try {
    if (!frame || !g || !out || mode > 1 || sample_rate == 0) return invalid("invalid argument");
    const RateParams p = make_rate_params((int)mode, sample_rate);
    if (window != (size_t)p.window) return invalid("window is not the sample rate's frame length");
    if (lo > hi || hi > window) return invalid("lo, hi must satisfy lo <= hi <= window");
    std::string why;
    if (!device_ok(why)) { set_error("%s", why.c_str()); return VISQOL_E_BACKEND; }
    std::lock_guard<std::mutex> lk(g_stage_mu);
    try {
        stage_engine(mode, false)->frame_grad(sample_rate, frame, g, (int)lo, (int)hi, out);
    } catch (const BackendError& e) {
        set_error("%s", e.msg.c_str());
        return e.code;
    }
    return VISQOL_OK;
} catch (...) {
    return exception_code();
}

int visqol_hip_filters(uint32_t mode, uint32_t sample_rate, double* bands8, double* cf) try {
    if (mode > 1 || sample_rate == 0) return invalid("invalid argument");
    const RateParams p = make_rate_params((int)mode, sample_rate);
    if (bands8) memcpy(bands8, p.raw_band.data(), sizeof(double) * p.raw_band.size());
    if (cf) memcpy(cf, p.cf.data(), sizeof(double) * p.cf.size());
    return VISQOL_OK;
} catch (...) {
    return exception_code();
}

int visqol_hip_best_lag(const double* a, size_t la, const double* b, size_t lb, int64_t* lag) try {
    if ((!a && la) || (!b && lb) || !lag) return invalid("invalid argument");
    std::string why;
    if (!device_ok(why)) { set_error("%s", why.c_str()); return VISQOL_E_BACKEND; }
    std::lock_guard<std::mutex> g(g_stage_mu);
    try {
        *lag = stage_engine(0, false)->best_lag(a, (long long)la, b, (long long)lb);
    } catch (const BackendError& e) {
        set_error("%s", e.msg.c_str());
        return e.code;
    }
    return VISQOL_OK;
} catch (...) {
    return exception_code();
}

int visqol_hip_sum(int kind, const double* x, size_t len, double plain_sum, int64_t lag, size_t half, int form,
                   double* out) try {
    if ((!x && len) || !out || kind < kSum || kind > kSumSqFrom || form < 0 || form > 2)
        return invalid("invalid argument");
    if (kind == kNyquist && len == 0) return invalid("the alternating sum needs at least one sample");
    std::string why;
    if (!device_ok(why)) { set_error("%s", why.c_str()); return VISQOL_E_BACKEND; }
    std::lock_guard<std::mutex> g(g_stage_mu);
    try {
        *out = stage_engine(0, false)->reference_sum(kind, x, (long long)len, plain_sum, (long long)lag,
                                                     (long long)half, form);
    } catch (const BackendError& e) {
        set_error("%s", e.msg.c_str());
        return e.code;
    }
    return VISQOL_OK;
} catch (...) {
    return exception_code();
}

int visqol_hip_best_lag_fast(const double* a, size_t la, const double* b, size_t lb, int cap, int64_t* lag,
                             int* nominated) try {
    if ((!a && la) || (!b && lb) || !lag) return invalid("invalid argument");
    if (cap > kMaxLagCands) return invalid("cap exceeds 64");
    std::string why;
    if (!device_ok(why)) { set_error("%s", why.c_str()); return VISQOL_E_BACKEND; }
    std::lock_guard<std::mutex> g(g_stage_mu);
    Engine* e = stage_engine(0, true);
    const int keep = e->options().lag_cap;
    if (cap >= 0) e->set_lag_cap(cap);
    bool nom = false;
    try {
        *lag = e->best_lag_fast(a, (long long)la, b, (long long)lb, &nom);
    } catch (...) {
        e->set_lag_cap(keep);
        throw;
    }
    e->set_lag_cap(keep);
    if (nominated) *nominated = nom ? 1 : 0;
    return VISQOL_OK;
} catch (...) {
    return exception_code();
}

int visqol_hip_xcorr32(const double* a, size_t la, const double* b, size_t lb, float* out, size_t cap,
                       size_t* n_out) try {
    if ((!a && la) || (!b && lb) || (!out && cap) || !n_out) return invalid("invalid argument");
    std::string why;
    if (!device_ok(why)) { set_error("%s", why.c_str()); return VISQOL_E_BACKEND; }
    std::lock_guard<std::mutex> g(g_stage_mu);
    std::vector<float> v;
    stage_engine(0, true)->xcorr32(a, (long long)la, b, (long long)lb, v);
    *n_out = v.size();
    if (!v.empty()) memcpy(out, v.data(), sizeof(float) * std::min(cap, v.size()));
    return VISQOL_OK;
} catch (...) {
    return exception_code();
}

int visqol_hip_xcorr(const double* a, size_t la, const double* b, size_t lb, double* out, size_t cap, size_t* n_out) try {
    if ((!a && la) || (!b && lb) || (!out && cap) || !n_out) return invalid("invalid argument");
    if (std::max(la, lb) < 2) return invalid("xcorr needs a signal of at least 2 samples");
    std::string why;
    if (!device_ok(why)) { set_error("%s", why.c_str()); return VISQOL_E_BACKEND; }
    std::lock_guard<std::mutex> g(g_stage_mu);
    try {
        std::vector<double> v;
        stage_engine(0, false)->xcorr(a, (long long)la, b, (long long)lb, v);
        *n_out = v.size();
        memcpy(out, v.data(), sizeof(double) * std::min(cap, v.size()));
    } catch (const BackendError& e) {
        set_error("%s", e.msg.c_str());
        return e.code;
    }
    return VISQOL_OK;
} catch (...) {
    return exception_code();
}

int visqol_hip_fft(double* z, size_t n, int dit) try {
    if (!z || n == 0 || (n & (n - 1))) return invalid("n must be a power of two");
    std::string why;
    if (!device_ok(why)) { set_error("%s", why.c_str()); return VISQOL_E_BACKEND; }
    std::lock_guard<std::mutex> g(g_stage_mu);
    try {
        stage_engine(0, false)->fft_probe(z, (long long)n, dit != 0);
    } catch (const BackendError& e) {
        set_error("%s", e.msg.c_str());
        return e.code;
    }
    return VISQOL_OK;
} catch (...) {
    return exception_code();
}

int visqol_hip_math_probe(int op, const double* a, const double* b, double* out, size_t n) try {
    if (!a || !out || op < 0 || op > 2 || (op == 2 && !b)) return invalid("invalid argument");
    std::string why;
    if (!device_ok(why)) { set_error("%s", why.c_str()); return VISQOL_E_BACKEND; }
    std::lock_guard<std::mutex> g(g_stage_mu);
    try {
        stage_engine(0, false)->math_probe(op, a, b, out, (long long)n);
    } catch (const BackendError& e) {
        set_error("%s", e.msg.c_str());
        return e.code;
    }
    return VISQOL_OK;
} catch (...) {
    return exception_code();
}

}  // extern "C"
