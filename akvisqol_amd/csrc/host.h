// Host-side planning and mapping for the gfx950 ViSQOL library: everything
// the device pipeline needs decided on the CPU (filter design, frame and
// patch geometry, VAD chunk decisions, realignment slices) and the MOS
// mapping models.  All f64 arithmetic here is plain IEEE double with no
// contraction; transcendental calls use the platform libm once per rate.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <exception>
#include <memory>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "dev_common.h"

namespace vq {

constexpr int kBandsAudio = 32;
constexpr int kBandsSpeech = 21;
constexpr int kPatchAudio = 30;
constexpr int kPatchSpeech = 20;

// Filterbank and framing of one (mode, sample rate) (visqol.zig:346-359).
struct RateParams {
    uint32_t sample_rate = 0;
    int nb = 0, patch = 0;
    int window = 0, hop = 0;
    double frame_duration = 0;
    std::vector<double> cf;        // centre frequencies, ascending
    std::vector<BandCoef> coef;    // per band, ascending
    std::vector<FirCoef> fir;      // the same bands in FIR + all-pole form
    std::vector<double> hann;      // window
    std::vector<double> raw_band;  // 8 per band: a0 a11 a12 a13 a14 b1 b2 gain
};
RateParams make_rate_params(int mode, uint32_t sample_rate);

// Twiddles W_n^k = (cos(-2πk/n), sin(-2πk/n)) for k < n/2, interleaved.
std::vector<double> make_twiddles(long long n);
// W_S^j for every span S = 2..n at complex index S/2 + j (2n doubles); the
// table for n begins with the table of every smaller power of two.
std::vector<double> make_level_twiddles(long long n, int threads = 1);

long long fft_size_for(long long biggest);  // xcorr.zig:64-69
int num_frames(long long len, int window, int hop);

// audio.zig:132-159 given Σx² and length; -inf for an empty signal.
double spl_of_energy(double energy, long long len);
double spl_factor(double ref_spl, double deg_spl);

// dp.zig:47-70; returns false when the reference is too short.
bool ref_patch_indices(int nf, int patch, std::vector<int>& out);
// vad.zig: the chunk count the VAD scans, and the patch choice from the
// per-chunk activity computed on the device.
struct VadPlan {
    int frame_len = 0, first = 0, patch_count = 0, n_chunks = 0;
};
VadPlan vad_plan(int nf, int patch, int window);
void vad_patches(const VadPlan& vp, const uint8_t* active, int patch, std::vector<int>& out);

// dp.zig:73-201: patch count after tail trimming and the search windows.
// Returns false for SignalsTooDifferent.
bool plan_search(const std::vector<int>& idx, int nf_ref, int nf_deg, int patch, int radius, int& np,
                 std::vector<int>& lo, std::vector<int>& hi);

// realign.zig:21-51
struct Layout {
    long long pad = 0, base = 0, len = 0;
};
Layout slice_layout(long long signal_len, uint32_t sr, double start, double end);
// realign.zig:221-250
struct Geometry {
    long long len = 0, ref_off = 0, deg_off = 0;
    double lag_seconds = 0;
};
Geometry trim_geometry(long long ref_len, long long deg_len, long long lag, uint32_t sr);

// ---------------------------------------------------------------- mapping
class SvrModel {
   public:
    // svr.zig:80-165; returns an error message on failure.
    static std::unique_ptr<SvrModel> parse(const std::string& text, std::string& err);
    double predict(const double* x) const;  // svr.zig:37-61 (f64 kernel)
    // ∂predict/∂x (visqol.zig:106-116), num_features() values
    void gradient(const double* x, double* g) const;
    int num_features() const { return nf_; }

   private:
    double gamma_ = 0.01, rho_ = 0.0;
    int nf_ = 0;
    std::vector<double> sv_, coef_;
};

class LatticeModel;  // models.cpp
struct LatticeDeleter {
    void operator()(LatticeModel*) const;
};
std::unique_ptr<LatticeModel, LatticeDeleter> load_lattice(const std::string& bytes, int nb, std::string& err);
// lattice.zig:49-71: the f32 graph output widened to f64; false on a
// graph error.
bool lattice_predict(const LatticeModel& m, const double* fvnsim, const double* fvnsim10, const double* fstdnsim,
                     const double* fvdegenergy, double& out);

// Runs f(t) for t = 0 .. n - 1: t >= 1 on threads of their own, t = 0 on
// the caller's, and any t whose thread cannot be created on the caller's
// too (after the threads already started), so a host short of threads
// runs slower instead of terminating.  The first exception of a stripe run
// on the caller's thread is rethrown once every thread has been joined.
template <typename F>
void run_striped(int n, F&& f) {
    std::vector<std::thread> pool;
    pool.reserve(n > 1 ? (size_t)(n - 1) : 0);
    std::exception_ptr err;
    auto here = [&](int t) {
        try {
            f(t);
        } catch (...) {
            if (!err) err = std::current_exception();
        }
    };
    for (int t = 1; t < n; t++) {
        try {
            pool.emplace_back(f, t);
        } catch (const std::system_error&) {
            here(t);
        }
    }
    here(0);
    for (auto& th : pool) th.join();
    if (err) std::rethrow_exception(err);
}

}  // namespace vq
