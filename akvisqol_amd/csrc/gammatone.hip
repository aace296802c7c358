// Gammatone ERB filterbank spectrogram on gfx950 (gammatone.zig:268-306,
// 432-467; accel.zig:116-124).
//
// Work decomposition: one lane per (frame, group of bands).  A wave holds 64
// consecutive frames of BPL bands, so the bands' coefficients are
// wave-uniform (scalar registers) and the four cascaded DF2T biquads run in
// vector registers, in the reference's exact operation order; each lane
// runs its BPL bands' recurrences side by side.  A block holds WAVES waves
// (WAVES·BPL bands) over the same 64 frames; the frames' Hann-windowed
// samples are staged CHUNK samples at a time in LDS (coalesced rows,
// x·hann computed once per block and shared by its bands), the next chunk's
// loads in flight while the current chunk's recurrences run.  The block
// shape (WAVES, BPL, CHUNK) only maps work to lanes: every shape computes
// each (frame, band) with the same operations in the same order, so all
// shapes give bit-identical spectrograms and the fastest can be chosen by
// timing them.
//
// Exact variant: mul and add rounded separately (-ffp-contract=off), the
// output bit-identical to the CPU recurrence.  Fast variant: the same
// recurrence with fused multiply-adds (17 instead of 30 operations per
// sample and band).
#include <array>
#include <utility>

#include "dev_common.h"
#include "kernels.h"

namespace vq {

constexpr int kFramesPerWave = 64;

// Block shapes the launcher offers, fastest-first guess at index 0.
constexpr GtShape kGtShapes[] = {
    {4, 4, 32}, {8, 4, 32}, {2, 4, 32}, {4, 2, 32}, {8, 2, 32}, {4, 4, 64}, {8, 4, 16},
};
constexpr int kGtShapeCount = (int)(sizeof(kGtShapes) / sizeof(kGtShapes[0]));

// The four cascaded DF2T biquads of one band on one frame, and the sum of
// squares of their output.
struct BandState {
    double s1c0 = 0, s1c1 = 0, s2c0 = 0, s2c1 = 0, s3c0 = 0, s3c1 = 0, s4c0 = 0, s4c1 = 0, sum_sq = 0;
};

template <bool FAST>
__device__ inline void band_step(BandState& b, const BandCoef& k, double s) {
    if (FAST) {
        const double f1 = __builtin_fma(k.s1n0, s, b.s1c0);
        b.s1c0 = __builtin_fma(-k.d1, f1, __builtin_fma(k.s1n1, s, b.s1c1));
        b.s1c1 = -k.d2 * f1;
        const double f2 = __builtin_fma(k.ca0, f1, b.s2c0);
        b.s2c0 = __builtin_fma(-k.d1, f2, __builtin_fma(k.s2n1, f1, b.s2c1));
        b.s2c1 = -k.d2 * f2;
        const double f3 = __builtin_fma(k.ca0, f2, b.s3c0);
        b.s3c0 = __builtin_fma(-k.d1, f3, __builtin_fma(k.s3n1, f2, b.s3c1));
        b.s3c1 = -k.d2 * f3;
        const double f4 = __builtin_fma(k.ca0, f3, b.s4c0);
        b.s4c0 = __builtin_fma(-k.d1, f4, __builtin_fma(k.s4n1, f3, b.s4c1));
        b.s4c1 = -k.d2 * f4;
        b.sum_sq = __builtin_fma(f4, f4, b.sum_sq);
    } else {
        const double f1 = k.s1n0 * s + b.s1c0;
        b.s1c0 = k.s1n1 * s + b.s1c1 - k.d1 * f1;
        b.s1c1 = -(k.d2 * f1);
        const double f2 = k.ca0 * f1 + b.s2c0;
        b.s2c0 = k.s2n1 * f1 + b.s2c1 - k.d1 * f2;
        b.s2c1 = -(k.d2 * f2);
        const double f3 = k.ca0 * f2 + b.s3c0;
        b.s3c0 = k.s3n1 * f2 + b.s3c1 - k.d1 * f3;
        b.s3c1 = -(k.d2 * f3);
        const double f4 = k.ca0 * f3 + b.s4c0;
        b.s4c0 = k.s4n1 * f3 + b.s4c1 - k.d1 * f4;
        b.s4c1 = -(k.d2 * f4);
        b.sum_sq += f4 * f4;
    }
}

template <bool FAST, int WAVES, int BPL, int CHUNK>
__global__ __launch_bounds__(64 * WAVES) void k_gammatone(const GtJob* __restrict__ jobs, const BandCoef* __restrict__ coef,
                                                          const double* __restrict__ hann, int nb, int window, int hop) {
    constexpr int kThreads = 64 * WAVES;
    constexpr int kStage = kFramesPerWave * CHUNK / kThreads;  // tile elements staged per thread
    static_assert(kStage >= 1 && kFramesPerWave * CHUNK % kThreads == 0, "tile must split evenly over the block");
    const GtJob job = jobs[blockIdx.z];
    const int f0 = blockIdx.x * kFramesPerWave;
    if (f0 >= job.nf) return;  // whole block idle: uniform exit
    __shared__ double tile[kFramesPerWave][CHUNK + 1];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int band0 = blockIdx.y * (WAVES * BPL) + wave * BPL;
    const int frame = f0 + lane;
    BandCoef k[BPL];
    BandState st[BPL];
#pragma unroll
    for (int q = 0; q < BPL; q++) {
        if (band0 + q < nb) k[q] = coef[band0 + q];
        else k[q] = BandCoef{0, 0, 0, 0, 0, 0, 0, 0};
    }
    const bool active = band0 < nb && frame < job.nf;

    // Staging: thread t stages sample i0 + (t mod CHUNK) of frames
    // f0 + t / CHUNK + u·(kThreads / CHUNK) (u < kStage), so each wave reads
    // contiguous runs of CHUNK samples and the Hann factor is one value per
    // thread and chunk.  The next chunk's loads are issued (unconditionally,
    // from clamped in-range addresses) before the current chunk's recurrence
    // runs; masking and x·hann happen at the LDS write.
    constexpr int kRowsPerPass = kThreads / CHUNK;
    const auto* sig = as_global(job.sig);
    const int ti = threadIdx.x % CHUNK;
    const int fr0 = threadIdx.x / CHUNK;
    double v[kStage];
    double h = 0.0;
    auto load_chunk = [&](int i0) {
        const int i = i0 + ti;
        const int ic = i < window ? i : 0;
        h = hann[ic];
#pragma unroll
        for (int u = 0; u < kStage; u++) {
            const int fr = fr0 + u * kRowsPerPass;
            const int fc = f0 + fr < job.nf ? f0 + fr : f0;
            v[u] = sig[(size_t)fc * hop + ic];
        }
    };
    load_chunk(0);
    for (int i0 = 0; i0 < window; i0 += CHUNK) {
        const int clen = min(CHUNK, window - i0);
        __syncthreads();
        const bool in = i0 + ti < window;
#pragma unroll
        for (int u = 0; u < kStage; u++) {
            const int fr = fr0 + u * kRowsPerPass;
            tile[fr][ti] = (in && f0 + fr < job.nf) ? v[u] * h : 0.0;
        }
        __syncthreads();
        if (i0 + CHUNK < window) load_chunk(i0 + CHUNK);
        if (active) {
            for (int i = 0; i < clen; i++) {
                const double s = tile[lane][i];
#pragma unroll
                for (int q = 0; q < BPL; q++) band_step<FAST>(st[q], k[q], s);
            }
        }
    }
    if (active) {
        const double inv_n = 1.0 / (double)window;
#pragma unroll
        for (int q = 0; q < BPL; q++)
            if (band0 + q < nb) job.out[(size_t)(band0 + q) * job.nf + frame] = sqrt(st[q].sum_sq * inv_n);
    }
}

namespace {
template <bool FAST, int V>
void launch_shape(const GtJob* d_jobs, int njobs, int max_nf, const BandCoef* d_coef, const double* d_hann, int nb,
                  int window, int hop, hipStream_t stream) {
    constexpr GtShape sh = kGtShapes[V];
    dim3 grid((max_nf + kFramesPerWave - 1) / kFramesPerWave, (nb + sh.waves * sh.bands_per_lane - 1) / (sh.waves * sh.bands_per_lane),
              njobs);
    hipLaunchKernelGGL((k_gammatone<FAST, sh.waves, sh.bands_per_lane, sh.chunk>), grid, dim3(64 * sh.waves), 0, stream,
                       d_jobs, d_coef, d_hann, nb, window, hop);
}

using GtLaunch = void (*)(const GtJob*, int, int, const BandCoef*, const double*, int, int, int, hipStream_t);

template <bool FAST, int... V>
constexpr std::array<GtLaunch, sizeof...(V)> shape_table(std::integer_sequence<int, V...>) {
    return {&launch_shape<FAST, V>...};
}
constexpr auto kFastLaunch = shape_table<true>(std::make_integer_sequence<int, kGtShapeCount>{});
constexpr auto kExactLaunch = shape_table<false>(std::make_integer_sequence<int, kGtShapeCount>{});
}  // namespace

int gammatone_shape_count() { return kGtShapeCount; }

GtShape gammatone_shape(int v) { return kGtShapes[v >= 0 && v < kGtShapeCount ? v : 0]; }

void launch_gammatone(const GtJob* d_jobs, int njobs, int max_nf, const BandCoef* d_coef, const double* d_hann, int nb,
                      int window, int hop, bool fast, int shape, hipStream_t stream) {
    if (njobs == 0 || max_nf == 0) return;
    const int v = shape >= 0 && shape < kGtShapeCount ? shape : 0;
    (fast ? kFastLaunch : kExactLaunch)[v](d_jobs, njobs, max_nf, d_coef, d_hann, nb, window, hop, stream);
}

}  // namespace vq
