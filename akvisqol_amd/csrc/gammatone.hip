// Gammatone ERB filterbank spectrogram on gfx950 (gammatone.zig:268-306,
// 432-467; accel.zig:116-124).
//
// Work decomposition: one lane per (frame, group of bands).  A wave holds 64
// consecutive frames of up to BPL bands (slot q of wave w in block row y is
// band y·WAVES·BPL + q·WAVES + w), so the bands' coefficients are
// wave-uniform (scalar registers) and the recurrences run in vector
// registers, each lane running its bands side by side.  A block holds WAVES
// waves (WAVES·BPL bands) over the same 64 frames; the frames'
// Hann-windowed samples are staged CHUNK samples at a time in LDS (coalesced
// rows, x·hann computed once per block and shared by its bands), the next
// chunk's loads in flight while the current chunk's recurrences run.  The
// block shape (WAVES, BPL, CHUNK) only maps work to lanes: every shape
// computes each (frame, band) with the same operations in the same order,
// so all shapes give bit-identical spectrograms and the fastest can be
// chosen by timing them.  A wave whose last slots fall past the band count
// runs a variant without them.
//
// Exact variant: the four cascaded DF2T biquads with mul and add rounded
// separately (-ffp-contract=off), in the reference's operation order: the
// output is bit-identical to the CPU recurrence.  Throughput variant:
// bands below kDf2tBands run the same cascade with fused multiply-adds (17
// operations per sample and band instead of 30); the other bands run the
// equivalent FIR + shared all-pole form (14 operations).  With WAVES
// dividing kDf2tBands, the DF2T bands are exactly the first kDf2tBands/WAVES
// slots of block row 0, fixed at compile time.
#include <array>
#include <utility>

#include "dev_common.h"
#include "kernels.h"

namespace vq {

constexpr int kFramesPerWave = 64;

// Block shapes the launcher offers; index 0 runs until a shape is chosen.
constexpr GtShape kGtShapes[] = {
    {4, 4, 32}, {8, 4, 32}, {4, 2, 32}, {8, 2, 32}, {4, 4, 64}, {8, 4, 16},
};
constexpr int kGtShapeCount = (int)(sizeof(kGtShapes) / sizeof(kGtShapes[0]));

// One band's recurrence state on one frame and the sum of squares of its
// output: the four DF2T biquads' two state values each, or in the FIR +
// all-pole form each all-pole section's last two outputs.
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

// Throughput form: u = Σ c_k x[t-k] (x the frame's windowed samples, zero
// before the frame) in transposed order, so u waits on one FMA after the new
// sample: u = c0·x + w1, w1' = c1·x + w2, w2' = c2·x + w3, w3' = c3·x +
// c4·x1 (x1 the previous sample); then the four sections
// y = u - d1·y[t-1] - d2·y[t-2].
struct FirState {
    double w1 = 0, w2 = 0, w3 = 0;
};

__device__ inline void band_step_fir(BandState& b, FirState& f, const FirCoef& c, double s, double x1) {
    const double u = __builtin_fma(c.c0, s, f.w1);
    f.w1 = __builtin_fma(c.c1, s, f.w2);
    f.w2 = __builtin_fma(c.c2, s, f.w3);
    f.w3 = __builtin_fma(c.c3, s, c.c4 * x1);
    const double y1 = __builtin_fma(-c.d1, b.s1c0, __builtin_fma(-c.d2, b.s1c1, u));
    b.s1c1 = b.s1c0;
    b.s1c0 = y1;
    const double y2 = __builtin_fma(-c.d1, b.s2c0, __builtin_fma(-c.d2, b.s2c1, y1));
    b.s2c1 = b.s2c0;
    b.s2c0 = y2;
    const double y3 = __builtin_fma(-c.d1, b.s3c0, __builtin_fma(-c.d2, b.s3c1, y2));
    b.s3c1 = b.s3c0;
    b.s3c0 = y3;
    const double y4 = __builtin_fma(-c.d1, b.s4c0, __builtin_fma(-c.d2, b.s4c1, y3));
    b.s4c1 = b.s4c0;
    b.s4c0 = y4;
    b.sum_sq = __builtin_fma(y4, y4, b.sum_sq);
}

// One wave's share of a block: V valid slots, the first M of them in DF2T
// form.  Every variant stages its part of the tile and meets the block's
// barriers the same number of times.
template <bool FAST, int WAVES, int BPL, int CHUNK, int M, int V>
__device__ __forceinline__ void gt_run(const GtJob& job, const BandCoef* __restrict__ coef,
                                       const FirCoef* __restrict__ fir, const double* __restrict__ hann, int band0,
                                       int window, int hop, double (*tile)[CHUNK + 1]) {
    constexpr int kThreads = 64 * WAVES;
    constexpr int kStage = kFramesPerWave * CHUNK / kThreads;  // tile elements staged per thread
    constexpr int kSlots = V > 0 ? V : 1;
    const int f0 = blockIdx.x * kFramesPerWave;
    const int lane = threadIdx.x & 63;
    const int frame = f0 + lane;
    BandCoef k[kSlots];
    FirCoef c[kSlots];
    BandState st[kSlots];
    FirState fs[kSlots];
#pragma unroll
    for (int q = 0; q < V; q++) {
        if (q < M) k[q] = coef[band0 + q * WAVES];
        else c[q] = fir[band0 + q * WAVES];
    }
    const bool active = V > 0 && frame < job.nf;
    double x1 = 0.0;  // previous windowed sample (FIR slots)

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
        if (V > 0 && active) {
#pragma unroll 4
            for (int i = 0; i < clen; i++) {
                const double s = tile[lane][i];
#pragma unroll
                for (int q = 0; q < V; q++) {
                    if (q < M) band_step<FAST>(st[q], k[q], s);
                    else band_step_fir(st[q], fs[q], c[q], s, x1);
                }
                x1 = s;
            }
        }
    }
    if (V > 0 && active) {
        const double inv_n = 1.0 / (double)window;
#pragma unroll
        for (int q = 0; q < V; q++)
            job.out[(size_t)(band0 + q * WAVES) * job.nf + frame] = sqrt(st[q].sum_sq * inv_n);
    }
}

// Runs gt_run for the wave's valid slot count v (wave-uniform), VV = BPL
// down to 0, with min(MM, VV) DF2T slots.
template <bool FAST, int WAVES, int BPL, int CHUNK, int MM, int VV = BPL>
__device__ __forceinline__ void gt_dispatch(int v, const GtJob& job, const BandCoef* __restrict__ coef,
                                            const FirCoef* __restrict__ fir, const double* __restrict__ hann,
                                            int band0, int window, int hop, double (*tile)[CHUNK + 1]) {
    if (v == VV) {
        gt_run<FAST, WAVES, BPL, CHUNK, (MM < VV ? MM : VV), VV>(job, coef, fir, hann, band0, window, hop, tile);
    } else if constexpr (VV > 0) {
        gt_dispatch<FAST, WAVES, BPL, CHUNK, MM, VV - 1>(v, job, coef, fir, hann, band0, window, hop, tile);
    }
}

template <bool FAST, int WAVES, int BPL, int CHUNK>
__global__ __launch_bounds__(64 * WAVES) void k_gammatone(const GtJob* __restrict__ jobs, const BandCoef* __restrict__ coef,
                                                          const FirCoef* __restrict__ fir,
                                                          const double* __restrict__ hann, int nb, int window, int hop) {
    static_assert(kFramesPerWave * CHUNK % (64 * WAVES) == 0, "tile must split evenly over the block");
    static_assert(kDf2tBands % WAVES == 0 && WAVES * BPL >= kDf2tBands,
                  "the DF2T bands must fill whole slots of block row 0");
    const GtJob job = jobs[blockIdx.z];
    if ((int)blockIdx.x * kFramesPerWave >= job.nf) return;  // whole block idle: uniform exit
    __shared__ double tile[kFramesPerWave][CHUNK + 1];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int band0 = blockIdx.y * (WAVES * BPL) + wave;
    const int v = band0 < nb ? min(BPL, (nb - band0 + WAVES - 1) / WAVES) : 0;
    constexpr int kM0 = FAST ? (BPL < kDf2tBands / WAVES ? BPL : kDf2tBands / WAVES) : BPL;
    if (!FAST || blockIdx.y == 0)
        gt_dispatch<FAST, WAVES, BPL, CHUNK, kM0>(v, job, coef, fir, hann, band0, window, hop, tile);
    else
        gt_dispatch<FAST, WAVES, BPL, CHUNK, 0>(v, job, coef, fir, hann, band0, window, hop, tile);
}

namespace {
template <bool FAST, int V>
void launch_shape(const GtJob* d_jobs, int njobs, int max_nf, const BandCoef* d_coef, const FirCoef* d_fir,
                  const double* d_hann, int nb, int window, int hop, hipStream_t stream) {
    constexpr GtShape sh = kGtShapes[V];
    dim3 grid((max_nf + kFramesPerWave - 1) / kFramesPerWave, (nb + sh.waves * sh.bands_per_lane - 1) / (sh.waves * sh.bands_per_lane),
              njobs);
    hipLaunchKernelGGL((k_gammatone<FAST, sh.waves, sh.bands_per_lane, sh.chunk>), grid, dim3(64 * sh.waves), 0, stream,
                       d_jobs, d_coef, d_fir, d_hann, nb, window, hop);
}

using GtLaunch = void (*)(const GtJob*, int, int, const BandCoef*, const FirCoef*, const double*, int, int, int,
                         hipStream_t);

template <bool FAST, int... V>
constexpr std::array<GtLaunch, sizeof...(V)> shape_table(std::integer_sequence<int, V...>) {
    return {&launch_shape<FAST, V>...};
}
constexpr auto kFastLaunch = shape_table<true>(std::make_integer_sequence<int, kGtShapeCount>{});
constexpr auto kExactLaunch = shape_table<false>(std::make_integer_sequence<int, kGtShapeCount>{});
}  // namespace

int gammatone_shape_count() { return kGtShapeCount; }

GtShape gammatone_shape(int v) { return kGtShapes[v >= 0 && v < kGtShapeCount ? v : 0]; }

void launch_gammatone(const GtJob* d_jobs, int njobs, int max_nf, const BandCoef* d_coef, const FirCoef* d_fir,
                      const double* d_hann, int nb, int window, int hop, bool fast, int shape, hipStream_t stream) {
    if (njobs == 0 || max_nf == 0) return;
    const int v = shape >= 0 && shape < kGtShapeCount ? shape : 0;
    (fast ? kFastLaunch : kExactLaunch)[v](d_jobs, njobs, max_nf, d_coef, d_fir, d_hann, nb, window, hop, stream);
}

}  // namespace vq
