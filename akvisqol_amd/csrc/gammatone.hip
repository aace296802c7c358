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
// slots of block row 0 (in blocks of fewer bands than kDf2tBands, the
// first block rows), fixed at compile time.
#include <array>
#include <utility>

#include "dev_common.h"
#include "kernels.h"

namespace vq {

constexpr int kFramesPerWave = kFramesPerUnit;

// Block shapes the launcher offers; index 0 runs until a shape is chosen.
// The last, one band per lane and four bands per block (one wave per SIMD),
// spreads a launch over the most blocks: launches too small to fill the
// device (a served pair or two) run it, since their time is one block's
// serial work.
constexpr GtShape kGtShapes[] = {
    {4, 4, 32}, {8, 4, 32}, {4, 2, 32}, {8, 2, 32}, {4, 4, 64}, {8, 4, 16}, {4, 1, 64},
};
constexpr int kGtShapeCount = (int)(sizeof(kGtShapes) / sizeof(kGtShapes[0]));
constexpr int kGtWideShape = kGtShapeCount - 1;
static_assert(kGtShapes[kGtWideShape].bands_per_lane == 1, "the wide shape runs one band per lane");

// One band's recurrence state on one frame and the sum of squares of its
// output: the four DF2T biquads' two state values each, or in the FIR +
// all-pole form each all-pole section's last two outputs.
struct BandState {
    double s1c0 = 0, s1c1 = 0, s2c0 = 0, s2c1 = 0, s3c0 = 0, s3c1 = 0, s4c0 = 0, s4c1 = 0, sum_sq = 0;
    // pipelined steps: sections 1-3's outputs of the previous step, the
    // next sections' inputs at this one
    double p1 = 0, p2 = 0, p3 = 0;
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

// Pipelined forms of band_step and band_step_fir for lanes that carry one
// band: at step n section k runs on section k-1's output of step n-1
// (section 1 on sample n), so the four sections' recurrences are
// independent within a step and a lone wave waits on one section's
// latency a step instead of the cascade's.  Each section performs the same
// operations on the same inputs in the same order as in the plain step, so
// every output is bit-identical; the sum of squares gains only the +0
// squares of the first three steps (zero-state sections on zero inputs
// stay zero) and needs three drain steps after the frame's last sample.
template <bool FAST>
__device__ inline void band_step_pipe(BandState& b, const BandCoef& k, double s) {
    const double in2 = b.p1, in3 = b.p2, in4 = b.p3;
    if (FAST) {
        const double f1 = __builtin_fma(k.s1n0, s, b.s1c0);
        b.s1c0 = __builtin_fma(-k.d1, f1, __builtin_fma(k.s1n1, s, b.s1c1));
        b.s1c1 = -k.d2 * f1;
        const double f2 = __builtin_fma(k.ca0, in2, b.s2c0);
        b.s2c0 = __builtin_fma(-k.d1, f2, __builtin_fma(k.s2n1, in2, b.s2c1));
        b.s2c1 = -k.d2 * f2;
        const double f3 = __builtin_fma(k.ca0, in3, b.s3c0);
        b.s3c0 = __builtin_fma(-k.d1, f3, __builtin_fma(k.s3n1, in3, b.s3c1));
        b.s3c1 = -k.d2 * f3;
        const double f4 = __builtin_fma(k.ca0, in4, b.s4c0);
        b.s4c0 = __builtin_fma(-k.d1, f4, __builtin_fma(k.s4n1, in4, b.s4c1));
        b.s4c1 = -k.d2 * f4;
        b.sum_sq = __builtin_fma(f4, f4, b.sum_sq);
        b.p1 = f1;
        b.p2 = f2;
        b.p3 = f3;
    } else {
        const double f1 = k.s1n0 * s + b.s1c0;
        b.s1c0 = k.s1n1 * s + b.s1c1 - k.d1 * f1;
        b.s1c1 = -(k.d2 * f1);
        const double f2 = k.ca0 * in2 + b.s2c0;
        b.s2c0 = k.s2n1 * in2 + b.s2c1 - k.d1 * f2;
        b.s2c1 = -(k.d2 * f2);
        const double f3 = k.ca0 * in3 + b.s3c0;
        b.s3c0 = k.s3n1 * in3 + b.s3c1 - k.d1 * f3;
        b.s3c1 = -(k.d2 * f3);
        const double f4 = k.ca0 * in4 + b.s4c0;
        b.s4c0 = k.s4n1 * in4 + b.s4c1 - k.d1 * f4;
        b.s4c1 = -(k.d2 * f4);
        b.sum_sq += f4 * f4;
        b.p1 = f1;
        b.p2 = f2;
        b.p3 = f3;
    }
}

__device__ inline void band_step_fir_pipe(BandState& b, FirState& f, const FirCoef& c, double s, double x1) {
    const double in2 = b.p1, in3 = b.p2, in4 = b.p3;
    const double u = __builtin_fma(c.c0, s, f.w1);
    f.w1 = __builtin_fma(c.c1, s, f.w2);
    f.w2 = __builtin_fma(c.c2, s, f.w3);
    f.w3 = __builtin_fma(c.c3, s, c.c4 * x1);
    const double y1 = __builtin_fma(-c.d1, b.s1c0, __builtin_fma(-c.d2, b.s1c1, u));
    b.s1c1 = b.s1c0;
    b.s1c0 = y1;
    const double y2 = __builtin_fma(-c.d1, b.s2c0, __builtin_fma(-c.d2, b.s2c1, in2));
    b.s2c1 = b.s2c0;
    b.s2c0 = y2;
    const double y3 = __builtin_fma(-c.d1, b.s3c0, __builtin_fma(-c.d2, b.s3c1, in3));
    b.s3c1 = b.s3c0;
    b.s3c0 = y3;
    const double y4 = __builtin_fma(-c.d1, b.s4c0, __builtin_fma(-c.d2, b.s4c1, in4));
    b.s4c1 = b.s4c0;
    b.s4c0 = y4;
    b.sum_sq = __builtin_fma(y4, y4, b.sum_sq);
    b.p1 = y1;
    b.p2 = y2;
    b.p3 = y3;
}

// One wave's share of a block: V valid slots, the first M of them in DF2T
// form.  Every variant stages its part of the tile and meets the block's
// barriers the same number of times.  PAIR: the block holds 32 frames of
// two jobs (tile rows and lanes 0-31 job a, 32-63 job b), for launches whose
// jobs have at most 32 frames; otherwise 64 frames of job a.
template <bool FAST, int WAVES, int BPL, int CHUNK, bool PAIR, int M, int V>
__device__ __forceinline__ void gt_run(const GtJob& ja, const GtJob& jb, const BandCoef* __restrict__ coef,
                                       const FirCoef* __restrict__ fir, const double* __restrict__ hann, int band0,
                                       int window, int hop, double (*tile)[CHUNK + 1], int fblock) {
    constexpr int kThreads = 64 * WAVES;
    constexpr int kStage = kFramesPerWave * CHUNK / kThreads;  // tile elements staged per thread
    // One band a lane: too few lanes to hide a cascade's latency, so the
    // sections run pipelined (with several bands a lane, the bands'
    // independent cascades already interleave)
    constexpr bool kPipe = BPL == 1;
    constexpr int kSlots = V > 0 ? V : 1;
    constexpr int kFramesPerJob = PAIR ? kFramesPerWave / 2 : kFramesPerWave;
    const int f0 = fblock * kFramesPerJob;
    // tile row (= lane) -> job and frame
    auto row_b = [](int row) { return PAIR && row >= kFramesPerWave / 2; };
    auto row_frame = [&](int row) { return f0 + (PAIR ? row & (kFramesPerWave / 2 - 1) : row); };
    // a job holding frame f0 (the kernel exits before any block that has none)
    const GtJob& jv = ja.nf > f0 ? ja : jb;
    const int lane = threadIdx.x & 63;
    const GtJob& job = row_b(lane) ? jb : ja;
    const int frame = row_frame(lane);
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
    const int ti = threadIdx.x % CHUNK;
    const int fr0 = threadIdx.x / CHUNK;
    // A staged row past its job's frames reads frame f0 of a job that has
    // it, and is masked at the LDS write.
    auto row_valid = [&](int fr) { return row_frame(fr) < (row_b(fr) ? jb.nf : ja.nf); };
    auto row_src = [&](int fr) -> const double* {
        if (!PAIR) return ja.sig + (size_t)(row_valid(fr) ? row_frame(fr) : f0) * hop;
        return row_valid(fr) ? (row_b(fr) ? jb.sig : ja.sig) + (size_t)row_frame(fr) * hop : jv.sig + (size_t)f0 * hop;
    };
    double v[kStage];
    double h = 0.0;
    auto load_chunk = [&](int i0) {
        const int i = i0 + ti;
        const int ic = i < window ? i : 0;
        h = hann[ic];
#pragma unroll
        for (int u = 0; u < kStage; u++) v[u] = as_global(row_src(fr0 + u * kRowsPerPass))[ic];
    };
    load_chunk(0);
    for (int i0 = 0; i0 < window; i0 += CHUNK) {
        const int clen = min(CHUNK, window - i0);
        __syncthreads();
        const bool in = i0 + ti < window;
#pragma unroll
        for (int u = 0; u < kStage; u++) {
            const int fr = fr0 + u * kRowsPerPass;
            tile[fr][ti] = (in && row_valid(fr)) ? v[u] * h : 0.0;
        }
        __syncthreads();
        if (i0 + CHUNK < window) load_chunk(i0 + CHUNK);
        if (V > 0 && active) {
#pragma unroll 4
            for (int i = 0; i < clen; i++) {
                const double s = tile[lane][i];
#pragma unroll
                for (int q = 0; q < V; q++) {
                    if (kPipe) {
                        if (q < M) band_step_pipe<FAST>(st[q], k[q], s);
                        else band_step_fir_pipe(st[q], fs[q], c[q], s, x1);
                    } else {
                        if (q < M) band_step<FAST>(st[q], k[q], s);
                        else band_step_fir(st[q], fs[q], c[q], s, x1);
                    }
                }
                x1 = s;
            }
        }
    }
    if (kPipe && V > 0 && active) {
        // drain: sections 2-4 take the last three samples' pending values
        // (section 1's outputs of these steps are never used)
        for (int d = 0; d < 3; d++) {
#pragma unroll
            for (int q = 0; q < V; q++) {
                if (q < M) band_step_pipe<FAST>(st[q], k[q], 0.0);
                else band_step_fir_pipe(st[q], fs[q], c[q], 0.0, x1);
            }
            x1 = 0.0;
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
template <bool FAST, int WAVES, int BPL, int CHUNK, bool PAIR, int MM, int VV = BPL>
__device__ __forceinline__ void gt_dispatch(int v, const GtJob& ja, const GtJob& jb, const BandCoef* __restrict__ coef,
                                            const FirCoef* __restrict__ fir, const double* __restrict__ hann,
                                            int band0, int window, int hop, double (*tile)[CHUNK + 1], int fblock) {
    if (v == VV) {
        gt_run<FAST, WAVES, BPL, CHUNK, PAIR, (MM < VV ? MM : VV), VV>(ja, jb, coef, fir, hann, band0, window, hop,
                                                                        tile, fblock);
    } else if constexpr (VV > 0) {
        gt_dispatch<FAST, WAVES, BPL, CHUNK, PAIR, MM, VV - 1>(v, ja, jb, coef, fir, hann, band0, window, hop, tile,
                                                              fblock);
    }
}

// One block's work: 64 frames (frame block fblock; PAIR: 32 of each job)
// of block row `row` of the bands.
template <bool FAST, int WAVES, int BPL, int CHUNK, bool PAIR>
__device__ __forceinline__ void gt_block(const GtJob& ja, const GtJob& jb, const BandCoef* __restrict__ coef,
                                         const FirCoef* __restrict__ fir, const double* __restrict__ hann, int nb,
                                         int window, int hop, int fblock, int row, double (*tile)[CHUNK + 1]) {
    // The DF2T bands fill whole slots of block row 0, or whole block rows
    // when a block holds fewer bands than there are DF2T bands.
    constexpr bool kNarrow = WAVES * BPL < kDf2tBands;
    static_assert(kNarrow ? kDf2tBands % (WAVES * BPL) == 0 : kDf2tBands % WAVES == 0,
                  "the DF2T bands must fill whole slots of block row 0, or whole block rows");
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int band0 = row * (WAVES * BPL) + wave;
    const int v = band0 < nb ? min(BPL, (nb - band0 + WAVES - 1) / WAVES) : 0;
    constexpr int kM0 = FAST ? (BPL < kDf2tBands / WAVES ? BPL : kDf2tBands / WAVES) : BPL;
    if (!FAST || (kNarrow ? row * WAVES * BPL < kDf2tBands : row == 0))
        gt_dispatch<FAST, WAVES, BPL, CHUNK, PAIR, kM0>(v, ja, jb, coef, fir, hann, band0, window, hop, tile, fblock);
    else
        gt_dispatch<FAST, WAVES, BPL, CHUNK, PAIR, 0>(v, ja, jb, coef, fir, hann, band0, window, hop, tile, fblock);
}

template <bool FAST, int WAVES, int BPL, int CHUNK, bool PAIR>
__global__ __launch_bounds__(64 * WAVES) void k_gammatone(const GtJob* __restrict__ jobs, int njobs,
                                                          const BandCoef* __restrict__ coef,
                                                          const FirCoef* __restrict__ fir,
                                                          const double* __restrict__ hann, int nb, int window, int hop) {
    static_assert(kFramesPerWave * CHUNK % (64 * WAVES) == 0, "tile must split evenly over the block");
    const int za = PAIR ? 2 * (int)blockIdx.z : (int)blockIdx.z;
    const GtJob ja = jobs[za];
    GtJob jb = ja;
    if (PAIR) {
        if (za + 1 < njobs) jb = jobs[za + 1];
        else jb.nf = 0;  // odd job count: lanes 32-63 idle
    }
    const int frames = PAIR ? kFramesPerWave / 2 : kFramesPerWave;
    if ((int)blockIdx.x * frames >= (ja.nf > jb.nf ? ja.nf : jb.nf)) return;  // whole block idle: uniform exit
    __shared__ double tile[kFramesPerWave][CHUNK + 1];
    gt_block<FAST, WAVES, BPL, CHUNK, PAIR>(ja, jb, coef, fir, hann, nb, window, hop, (int)blockIdx.x,
                                            (int)blockIdx.y, tile);
}

// Persistent form for launches whose jobs differ in length: a grid of at
// most the blocks the device holds at once walks a compact list of the
// (job, frame block, band row) units that have frames, unit u on block
// u mod gridDim.x, instead of a grid sized by the longest job whose blocks
// past a shorter job's end exit at once.  Each unit runs exactly as that
// block of the grid form, so the bits are the same.
template <bool FAST, int WAVES, int BPL, int CHUNK>
__global__ __launch_bounds__(64 * WAVES) void k_gammatone_units(const GtJob* __restrict__ jobs,
                                                                const GtUnit* __restrict__ units, int nunits,
                                                                const BandCoef* __restrict__ coef,
                                                                const FirCoef* __restrict__ fir,
                                                                const double* __restrict__ hann, int nb, int window,
                                                                int hop) {
    __shared__ double tile[kFramesPerWave][CHUNK + 1];
    for (int u = blockIdx.x; u < nunits; u += gridDim.x) {
        const GtUnit w = units[u];
        const GtJob ja = jobs[w.job];
        // the previous unit's last reads of the tile precede this one's writes
        __syncthreads();
        gt_block<FAST, WAVES, BPL, CHUNK, false>(ja, ja, coef, fir, hann, nb, window, hop, w.block, w.row, tile);
    }
}

namespace {
template <bool FAST, int V, bool PAIR = false>
void launch_shape(const GtJob* d_jobs, int njobs, int max_nf, const BandCoef* d_coef, const FirCoef* d_fir,
                  const double* d_hann, int nb, int window, int hop, hipStream_t stream) {
    constexpr GtShape sh = kGtShapes[V];
    constexpr int frames = PAIR ? kFramesPerWave / 2 : kFramesPerWave;
    // blocks past a job's last frame exit before any work
    dim3 grid(xcd_spread((max_nf + frames - 1) / frames),
              (nb + sh.waves * sh.bands_per_lane - 1) / (sh.waves * sh.bands_per_lane), PAIR ? (njobs + 1) / 2 : njobs);
    hipLaunchKernelGGL((k_gammatone<FAST, sh.waves, sh.bands_per_lane, sh.chunk, PAIR>), grid, dim3(64 * sh.waves), 0,
                       stream, d_jobs, njobs, d_coef, d_fir, d_hann, nb, window, hop);
}

using GtLaunch = void (*)(const GtJob*, int, int, const BandCoef*, const FirCoef*, const double*, int, int, int,
                         hipStream_t);

template <bool FAST, int V>
void launch_units_shape(const GtJob* d_jobs, const GtUnit* d_units, int nunits, const BandCoef* d_coef,
                        const FirCoef* d_fir, const double* d_hann, int nb, int window, int hop, hipStream_t stream) {
    constexpr GtShape sh = kGtShapes[V];
    auto kern = k_gammatone_units<FAST, sh.waves, sh.bands_per_lane, sh.chunk>;
    // as many blocks as the device holds at once (queried once per kernel)
    static int resident = [&] {
        int dev = 0, cus = 0, per = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, 64 * sh.waves, 0) != hipSuccess || cus <= 0 ||
            per <= 0) {
            (void)hipGetLastError();
            return 256;
        }
        return cus * per;
    }();
    const int grid = nunits < resident ? nunits : resident;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * sh.waves), 0, stream, d_jobs, d_units, nunits, d_coef, d_fir,
                       d_hann, nb, window, hop);
}

using GtUnitsLaunch = void (*)(const GtJob*, const GtUnit*, int, const BandCoef*, const FirCoef*, const double*, int,
                               int, int, hipStream_t);

template <bool FAST, int... V>
constexpr std::array<GtUnitsLaunch, sizeof...(V)> units_table(std::integer_sequence<int, V...>) {
    return {&launch_units_shape<FAST, V>...};
}
constexpr auto kFastUnits = units_table<true>(std::make_integer_sequence<int, kGtShapeCount>{});
constexpr auto kExactUnits = units_table<false>(std::make_integer_sequence<int, kGtShapeCount>{});

template <bool FAST, int... V>
constexpr std::array<GtLaunch, sizeof...(V)> shape_table(std::integer_sequence<int, V...>) {
    return {&launch_shape<FAST, V>...};
}
constexpr auto kFastLaunch = shape_table<true>(std::make_integer_sequence<int, kGtShapeCount>{});
constexpr auto kExactLaunch = shape_table<false>(std::make_integer_sequence<int, kGtShapeCount>{});
}  // namespace

int gammatone_shape_count() { return kGtShapeCount; }

int gammatone_wide_shape() { return kGtWideShape; }

GtShape gammatone_shape(int v) { return kGtShapes[v >= 0 && v < kGtShapeCount ? v : 0]; }

void launch_gammatone_units(const GtJob* d_jobs, const GtUnit* d_units, int nunits, const BandCoef* d_coef,
                            const FirCoef* d_fir, const double* d_hann, int nb, int window, int hop, bool fast,
                            int shape, hipStream_t stream) {
    if (nunits <= 0) return;
    const int v = shape >= 0 && shape < kGtShapeCount ? shape : 0;
    (fast ? kFastUnits : kExactUnits)[v](d_jobs, d_units, nunits, d_coef, d_fir, d_hann, nb, window, hop, stream);
}

void launch_gammatone(const GtJob* d_jobs, int njobs, int max_nf, const BandCoef* d_coef, const FirCoef* d_fir,
                      const double* d_hann, int nb, int window, int hop, bool fast, int shape, hipStream_t stream) {
    if (njobs == 0 || max_nf == 0) return;
    if (max_nf <= kFramesPerWave / 2 && njobs > 1) {  // short jobs (realignment patches): two per block
        if (shape == kGtWideShape) {
            if (fast) launch_shape<true, kGtWideShape, true>(d_jobs, njobs, max_nf, d_coef, d_fir, d_hann, nb, window, hop, stream);
            else launch_shape<false, kGtWideShape, true>(d_jobs, njobs, max_nf, d_coef, d_fir, d_hann, nb, window, hop, stream);
        } else {
            if (fast) launch_shape<true, 0, true>(d_jobs, njobs, max_nf, d_coef, d_fir, d_hann, nb, window, hop, stream);
            else launch_shape<false, 0, true>(d_jobs, njobs, max_nf, d_coef, d_fir, d_hann, nb, window, hop, stream);
        }
        return;
    }
    const int v = shape >= 0 && shape < kGtShapeCount ? shape : 0;
    (fast ? kFastLaunch : kExactLaunch)[v](d_jobs, njobs, max_nf, d_coef, d_fir, d_hann, nb, window, hop, stream);
}

}  // namespace vq
