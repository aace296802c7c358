// Gammatone ERB filterbank spectrogram on gfx950 (gammatone.zig:268-306,
// 432-467; accel.zig:116-124).
//
// Work decomposition: one lane per (frame, band).  A wave holds 64
// consecutive frames of ONE band, so the band's eight coefficients are
// wave-uniform (scalar registers) and the four cascaded DF2T biquads run in
// vector registers, in the reference's exact operation order.  A block
// holds `bands_per_block` waves over the same 64 frames; the frames'
// Hann-windowed samples are staged chunk by chunk in LDS (coalesced 512 B
// rows from HBM, x·hann computed once per block and shared by its bands).
//
// Exact variant: mul and add rounded separately (-ffp-contract=off), the
// output bit-identical to the CPU recurrence.  Fast variant: the same
// recurrence with fused multiply-adds (17 instead of 30 operations per
// sample and band).
#include "dev_common.h"
#include "kernels.h"

namespace vq {

constexpr int kFramesPerWave = 64;
constexpr int kChunk = 64;

template <bool FAST>
__global__ __launch_bounds__(512) void k_gammatone(const GtJob* __restrict__ jobs, const BandCoef* __restrict__ coef,
                                                   const double* __restrict__ hann, int nb, int window, int hop,
                                                   int bands_per_block) {
    const GtJob job = jobs[blockIdx.z];
    const int f0 = blockIdx.x * kFramesPerWave;
    if (f0 >= job.nf) return;  // whole block idle: uniform exit
    __shared__ double tile[kFramesPerWave][kChunk + 1];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int band = blockIdx.y * bands_per_block + wave;
    const bool active_band = band < nb;
    const int frame = f0 + lane;
    const bool active = active_band && frame < job.nf;

    BandCoef k;
    if (active_band) k = coef[band];
    else k = BandCoef{0, 0, 0, 0, 0, 0, 0, 0};
    double s1c0 = 0, s1c1 = 0, s2c0 = 0, s2c1 = 0, s3c0 = 0, s3c1 = 0, s4c0 = 0, s4c1 = 0, sum_sq = 0;

    const double* sig = job.sig;
    const int nthreads = blockDim.x;
    for (int i0 = 0; i0 < window; i0 += kChunk) {
        const int clen = min(kChunk, window - i0);
        __syncthreads();
        for (int e = threadIdx.x; e < kFramesPerWave * kChunk; e += nthreads) {
            const int fr = e / kChunk, i = e % kChunk;
            double v = 0.0;
            if (f0 + fr < job.nf && i < clen) v = sig[(size_t)(f0 + fr) * hop + i0 + i] * hann[i0 + i];
            tile[fr][i] = v;
        }
        __syncthreads();
        if (active) {
            for (int i = 0; i < clen; i++) {
                const double s = tile[lane][i];
                if (FAST) {
                    const double f1 = __builtin_fma(k.s1n0, s, s1c0);
                    s1c0 = __builtin_fma(-k.d1, f1, __builtin_fma(k.s1n1, s, s1c1));
                    s1c1 = -k.d2 * f1;
                    const double f2 = __builtin_fma(k.ca0, f1, s2c0);
                    s2c0 = __builtin_fma(-k.d1, f2, __builtin_fma(k.s2n1, f1, s2c1));
                    s2c1 = -k.d2 * f2;
                    const double f3 = __builtin_fma(k.ca0, f2, s3c0);
                    s3c0 = __builtin_fma(-k.d1, f3, __builtin_fma(k.s3n1, f2, s3c1));
                    s3c1 = -k.d2 * f3;
                    const double f4 = __builtin_fma(k.ca0, f3, s4c0);
                    s4c0 = __builtin_fma(-k.d1, f4, __builtin_fma(k.s4n1, f3, s4c1));
                    s4c1 = -k.d2 * f4;
                    sum_sq = __builtin_fma(f4, f4, sum_sq);
                } else {
                    const double f1 = k.s1n0 * s + s1c0;
                    s1c0 = k.s1n1 * s + s1c1 - k.d1 * f1;
                    s1c1 = -(k.d2 * f1);
                    const double f2 = k.ca0 * f1 + s2c0;
                    s2c0 = k.s2n1 * f1 + s2c1 - k.d1 * f2;
                    s2c1 = -(k.d2 * f2);
                    const double f3 = k.ca0 * f2 + s3c0;
                    s3c0 = k.s3n1 * f2 + s3c1 - k.d1 * f3;
                    s3c1 = -(k.d2 * f3);
                    const double f4 = k.ca0 * f3 + s4c0;
                    s4c0 = k.s4n1 * f3 + s4c1 - k.d1 * f4;
                    s4c1 = -(k.d2 * f4);
                    sum_sq += f4 * f4;
                }
            }
        }
    }
    if (active) {
        const double inv_n = 1.0 / (double)window;
        job.out[(size_t)band * job.nf + frame] = sqrt(sum_sq * inv_n);
    }
}

void launch_gammatone(const GtJob* d_jobs, int njobs, int max_nf, const BandCoef* d_coef, const double* d_hann, int nb,
                      int window, int hop, bool fast, hipStream_t stream) {
    if (njobs == 0 || max_nf == 0) return;
    const int bpb = nb % 4 == 0 ? 4 : (nb % 3 == 0 ? 3 : 4);
    dim3 grid((max_nf + kFramesPerWave - 1) / kFramesPerWave, (nb + bpb - 1) / bpb, njobs);
    dim3 block(64 * bpb);
    if (fast)
        hipLaunchKernelGGL(k_gammatone<true>, grid, block, 0, stream, d_jobs, d_coef, d_hann, nb, window, hop, bpb);
    else
        hipLaunchKernelGGL(k_gammatone<false>, grid, block, 0, stream, d_jobs, d_coef, d_hann, nb, window, hop, bpb);
}

}  // namespace vq
