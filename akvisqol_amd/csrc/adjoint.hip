// Backward of the gammatone spectrogram on gfx950 (gammatone_bwd.slang:84-183,
// pipeline.zig:1020-1060): each frame's slope with respect to its samples,
// then the frames' overlap summed into the signal's slope.
//
// A spectrogram cell is rms_b = sqrt(Σ_t y_b[t]² / W) with y_b = H_b u, u the
// Hann-windowed frame and H_b band b's cascade of four biquads from zero
// state: a causal, time-invariant map, lower-triangular Toeplitz.  With
// v_b = g_b · y_b / (W · rms_b) the frame's slope is hann ⊙ Σ_b H_bᵀ v_b, and
// H_bᵀ v is the same cascade run over v backwards in time from zero state
// (reversal turns the Toeplitz matrix into its transpose).  So the adjoint
// needs no new recurrence, only y_b in reverse order: a first pass runs the
// cascade forward and checkpoints its state every kAdjSpan samples in
// scratch; the second walks the spans from the last, recomputes each span's
// outputs from its checkpoint into LDS, and runs the cascade backwards over
// them.  The bands' slopes of each sample are then summed from LDS in band
// order, so the result does not depend on the launch.  f64 with fused
// multiply-adds throughout (the slope is not compared bit for bit).
//
// Work decomposition: one wave per block, two frames per wave, one band per
// lane (32 lanes per frame); the frames' windowed samples are staged in LDS
// one span at a time and read by their 32 band lanes.
#include "dev_common.h"
#include "kernels.h"

namespace vq {

namespace {

constexpr int kAdjFrames = 2;
constexpr int kAdjLanes = 64;
static_assert(kAdjFrames * kMaxBands == kAdjLanes, "two frames of 32 band lanes per wave");

struct Cascade {
    double s1c0 = 0, s1c1 = 0, s2c0 = 0, s2c1 = 0, s3c0 = 0, s3c1 = 0, s4c0 = 0, s4c1 = 0;
};

// One sample through the four DF2T biquads; returns the last one's output.
__device__ inline double cascade_step(Cascade& z, const BandCoef& k, double s) {
    const double f1 = __builtin_fma(k.s1n0, s, z.s1c0);
    z.s1c0 = __builtin_fma(-k.d1, f1, __builtin_fma(k.s1n1, s, z.s1c1));
    z.s1c1 = -k.d2 * f1;
    const double f2 = __builtin_fma(k.ca0, f1, z.s2c0);
    z.s2c0 = __builtin_fma(-k.d1, f2, __builtin_fma(k.s2n1, f1, z.s2c1));
    z.s2c1 = -k.d2 * f2;
    const double f3 = __builtin_fma(k.ca0, f2, z.s3c0);
    z.s3c0 = __builtin_fma(-k.d1, f3, __builtin_fma(k.s3n1, f2, z.s3c1));
    z.s3c1 = -k.d2 * f3;
    const double f4 = __builtin_fma(k.ca0, f3, z.s4c0);
    z.s4c0 = __builtin_fma(-k.d1, f4, __builtin_fma(k.s4n1, f3, z.s4c1));
    z.s4c1 = -k.d2 * f4;
    return f4;
}

__device__ inline void save(double* p, const Cascade& z) {
    p[0] = z.s1c0; p[1] = z.s1c1; p[2] = z.s2c0; p[3] = z.s2c1;
    p[4] = z.s3c0; p[5] = z.s3c1; p[6] = z.s4c0; p[7] = z.s4c1;
}

__device__ inline Cascade restore(const double* p) {
    Cascade z;
    z.s1c0 = p[0]; z.s1c1 = p[1]; z.s2c0 = p[2]; z.s2c1 = p[3];
    z.s3c0 = p[4]; z.s3c1 = p[5]; z.s4c0 = p[6]; z.s4c1 = p[7];
    return z;
}

__global__ __launch_bounds__(kAdjLanes) void k_gammatone_adjoint(const FrameGradJob* __restrict__ jobs, int njobs,
                                                                const BandCoef* __restrict__ coef,
                                                                const double* __restrict__ hann, int nb, int window,
                                                                double* __restrict__ ckpt) {
    __shared__ double u_s[kAdjFrames][kAdjSpan];
    __shared__ double w_s[kAdjLanes][kAdjSpan + 1];
    const int lane = threadIdx.x;
    const int fs = lane / kMaxBands, b = lane % kMaxBands;
    const int j0 = blockIdx.x * kAdjFrames;
    // both frames' jobs (the second may be past the end: then it is unused)
    FrameGradJob jb[kAdjFrames];
    bool has[kAdjFrames];
    for (int f = 0; f < kAdjFrames; f++) {
        has[f] = j0 + f < njobs;
        jb[f] = jobs[has[f] ? j0 + f : j0];
    }
    const bool live = has[fs] && b < nb;
    const BandCoef k = coef[b < nb ? b : 0];
    const int nspan = (window + kAdjSpan - 1) / kAdjSpan;
    double* ck = ckpt + ((size_t)blockIdx.x * kAdjLanes + lane) * (size_t)nspan * 8;

    // u_s[f][t] = hann · sample of span `sp`, zero outside the frame's [lo, hi)
    auto stage = [&](int sp) {
        for (int i = lane; i < kAdjFrames * kAdjSpan; i += kAdjLanes) {
            const int f = i / kAdjSpan, t = sp * kAdjSpan + i % kAdjSpan;
            double v = 0.0;
            if (has[f] && t < window && t >= jb[f].lo && t < jb[f].hi) v = hann[t] * jb[f].sig[jb[f].start + t];
            u_s[f][i % kAdjSpan] = v;
        }
    };

    // pass 1: forward, checkpoints, sum of squares
    Cascade z;
    double ss = 0.0;
    for (int sp = 0; sp < nspan; sp++) {
        __syncthreads();
        stage(sp);
        __syncthreads();
        save(ck + (size_t)sp * 8, z);
        const int n = min(kAdjSpan, window - sp * kAdjSpan);
        for (int t = 0; t < n; t++) {
            const double y = cascade_step(z, k, u_s[fs][t]);
            ss = __builtin_fma(y, y, ss);
        }
    }
    // ∂(g · rms)/∂y[t] = c · y[t]
    double c = 0.0;
    if (live && ss > 0.0) {
        const double rms = sqrt(ss / (double)window);
        c = jb[fs].g_col[(size_t)b * jb[fs].stride] / ((double)window * rms);
    }

    // pass 2: spans from the last, the cascade run backwards over c · y
    Cascade a;
    for (int sp = nspan - 1; sp >= 0; sp--) {
        __syncthreads();
        stage(sp);
        __syncthreads();
        const int n = min(kAdjSpan, window - sp * kAdjSpan);
        z = restore(ck + (size_t)sp * 8);
        for (int t = 0; t < n; t++) w_s[lane][t] = cascade_step(z, k, u_s[fs][t]);
        for (int t = n - 1; t >= 0; t--) w_s[lane][t] = cascade_step(a, k, c * w_s[lane][t]);
        __syncthreads();
        // Σ over bands in band order, times the window, masked to [lo, hi)
        for (int i = lane; i < kAdjFrames * kAdjSpan; i += kAdjLanes) {
            const int f = i / kAdjSpan, tt = i % kAdjSpan, t = sp * kAdjSpan + tt;
            if (!has[f] || tt >= n) continue;
            double acc = 0.0;
            for (int bb = 0; bb < nb; bb++) acc += w_s[f * kMaxBands + bb][tt];
            jb[f].out[t] = (t >= jb[f].lo && t < jb[f].hi) ? hann[t] * acc : 0.0;
        }
    }
}

__global__ __launch_bounds__(256) void k_frame_accumulate(const AccRun* __restrict__ runs,
                                                          const int* __restrict__ list,
                                                          const FrameGradJob* __restrict__ frames) {
    const AccRun r = runs[blockIdx.x];
    for (int i = threadIdx.x; i < r.len; i += blockDim.x) {
        const long long t = r.pos + i;
        double acc = 0.0;
        for (int e = r.first; e < r.first + r.count; e++) {
            const FrameGradJob& f = frames[list[e]];
            const long long k = t - f.start;
            if (k >= f.lo && k < f.hi) acc += f.out[k];
        }
        r.dst[i] = acc;
    }
}

}  // namespace

size_t gammatone_adjoint_scratch(int njobs, int window) {
    const size_t blocks = (size_t)(njobs + kAdjFrames - 1) / kAdjFrames;
    return blocks * kAdjLanes * (size_t)((window + kAdjSpan - 1) / kAdjSpan) * 8;
}

void launch_gammatone_adjoint(const FrameGradJob* d_jobs, int njobs, const BandCoef* d_coef, const double* d_hann,
                              int nb, int window, double* d_ckpt, hipStream_t stream) {
    if (njobs <= 0) return;
    const int blocks = (njobs + kAdjFrames - 1) / kAdjFrames;
    hipLaunchKernelGGL(k_gammatone_adjoint, dim3(blocks), dim3(kAdjLanes), 0, stream, d_jobs, njobs, d_coef, d_hann,
                       nb, window, d_ckpt);
}

void launch_frame_accumulate(const AccRun* d_runs, int nruns, const int* d_list, const FrameGradJob* d_frames,
                             hipStream_t stream) {
    if (nruns <= 0) return;
    hipLaunchKernelGGL(k_frame_accumulate, dim3(nruns), dim3(256), 0, stream, d_runs, d_list, d_frames);
}

}  // namespace vq
