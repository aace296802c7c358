// FFT cross-correlation lag search on gfx950 (xcorr.zig:180-251,
// envelope.zig:23-53, dsp/fft.zig:45-166).
//
// z = env(a) + i·env(b), zero padded to n, goes through a radix-2 DIF
// forward transform (bit-reversed out), the cross-spectrum is formed in
// bit-reversed order, and a DIT inverse (bit-reversed in) gives the circular
// correlation; the lag is the first maximum of |Re| scanning lags
// -max..-1 then 0..max.
//
// Each HBM pass runs up to 8 radix-2 stages on 2048-element groups in LDS:
// the stages with spans S_hi .. S_hi/2^(k-1) only mix elements
// blk·S_hi + r + m·q (q = S_hi/2^k) of one residue r, so a workgroup loads R
// residues of Bn span blocks (R·16 B contiguous per m).  Every butterfly is
// the radix-2 butterfly of its stage with the same twiddle, so results equal
// the stage-by-stage transform bit for bit.  The first DIF pass builds its
// input from the two signals (the envelope is never stored); the last DIT
// pass keeps no output, only each workgroup's first maximum in scan order.
#include "dev_common.h"
#include "kernels.h"

namespace vq {

__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// fft.zig:181-188: cmul and the conjugate form, in the reference's order.
__device__ inline double2 cmul(double2 x, double2 w) {
    return make_double2(x.x * w.x - x.y * w.y, x.x * w.y + x.y * w.x);
}
__device__ inline double2 cmul_conj(double2 x, double2 w) {
    const double wi = -w.y;
    return make_double2(x.x * w.x - x.y * wi, x.x * wi + x.y * w.x);
}

__device__ inline long long env_fft_size(long long l) {
    long long n = 1;
    while (n < l) n <<= 1;
    return n < 32 ? 32 : n;
}

// envelope.zig:23-53 closed form for sample i of a signal with the given
// plain sum and alternating (Nyquist) sum.
__device__ inline double env_at(const double* x, long long i, long long len, double sum, double nyq) {
    const double mean = sum / (double)len;
    const double nq = nyq / (double)env_fft_size(len);
    const double term = (i & 1) ? -nq : nq;
    return fabs(2.0 * ((x[i] - mean) - term) - 0.000001) + mean;
}

__device__ inline double2 packed_env(const LagJob& j, const double* sums, long long i) {
    double re = 0.0, im = 0.0;
    if (i < j.la) re = env_at(j.a, i, j.la, sums[j.sum_a], sums[j.nyq_a]);
    if (i < j.lb) im = env_at(j.b, i, j.lb, sums[j.sum_b], sums[j.nyq_b]);
    return make_double2(re, im);
}

constexpr int kTile = 2048;  // complex doubles per workgroup (32 KiB LDS)
constexpr int kThreads = 256;
constexpr int kBfPerThread = kTile / 2 / kThreads;

enum : int { kSrcZ = 0, kSrcEnvelope = 1 };

// Scan position of circular index i (xcorr.zig:233-250), or -1 outside
// the lag range.
__device__ inline long long scan_order(long long i, long long n, long long max_lag) {
    if (i >= n - max_lag) return i - (n - max_lag);
    if (i <= max_lag) return max_lag + i;
    return -1;
}

template <bool DIT, int SRC, bool ARGMAX>
__global__ __launch_bounds__(kThreads) void k_fft_pass(const LagJob* __restrict__ jobs, long long n, long long S_hi, int k,
                                                       int R, int Bn, const double2* __restrict__ tw,
                                                       const double* __restrict__ sums,
                                                       unsigned long long* __restrict__ partials) {
    __shared__ double2 sm[kTile];
    const LagJob job = jobs[blockIdx.y];
    const int G = 1 << k;
    const int GR = G * R;
    const long long q = S_hi >> k;
    const long long per_blk = q / R;
    const long long wg = blockIdx.x;
    const long long blk0 = (wg / per_blk) * Bn;
    const long long r0 = (wg - (wg / per_blk) * per_blk) * R;
    const long long base = blk0 * S_hi + r0;
    double2* z = job.z;
    for (int e = threadIdx.x; e < GR * Bn; e += kThreads) {
        const int bb = e / GR, rem = e - bb * GR;
        const int m = rem / R, rr = rem - m * R;
        const long long i = base + (long long)bb * S_hi + (long long)m * q + rr;
        sm[e] = SRC == kSrcEnvelope ? packed_env(job, sums, i) : z[i];
    }
    // Butterfly u of this thread at level li: (i0, i1) in LDS and its twiddle.
    const int half = GR / 2;
    auto butterfly = [&](int li, int u, int& i0, int& i1) -> long long {
        const int l = DIT ? k - 1 - li : li;
        const int hm = G >> (l + 1);
        const int t = threadIdx.x + u * kThreads;
        const int bb = t / half, rem = t - bb * half;
        const int rr = rem % R, bm = rem / R;
        const int m0 = (bm / hm) * 2 * hm + (bm % hm);
        i0 = bb * GR + m0 * R + rr;
        i1 = i0 + hm * R;
        const long long j = (long long)(bm % hm) * q + r0 + rr;
        return j * (n / (S_hi >> l));
    };
    const int nbf = half * Bn;  // == kTile / 2 whenever the tile is full
    double2 w[kBfPerThread];
    {
        int a, b;
        for (int u = 0; u < kBfPerThread; u++)
            if (threadIdx.x + u * kThreads < nbf) w[u] = tw[butterfly(0, u, a, b)];
    }
    __syncthreads();
    for (int li = 0; li < k; li++) {
        double2 wn[kBfPerThread];
        if (li + 1 < k) {  // prefetch the next level's twiddles
            int a, b;
            for (int u = 0; u < kBfPerThread; u++)
                if (threadIdx.x + u * kThreads < nbf) wn[u] = tw[butterfly(li + 1, u, a, b)];
        }
        for (int u = 0; u < kBfPerThread; u++) {
            if (threadIdx.x + u * kThreads >= nbf) continue;
            int i0, i1;
            butterfly(li, u, i0, i1);
            const double2 a = sm[i0], b = sm[i1];
            if (DIT) {
                const double2 bw = cmul_conj(b, w[u]);
                sm[i0] = cadd(a, bw);
                sm[i1] = csub(a, bw);
            } else {
                sm[i0] = cadd(a, b);
                sm[i1] = cmul(csub(a, b), w[u]);
            }
        }
        __syncthreads();
        if (li + 1 < k)
            for (int u = 0; u < kBfPerThread; u++) w[u] = wn[u];
    }
    if (!ARGMAX) {
        for (int e = threadIdx.x; e < GR * Bn; e += kThreads) {
            const int bb = e / GR, rem = e - bb * GR;
            const int m = rem / R, rr = rem - m * R;
            z[base + (long long)bb * S_hi + (long long)m * q + rr] = sm[e];
        }
        return;
    }
    // First maximum of |Re| in scan order within this workgroup.
    const long long big = job.la > job.lb ? job.la : job.lb;
    const long long max_lag = big - 1;
    unsigned long long best_v = 0, best_o = ~0ULL;
    for (int e = threadIdx.x; e < GR * Bn; e += kThreads) {
        const int bb = e / GR, rem = e - bb * GR;
        const int m = rem / R, rr = rem - m * R;
        const long long i = base + (long long)bb * S_hi + (long long)m * q + rr;
        const long long o = scan_order(i, n, max_lag);
        if (o < 0) continue;
        const unsigned long long v = f64_bits(fabs(sm[e].x));
        if (v > best_v || (v == best_v && (unsigned long long)o < best_o)) {
            best_v = v;
            best_o = (unsigned long long)o;
        }
    }
    __shared__ unsigned long long rv[kThreads], ro[kThreads];
    rv[threadIdx.x] = best_v;
    ro[threadIdx.x] = best_o;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            const unsigned long long v2 = rv[threadIdx.x + s], o2 = ro[threadIdx.x + s];
            if (v2 > rv[threadIdx.x] || (v2 == rv[threadIdx.x] && o2 < ro[threadIdx.x])) {
                rv[threadIdx.x] = v2;
                ro[threadIdx.x] = o2;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[2 * ((long long)blockIdx.y * gridDim.x + blockIdx.x)] = rv[0];
        partials[2 * ((long long)blockIdx.y * gridDim.x + blockIdx.x) + 1] = ro[0];
    }
}

// xcorr.zig:211-229 cross terms, in place in bit-reversed order.
__device__ inline double2 cross_term(double2 x, double2 zy) {
    const double2 y = make_double2(zy.x, -zy.y);
    const double2 sum = make_double2(x.x + y.x, x.y + y.y);
    const double2 dif_c = make_double2(x.x - y.x, -(x.y - y.y));
    const double2 idc = make_double2(-dif_c.y, dif_c.x);
    return make_double2(0.25 * (sum.x * idc.x - sum.y * idc.y), 0.25 * (sum.x * idc.y + sum.y * idc.x));
}

__global__ __launch_bounds__(256) void k_cross_terms(const LagJob* __restrict__ jobs, long long n) {
    double2* z = jobs[blockIdx.y].z;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (i < 2) {
        const double2 v = z[i];
        z[i] = cross_term(v, v);
        return;
    }
    const long long h = 1LL << (63 - __builtin_clzll((unsigned long long)i));
    const long long p = 3 * h - 1 - i;
    if (i < p) {
        const double2 za = z[i], zb = z[p];
        z[i] = cross_term(za, zb);
        z[p] = cross_term(zb, za);
    }
}

// Exact reduction of the per-workgroup (max, first position) partials.
__global__ __launch_bounds__(256) void k_lag_reduce(const LagJob* __restrict__ jobs, const unsigned long long* partials,
                                                    int nparts, long long* __restrict__ lags) {
    const LagJob j = jobs[blockIdx.x];
    const long long big = j.la > j.lb ? j.la : j.lb;
    unsigned long long bv = 0, bo = ~0ULL;
    for (int p = threadIdx.x; p < nparts; p += blockDim.x) {
        const unsigned long long v = partials[2 * ((long long)blockIdx.x * nparts + p)];
        const unsigned long long o = partials[2 * ((long long)blockIdx.x * nparts + p) + 1];
        if (v > bv || (v == bv && o < bo)) {
            bv = v;
            bo = o;
        }
    }
    __shared__ unsigned long long rv[256], ro[256];
    rv[threadIdx.x] = bv;
    ro[threadIdx.x] = bo;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            const unsigned long long v2 = rv[threadIdx.x + s], o2 = ro[threadIdx.x + s];
            if (v2 > rv[threadIdx.x] || (v2 == rv[threadIdx.x] && o2 < ro[threadIdx.x])) {
                rv[threadIdx.x] = v2;
                ro[threadIdx.x] = o2;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) lags[blockIdx.x] = big <= 1 ? 0 : (long long)ro[0] - (big - 1);
}

namespace {
constexpr int kMaxFuse = 8;

// Stage counts per pass, as even as possible.
int plan_passes(long long n, int* ks) {
    int L = 0;
    while ((1LL << L) < n) L++;
    const int np = (L + kMaxFuse - 1) / kMaxFuse;
    for (int p = 0; p < np; p++) ks[p] = L / np + (p < L % np ? 1 : 0);
    return np;
}

struct PassShape {
    int R, Bn;
    long long wgs;
};

PassShape pass_shape(long long n, long long S_hi, int k) {
    const long long G = 1LL << k, q = S_hi >> k;
    const long long want = kTile / G;
    long long R = want < q ? want : q;
    long long Bn = want / R;
    if (Bn > n / S_hi) Bn = n / S_hi;
    if (Bn < 1) Bn = 1;
    return PassShape{(int)R, (int)Bn, (n / S_hi / Bn) * (q / R)};
}

template <bool DIT, int SRC, bool ARGMAX>
void pass(const LagJob* d_jobs, int njobs, long long n, long long S_hi, int k, const double2* tw, const double* sums,
          unsigned long long* partials, hipStream_t stream) {
    const PassShape s = pass_shape(n, S_hi, k);
    hipLaunchKernelGGL((k_fft_pass<DIT, SRC, ARGMAX>), dim3((unsigned)s.wgs, njobs), dim3(kThreads), 0, stream, d_jobs, n,
                       S_hi, k, s.R, s.Bn, tw, sums, partials);
}
}  // namespace

long long lag_partials_per_job(long long n) {
    if (n < 2) return 1;
    int ks[8];
    plan_passes(n, ks);
    return pass_shape(n, n, ks[0]).wgs;  // the last DIT pass covers the spans up to n
}

void launch_lag_search(const LagJob* d_jobs, int njobs, long long n, const double2* d_tw, const double* d_sums,
                       unsigned long long* d_partials, long long* d_lags, hipStream_t stream) {
    if (njobs == 0) return;
    if (n < 2) {
        // n == 1: one sample per side, lag 0.
        hipLaunchKernelGGL(k_lag_reduce, dim3(njobs), dim3(256), 0, stream, d_jobs, (const unsigned long long*)d_partials,
                           0, d_lags);
        return;
    }
    int ks[8];
    const int np = plan_passes(n, ks);
    long long S = n;
    for (int p = 0; p < np; p++) {
        if (p == 0) pass<false, kSrcEnvelope, false>(d_jobs, njobs, n, S, ks[p], d_tw, d_sums, nullptr, stream);
        else pass<false, kSrcZ, false>(d_jobs, njobs, n, S, ks[p], d_tw, d_sums, nullptr, stream);
        S >>= ks[p];
    }
    hipLaunchKernelGGL(k_cross_terms, dim3((unsigned)((n + 255) / 256), njobs), dim3(256), 0, stream, d_jobs, n);
    S = 1;
    for (int p = np - 1; p >= 0; p--) {
        S <<= ks[p];
        if (S == n) pass<true, kSrcZ, true>(d_jobs, njobs, n, S, ks[p], d_tw, d_sums, d_partials, stream);
        else pass<true, kSrcZ, false>(d_jobs, njobs, n, S, ks[p], d_tw, d_sums, nullptr, stream);
    }
    const long long parts = pass_shape(n, n, ks[0]).wgs;
    hipLaunchKernelGGL(k_lag_reduce, dim3(njobs), dim3(256), 0, stream, d_jobs, (const unsigned long long*)d_partials,
                       (int)parts, d_lags);
}

void launch_fft_dif(const LagJob* d_jobs, int njobs, long long n, const double2* d_tw, hipStream_t stream) {
    if (njobs == 0 || n < 2) return;
    int ks[8];
    const int np = plan_passes(n, ks);
    long long S = n;
    for (int p = 0; p < np; p++) {
        pass<false, kSrcZ, false>(d_jobs, njobs, n, S, ks[p], d_tw, nullptr, nullptr, stream);
        S >>= ks[p];
    }
}

void launch_fft_dit(const LagJob* d_jobs, int njobs, long long n, const double2* d_tw, hipStream_t stream) {
    if (njobs == 0 || n < 2) return;
    int ks[8];
    const int np = plan_passes(n, ks);
    long long S = 1;
    for (int p = np - 1; p >= 0; p--) {
        S <<= ks[p];
        pass<true, kSrcZ, false>(d_jobs, njobs, n, S, ks[p], d_tw, nullptr, nullptr, stream);
    }
}

}  // namespace vq
