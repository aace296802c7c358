// Signal-level kernels: reductions (serial reference order or tree), upper
// envelope + packing, radix-2 complex FFT with LDS-resident inner stages,
// the bit-reversed cross-spectrum, the first-max lag scan, lag/SPL
// application and patch slicing.
//
// References: envelope.zig:23-53, xcorr.zig:180-251, dsp/fft.zig:45-166,
// realign.zig:21-51 and 91-110, audio.zig:132-159.
#include "dev_common.h"
#include "kernels.h"

namespace vq {

// ---------------------------------------------------------------- reductions
__device__ inline double sum_term(const SumJob& j, long long i, double mean) {
    const double x = j.x[i];
    switch (j.kind) {
        case kSum: return x;
        case kNyquist: {
            const double d = x - mean;
            return (i & 1) ? -d : d;  // sign * (x - mean), sign = ±1 exactly
        }
        default: return x * x;
    }
}

constexpr int kSerialTile = 2048;  // terms per LDS tile (16 KiB)

// One block per job.  serial: the terms are added in sample order by one
// lane (the reference's rounding sequence).  Each term (x, x·x or
// ±(x - mean)) rounds on its own, so all lanes compute terms for the next
// tile into a second LDS buffer while lane 0 runs the add chain over the
// current one.  tree: fixed-shape block reduction.
__global__ __launch_bounds__(256) void k_sums(const SumJob* __restrict__ jobs, double* __restrict__ out, int serial) {
    const SumJob j = jobs[blockIdx.x];
    long long start = 0;
    if (j.kind == kSumSqFrom) {
        const long long lag = *j.lag;
        const long long a = lag < 0 ? -lag : lag;
        if (lag < 0 && a <= j.half) start = a < j.len ? a : j.len;
    }
    double mean = 0.0;
    if (j.kind == kNyquist) mean = out[j.aux] / (double)j.len;
    if (serial) {
        __shared__ double tile[2][kSerialTile];
        constexpr int per = kSerialTile / 256;
        const long long total = j.len - start;
        const long long ntiles = (total + kSerialTile - 1) / kSerialTile;
        double s = 0.0;
        double v[per];
        // prologue: tile 0
        for (int q = 0; q < per; q++) {
            const long long i = start + (long long)q * 256 + threadIdx.x;
            v[q] = i < j.len ? sum_term(j, i, mean) : 0.0;
        }
        for (long long t = 0; t < ntiles; t++) {
            const int cur = (int)(t & 1);
            for (int q = 0; q < per; q++) tile[cur][q * 256 + threadIdx.x] = v[q];
            __syncthreads();
            // issue the next tile's loads before the serial chain
            const long long base = start + (t + 1) * kSerialTile;
            for (int q = 0; q < per; q++) {
                const long long i = base + (long long)q * 256 + threadIdx.x;
                v[q] = i < j.len ? sum_term(j, i, mean) : 0.0;
            }
            if (threadIdx.x == 0) {
                const long long rem = total - t * kSerialTile;
                const int cnt = rem < kSerialTile ? (int)rem : kSerialTile;
                const double* b = tile[cur];
                int k = 0;
                for (; k + 8 <= cnt; k += 8) {
                    const double a0 = b[k], a1 = b[k + 1], a2 = b[k + 2], a3 = b[k + 3];
                    const double a4 = b[k + 4], a5 = b[k + 5], a6 = b[k + 6], a7 = b[k + 7];
                    s += a0; s += a1; s += a2; s += a3; s += a4; s += a5; s += a6; s += a7;
                }
                for (; k < cnt; k++) s += b[k];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) out[j.slot] = s;
        return;
    }
    __shared__ double part[256];
    double s = 0.0;
    for (long long i = start + threadIdx.x; i < j.len; i += blockDim.x) s += sum_term(j, i, mean);
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[j.slot] = part[0];
}

void launch_sums(const SumJob* d_jobs, int njobs, double* d_out, bool serial, hipStream_t stream) {
    if (njobs == 0) return;
    hipLaunchKernelGGL(k_sums, dim3(njobs), dim3(256), 0, stream, d_jobs, d_out, serial ? 1 : 0);
}

// ---------------------------------------------------------------- envelope
__device__ inline long long env_fft_size(long long l) {
    long long n = 1;
    while (n < l) n <<= 1;
    return n < 32 ? 32 : n;
}

// envelope.zig:23-53 for both signals of a lag job, packed as z = a + i·b.
__global__ __launch_bounds__(256) void k_env_pack(const LagJob* __restrict__ jobs, const double* __restrict__ sums) {
    const LagJob j = jobs[blockIdx.y];
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= j.n) return;
    const double eps = 0.000001;
    double re = 0.0, im = 0.0;
    if (i < j.la) {
        const double mean = sums[j.sum_a] / (double)j.la;
        const double nyq = sums[j.nyq_a] / (double)env_fft_size(j.la);
        const double term = (i & 1) ? -nyq : nyq;
        re = fabs(2.0 * ((j.a[i] - mean) - term) - eps) + mean;
    }
    if (i < j.lb) {
        const double mean = sums[j.sum_b] / (double)j.lb;
        const double nyq = sums[j.nyq_b] / (double)env_fft_size(j.lb);
        const double term = (i & 1) ? -nyq : nyq;
        im = fabs(2.0 * ((j.b[i] - mean) - term) - eps) + mean;
    }
    j.z[i] = make_double2(re, im);
}

void launch_env_pack(const LagJob* d_jobs, int njobs, long long max_n, const double* d_sums, hipStream_t stream) {
    if (njobs == 0) return;
    dim3 grid((unsigned)((max_n + 255) / 256), njobs);
    hipLaunchKernelGGL(k_env_pack, grid, dim3(256), 0, stream, d_jobs, d_sums);
}

// ---------------------------------------------------------------- FFT
// Butterflies with the reference's operation order (fft.zig:181-188, 314-323).
__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ inline double2 cmul(double2 x, double2 w) {
    return make_double2(x.x * w.x - x.y * w.y, x.x * w.y + x.y * w.x);
}
__device__ inline double2 cmul_conj(double2 x, double2 w) {
    const double wi = -w.y;
    return make_double2(x.x * w.x - x.y * wi, x.x * wi + x.y * w.x);
}

constexpr int kFftBlock = 4096;  // complex doubles per LDS block (64 KiB)

// One span-S stage over global memory: S/2 butterflies per span.
template <bool DIT>
__global__ __launch_bounds__(256) void k_fft_stage(double2* const* __restrict__ rows, long long n, long long S,
                                                   const double2* __restrict__ tw) {
    double2* z = rows[blockIdx.y];
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n / 2) return;
    const long long half = S / 2;
    const long long blk = t / half, j = t - blk * half;
    const long long i0 = blk * S + j, i1 = i0 + half;
    const double2 w = tw[j * (n / S)];
    const double2 a = z[i0], b = z[i1];
    if (DIT) {
        const double2 bw = cmul_conj(b, w);
        z[i0] = cadd(a, bw);
        z[i1] = csub(a, bw);
    } else {
        z[i0] = cadd(a, b);
        z[i1] = cmul(csub(a, b), w);
    }
}

// All stages with span <= B inside one B-element block held in LDS.
template <bool DIT>
__global__ __launch_bounds__(256) void k_fft_block(double2* const* __restrict__ rows, long long n, int B,
                                                   const double2* __restrict__ tw) {
    __shared__ double2 s[kFftBlock];
    double2* z = rows[blockIdx.y] + (long long)blockIdx.x * B;
    for (int i = threadIdx.x; i < B; i += blockDim.x) s[i] = z[i];
    __syncthreads();
    if (DIT) {
        for (int S = 2; S <= B; S <<= 1) {
            const int half = S / 2;
            const long long stride = n / S;
            for (int t = threadIdx.x; t < B / 2; t += blockDim.x) {
                const int blk = t / half, j = t - blk * half;
                const int i0 = blk * S + j, i1 = i0 + half;
                const double2 a = s[i0];
                const double2 bw = cmul_conj(s[i1], tw[j * stride]);
                s[i0] = cadd(a, bw);
                s[i1] = csub(a, bw);
            }
            __syncthreads();
        }
    } else {
        for (int S = B; S >= 2; S >>= 1) {
            const int half = S / 2;
            const long long stride = n / S;
            for (int t = threadIdx.x; t < B / 2; t += blockDim.x) {
                const int blk = t / half, j = t - blk * half;
                const int i0 = blk * S + j, i1 = i0 + half;
                const double2 a = s[i0], b = s[i1];
                s[i0] = cadd(a, b);
                s[i1] = cmul(csub(a, b), tw[j * stride]);
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < B; i += blockDim.x) z[i] = s[i];
}

// k consecutive radix-2 stages in one HBM round trip.  The stages with spans
// S_hi, S_hi/2, ..., S_hi/2^(k-1) only mix elements i = blk·S_hi + r + m·q
// (q = S_hi/2^k, m < 2^k) of one residue r, so a workgroup loads R
// consecutive residues (R·16 B contiguous per m) of one span block into LDS,
// runs the k stages on them and writes them back.  Every butterfly is the
// radix-2 butterfly of the stage it belongs to, with the same twiddle, so
// the result equals the stage-by-stage transform bit for bit.
// A workgroup holds Bn consecutive span blocks × R residues × 2^k elements
// (Bn > 1 only when a block has fewer than R residues, so short spans still
// fill a workgroup with contiguous data).
template <bool DIT>
__global__ __launch_bounds__(256) void k_fft_pass(double2* const* __restrict__ rows, long long n, long long S_hi, int k,
                                                  int R, int Bn, const double2* __restrict__ tw) {
    extern __shared__ double2 sm[];
    const int G = 1 << k;
    const int GR = G * R;
    const long long q = S_hi >> k;
    const long long per_blk = q / R;
    const long long wg = blockIdx.x;
    const long long blk0 = (wg / per_blk) * Bn;
    const long long r0 = (wg - (wg / per_blk) * per_blk) * R;
    double2* z = rows[blockIdx.y] + blk0 * S_hi + r0;
    for (int e = threadIdx.x; e < GR * Bn; e += blockDim.x) {
        const int bb = e / GR, rem = e - bb * GR;
        const int m = rem / R, rr = rem - m * R;
        sm[e] = z[(long long)bb * S_hi + (long long)m * q + rr];
    }
    __syncthreads();
    for (int li = 0; li < k; li++) {
        const int l = DIT ? k - 1 - li : li;
        const int hm = G >> (l + 1);
        const long long stride = n / (S_hi >> l);
        const int half = GR / 2;
        for (int t = threadIdx.x; t < half * Bn; t += blockDim.x) {
            const int bb = t / half, rem = t - bb * half;
            const int rr = rem % R, bm = rem / R;
            const int m0 = (bm / hm) * 2 * hm + (bm % hm);
            const int i0 = bb * GR + m0 * R + rr, i1 = i0 + hm * R;
            const long long j = (long long)(bm % hm) * q + r0 + rr;
            const double2 w = tw[j * stride];
            const double2 a = sm[i0], b = sm[i1];
            if (DIT) {
                const double2 bw = cmul_conj(b, w);
                sm[i0] = cadd(a, bw);
                sm[i1] = csub(a, bw);
            } else {
                sm[i0] = cadd(a, b);
                sm[i1] = cmul(csub(a, b), w);
            }
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < GR * Bn; e += blockDim.x) {
        const int bb = e / GR, rem = e - bb * GR;
        const int m = rem / R, rr = rem - m * R;
        z[(long long)bb * S_hi + (long long)m * q + rr] = sm[e];
    }
}

namespace {
constexpr int kMaxFuse = 8;      // stages per pass: 256-element groups
constexpr int kMaxGroupR = 4096;  // G·R complex values per workgroup (64 KiB LDS)

// Stage counts per pass, as even as possible.
void plan_passes(int L, int* ks, int& np) {
    np = (L + kMaxFuse - 1) / kMaxFuse;
    for (int p = 0; p < np; p++) ks[p] = L / np + (p < L % np ? 1 : 0);
}

template <bool DIT>
void launch_pass(double2* const* d_rows, int nrows, long long n, long long S_hi, int k, const double2* d_tw,
                 hipStream_t stream) {
    const long long G = 1LL << k, q = S_hi >> k;
    const long long want = kMaxGroupR / G;  // residues (times blocks) per workgroup
    long long R = want < q ? want : q;
    long long Bn = want / R;
    if (Bn > n / S_hi) Bn = n / S_hi;
    if (Bn < 1) Bn = 1;
    const long long wgs = (n / S_hi / Bn) * (q / R);
    hipLaunchKernelGGL(k_fft_pass<DIT>, dim3((unsigned)wgs, nrows), dim3(256), sizeof(double2) * G * R * Bn, stream, d_rows,
                       n, S_hi, k, (int)R, (int)Bn, d_tw);
}
}  // namespace

void launch_fft_dif(double2* const* d_rows, int nrows, long long n, const double2* d_tw, hipStream_t stream) {
    if (nrows == 0 || n < 2) return;
    int L = 0;
    while ((1LL << L) < n) L++;
    int ks[8], np;
    plan_passes(L, ks, np);
    long long S = n;
    for (int p = 0; p < np; p++) {
        launch_pass<false>(d_rows, nrows, n, S, ks[p], d_tw, stream);
        S >>= ks[p];
    }
}

void launch_fft_dit(double2* const* d_rows, int nrows, long long n, const double2* d_tw, hipStream_t stream) {
    if (nrows == 0 || n < 2) return;
    int L = 0;
    while ((1LL << L) < n) L++;
    int ks[8], np;
    plan_passes(L, ks, np);
    long long S = 1;
    for (int p = np - 1; p >= 0; p--) {
        S <<= ks[p];
        launch_pass<true>(d_rows, nrows, n, S, ks[p], d_tw, stream);
    }
}

// xcorr.zig:182-192: with X = Z[k], Y = conj(Z[-k]): (X+Y)·i·conj(X-Y)/4.
__device__ inline double2 cross_term(double2 x, double2 zy) {
    const double2 y = make_double2(zy.x, -zy.y);
    const double2 sum = make_double2(x.x + y.x, x.y + y.y);
    const double2 dif_c = make_double2(x.x - y.x, -(x.y - y.y));
    const double2 idc = make_double2(-dif_c.y, dif_c.x);
    return make_double2(0.25 * (sum.x * idc.x - sum.y * idc.y), 0.25 * (sum.x * idc.y + sum.y * idc.x));
}

// In DIF bit-reversed order Z[-k] of position i (i >= 2) sits at 3h-1-i,
// h the largest power of two <= i; positions 0 and 1 pair with themselves
// (xcorr.zig:211-229).
__global__ __launch_bounds__(256) void k_cross_terms(double2* const* __restrict__ rows, long long n) {
    double2* z = rows[blockIdx.y];
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

void launch_cross_terms(double2* const* d_rows, int nrows, long long n, hipStream_t stream) {
    if (nrows == 0) return;
    hipLaunchKernelGGL(k_cross_terms, dim3((unsigned)((n + 255) / 256), nrows), dim3(256), 0, stream, d_rows, n);
}

// ---------------------------------------------------------------- lag scan
// Scan position o in [0, 2·max_lag] maps to circular index
// o < max_lag ? n - max_lag + o : o - max_lag, and to lag o - max_lag.
__device__ inline long long scan_index(long long o, long long n, long long max_lag) {
    return o < max_lag ? n - max_lag + o : o - max_lag;
}

__global__ void k_lag_init(unsigned long long* scratch, int njobs) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < njobs) {
        scratch[2 * j] = 0ULL;
        scratch[2 * j + 1] = ~0ULL;
    }
}

// Pass 1: max |Re z| over the scan range (non-negative doubles compare as
// their bit patterns).
__global__ __launch_bounds__(256) void k_lag_max(const LagJob* __restrict__ jobs, unsigned long long* scratch) {
    const LagJob j = jobs[blockIdx.y];
    const long long big = j.la > j.lb ? j.la : j.lb;
    if (big == 0) return;
    const long long max_lag = big - 1, total = 2 * max_lag + 1;
    unsigned long long best = 0;
    for (long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x; o < total; o += (long long)gridDim.x * blockDim.x) {
        const unsigned long long b = f64_bits(fabs(j.z[scan_index(o, j.n, max_lag)].x));
        best = b > best ? b : best;
    }
    __shared__ unsigned long long red[256];
    red[threadIdx.x] = best;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = red[threadIdx.x] > red[threadIdx.x + w] ? red[threadIdx.x] : red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(&scratch[2 * blockIdx.y], red[0]);
}

// Pass 2: the first scan position attaining the max.
__global__ __launch_bounds__(256) void k_lag_first(const LagJob* __restrict__ jobs, unsigned long long* scratch) {
    const LagJob j = jobs[blockIdx.y];
    const long long big = j.la > j.lb ? j.la : j.lb;
    if (big == 0) return;
    const long long max_lag = big - 1, total = 2 * max_lag + 1;
    const unsigned long long target = scratch[2 * blockIdx.y];
    unsigned long long first = ~0ULL;
    for (long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x; o < total; o += (long long)gridDim.x * blockDim.x) {
        if (f64_bits(fabs(j.z[scan_index(o, j.n, max_lag)].x)) == target) {
            first = (unsigned long long)o;
            break;  // later o of this thread are larger
        }
    }
    __shared__ unsigned long long red[256];
    red[threadIdx.x] = first;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = red[threadIdx.x] < red[threadIdx.x + w] ? red[threadIdx.x] : red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMin(&scratch[2 * blockIdx.y + 1], red[0]);
}

__global__ void k_lag_final(const LagJob* __restrict__ jobs, const unsigned long long* scratch, long long* lags, int njobs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= njobs) return;
    const LagJob j = jobs[i];
    const long long big = j.la > j.lb ? j.la : j.lb;
    if (big == 0) { lags[i] = 0; return; }
    lags[i] = (long long)scratch[2 * i + 1] - (big - 1);
}

void launch_lag_argmax(const LagJob* d_jobs, int njobs, long long max_n, unsigned long long* d_scratch, long long* d_lags,
                       hipStream_t stream) {
    if (njobs == 0) return;
    const unsigned tb = (unsigned)((njobs + 255) / 256);
    hipLaunchKernelGGL(k_lag_init, dim3(tb), dim3(256), 0, stream, d_scratch, njobs);
    long long blocks = (max_n + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(k_lag_max, dim3((unsigned)blocks, njobs), dim3(256), 0, stream, d_jobs, d_scratch);
    hipLaunchKernelGGL(k_lag_first, dim3((unsigned)blocks, njobs), dim3(256), 0, stream, d_jobs, d_scratch);
    hipLaunchKernelGGL(k_lag_final, dim3(tb), dim3(256), 0, stream, d_jobs, d_scratch, d_lags, njobs);
}

// ---------------------------------------------------------------- signal prep
__global__ __launch_bounds__(256) void k_scale(const ScaleJob* __restrict__ jobs) {
    const ScaleJob j = jobs[blockIdx.y];
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= j.dst_len) return;
    const long long s = i - j.shift;
    const double x = (s >= 0 && s < j.src_len) ? j.src[s] : 0.0;
    j.dst[i] = x * j.factor;
}

void launch_scale(const ScaleJob* d_jobs, int njobs, long long max_len, hipStream_t stream) {
    if (njobs == 0 || max_len == 0) return;
    hipLaunchKernelGGL(k_scale, dim3((unsigned)((max_len + 255) / 256), njobs), dim3(256), 0, stream, d_jobs);
}

__global__ __launch_bounds__(256) void k_slices(const SliceJob* __restrict__ jobs) {
    const SliceJob j = jobs[blockIdx.y];
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= j.total) return;
    j.dst[i] = (i >= j.pad && i < j.pad + j.len) ? j.src[j.base + i - j.pad] : 0.0;
}

void launch_slices(const SliceJob* d_jobs, int njobs, long long max_len, hipStream_t stream) {
    if (njobs == 0 || max_len == 0) return;
    hipLaunchKernelGGL(k_slices, dim3((unsigned)((max_len + 255) / 256), njobs), dim3(256), 0, stream, d_jobs);
}

__global__ __launch_bounds__(256) void k_frame_copies(const FrameCopyJob* __restrict__ jobs, int nb) {
    const FrameCopyJob j = jobs[blockIdx.y];
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nb * j.nf) return;
    const int b = e / j.nf, f = e % j.nf;
    j.dst[e] = j.src[(long long)b * j.src_stride + f];
}

void launch_frame_copies(const FrameCopyJob* d_jobs, int njobs, int nb, int max_nf, hipStream_t stream) {
    if (njobs == 0 || max_nf == 0) return;
    hipLaunchKernelGGL(k_frame_copies, dim3((nb * max_nf + 255) / 256, njobs), dim3(256), 0, stream, d_jobs, nb);
}

// ---------------------------------------------------------------- VAD
__global__ void k_vad_init(unsigned long long* scratch, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) scratch[i] = 0ULL;
}

// max(ref) as the larger orderable key (vad.zig:31-32).
__global__ __launch_bounds__(256) void k_vad_max(const VadJob* __restrict__ jobs, unsigned long long* scratch) {
    const VadJob j = jobs[blockIdx.y];
    unsigned long long best = 0ULL;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < j.len; i += (long long)gridDim.x * blockDim.x) {
        const unsigned long long k = order_key(j.ref[i]);
        best = k > best ? k : best;
    }
    __shared__ unsigned long long red[256];
    red[threadIdx.x] = best;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = red[threadIdx.x] > red[threadIdx.x + w] ? red[threadIdx.x] : red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(&scratch[blockIdx.y], red[0]);
}

// One chunk per lane, its squares summed in sample order (vad.zig:37-48).
__global__ __launch_bounds__(256) void k_vad_chunks(const VadJob* __restrict__ jobs, const unsigned long long* scratch) {
    const VadJob j = jobs[blockIdx.y];
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= j.n_chunks) return;
    const double max_val = key_value(scratch[blockIdx.y]);
    double square = 0.0;
    for (int q = 0; q < j.frame_len; q++) {
        const long long idx = (long long)j.first + (long long)c * j.frame_len + q;
        const double x = idx < j.len ? j.ref[idx] / max_val : 0.0;
        const double scaled = fmax(-32768.0, fmin(x * 32768.0, 32767.0));
        const double s = trunc(scaled);
        square += s * s;
    }
    const double rms = sqrt(square / (double)j.frame_len);
    j.active[c] = rms >= 5000.0 ? 1 : 0;
}

void launch_vad(const VadJob* d_jobs, int njobs, int max_chunks, unsigned long long* d_scratch, hipStream_t stream) {
    if (njobs == 0) return;
    hipLaunchKernelGGL(k_vad_init, dim3((njobs + 255) / 256), dim3(256), 0, stream, d_scratch, njobs);
    hipLaunchKernelGGL(k_vad_max, dim3(64, njobs), dim3(256), 0, stream, d_jobs, d_scratch);
    if (max_chunks > 0)
        hipLaunchKernelGGL(k_vad_chunks, dim3((max_chunks + 255) / 256, njobs), dim3(256), 0, stream, d_jobs, d_scratch);
}

// ---------------------------------------------------------------- math probe
__global__ void k_math_probe(int op, const double* a, const double* b, double* out, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double r;
    switch (op) {
        case 0: r = fdlibm_log10(a[i]); break;
        case 1: r = sqrt(a[i]); break;
        default: r = a[i] / b[i]; break;
    }
    out[i] = r;
}

void launch_math_probe(int op, const double* a, const double* b, double* out, long long n, hipStream_t stream) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_math_probe, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, op, a, b, out, n);
}

}  // namespace vq
