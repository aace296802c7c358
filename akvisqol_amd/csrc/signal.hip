// Signal-level kernels: reductions (serial reference order or tree), lag /
// SPL application, patch slicing, frame copies, VAD chunk activity and the
// device arithmetic probe.
//
// References: envelope.zig:23-53 (sums), realign.zig:21-51 and 91-110,
// audio.zig:132-159, vad.zig:31-48.
#include "dev_common.h"
#include "kernels.h"

namespace vq {

// ---------------------------------------------------------------- reductions
__device__ inline double sum_term(const SumJob& j, long long i, double mean) {
    const double x = as_global(j.x)[i];
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

// Shares of a tree reduction over a signal of `len` samples: about 8 K
// terms each, at most kSumParts.
__host__ __device__ inline int sum_parts_for(long long len) {
    const long long p = (len + 8191) / 8192;
    return p < 1 ? 1 : (p > kSumParts ? kSumParts : (int)p);
}

// serial: one block per job; the terms are added in sample order by one
// lane (the reference's rounding sequence).  Each term (x, x·x or
// ±(x - mean)) rounds on its own, so all lanes compute terms for the next
// tile into a second LDS buffer while lane 0 runs the add chain over the
// current one.  tree: kSumParts blocks per job, each a fixed-shape block
// reduction of one contiguous share, then k_sums_finish adds the shares.
__global__ __launch_bounds__(256) void k_sums(const SumJob* __restrict__ jobs, double* __restrict__ out,
                                              double* __restrict__ partials, int serial, int parts) {
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
    // tree: this block's share of [start, len) into
    // partials[(job·kSumOutputs + output)·kSumParts + blockIdx.y].  The
    // share count depends on the job's own length only, so a signal's sums
    // do not depend on the batch it is in; blocks past it write zeros.
    const long long total = j.len - start;
    const int jparts = sum_parts_for(j.len);
    if ((int)blockIdx.y >= jparts) {
        if (threadIdx.x < kSumOutputs)
            partials[((long long)blockIdx.x * kSumOutputs + threadIdx.x) * kSumParts + blockIdx.y] = 0.0;
        return;
    }
    const long long per = (total + jparts - 1) / jparts;
    const long long lo = start + per * blockIdx.y;
    const long long hi = lo + per < j.len ? lo + per : j.len;
    __shared__ double part[kSumOutputs][256];
    // Each thread adds its terms lo + t, lo + t + 256, ... in order; eight
    // loads are issued before they are added, which keeps more loads in
    // flight without changing any sum's order.
    constexpr int kUnroll = 8;
    const long long stride = blockDim.x;
    double s = 0.0, alt = 0.0, sq = 0.0;
    long long i = lo + threadIdx.x;
    if (j.kind == kMoments) {
        const auto* x = as_global(j.x);
        for (; i + (kUnroll - 1) * stride < hi; i += kUnroll * stride) {
            double v[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; u++) v[u] = x[i + u * stride];
#pragma unroll
            for (int u = 0; u < kUnroll; u++) {
                s += v[u];
                alt += ((i + u * stride) & 1) ? -v[u] : v[u];
                sq = __builtin_fma(v[u], v[u], sq);
            }
        }
        for (; i < hi; i += stride) {
            const double v = x[i];
            s += v;
            alt += (i & 1) ? -v : v;
            sq = __builtin_fma(v, v, sq);
        }
    } else {
        // one loop per kind, each term as sum_term forms it
        const auto* x = as_global(j.x);
        auto run = [&](auto term) {
            for (; i + (kUnroll - 1) * stride < hi; i += kUnroll * stride) {
                double v[kUnroll];
#pragma unroll
                for (int u = 0; u < kUnroll; u++) v[u] = x[i + u * stride];
#pragma unroll
                for (int u = 0; u < kUnroll; u++) s += term(v[u], i + u * stride);
            }
            for (; i < hi; i += stride) s += term(x[i], i);
        };
        if (j.kind == kSum) {
            run([](double v, long long) { return v; });
        } else if (j.kind == kNyquist) {
            run([mean](double v, long long k) {
                const double d = v - mean;
                return (k & 1) ? -d : d;
            });
        } else {
            run([](double v, long long) { return v * v; });
        }
    }
    part[0][threadIdx.x] = s;
    part[1][threadIdx.x] = alt;
    part[2][threadIdx.x] = sq;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int o = 0; o < kSumOutputs; o++) part[o][threadIdx.x] += part[o][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < kSumOutputs)
        partials[((long long)blockIdx.x * kSumOutputs + threadIdx.x) * kSumParts + blockIdx.y] = part[threadIdx.x][0];
}

// Fixed-order sums of a job's partials (one wave per job), written to the
// job's output slots.
__global__ __launch_bounds__(64) void k_sums_finish(const SumJob* __restrict__ jobs, const double* __restrict__ partials,
                                                    double* __restrict__ out, int parts) {
    const SumJob j = jobs[blockIdx.x];
    double r[kSumOutputs];
    for (int o = 0; o < kSumOutputs; o++) {
        double v = (int)threadIdx.x < parts ? partials[((long long)blockIdx.x * kSumOutputs + o) * kSumParts + threadIdx.x]
                                            : 0.0;
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        r[o] = v;
    }
    if (threadIdx.x != 0) return;
    out[j.slot] = r[0];
    if (j.kind == kMoments) {
        const double mean = j.len > 0 ? r[0] / (double)j.len : 0.0;
        out[j.aux] = (j.len & 1) ? r[1] - mean : r[1];
        if (j.slot2 >= 0) out[j.slot2] = r[2];
    }
}

void launch_sums(const SumJob* d_jobs, int njobs, long long max_len, double* d_out, bool serial, double* d_partials,
                 hipStream_t stream) {
    if (njobs == 0) return;
    if (serial) {
        hipLaunchKernelGGL(k_sums, dim3(njobs, 1), dim3(256), 0, stream, d_jobs, d_out, d_partials, 1, 1);
        return;
    }
    const int parts = sum_parts_for(max_len);  // the longest job's share count
    hipLaunchKernelGGL(k_sums, dim3(njobs, parts), dim3(256), 0, stream, d_jobs, d_out, d_partials, 0, parts);
    hipLaunchKernelGGL(k_sums_finish, dim3(njobs), dim3(64), 0, stream, d_jobs, (const double*)d_partials, d_out, parts);
}

// ---------------------------------------------------------------- signal prep
// Elementwise copies: each block covers kCopyPer·256 consecutive elements
// (each thread kCopyPer of them, 256 apart, loads first).
constexpr int kCopyPer = 4;
constexpr int kCopyBlock = 256 * kCopyPer;

__global__ __launch_bounds__(256) void k_scale(const ScaleJob* __restrict__ jobs) {
    const ScaleJob j = jobs[blockIdx.y];
    const long long i0 = (long long)blockIdx.x * kCopyBlock + threadIdx.x;
    if (i0 >= j.dst_len) return;
    const auto* src = as_global(j.src);
    double x[kCopyPer];
#pragma unroll
    for (int u = 0; u < kCopyPer; u++) {
        const long long s = i0 + u * 256 - j.shift;
        x[u] = (s >= 0 && s < j.src_len) ? src[s] : 0.0;
    }
    auto* dst = as_global(j.dst);
#pragma unroll
    for (int u = 0; u < kCopyPer; u++)
        if (i0 + u * 256 < j.dst_len) dst[i0 + u * 256] = x[u] * j.factor;
}

void launch_scale(const ScaleJob* d_jobs, int njobs, long long max_len, hipStream_t stream) {
    if (njobs == 0 || max_len == 0) return;
    hipLaunchKernelGGL(k_scale, dim3((unsigned)((max_len + kCopyBlock - 1) / kCopyBlock), njobs), dim3(256), 0, stream,
                       d_jobs);
}

__global__ __launch_bounds__(256) void k_slices(const SliceJob* __restrict__ jobs) {
    const SliceJob j = jobs[blockIdx.y];
    const long long i0 = (long long)blockIdx.x * kCopyBlock + threadIdx.x;
    if (i0 >= j.total) return;
    const auto* src = as_global(j.src);
    double x[kCopyPer];
#pragma unroll
    for (int u = 0; u < kCopyPer; u++) {
        const long long i = i0 + u * 256;
        x[u] = (i >= j.pad && i < j.pad + j.len) ? src[j.base + i - j.pad] : 0.0;
    }
    auto* dst = as_global(j.dst);
#pragma unroll
    for (int u = 0; u < kCopyPer; u++)
        if (i0 + u * 256 < j.total) dst[i0 + u * 256] = x[u];
}

void launch_slices(const SliceJob* d_jobs, int njobs, long long max_len, hipStream_t stream) {
    if (njobs == 0 || max_len == 0) return;
    hipLaunchKernelGGL(k_slices, dim3((unsigned)((max_len + kCopyBlock - 1) / kCopyBlock), njobs), dim3(256), 0, stream,
                       d_jobs);
}

__global__ __launch_bounds__(256) void k_frame_copies(const FrameCopyJob* __restrict__ jobs, int nb) {
    const FrameCopyJob j = jobs[blockIdx.y];
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nb * j.nf) return;
    const int b = e / j.nf, f = e % j.nf;
    as_global(j.dst)[e] = as_global(j.src)[(long long)b * j.src_stride + f];
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
