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
// Term i of a sum of the given kind over sample value x: x, x·x, or
// ±(x - mean) for the alternating (Nyquist) sum.
__device__ inline double sum_term(int kind, double x, long long i, double mean) {
    switch (kind) {
        case kSum: return x;
        case kNyquist: {
            const double d = x - mean;
            return (i & 1) ? -d : d;  // sign * (x - mean), sign = ±1 exactly
        }
        default: return x * x;
    }
}

// s + t as one fused multiply-add, t·1 + s: the product is exact, so the
// single rounding gives the bits of s + t (signed zeros, infinities and
// NaNs included).  A chain of v_fmac_f64 measured about 10 cycles a step
// on gfx950 against 12.5 for v_add_f64 (one wave, register operands).
// `one` must be opaque to the optimiser (one_opaque), or the fma folds
// back to an add.
__device__ inline double one_opaque() {
    double one = 1.0;
    asm volatile("" : "+v"(one));
    return one;
}
__device__ inline double acc_add(double s, double t, double one) { return __builtin_fma(t, one, s); }

constexpr int kSerialTile = 2048;  // terms per LDS tile (16 KiB)

// Shares of a tree reduction over a signal of `len` samples: about 8 K
// terms each, at most kSumParts.
__host__ __device__ inline int sum_parts_for(long long len) {
    const long long p = (len + 8191) / 8192;
    return p < 1 ? 1 : (p > kSumParts ? kSumParts : (int)p);
}

// Start of a job's terms: kSumSqFrom begins at the lag's sample when the
// lag is negative and within half the reference (realign.zig's aligned
// degraded signal), every other kind at 0.
__device__ inline long long sum_start(const SumJob& j) {
    if (j.kind != kSumSqFrom) return 0;
    const long long lag = *j.lag;
    const long long a = lag < 0 ? -lag : lag;
    return (lag < 0 && a <= j.half) ? (a < j.len ? a : j.len) : 0;
}
__device__ inline double sum_mean(const SumJob& j, const double* out) {
    return j.kind == kNyquist ? out[j.aux] / (double)j.len : 0.0;
}

// The terms [lo, hi) of job j added in sample order onto s0 by lane 0 of
// the block (the result is lane 0's).  Each term (x, x·x or ±(x - mean))
// rounds on its own, so all lanes stage the next tile of terms into the
// other half of `tile` while lane 0 runs the add chain over the current
// one.  The samples of tile t are loaded raw (from a clamped, always
// valid index, so the loads carry no branch and all issue at once) while
// tile t - 1's chain runs, and their terms are formed only when written to
// LDS: forming them at load time would make every wave, lane 0's included,
// wait for its loads before the chain could start.  Every lane of the
// block must call this (it holds barriers).
__device__ double serial_chain(const SumJob& j, double mean, long long lo, long long hi, double s0,
                               double (*tile)[kSerialTile]) {
    constexpr int per = kSerialTile / 256;
    const long long total = hi - lo;
    const long long ntiles = total > 0 ? (total + kSerialTile - 1) / kSerialTile : 0;
    const auto* x = as_global(j.x);
    const double one = one_opaque();
    double s = s0;
    double v[per];
    auto fetch = [&](long long base) {
        for (int q = 0; q < per; q++) {
            const long long i = base + (long long)q * 256 + threadIdx.x;
            v[q] = x[i < hi ? i : lo];
        }
    };
    auto put = [&](int buf, long long base) {
        for (int q = 0; q < per; q++) {
            const long long i = base + (long long)q * 256 + threadIdx.x;
            tile[buf][q * 256 + threadIdx.x] = i < hi ? sum_term(j.kind, v[q], i, mean) : 0.0;
        }
    };
    if (ntiles > 0) fetch(lo);  // prologue: tile 0 (lo < hi)
    for (long long t = 0; t < ntiles; t++) {
        const int cur = (int)(t & 1);
        put(cur, lo + t * kSerialTile);
        __syncthreads();
        // issue the next tile's loads before the serial chain
        if (t + 1 < ntiles) fetch(lo + (t + 1) * kSerialTile);
        if (threadIdx.x == 0) {
            const long long rem = total - t * kSerialTile;
            const int cnt = rem < kSerialTile ? (int)rem : kSerialTile;
            const double* b = tile[cur];
            int k = 0;
            for (; k + 8 <= cnt; k += 8) {
                const double a0 = b[k], a1 = b[k + 1], a2 = b[k + 2], a3 = b[k + 3];
                const double a4 = b[k + 4], a5 = b[k + 5], a6 = b[k + 6], a7 = b[k + 7];
                s = acc_add(s, a0, one); s = acc_add(s, a1, one); s = acc_add(s, a2, one);
                s = acc_add(s, a3, one); s = acc_add(s, a4, one); s = acc_add(s, a5, one);
                s = acc_add(s, a6, one); s = acc_add(s, a7, one);
            }
            for (; k < cnt; k++) s = acc_add(s, b[k], one);
        }
        __syncthreads();
    }
    return s;
}

// serial: one block per job, serial_chain over the job's terms from 0.0
// (the reference's rounding sequence).  tree: kSumParts blocks per job,
// each a fixed-shape block reduction of one contiguous share, then
// k_sums_finish adds the shares.
__global__ __launch_bounds__(256) void k_sums(const SumJob* __restrict__ jobs, double* __restrict__ out,
                                              double* __restrict__ partials, int serial, int parts) {
    const SumJob j = jobs[blockIdx.x];
    const long long start = sum_start(j);
    const double mean = sum_mean(j, out);
    if (serial) {
        __shared__ double tile[2][kSerialTile];
        const double s = serial_chain(j, mean, start, j.len, 0.0, tile);
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

// ---------------------------------------------------------------- speculative serial sums
// The reference-order sum of a long signal split into segments that run at
// once, with the same bits as serial_chain.  Rounded f64 addition is
// monotonic and a chain's value after a term depends only on its value
// before it, so segment p is run from C consecutive start values
// (kSpecPerLane to a lane: est_p - C/2 ulps .. est_p + C/2 - 1 ulps)
// centred on an accurate estimate est_p of the true running sum at its
// first term; once segment p - 1's end value e is known, the lane whose
// start equals e holds segment p's true end value.  When e lies outside
// the candidates (or is not finite), segment p is run again serially
// from e.  The result never depends on the estimates, only the speed does.
//
//   k_spec_chunks   double-double sums of 64-term chunks
//   k_spec_plan     their prefix; segment boundaries at chunk edges, each
//                   the edge of largest |prefix| within a window of its
//                   nominal place (the ulp gap between the rounded chain
//                   and the exact prefix is widest where the running sum
//                   passes near zero); est_p = that prefix
//   k_spec_chains   every candidate chain of every segment
//   k_spec_resolve  segment by segment: lookup, or the serial chain
struct DD {
    double hi, lo;
};
__device__ inline DD two_sum(double a, double b) {
    const double s = a + b;
    const double v = s - a;
    return {s, (a - (s - v)) + (b - v)};
}
__device__ inline DD fast_two_sum(double a, double b) {
    const double s = a + b;
    return {s, b - (s - a)};
}
// Double-double addition with both error terms carried (relative error
// about 2^-104; -ffp-contract=off keeps the error-free transforms exact).
__device__ inline DD dd_add(DD a, DD b) {
    DD s = two_sum(a.hi, b.hi);
    const DD t = two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = fast_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return fast_two_sum(s.hi, s.lo);
}
// Doubles as integers in value order (consecutive finite values differ by
// one; +0.0 and -0.0 both map to 0, and 0 maps back to +0.0).
__device__ inline long long ordered_bits(double v) {
    const long long b = __double_as_longlong(v);
    return b < 0 ? -(b & 0x7fffffffffffffffLL) : b;
}
__device__ inline double from_ordered_bits(long long o) {
    return __longlong_as_double(o < 0 ? ((-o) | (long long)0x8000000000000000ULL) : o);
}

constexpr int kSpecChunk = 64;  // terms per estimate chunk
constexpr int kSpecPlanThreads = 1024;
constexpr int kSpecPerLane = 2;  // candidate starts per lane

// Scratch of one launch: chunk sums, segment starts and estimates, segment
// counts, and every candidate chain's end value.
struct SpecLayout {
    long long nchunks;  // chunks of the longest job
    long long tstride;  // terms per job in the terms array (the longest job's)
    int seg_chunks;     // nominal segment length in chunks
    int window;         // boundary search: ± chunks around the nominal edge
    int pmax;           // segments of the longest job
    int cands;          // candidate starts per segment
    size_t chunks, terms, begin, est, nseg, starts, ends, bytes;
};
__host__ __device__ inline long long spec_chunks_for(long long terms) {
    return terms > 0 ? (terms + kSpecChunk - 1) / kSpecChunk : 0;
}
__host__ __device__ inline int spec_segments_for(long long nchunks, int seg_chunks) {
    return nchunks <= seg_chunks ? 1 : (int)((nchunks + seg_chunks - 1) / seg_chunks);
}
// Segment length, boundary window and candidate count per launch.  The
// rounded chain's distance from the estimate at a boundary grows about as
// the square root of the position: on ten-second music it reached 1776
// ulps (1024 candidates then missed up to 3 of a signal's 29 boundaries),
// on 30-45 K-sample slices with 4 K-term segments 487 (2 of 690 beyond
// ±256, each miss one 4 K-term serial segment).  Windows stay within
// a quarter of the segment, so neighbouring windows never meet (nominal
// edges lie at least half a segment apart).
inline SpecLayout spec_layout(int njobs, long long max_len) {
    SpecLayout L{};
    L.nchunks = spec_chunks_for(max_len);
    const bool shortjobs = max_len < kSpecLongLen;
    L.seg_chunks = shortjobs ? 64 : 128;
    L.window = shortjobs ? 16 : 32;
    L.cands = shortjobs ? 512 : 2048;  // whole blocks: multiples of 256 · kSpecPerLane
    L.pmax = spec_segments_for(L.nchunks, L.seg_chunks);
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    size_t off = 0;
    L.tstride = max_len > 0 ? max_len : 1;
    L.chunks = off;
    off += up(sizeof(DD) * (size_t)njobs * (size_t)(L.nchunks > 0 ? L.nchunks : 1));
    L.terms = off;
    off += up(sizeof(double) * (size_t)njobs * (size_t)L.tstride);
    L.begin = off;
    off += up(sizeof(long long) * (size_t)njobs * (size_t)(L.pmax + 1));
    L.est = off;
    off += up(sizeof(double) * (size_t)njobs * (size_t)L.pmax);
    L.nseg = off;
    off += up(sizeof(int) * (size_t)njobs);
    L.starts = off;
    off += up(sizeof(long long) * (size_t)njobs);
    L.ends = off;
    off += up(sizeof(double) * (size_t)njobs * (size_t)L.pmax * (size_t)L.cands);
    L.bytes = off;
    return L;
}

// One wave per chunk: lane l forms term 64c + l (coalesced), stores it in
// the terms array, and the wave adds its 64 terms in double-double.
__device__ inline DD dd_shfl_down(DD v, int off) {
    return {__shfl_down(v.hi, off, 64), __shfl_down(v.lo, off, 64)};
}
__global__ __launch_bounds__(256) void k_spec_chunks(const SumJob* __restrict__ jobs, const double* __restrict__ out,
                                                     SpecLayout L, char* __restrict__ scratch) {
    static_assert(kSpecChunk == 64, "one wave per chunk");
    const SumJob j = jobs[blockIdx.y];
    const long long start = sum_start(j);
    const long long c = (long long)blockIdx.x * 4 + threadIdx.x / 64;
    if (c >= spec_chunks_for(j.len - start)) return;  // whole wave
    const int lane = threadIdx.x & 63;
    const long long i = start + c * kSpecChunk + lane;
    double v = 0.0;
    if (i < j.len) {
        v = sum_term(j.kind, as_global(j.x)[i], i, sum_mean(j, out));
        reinterpret_cast<double*>(scratch + L.terms)[(long long)blockIdx.y * L.tstride + (i - start)] = v;
    }
    DD acc{v, 0.0};
    for (int off = 32; off > 0; off >>= 1) acc = dd_add(acc, dd_shfl_down(acc, off));
    if (lane == 0) reinterpret_cast<DD*>(scratch + L.chunks)[(long long)blockIdx.y * L.nchunks + c] = acc;
}

__global__ __launch_bounds__(kSpecPlanThreads) void k_spec_plan(const SumJob* __restrict__ jobs, SpecLayout L,
                                                                char* __restrict__ scratch) {
    const int job = blockIdx.x;
    const SumJob j = jobs[job];
    const long long start = sum_start(j);
    const long long nch = spec_chunks_for(j.len - start);
    const int P = spec_segments_for(nch, L.seg_chunks);  // <= L.pmax: the job is no longer than the longest
    const DD* ch = reinterpret_cast<const DD*>(scratch + L.chunks) + (long long)job * L.nchunks;
    long long* begin = reinterpret_cast<long long*>(scratch + L.begin) + (long long)job * (L.pmax + 1);
    double* est = reinterpret_cast<double*>(scratch + L.est) + (long long)job * L.pmax;
    if (threadIdx.x == 0) {
        begin[0] = start;
        est[0] = 0.0;
        begin[P] = j.len;
        reinterpret_cast<int*>(scratch + L.nseg)[job] = P;
        reinterpret_cast<long long*>(scratch + L.starts)[job] = start;
    }
    if (P == 1) return;
    // exclusive double-double prefix of each thread's run of chunks
    __shared__ DD scan[kSpecPlanThreads];
    __shared__ unsigned long long key[kSpecMaxSegs];
    const long long per = (nch + kSpecPlanThreads - 1) / kSpecPlanThreads;
    const long long c0 = per * threadIdx.x;
    const long long c1 = c0 + per < nch ? c0 + per : nch;
    // a thread's run of chunk sums, held in registers when it is short
    // (all loads at once; static indices keep the array in VGPRs)
    constexpr int kCache = 8;
    const bool cached = per <= kCache;
    DD r[kCache];
    if (cached) {
#pragma unroll
        for (int u = 0; u < kCache; u++) r[u] = c0 + u < c1 ? ch[c0 + u] : DD{0.0, 0.0};
    }
    auto each = [&](auto&& f) {  // f(c, sum of chunk c) for the thread's chunks in order
        if (cached) {
#pragma unroll
            for (int u = 0; u < kCache; u++)
                if (c0 + u < c1) f(c0 + u, r[u]);
        } else {
            for (long long c = c0; c < c1; c++) f(c, ch[c]);
        }
    };
    DD tot{0.0, 0.0};
    each([&](long long, DD v) { tot = dd_add(tot, v); });
    scan[threadIdx.x] = tot;
    for (int q = threadIdx.x; q < P; q += kSpecPlanThreads) key[q] = 0;
    __syncthreads();
    for (int off = 1; off < kSpecPlanThreads; off <<= 1) {
        const DD v = (int)threadIdx.x >= off ? scan[threadIdx.x - off] : DD{0.0, 0.0};
        __syncthreads();
        scan[threadIdx.x] = dd_add(scan[threadIdx.x], v);
        __syncthreads();
    }
    const DD excl = threadIdx.x > 0 ? scan[threadIdx.x - 1] : DD{0.0, 0.0};
    // The boundary whose window holds chunk edge c (edge c: the prefix of
    // chunks [0, c)), or 0.  Nominal edges q·nch/P lie at least half a
    // segment (twice a window) apart, so at most one window holds c, and
    // the nearest q is it.
    const int W = L.window;
    auto window = [&](long long c, long long& lo_edge) -> int {
        const long long q = (c * P + nch / 2) / nch;
        if (q < 1 || q > P - 1) return 0;
        lo_edge = q * nch / P - W;
        return (c >= lo_edge && c < lo_edge + 2 * W) ? (int)q : 0;
    };
    // An edge's key: |prefix| with its 7 lowest bits replaced by the
    // edge's place in the window counted from the far end, so the largest
    // key is an edge of (all but) the largest |prefix|, the first one on a
    // tie.  One walk takes each window's largest key; a second writes the
    // boundary at the edge that holds it.

    for (int pass = 0; pass < 2; pass++) {
        DD pre = excl;
        each([&](long long c, DD v) {
            long long lo_edge = 0;
            const int q = window(c, lo_edge);
            if (q) {
                const unsigned long long k =
                    ((unsigned long long)__double_as_longlong(fabs(pre.hi)) & ~0x7FULL) |
                    (unsigned long long)(2 * W - 1 - (c - lo_edge));  // < 128: 2W <= 128
                if (pass == 0) {
                    atomicMax(&key[q], k);
                } else if (k == key[q]) {
                    begin[q] = start + c * kSpecChunk;
                    est[q] = pre.hi;  // the rounded double-double prefix
                }
            }
            pre = dd_add(pre, v);
        });
        __syncthreads();
    }
}

// Every candidate of a segment adds the same term at the same step: the
// segment's terms are staged through LDS a tile at a time (the next tile's
// loads in flight while the current one is added) and each term read as
// one LDS broadcast feeds the lane's kSpecPerLane chains.  A broadcast
// returns 8 bytes to every lane, so at one candidate a lane the LDS return
// path bounds this kernel once a few waves share a CU; two independent
// chains a lane halve that traffic and still fit in the ~10-cycle latency
// of each chain's step (a wave issues one f64 fma in about 5 cycles).
// Measured on one ten-second pair's three launches: two a lane 2.24-2.27
// ms per exact pair, one 2.25-2.35, four 2.47.
__global__ __launch_bounds__(256) void k_spec_chains(const int* __restrict__ nseg, const long long* __restrict__ seg_begin,
                                                     const double* __restrict__ seg_est,
                                                     const double* __restrict__ terms, double* __restrict__ ends,
                                                     const long long* __restrict__ starts, SpecLayout L) {
    constexpr int K = kSpecPerLane;
    const int job = blockIdx.y;
    const int per_seg = L.cands / (256 * K);
    const int p = blockIdx.x / per_seg;
    if (p >= nseg[job]) return;  // whole block: uniform exit
    const long long* begin = seg_begin + (long long)job * (L.pmax + 1);
    const long long lo = begin[p], hi = begin[p + 1];
    const int cand = ((blockIdx.x % per_seg) * 256 + threadIdx.x) * K;  // the lane's first candidate
    const long long o = ordered_bits(seg_est[(long long)job * L.pmax + p]) + (cand - L.cands / 2);
    const double one = one_opaque();
    double s[K];
#pragma unroll
    for (int k = 0; k < K; k++) s[k] = from_ordered_bits(o + k);
    const double* t = terms + (long long)job * L.tstride + (lo - starts[job]);
    __shared__ double tile[2][kSerialTile];
    constexpr int per = kSerialTile / 256;
    const long long total = hi - lo;
    const long long ntiles = total > 0 ? (total + kSerialTile - 1) / kSerialTile : 0;
    double v[per];
    auto fetch = [&](long long base) {
        for (int q = 0; q < per; q++) {
            const long long i = base + (long long)q * 256 + threadIdx.x;
            v[q] = t[i < total ? i : 0];
        }
    };
    if (ntiles > 0) fetch(0);
    for (long long tt = 0; tt < ntiles; tt++) {
        const int cur = (int)(tt & 1);
        for (int q = 0; q < per; q++) tile[cur][q * 256 + threadIdx.x] = v[q];
        __syncthreads();
        if (tt + 1 < ntiles) fetch((tt + 1) * kSerialTile);
        const long long rem = total - tt * kSerialTile;
        const int cnt = rem < kSerialTile ? (int)rem : kSerialTile;
        const double* b = tile[cur];
        int k = 0;
        for (; k + 16 <= cnt; k += 16) {
            double a[16];
#pragma unroll
            for (int u = 0; u < 16; u++) a[u] = b[k + u];
#pragma unroll
            for (int u = 0; u < 16; u++)
#pragma unroll
                for (int c = 0; c < K; c++) s[c] = acc_add(s[c], a[u], one);
        }
        for (; k < cnt; k++) {
#pragma unroll
            for (int c = 0; c < K; c++) s[c] = acc_add(s[c], b[k], one);
        }
        __syncthreads();
    }
    double* e = ends + ((long long)job * L.pmax + p) * L.cands + cand;
#pragma unroll
    for (int c = 0; c < K; c++) e[c] = s[c];
}

__global__ __launch_bounds__(256) void k_spec_resolve(const SumJob* __restrict__ jobs, double* __restrict__ out,
                                                      SpecLayout L, const char* __restrict__ scratch) {
    const int job = blockIdx.x;
    const SumJob j = jobs[job];
    const double mean = sum_mean(j, out);
    const int P = reinterpret_cast<const int*>(scratch + L.nseg)[job];
    const long long* begin = reinterpret_cast<const long long*>(scratch + L.begin) + (long long)job * (L.pmax + 1);
    const double* est = reinterpret_cast<const double*>(scratch + L.est) + (long long)job * L.pmax;
    const double* ends = reinterpret_cast<const double*>(scratch + L.ends) + (long long)job * L.pmax * L.cands;
    const int half = L.cands / 2;
    __shared__ double tile[2][kSerialTile];
    __shared__ double e_sh;
    __shared__ int stop_sh;
    // segment 0 starts from 0.0 exactly: candidate C/2 (offset 0)
    int p = 1;
    double e = ends[half];
    for (;;) {
        if (threadIdx.x == 0) {
            int q = p;
            for (; q < P; q++) {
                const long long d = ordered_bits(e) - ordered_bits(est[q]);
                if (!isfinite(e) || (e == 0.0 && signbit(e)) || d < -half || d >= half) break;
                e = ends[(long long)q * L.cands + (d + half)];
            }
            e_sh = e;
            stop_sh = q;
        }
        __syncthreads();
        const int q = stop_sh;
        e = e_sh;
        if (q >= P) break;
        // segment q's true start lies outside its candidates: run it serially
        e = serial_chain(j, mean, begin[q], begin[q + 1], e, tile);  // lane 0's value
        p = q + 1;
        __syncthreads();  // every lane has read e_sh and stop_sh before lane 0 writes them again
    }
    if (threadIdx.x == 0) out[j.slot] = e;
}

size_t spec_sums_scratch_bytes(int njobs, long long max_len) { return spec_layout(njobs, max_len).bytes; }

bool spec_sums_fit(long long max_len) { return spec_layout(1, max_len).pmax <= kSpecMaxSegs; }

double spec_sums_cost_ratio(int njobs, long long max_len, long long total_len) {
    // time model in cycles: the speculative form's longest candidate chain
    // (~10 cycles a term) or its LDS broadcasts (8 bytes per lane per term
    // to C / (64 · kSpecPerLane) waves a segment, 128 bytes a cycle on each
    // of 256 CUs), whichever is longer, plus ~40 us of fixed launches;
    // against one serial chain of the longest job
    const SpecLayout L = spec_layout(njobs, max_len);
    const double seg = 10.0 * (double)L.seg_chunks * kSpecChunk * 1.5;
    const double lds = (double)total_len * L.cands * 8.0 / kSpecPerLane / (256.0 * 128.0);
    const double spec = (seg > lds ? seg : lds) + 40e-6 * 2.4e9;
    return spec / (10.0 * (double)max_len);
}

void launch_sums_spec(const SumJob* d_jobs, int njobs, long long max_len, double* d_out, void* d_scratch,
                      hipStream_t stream) {
    if (njobs == 0) return;
    const SpecLayout L = spec_layout(njobs, max_len);
    if (L.pmax > kSpecMaxSegs) {  // more segments than the plan's LDS holds: the serial form
        launch_sums(d_jobs, njobs, max_len, d_out, true, nullptr, stream);
        return;
    }
    char* sc = static_cast<char*>(d_scratch);
    const unsigned cblocks = (unsigned)((L.nchunks + 3) / 4);  // four chunks (waves) per block
    if (cblocks > 0)
        hipLaunchKernelGGL(k_spec_chunks, dim3(cblocks, njobs), dim3(256), 0, stream, d_jobs, (const double*)d_out, L,
                           sc);
    hipLaunchKernelGGL(k_spec_plan, dim3(njobs), dim3(kSpecPlanThreads), 0, stream, d_jobs, L, sc);
    hipLaunchKernelGGL(k_spec_chains, dim3((unsigned)L.pmax * (unsigned)(L.cands / (256 * kSpecPerLane)), njobs),
                       dim3(256), 0, stream,
                       (const int*)(sc + L.nseg), (const long long*)(sc + L.begin), (const double*)(sc + L.est),
                       (const double*)(sc + L.terms), (double*)(sc + L.ends), (const long long*)(sc + L.starts), L);
    hipLaunchKernelGGL(k_spec_resolve, dim3(njobs), dim3(256), 0, stream, d_jobs, d_out, L, (const char*)sc);
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
