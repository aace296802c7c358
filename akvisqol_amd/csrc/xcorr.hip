// FFT cross-correlation lag search on gfx950 (xcorr.zig:180-251,
// envelope.zig:23-53, dsp/fft.zig:45-166).
//
// z = env(a) + i·env(b), zero padded to n, goes through a radix-2 DIF
// forward transform (bit-reversed out), the cross-spectrum is formed in
// bit-reversed order, and a DIT inverse (bit-reversed in) gives the circular
// correlation; the lag is the first maximum of |Re| scanning lags
// -max..-1 then 0..max.
//
// Each HBM pass runs radix-2 stages on 2048- or 4096-element groups in LDS:
// the stages with spans S_hi .. S_hi/2^(k-1) only mix elements
// blk·S_hi + r + m·q (q = S_hi/2^k) of one residue r, so a workgroup loads R
// residues of Bn span blocks (R·16 B contiguous per m).  Every butterfly is
// the radix-2 butterfly of its stage with the same twiddle, so results equal
// the stage-by-stage transform bit for bit.  The first DIF pass builds its
// input from the two signals (the envelope is never stored); the last DIT
// pass keeps no output, only each workgroup's first maximum in scan order.
#include <float.h>

#include <type_traits>

#include "dev_common.h"
#include "kernels.h"

namespace vq {

// Complex arithmetic on double2 (the exact transform) and float2 (the
// throughput mode's nominating transform).
template <typename C> struct Cx;
template <> struct Cx<double2> {
    using R = double;
    __device__ static double2 make(double x, double y) { return make_double2(x, y); }
};
template <> struct Cx<float2> {
    using R = float;
    __device__ static float2 make(float x, float y) { return make_float2(x, y); }
};
template <typename C>
__device__ inline C cadd(C a, C b) { return Cx<C>::make(a.x + b.x, a.y + b.y); }
template <typename C>
__device__ inline C csub(C a, C b) { return Cx<C>::make(a.x - b.x, a.y - b.y); }
// fft.zig:181-188: cmul and the conjugate form, in the reference's order.
template <typename C>
__device__ inline C cmul(C x, C w) {
    return Cx<C>::make(x.x * w.x - x.y * w.y, x.x * w.y + x.y * w.x);
}
template <typename C>
__device__ inline C cmul_conj(C x, C w) {
    const typename Cx<C>::R wi = -w.y;
    return Cx<C>::make(x.x * w.x - x.y * wi, x.x * wi + x.y * w.x);
}
// The nominating transform's f32 products with fused multiply-adds (its
// error bound covers the separately rounded form, so it covers these); the
// f64 transform keeps the reference's separate roundings above.
__device__ inline float2 cmul(float2 x, float2 w) {
    return make_float2(__builtin_fmaf(x.x, w.x, -(x.y * w.y)), __builtin_fmaf(x.x, w.y, x.y * w.x));
}
__device__ inline float2 cmul_conj(float2 x, float2 w) {
    return make_float2(__builtin_fmaf(x.x, w.x, x.y * w.y), __builtin_fmaf(x.y, w.x, -(x.x * w.y)));
}

__device__ inline long long env_fft_size(long long l) {
    long long n = 1;
    while (n < l) n <<= 1;
    return n < 32 ? 32 : n;
}

// envelope.zig:23-53 closed form for sample i of a signal with the given
// plain sum and alternating (Nyquist) sum: the per-signal constants once,
// then one value per sample.
struct EnvSide {
    const double* x;
    long long len;
    double mean, nq;
};
__device__ inline EnvSide env_side(const double* x, long long len, double sum, double nyq) {
    EnvSide e{x, len, 0.0, 0.0};
    if (len > 0) {
        e.mean = sum / (double)len;
        e.nq = nyq / (double)env_fft_size(len);
    }
    return e;
}
__device__ inline double env_value(const EnvSide& e, double xi, long long i) {
    const double term = (i & 1) ? -e.nq : e.nq;
    return fabs(2.0 * ((xi - e.mean) - term) - 0.000001) + e.mean;
}
// The sample a lane needs for index i, loaded unconditionally from an
// in-range address (the value is discarded past the end).
__device__ inline double env_load(const EnvSide& e, long long i) {
    return e.len > 0 ? as_global(e.x)[i < e.len ? i : 0] : 0.0;
}

__device__ inline double2 packed_env(const LagJob& j, const double* sums, long long i) {
    const EnvSide a = env_side(j.a, j.la, sums[j.sum_a], sums[j.nyq_a]);
    const EnvSide b = env_side(j.b, j.lb, sums[j.sum_b], sums[j.nyq_b]);
    return make_double2(i < a.len ? env_value(a, env_load(a, i), i) : 0.0,
                        i < b.len ? env_value(b, env_load(b, i), i) : 0.0);
}

constexpr int kTile = 2048;  // complex doubles per workgroup (32 KiB LDS)
constexpr int kThreads = 256;
constexpr int kBfPerThread = kTile / 2 / kThreads;

// Tile sources: the row z, the two signals' envelopes, or a row holding
// the nominating transform's half-length inverse (half position h at
// h + (h & ~(kHalfChunk32 - 1)), value c[2h] + i·c[2h + 1]).
enum : int { kSrcZ = 0, kSrcEnvelope = 1, kSrcHalf = 2 };
constexpr long long kHalfChunk32 = 2048;  // lag32_xcorr_slot's chunk (kernels.h)

// Scan position of circular index i (xcorr.zig:233-250), or -1 outside
// the lag range.
__device__ inline long long scan_order(long long i, long long n, long long max_lag) {
    if (i >= n - max_lag) return i - (n - max_lag);
    if (i <= max_lag) return max_lag + i;
    return -1;
}

// First maximum of |Re| in scan order over a tile in LDS; writes the
// (max bits, scan position) pair of this workgroup to part[0..1].  The
// reduction reuses the tile's LDS (at least 2·THREADS words) once every
// thread has read its elements.
template <int THREADS, typename GIDX>
__device__ inline void tile_argmax(double2* sm, int count, GIDX gidx, long long n, const LagJob& job,
                                   unsigned long long* part) {
    const long long big = job.la > job.lb ? job.la : job.lb;
    const long long max_lag = big - 1;
    unsigned long long best_v = 0, best_o = ~0ULL;
    for (int e = threadIdx.x; e < count; e += THREADS) {
        const long long o = scan_order(gidx(e), n, max_lag);
        if (o < 0) continue;
        const unsigned long long v = f64_bits(fabs(sm[e].x));
        if (v > best_v || (v == best_v && (unsigned long long)o < best_o)) {
            best_v = v;
            best_o = (unsigned long long)o;
        }
    }
    __syncthreads();
    unsigned long long* rv = reinterpret_cast<unsigned long long*>(sm);
    unsigned long long* ro = rv + THREADS;
    rv[threadIdx.x] = best_v;
    ro[threadIdx.x] = best_o;
    __syncthreads();
    for (int s = THREADS / 2; s > 0; s >>= 1) {
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
        part[0] = rv[0];
        part[1] = ro[0];
    }
}

// DIF butterfly (fft.zig:45-90): (a, b) -> (a + b, (a - b)·w).
template <typename C>
__device__ inline void bf_dif(C& a, C& b, C w) {
    const C t = csub(a, b);
    a = cadd(a, b);
    b = cmul(t, w);
}
// DIT butterfly with the conjugate twiddle (fft.zig:92-166).
template <typename C>
__device__ inline void bf_dit(C& a, C& b, C w) {
    const C bw = cmul_conj(b, w);
    b = csub(a, bw);
    a = cadd(a, bw);
}

// Every size is a power of two: lgn = log2 n, lgS = log2 S_hi, k stages,
// R = 2^lgR residues, Bn = 2^lgBn span blocks per workgroup.
template <bool DIT, int SRC, bool ARGMAX>
__global__ __launch_bounds__(kThreads) void k_fft_pass(const LagJob* __restrict__ jobs, int lgn, int lgS, int k, int lgR,
                                                       int lgBn, const double2* __restrict__ tw,
                                                       const double* __restrict__ sums,
                                                       unsigned long long* __restrict__ partials, int njobs) {
    __shared__ double2 sm[kTile];
    // Job index fastest: workgroups in flight together share residues, so
    // their twiddle reads hit the same cache lines.
    const int jb = (int)(blockIdx.x % (unsigned)njobs);
    const long long wg = blockIdx.x / (unsigned)njobs;
    const long long nwg = gridDim.x / (unsigned)njobs;
    const LagJob job = jobs[jb];
    const long long n = 1LL << lgn;
    const int lgq = lgS - k;                 // q = S_hi / 2^k
    const int lgGR = k + lgR;                // elements per span block in the tile
    const int lgPer = lgq - lgR;             // workgroups per span-block group
    const long long blk0 = (wg >> lgPer) << lgBn;
    const long long r0 = (wg & ((1LL << lgPer) - 1)) << lgR;
    const long long base = (blk0 << lgS) + r0;
    const int count = 1 << (lgGR + lgBn);   // tile elements
    double2* z = job.z;
    // tile element e = (bb, m, rr) <-> z[base + bb·S_hi + m·q + rr]
    auto gidx = [&](int e) -> long long {
        const int bb = e >> lgGR;
        const int m = (e >> lgR) & ((1 << k) - 1);
        const int rr = e & ((1 << lgR) - 1);
        return base + ((long long)bb << lgS) + ((long long)m << lgq) + rr;
    };
    for (int e = threadIdx.x; e < count; e += kThreads) {
        const long long i = gidx(e);
        sm[e] = SRC == kSrcEnvelope ? packed_env(job, sums, i) : load2(z, i);
    }
    // Butterfly u of this thread at level li: (i0, i1) in LDS and its twiddle
    // index.  Stage l has distance hm·q (hm = 2^(k-1-l)) and twiddle
    // W_{S_l}^j (S_l = S_hi/2^l) at tw[S_l/2 + j].
    const int lgHalf = lgGR - 1;
    auto butterfly = [&](int li, int u, int& i0, int& i1) -> long long {
        const int l = DIT ? k - 1 - li : li;
        const int lgHm = k - 1 - l;
        const int t = threadIdx.x + u * kThreads;
        const int bb = t >> lgHalf, rem = t & ((1 << lgHalf) - 1);
        const int rr = rem & ((1 << lgR) - 1), bm = rem >> lgR;
        const int lo = bm & ((1 << lgHm) - 1);
        const int m0 = ((bm >> lgHm) << (lgHm + 1)) | lo;
        i0 = (bb << lgGR) + (m0 << lgR) + rr;
        i1 = i0 + (1 << (lgHm + lgR));
        const long long j = ((long long)lo << lgq) + r0 + rr;
        return (1LL << (lgS - l - 1)) + j;
    };
    const int nbf = count >> 1;  // == kTile / 2 whenever the tile is full
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
        for (int e = threadIdx.x; e < count; e += kThreads) store2(z, gidx(e), sm[e]);
        return;
    }
    tile_argmax<kThreads>(sm, count, gidx, n, job, partials + 2 * (jb * nwg + wg));
}

template <int K, bool DIT>
struct TileStep {
    // processing index li -> stage l (distance 2^(K-1-l) in units of q)
    static constexpr int stage(int li) { return DIT ? K - 1 - li : li; }
};

// The K radix-2 stages of one full tile in LDS (element (bb, m, rr) at
// (bb << (K + lgR)) + (m << lgR) + rr, m the index along the stages'
// spans), fully unrolled.  Stages are taken two at a time: a thread holds
// the four elements m0, m0+h, m0+2h, m0+3h of one residue (h the smaller of
// the two stage distances, in units of q) and runs both stages'
// butterflies on them in registers with those stages' twiddles — the same
// radix-2 operations as the stage-by-stage transform.  An odd K ends with
// one radix-2 stage.  Twiddles of the next step load while the current one
// computes.  Twiddle of stage l (span S_l = 2^(lgq + K - l)) at position
// j = p·q + r0 + rr: tw[S_l/2 + j].  Starts and ends with a barrier.
template <int K, bool DIT, int THREADS, typename C>
__device__ inline void tile_stages(C* sm, int lgR, int lgq, long long r0, const C* __restrict__ tw) {
    const int lgGR = K + lgR;
    // Radix-4 group g (two per thread): residue rr, quarter index mq, block bb.
    // lgH = log2 of the smaller distance.  Element x_c sits at m0 + c·h.
    auto quad = [&](int g, int lgH, int& i0, int& lo, int& rr) {
        rr = g & ((1 << lgR) - 1);
        const int mq = (g >> lgR) & ((1 << (K - 2)) - 1);
        const int bb = g >> (lgR + K - 2);
        lo = mq & ((1 << lgH) - 1);
        const int m0 = ((mq >> lgH) << (lgH + 2)) | lo;
        i0 = (bb << lgGR) + (m0 << lgR) + rr;
    };
    // Twiddle of stage l at position p (< distance) for residue r0 + rr.
    auto twi = [&](int l, int p, int rr) -> C {
        const long long j = ((long long)p << lgq) + r0 + rr;
        return tw[(1LL << (lgq + K - l - 1)) + j];
    };
    constexpr int kSteps = (K + 1) / 2;
    // w[u][0..1]: the larger-distance stage at lo and lo + h; w[u][2]: the
    // smaller-distance stage at lo.  Radix-2 step: w[u][0] for butterfly u.
    C w[4][3];
    auto load_tw = [&](int st, C (&dst)[4][3]) {
        const int li = 2 * st;
        if (li + 1 < K) {
            const int lgH = DIT ? li : K - 2 - li;  // smaller distance 2^lgH
            const int l_big = K - 2 - lgH, l_small = K - 1 - lgH;
#pragma unroll
            for (int u = 0; u < 2; u++) {
                int i0, lo, rr;
                quad(threadIdx.x + u * THREADS, lgH, i0, lo, rr);
                dst[u][0] = twi(l_big, lo, rr);
                dst[u][1] = twi(l_big, lo + (1 << lgH), rr);
                dst[u][2] = twi(l_small, lo, rr);
            }
        } else {
            const int l = TileStep<K, DIT>::stage(li);
            const int lgHm = K - 1 - l;
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int t = threadIdx.x + u * THREADS;
                const int rem = t & ((1 << (lgGR - 1)) - 1);
                const int rr = rem & ((1 << lgR) - 1), bm = rem >> lgR;
                dst[u][0] = twi(l, bm & ((1 << lgHm) - 1), rr);
            }
        }
    };
    load_tw(0, w);
    __syncthreads();
#pragma unroll
    for (int st = 0; st < kSteps; st++) {
        const int li = 2 * st;
        C wn[4][3];
        if (li + 1 < K) {
            const int lgH = DIT ? li : K - 2 - li;
            C x[2][4];
            int ib[2];
#pragma unroll
            for (int u = 0; u < 2; u++) {
                int lo, rr;
                quad(threadIdx.x + u * THREADS, lgH, ib[u], lo, rr);
#pragma unroll
                for (int c = 0; c < 4; c++) x[u][c] = sm[ib[u] + ((c << lgH) << lgR)];
            }
            if (st + 1 < kSteps) load_tw(st + 1, wn);
#pragma unroll
            for (int u = 0; u < 2; u++) {
                if (DIT) {  // smaller distance first, then the larger
                    bf_dit(x[u][0], x[u][1], w[u][2]);
                    bf_dit(x[u][2], x[u][3], w[u][2]);
                    bf_dit(x[u][0], x[u][2], w[u][0]);
                    bf_dit(x[u][1], x[u][3], w[u][1]);
                } else {    // larger distance first, then the smaller
                    bf_dif(x[u][0], x[u][2], w[u][0]);
                    bf_dif(x[u][1], x[u][3], w[u][1]);
                    bf_dif(x[u][0], x[u][1], w[u][2]);
                    bf_dif(x[u][2], x[u][3], w[u][2]);
                }
#pragma unroll
                for (int c = 0; c < 4; c++) sm[ib[u] + ((c << lgH) << lgR)] = x[u][c];
            }
        } else {
            const int l = TileStep<K, DIT>::stage(li);
            const int lgHm = K - 1 - l;
            C x[4][2];
            int ia[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int t = threadIdx.x + u * THREADS;
                const int bb = t >> (lgGR - 1), rem = t & ((1 << (lgGR - 1)) - 1);
                const int rr = rem & ((1 << lgR) - 1), bm = rem >> lgR;
                const int m0 = ((bm >> lgHm) << (lgHm + 1)) | (bm & ((1 << lgHm) - 1));
                ia[u] = (bb << lgGR) + (m0 << lgR) + rr;
                x[u][0] = sm[ia[u]];
                x[u][1] = sm[ia[u] + ((1 << lgHm) << lgR)];
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                if (DIT) bf_dit(x[u][0], x[u][1], w[u][0]);
                else bf_dif(x[u][0], x[u][1], w[u][0]);
                sm[ia[u]] = x[u][0];
                sm[ia[u] + ((1 << lgHm) << lgR)] = x[u][1];
            }
        }
        __syncthreads();
        if (st + 1 < kSteps) {
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int c = 0; c < 3; c++) w[u][c] = wn[u][c];
        }
    }
}

// Full-tile pass (every tile holds TILE elements, eight per thread) with K
// stages: load the tile, run tile_stages, then store it, keep its first
// maximum of |Re| in scan order (the exact inverse transform's last pass),
// or keep its lag candidates (the nominating transform's last pass).
enum : int { kEpiStore = 0, kEpiArgmax = 1, kEpiCandidates = 2 };

// Lag candidates of one workgroup of the nominating transform's last pass:
// the tile's largest in-range |Re|, how many in-range elements lie within
// the row's slack of it, and the first kTileCands of them (scan position
// and |Re|).  The rows' candidates are a subset of the union of these.
constexpr int kTileCands = 30;
struct TileCands {
    float top;
    int count;
    int pos[kTileCands];
    float val[kTileCands];
};

// Block-wide sum of one double per thread in a fixed order (the same
// bits on every run): a shuffle tree per wave, then the waves in order.
// red holds THREADS / 64 doubles; the sum is valid in thread 0.
template <int THREADS>
__device__ inline double block_sum(double v, double* red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < THREADS / 64; w++) r += red[w];
    return r;
}

// The candidates epilogue (kEpiCandidates): in-range |Re| against the
// tile's own top.  The row's floor (row top - slack) is at least this
// tile's floor, so every row candidate in this tile is among those counted
// here; their order in out->pos is arbitrary (the row merge sorts them).
// HALF: element e holds c[2i] + i·c[2i + 1] of half position i = gidx(e),
// two lags each; n is the full transform size either way.
template <int THREADS, bool HALF, typename GIDX>
__device__ inline void tile_candidates(const float2* sm, GIDX gidx, long long n, const LagJob& job, float slack,
                                       TileCands* out) {
    constexpr int kPer = 8;
    constexpr int kVals = HALF ? 2 : 1;
    // 32-bit scan positions (n <= 2^30): scan_order's two ranges
    const int max_lag = (int)((job.la > job.lb ? job.la : job.lb) - 1);
    const int neg0 = (int)n - max_lag;  // first index of the negative lags
    auto order = [&](int j) -> int { return j >= neg0 ? j - neg0 : (j <= max_lag ? max_lag + j : -1); };
    auto value = [&](int e, int h) -> float { return fabsf(h ? sm[e].y : sm[e].x); };
    auto index = [&](int e, int h) -> int { return HALF ? 2 * (int)gidx(e) + h : (int)gidx(e); };
    unsigned top_bits = 0;
#pragma unroll
    for (int u = 0; u < kPer; u++) {
        const int e = threadIdx.x + u * THREADS;
#pragma unroll
        for (int h = 0; h < kVals; h++)
            if (order(index(e, h)) >= 0) top_bits = max(top_bits, __float_as_uint(value(e, h)));
    }
    __shared__ unsigned s_top;
    __shared__ int s_count;
    if (threadIdx.x == 0) {
        s_top = 0;
        s_count = 0;
    }
    __syncthreads();
    for (int off = 32; off > 0; off >>= 1) top_bits = max(top_bits, (unsigned)__shfl_down((int)top_bits, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(&s_top, top_bits);  // |Re| >= 0: bit order is value order
    __syncthreads();
    const float top = __uint_as_float(s_top);
    const float floor = top - slack;
#pragma unroll
    for (int u = 0; u < kPer; u++) {
        const int e = threadIdx.x + u * THREADS;
#pragma unroll
        for (int h = 0; h < kVals; h++) {
            const float v = value(e, h);
            if (!(v >= floor)) continue;
            const int so = order(index(e, h));
            if (so < 0) continue;
            const int slot = atomicAdd(&s_count, 1);
            if (slot < kTileCands) {
                out->pos[slot] = so;
                out->val[slot] = v;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        out->top = top;
        out->count = s_count;
    }
}

// Workgroups per CU: f64 tiles at least two (TILE / 1024 waves per SIMD;
// 4096-element tiles of 8 waves at most 128 VGPRs); f32 tiles of 8192
// elements (64 KiB, 16 waves) two, within 64 VGPRs.  With fused complex
// products and 32-bit row positions the f32 passes need 51-67 VGPRs, and
// two workgroups per CU took one 279-pair chunk's passes from 7.4 to 6.7 ms
// (the candidate passes 0.78 -> 0.52 and 0.60 -> 0.43 ms); the earlier
// 105-120 VGPR form spilled 70-160 bytes per lane under the same cap.
template <typename C, int TILE>
struct TileWaves {
    static constexpr int value = TILE / 1024;
};

// SRC == kSrcEnvelope builds the tile from the two signals' envelopes
// (computed in f64; a float2 tile rounds them to f32 and adds the tile's
// share of ‖env_a‖² + ‖env_b‖² to aux[job·nwg + wg], as doubles);
// kSrcHalf works on the half-length row of the nominating transform's
// inverse (lgn its length, half the transform's), loading and storing it
// where the middle pass put it.
// EPI == kEpiArgmax writes the (bits, position) pair of the tile's first
// maximum to aux; kEpiCandidates writes a TileCands to aux against the
// row's slack slack[job].
template <int K, bool DIT, int SRC, int EPI, int TILE, typename C>
__global__ __launch_bounds__(TILE / 8) __attribute__((amdgpu_waves_per_eu(TileWaves<C, TILE>::value))) void k_fft_tile(
    const LagJob* __restrict__ jobs, int lgn, int lgS, int lgR, int lgBn, const C* __restrict__ tw,
    const double* __restrict__ sums, void* __restrict__ aux, const float* __restrict__ slack, int njobs) {
    constexpr int THREADS = TILE / 8;
    __shared__ C sm[TILE];
    const int jb = (int)(blockIdx.x % (unsigned)njobs);
    const long long wg = blockIdx.x / (unsigned)njobs;
    const long long nwg = gridDim.x / (unsigned)njobs;
    const LagJob job = jobs[jb];
    const long long n = 1LL << lgn;
    const int lgq = lgS - K;
    const int lgGR = K + lgR;
    const int lgPer = lgq - lgR;
    const long long blk0 = (wg >> lgPer) << lgBn;
    const long long r0 = (wg & ((1LL << lgPer) - 1)) << lgR;
    // row positions: 32-bit for the nominating transform (n <= 2^30)
    using Idx = typename std::conditional<sizeof(C) == 8, int, long long>::type;
    const Idx base = (Idx)((blk0 << lgS) + r0);
    C* z = reinterpret_cast<C*>(job.z);
    auto gidx = [&](int e) -> Idx {
        const int bb = e >> lgGR;
        const int m = (e >> lgR) & ((1 << K) - 1);
        const int rr = e & ((1 << lgR) - 1);
        return base + ((Idx)bb << lgS) + ((Idx)m << lgq) + rr;
    };
    // where row position g is stored
    auto addr = [&](Idx g) -> Idx {
        if constexpr (SRC == kSrcHalf) return g + (g & ~(Idx)(kHalfChunk32 - 1));
        else return g;
    };
    constexpr int kPer = 8;
    if (SRC == kSrcEnvelope) {
        // both signals' samples for the tile, loads first, envelope after
        const EnvSide ea = env_side(job.a, job.la, sums[job.sum_a], sums[job.nyq_a]);
        const EnvSide eb = env_side(job.b, job.lb, sums[job.sum_b], sums[job.nyq_b]);
        // f64 tiles load all eight pairs first; f32 tiles (16 waves per
        // workgroup, two workgroups per CU) four at a time, within 64 VGPRs
        constexpr int kGroup = sizeof(C) == 16 ? kPer : kPer / 2;
        double energy = 0.0;
#pragma unroll
        for (int g = 0; g < kPer; g += kGroup) {
            if (sizeof(C) == 8 && g >= kPer / 2) {
                // the first pass's tile spans the whole row, its second
                // half the positions >= n/2, past both signals (n >= 2·len - 1)
#pragma unroll
                for (int u = 0; u < kGroup; u++) sm[threadIdx.x + (g + u) * THREADS] = Cx<C>::make(0, 0);
                continue;
            }
            double xa[kGroup], xb[kGroup];
#pragma unroll
            for (int u = 0; u < kGroup; u++) {
                const long long i = gidx(threadIdx.x + (g + u) * THREADS);
                xa[u] = env_load(ea, i);
                xb[u] = env_load(eb, i);
            }
#pragma unroll
            for (int u = 0; u < kGroup; u++) {
                const long long i = gidx(threadIdx.x + (g + u) * THREADS);
                const double va = i < ea.len ? env_value(ea, xa[u], i) : 0.0;
                const double vb = i < eb.len ? env_value(eb, xb[u], i) : 0.0;
                sm[threadIdx.x + (g + u) * THREADS] = Cx<C>::make((typename Cx<C>::R)va, (typename Cx<C>::R)vb);
                if constexpr (sizeof(C) == 8) energy = __builtin_fma(va, va, __builtin_fma(vb, vb, energy));
            }
        }
        if constexpr (sizeof(C) == 8) {
            __shared__ double red[THREADS / 64];
            const double e = block_sum<THREADS>(energy, red);
            if (threadIdx.x == 0) static_cast<double*>(aux)[jb * nwg + wg] = e;
        }
    } else {
        C v[kPer];
#pragma unroll
        for (int u = 0; u < kPer; u++) v[u] = load2(z, addr(gidx(threadIdx.x + u * THREADS)));
#pragma unroll
        for (int u = 0; u < kPer; u++) sm[threadIdx.x + u * THREADS] = v[u];
    }

    tile_stages<K, DIT, THREADS>(sm, lgR, lgq, r0, tw);
    if constexpr (EPI == kEpiStore) {
#pragma unroll
        for (int u = 0; u < kPer; u++) {
            const int e = threadIdx.x + u * THREADS;
            store2(z, addr(gidx(e)), sm[e]);
        }
        return;
    } else if constexpr (EPI == kEpiArgmax) {
        tile_argmax<THREADS>(sm, TILE, gidx, n, job, static_cast<unsigned long long*>(aux) + 2 * (jb * nwg + wg));
        return;
    } else {
        tile_candidates<THREADS, SRC == kSrcHalf>(sm, gidx, SRC == kSrcHalf ? 2 * n : n, job, slack[jb],
                                                  static_cast<TileCands*>(aux) + (jb * nwg + wg));
    }
}

// xcorr.zig:211-229 cross terms, in place in bit-reversed order.
__device__ inline float2 cross_term(float2 x, float2 zy) {
    const float2 y = make_float2(zy.x, -zy.y);
    const float2 sum = make_float2(x.x + y.x, x.y + y.y);
    const float2 dif_c = make_float2(x.x - y.x, -(x.y - y.y));
    const float2 idc = make_float2(-dif_c.y, dif_c.x);
    return make_float2(0.25f * (sum.x * idc.x - sum.y * idc.y), 0.25f * (sum.x * idc.y + sum.y * idc.x));
}
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
        const double2 v = load2(z, i);
        store2(z, i, cross_term(v, v));
        return;
    }
    const long long h = 1LL << (63 - __builtin_clzll((unsigned long long)i));
    const long long p = 3 * h - 1 - i;
    if (i < p) {
        const double2 za = load2(z, i), zb = load2(z, p);
        store2(z, i, cross_term(za, zb));
        store2(z, p, cross_term(zb, za));
    }
}

// Middle of the lag search (n >= TILE): the forward transform's last pass,
// the cross terms and the inverse transform's first pass in one kernel.
// Both passes work on aligned contiguous blocks of 2^K elements (K <=
// log2(TILE/2)), and the cross term of bit-reversed position i pairs it
// with 3h-1-i (h the largest power of two <= i), which maps chunk c of
// C = TILE/2 elements in [h, 2h) onto the whole chunk 2h/C - 1 - c; chunks
// 0 and 1 pair within themselves.
// Workgroup w holds the chunk pair (0, 1), or for w >= 1 with
// 2^k <= w < 2^(k+1): chunks 2^(k+1) + t and 2^(k+2) - 1 - t, t = w - 2^k.

// One radix-2^B step (B = 1, 2, 3) over the N (eight or four) elements a
// thread holds in a pass over contiguous blocks (one residue): N >> B
// groups, element c of
// group gi in x[gi·2^B + c] at block position m0 + c·2^lgH.  The stage with
// distance d·2^lgH (d = 2^s, s < B) pairs c with c + d (bit s of c clear)
// and takes W_{2d·2^lgH} at position lo + (c mod d)·2^lgH, i.e. table index
// (d << lgH) + lo + ((c mod d) << lgH).  DIF runs the stages largest
// distance first, DIT smallest first: the radix-2 butterflies of the
// stage-by-stage transform.
template <int B, bool DIT, typename C, int N>
__device__ inline void radix_step(C (&x)[N], const int (&lo)[N >> B], int lgH, const C* __restrict__ tw) {
    constexpr int G = 1 << B;
    constexpr int NG = N >> B;
    C w[NG][G];  // w[gi][d + t]: stage distance d, position t (t < d)
#pragma unroll
    for (int gi = 0; gi < NG; gi++)
#pragma unroll
        for (int s = 0; s < B; s++) {
            const int d = 1 << s;
#pragma unroll
            for (int t = 0; t < d; t++) w[gi][d + t] = tw[(d << lgH) + lo[gi] + (t << lgH)];
        }
#pragma unroll
    for (int si = 0; si < B; si++) {
        const int s = DIT ? si : B - 1 - si;
        const int d = 1 << s;
#pragma unroll
        for (int gi = 0; gi < NG; gi++)
#pragma unroll
            for (int c = 0; c < G; c++) {
                if (c & d) continue;
                C& a = x[gi * G + c];
                C& b = x[gi * G + c + d];
                if (DIT) bf_dit(a, b, w[gi][d + (c & (d - 1))]);
                else bf_dif(a, b, w[gi][d + (c & (d - 1))]);
            }
    }
}

// Tile positions (block bb of 2^K elements, position m) of the thread's
// N elements for a radix-2^B step with smallest distance 2^lgH: group
// G = thread + gi·THREADS, q = G mod 2^(K-B) its index in the block, lo = q
// mod 2^lgH, m0 = q with B zero bits inserted at bit lgH.
template <int K, int B, int THREADS, int N = 8>
__device__ inline void group_layout(int lgH, int (&idx)[N], int (&lo)[N >> B]) {
#pragma unroll
    for (int gi = 0; gi < (N >> B); gi++) {
        const int g = (int)threadIdx.x + gi * THREADS;
        const int q = g & ((1 << (K - B)) - 1);
        const int bb = g >> (K - B);
        lo[gi] = q & ((1 << lgH) - 1);
        const int m0 = ((q >> lgH) << (lgH + B)) | lo[gi];
#pragma unroll
        for (int c = 0; c < (1 << B); c++) idx[gi * (1 << B) + c] = (bb << K) + m0 + (c << lgH);
    }
}

// The nominating transform's steps of smallest distance (lgH = 0), whose
// twiddles are W_2^0, W_4^0..1 and W_8^0..3: as constants, 1 and -i exact
// (the table's W_4^1 has a 6e-17 real part) and W_8^1, W_8^3 = (±r, -r)
// with r = cos(pi/4) rounded to f32.  The same butterflies as radix_step
// with lgH = 0, the trivial products left out.
constexpr float kR8 = 0.70710677f;
__device__ inline float2 mul_neg_i(float2 t) { return make_float2(t.y, -t.x); }  // t·(-i)
__device__ inline float2 mul_pos_i(float2 t) { return make_float2(-t.y, t.x); }  // t·conj(-i)
__device__ inline void dif8_central(float2 (&x)[8]) {
    float2 t;
    t = csub(x[0], x[4]); x[0] = cadd(x[0], x[4]); x[4] = t;
    t = csub(x[1], x[5]); x[1] = cadd(x[1], x[5]); x[5] = cmul(t, make_float2(kR8, -kR8));
    t = csub(x[2], x[6]); x[2] = cadd(x[2], x[6]); x[6] = mul_neg_i(t);
    t = csub(x[3], x[7]); x[3] = cadd(x[3], x[7]); x[7] = cmul(t, make_float2(-kR8, -kR8));
#pragma unroll
    for (int b = 0; b < 8; b += 4) {
        t = csub(x[b], x[b + 2]); x[b] = cadd(x[b], x[b + 2]); x[b + 2] = t;
        t = csub(x[b + 1], x[b + 3]); x[b + 1] = cadd(x[b + 1], x[b + 3]); x[b + 3] = mul_neg_i(t);
    }
#pragma unroll
    for (int c = 0; c < 8; c += 2) {
        t = csub(x[c], x[c + 1]); x[c] = cadd(x[c], x[c + 1]); x[c + 1] = t;
    }
}
// DIT butterfly (a, b) -> (a + b·conj(w), a - b·conj(w)) given bw = b·conj(w)
__device__ inline void dit_with(float2& a, float2& b, float2 bw) {
    b = csub(a, bw);
    a = cadd(a, bw);
}
__device__ inline void dit8_central(float2 (&x)[8]) {
#pragma unroll
    for (int c = 0; c < 8; c += 2) dit_with(x[c], x[c + 1], x[c + 1]);
#pragma unroll
    for (int b = 0; b < 8; b += 4) {
        dit_with(x[b], x[b + 2], x[b + 2]);
        dit_with(x[b + 1], x[b + 3], mul_pos_i(x[b + 3]));
    }
    dit_with(x[0], x[4], x[4]);
    dit_with(x[1], x[5], cmul_conj(x[5], make_float2(kR8, -kR8)));
    dit_with(x[2], x[6], mul_pos_i(x[6]));
    dit_with(x[3], x[7], cmul_conj(x[7], make_float2(-kR8, -kR8)));
}
__device__ inline void dit4_central(float2 (&y)[4]) {
    dit_with(y[0], y[1], y[1]);
    dit_with(y[2], y[3], y[3]);
    dit_with(y[0], y[2], y[2]);
    dit_with(y[1], y[3], mul_pos_i(y[3]));
}

// The half-size inverse of the nominating transform's middle pass (see
// k_fft_mid, HALF): x holds the cross terms of the aligned run of eight at
// row position g0; makes the thread's four Y values, runs the n/2-point
// inverse stages of spans 2 .. 2^(K-1) over the workgroup's two half
// chunks (4 + 4 + 4 + ... + a last radix-2 step when K - 1 is odd) and
// stores them.  LDS holds the half chunks at h ^ (((h >> 5)·5) & 31): every
// radix-4/2 access of 32 lanes then falls on distinct 8-byte columns of a
// 256-byte bank row, but for a few of the smallest-distance steps.
template <int K, int THREADS>
__device__ inline void mid_half_inverse(float2 (&x)[8], int g0, int ca, int cb, int lgn,
                                        float2* sm, float2* z, const float2* __restrict__ tw) {
    constexpr int KH = K - 1;
    constexpr int kQuarter = 1 << KH;  // half-array elements per chunk
    float2 y[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const unsigned k = __builtin_bitreverse32((unsigned)(g0 + 2 * c)) >> (32 - lgn);
        const float2 w = tw[(1u << (lgn - 1)) + k];
        const float2 sum = cadd(x[2 * c], x[2 * c + 1]);
        const float2 dif = cmul_conj(csub(x[2 * c], x[2 * c + 1]), w);
        y[c] = make_float2(sum.x - dif.y, sum.y + dif.x);  // sum + i·dif
    }
    auto hswz = [](int h) { return h ^ (((h >> 5) * 5) & 31); };
    int hidx[4];
    {
        int lo[1];
        group_layout<KH, 2, THREADS, 4>(0, hidx, lo);  // the run of eight's four half positions
        dit4_central(y);
    }
#pragma unroll
    for (int lgH = 2; lgH < KH; lgH += 2) {
#pragma unroll
        for (int c = 0; c < 4; c++) sm[hswz(hidx[c])] = y[c];
        __syncthreads();
        if (lgH + 2 <= KH) {
            int lo[1];
            group_layout<KH, 2, THREADS, 4>(lgH, hidx, lo);
#pragma unroll
            for (int c = 0; c < 4; c++) y[c] = sm[hswz(hidx[c])];
            radix_step<2, true>(y, lo, lgH, tw);
        } else {
            int lo[2];
            group_layout<KH, 1, THREADS, 4>(lgH, hidx, lo);
#pragma unroll
            for (int c = 0; c < 4; c++) y[c] = sm[hswz(hidx[c])];
            radix_step<1, true>(y, lo, lgH, tw);
        }
    }
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const int h = hidx[c];
        store2(z, (h < kQuarter ? ca : cb) * (2 * kQuarter) + (h & (kQuarter - 1)), y[c]);
    }
}

// Stage counts of the steps before the central one (K - 3 stages in steps of
// at most three, larger first): K = 7: 2,2; 8: 3,2; 9: 3,3; 10: 3,2,2;
// 11: 3,3,2.
template <int K>
struct MidPlan {
    static constexpr int kPre = K - 3;
    static constexpr int kSteps = (kPre + 2) / 3;
    static constexpr int b(int i) { return kPre / kSteps + (i < kPre % kSteps ? 1 : 0); }
};

// Middle of the lag search (n >= TILE): the forward transform's last pass,
// the cross terms and the inverse transform's first pass in one kernel.
// Both passes work on aligned contiguous blocks of 2^K elements (K <=
// log2(TILE/2)), and the cross term of bit-reversed position i pairs it
// with 3h-1-i (h the largest power of two <= i), which maps chunk c of
// C = TILE/2 elements in [h, 2h) onto the whole chunk 2h/C - 1 - c; chunks
// 0 and 1 pair within themselves.
// Workgroup w holds the chunk pair (0, 1), or for w >= 1 with
// 2^k <= w < 2^(k+1): chunks 2^(k+1) + t and 2^(k+2) - 1 - t, t = w - 2^k.
//
// A thread holds eight elements throughout.  The forward stages run in
// radix-8/4 steps (the first loading straight from memory); the central
// step holds an aligned run of eight, runs the forward stages of distance
// 4, 2, 1, reads the mirror run (the cross-term partners of an aligned run
// of eight form another aligned run of eight) through LDS, forms the cross
// terms and runs the inverse stages of distance 1, 2, 4 in registers; the
// remaining inverse steps mirror the forward ones, the last storing straight
// to memory.  Every butterfly and twiddle is the stage-by-stage one.
// Four waves per SIMD (at most 128 VGPRs): 2048-element tiles then fit four
// workgroups per CU instead of three (a few spilled registers cost less
// than the lost occupancy); 4096-element tiles are held to two by LDS.
//
// HALF (the nominating transform, blocks of one chunk): the correlation c
// is real, so its unscaled n-point inverse is taken as an n/2-point one.
// With C the cross-term spectrum and, for k < n/2,
//   Y[k] = (C[k] + C[k + n/2]) + i·conj(W_n^k)·(C[k] - C[k + n/2]),
// the unscaled n/2-point inverse of Y is y[m] = c[2m] + i·c[2m + 1].  In
// the forward output's bit-reversed order C[k] and C[k + n/2] sit side by
// side, at positions 2p and 2p + 1 with k = rev(2p), and Y[k] belongs at
// position p of a half-length row in the bit-reversed order of n/2 points
// (rev_n(2p) = rev_{n/2}(p)), so the inverse goes on as an n/2-point DIT:
// the stages of spans 2 .. 2^(K-1) here, four elements per thread, the rest
// in the strided passes.  Chunk c's 2^(K-1) results go to the first half of
// the chunk's own storage: half position h at h + (h & ~(2^(K-1) - 1)).
template <int K, int TILE, typename C, bool HALF = false>
__global__ __launch_bounds__(TILE / 8) __attribute__((amdgpu_waves_per_eu(4))) void k_fft_mid(const LagJob* __restrict__ jobs, const C* __restrict__ tw,
                                                      int njobs, int lgn) {
    constexpr int THREADS = TILE / 8;
    constexpr int kHalfTile = TILE / 2;
    static_assert((1 << K) <= kHalfTile, "blocks of 2^K elements must fit a chunk");
    static_assert(!HALF || (1 << K) == kHalfTile, "the half-size inverse works on whole chunks");
    __shared__ C sm[TILE];
    using P = MidPlan<K>;
    const int jb = (int)(blockIdx.x % (unsigned)njobs);
    const long long wg = blockIdx.x / (unsigned)njobs;
    C* z = reinterpret_cast<C*>(jobs[jb].z);
    // row positions: 32-bit for the nominating transform (n <= 2^30)
    using Idx = typename std::conditional<sizeof(C) == 8, int, long long>::type;
    Idx ca = 0, cb = 1;
    if (wg > 0) {
        const int k = 63 - __builtin_clzll((unsigned long long)wg);
        const Idx t = (Idx)(wg - (1LL << k));
        ca = ((Idx)2 << k) + t;
        cb = ((Idx)4 << k) - 1 - t;
    }
    auto gidx = [&](int e) -> Idx { return (e < kHalfTile ? ca : cb) * kHalfTile + (e & (kHalfTile - 1)); };
    // Tile element e sits in LDS at e ^ h(e), h taking bits 0-2 from e's
    // bits 3-5 and bit 3 from bits 5 ^ 6: a permutation within each run of
    // 16 (bit 3 of the result recovers e's bit 3 from bits 5, 6, then bits
    // 0-2 from bits 3-5).  Every access of a step — eight lanes or sixteen
    // (16-byte elements over 128- or 256-byte bank rows): consecutive runs
    // (distances of 32 and up), two runs of eight 32 apart (distance 8),
    // stride 8 (the central step and its mirror reads) — then falls on
    // distinct banks.
    auto swz = [](int e) { return e ^ (((e >> 3) & 7) | ((((e >> 5) ^ (e >> 6)) & 1) << 3)); };
    C x[8];
    int idx[8];
    // The swizzle is linear over the index bits (each result bit an XOR of
    // index bits), and a radix-2^B group's positions are its first position
    // with c << lgH OR-ed into bits group_layout leaves zero, so slot c is
    // swz(first) ^ swz(c << lgH): one XOR with a constant per element.
    int sw[8];
    auto slots = [&](int B, int lgH) {
#pragma unroll
        for (int g0 = 0; g0 < 8; g0 += 1 << B) {
            const int s0 = swz(idx[g0]);
#pragma unroll
            for (int c = 0; c < (1 << B); c++) sw[g0 + c] = s0 ^ swz(c << lgH);
        }
    };
    // LDS hand-over between steps: store at the current layout, barrier,
    // load at the next.
    auto put = [&]() {
#pragma unroll
        for (int c = 0; c < 8; c++) sm[sw[c]] = x[c];
    };
    auto get = [&]() {
#pragma unroll
        for (int c = 0; c < 8; c++) x[c] = sm[sw[c]];
    };
    // forward steps before the central one
    int top = K;
#pragma unroll
    for (int i = 0; i < P::kSteps; i++) {
        const int B = P::b(i);
        const int lgH = top - B;
        top -= B;
        if (B == 3) {
            int lo[1];
            group_layout<K, 3, THREADS>(lgH, idx, lo);
            slots(3, lgH);
            if (i == 0) {
#pragma unroll
                for (int c = 0; c < 8; c++) x[c] = load2(z, gidx(idx[c]));
            } else {
                get();
            }
            radix_step<3, false>(x, lo, lgH, tw);
        } else {
            int lo[2];
            group_layout<K, 2, THREADS>(lgH, idx, lo);
            slots(2, lgH);
            if (i == 0) {
#pragma unroll
                for (int c = 0; c < 8; c++) x[c] = load2(z, gidx(idx[c]));
            } else {
                get();
            }
            radix_step<2, false>(x, lo, lgH, tw);
        }
        put();
        __syncthreads();
    }
    // central step: forward distances 4, 2, 1; cross terms; inverse 1, 2, 4
    {
        int lo[1];
        group_layout<K, 3, THREADS>(0, idx, lo);
        slots(3, 0);
        get();
        if constexpr (sizeof(C) == 8) dif8_central(x);
        else radix_step<3, false>(x, lo, 0, tw);
        put();
        __syncthreads();
        C mirror[8];
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const Idx g = gidx(idx[c]);
            Idx pg = g;
            if (g >= 2) {
                const Idx h = (Idx)1 << (63 - __builtin_clzll((unsigned long long)g));
                pg = 3 * h - 1 - g;
            }
            const int ep = ((pg / kHalfTile) == ca ? 0 : kHalfTile) + (int)(pg & (kHalfTile - 1));
            mirror[c] = sm[swz(ep)];
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 8; c++) x[c] = cross_term(x[c], mirror[c]);
        if constexpr (HALF) {
            mid_half_inverse<K, THREADS>(x, gidx(idx[0]), ca, cb, lgn, sm, z, tw);
            return;
        } else if constexpr (sizeof(C) == 8) {
            dit8_central(x);
        } else {
            radix_step<3, true>(x, lo, 0, tw);
        }
    }
    // inverse steps, mirroring the forward ones
    int bottom = 3;
#pragma unroll
    for (int i = P::kSteps - 1; i >= 0; i--) {
        const int B = P::b(i);
        const int lgH = bottom;
        bottom += B;
        put();
        __syncthreads();
        if (B == 3) {
            int lo[1];
            group_layout<K, 3, THREADS>(lgH, idx, lo);
            slots(3, lgH);
            get();
            radix_step<3, true>(x, lo, lgH, tw);
        } else {
            int lo[2];
            group_layout<K, 2, THREADS>(lgH, idx, lo);
            slots(2, lgH);
            get();
            radix_step<2, true>(x, lo, lgH, tw);
        }
    }
#pragma unroll
    for (int c = 0; c < 8; c++) store2(z, gidx(idx[c]), x[c]);
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
constexpr int kMaxFuse = 8;        // stages per pass below one full tile
constexpr int kBigTile = 4096;     // complex doubles per workgroup (64 KiB LDS, 512 threads)
constexpr int kMinStrided = 4;     // strided passes: 4..6 stages on 2048-element tiles (runs of >= 32
constexpr int kMaxSmallStrided = 6;  // residues, 512 B), 7..9 on 4096-element tiles (>= 8 residues, 128 B)
constexpr int kMaxStrided = 9;
constexpr int kMinLast = 7;        // the contiguous last pass: 7..10 stages on 2048-element tiles,
constexpr int kMaxSmallLast = 10;  // 11 on 4096-element tiles
constexpr int kMaxLast = 11;

// Stage counts per pass, the pass over the largest spans first.  Below one
// full tile: as even as possible.  With full tiles: the fewest passes
// (strided passes of at most 9 stages, the last pass of at most 11).  A
// strided pass of up to 6 stages on 2048-element tiles costs about the
// same whatever its stage count (HBM-bound, runs of >= 512 B), one of 7-9
// stages on 4096-element tiles reads shorter runs and costs up to ~1.7x
// that, while the fused last pass grows slowly with its stages: so the
// strided passes take up to 6 stages each, the last pass the rest up to 11,
// and only what is left beyond that goes to wider strided passes; split as
// evenly as possible, smaller first.
int plan_passes(long long n, int* ks) {
    int L = 0;
    while ((1LL << L) < n) L++;
    if (n < kTile) {
        const int np = (L + kMaxFuse - 1) / kMaxFuse;
        for (int p = 0; p < np; p++) ks[p] = L / np + (p >= np - L % np ? 1 : 0);
        return np;
    }
    int ns = (L - kMaxLast + kMaxStrided - 1) / kMaxStrided;
    if (ns < 1) ns = 1;
    int T = kMaxSmallStrided * ns < L - kMinLast ? kMaxSmallStrided * ns : L - kMinLast;
    if (T < kMinStrided * ns) T = kMinStrided * ns;
    if (L - T > kMaxLast) T = L - kMaxLast;
    for (int p = 0; p < ns; p++) ks[p] = T / ns + (p >= ns - T % ns ? 1 : 0);
    ks[ns] = L - T;
    return ns + 1;
}

int log2i(long long v) {
    int l = 0;
    while ((1LL << l) < v) l++;
    return l;
}

// Tile of a full-tile pass of k stages over n elements.
int tile_for(long long n, int k) { return k > kMaxSmallStrided && k <= kMaxStrided && n >= kBigTile ? kBigTile : kTile; }

struct PassShape {
    int lgR, lgBn;
    long long wgs;
};

// R residues (<= q) and Bn span blocks (<= n / S_hi) filling one tile.
PassShape pass_shape(long long n, long long S_hi, int k, int tile) {
    const int lgn = log2i(n), lgS = log2i(S_hi), lgq = lgS - k;
    const int lgWant = log2i(tile) - k;
    const int lgR = lgWant < lgq ? lgWant : lgq;
    int lgBn = lgWant - lgR;
    if (lgBn > lgn - lgS) lgBn = lgn - lgS;
    return PassShape{lgR, lgBn, 1LL << ((lgn - lgS - lgBn) + (lgq - lgR))};
}

template <int K, bool DIT, int SRC, bool ARGMAX, int TILE>
void tile_pass(const PassShape& s, const LagJob* d_jobs, int njobs, long long n, long long S_hi, const double2* tw,
               const double* sums, unsigned long long* partials, hipStream_t stream) {
    hipLaunchKernelGGL((k_fft_tile<K, DIT, SRC, ARGMAX ? kEpiArgmax : kEpiStore, TILE, double2>),
                       dim3((unsigned)(s.wgs * njobs)), dim3(TILE / 8), 0, stream, d_jobs, log2i(n), log2i(S_hi), s.lgR,
                       s.lgBn, tw, sums, (void*)partials, (const float*)nullptr, njobs);
}

template <bool DIT, int SRC, bool ARGMAX>
void pass(const LagJob* d_jobs, int njobs, long long n, long long S_hi, int k, const double2* tw, const double* sums,
          unsigned long long* partials, hipStream_t stream) {
    const int tile = tile_for(n, k);
    const PassShape s = pass_shape(n, S_hi, k, tile);
    if (n >= tile && k >= kMinStrided && k <= kMaxSmallLast) {  // every tile full
        if (tile == kBigTile) {
            switch (k) {
                case 7: tile_pass<7, DIT, SRC, ARGMAX, kBigTile>(s, d_jobs, njobs, n, S_hi, tw, sums, partials, stream); return;
                case 8: tile_pass<8, DIT, SRC, ARGMAX, kBigTile>(s, d_jobs, njobs, n, S_hi, tw, sums, partials, stream); return;
                default: tile_pass<9, DIT, SRC, ARGMAX, kBigTile>(s, d_jobs, njobs, n, S_hi, tw, sums, partials, stream); return;
            }
        }
        switch (k) {
            case 4: tile_pass<4, DIT, SRC, ARGMAX, kTile>(s, d_jobs, njobs, n, S_hi, tw, sums, partials, stream); return;
            case 5: tile_pass<5, DIT, SRC, ARGMAX, kTile>(s, d_jobs, njobs, n, S_hi, tw, sums, partials, stream); return;
            case 6: tile_pass<6, DIT, SRC, ARGMAX, kTile>(s, d_jobs, njobs, n, S_hi, tw, sums, partials, stream); return;
            case 7: tile_pass<7, DIT, SRC, ARGMAX, kTile>(s, d_jobs, njobs, n, S_hi, tw, sums, partials, stream); return;
            case 8: tile_pass<8, DIT, SRC, ARGMAX, kTile>(s, d_jobs, njobs, n, S_hi, tw, sums, partials, stream); return;
            case 9: tile_pass<9, DIT, SRC, ARGMAX, kTile>(s, d_jobs, njobs, n, S_hi, tw, sums, partials, stream); return;
            default: tile_pass<10, DIT, SRC, ARGMAX, kTile>(s, d_jobs, njobs, n, S_hi, tw, sums, partials, stream); return;
        }
    }
    hipLaunchKernelGGL((k_fft_pass<DIT, SRC, ARGMAX>), dim3((unsigned)(s.wgs * njobs)), dim3(kThreads), 0, stream, d_jobs,
                       log2i(n), log2i(S_hi), k, s.lgR, s.lgBn, tw, sums, partials, njobs);
}
}  // namespace

long long lag_partials_per_job(long long n) {
    if (n < 2) return 1;
    int ks[8];
    plan_passes(n, ks);
    return pass_shape(n, n, ks[0], tile_for(n, ks[0])).wgs;  // the last DIT pass covers the spans up to n
}

void launch_lag_search(const LagJob* d_jobs, int njobs, long long n, const double2* d_tw, const double* d_sums,
                       unsigned long long* d_partials, long long* d_lags, hipStream_t stream, bool keep_xcorr) {
    if (njobs == 0) return;
    if (n < 2) {
        // n == 1: one sample per side, lag 0.
        if (keep_xcorr) return;
        hipLaunchKernelGGL(k_lag_reduce, dim3(njobs), dim3(256), 0, stream, d_jobs, (const unsigned long long*)d_partials,
                           0, d_lags);
        return;
    }
    int ks[8];
    const int np = plan_passes(n, ks);
    // Fused middle for full tiles: the last forward pass, the cross terms and
    // the first inverse pass in one kernel (np >= 2 whenever n >= kTile).
    const bool fuse = n >= kTile && np >= 2 && ks[np - 1] >= kMinLast && ks[np - 1] <= kMaxLast;
    long long S = n;
    for (int p = 0; p < (fuse ? np - 1 : np); p++) {
        if (p == 0) pass<false, kSrcEnvelope, false>(d_jobs, njobs, n, S, ks[p], d_tw, d_sums, nullptr, stream);
        else pass<false, kSrcZ, false>(d_jobs, njobs, n, S, ks[p], d_tw, d_sums, nullptr, stream);
        S >>= ks[p];
    }
    if (fuse) {
        const dim3 g((unsigned)((n / kTile) * njobs));
        switch (ks[np - 1]) {
            case 7: hipLaunchKernelGGL((k_fft_mid<7, kTile, double2>), g, dim3(kThreads), 0, stream, d_jobs, d_tw, njobs, 0); break;
            case 8: hipLaunchKernelGGL((k_fft_mid<8, kTile, double2>), g, dim3(kThreads), 0, stream, d_jobs, d_tw, njobs, 0); break;
            case 9: hipLaunchKernelGGL((k_fft_mid<9, kTile, double2>), g, dim3(kThreads), 0, stream, d_jobs, d_tw, njobs, 0); break;
            case 10: hipLaunchKernelGGL((k_fft_mid<10, kTile, double2>), g, dim3(kThreads), 0, stream, d_jobs, d_tw, njobs, 0); break;
            default:
                hipLaunchKernelGGL((k_fft_mid<11, kBigTile, double2>), dim3((unsigned)((n / kBigTile) * njobs)),
                                   dim3(kBigTile / 8), 0, stream, d_jobs, d_tw, njobs, 0);
                break;
        }
    } else {
        hipLaunchKernelGGL(k_cross_terms, dim3((unsigned)((n + 255) / 256), njobs), dim3(256), 0, stream, d_jobs, n);
    }
    S = fuse ? (1LL << ks[np - 1]) : 1;
    for (int p = fuse ? np - 2 : np - 1; p >= 0; p--) {
        S <<= ks[p];
        if (S == n && !keep_xcorr) pass<true, kSrcZ, true>(d_jobs, njobs, n, S, ks[p], d_tw, d_sums, d_partials, stream);
        else pass<true, kSrcZ, false>(d_jobs, njobs, n, S, ks[p], d_tw, d_sums, nullptr, stream);
    }
    if (keep_xcorr) return;
    const long long parts = pass_shape(n, n, ks[0], tile_for(n, ks[0])).wgs;
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

// ---------------------------------------------------------------- nominating (f32) lag search
// The throughput mode's lag search (xcorr.zig:269-340 refineLag and rescore,
// engine.zig:800-822): the same packed-envelope correlation computed in f32
// nominates every lag whose |Re| lies within the f32 error bound of the
// row's maximum; those are rescored by f64 dot products of the f64
// envelopes and the first maximum in scan order wins.  The transform is not
// bound to the exact path's operation order, so its passes use 8192-element
// f32 tiles: one strided DIF pass (two past 2^21 points) from the signals,
// the fused middle on 4096-element blocks, and the mirrored DIT passes, the
// last one keeping only each tile's candidates.
namespace {
constexpr int kTile32 = 8192;   // float2 per workgroup (64 KiB LDS, 1024 threads)
constexpr int kMid32 = 12;      // stages of the fused middle (blocks of 4096)
static_assert(kHalfChunk32 == (1 << (kMid32 - 1)), "half-length chunks are the middle blocks' halves");
constexpr int kMaxStrided32 = 9;
constexpr int kRescoreThreads = 256;
constexpr int kRescorePer = 16;
constexpr long long kRescoreSeg = (long long)kRescoreThreads * kRescorePer;  // t values per rescoring block

// Strided stage counts (over the largest spans first) of an n-point
// nominating transform; returns the number of strided passes (1 or 2), or
// 0 when n is outside the range this path handles.
int plan32(long long n, int* ks) {
    const int L = log2i(n);
    if ((1LL << L) != n || L < kMid32 + 2 || L > kMid32 + 2 * kMaxStrided32) return 0;
    const int T = L - kMid32;
    if (T <= kMaxStrided32) {
        ks[0] = T;
        return 1;
    }
    ks[0] = (T + 1) / 2;
    ks[1] = T - ks[0];
    return 2;
}

template <int K, bool DIT, int SRC, int EPI>
void tile32(const LagJob* d_jobs, int njobs, long long n, long long S_hi, const float2* tw, const double* sums, void* aux,
            const float* slack, hipStream_t stream) {
    const PassShape s = pass_shape(n, S_hi, K, kTile32);
    hipLaunchKernelGGL((k_fft_tile<K, DIT, SRC, EPI, kTile32, float2>), dim3((unsigned)(s.wgs * njobs)),
                       dim3(kTile32 / 8), 0, stream, d_jobs, log2i(n), log2i(S_hi), s.lgR, s.lgBn, tw, sums, aux, slack,
                       njobs);
}

template <bool DIT, int SRC, int EPI>
void pass32(int k, const LagJob* d_jobs, int njobs, long long n, long long S_hi, const float2* tw, const double* sums,
            void* aux, const float* slack, hipStream_t stream) {
    switch (k) {
        case 2: tile32<2, DIT, SRC, EPI>(d_jobs, njobs, n, S_hi, tw, sums, aux, slack, stream); return;
        case 3: tile32<3, DIT, SRC, EPI>(d_jobs, njobs, n, S_hi, tw, sums, aux, slack, stream); return;
        case 4: tile32<4, DIT, SRC, EPI>(d_jobs, njobs, n, S_hi, tw, sums, aux, slack, stream); return;
        case 5: tile32<5, DIT, SRC, EPI>(d_jobs, njobs, n, S_hi, tw, sums, aux, slack, stream); return;
        case 6: tile32<6, DIT, SRC, EPI>(d_jobs, njobs, n, S_hi, tw, sums, aux, slack, stream); return;
        case 7: tile32<7, DIT, SRC, EPI>(d_jobs, njobs, n, S_hi, tw, sums, aux, slack, stream); return;
        case 8: tile32<8, DIT, SRC, EPI>(d_jobs, njobs, n, S_hi, tw, sums, aux, slack, stream); return;
        default: tile32<9, DIT, SRC, EPI>(d_jobs, njobs, n, S_hi, tw, sums, aux, slack, stream); return;
    }
}

// Scratch of one launch, carved in this order (each part 256-byte aligned).
struct Lag32Layout {
    long long wg_first, wg_last, nseg;
    size_t energy, slack, tiles, count, cands, dots, total;
};
size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
Lag32Layout lag32_layout(int njobs, long long n) {
    int ks[2];
    const int np = plan32(n, ks);
    Lag32Layout l{};
    if (np == 0) return l;
    l.wg_first = pass_shape(n, n, ks[0], kTile32).wgs;
    l.wg_last = pass_shape(n / 2, n / 2, ks[0], kTile32).wgs;  // over the half-length inverse
    l.nseg = (n / 2 + kRescoreSeg - 1) / kRescoreSeg;
    size_t at = 0;
    l.energy = at;
    at += align256(sizeof(double) * njobs * l.wg_first);
    l.slack = at;
    at += align256(sizeof(float) * njobs);
    l.tiles = at;
    at += align256(sizeof(TileCands) * njobs * l.wg_last);
    l.count = at;
    at += align256(sizeof(int) * njobs);
    l.cands = at;
    at += align256(sizeof(long long) * njobs * kMaxLagCands);
    l.dots = at;
    at += align256(sizeof(double) * njobs * kMaxLagCands * l.nseg);
    l.total = at;
    return l;
}
}  // namespace

// Per row: the f32 slack 2·tol·(‖env_a‖² + ‖env_b‖²)·n with tol =
// 8·eps_f32·log2 n (engine.zig:819-822 lagTolerance, xcorr.zig:320-326
// refineSlack; the correlation here carries the factor n of the unscaled
// inverse transform), the energy summed from the first pass's tile
// partials in tile order.
__global__ __launch_bounds__(64) void k_lag32_slack(const double* __restrict__ energy, int nparts, int lgn,
                                                    float* __restrict__ slack) {
    // the row's energy: lane l sums partials l, l + 64, ..., then a fixed
    // shuffle tree (the same order on every run, whatever the batch)
    const double* e = energy + (long long)blockIdx.x * nparts;
    double acc = 0.0;
    for (int p = (int)threadIdx.x; p < nparts; p += 64) acc += e[p];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if (threadIdx.x != 0) return;
    const double tol = 8.0 * (double)FLT_EPSILON * (double)lgn;
    slack[blockIdx.x] = (float)(2.0 * tol * acc * (double)(1LL << lgn));
}

// Per row: the row's top (the largest tile top), floor = top - slack in
// f32, and every in-range lag at or above it, gathered from the tiles whose
// own top reaches the floor and sorted into scan order.  count[row] = -1
// when the row is not nominated: a tile that reaches the floor overflowed
// its list, more than cap lags qualify, or the top is not finite.
__global__ __launch_bounds__(256) void k_lag32_select(const LagJob* __restrict__ jobs, const TileCands* __restrict__ tiles,
                                                      int ntiles, const float* __restrict__ slack, int cap,
                                                      int* __restrict__ count, long long* __restrict__ cands) {
    const int row = blockIdx.x;
    const TileCands* t = tiles + (long long)row * ntiles;
    __shared__ unsigned s_top;
    __shared__ int s_n, s_bad;
    __shared__ int s_pos[kMaxLagCands];
    if (threadIdx.x == 0) {
        s_top = 0;
        s_n = 0;
        s_bad = 0;
    }
    __syncthreads();
    unsigned top = 0;
    for (int w = threadIdx.x; w < ntiles; w += blockDim.x) top = max(top, __float_as_uint(t[w].top));
    atomicMax(&s_top, top);
    __syncthreads();
    const float row_top = __uint_as_float(s_top);
    const float floor = row_top - slack[row];
    if (!isfinite(row_top) || !isfinite(floor)) {
        if (threadIdx.x == 0) count[row] = -1;
        return;
    }
    for (int w = threadIdx.x; w < ntiles; w += blockDim.x) {
        if (!(t[w].top >= floor)) continue;
        if (t[w].count > kTileCands) {
            atomicExch(&s_bad, 1);
            continue;
        }
        for (int c = 0; c < t[w].count; c++) {
            if (!(t[w].val[c] >= floor)) continue;
            const int slot = atomicAdd(&s_n, 1);
            if (slot < kMaxLagCands) s_pos[slot] = t[w].pos[c];
        }
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    const int nc = s_n;
    if (s_bad || nc > cap || nc == 0) {
        count[row] = -1;
        return;
    }
    // insertion sort into scan order (at most kMaxLagCands entries)
    for (int i = 1; i < nc; i++) {
        const int v = s_pos[i];
        int j = i - 1;
        while (j >= 0 && s_pos[j] > v) {
            s_pos[j + 1] = s_pos[j];
            j--;
        }
        s_pos[j + 1] = v;
    }
    const LagJob j = jobs[row];
    const long long max_lag = (j.la > j.lb ? j.la : j.lb) - 1;
    for (int i = 0; i < nc; i++) cands[(long long)row * kMaxLagCands + i] = (long long)s_pos[i] - max_lag;
    count[row] = nc;
}

// Partial f64 dot products Σ_t env_a[t + lag]·env_b[t] (xcorr.zig:328-340
// dotAtLag) of every candidate of a row over one segment of t, with the
// envelopes formed in f64 from the signals as the transform's first pass
// forms them.  Fixed segments, per-thread order and reduction tree: the
// same bits for a row whatever the batch.
__global__ __launch_bounds__(kRescoreThreads) void k_lag32_rescore(const LagJob* __restrict__ jobs,
                                                                   const double* __restrict__ sums,
                                                                   const int* __restrict__ count,
                                                                   const long long* __restrict__ cands, int nseg,
                                                                   double* __restrict__ dots) {
    const int row = blockIdx.y;
    const int nc = count[row];
    if (nc <= 1) return;  // no row, or a lone nominee (k_lag32_pick takes it as is)
    const LagJob j = jobs[row];
    const long long t0 = (long long)blockIdx.x * kRescoreSeg;
    if (t0 >= j.lb) {
        if (threadIdx.x < nc) dots[((long long)row * kMaxLagCands + threadIdx.x) * nseg + blockIdx.x] = 0.0;
        return;
    }
    const EnvSide ea = env_side(j.a, j.la, sums[j.sum_a], sums[j.nyq_a]);
    const EnvSide eb = env_side(j.b, j.lb, sums[j.sum_b], sums[j.nyq_b]);
    double vb[kRescorePer];
#pragma unroll
    for (int u = 0; u < kRescorePer; u++) {
        const long long t = t0 + u * kRescoreThreads + threadIdx.x;
        vb[u] = t < eb.len ? env_value(eb, env_load(eb, t), t) : 0.0;
    }
    __shared__ double red[kRescoreThreads / 64];
    for (int c = 0; c < nc; c++) {
        const long long lag = cands[(long long)row * kMaxLagCands + c];
        double acc = 0.0;
#pragma unroll
        for (int u = 0; u < kRescorePer; u++) {
            const long long t = t0 + u * kRescoreThreads + threadIdx.x;
            const long long i = t + lag;
            const bool in = t < eb.len && i >= 0 && i < ea.len;
            const double va = env_value(ea, env_load(ea, in ? i : 0), i);
            acc = __builtin_fma(in ? va : 0.0, vb[u], acc);
        }
        const double d = block_sum<kRescoreThreads>(acc, red);
        if (threadIdx.x == 0) dots[((long long)row * kMaxLagCands + c) * nseg + blockIdx.x] = d;
        __syncthreads();
    }
}

// Per row: each candidate's dot summed over the segments in order, then the
// first maximum of |dot| in scan order (xcorr.zig:313-326 rescore).
__global__ __launch_bounds__(kMaxLagCands) void k_lag32_pick(const int* __restrict__ count,
                                                             const long long* __restrict__ cands,
                                                             const double* __restrict__ dots, int nseg,
                                                             long long* __restrict__ lags, int* __restrict__ fallback) {
    const int row = blockIdx.x;
    const int nc = count[row];
    __shared__ double s_dot[kMaxLagCands];
    if (nc <= 0) {
        if (threadIdx.x == 0) fallback[row] = 1;
        return;
    }
    if (nc == 1) {
        // the row's argmax is always nominated (|error| <= slack / 2), so a
        // lone nominee is it; rescoring one candidate picks it too
        if (threadIdx.x == 0) {
            lags[row] = cands[(long long)row * kMaxLagCands];
            fallback[row] = 0;
        }
        return;
    }
    if ((int)threadIdx.x < nc) {
        const double* d = dots + ((long long)row * kMaxLagCands + threadIdx.x) * nseg;
        double acc = 0.0;
        for (int g = 0; g < nseg; g++) acc += d[g];
        s_dot[threadIdx.x] = fabs(acc);
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    int best = 0;
    for (int c = 1; c < nc; c++)
        if (s_dot[c] > s_dot[best]) best = c;
    lags[row] = cands[(long long)row * kMaxLagCands + best];
    fallback[row] = 0;
}

bool lag32_handles(long long n) {
    int ks[2];
    return plan32(n, ks) != 0;
}

size_t lag32_scratch_bytes(int njobs, long long n) { return lag32_layout(njobs, n).total; }

void launch_lag_search32(const LagJob* d_jobs, int njobs, long long n, const float2* d_tw, const double* d_sums,
                         void* d_scratch, int cap, long long* d_lags, int* d_fallback, hipStream_t stream,
                         bool keep_xcorr) {
    if (njobs == 0) return;
    int ks[2];
    const int np = plan32(n, ks);
    const Lag32Layout l = lag32_layout(njobs, n);
    char* base = static_cast<char*>(d_scratch);
    double* energy = reinterpret_cast<double*>(base + l.energy);
    float* slack = reinterpret_cast<float*>(base + l.slack);
    TileCands* tiles = reinterpret_cast<TileCands*>(base + l.tiles);
    int* count = reinterpret_cast<int*>(base + l.count);
    long long* cands = reinterpret_cast<long long*>(base + l.cands);
    double* dots = reinterpret_cast<double*>(base + l.dots);
    const int lgn = log2i(n);
    // forward strided pass(es), the first from the signals
    pass32<false, kSrcEnvelope, kEpiStore>(ks[0], d_jobs, njobs, n, n, d_tw, d_sums, energy, nullptr, stream);
    if (np == 2) pass32<false, kSrcZ, kEpiStore>(ks[1], d_jobs, njobs, n, n >> ks[0], d_tw, d_sums, nullptr, nullptr, stream);
    hipLaunchKernelGGL(k_lag32_slack, dim3(njobs), dim3(64), 0, stream, (const double*)energy, (int)l.wg_first, lgn,
                       slack);
    // forward middle, cross terms, and the first stages of the half-length inverse
    hipLaunchKernelGGL((k_fft_mid<kMid32, kTile32, float2, true>), dim3((unsigned)((n / kTile32) * njobs)),
                       dim3(kTile32 / 8), 0, stream, d_jobs, d_tw, njobs, lgn);
    // the half-length inverse's strided passes (the same stage counts)
    const long long h = n / 2;
    if (np == 2)
        pass32<true, kSrcHalf, kEpiStore>(ks[1], d_jobs, njobs, h, h >> ks[0], d_tw, d_sums, nullptr, nullptr, stream);
    if (keep_xcorr) {
        pass32<true, kSrcHalf, kEpiStore>(ks[0], d_jobs, njobs, h, h, d_tw, d_sums, nullptr, nullptr, stream);
        return;
    }
    pass32<true, kSrcHalf, kEpiCandidates>(ks[0], d_jobs, njobs, h, h, d_tw, d_sums, tiles, slack, stream);
    hipLaunchKernelGGL(k_lag32_select, dim3(njobs), dim3(256), 0, stream, d_jobs, (const TileCands*)tiles,
                       (int)l.wg_last, (const float*)slack, cap < kMaxLagCands ? cap : kMaxLagCands, count, cands);
    hipLaunchKernelGGL(k_lag32_rescore, dim3((unsigned)l.nseg, njobs), dim3(kRescoreThreads), 0, stream, d_jobs,
                       d_sums, (const int*)count, (const long long*)cands, (int)l.nseg, dots);
    hipLaunchKernelGGL(k_lag32_pick, dim3(njobs), dim3(kMaxLagCands), 0, stream, (const int*)count,
                       (const long long*)cands, (const double*)dots, (int)l.nseg, d_lags, d_fallback);
}

}  // namespace vq
