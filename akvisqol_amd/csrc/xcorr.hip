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

enum : int { kSrcZ = 0, kSrcEnvelope = 1 };

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
__device__ inline void bf_dif(double2& a, double2& b, double2 w) {
    const double2 t = csub(a, b);
    a = cadd(a, b);
    b = cmul(t, w);
}
// DIT butterfly with the conjugate twiddle (fft.zig:92-166).
__device__ inline void bf_dit(double2& a, double2& b, double2 w) {
    const double2 bw = cmul_conj(b, w);
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
template <int K, bool DIT, int THREADS>
__device__ inline void tile_stages(double2* sm, int lgR, int lgq, long long r0, const double2* __restrict__ tw) {
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
    auto twi = [&](int l, int p, int rr) -> double2 {
        const long long j = ((long long)p << lgq) + r0 + rr;
        return tw[(1LL << (lgq + K - l - 1)) + j];
    };
    constexpr int kSteps = (K + 1) / 2;
    // w[u][0..1]: the larger-distance stage at lo and lo + h; w[u][2]: the
    // smaller-distance stage at lo.  Radix-2 step: w[u][0] for butterfly u.
    double2 w[4][3];
    auto load_tw = [&](int st, double2 (&dst)[4][3]) {
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
        double2 wn[4][3];
        if (li + 1 < K) {
            const int lgH = DIT ? li : K - 2 - li;
            double2 x[2][4];
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
            double2 x[4][2];
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
// stages: load the tile, run tile_stages, store it (or, on the inverse
// transform's last pass, keep only the tile's first maximum).

// At least two workgroups per CU (TILE / 1024 waves per SIMD): for
// 4096-element tiles (8 waves each) at most 128 VGPRs.
template <int K, bool DIT, int SRC, bool ARGMAX, int TILE>
__global__ __launch_bounds__(TILE / 8) __attribute__((amdgpu_waves_per_eu(TILE / 1024))) void k_fft_tile(const LagJob* __restrict__ jobs, int lgn, int lgS, int lgR, int lgBn,
                                                       const double2* __restrict__ tw, const double* __restrict__ sums,
                                                       unsigned long long* __restrict__ partials, int njobs) {
    constexpr int THREADS = TILE / 8;
    __shared__ double2 sm[TILE];
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
    const long long base = (blk0 << lgS) + r0;
    double2* z = job.z;
    auto gidx = [&](int e) -> long long {
        const int bb = e >> lgGR;
        const int m = (e >> lgR) & ((1 << K) - 1);
        const int rr = e & ((1 << lgR) - 1);
        return base + ((long long)bb << lgS) + ((long long)m << lgq) + rr;
    };
    constexpr int kPer = 8;
    if (SRC == kSrcEnvelope) {
        // both signals' samples for the tile, loads first, envelope after
        const EnvSide ea = env_side(job.a, job.la, sums[job.sum_a], sums[job.nyq_a]);
        const EnvSide eb = env_side(job.b, job.lb, sums[job.sum_b], sums[job.nyq_b]);
        double xa[kPer], xb[kPer];
#pragma unroll
        for (int u = 0; u < kPer; u++) {
            const long long i = gidx(threadIdx.x + u * THREADS);
            xa[u] = env_load(ea, i);
            xb[u] = env_load(eb, i);
        }
#pragma unroll
        for (int u = 0; u < kPer; u++) {
            const long long i = gidx(threadIdx.x + u * THREADS);
            sm[threadIdx.x + u * THREADS] = make_double2(i < ea.len ? env_value(ea, xa[u], i) : 0.0,
                                                          i < eb.len ? env_value(eb, xb[u], i) : 0.0);
        }
    } else {
        double2 v[kPer];
#pragma unroll
        for (int u = 0; u < kPer; u++) v[u] = load2(z, gidx(threadIdx.x + u * THREADS));
#pragma unroll
        for (int u = 0; u < kPer; u++) sm[threadIdx.x + u * THREADS] = v[u];
    }

    tile_stages<K, DIT, THREADS>(sm, lgR, lgq, r0, tw);
    if (!ARGMAX) {
#pragma unroll
        for (int u = 0; u < kPer; u++) {
            const int e = threadIdx.x + u * THREADS;
            store2(z, gidx(e), sm[e]);
        }
        return;
    }
    tile_argmax<THREADS>(sm, TILE, gidx, n, job, partials + 2 * (jb * nwg + wg));
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

// One radix-2^B step (B = 1, 2, 3) over the eight elements a thread holds in
// a pass over contiguous blocks (one residue): 8 >> B groups, element c of
// group gi in x[gi·2^B + c] at block position m0 + c·2^lgH.  The stage with
// distance d·2^lgH (d = 2^s, s < B) pairs c with c + d (bit s of c clear)
// and takes W_{2d·2^lgH} at position lo + (c mod d)·2^lgH, i.e. table index
// (d << lgH) + lo + ((c mod d) << lgH).  DIF runs the stages largest
// distance first, DIT smallest first: the radix-2 butterflies of the
// stage-by-stage transform.
template <int B, bool DIT>
__device__ inline void radix_step(double2 (&x)[8], const int (&lo)[8 >> B], int lgH, const double2* __restrict__ tw) {
    constexpr int G = 1 << B;
    constexpr int NG = 8 >> B;
    double2 w[NG][G];  // w[gi][d + t]: stage distance d, position t (t < d)
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
                double2& a = x[gi * G + c];
                double2& b = x[gi * G + c + d];
                if (DIT) bf_dit(a, b, w[gi][d + (c & (d - 1))]);
                else bf_dif(a, b, w[gi][d + (c & (d - 1))]);
            }
    }
}

// Tile positions (block bb of 2^K elements, position m) of the thread's
// eight elements for a radix-2^B step with smallest distance 2^lgH: group
// G = thread + gi·THREADS, q = G mod 2^(K-B) its index in the block, lo = q
// mod 2^lgH, m0 = q with B zero bits inserted at bit lgH.
template <int K, int B, int THREADS>
__device__ inline void group_layout(int lgH, int (&idx)[8], int (&lo)[8 >> B]) {
#pragma unroll
    for (int gi = 0; gi < (8 >> B); gi++) {
        const int g = (int)threadIdx.x + gi * THREADS;
        const int q = g & ((1 << (K - B)) - 1);
        const int bb = g >> (K - B);
        lo[gi] = q & ((1 << lgH) - 1);
        const int m0 = ((q >> lgH) << (lgH + B)) | lo[gi];
#pragma unroll
        for (int c = 0; c < (1 << B); c++) idx[gi * (1 << B) + c] = (bb << K) + m0 + (c << lgH);
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
template <int K, int TILE>
__global__ __launch_bounds__(TILE / 8) __attribute__((amdgpu_waves_per_eu(4))) void k_fft_mid(const LagJob* __restrict__ jobs, const double2* __restrict__ tw,
                                                      int njobs) {
    constexpr int THREADS = TILE / 8;
    constexpr int kHalfTile = TILE / 2;
    static_assert((1 << K) <= kHalfTile, "blocks of 2^K elements must fit a chunk");
    __shared__ double2 sm[TILE];
    using P = MidPlan<K>;
    const int jb = (int)(blockIdx.x % (unsigned)njobs);
    const long long wg = blockIdx.x / (unsigned)njobs;
    double2* z = jobs[jb].z;
    long long ca = 0, cb = 1;
    if (wg > 0) {
        const int k = 63 - __builtin_clzll((unsigned long long)wg);
        const long long t = wg - (1LL << k);
        ca = (2LL << k) + t;
        cb = (4LL << k) - 1 - t;
    }
    auto gidx = [&](int e) -> long long { return (e < kHalfTile ? ca : cb) * kHalfTile + (e & (kHalfTile - 1)); };
    // Tile element e sits in LDS at e ^ h(e), h taking bits 0-2 from e's
    // bits 3-5 and bit 3 from bits 5 ^ 6: a permutation within each run of
    // 16 (bit 3 of the result recovers e's bit 3 from bits 5, 6, then bits
    // 0-2 from bits 3-5).  Every access of a step — eight lanes or sixteen
    // (16-byte elements over 128- or 256-byte bank rows): consecutive runs
    // (distances of 32 and up), two runs of eight 32 apart (distance 8),
    // stride 8 (the central step and its mirror reads) — then falls on
    // distinct banks.
    auto swz = [](int e) { return e ^ (((e >> 3) & 7) | ((((e >> 5) ^ (e >> 6)) & 1) << 3)); };
    double2 x[8];
    int idx[8];
    // LDS hand-over between steps: store at the current layout, barrier,
    // load at the next.
    auto put = [&]() {
#pragma unroll
        for (int c = 0; c < 8; c++) sm[swz(idx[c])] = x[c];
    };
    auto get = [&]() {
#pragma unroll
        for (int c = 0; c < 8; c++) x[c] = sm[swz(idx[c])];
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
        get();
        radix_step<3, false>(x, lo, 0, tw);
        put();
        __syncthreads();
        double2 mirror[8];
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const long long g = gidx(idx[c]);
            long long pg = g;
            if (g >= 2) {
                const long long h = 1LL << (63 - __builtin_clzll((unsigned long long)g));
                pg = 3 * h - 1 - g;
            }
            const int ep = ((pg / kHalfTile) == ca ? 0 : kHalfTile) + (int)(pg & (kHalfTile - 1));
            mirror[c] = sm[swz(ep)];
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 8; c++) x[c] = cross_term(x[c], mirror[c]);
        radix_step<3, true>(x, lo, 0, tw);
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
            get();
            radix_step<3, true>(x, lo, lgH, tw);
        } else {
            int lo[2];
            group_layout<K, 2, THREADS>(lgH, idx, lo);
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
    hipLaunchKernelGGL((k_fft_tile<K, DIT, SRC, ARGMAX, TILE>), dim3((unsigned)(s.wgs * njobs)), dim3(TILE / 8), 0,
                       stream, d_jobs, log2i(n), log2i(S_hi), s.lgR, s.lgBn, tw, sums, partials, njobs);
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
            case 7: hipLaunchKernelGGL((k_fft_mid<7, kTile>), g, dim3(kThreads), 0, stream, d_jobs, d_tw, njobs); break;
            case 8: hipLaunchKernelGGL((k_fft_mid<8, kTile>), g, dim3(kThreads), 0, stream, d_jobs, d_tw, njobs); break;
            case 9: hipLaunchKernelGGL((k_fft_mid<9, kTile>), g, dim3(kThreads), 0, stream, d_jobs, d_tw, njobs); break;
            case 10: hipLaunchKernelGGL((k_fft_mid<10, kTile>), g, dim3(kThreads), 0, stream, d_jobs, d_tw, njobs); break;
            default:
                hipLaunchKernelGGL((k_fft_mid<11, kBigTile>), dim3((unsigned)((n / kBigTile) * njobs)),
                                   dim3(kBigTile / 8), 0, stream, d_jobs, d_tw, njobs);
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

}  // namespace vq
