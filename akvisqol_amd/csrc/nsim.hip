// Spectrogram preparation, NSIM similarity table, DP patch search and
// per-band patch statistics on gfx950.
//
// References: spectrogram.zig:14-80 (dB / floors / joint minimum),
// nsim.zig:119-510 (taps, DegPlane, measureSimilarityAt,
// measurePatchSimilarity), dp.zig:205-355 (DP + backtrack),
// accel.zig:126-146 (table over the search windows).
#include <float.h>
#include <limits.h>

#include <type_traits>

#include "dev_common.h"
#include "kernels.h"

namespace vq {

// ---------------------------------------------------------------- dB prep
__global__ __launch_bounds__(256) void k_db(const PrepJob* __restrict__ jobs, int nb) {
    const PrepJob j = jobs[blockIdx.y];
    const long long nr = (long long)nb * j.nfr, nd = (long long)nb * j.nfd;
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nr + nd) return;
    double* p = e < nr ? j.ref + e : j.deg + (e - nr);
    const double v = *p;
    const double x = v == 0.0 ? DBL_EPSILON : fabs(v);
    const double db = 10.0 * fdlibm_log10(x);
    *p = fmax(db, -45.0);
}

__global__ __launch_bounds__(256) void k_frame_floor(const PrepJob* __restrict__ jobs, int nb) {
    const PrepJob j = jobs[blockIdx.y];
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int mf = j.nfr < j.nfd ? j.nfr : j.nfd;
    if (c >= mf) return;
    // the column's bands held in registers (nb <= kMaxBands): every load
    // issued before the first is used, each value read once
    double rv[kMaxBands], dv[kMaxBands];
#pragma unroll
    for (int r = 0; r < kMaxBands; r++)
        if (r < nb) {
            rv[r] = j.ref[(long long)r * j.nfr + c];
            dv[r] = j.deg[(long long)r * j.nfd + c];
        }
    double col_max = -INFINITY;
#pragma unroll
    for (int r = 0; r < kMaxBands; r++)
        if (r < nb) {
            const double m = fmax(rv[r], dv[r]);
            if (m > col_max) col_max = m;
        }
    const double floor_db = col_max - 45.0;
#pragma unroll
    for (int r = 0; r < kMaxBands; r++)
        if (r < nb) {
            if (rv[r] < floor_db) j.ref[(long long)r * j.nfr + c] = floor_db;
            if (dv[r] < floor_db) j.deg[(long long)r * j.nfd + c] = floor_db;
        }
}

__global__ void k_min_init(unsigned long long* scratch, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) scratch[i] = ~0ULL;
}

__global__ __launch_bounds__(256) void k_min(const PrepJob* __restrict__ jobs, int nb, unsigned long long* scratch) {
    const PrepJob j = jobs[blockIdx.y];
    const long long nr = (long long)nb * j.nfr, nd = (long long)nb * j.nfd;
    unsigned long long best = ~0ULL;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < nr + nd; e += (long long)gridDim.x * blockDim.x) {
        const unsigned long long k = order_key(e < nr ? j.ref[e] : j.deg[e - nr]);
        best = k < best ? k : best;
    }
    __shared__ unsigned long long red[256];
    red[threadIdx.x] = best;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = red[threadIdx.x] < red[threadIdx.x + w] ? red[threadIdx.x] : red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMin(&scratch[blockIdx.y], red[0]);
}

__global__ __launch_bounds__(256) void k_sub_min(const PrepJob* __restrict__ jobs, int nb, const unsigned long long* scratch) {
    const PrepJob j = jobs[blockIdx.y];
    const long long nr = (long long)nb * j.nfr, nd = (long long)nb * j.nfd;
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nr + nd) return;
    const double lo = key_value(scratch[blockIdx.y]);
    double* p = e < nr ? j.ref + e : j.deg + (e - nr);
    *p = *p - lo;
}

void launch_prep(const PrepJob* d_jobs, int njobs, int nb, int max_nf, unsigned long long* d_scratch, hipStream_t stream) {
    if (njobs == 0) return;
    const long long cells = 2LL * nb * max_nf;
    const dim3 ge((unsigned)((cells + 255) / 256), njobs);
    hipLaunchKernelGGL(k_db, ge, dim3(256), 0, stream, d_jobs, nb);
    hipLaunchKernelGGL(k_frame_floor, dim3((unsigned)((max_nf + 255) / 256), njobs), dim3(256), 0, stream, d_jobs, nb);
    hipLaunchKernelGGL(k_min_init, dim3((njobs + 255) / 256), dim3(256), 0, stream, d_scratch, njobs);
    unsigned mb = (unsigned)((cells + 255) / 256);
    if (mb > 64) mb = 64;
    hipLaunchKernelGGL(k_min, dim3(mb, njobs), dim3(256), 0, stream, d_jobs, nb, d_scratch);
    hipLaunchKernelGGL(k_sub_min, ge, dim3(256), 0, stream, d_jobs, nb, d_scratch);
}

// ---------------------------------------------------------------- deg plane
// nsim.zig:236-263: the deg spectrogram zero padded to `width` columns and
// its square (planes kPlanePad, kPlaneSq).
__global__ __launch_bounds__(256) void k_plane_fill(const PlaneJob* __restrict__ jobs, int nb) {
    const PlaneJob j = jobs[blockIdx.y];
    const long long n = (long long)nb * j.width;
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int r = (int)(e / j.width), c = (int)(e % j.width);
    const double v = c < j.nf ? j.spec[(long long)r * j.nf + c] : 0.0;
    j.plane[e] = v;
    j.plane[n + e] = v * v;
}

// Signed square root: the sign of v with the root of |v|, so that the
// product of two of them is ±sqrt(|product|) with the product's sign.
__device__ inline double signed_sqrt(double v) { return v < 0.0 ? -sqrt(-v) : sqrt(v); }

// Means and variances of the 3x3 window at every column u, for the three
// column sets a patch cell can see: interior (u-1, u, u+1), a patch's first
// column (u, u, u+1) and its last (u-1, u, u) — the patch's own replicate
// padding (nsim.zig:380-395); rows replicate at the spectrogram edges.
// Variance planes hold conv(x²) - mean², or its signed square root (fast).
template <bool FAST>
__global__ __launch_bounds__(256) void k_plane_conv(const PlaneJob* __restrict__ jobs, int nb, int S) {
    const PlaneJob j = jobs[blockIdx.y];
    const long long n = (long long)nb * j.width;
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int w = j.width;
    const int r = (int)(e / w), c = (int)(e % w);
    const int rm = r == 0 ? 0 : r - 1, rp = r + 1 >= nb ? nb - 1 : r + 1;
    const int cl = c == 0 ? 0 : c - 1, cr = c + 1 >= w ? w - 1 : c + 1;
    const double* pad = j.plane;
    const double* sq = j.plane + n;
    auto put = [&](int plane_mu, int plane_v, int cm, int cp) {
        const double mu = tap9([&](int rr, int cc) { return pad[(long long)rr * w + cc]; }, rm, r, rp, cm, c, cp);
        const double csq = tap9([&](int rr, int cc) { return sq[(long long)rr * w + cc]; }, rm, r, rp, cm, c, cp);
        const double v = csq - mu * mu;
        j.plane[plane_mu * n + e] = mu;
        j.plane[plane_v * n + e] = FAST ? signed_sqrt(v) : v;
    };
    put(kPlaneMu, kPlaneVar, cl, cr);
    put(kPlaneMuFirst, kPlaneVarFirst, c, S > 1 ? cr : c);
    put(kPlaneMuLast, kPlaneVarLast, S > 1 ? cl : c, c);
}

void launch_planes(const PlaneJob* d_jobs, int njobs, int nb, int patch, int max_width, bool fast, hipStream_t stream) {
    if (njobs == 0) return;
    const dim3 g((unsigned)(((long long)nb * max_width + 255) / 256), njobs);
    hipLaunchKernelGGL(k_plane_fill, g, dim3(256), 0, stream, d_jobs, nb);
    if (fast) hipLaunchKernelGGL(k_plane_conv<true>, g, dim3(256), 0, stream, d_jobs, nb, patch);
    else hipLaunchKernelGGL(k_plane_conv<false>, g, dim3(256), 0, stream, d_jobs, nb, patch);
}

// ---------------------------------------------------------------- similarity table
// Ref-patch terms (nsim.zig:318-329): mu_ref, mu_ref² (fast: mu_ref² + c1,
// the sum every cell forms), conv(ref²) - mu_ref² (fast: its signed square
// root), with the patch's own replicate padding, and the patch's values;
// terms[(p*4 + t)*nb*S + r*S + c].
template <bool FAST>
__global__ __launch_bounds__(256) void k_ref_terms(const SimJob* __restrict__ jobs, int nb, int S, double* __restrict__ terms,
                                                   int max_np) {
    const SimJob j = jobs[blockIdx.z];
    const int p = blockIdx.y;
    if (p >= j.np) return;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nb * S) return;
    const int r = e / S, c = e % S;
    const int rm = r == 0 ? 0 : r - 1, rp = r + 1 >= nb ? nb - 1 : r + 1;
    const int cm = S == 1 ? 0 : (c == 0 ? 0 : c - 1), cp = S == 1 ? 0 : (c + 1 >= S ? S - 1 : c + 1);
    const double* ref = j.ref_spec + j.ref_idx[p];
    const long long st = j.nf_ref;
    const double mu = tap9([&](int rr, int cc) { return ref[rr * st + cc]; }, rm, r, rp, cm, c, cp);
    const double crs = tap9([&](int rr, int cc) { const double v = ref[rr * st + cc]; return v * v; }, rm, r, rp, cm, c, cp);
    const double rms = mu * mu;
    double* t = terms + ((long long)blockIdx.z * max_np + p) * 4 * nb * S;
    if (FAST) {  // cell-major: a cell's four terms side by side
        t[4 * e] = mu;
        t[4 * e + 1] = rms + nsim_c1();
        t[4 * e + 2] = signed_sqrt(crs - rms);
        t[4 * e + 3] = ref[r * st + c];
    } else {
        t[e] = mu;
        t[nb * S + e] = rms;
        t[2 * nb * S + e] = crs - rms;
        t[3 * nb * S + e] = ref[r * st + c];
    }
}

// One similarity cell in the throughput mode: simCell's value (nsim.zig:
// 293-308) from the signed roots of the two variances (their product is
// ±sqrt(|srs·sds|)), with fused multiply-adds and one reciprocal refined by
// two Newton steps in place of the two divisions.  rmsc1 = mu_ref² + c1;
// crd3 = the windowed covariance term plus c3 (accumulated from c3).
__device__ inline double sim_cell_fast(double mr, double rmsc1, double ssr, double md, double ssd, double crd3) {
    const double c1 = nsim_c1(), c3 = nsim_c3();
    const double mrd = mr * md;
    const double num = __builtin_fma(2.0, mrd, c1) * (crd3 - mrd);
    const double t = ssr * ssd;
    const double den = fmax(t, 0.0) + c3;  // the same value as (t < 0 ? 0 : t) + c3 for every non-NaN t
    const double d = __builtin_fma(md, md, rmsc1) * den;
    double q = __builtin_amdgcn_rcp(d);
    double err = __builtin_fma(-d, q, 1.0);
    q = __builtin_fma(q, err, q);
    err = __builtin_fma(-d, q, 1.0);
    q = __builtin_fma(q, err, q);
    return num * q;
}

constexpr int kSimThreads = 512;
// Launches of fewer blocks of kSimPatchGroup patches and kSimThreads lanes
// than this (up to four per CU) take 64-lane one-patch blocks (one 60 s
// pair alone: 2.47 -> 2.38 ms against a threshold of one block per CU).
constexpr long long kSimSmallLaunch = 1024;

// measureSimilarityAt (nsim.zig:364-424) for kSimPatchGroup patches of one
// pair at one offset per lane, lanes on consecutive offsets of the union of
// the patches' search windows (one window for every patch of a signal
// shorter than the search radius).  The patches' terms and values sit in
// LDS (read by broadcast).  Along a row the 3x3 taps slide one column per
// cell: the products X = ref·deg of rows rm, r, rp at columns c-1, c, c+1
// (patch-clamped) stay in registers, so each cell loads three degraded
// values plus its window's mean and variance (from the plane that matches
// its column: interior, a patch's first or its last) — once for all the
// group's patches, which see the same degraded columns at the same offset.
// Each (patch, offset) keeps its own sum in row-major order.  Exact: the
// reference's tap order and serial sum, every product rounded on its own.
// FAST: per column the products keep only T = X[rm] + X[rp] (one FMA) and
// M = X[r]; the window's weights (corner WC, edge WE, middle WM) then give
// c3 + WC·(T[c-1] + T[c+1]) + WE·(T[c] + M[c-1] + M[c+1]) + WM·M[c]:
// six operations a cell instead of nine, with fused multiply-adds.
//
// PP patches per block and THREADS lanes: launches that fill the device run
// kSimPatchGroup patches on kSimThreads lanes; launches of a few pairs, whose
// time is one block's serial cell loop, run one patch per 64-lane block, so
// their (patch, offset) lanes spread over the most CUs.  Every grouping
// gives the same bits (each lane's sum keeps its order).
template <bool FAST, int PP, int THREADS>
__global__ __launch_bounds__(THREADS) void k_sim_table(const SimJob* __restrict__ jobs, int nb, int S,
                                                       const double* __restrict__ terms, int max_np) {
    extern __shared__ double sh[];  // [PP][4][nb*S]: mu_ref, mu_ref², sigma_ref² (or root), ref
    const SimJob j = jobs[blockIdx.z];
    const int p0 = blockIdx.y * PP;
    if (p0 >= j.np) return;
    const int npres = j.np - p0 < PP ? j.np - p0 : PP;  // patches present in this group
    int lo_u = j.lo[p0], hi_u = j.hi[p0];
    int lo[PP], hi[PP];
#pragma unroll
    for (int q = 0; q < PP; q++) {
        lo[q] = q < npres ? j.lo[p0 + q] : 1;
        hi[q] = q < npres ? j.hi[p0 + q] : 0;  // absent: empty window
        if (q < npres) {
            lo_u = lo[q] < lo_u ? lo[q] : lo_u;
            hi_u = hi[q] > hi_u ? hi[q] : hi_u;
        }
    }
    const int first = lo_u + (int)blockIdx.x * THREADS;
    if (first > hi_u) return;  // uniform
    const int cells = nb * S;
    const double* t = terms + ((long long)blockIdx.z * max_np + p0) * 4 * cells;
    if (!FAST) {
        for (int e = threadIdx.x; e < 4 * cells * npres; e += THREADS) sh[e] = t[e];
        __syncthreads();
    }
    // Term k of cell e of the group's patch q: exact mode from LDS (four
    // planes per patch), fast mode straight from the cell-major global
    // table with wave-uniform indices (scalar loads, no LDS traffic).
    auto term = [&](int q, int k, int e) -> double {
        if (FAST) return t[((long long)q * cells + e) * 4 + k];
        return sh[(q * 4 + k) * cells + e];
    };

    const int off = first + (int)threadIdx.x;
    const int o = off <= hi_u ? off : lo_u;  // idle lanes shadow a valid column
    const int w = j.width;
    const long long n = (long long)nb * w;
    // Loads are (uniform column pointer)[per-lane 32-bit index]: plane sizes
    // stay far below 2^29 elements.
    const double* pad = j.plane;
    const double* mu = j.plane + kPlaneMu * n;
    const double* var = j.plane + kPlaneVar * n;
    const double* mu_first = j.plane + kPlaneMuFirst * n;
    const double* var_first = j.plane + kPlaneVarFirst * n;
    const double* mu_last = j.plane + kPlaneMuLast * n;
    const double* var_last = j.plane + kPlaneVarLast * n;
    double sum[PP];
#pragma unroll
    for (int q = 0; q < PP; q++) sum[q] = 0.0;
    enum : int { kFirst, kMid, kLast };
    for (int r = 0; r < nb; r++) {
        const int rm = r == 0 ? 0 : r - 1, rp = r + 1 >= nb ? nb - 1 : r + 1;
        const unsigned q0 = (unsigned)(rm * w + o), q1 = (unsigned)(r * w + o), q2 = (unsigned)(rp * w + o);
        // Three column slots of products per patch; cell c uses slots
        // (a, b, c) as columns c-1, c, c+1, then refills slot a with column
        // c+2 (the next cell's c+1): the roles rotate.
        double X[3][PP][3];
        auto fill = [&](double (&Xd)[PP][3], int col) {
            const auto* g = as_global(pad + col);
            const double d0 = g[q0], d1 = g[q1], d2 = g[q2];
#pragma unroll
            for (int q = 0; q < PP; q++) {
                const double fm = term(q, 3, rm * S + col), f0 = term(q, 3, r * S + col), fp = term(q, 3, rp * S + col);
                if (FAST) {
                    Xd[q][0] = __builtin_fma(fp, d2, fm * d0);  // T
                    Xd[q][1] = f0 * d1;                         // M
                } else {
                    Xd[q][0] = fm * d0;
                    Xd[q][1] = f0 * d1;
                    Xd[q][2] = fp * d2;
                }
            }
        };
        fill(X[1], 0);
#pragma unroll
        for (int q = 0; q < PP; q++)
#pragma unroll
            for (int k = 0; k < (FAST ? 2 : 3); k++) X[0][q][k] = X[1][q][k];
        fill(X[2], S > 1 ? 1 : 0);
        auto cell = [&](auto kind, int c, double (&Xa)[PP][3], double (&Xb)[PP][3], double (&Xc)[PP][3]) {
            constexpr int K = decltype(kind)::value;
            const double* pm = K == kFirst ? mu_first : (K == kLast ? mu_last : mu);
            const double* pv = K == kFirst ? var_first : (K == kLast ? var_last : var);
            const double md = as_global(pm + c)[q1];
            const double v = as_global(pv + c)[q1];
            const int e = r * S + c;
#pragma unroll
            for (int q = 0; q < PP; q++) {
                double crd;
                if (FAST) {
                    const double corners = Xa[q][0] + Xc[q][0];
                    const double edges = Xb[q][0] + (Xa[q][1] + Xc[q][1]);
                    crd = __builtin_fma(VQ_WM, Xb[q][1], nsim_c3());
                    crd = __builtin_fma(VQ_WE, edges, crd);
                    crd = __builtin_fma(VQ_WC, corners, crd);
                } else {
                    crd = VQ_WC * Xa[q][0] + VQ_WE * Xb[q][0] + VQ_WC * Xc[q][0] + VQ_WE * Xa[q][1] +
                          VQ_WM * Xb[q][1] + VQ_WE * Xc[q][1] + VQ_WC * Xa[q][2] + VQ_WE * Xb[q][2] +
                          VQ_WC * Xc[q][2];
                }
                if (FAST) sum[q] += sim_cell_fast(term(q, 0, e), term(q, 1, e), term(q, 2, e), md, v, crd);
                else sum[q] += sim_cell(term(q, 0, e), term(q, 1, e), term(q, 2, e), md, md * md, v, crd);
            }
            if (K != kLast) fill(Xa, c + 2 < S ? c + 2 : S - 1);
        };
        using First = std::integral_constant<int, kFirst>;
        using Mid = std::integral_constant<int, kMid>;
        using Last = std::integral_constant<int, kLast>;
        // S == 1: the one cell's window is (0, 0, 0), the first-column plane.
        cell(First(), 0, X[0], X[1], X[2]);
        if (S > 1) {
            int c = 1;
            for (; c + 3 <= S - 1; c += 3) {
                cell(Mid(), c, X[1], X[2], X[0]);
                cell(Mid(), c + 1, X[2], X[0], X[1]);
                cell(Mid(), c + 2, X[0], X[1], X[2]);
            }
            const int rem = S - 1 - c;  // interior cells left: 0, 1 or 2
            if (rem == 0) {
                cell(Last(), c, X[1], X[2], X[0]);
            } else if (rem == 1) {
                cell(Mid(), c, X[1], X[2], X[0]);
                cell(Last(), c + 1, X[2], X[0], X[1]);
            } else {
                cell(Mid(), c, X[1], X[2], X[0]);
                cell(Mid(), c + 1, X[2], X[0], X[1]);
                cell(Last(), c + 2, X[0], X[1], X[2]);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < PP; q++)
        if (off >= lo[q] && off <= hi[q]) j.sims[(long long)(p0 + q) * j.nf_deg + off] = sum[q] / (double)(nb * S);
}

void launch_sim_table(const SimJob* d_jobs, int njobs, int nb, int patch, int max_np, int max_window, double* d_terms,
                      bool fast, hipStream_t stream) {
    if (njobs == 0 || max_np == 0 || max_window == 0) return;
    const dim3 gt((nb * patch + 255) / 256, max_np, njobs);
    if (fast) hipLaunchKernelGGL(k_ref_terms<true>, gt, dim3(256), 0, stream, d_jobs, nb, patch, d_terms, max_np);
    else hipLaunchKernelGGL(k_ref_terms<false>, gt, dim3(256), 0, stream, d_jobs, nb, patch, d_terms, max_np);
    // blocks past a job's patches or search windows exit before any work
    auto launch = [&](auto pp, auto threads) {
        constexpr int PP = decltype(pp)::value, T = decltype(threads)::value;
        const dim3 grid(xcd_spread((max_window + T - 1) / T), xcd_spread((max_np + PP - 1) / PP), njobs);
        const size_t lds = fast ? 0 : sizeof(double) * 4 * nb * patch * PP;
        if (fast)
            hipLaunchKernelGGL((k_sim_table<true, PP, T>), grid, dim3(T), lds, stream, d_jobs, nb, patch,
                               (const double*)d_terms, max_np);
        else
            hipLaunchKernelGGL((k_sim_table<false, PP, T>), grid, dim3(T), lds, stream, d_jobs, nb, patch,
                               (const double*)d_terms, max_np);
    };
    const long long big_blocks = (long long)njobs * ((max_np + kSimPatchGroup - 1) / kSimPatchGroup) *
                                 ((max_window + kSimThreads - 1) / kSimThreads);
    if (big_blocks < kSimSmallLaunch)
        launch(std::integral_constant<int, 1>(), std::integral_constant<int, 64>());
    else
        launch(std::integral_constant<int, kSimPatchGroup>(), std::integral_constant<int, kSimThreads>());
}

// ---------------------------------------------------------------- DP
// dp.zig:227-352 for one pair per block.  Row p's running max over
// dp[p-1][lo_{p-1} .. s) with `>=` (the later index wins ties) is an
// exclusive scan under the associative "last argmax" operator, so each
// row is one block-wide scan plus an elementwise update; the decisions are
// exact because max never rounds.
struct ArgMax {
    double v;
    int i;
};
__device__ inline ArgMax later_max(ArgMax a, ArgMax b) { return b.v >= a.v ? b : a; }

// Exclusive scan of one ArgMax per thread of a block under later_max, in
// thread order: shuffles within each wave, then the waves' totals through
// LDS.  later_max is associative (max never rounds; ties go
// to the later operand), so the result is the serial left-to-right one.
__device__ inline ArgMax block_exclusive_later_max(ArgMax x, double* wv, int* wi) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const double v = __shfl_up(x.v, off, 64);
        const int i = __shfl_up(x.i, off, 64);
        if (lane >= off) x = later_max(ArgMax{v, i}, x);
    }
    if (lane == 63) {
        wv[wave] = x.v;
        wi[wave] = x.i;
    }
    const double ev = __shfl_up(x.v, 1, 64);
    const int ei = __shfl_up(x.i, 1, 64);
    __syncthreads();
    ArgMax pre = {-INFINITY, -1};
    for (int w = 0; w < wave; w++) pre = later_max(pre, ArgMax{wv[w], wi[w]});
    return lane == 0 ? pre : later_max(pre, ArgMax{ev, ei});
}

// Rows whose scan range (lo of the row before .. hi) spans at most kDpLds
// frames keep the previous DP row and the current similarity row in LDS,
// in rings indexed by frame mod kDpLds: the similarity row comes in with
// coalesced loads, the scan and update read LDS, and the DP row and the
// backtrack pointers still go to global memory (the backtrack and the last
// row's search read them there).  Longer rows (search radii far above the
// default) read and write global memory only.
constexpr int kDpLds = 4096;

//
// THREADS lanes per pair: 256, or 1024 when the launch's longest scan range
// is long enough (long pairs, wide search windows) that a quarter of the
// per-thread serial run outweighs the longer cross-wave step.
template <int THREADS>
__global__ __launch_bounds__(THREADS) void k_dp(const DpJob* __restrict__ jobs) {
    constexpr int kWaves = THREADS / 64;
    const DpJob j = jobs[blockIdx.x];
    const int cols = j.nf_deg;
    const auto* wlo = as_global(j.lo);
    const auto* whi = as_global(j.hi);
    __shared__ double wv[kWaves];
    __shared__ int wi[kWaves];
    __shared__ double ring[2][kDpLds];
    __shared__ int s_span;
    constexpr int kMask = kDpLds - 1;
    // widest scan range of the job: every row fits the rings or none uses them
    if (threadIdx.x == 0) s_span = 0;
    __syncthreads();
    {
        int span = 0;
        for (int p = threadIdx.x; p < j.np; p += blockDim.x) {
            const int from = p == 0 ? wlo[0] : min(wlo[p - 1], wlo[p]);
            span = max(span, whi[p] - from + 1);
        }
        atomicMax(&s_span, span);
    }
    __syncthreads();
    const bool in_lds = s_span <= kDpLds;
    // LDS rows: each thread's share of the next similarity row is loaded
    // into registers while the current row is scanned
    constexpr int kPre = kDpLds / THREADS;
    double pre[kPre];
    // window bounds run one row ahead, so their loads overlap a row's work
    int lo_prev = 0, hi_prev = 0, lo_cur = wlo[0], hi_cur = whi[0];
    int lo_next = j.np > 1 ? wlo[1] : 0, hi_next = j.np > 1 ? whi[1] : 0;
    auto prefetch = [&](int q, int qlo, int qhi) {
        if (!in_lds || q >= j.np) return;
        const auto* sr = as_global(j.sims + (long long)q * cols);
#pragma unroll
        for (int k = 0; k < kPre; k++) {
            const int s = qlo + (int)threadIdx.x + k * THREADS;
            pre[k] = s <= qhi ? sr[s] : 0.0;
        }
    };
    prefetch(1, lo_next, hi_next);
    for (int p = 0; p < j.np; p++) {
        const int lo = lo_cur, hi = hi_cur;
        const int lo_after = p + 2 < j.np ? wlo[p + 2] : 0, hi_after = p + 2 < j.np ? whi[p + 2] : 0;
        // global views, so these are global (not flat) memory instructions
        auto* row = as_global(j.dp + (long long)p * cols);
        auto* brow = as_global(j.bt + (long long)p * cols);
        const auto* srow = as_global(j.sims + (long long)p * cols);
        double* cur = ring[p & 1];         // this row: similarities, then DP values
        const double* prv = ring[(p + 1) & 1];  // the previous row's DP values
        if (p == 0) {
            for (int s = lo + threadIdx.x; s <= hi; s += blockDim.x) {
                const double v = srow[s];
                row[s] = v;
                brow[s] = -1;
                if (in_lds) cur[s & kMask] = v;
            }
        } else {
            const auto* prev = as_global(j.dp + (long long)(p - 1) * cols);
            const int t0 = lo_prev;
            const int m = hi - t0 + 1;  // scan elements t0..hi
            if (in_lds) {
                // this row's similarities (prefetched); the previous row is
                // zero past its hi
                const int phi = hi_prev;
#pragma unroll
                for (int k = 0; k < kPre; k++) {
                    const int s = lo + (int)threadIdx.x + k * THREADS;
                    if (s <= hi) cur[s & kMask] = pre[k];
                }
                for (int s = phi + 1 + threadIdx.x; s <= hi; s += blockDim.x) ring[(p + 1) & 1][s & kMask] = 0.0;
                __syncthreads();
                prefetch(p + 1, lo_next, hi_next);
            }
            // one body per source, so LDS reads stay LDS instructions
            auto body = [&](auto prev_at, auto sim_at) {
                const int per = (m + blockDim.x - 1) / blockDim.x;
                const int a = t0 + threadIdx.x * per;
                const int b = min(a + per, hi + 1);
                ArgMax acc = {-INFINITY, -1};
                for (int t = a; t < b; t++) acc = later_max(acc, ArgMax{prev_at(t), t});
                ArgMax run = block_exclusive_later_max(acc, wv, wi);
                for (int t = a; t < b; t++) {
                    const double pv = prev_at(t);
                    if (t >= lo) {
                        double sim = sim_at(t) + run.v;
                        int past = run.i;
                        if (pv > sim) {
                            sim = pv;
                            past = t;
                        }
                        row[t] = sim;
                        brow[t] = past < 0 ? -1 : past;
                        if (in_lds) cur[t & kMask] = sim;
                    }
                    run = later_max(run, ArgMax{pv, t});
                }
            };
            if (in_lds)
                body([&](int t) { return prv[t & kMask]; }, [&](int t) { return cur[t & kMask]; });
            else
                body([&](int t) { return prev[t]; }, [&](int t) { return srow[t]; });
        }
        __syncthreads();  // the row is complete (and wv free) before the next reads it
        lo_prev = lo_cur;
        hi_prev = hi_cur;
        lo_cur = lo_next;
        hi_cur = hi_next;
        lo_next = lo_after;
        hi_next = hi_after;
    }
    // first maximum of the last row (strictly greater wins), over
    // contiguous per-thread runs combined in thread order
    const int last = j.np - 1;
    const auto* lrow = as_global(j.dp + (long long)last * cols);
    const int llo = wlo[last], lhi = whi[last];
    const int lm = lhi - llo + 1;
    const int lper = (lm + blockDim.x - 1) / blockDim.x;
    const int la = llo + threadIdx.x * lper;
    const int lb = min(la + lper, lhi + 1);
    ArgMax best = {-INFINITY, INT_MAX};
    for (int s = la; s < lb; s++)
        if (lrow[s] > best.v) best = ArgMax{lrow[s], s};
    for (int off = 32; off > 0; off >>= 1) {
        const double v = __shfl_down(best.v, off, 64);
        const int i = __shfl_down(best.i, off, 64);
        // the partner covers later elements: it wins only when strictly greater
        if (v > best.v || (v == best.v && i < best.i)) best = ArgMax{v, i};
    }
    if ((threadIdx.x & 63) == 0) {
        wv[threadIdx.x >> 6] = best.v;
        wi[threadIdx.x >> 6] = best.i;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        ArgMax r = {wv[0], wi[0]};
        for (int w = 1; w < kWaves; w++)
            if (wv[w] > r.v) r = ArgMax{wv[w], wi[w]};
        int cur = r.i == INT_MAX ? 0 : r.i;
        for (int p = 0; p < j.np; p++) {
            j.offsets[p] = 0;
            j.state[p] = 0;
        }
        for (int p = last; p >= 0; p--) {
            const int back = j.bt[(long long)p * cols + cur];
            j.offsets[p] = cur;
            j.state[p] = back == cur ? 2 : 1;
            if (back < 0) break;
            cur = back;
        }
    }
}

void launch_dp(const DpJob* d_jobs, int njobs, int max_span, hipStream_t stream) {
    if (njobs == 0) return;
    if (max_span >= kDpWideSpan) hipLaunchKernelGGL(k_dp<1024>, dim3(njobs), dim3(1024), 0, stream, d_jobs);
    else hipLaunchKernelGGL(k_dp<256>, dim3(njobs), dim3(256), 0, stream, d_jobs);
}

// ---------------------------------------------------------------- patch statistics
// nsim.zig:427-510 on one (ref patch, deg patch) pair per block; lane r
// owns band r and keeps the reference's per-band summation order.
__global__ __launch_bounds__(64) void k_stats(const StatJob* __restrict__ jobs, int nb) {
    extern __shared__ double lds[];
    const StatJob j = jobs[blockIdx.x];
    const int S = j.ncols;
    double* R = lds;
    double* D = lds + nb * S;
    double* SIM = lds + 2 * nb * S;
    // staged eight elements a thread at a time, every load of a group
    // issued before the group's first is stored
    constexpr int kGroup = 8;
    for (int e0 = threadIdx.x; e0 < nb * S; e0 += kGroup * blockDim.x) {
        double rv[kGroup], dv[kGroup];
#pragma unroll
        for (int u = 0; u < kGroup; u++) {
            const int e = e0 + u * (int)blockDim.x;
            rv[u] = 0.0;
            dv[u] = 0.0;
            if (e < nb * S) {
                const int r = e / S, c = e - r * S;
                rv[u] = j.ref[(long long)r * j.ref_stride + j.ref_col + c];
                const int dc = j.deg_col + c;
                if (dc < j.deg_valid) dv[u] = j.deg[(long long)r * j.deg_stride + dc];
            }
        }
#pragma unroll
        for (int u = 0; u < kGroup; u++) {
            const int e = e0 + u * (int)blockDim.x;
            if (e < nb * S) {
                R[e] = rv[u];
                D[e] = dv[u];
            }
        }
    }
    __syncthreads();
    const int r = threadIdx.x;
    if (r < nb) {
        const double inv_ncols = 1.0 / (double)S;
        double deg_sum = 0.0;
        for (int c = 0; c < S; c++) deg_sum += D[r * S + c];
        j.out[2 * nb + r] = deg_sum * inv_ncols;
        const int rm = r == 0 ? 0 : r - 1, rp = r + 1 >= nb ? nb - 1 : r + 1;
        double sim_sum = 0.0;
        for (int c = 0; c < S; c++) {
            const int cm = S == 1 ? 0 : (c == 0 ? 0 : c - 1), cp = S == 1 ? 0 : (c + 1 >= S ? S - 1 : c + 1);
            const double mr = tap9([&](int rr, int cc) { return R[rr * S + cc]; }, rm, r, rp, cm, c, cp);
            const double md = tap9([&](int rr, int cc) { return D[rr * S + cc]; }, rm, r, rp, cm, c, cp);
            const double crs = tap9([&](int rr, int cc) { const double v = R[rr * S + cc]; return v * v; }, rm, r, rp, cm, c, cp);
            const double cds = tap9([&](int rr, int cc) { const double v = D[rr * S + cc]; return v * v; }, rm, r, rp, cm, c, cp);
            const double crd = tap9([&](int rr, int cc) { return R[rr * S + cc] * D[rr * S + cc]; }, rm, r, rp, cm, c, cp);
            const double rms = mr * mr, dms = md * md, mrd = mr * md;
            const double srs = crs - rms, sds = cds - dms, srd = crd - mrd;
            const double c1 = nsim_c1(), c3 = nsim_c3();
            const double intensity = (2.0 * mrd + c1) / (rms + dms + c1);
            const double prod = srs * sds;
            const double den = prod < 0.0 ? c3 : sqrt(prod) + c3;
            const double structure = (srd + c3) / den;
            const double sim = intensity * structure;
            sim_sum += sim;
            SIM[r * S + c] = sim;
        }
        const double mean = sim_sum * inv_ncols;
        j.out[r] = mean;
        if (S <= 1) {
            j.out[nb + r] = 0.0;
        } else {
            double ss = 0.0;
            for (int c = 0; c < S; c++) {
                const double diff = SIM[r * S + c] - mean;
                ss += diff * diff;
            }
            j.out[nb + r] = sqrt(ss / ((double)S - 1.0));
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double acc = 0.0;
        for (int b = 0; b < nb; b++) acc += j.out[b];
        j.out[3 * nb] = acc / (double)nb;
    }
}

void launch_stats(const StatJob* d_jobs, int njobs, int nb, int max_cols, hipStream_t stream) {
    if (njobs == 0) return;
    const size_t lds = sizeof(double) * 3 * nb * (size_t)max_cols;
    hipLaunchKernelGGL(k_stats, dim3(njobs), dim3(64), lds, stream, d_jobs, nb);
}


}  // namespace vq
