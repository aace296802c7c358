// Shared device/host helpers for the gfx950 ViSQOL kernels.
//
// Exact-mode kernels are compiled with -ffp-contract=off, so every a*b+c
// below rounds twice, as the reference CPU path does (Zig does not contract
// by default).  Fast-mode kernels spell their fused multiply-adds out with
// __builtin_fma, so they never depend on the compiler's contraction.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vq {

constexpr int kMaxBands = 32;

// A device pointer read from a job table is a generic (flat) pointer; its
// loads count against both the vector-memory and the LDS wait counters, so a
// wait for an LDS read also waits for them.  Every buffer here is global
// memory: this view makes the loads global loads.
template <typename T>
__device__ inline const __attribute__((address_space(1))) T* as_global(const T* p) {
    return (const __attribute__((address_space(1))) T*)p;
}
template <typename T>
__device__ inline __attribute__((address_space(1))) T* as_global(T* p) {
    return (__attribute__((address_space(1))) T*)p;
}
// double2 element i of a global buffer, as one 16-byte global access.
typedef double vq_d2 __attribute__((ext_vector_type(2)));
__device__ inline double2 load2(const double2* p, long long i) {
    const vq_d2 v = ((const __attribute__((address_space(1))) vq_d2*)p)[i];
    return make_double2(v.x, v.y);
}
__device__ inline void store2(double2* p, long long i, double2 v) {
    vq_d2 w;
    w.x = v.x;
    w.y = v.y;
    ((__attribute__((address_space(1))) vq_d2*)p)[i] = w;
}
// float2 element i, one 8-byte global access.
typedef float vq_f2 __attribute__((ext_vector_type(2)));
__device__ inline float2 load2(const float2* p, long long i) {
    const vq_f2 v = ((const __attribute__((address_space(1))) vq_f2*)p)[i];
    return make_float2(v.x, v.y);
}
__device__ inline void store2(float2* p, long long i, float2 v) {
    vq_f2 w;
    w.x = v.x;
    w.y = v.y;
    ((__attribute__((address_space(1))) vq_f2*)p)[i] = w;
}

// 2 × 8 coefficients per band, precomputed on the host exactly as
// gammatone.zig:268-278 derives them from Filters: s1n0, s1n1, s2n1, s3n1,
// s4n1, ca0, d1, d2.
struct BandCoef {
    double s1n0, s1n1, s2n1, s3n1, s4n1, ca0, d1, d2;
};

// The same band as one 5-tap FIR (the product of the four biquads'
// numerators) followed by the four biquads' shared all-pole sections
// 1 / (1 + d1 z^-1 + d2 z^-2): the throughput mode's form above the lowest
// bands (kDf2tBands), 14 instead of 17 operations per sample.
struct FirCoef {
    double c0, c1, c2, c3, c4, d1, d2, pad;
};
// Bands below this index keep the cascaded DF2T form in the throughput
// mode: their numerator zeros and poles crowd z = 1, where the FIR form
// loses up to ~4e-9 relative accuracy (band 0-4 at 48 kHz); from band 8 up
// (cf >= 622 Hz at 48 kHz, 783 Hz at 16 kHz) it stays within ~3e-12.
constexpr int kDf2tBands = 8;

__host__ __device__ inline double bits_f64(uint64_t b) {
    union { uint64_t u; double d; } x;
    x.u = b;
    return x.d;
}
__host__ __device__ inline uint64_t f64_bits(double d) {
    union { uint64_t u; double d; } x;
    x.d = d;
    return x.u;
}

// log10 by the published fdlibm / FreeBSD e_log10.c algorithm (as carried
// by musl): reduce x to 2^k·(1+f) with 1+f in [sqrt(2)/2, sqrt(2)], evaluate
// log(1+f) with the Lg1..Lg7 minimax polynomial in s = f/(2+f), and scale by
// 1/ln(10) and log10(2) split into hi/lo parts.  The reference's Zig
// std.math.log10 implements this same algorithm; host (SPL) and device
// (dB conversion) evaluate this one definition, so they agree bit for bit.
__host__ __device__ inline double fdlibm_log10(double x) {
    const double ivln10hi = bits_f64(0x3fdbcb7b15200000ULL);
    const double ivln10lo = bits_f64(0x3dbb9438ca9aadd5ULL);
    const double log10_2hi = bits_f64(0x3fd34413509f6000ULL);
    const double log10_2lo = bits_f64(0x3d59fef311f12b36ULL);
    const double Lg1 = bits_f64(0x3fe5555555555593ULL), Lg2 = bits_f64(0x3fd999999997fa04ULL);
    const double Lg3 = bits_f64(0x3fd2492494229359ULL), Lg4 = bits_f64(0x3fcc71c51d8e78afULL);
    const double Lg5 = bits_f64(0x3fc7466496cb03deULL), Lg6 = bits_f64(0x3fc39a09d078c69fULL);
    const double Lg7 = bits_f64(0x3fc2f112df3e5244ULL);
    uint64_t ix = f64_bits(x);
    uint32_t hx = (uint32_t)(ix >> 32);
    int k = 0;
    if (hx < 0x00100000u || (hx >> 31)) {
        if ((ix << 1) == 0) return -1.0 / (x * x);
        if (hx >> 31) return (x - x) / 0.0;
        k -= 54;
        x *= bits_f64(0x4350000000000000ULL);
        ix = f64_bits(x);
        hx = (uint32_t)(ix >> 32);
    } else if (hx >= 0x7ff00000u) {
        return x;
    } else if (hx == 0x3ff00000u && (ix << 32) == 0) {
        return 0.0;
    }
    hx += 0x3ff00000u - 0x3fe6a09eu;
    k += (int)(hx >> 20) - 0x3ff;
    hx = (hx & 0x000fffffu) + 0x3fe6a09eu;
    x = bits_f64(((uint64_t)hx << 32) | (ix & 0xffffffffULL));
    const double f = x - 1.0;
    const double hfsq = 0.5 * f * f;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double w = z * z;
    const double t1 = w * (Lg2 + w * (Lg4 + w * Lg6));
    const double t2 = z * (Lg1 + w * (Lg3 + w * (Lg5 + w * Lg7)));
    const double R = t2 + t1;
    double hi = f - hfsq;
    hi = bits_f64(f64_bits(hi) & 0xffffffff00000000ULL);
    const double lo = f - hi - hfsq + s * (hfsq + R);
    double val_hi = hi * ivln10hi;
    const double dk = (double)k;
    const double y = dk * log10_2hi;
    double val_lo = dk * log10_2lo + (lo + hi) * ivln10lo + lo * ivln10hi;
    w = y + val_hi;
    val_lo += (y - w) + val_hi;
    val_hi = w;
    return val_lo + val_hi;
}

// Orderable 64-bit key of a double: unsigned order equals numeric order
// (NaN-free inputs), so atomicMin / atomicMax on the key give exact
// min / max independent of the reduction order.
__host__ __device__ inline uint64_t order_key(double d) {
    uint64_t u = f64_bits(d);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ULL);
}
__host__ __device__ inline double key_value(uint64_t k) {
    return (k >> 63) ? bits_f64(k & 0x7fffffffffffffffULL) : bits_f64(~k);
}

// NSIM constants (nsim.zig:340-341 with intensity_range 1).
__host__ __device__ inline double nsim_c1() { return (0.01 * 1.0) * (0.01 * 1.0); }
__host__ __device__ inline double nsim_c3() { return (0.03 * 1.0) * (0.03 * 1.0) / 2.0; }

// The 3x3 Gaussian window's three distinct weights (nsim.zig:111-115); the
// window is symmetric, so its flipped taps equal it.
#define VQ_WC 0.0113033910173052
#define VQ_WE 0.0838251475442633
#define VQ_WM 0.619485845753726

// One output cell of the 3x3 true convolution in the reference's single
// tap order (nsim.zig:145-155): rows rm/r0/rp, columns cm/c/cp.
template <typename F>
__device__ inline double tap9(F at, int rm, int r0, int rp, int cm, int c, int cp) {
    return VQ_WC * at(rm, cm) + VQ_WE * at(rm, c) + VQ_WC * at(rm, cp) + VQ_WE * at(r0, cm) + VQ_WM * at(r0, c) +
           VQ_WE * at(r0, cp) + VQ_WC * at(rp, cm) + VQ_WE * at(rp, c) + VQ_WC * at(rp, cp);
}

// nsim.zig:293-308 (simCell)
__device__ inline double sim_cell(double mr, double rms, double srs, double md, double dms, double sds, double crd) {
    const double c1 = nsim_c1(), c3 = nsim_c3();
    const double mrd = mr * md;
    const double srd = crd - mrd;
    const double intensity = (2.0 * mrd + c1) / (rms + dms + c1);
    const double prod = srs * sds;
    const double den = prod < 0.0 ? c3 : sqrt(prod) + c3;
    const double structure = (srd + c3) / den;
    return intensity * structure;
}

}  // namespace vq
