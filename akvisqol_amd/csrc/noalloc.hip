// visqol_hip_best_lag_noalloc: the exact FFT lag search of
// visqol_hip_best_lag (xcorr.zig:180-251) on signals already in device
// memory, for callers that cannot allocate device memory (code under graph
// capture, serving loops that forbid hipMalloc).  Every buffer the search
// needs beyond the caller's inputs lives in this file's static device
// arrays: `scratch` (the complex work row, the reductions, the per-tile
// maxima, the lag and the job tables) and `noalloc_tw` (the twiddle table
// of the largest transform, which holds every smaller transform's twiddles
// bit for bit: W_S^j sits at S/2 + j whatever the size, and its angle
// -2π·(j·n/S)/n rounds the same for every power of two n).  Each device
// holds its own copy until the process ends; every call on a device shares
// it, so calls that overlap on one device race.
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <string.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/visqol.h"
#include "../../include/visqol_hip.h"
#include "capi_internal.h"
#include "host.h"
#include "kernels.h"

using namespace vq;
using namespace vq_capi;

namespace {

constexpr long long kMaxN = 1LL << 20;  // largest transform: signals of up to 2^19 samples
constexpr size_t kZBytes = sizeof(double2) * kMaxN;
constexpr size_t kPartsBytes = 16 << 10;  // per-tile (max, position) pairs of the last pass (<= 256 tiles)
constexpr size_t kTableBytes = 4 << 10;   // job tables, sums and the lag
constexpr size_t kScratchBytes = kZBytes + kPartsBytes + kTableBytes;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];
static __device__ __align__(256) double noalloc_tw[2 * kMaxN];

namespace {

// The job tables written by a kernel from its arguments, so a call copies
// nothing from the host.
struct NoallocJobs {
    SumJob plain[2], nyquist[2];
    LagJob lag;
};

__global__ void k_noalloc_jobs(NoallocJobs* dst, NoallocJobs jobs) { *dst = jobs; }

struct DeviceScratch {
    bool ready = false;
    unsigned char* scratch = nullptr;  // this device's copy of `scratch`
    double2* tw = nullptr;             // and of `noalloc_tw`, filled
};

std::mutex g_mu;
std::vector<DeviceScratch> g_dev;

// Once per device, with it current: the arrays' device addresses (the
// host-side names are stand-ins, never device addresses) and the twiddle
// table.  Returns a visqol.h code.
int setup(int dev, DeviceScratch*& out) {
    std::lock_guard<std::mutex> lk(g_mu);
    if ((int)g_dev.size() <= dev) g_dev.resize(dev + 1);
    DeviceScratch& d = g_dev[dev];
    if (!d.ready) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            set_error("hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
            return VISQOL_E_BACKEND;
        }
        void* t = nullptr;
        e = hipGetSymbolAddress(&t, HIP_SYMBOL(noalloc_tw));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            set_error("hipGetSymbolAddress(noalloc_tw): %s", hipGetErrorString(e));
            return VISQOL_E_BACKEND;
        }
        const std::vector<double> tw = make_level_twiddles(kMaxN);
        e = hipMemcpy(t, tw.data(), sizeof(double) * tw.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            set_error("twiddle upload: %s", hipGetErrorString(e));
            return VISQOL_E_BACKEND;
        }
        d.scratch = static_cast<unsigned char*>(p);
        d.tw = static_cast<double2*>(t);
        d.ready = true;
    }
    out = &d;
    return VISQOL_OK;
}

}  // namespace

extern "C" {

int visqol_hip_best_lag_noalloc(const double* dev_a, size_t la, const double* dev_b, size_t lb, int64_t* lag) try {
    if ((!dev_a && la) || (!dev_b && lb) || !lag) return invalid("invalid argument");
    if (la > (size_t)(kMaxN / 2) || lb > (size_t)(kMaxN / 2)) {
        set_error("visqol_hip_best_lag_noalloc: signals of %zu and %zu samples exceed its limit of %lld "
                  "(a transform of at most 2^20 points)",
                  la, lb, kMaxN / 2);
        return VISQOL_E_INVALID_ARGUMENT;
    }
    std::string why;
    if (!device_ok(why)) {
        set_error("%s", why.c_str());
        return VISQOL_E_BACKEND;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        (void)hipGetLastError();
        set_error("hipGetDevice failed");
        return VISQOL_E_BACKEND;
    }
    DeviceScratch* ds = nullptr;
    int rc = setup(dev, ds);
    if (rc != VISQOL_OK) return rc;
    const long long big = (long long)(la > lb ? la : lb);
    if (big == 0) {
        *lag = 0;
        return VISQOL_OK;
    }
    const long long n = fft_size_for(big);
    const long long parts = lag_partials_per_job(n);
    if ((size_t)(2 * parts) * sizeof(unsigned long long) > kPartsBytes) {
        set_error("visqol_hip_best_lag_noalloc: internal scratch layout too small");
        return VISQOL_E_INTERNAL;
    }
    // scratch: z row | per-tile partials | job tables | sums | lag
    double2* z = reinterpret_cast<double2*>(ds->scratch);
    auto* d_parts = reinterpret_cast<unsigned long long*>(ds->scratch + kZBytes);
    auto* d_jobs = reinterpret_cast<NoallocJobs*>(ds->scratch + kZBytes + kPartsBytes);
    auto* d_sums = reinterpret_cast<double*>(ds->scratch + kZBytes + kPartsBytes + 1024);
    auto* d_lag = reinterpret_cast<long long*>(ds->scratch + kZBytes + kPartsBytes + 2048);
    static_assert(sizeof(NoallocJobs) <= 1024, "job tables fit their slot");
    NoallocJobs jobs{};
    jobs.plain[0] = SumJob{dev_a, (long long)la, kSum, 0, 0, 0, nullptr, 0};
    jobs.plain[1] = SumJob{dev_b, (long long)lb, kSum, 1, 0, 0, nullptr, 0};
    jobs.nyquist[0] = SumJob{dev_a, (long long)la, kNyquist, 2, 0, 0, nullptr, 0};
    jobs.nyquist[1] = SumJob{dev_b, (long long)lb, kNyquist, 3, 1, 0, nullptr, 0};
    jobs.lag = LagJob{dev_a, dev_b, (long long)la, (long long)lb, z, n, 0, 1, 2, 3};
    // Every value below is written before it is read (the sums by their
    // launches, the work row by the first pass, each partial by its tile,
    // the lag by the reduction), so nothing is cleared between calls.
    const hipStream_t s = nullptr;
    hipLaunchKernelGGL(k_noalloc_jobs, dim3(1), dim3(1), 0, s, d_jobs, jobs);
    auto* jb = reinterpret_cast<unsigned char*>(d_jobs);
    launch_sums(reinterpret_cast<const SumJob*>(jb + offsetof(NoallocJobs, plain)), 2, big, d_sums, true, nullptr, s);
    launch_sums(reinterpret_cast<const SumJob*>(jb + offsetof(NoallocJobs, nyquist)), 2, big, d_sums, true, nullptr, s);
    launch_lag_search(reinterpret_cast<const LagJob*>(jb + offsetof(NoallocJobs, lag)), 1, n, ds->tw, d_sums, d_parts,
                      d_lag, s);
    long long h = 0;
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(&h, d_lag, sizeof h, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        set_error("visqol_hip_best_lag_noalloc: %s", hipGetErrorString(e));
        return VISQOL_E_BACKEND;
    }
    *lag = h;
    return VISQOL_OK;
} catch (...) {
    return exception_code();
}

}  // extern "C"
