// Continuous-batching scoring server on one gfx950 device (visqol_hip.h).
//
// submit() copies a request's signals into a device slab right away, so the
// queued work is device resident, and admits it only while the summed
// device footprint of queued and running requests fits the budget (the
// caller blocks until it does).  Worker threads, each with one engine per
// mode on its own HIP stream, repeatedly take the queued requests of the
// class (mode, sample rate) at the head of the queue and score them as one
// batch; while one worker's batch runs, the next takes whatever has arrived,
// so batches overlap on the device.  Each request keeps its submit and
// completion times on the steady clock.
//
// A request's result does not depend on the batch it lands in: every stage
// works per pair, and the fast mode's reductions split each signal by its
// own length.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <system_error>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/visqol.h"
#include "../../include/visqol_hip.h"
#include "capi_internal.h"
#include "engine.h"

using namespace vq;
using namespace vq_capi;

namespace {

using Clock = std::chrono::steady_clock;

// First-fit allocator over one device allocation, 256-byte granules.
class DeviceSlab {
   public:
    bool init(size_t bytes) {
        size_ = (bytes + 255) & ~(size_t)255;
        if (hipMalloc(&base_, size_) != hipSuccess) {
            (void)hipGetLastError();
            base_ = nullptr;
            return false;
        }
        free_[0] = size_;
        return true;
    }
    ~DeviceSlab() {
        if (base_) (void)hipFree(base_);
    }
    size_t size() const { return size_; }
    // Offset of a free range of `bytes`, or SIZE_MAX.
    size_t alloc(size_t bytes) {
        bytes = std::max<size_t>(256, (bytes + 255) & ~(size_t)255);
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= bytes) {
                const size_t off = it->first, rest = it->second - bytes;
                free_.erase(it);
                if (rest) free_[off + bytes] = rest;
                return off;
            }
        return SIZE_MAX;
    }
    void release(size_t off, size_t bytes) {
        bytes = std::max<size_t>(256, (bytes + 255) & ~(size_t)255);
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + bytes == next->first) {
            bytes += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += bytes;
                return;
            }
        }
        free_[off] = bytes;
    }
    char* at(size_t off) const { return base_ + off; }

   private:
    char* base_ = nullptr;
    size_t size_ = 0;
    std::map<size_t, size_t> free_;  // offset -> length
};

// Pinned staging for request uploads: kStageBuffers buffers of
// kStageBytes, each held by one submitting thread for the length of its
// upload.  A request's bytes pass through its buffer in pieces of half the
// buffer, the halves alternating, so the thread copies one piece while the
// DMA of the previous one runs; the DMAs go on the server's copy streams
// (pinned sources copy at ~45 GB/s in such pieces, pageable ones at
// ~26 GB/s).  Requests of 32 MiB and up are cut into stripes uploaded by
// several threads at once (upload_striped): one thread's host copy
// (~30 GB/s) would otherwise bound them.
class Stager {
   public:
    static constexpr int kStageBuffers = 8;
    static constexpr size_t kStageBytes = (size_t)16 << 20;
    bool init() {
        for (int i = 0; i < kStageBuffers; i++) {
            Buf b;
            if (hipHostMalloc(&b.host, kStageBytes, hipHostMallocDefault) != hipSuccess) {
                (void)hipGetLastError();
                return false;
            }
            bufs_.push_back(b);
            for (auto& e : bufs_.back().done)
                if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
                    (void)hipGetLastError();
                    e = nullptr;
                    return false;
                }
            free_.push_back(i);
        }
        return true;
    }
    ~Stager() {
        for (auto& b : bufs_) {
            for (auto e : b.done)
                if (e) (void)hipEventDestroy(e);
            if (b.host) (void)hipHostFree(b.host);
        }
    }
    // Copies the segments (src, bytes) one after another to dst; returns
    // once every byte is on the device.
    hipError_t upload(char* dst, const std::pair<const char*, size_t>* seg, int nseg, hipStream_t stream) {
        int k;
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&] { return !free_.empty(); });
            k = free_.back();
            free_.pop_back();
        }
        Buf& b = bufs_[k];
        constexpr size_t half = kStageBytes / 2;
        hipError_t e = hipSuccess;
        bool pending[2] = {false, false};
        int h = 0;
        size_t out = 0;
        for (int i = 0; i < nseg && e == hipSuccess; i++)
            for (size_t at = 0; at < seg[i].second && e == hipSuccess;) {
                const size_t len = std::min(half, seg[i].second - at);
                if (pending[h]) e = hipEventSynchronize(b.done[h]);  // that half's last DMA is done
                if (e != hipSuccess) break;
                memcpy(b.host + h * half, seg[i].first + at, len);
                e = hipMemcpyAsync(dst + out, b.host + h * half, len, hipMemcpyHostToDevice, stream);
                if (e == hipSuccess) e = hipEventRecord(b.done[h], stream);
                pending[h] = e == hipSuccess;
                at += len;
                out += len;
                h ^= 1;
            }
        // the stream runs the DMAs in order: the last one recorded ends last;
        // after a failed call nothing is known of the pieces, so the buffer
        // is returned only once the whole stream has drained
        if (e != hipSuccess) {
            (void)hipGetLastError();
            (void)hipStreamSynchronize(stream);
        } else if (pending[h ^ 1]) {
            e = hipEventSynchronize(b.done[h ^ 1]);
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            free_.push_back(k);
        }
        cv_.notify_one();
        return e;
    }

    // A large request: its bytes cut into `parts` stripes (256-byte
    // aligned), each uploaded as above through a staging buffer of its own
    // on a host thread of its own (the caller's the first), the stripes'
    // DMAs alternating between two copy streams, so several host copies
    // feed the link at once.  Returns once every stripe is on the device.
    hipError_t upload_striped(char* dst, const std::pair<const char*, size_t>* seg, int nseg,
                              const hipStream_t* streams, int parts) {
        size_t total = 0;
        for (int i = 0; i < nseg; i++) total += seg[i].second;
        std::vector<hipError_t> err(parts, hipSuccess);
        auto stripe = [&](int t) {
            const size_t a = t == 0 ? 0 : (total * t / parts) & ~(size_t)255;
            const size_t b = t + 1 == parts ? total : (total * (t + 1) / parts) & ~(size_t)255;
            std::vector<std::pair<const char*, size_t>> sub;
            size_t pos = 0;
            for (int i = 0; i < nseg; i++) {
                const size_t s0 = pos, s1 = pos + seg[i].second;
                const size_t lo = std::max(a, s0), hi = std::min(b, s1);
                if (lo < hi) sub.emplace_back(seg[i].first + (lo - s0), hi - lo);
                pos = s1;
            }
            if (!sub.empty()) err[t] = upload(dst + a, sub.data(), (int)sub.size(), streams[t & 1]);
        };
        std::vector<std::thread> ts;
        ts.reserve(parts);  // only thread creation can throw below
        for (int t = 1; t < parts; t++) {
            try {
                ts.emplace_back(stripe, t);
            } catch (const std::system_error&) {
                stripe(t);
            }
        }
        stripe(0);
        for (auto& th : ts) th.join();
        for (hipError_t e : err)
            if (e != hipSuccess) return e;
        return hipSuccess;
    }

   private:
    struct Buf {
        char* host = nullptr;
        hipEvent_t done[2] = {nullptr, nullptr};
    };
    std::vector<Buf> bufs_;
    std::vector<int> free_;
    std::mutex mu_;
    std::condition_variable cv_;
};

struct Request {
    uint64_t id = 0;
    int mode = 0;
    uint32_t sr = 0;
    size_t ref_len = 0, deg_len = 0;
    size_t slab_off = 0, slab_bytes = 0, footprint = 0;
    // submit called, admitted (footprint and slab space granted), signals
    // on the device and queued, batch started, batch finished
    Clock::time_point t_submit, t_admit, t_queued, t_run, t_done;
    uint32_t batch_size = 0;
    PairResult res;
    bool done = false;
};

}  // namespace

struct visqol_hip_server {
    visqol_options opts{};
    visqol_hip_server_config cfg{};
    DeviceSlab slab;
    Stager stager;
    hipStream_t copy_stream = nullptr, copy_stream2 = nullptr;
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_space;
    std::deque<Request*> queue;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> all;
    size_t inflight = 0;  // summed footprint of queued and running requests
    uint64_t next_id = 1;
    bool stop = false;
    int device = 0;
    std::vector<std::thread> workers;
    std::string worker_error;
    uint64_t batches = 0, batched_pairs = 0;

    void work();
};

namespace {

// Device bytes a request holds while queued or running: its signals plus
// the scoring working set, dominated by the complex lag-search buffer of up
// to 4·max(len) elements and the per-stage scratch (an upper estimate).
size_t footprint_of(size_t ref_len, size_t deg_len) {
    const size_t big = std::max(ref_len, deg_len);
    return 8 * (ref_len + deg_len) + 16 * 4 * big + 16 * (ref_len + deg_len);
}

}  // namespace

void visqol_hip_server::work() {
    (void)hipSetDevice(device);
    std::unique_ptr<Engine> eng[2];
    EngineOptions eo[2];
    for (;;) {
        std::vector<Request*> batch;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] { return stop || !queue.empty(); });
            if (stop && queue.empty()) return;
            const int mode = queue.front()->mode;
            const uint32_t sr = queue.front()->sr;
            for (auto it = queue.begin(); it != queue.end() && batch.size() < cfg.max_batch;) {
                if ((*it)->mode == mode && (*it)->sr == sr) {
                    batch.push_back(*it);
                    it = queue.erase(it);
                } else {
                    ++it;
                }
            }
            if (!queue.empty()) cv_work.notify_one();
        }
        const int mode = batch.front()->mode;
        int rc = VISQOL_OK;
        std::string err;
        std::vector<PairResult> out;
        {
            try {
                if (!eng[mode]) {
                    visqol_options o = opts;
                    o.mode = (uint32_t)mode;
                    rc = build_engine(o, eo[mode], eng[mode], false);
                    if (rc != VISQOL_OK) throw BackendError{rc, "cannot build the scoring engine"};
                }
                std::vector<PairInput> in(batch.size());
                out.resize(batch.size());
                for (size_t i = 0; i < batch.size(); i++) {
                    const Request* r = batch[i];
                    const double* ref = (const double*)slab.at(r->slab_off);
                    in[i] = PairInput{ref, (long long)r->ref_len, ref + r->ref_len, (long long)r->deg_len, true};
                }
                const Clock::time_point t_run = Clock::now();
                for (Request* r : batch) r->t_run = t_run;
                eng[mode]->run(in.data(), (int)in.size(), batch.front()->sr, out.data(), false);
            } catch (const BackendError& e) {
                rc = e.code;
                err = e.msg;
            } catch (const std::bad_alloc&) {
                rc = VISQOL_E_OUT_OF_MEMORY;
                err = "OutOfMemory";
            } catch (const std::exception& e) {
                // e.g. std::system_error from the engine's host threads: the
                // batch fails, the worker and its waiters carry on
                rc = VISQOL_E_BACKEND;
                err = e.what();
            }
        }
        const Clock::time_point now = Clock::now();
        {
            std::lock_guard<std::mutex> lk(mu);
            for (size_t i = 0; i < batch.size(); i++) {
                Request* r = batch[i];
                if (rc == VISQOL_OK) r->res = std::move(out[i]);
                r->batch_size = (uint32_t)batch.size();
                if (rc != VISQOL_OK) {
                    r->res = PairResult();
                    r->res.status = rc;
                    r->res.error = err;
                }
                r->t_done = now;
                r->done = true;
                slab.release(r->slab_off, r->slab_bytes);
                inflight -= r->footprint;
            }
            batches++;
            batched_pairs += batch.size();
        }
        cv_done.notify_all();
        cv_space.notify_all();
    }
}

extern "C" {

void visqol_hip_server_config_default(visqol_hip_server_config* c) {
    if (!c) return;
    c->workers = 3;
    c->max_batch = 256;
    c->slab_bytes = (uint64_t)16 << 30;
    c->max_inflight_bytes = (uint64_t)96 << 30;
}

int visqol_hip_server_create(const visqol_options* opts, const visqol_hip_server_config* cfg,
                             visqol_hip_server** out) try {
    if (!out) return invalid("out is NULL");
    *out = nullptr;
    visqol_options o;
    visqol_options_default(&o);
    if (opts) o = *opts;
    o.backend = VISQOL_BACKEND_HIP;
    if (o.precision > VISQOL_PRECISION_AUTO) return invalid("invalid precision");
    visqol_hip_server_config c;
    visqol_hip_server_config_default(&c);
    if (cfg) c = *cfg;
    if (c.workers == 0 || c.workers > 16 || c.max_batch == 0 || c.slab_bytes == 0 || c.max_inflight_bytes == 0)
        return invalid("invalid server configuration");
    if (o.model_path && o.model_path[0])
        return invalid("a server scores both modes with the embedded models; model_path must be NULL");
    std::string why;
    if (!device_ok(why)) {
        set_error("%s", why.c_str());
        return VISQOL_E_BACKEND;
    }
    // Both modes' models must load before any request is taken.
    for (uint32_t mode = 0; mode < 2; mode++) {
        visqol_options m = o;
        m.mode = mode;
        EngineOptions eo;
        std::unique_ptr<Engine> probe;
        const int rc = build_engine(m, eo, probe);
        if (rc != VISQOL_OK) return rc;
    }
    auto* s = new (std::nothrow) visqol_hip_server();
    if (!s) {
        set_error("OutOfMemory");
        return VISQOL_E_OUT_OF_MEMORY;
    }
    s->opts = o;
    s->opts.model_path = nullptr;
    s->cfg = c;
    (void)hipGetDevice(&s->device);
    if (!s->slab.init(c.slab_bytes)) {
        set_error("cannot allocate a %llu-byte request slab", (unsigned long long)c.slab_bytes);
        delete s;
        return VISQOL_E_OUT_OF_MEMORY;
    }
    if (!s->stager.init()) {
        set_error("cannot allocate pinned staging buffers");
        delete s;
        return VISQOL_E_OUT_OF_MEMORY;
    }
    if (hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&s->copy_stream2, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        set_error("cannot create a HIP stream");
        delete s;
        return VISQOL_E_BACKEND;
    }
    try {
        for (uint32_t w = 0; w < c.workers; w++) s->workers.emplace_back([s] { s->work(); });
    } catch (const std::system_error&) {
        visqol_hip_server_destroy(s);  // stops and joins the workers already started
        set_error("cannot start %u server worker threads", c.workers);
        return VISQOL_E_BACKEND;
    }
    *out = s;
    return VISQOL_OK;
} catch (...) {
    return exception_code();
}

int visqol_hip_server_submit(visqol_hip_server* s, uint32_t mode, const double* ref, size_t ref_len,
                             const double* deg, size_t deg_len, uint32_t sample_rate, uint64_t* ticket) try {
    const Clock::time_point t0 = Clock::now();
    if (!s || !ticket) return invalid("NULL argument");
    if (mode > 1 || sample_rate == 0) return invalid("invalid mode or sample rate");
    if ((!ref && ref_len) || (!deg && deg_len)) return invalid("NULL signal with nonzero length");
    auto r = std::make_unique<Request>();
    r->mode = (int)mode;
    r->sr = sample_rate;
    r->ref_len = ref_len;
    r->deg_len = deg_len;
    r->slab_bytes = sizeof(double) * (ref_len + deg_len);
    r->footprint = footprint_of(ref_len, deg_len);
    r->t_submit = t0;
    if (r->slab_bytes > s->slab.size()) {
        set_error("request of %zu bytes exceeds the %zu-byte slab", r->slab_bytes, s->slab.size());
        return VISQOL_E_OUT_OF_MEMORY;
    }
    Request* rp = r.get();
    {
        // Admission: wait until the request's footprint and signals fit (a
        // request is always admitted when nothing else is in flight).
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_space.wait(lk, [&] {
            if (s->inflight != 0 && s->inflight + rp->footprint > s->cfg.max_inflight_bytes) return false;
            rp->slab_off = s->slab.alloc(rp->slab_bytes);
            return rp->slab_off != SIZE_MAX;
        });
        s->inflight += rp->footprint;
        rp->id = s->next_id++;
    }
    rp->t_admit = Clock::now();
    // Upload outside the lock; the request is queued once its signals are
    // on the device.
    (void)hipSetDevice(s->device);
    const std::pair<const char*, size_t> seg[2] = {{(const char*)ref, sizeof(double) * ref_len},
                                                   {(const char*)deg, sizeof(double) * deg_len}};
    // requests of 32 MiB and up (pairs longer than ~35 s at 48 kHz) go up
    // in stripes of at least 16 MiB on up to four threads
    const size_t bytes = seg[0].second + seg[1].second;
    const hipStream_t streams[2] = {s->copy_stream, s->copy_stream2};
    const int parts = (int)std::min<size_t>(4, bytes / ((size_t)16 << 20));
    const hipError_t e = parts >= 2 ? s->stager.upload_striped(s->slab.at(rp->slab_off), seg, 2, streams, parts)
                                    : s->stager.upload(s->slab.at(rp->slab_off), seg, 2, s->copy_stream);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        std::lock_guard<std::mutex> lk(s->mu);
        s->slab.release(rp->slab_off, rp->slab_bytes);
        s->inflight -= rp->footprint;
        s->cv_space.notify_all();
        set_error("upload failed: %s", hipGetErrorString(e));
        return VISQOL_E_BACKEND;
    }
    rp->t_queued = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        *ticket = rp->id;
        s->queue.push_back(rp);
        s->all[rp->id] = std::move(r);
    }
    s->cv_work.notify_one();
    return VISQOL_OK;
} catch (...) {
    return exception_code();
}

int visqol_hip_server_wait(visqol_hip_server* s, uint64_t ticket, visqol_result** out, double* latency_ms) try {
    visqol_hip_request_timing t;
    const int rc = visqol_hip_server_wait_timed(s, ticket, out, &t);
    if (latency_ms) *latency_ms = t.total_ms;
    return rc;
} catch (...) {
    return exception_code();
}

int visqol_hip_server_wait_timed(visqol_hip_server* s, uint64_t ticket, visqol_result** out,
                                 visqol_hip_request_timing* timing) try {
    if (timing) memset(timing, 0, sizeof *timing);
    if (!s || !out) return invalid("NULL argument");
    *out = nullptr;
    std::unique_ptr<Request> r;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->all.find(ticket);
        if (it == s->all.end()) return invalid("unknown ticket");
        Request* rp = it->second.get();
        s->cv_done.wait(lk, [&] { return rp->done; });
        r = std::move(it->second);
        s->all.erase(it);
    }
    if (timing) {
        auto ms = [](Clock::time_point a, Clock::time_point b) {
            return std::chrono::duration<double, std::milli>(b - a).count();
        };
        timing->admit_ms = ms(r->t_submit, r->t_admit);
        timing->upload_ms = ms(r->t_admit, r->t_queued);
        // a batch that failed before it started has no start time
        const Clock::time_point run = r->t_run < r->t_queued ? r->t_done : r->t_run;
        timing->queue_ms = ms(r->t_queued, run);
        timing->run_ms = ms(run, r->t_done);
        timing->total_ms = ms(r->t_submit, r->t_done);
        timing->batch_size = r->batch_size;
    }
    if (r->res.status != VISQOL_OK) {
        set_error("%s", r->res.error.c_str());
        return r->res.status;
    }
    *out = box(r->res);
    if (!*out) {
        set_error("OutOfMemory");
        return VISQOL_E_OUT_OF_MEMORY;
    }
    return VISQOL_OK;
} catch (...) {
    return exception_code();
}

int visqol_hip_server_stats(visqol_hip_server* s, uint64_t* batches, uint64_t* pairs, uint64_t* inflight_bytes) try {
    if (!s) return invalid("NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    if (batches) *batches = s->batches;
    if (pairs) *pairs = s->batched_pairs;
    if (inflight_bytes) *inflight_bytes = s->inflight;
    return VISQOL_OK;
} catch (...) {
    return exception_code();
}

void visqol_hip_server_destroy(visqol_hip_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto& t : s->workers) t.join();
    if (s->copy_stream) (void)hipStreamDestroy(s->copy_stream);
    if (s->copy_stream2) (void)hipStreamDestroy(s->copy_stream2);
    delete s;
}

}  // extern "C"
