/*
 * jxs_server.cpp -- the serving loop: frames of many live streams, already
 * in HBM, are admitted by HBM footprint into a budget allocated once, grouped
 * by continuous batching into launches of the batched encode, and run on up
 * to four HIP streams at once.  A worker thread owns the device side: it
 * reaps finished batches (releasing their memory and publishing completion
 * records), then forms the next batch from the queue as soon as a stream is
 * free.
 *
 * Batch memory: one block per batch, first fit from the budget, holding the
 * descriptor image (frame descriptors and transform job lists, uploaded with
 * one asynchronous copy from pinned staging), the error words, the packing
 * units and every frame's work buffers.  A batch's block returns to the
 * budget when its completion event has fired.
 */
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "jxs_enc.h"

namespace jxs {

static uint64_t clock_ns() {
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

/* First-fit allocator over byte offsets [0, size), every block a multiple of
 * 256 bytes; free neighbours merge. */
class Arena {
  public:
    void init(uint64_t size) {
        free_.clear();
        if (size)
            free_[0] = size;
    }
    bool alloc(uint64_t n, uint64_t& off) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n)
                continue;
            off = it->first;
            const uint64_t rest = it->second - n;
            free_.erase(it);
            if (rest)
                free_[off + n] = rest;
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t n) {
        auto it = free_.emplace(off, n).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
    uint64_t largest() const {
        uint64_t m = 0;
        for (const auto& f : free_)
            m = f.second > m ? f.second : m;
        return m;
    }

  private:
    std::map<uint64_t, uint64_t> free_;
};

struct Request {
    svt_jpeg_xs_amd_frame_spec_t spec;
    const DevGeom* geom;
    uint64_t tag, submit_ns;
};

/* One launch slot: a stream and what its batch in flight holds. */
struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    uint8_t* staging = nullptr; /* pinned: the descriptor image */
    uint64_t staging_cap = 0;
    uint32_t* h_status = nullptr; /* pinned: error word of every frame */
    std::vector<uint8_t> img;
    std::vector<Request> reqs;
    uint64_t off = 0, bytes = 0, launch_ns = 0;
    bool busy = false;
    bool failed = false;
};

}  // namespace jxs

struct svt_jpeg_xs_amd_server {
    svt_jpeg_xs_amd_server_config_t cfg;
    int device = 0;
    uint8_t* d_base = nullptr;
    jxs::Arena arena;
    std::vector<jxs::Slot> slots;
    std::deque<jxs::Request> queue;
    std::deque<svt_jpeg_xs_amd_completion_t> done;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::thread worker;
    bool stop = false;
    uint32_t inflight = 0;
    svt_jpeg_xs_amd_server_stats_t st = {};
};

namespace jxs {

/* Publishes the completions of a finished slot and frees its memory (lock
 * held). */
static void finish_slot(svt_jpeg_xs_amd_server_t* s, Slot& sl, bool ok) {
    const uint64_t t = clock_ns();
    const uint32_t n = (uint32_t)sl.reqs.size();
    for (uint32_t k = 0; k < n; k++) {
        svt_jpeg_xs_amd_completion_t c;
        c.tag = sl.reqs[k].tag;
        c.status = !ok || sl.failed ? (int32_t)SvtJxsErrorUndefined
                   : sl.h_status[k] ? (int32_t)SvtJxsErrorEncodeFrameError
                                    : (int32_t)SvtJxsErrorNone;
        c.batch_frames = n;
        c.submit_ns = sl.reqs[k].submit_ns;
        c.launch_ns = sl.launch_ns;
        c.done_ns = t;
        s->done.push_back(c);
    }
    s->st.completed += n;
    if (sl.bytes)
        s->arena.release(sl.off, sl.bytes);
    s->st.hbm_in_use -= sl.bytes;
    sl.reqs.clear();
    sl.bytes = 0;
    sl.busy = false;
    sl.failed = false;
    s->inflight--;
}

/* Checks every busy slot's event; returns whether any finished (lock held). */
static bool reap(svt_jpeg_xs_amd_server_t* s) {
    bool any = false;
    for (auto& sl : s->slots) {
        if (!sl.busy)
            continue;
        const hipError_t q = hipEventQuery(sl.done);
        if (q == hipErrorNotReady)
            continue;
        finish_slot(s, sl, q == hipSuccess);
        any = true;
    }
    if (any)
        s->cv_done.notify_all();
    return any;
}

/* Takes queued frames in arrival order while they fit in one block of the
 * budget (lock held).  Returns the block size, 0 when none fits now. */
static uint64_t take_batch(svt_jpeg_xs_amd_server_t* s, std::vector<Request>& out) {
    BatchSizer sz;
    const uint64_t room = s->arena.largest();
    while (!s->queue.empty() && out.size() < s->cfg.max_batch) {
        BatchSizer t = sz;
        t.add(*s->queue.front().geom);
        if (t.bytes() > room)
            break;
        sz = t;
        out.push_back(s->queue.front());
        s->queue.pop_front();
    }
    return out.empty() ? 0 : sz.bytes();
}

/* Lays out and launches the slot's batch on its stream (lock not held). */
static bool launch_slot(Slot& sl, uint8_t* base) {
    std::vector<svt_jpeg_xs_amd_frame_spec_t> specs(sl.reqs.size());
    for (size_t k = 0; k < sl.reqs.size(); k++)
        specs[k] = sl.reqs[k].spec;
    BatchLaunch L;
    if (!batch_place(specs.data(), (uint32_t)specs.size(), base, L, sl.img))
        return false;
    if (sl.img.size() > sl.staging_cap) {
        if (sl.staging)
            (void)hipHostFree(sl.staging);
        sl.staging = nullptr;
        sl.staging_cap = 0;
        const uint64_t cap = sl.img.size() * 2;
        if (hipHostMalloc((void**)&sl.staging, cap, 0) != hipSuccess)
            return false;
        sl.staging_cap = cap;
    }
    memcpy(sl.staging, sl.img.data(), sl.img.size());
    return hipMemcpyAsync(base, sl.staging, sl.img.size(), hipMemcpyHostToDevice, sl.stream) == hipSuccess &&
           batch_launch(L, sl.stream, nullptr) == hipSuccess &&
           hipMemcpyAsync(sl.h_status, L.status, L.nfr * 4, hipMemcpyDeviceToHost, sl.stream) == hipSuccess;
}

static void worker_main(svt_jpeg_xs_amd_server_t* s) {
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        bool progress = reap(s);
        Slot* free_slot = nullptr;
        for (auto& sl : s->slots)
            if (!sl.busy) {
                free_slot = &sl;
                break;
            }
        if (free_slot && !s->queue.empty()) {
            std::vector<Request> reqs;
            const uint64_t bytes = take_batch(s, reqs);
            uint64_t off = 0;
            if (bytes && s->arena.alloc(bytes, off)) {
                Slot& sl = *free_slot;
                sl.reqs.swap(reqs);
                sl.off = off;
                sl.bytes = bytes;
                sl.busy = true;
                s->inflight++;
                s->st.batches++;
                if (sl.reqs.size() > s->st.largest_batch)
                    s->st.largest_batch = sl.reqs.size();
                s->st.hbm_in_use += bytes;
                if (s->st.hbm_in_use > s->st.hbm_peak)
                    s->st.hbm_peak = s->st.hbm_in_use;
                lk.unlock();
                sl.launch_ns = clock_ns();
                const bool ok = launch_slot(sl, s->d_base + off);
                /* the event is recorded even after a failed launch so the
                 * slot is reaped normally */
                sl.failed = !ok;
                (void)hipEventRecord(sl.done, sl.stream);
                lk.lock();
                progress = true;
            } else if (!reqs.empty()) {
                /* cannot happen (take_batch only takes what fits); put the
                 * frames back in order */
                for (auto it = reqs.rbegin(); it != reqs.rend(); ++it)
                    s->queue.push_front(*it);
            }
        }
        if (s->stop && s->queue.empty() && s->inflight == 0)
            break;
        if (progress)
            continue;
        if (s->inflight)
            s->cv_work.wait_for(lk, std::chrono::microseconds(20));
        else
            s->cv_work.wait(lk, [&] { return s->stop || !s->queue.empty(); });
    }
    s->cv_done.notify_all();
}

static void server_free(svt_jpeg_xs_amd_server_t* s) {
    for (auto& sl : s->slots) {
        if (sl.stream)
            (void)hipStreamSynchronize(sl.stream);
        if (sl.done)
            (void)hipEventDestroy(sl.done);
        if (sl.stream)
            (void)hipStreamDestroy(sl.stream);
        if (sl.staging)
            (void)hipHostFree(sl.staging);
        if (sl.h_status)
            (void)hipHostFree(sl.h_status);
    }
    if (s->d_base)
        (void)hipFree(s->d_base);
    delete s;
}

}  // namespace jxs

using namespace jxs;

extern "C" {

PREFIX_API svt_jpeg_xs_amd_server_t* svt_jpeg_xs_amd_server_create(const svt_jpeg_xs_amd_server_config_t* cfg) {
    if (!cfg || cfg->hbm_budget < 256 || cfg->max_batch < 1 || cfg->max_batch > 4096 || cfg->max_inflight < 1 ||
        cfg->max_inflight > 4)
        return nullptr;
    auto* s = new (std::nothrow) svt_jpeg_xs_amd_server;
    if (!s)
        return nullptr;
    s->cfg = *cfg;
    s->cfg.hbm_budget &= ~(uint64_t)255;
    s->st.hbm_budget = s->cfg.hbm_budget;
    if (hipGetDevice(&s->device) != hipSuccess || hipMalloc(&s->d_base, s->cfg.hbm_budget) != hipSuccess) {
        s->d_base = nullptr;
        server_free(s);
        return nullptr;
    }
    s->arena.init(s->cfg.hbm_budget);
    s->slots.resize(cfg->max_inflight);
    for (auto& sl : s->slots) {
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess ||
            hipHostMalloc((void**)&sl.h_status, (size_t)cfg->max_batch * 4, 0) != hipSuccess) {
            server_free(s);
            return nullptr;
        }
        /* staging for the descriptor image of a full batch of frames with up
         * to four components (one descriptor and two transform jobs each);
         * it grows if a batch needs more */
        sl.staging_cap = align256((uint64_t)cfg->max_batch * sizeof(JxsFrameDesc)) +
                         2 * align256((uint64_t)cfg->max_batch * 8 * (sizeof(JxsDwtJob) + 4)) + 12 * 256;
        if (hipHostMalloc((void**)&sl.staging, sl.staging_cap, 0) != hipSuccess) {
            sl.staging = nullptr;
            server_free(s);
            return nullptr;
        }
    }
    try {
        s->worker = std::thread(worker_main, s);
    } catch (...) {
        server_free(s);
        return nullptr;
    }
    return s;
}

PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_amd_server_submit(svt_jpeg_xs_amd_server_t* s,
                                                           const svt_jpeg_xs_amd_frame_spec_t* frame, uint64_t tag) {
    if (!s || !frame)
        return SvtJxsErrorBadParameter;
    const DevGeom* g = spec_geom(*frame);
    if (!g)
        return SvtJxsErrorBadParameter;
    BatchSizer one;
    one.add(*g);
    if (one.bytes() > s->cfg.hbm_budget)
        return SvtJxsErrorInsufficientResources;
    Request r;
    r.spec = *frame;
    r.geom = g;
    r.tag = tag;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop)
            return SvtJxsErrorBadParameter;
        r.submit_ns = clock_ns();
        s->queue.push_back(r);
        s->st.submitted++;
    }
    s->cv_work.notify_one();
    return SvtJxsErrorNone;
}

PREFIX_API uint32_t svt_jpeg_xs_amd_server_poll(svt_jpeg_xs_amd_server_t* s, svt_jpeg_xs_amd_completion_t* out,
                                                uint32_t max, uint32_t blocking) {
    if (!s || !out || !max)
        return 0;
    std::unique_lock<std::mutex> lk(s->mu);
    if (blocking)
        s->cv_done.wait(lk, [&] { return !s->done.empty() || (s->queue.empty() && s->inflight == 0); });
    uint32_t n = 0;
    while (n < max && !s->done.empty()) {
        out[n++] = s->done.front();
        s->done.pop_front();
    }
    return n;
}

PREFIX_API SvtJxsErrorType_t svt_jpeg_xs_amd_server_stats(svt_jpeg_xs_amd_server_t* s,
                                                          svt_jpeg_xs_amd_server_stats_t* stats) {
    if (!s || !stats)
        return SvtJxsErrorBadParameter;
    std::lock_guard<std::mutex> lk(s->mu);
    *stats = s->st;
    stats->queued = (uint32_t)s->queue.size();
    stats->inflight = s->inflight;
    return SvtJxsErrorNone;
}

PREFIX_API uint64_t svt_jpeg_xs_amd_server_now_ns(void) { return clock_ns(); }

PREFIX_API void svt_jpeg_xs_amd_server_destroy(svt_jpeg_xs_amd_server_t* s) {
    if (!s)
        return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable())
        s->worker.join();
    server_free(s);
}

}  // extern "C"
