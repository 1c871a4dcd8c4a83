// Serving layer: continuous batching of sort requests on one GPU.
//
// Requests (host buffers) are queued in arrival order.  A scheduler thread
// forms a ragged batch whenever one of `slots` in-flight slots is free: it
// scans the pending requests in order and takes every one that keeps the
// batch within the slot's element / request caps and within the largest free
// region of the device pool (sized to the HBM admission budget; a request
// longer than one tile also needs merge scratch of its own size).  A request
// that does not fit is passed over, so small requests behind a large one are
// not held up, but only a bounded number of times: one passed over
// kBypassLimit times admits nothing behind it until it fits, so it cannot
// starve.  The batch is packed into the slot's pinned staging buffer, copied
// to its pool region, sorted as one ragged batch (one persistent launch over
// all chunks, merge passes for long requests) and copied back, all on the
// slot's own stream, so batches of different slots overlap on the device.
// Each slot has its own thread that packs and launches the batches handed
// to it, waits for each batch's event, copies every result to its request's
// output buffer, stamps the completion time and releases the pool region:
// batches complete as they finish, not in launch order, and packing a large
// batch holds up neither the scheduler nor the other slots.  A batch whose launch failed after work was queued on the
// slot's stream drains that stream before its region and staging buffers
// are reused.  A request larger than the whole budget or the slot caps is
// rejected at submission.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/cusort_hip.h"
#include "admission.hpp"
#include "gen.hpp"
#include "pool.hpp"
#include "ragged.hpp"

using namespace cs;

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Request {
    uint64_t id = 0;
    const void* in = nullptr;
    void* out = nullptr;
    int64_t n = 0;
    double t_arrival = 0;
    double t_done = 0;
    std::atomic<int> done{0};
    int status = CS_OK;
    int bypassed = 0;   // times a later request was admitted ahead of this one
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    unsigned char* h_in = nullptr;      // pinned staging, capacity max_elems (rounded to 16 bytes)
    unsigned char* h_out = nullptr;
    unsigned char* h_meta = nullptr;    // pinned metadata of one batch (MetaLayout)
    unsigned char* d_meta = nullptr;    // its device copy
    int* d_counter = nullptr;   // two item counters of the persistent kernels
    size_t pool_off = 0;
    std::vector<std::shared_ptr<Request>> reqs;
    std::vector<int64_t> req_off;
    int64_t elems = 0;
    bool busy = false;       // holds a batch (from forming to completion)
    bool enqueued = false;   // some work of the batch was queued on the stream
    int inject = CS_OK;      // failure to inject into this batch's launch
    int status = CS_OK;
    std::thread worker;
};

constexpr int kBypassLimit = 16;       // times a request may be passed over
constexpr size_t kScanWindow = 1024;   // pending requests examined per batch formation

size_t align16(size_t b) { return (b + 15) & ~size_t(15); }

// The metadata of one batch (chunk offsets, bucket-ordered chunk list, work
// items, merge tiles) packed into one block, so that one copy moves it.
struct MetaLayout {
    size_t off = 0, list = 0, items = 0, tiles = 0, bytes = 0;
};

MetaLayout meta_layout(size_t offs, size_t list, size_t items, size_t tiles) {
    MetaLayout m;
    m.list = align16(offs * sizeof(int64_t));
    m.items = m.list + align16(std::max<size_t>(list, 1) * sizeof(int));
    m.tiles = m.items + align16(std::max<size_t>(items, 1) * sizeof(SegItem));
    m.bytes = m.tiles + align16(tiles * sizeof(MergeTile));
    return m;
}

}  // namespace

struct cs_server {
    int dtype = 0;
    size_t esz = 0;
    int64_t max_elems = 0;
    int max_reqs = 0;
    int64_t max_chunks = 0;
    int64_t max_tiles = 0;
    unsigned char* d_pool = nullptr;
    OffsetPool pool;
    std::vector<Slot> slots;

    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::shared_ptr<Request>> pending;
    int inflight = 0;                               // batches formed and not yet completed
    std::unordered_map<uint64_t, std::shared_ptr<Request>> all;
    uint64_t next_id = 1;
    bool stop = false;
    int inject_status = CS_OK;                      // test hook: fail the next launch with this status
    std::thread scheduler;

    // statistics
    std::vector<double> latencies;
    int64_t keys_done = 0;
    int64_t completed = 0;
    int64_t rejected = 0;
    int64_t batches = 0;
    int64_t batched_requests = 0;
    double first_arrival = -1;
    double last_done = 0;
    int64_t max_inflight_seen = 0;
    int64_t batches_done = 0;
    int64_t bypasses = 0;
};

namespace {

bool valid_dtype(int d) { return d >= CS_I32 && d <= CS_F64; }

// Called with s->mu held; returns the batch in `sl` or false.
bool form_batch(cs_server* s, Slot& sl) {
    AdmitLimits lim;
    lim.max_elems = s->max_elems;
    lim.max_reqs = s->max_reqs;
    lim.chunk = large_kernels_for(s->dtype).chunk;
    lim.esz = s->esz;
    lim.room = s->pool.largest_free();
    lim.window = kScanWindow;
    lim.bypass_limit = kBypassLimit;
    const AdmitResult r = select_batch(
        s->pending.size(), [&](size_t i) { return s->pending[i]->n; },
        [&](size_t i) { return s->pending[i]->bypassed; }, lim);
    if (r.taken.empty() || !s->pool.alloc(r.bytes, &sl.pool_off)) return false;
    sl.reqs.clear();
    sl.req_off.assign(1, 0);
    for (size_t i : r.taken) {
        sl.reqs.push_back(s->pending[i]);
        sl.req_off.push_back(sl.req_off.back() + s->pending[i]->n);
    }
    for (size_t i : r.bypassed) s->pending[i]->bypassed += 1;
    s->bypasses += int64_t(r.bypassed.size());
    for (size_t k = r.taken.size(); k-- > 0;) s->pending.erase(s->pending.begin() + long(r.taken[k]));
    sl.elems = r.elems;
    return true;
}

// Enqueues the batch staged in sl.h_in (sl.req_off, sl.elems, sl.pool_off
// set) on the slot's stream: copy-in, sort, copy-out, completion event.
// The copies are kernels reading and writing the host-coherent staging
// directly, so they run on the slot's own queue: a DMA engine shared by the
// slots would serialise one batch's copies behind another batch's.  Sets
// sl.enqueued once any work is queued; `inject` != CS_OK fails the batch
// right after its input copy (failure-injection hook).
int enqueue_batch(cs_server* s, Slot& sl, int inject) {
    sl.enqueued = false;
    RaggedPrep prep;
    int rc = ragged_prepare(s->dtype, sl.req_off.data(), int64_t(sl.req_off.size()) - 1, &prep);
    if (rc == CS_OK && (int64_t(prep.chunk_off.size()) > s->max_chunks || int64_t(prep.tiles.size()) > s->max_tiles))
        rc = CS_E_CAPACITY;
    if (rc != CS_OK) return rc;
    const MetaLayout m = meta_layout(prep.chunk_off.size(), prep.list.size(), prep.items.size(), prep.tiles.size());
    std::memcpy(sl.h_meta + m.off, prep.chunk_off.data(), prep.chunk_off.size() * sizeof(int64_t));
    std::memcpy(sl.h_meta + m.list, prep.list.data(), prep.list.size() * sizeof(int));
    std::memcpy(sl.h_meta + m.items, prep.items.data(), prep.items.size() * sizeof(SegItem));
    if (!prep.tiles.empty()) std::memcpy(sl.h_meta + m.tiles, prep.tiles.data(), prep.tiles.size() * sizeof(MergeTile));
    unsigned char* data = s->d_pool + sl.pool_off;
    const size_t bytes = size_t(sl.elems) * s->esz;
    unsigned char* scratch = data + bytes;
    // Rounded to 16 bytes for the wide copy kernel: the staging buffers are
    // allocated that large, and the pool block holds at least that much
    // (blocks are 256-byte multiples; bytes past the data land in the merge
    // scratch, which is written before it is read).
    const size_t copy = align16(bytes);
    hipError_t e = device_copy(data, sl.h_in, copy, sl.stream);
    sl.enqueued = e == hipSuccess;
    if (e == hipSuccess && inject != CS_OK) return inject;
    if (e == hipSuccess) e = device_copy(sl.d_meta, sl.h_meta, m.bytes, sl.stream);
    if (e == hipSuccess)
        e = ragged_launch(s->dtype, data, data, scratch, reinterpret_cast<const int64_t*>(sl.d_meta + m.off),
                          reinterpret_cast<const int*>(sl.d_meta + m.list),
                          reinterpret_cast<const SegItem*>(sl.d_meta + m.items), int(prep.items.size()),
                          sl.d_counter, prep.shape, reinterpret_cast<MergeTile*>(sl.d_meta + m.tiles),
                          int(prep.tiles.size()), prep.passes, sl.stream);
    if (e == hipSuccess) e = device_copy(sl.h_out, data, copy, sl.stream);
    if (e == hipSuccess) e = hipEventRecord(sl.done, sl.stream);
    return e == hipSuccess ? CS_OK : int(e);
}

void launch_batch(cs_server* s, Slot& sl, int inject) {
    // pack inputs into pinned staging
    for (size_t i = 0; i < sl.reqs.size(); ++i)
        std::memcpy(sl.h_in + size_t(sl.req_off[i]) * s->esz, sl.reqs[i]->in, size_t(sl.reqs[i]->n) * s->esz);
    sl.status = enqueue_batch(s, sl, inject);
}

// Forms batches and hands each to a free slot; the slot's own thread packs,
// launches and completes it, so packing a large batch does not hold up the
// forming of the next one.
void scheduler_loop(cs_server* s) {
    std::unique_lock<std::mutex> lock(s->mu);
    for (;;) {
        int idx = -1;
        s->cv.wait(lock, [&] {
            if (s->pending.empty()) return s->stop;
            for (size_t i = 0; i < s->slots.size(); ++i)
                if (!s->slots[i].busy) { idx = int(i); return true; }
            return false;   // every slot holds a batch: wait for a completion
        });
        if (s->stop && s->pending.empty()) return;
        if (idx < 0) continue;
        Slot& sl = s->slots[size_t(idx)];
        if (!form_batch(s, sl)) {
            if (s->inflight == 0) {
                // An empty pool always holds an admitted request (checked at
                // submission); fail the head rather than wait forever.
                auto r = s->pending.front();
                s->pending.pop_front();
                r->status = CS_E_BUDGET;
                r->t_done = now_s();
                r->done.store(1);
                s->cv.notify_all();
                continue;
            }
            // back-pressure: wait until a batch completes (the pool frees up)
            // or a new request arrives (it may fit what is free now)
            const int64_t seen = s->batches_done;
            const uint64_t seen_id = s->next_id;
            s->cv.wait(lock, [&] { return s->batches_done != seen || s->next_id != seen_id; });
            continue;
        }
        sl.busy = true;
        sl.inject = s->inject_status;
        s->inject_status = CS_OK;
        s->inflight += 1;
        s->batches += 1;
        s->batched_requests += int64_t(sl.reqs.size());
        s->max_inflight_seen = std::max<int64_t>(s->max_inflight_seen, s->inflight);
        s->cv.notify_all();
    }
}

// Thread of one slot: packs and launches each batch handed to the slot, then
// completes it as soon as its own event fires, whatever other slots do.
void slot_loop(cs_server* s, size_t idx) {
    Slot& sl = s->slots[idx];
    std::unique_lock<std::mutex> lock(s->mu);
    for (;;) {
        s->cv.wait(lock, [&] { return sl.busy || (s->stop && s->pending.empty()); });
        if (!sl.busy) return;
        const int inject = sl.inject;
        lock.unlock();
        launch_batch(s, sl, inject);
        int status = sl.status;
        if (status == CS_OK) {
            const hipError_t e = hipEventSynchronize(sl.done);
            if (e != hipSuccess) status = int(e);
        }
        // A failed batch may still have copies queued that read the staging
        // buffer or write its pool region: let them finish before either is
        // handed to another batch.
        if (status != CS_OK && sl.enqueued) (void)hipStreamSynchronize(sl.stream);
        if (status == CS_OK) {
            for (size_t i = 0; i < sl.reqs.size(); ++i)
                std::memcpy(sl.reqs[i]->out, sl.h_out + size_t(sl.req_off[i]) * s->esz, size_t(sl.reqs[i]->n) * s->esz);
        }
        const double t = now_s();
        lock.lock();
        for (auto& r : sl.reqs) {
            r->status = status;
            r->t_done = t;
            r->done.store(1);
            if (status == CS_OK) {
                s->latencies.push_back(t - r->t_arrival);
                s->keys_done += r->n;
                s->completed += 1;
                s->last_done = std::max(s->last_done, t);
            }
        }
        s->pool.release(sl.pool_off);
        sl.reqs.clear();
        sl.busy = false;
        s->inflight -= 1;
        s->batches_done += 1;
        s->cv.notify_all();
    }
}

void destroy_server(cs_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lock(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    if (s->scheduler.joinable()) s->scheduler.join();
    s->cv.notify_all();
    for (auto& sl : s->slots)
        if (sl.worker.joinable()) sl.worker.join();
    for (auto& sl : s->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.h_in) (void)hipHostFree(sl.h_in);
        if (sl.h_out) (void)hipHostFree(sl.h_out);
        if (sl.h_meta) (void)hipHostFree(sl.h_meta);
        if (sl.d_meta) (void)hipFree(sl.d_meta);
        if (sl.d_counter) (void)hipFree(sl.d_counter);
    }
    if (s->d_pool) (void)hipFree(s->d_pool);
    delete s;
}

}  // namespace

extern "C" {

double cs_server_now(void) { return now_s(); }

int cs_server_create(cs_server** out, int dtype, size_t budget_bytes, int32_t slots, int64_t max_batch_elems,
                     int32_t max_batch_requests) {
    if (!out || !valid_dtype(dtype) || budget_bytes == 0 || slots < 1 || slots > 64) return CS_E_INVALID_VALUE;
    if (max_batch_elems < 0 || max_batch_requests < 0) return CS_E_INVALID_VALUE;
    cs_server* s = new cs_server();
    s->dtype = dtype;
    s->esz = cs_dtype_size(dtype);
    s->max_elems = max_batch_elems ? max_batch_elems : (int64_t(1) << 22);
    s->max_reqs = max_batch_requests ? max_batch_requests : 4096;
    const int64_t chunk = large_kernels_for(dtype).chunk;
    const int64_t tile = large_kernels_for(dtype).tile;
    s->max_chunks = s->max_reqs + s->max_elems / chunk + 2;
    s->max_tiles = s->max_elems / tile + s->max_reqs + 2;
    s->pool.reset(budget_bytes);
    hipError_t e = hipMalloc(&s->d_pool, s->pool.capacity());
    s->slots.resize(size_t(slots));
    // Staging is host-coherent memory: the copy kernels read and write it
    // directly, with no host-side cache state to flush.
    const size_t staging = align16(size_t(s->max_elems) * s->esz);
    const size_t meta = meta_layout(size_t(s->max_chunks) + 1, size_t(s->max_chunks) + 1, size_t(s->max_chunks) + 1,
                                    size_t(s->max_tiles) + 1)
                            .bytes;
    for (auto& sl : s->slots) {
        if (e != hipSuccess) break;
        e = hipStreamCreate(&sl.stream);   // default priority
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
        if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&sl.h_in), staging, hipHostMallocCoherent);
        if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&sl.h_out), staging, hipHostMallocCoherent);
        if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&sl.h_meta), meta, hipHostMallocCoherent);
        // The host's first write to a page of pinned memory populates it
        // (about 0.35 ms per MiB): do that here rather than in the batches.
        if (e == hipSuccess) {
            std::memset(sl.h_in, 0, staging);
            std::memset(sl.h_out, 0, staging);
            std::memset(sl.h_meta, 0, meta);
        }
        if (e == hipSuccess) e = hipMalloc(&sl.d_meta, meta);
        if (e == hipSuccess) e = hipMalloc(&sl.d_counter, 2 * sizeof(int));
    }
    if (e == hipSuccess && (seg_persistent_grid(dtype, kSegMulti, 1) <= 0 || seg_persistent_grid(dtype, kSegSingle, 1) <= 0))
        e = hipErrorNoDevice;
    // One small batch per slot through the whole path, so that the first
    // requests do not pay one-time costs: loading each kernel's code (about
    // 0.7 ms at a kernel's first launch) and setting up the scratch of each
    // hardware queue (about 0.13 ms at the first launch there).  Lengths
    // cover a short segment, a one-tile segment and, when the pool holds it,
    // a segment just over one chunk (merge passes).
    std::vector<int64_t> warm = {0, 1, 1 + std::min<int64_t>(3000, s->max_elems - 1)};
    auto warm_bytes = [&](const std::vector<int64_t>& w) {
        bool lr = false;
        for (size_t i = 1; i < w.size(); ++i) lr = lr || w[i] - w[i - 1] > chunk;
        return OffsetPool::round_up(admit_bytes(w.back(), s->esz, lr));
    };
    warm.push_back(warm.back() + chunk + 1);
    if (warm.back() > s->max_elems || warm_bytes(warm) > s->pool.capacity()) warm.pop_back();
    if (warm_bytes(warm) > s->pool.capacity()) warm.resize(2);
    for (auto& sl : s->slots) {
        if (e != hipSuccess || warm_bytes(warm) > s->pool.capacity()) break;
        sl.req_off = warm;
        sl.elems = warm.back();
        sl.pool_off = 0;
        std::memset(sl.h_in, 0, align16(size_t(sl.elems) * s->esz));
        const int rc = enqueue_batch(s, sl, CS_OK);
        if (sl.enqueued) e = hipStreamSynchronize(sl.stream);
        // a runtime error is fatal; a batch outside small caps only skips the warm-up
        if (rc > 0) e = hipError_t(rc);
    }
    if (e != hipSuccess) {
        destroy_server(s);
        return int(e);
    }
    s->scheduler = std::thread(scheduler_loop, s);
    for (size_t i = 0; i < s->slots.size(); ++i) s->slots[i].worker = std::thread(slot_loop, s, i);
    *out = s;
    return CS_OK;
}

int cs_server_submit(cs_server* s, const void* host_in, void* host_out, int64_t n, double arrival_s,
                     uint64_t* ticket) {
    if (!s || n < 0 || (n > 0 && (!host_in || !host_out)) || !ticket) return CS_E_INVALID_VALUE;
    const bool long_req = n > large_kernels_for(s->dtype).chunk;
    if (n > s->max_elems || !s->pool.fits_ever(size_t(n) * s->esz * (long_req ? 2 : 1))) {
        std::lock_guard<std::mutex> lock(s->mu);
        s->rejected += 1;
        return CS_E_BUDGET;
    }
    auto r = std::make_shared<Request>();
    r->in = host_in;
    r->out = host_out;
    r->n = n;
    r->t_arrival = arrival_s >= 0 ? arrival_s : now_s();
    {
        std::lock_guard<std::mutex> lock(s->mu);
        r->id = s->next_id++;
        if (s->first_arrival < 0 || r->t_arrival < s->first_arrival) s->first_arrival = r->t_arrival;
        s->all[r->id] = r;
        s->pending.push_back(r);
    }
    *ticket = r->id;
    s->cv.notify_all();
    return CS_OK;
}

int cs_server_poll(cs_server* s, uint64_t ticket, int32_t* done, int32_t* status) {
    if (!s || !done) return CS_E_INVALID_VALUE;
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return CS_E_INVALID_VALUE;
    *done = it->second->done.load();
    if (status) *status = it->second->status;
    return CS_OK;
}

int cs_server_wait(cs_server* s, uint64_t ticket, double timeout_s) {
    if (!s) return CS_E_INVALID_VALUE;
    std::unique_lock<std::mutex> lock(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return CS_E_INVALID_VALUE;
    auto r = it->second;
    auto pred = [&] { return r->done.load() != 0; };
    if (timeout_s < 0) s->cv.wait(lock, pred);
    else if (!s->cv.wait_for(lock, std::chrono::duration<double>(timeout_s), pred)) return CS_E_AGAIN;
    return r->status;
}

int cs_server_latency(cs_server* s, uint64_t ticket, double* arrival_s, double* done_s) {
    if (!s) return CS_E_INVALID_VALUE;
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return CS_E_INVALID_VALUE;
    if (arrival_s) *arrival_s = it->second->t_arrival;
    if (done_s) *done_s = it->second->t_done;
    return CS_OK;
}

int cs_server_release(cs_server* s, uint64_t ticket) {
    if (!s) return CS_E_INVALID_VALUE;
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return CS_E_INVALID_VALUE;
    if (!it->second->done.load()) return CS_E_AGAIN;
    s->all.erase(it);
    return CS_OK;
}

int cs_server_drain(cs_server* s) {
    if (!s) return CS_E_INVALID_VALUE;
    std::unique_lock<std::mutex> lock(s->mu);
    s->cv.wait(lock, [&] {
        if (!s->pending.empty() || s->inflight != 0) return false;
        for (const auto& sl : s->slots)
            if (sl.busy) return false;
        return true;
    });
    return CS_OK;
}

int cs_server_get_stats(cs_server* s, cs_server_stats* st) {
    if (!s || !st) return CS_E_INVALID_VALUE;
    std::lock_guard<std::mutex> lock(s->mu);
    std::vector<double> l = s->latencies;
    std::sort(l.begin(), l.end());
    auto pct = [&](double q) { return l.empty() ? 0.0 : l[std::min(l.size() - 1, size_t(q * double(l.size())))]; };
    st->completed = s->completed;
    st->rejected = s->rejected;
    st->batches = s->batches;
    st->keys = s->keys_done;
    st->p50_s = pct(0.50);
    st->p99_s = pct(0.99);
    st->max_s = l.empty() ? 0.0 : l.back();
    double sum = 0;
    for (double x : l) sum += x;
    st->mean_s = l.empty() ? 0.0 : sum / double(l.size());
    const double span = s->last_done - s->first_arrival;
    st->keys_per_s = (span > 0 && s->first_arrival >= 0) ? double(s->keys_done) / span : 0.0;
    st->mean_batch_requests = s->batches ? double(s->batched_requests) / double(s->batches) : 0.0;
    st->peak_pool_bytes = int64_t(s->pool.peak());
    st->max_inflight_batches = s->max_inflight_seen;
    st->bypasses = s->bypasses;
    return CS_OK;
}

int cs_server_destroy(cs_server* s) {
    destroy_server(s);
    return CS_OK;
}

int cs_server_inject_launch_failure(cs_server* s, int32_t status) {
    if (!s || status == CS_OK) return CS_E_INVALID_VALUE;
    std::lock_guard<std::mutex> lock(s->mu);
    s->inject_status = status;
    return CS_OK;
}

int cs_server_pool_in_use(cs_server* s, int64_t* bytes) {
    if (!s || !bytes) return CS_E_INVALID_VALUE;
    std::lock_guard<std::mutex> lock(s->mu);
    *bytes = int64_t(s->pool.in_use());
    return CS_OK;
}

}  // extern "C"
