// Serving layer: continuous batching of sort requests on one GPU.
//
// Requests (host buffers) are queued in arrival order.  A scheduler thread
// forms a ragged batch whenever one of `slots` in-flight slots is free: it
// takes pending requests in order while the batch stays within the slot's
// element / request caps and while the device pool (sized to the HBM
// admission budget) can hold the batch (data, plus merge scratch when a
// request is longer than one tile).  The batch is packed into the slot's
// pinned staging buffer, copied to its pool region, sorted as one ragged
// batch (one persistent launch over all chunks, merge passes for long
// requests) and copied back, all on the slot's own stream, so batches of
// different slots overlap on the device.  A completion thread waits for
// each batch's event in order, copies every result to its request's output
// buffer, stamps the completion time and releases the pool region, which
// admits the next requests.  A request larger than the whole budget or the
// slot caps is rejected at submission.
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
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    unsigned char* h_in = nullptr;      // pinned staging, capacity max_elems
    unsigned char* h_out = nullptr;
    int64_t* h_off = nullptr;           // pinned metadata
    int* h_list = nullptr;
    SegItem* h_items = nullptr;
    MergeTile* h_tiles = nullptr;
    int64_t* d_off = nullptr;           // device metadata
    int* d_list = nullptr;
    SegItem* d_items = nullptr;
    MergeTile* d_tiles = nullptr;
    int* d_counter = nullptr;   // two item counters of the persistent kernels
    size_t pool_off = 0;
    std::vector<std::shared_ptr<Request>> reqs;
    std::vector<int64_t> req_off;
    int64_t elems = 0;
    bool busy = false;
    int status = CS_OK;
};

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
    std::deque<int> inflight;                       // slot indices in launch order
    std::unordered_map<uint64_t, std::shared_ptr<Request>> all;
    uint64_t next_id = 1;
    bool stop = false;
    std::thread scheduler, completer;

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
};

namespace {

bool valid_dtype(int d) { return d >= CS_I32 && d <= CS_F64; }

// Called with s->mu held; returns the batch in `sl` or false.
bool form_batch(cs_server* s, Slot& sl) {
    sl.reqs.clear();
    sl.req_off.assign(1, 0);
    int64_t elems = 0;
    bool long_req = false;
    const int64_t chunk = large_kernels_for(s->dtype).chunk;
    for (const auto& r : s->pending) {
        if (int(sl.reqs.size()) >= s->max_reqs) break;
        if (elems + r->n > s->max_elems) break;
        const bool lr = long_req || r->n > chunk;
        const size_t need = size_t(elems + r->n) * s->esz * (lr ? 2 : 1);
        if (!s->pool.fits_ever(need) || need > s->pool.largest_free()) break;
        elems += r->n;
        long_req = lr;
        sl.reqs.push_back(r);
        sl.req_off.push_back(elems);
    }
    if (sl.reqs.empty()) return false;
    const size_t need = size_t(elems) * s->esz * (long_req ? 2 : 1);
    if (!s->pool.alloc(need, &sl.pool_off)) {
        sl.reqs.clear();
        return false;
    }
    for (size_t i = 0; i < sl.reqs.size(); ++i) s->pending.pop_front();
    sl.elems = elems;
    return true;
}

void launch_batch(cs_server* s, Slot& sl) {
    // pack inputs into pinned staging
    for (size_t i = 0; i < sl.reqs.size(); ++i)
        std::memcpy(sl.h_in + size_t(sl.req_off[i]) * s->esz, sl.reqs[i]->in, size_t(sl.reqs[i]->n) * s->esz);
    RaggedPrep prep;
    int rc = ragged_prepare(s->dtype, sl.req_off.data(), int64_t(sl.reqs.size()), &prep);
    if (rc == CS_OK && (int64_t(prep.chunk_off.size()) > s->max_chunks || int64_t(prep.tiles.size()) > s->max_tiles))
        rc = CS_E_CAPACITY;
    if (rc != CS_OK) { sl.status = rc; return; }
    std::memcpy(sl.h_off, prep.chunk_off.data(), prep.chunk_off.size() * sizeof(int64_t));
    std::memcpy(sl.h_list, prep.list.data(), prep.list.size() * sizeof(int));
    std::memcpy(sl.h_items, prep.items.data(), prep.items.size() * sizeof(SegItem));
    if (!prep.tiles.empty()) std::memcpy(sl.h_tiles, prep.tiles.data(), prep.tiles.size() * sizeof(MergeTile));
    unsigned char* data = s->d_pool + sl.pool_off;
    unsigned char* scratch = data + size_t(sl.elems) * s->esz;
    const size_t bytes = size_t(sl.elems) * s->esz;
    hipError_t e = hipMemcpyAsync(data, sl.h_in, bytes, hipMemcpyHostToDevice, sl.stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(sl.d_off, sl.h_off, prep.chunk_off.size() * sizeof(int64_t), hipMemcpyHostToDevice,
                           sl.stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(sl.d_list, sl.h_list, std::max<size_t>(prep.list.size(), 1) * sizeof(int),
                           hipMemcpyHostToDevice, sl.stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(sl.d_items, sl.h_items, std::max<size_t>(prep.items.size(), 1) * sizeof(SegItem),
                           hipMemcpyHostToDevice, sl.stream);
    if (e == hipSuccess && !prep.tiles.empty())
        e = hipMemcpyAsync(sl.d_tiles, sl.h_tiles, prep.tiles.size() * sizeof(MergeTile), hipMemcpyHostToDevice,
                           sl.stream);
    if (e == hipSuccess)
        e = ragged_launch(s->dtype, data, data, scratch, sl.d_off, sl.d_list, sl.d_items, int(prep.items.size()),
                          sl.d_counter, prep.shape, sl.d_tiles, int(prep.tiles.size()), prep.passes, sl.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(sl.h_out, data, bytes, hipMemcpyDeviceToHost, sl.stream);
    if (e == hipSuccess) e = hipEventRecord(sl.done, sl.stream);
    sl.status = e == hipSuccess ? CS_OK : int(e);
}

void scheduler_loop(cs_server* s) {
    std::unique_lock<std::mutex> lock(s->mu);
    for (;;) {
        int idx = -1;
        s->cv.wait(lock, [&] {
            if (s->pending.empty()) return s->stop;
            for (size_t i = 0; i < s->slots.size(); ++i)
                if (!s->slots[i].busy) { idx = int(i); return true; }
            return false;   // every slot in flight: wait for a completion
        });
        if (s->stop && s->pending.empty()) return;
        if (idx < 0) continue;
        Slot& sl = s->slots[size_t(idx)];
        if (!form_batch(s, sl)) {
            if (s->inflight.empty()) {
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
            // back-pressure: the pool frees up when a batch completes
            const int64_t seen = s->batches_done;
            s->cv.wait(lock, [&] { return s->batches_done != seen; });
            continue;
        }
        sl.busy = true;
        lock.unlock();
        launch_batch(s, sl);
        lock.lock();
        s->inflight.push_back(idx);
        s->batches += 1;
        s->batched_requests += int64_t(sl.reqs.size());
        s->max_inflight_seen = std::max<int64_t>(s->max_inflight_seen, int64_t(s->inflight.size()));
        s->cv.notify_all();
    }
}

void completer_loop(cs_server* s) {
    std::unique_lock<std::mutex> lock(s->mu);
    for (;;) {
        s->cv.wait(lock, [&] { return !s->inflight.empty() || (s->stop && s->pending.empty()); });
        if (s->inflight.empty()) return;
        const int idx = s->inflight.front();
        Slot& sl = s->slots[size_t(idx)];
        lock.unlock();
        int status = sl.status;
        if (status == CS_OK) {
            const hipError_t e = hipEventSynchronize(sl.done);
            if (e != hipSuccess) status = int(e);
        }
        const double t = now_s();
        if (status == CS_OK) {
            for (size_t i = 0; i < sl.reqs.size(); ++i)
                std::memcpy(sl.reqs[i]->out, sl.h_out + size_t(sl.req_off[i]) * s->esz, size_t(sl.reqs[i]->n) * s->esz);
        }
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
        s->inflight.pop_front();
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
    {
        std::lock_guard<std::mutex> lock(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    if (s->completer.joinable()) s->completer.join();
    for (auto& sl : s->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.h_in) (void)hipHostFree(sl.h_in);
        if (sl.h_out) (void)hipHostFree(sl.h_out);
        if (sl.h_off) (void)hipHostFree(sl.h_off);
        if (sl.h_list) (void)hipHostFree(sl.h_list);
        if (sl.h_items) (void)hipHostFree(sl.h_items);
        if (sl.h_tiles) (void)hipHostFree(sl.h_tiles);
        if (sl.d_off) (void)hipFree(sl.d_off);
        if (sl.d_list) (void)hipFree(sl.d_list);
        if (sl.d_items) (void)hipFree(sl.d_items);
        if (sl.d_tiles) (void)hipFree(sl.d_tiles);
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
    for (auto& sl : s->slots) {
        if (e != hipSuccess) break;
        e = hipStreamCreate(&sl.stream);   // default priority
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
        if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&sl.h_in), size_t(s->max_elems) * s->esz, 0);
        if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&sl.h_out), size_t(s->max_elems) * s->esz, 0);
        if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&sl.h_off), size_t(s->max_chunks + 1) * 8, 0);
        if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&sl.h_list), size_t(s->max_chunks + 1) * 4, 0);
        if (e == hipSuccess)
            e = hipHostMalloc(reinterpret_cast<void**>(&sl.h_items), size_t(s->max_chunks + 1) * sizeof(SegItem), 0);
        if (e == hipSuccess)
            e = hipHostMalloc(reinterpret_cast<void**>(&sl.h_tiles), size_t(s->max_tiles + 1) * sizeof(MergeTile), 0);
        if (e == hipSuccess) e = hipMalloc(&sl.d_off, size_t(s->max_chunks + 1) * 8);
        if (e == hipSuccess) e = hipMalloc(&sl.d_list, size_t(s->max_chunks + 1) * 4);
        if (e == hipSuccess) e = hipMalloc(&sl.d_items, size_t(s->max_chunks + 1) * sizeof(SegItem));
        if (e == hipSuccess) e = hipMalloc(&sl.d_tiles, size_t(s->max_tiles + 1) * sizeof(MergeTile));
        if (e == hipSuccess) e = hipMalloc(&sl.d_counter, 2 * sizeof(int));
    }
    if (e == hipSuccess && (seg_persistent_grid(dtype, kSegMulti, 1) <= 0 || seg_persistent_grid(dtype, kSegSingle, 1) <= 0))
        e = hipErrorNoDevice;
    if (e != hipSuccess) {
        destroy_server(s);
        return int(e);
    }
    s->scheduler = std::thread(scheduler_loop, s);
    s->completer = std::thread(completer_loop, s);
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
        if (!s->pending.empty() || !s->inflight.empty()) return false;
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
    return CS_OK;
}

int cs_server_destroy(cs_server* s) {
    destroy_server(s);
    return CS_OK;
}

}  // extern "C"
