// Serving form: a request loop that keeps meshes and textures resident in HBM,
// admits render requests against a budget of colour + depth target memory,
// batches whatever is admitted into one set of launches, and reads each frame
// back into pinned host memory on a second stream while the next batch draws.
//
// No counterpart in the reference (its main loop draws one frame at a time,
// main.go:222-227); the frame itself is the reference's Draw, bit for bit,
// whether a request is drawn alone or in a batch.
//
// Threads: callers submit and wait; one worker thread owns the context (every
// HIP call on it after gr_server_create). The context must not be used
// directly while a server runs on it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "gorender_hip.h"
#include "gr_error.h"
#include "gr_internal.h"
#include "gr_launch.h"

namespace {

using gr_internal::fail;

int64_t now_ns() {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

// Colour + depth of one request: device target (W*H*8 bytes) and pinned host
// copy (W*H*8 bytes, or W*H*4 when the server reads back colour only).
struct Target {
    void* dev = nullptr;
    void* host = nullptr;
    size_t bytes = 0;       // device bytes (what the budget counts)
    size_t host_bytes = 0;
};

struct Request {
    int64_t ticket = 0;
    gr_frame frame{};
    std::vector<gr_object> objects;
    Target target;
    gr_serve_result result{};
    std::string error;  // message of a failed request (the worker's, for the waiter)
    bool done = false;
};

// One batched launch in flight.
struct Batch {
    std::vector<Request*> reqs;
    hipEvent_t rendered = nullptr, copied = nullptr;
    unsigned char* h_stats = nullptr;  // pinned: stats of every frame
    unsigned char* h_ctrl = nullptr;   // pinned: control words after the batch
    int64_t start_ns = 0;
};

}  // namespace

struct gr_server {
    gr_context* ctx = nullptr;
    gr_server_config cfg{};
    hipStream_t copy_stream = nullptr;
    std::thread worker;

    std::mutex m;
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    int64_t next_ticket = 1;
    std::deque<Request*> queue;                                  // submitted, not admitted
    std::unordered_map<int64_t, std::unique_ptr<Request>> all;   // until released
    gr_status fatal = GR_OK;                                     // device error: server stopped

    // Under m: device target bytes admitted, and the target pools by size
    // (device halves return when a frame is in host memory, pinned halves
    // when its result is released).
    size_t used = 0;
    std::multimap<size_t, void*> free_dev, free_host;
    // Worker only.
    std::deque<Batch> inflight;
    std::vector<Batch> spare;  // batches with events and pinned blocks to reuse

    ~gr_server();
    void run();
    gr_status take_target(size_t bytes, size_t host_bytes, Target* t);
    gr_status launch(std::vector<Request*>& reqs);
    gr_status reap(bool wait);
    void finish(Request* r, gr_status st, const char* msg = nullptr);
};

namespace {

size_t target_bytes(const gr_frame& f) { return (size_t)f.width * f.height * 8; }

}  // namespace

// Caller holds m.
gr_status gr_server::take_target(size_t bytes, size_t host_bytes, Target* t) {
    t->bytes = bytes;
    t->host_bytes = host_bytes;
    auto d = free_dev.find(bytes);
    if (d != free_dev.end()) {
        t->dev = d->second;
        free_dev.erase(d);
    } else if (hipMalloc(&t->dev, bytes) != hipSuccess) {
        t->dev = nullptr;
        return fail(GR_ERR_OUT_OF_MEMORY, "device target");
    }
    auto h = free_host.find(host_bytes);
    if (h != free_host.end()) {
        t->host = h->second;
        free_host.erase(h);
    } else if (hipHostMalloc(&t->host, host_bytes, hipHostMallocDefault) != hipSuccess) {
        free_dev.emplace(bytes, t->dev);
        t->dev = t->host = nullptr;
        return fail(GR_ERR_OUT_OF_MEMORY, "pinned host target");
    }
    return GR_OK;
}

// Caller holds m. A failure keeps `msg`, else the message of the fail() on
// this (the worker's) thread that produced `st`.
void gr_server::finish(Request* r, gr_status st, const char* msg) {
    if (st) r->error = msg ? msg : gr_last_error();
    r->result.status = (int32_t)st;
    r->result.done_ns = now_ns();
    r->done = true;
}

gr_status gr_server::launch(std::vector<Request*>& reqs) {
    Batch b;
    if (!spare.empty()) {
        b = spare.back();
        spare.pop_back();
        b.reqs.clear();
    } else {
        if (hipEventCreateWithFlags(&b.rendered, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&b.copied, hipEventDisableTiming) != hipSuccess ||
            hipHostMalloc((void**)&b.h_stats,
                          gr_internal::stats_bytes_per_frame() * (size_t)cfg.max_batch,
                          hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void**)&b.h_ctrl, gr_internal::control_bytes(), hipHostMallocDefault) !=
                hipSuccess) {
            // The parts that were created go to the spare list (freed with the server).
            spare.push_back(b);
            return fail(GR_ERR_DEVICE, "server batch resources");
        }
    }
    std::vector<gr_batch_frame> frames(reqs.size());
    std::vector<gr_object> objs;
    for (size_t k = 0; k < reqs.size(); k++) {
        Request* r = reqs[k];
        gr_batch_frame& f = frames[k];
        f.frame = r->frame;
        f.first_object = (int32_t)objs.size();
        f.num_objects = (int32_t)r->objects.size();
        f.d_rgba = r->target.dev;
        f.d_depth = (char*)r->target.dev + (size_t)r->frame.width * r->frame.height * 4;
        objs.insert(objs.end(), r->objects.begin(), r->objects.end());
    }
    b.reqs = reqs;
    b.start_ns = now_ns();
    gr_status st = gr_internal::enqueue_batch(ctx, frames.data(), (int32_t)frames.size(),
                                              objs.data(), (int32_t)objs.size());
    if (st) {
        spare.push_back(b);
        return st;
    }
    hipStream_t s = gr_internal::context_stream(ctx);
    // Stats and control words of this batch, before the next batch reuses them.
    if (hipMemcpyAsync(b.h_stats, gr_internal::last_stats_device(ctx),
                       gr_internal::stats_bytes_per_frame() * reqs.size(), hipMemcpyDeviceToHost,
                       s) != hipSuccess ||
        hipMemcpyAsync(b.h_ctrl, gr_internal::control_device(ctx), gr_internal::control_bytes(),
                       hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemsetAsync(gr_internal::control_device(ctx), 0, gr_internal::control_bytes(), s) !=
            hipSuccess ||
        hipEventRecord(b.rendered, s) != hipSuccess ||
        hipStreamWaitEvent(copy_stream, b.rendered, 0) != hipSuccess) {
        spare.push_back(b);
        return fail(GR_ERR_DEVICE, "server batch readback");
    }
    // Frames to pinned host memory on the copy stream, overlapping the next
    // batch: colour, then depth unless the server reads back colour only. A
    // copy kernel writes the pinned pages over PCIe: while the next batch
    // renders, DMA copies of the same frames ran at about half the link's
    // rate, the kernel near its full rate.
    for (Request* r : reqs)
        gr::launch_copy(r->target.host, r->target.dev, r->target.host_bytes, copy_stream);
    if (hipGetLastError() != hipSuccess) {
        spare.push_back(b);
        return fail(GR_ERR_DEVICE, "server frame readback");
    }
    if (hipEventRecord(b.copied, copy_stream) != hipSuccess) {
        spare.push_back(b);
        return fail(GR_ERR_DEVICE, "server batch event");
    }
    for (Request* r : reqs) {
        r->result.start_ns = b.start_ns;
        r->result.batch_size = (int32_t)reqs.size();
    }
    inflight.push_back(b);
    return GR_OK;
}

// Completes the oldest batch if its frames are in host memory (waits for it
// when `wait`). A batch that overflowed its tile lists is queued again first
// (after growing them), whole.
gr_status gr_server::reap(bool wait) {
    if (inflight.empty()) return GR_OK;
    Batch& b = inflight.front();
    if (wait) {
        if (hipEventSynchronize(b.copied) != hipSuccess) return fail(GR_ERR_DEVICE, "server wait");
    } else {
        const hipError_t q = hipEventQuery(b.copied);
        if (q == hipErrorNotReady) return GR_OK;
        if (q != hipSuccess) return fail(GR_ERR_DEVICE, "server event query");
    }
    bool overflow = false, unsupported = false;
    unsigned long long needed = 0;
    gr_internal::decode_control(b.h_ctrl, &overflow, &needed, &unsupported);
    std::lock_guard<std::mutex> lk(m);
    if (overflow) {
        gr_internal::grow_refs_to(ctx, needed);
        for (auto it = b.reqs.rbegin(); it != b.reqs.rend(); ++it) queue.push_front(*it);
    } else {
        for (size_t k = 0; k < b.reqs.size(); k++) {
            Request* r = b.reqs[k];
            const unsigned char* fs = b.h_stats + k * gr_internal::stats_bytes_per_frame();
            gr_internal::decode_stats(fs, &r->result.stats);
            used -= r->target.bytes;
            r->result.rgba = (const uint8_t*)r->target.host;
            r->result.depth =
                cfg.colour_only ? nullptr
                                : (const float*)((const char*)r->target.host +
                                                 (size_t)r->frame.width * r->frame.height * 4);
            // The device half returns to the pool; the host half stays with the
            // result until it is released.
            free_dev.emplace(r->target.bytes, r->target.dev);
            r->target.dev = nullptr;
            // Only the frames with an overlay line too long fail; the others
            // of the batch are complete.
            if (gr_internal::frame_long_lines(fs))
                finish(r, GR_ERR_UNSUPPORTED, "overlay line longer than 65536 pixels");
            else
                finish(r, GR_OK);
        }
        cv_done.notify_all();
    }
    spare.push_back(b);
    inflight.pop_front();
    return GR_OK;
}

void gr_server::run() {
    (void)hipSetDevice(gr_internal::context_device(ctx));
    std::vector<Request*> batch;
    for (;;) {
        gr_status st = reap(false);
        batch.clear();
        {
            std::unique_lock<std::mutex> lk(m);
            if (st) {
                fatal = st;
                break;
            }
            if (stop && queue.empty() && inflight.empty()) break;
            if (queue.empty() && inflight.empty()) {
                cv_work.wait(lk, [&] { return stop || !queue.empty(); });
                continue;
            }
            // Admission: in arrival order, while the target budget and the
            // batch and in-flight limits allow (a request that does not fit
            // waits for earlier ones to complete).
            const int64_t now = now_ns();
            while (!queue.empty() && (int)batch.size() < cfg.max_batch &&
                   (int)inflight.size() < cfg.max_inflight) {
                Request* r = queue.front();
                // Shed a request that waited past the deadline before it
                // could be admitted (a redraw after an overflow is kept).
                if (cfg.deadline_ms > 0 && !r->target.dev &&
                    now - r->result.submit_ns > (int64_t)cfg.deadline_ms * 1000000) {
                    queue.pop_front();
                    finish(r, GR_ERR_BUSY, "shed: waited past the deadline before admission");
                    cv_done.notify_all();
                    continue;
                }
                const size_t need = r->target.dev ? 0 : target_bytes(r->frame);
                if (need && used + need > (size_t)cfg.target_budget) break;
                queue.pop_front();
                if (need) {
                    Target t;
                    gr_status ts = take_target(need, cfg.colour_only ? need / 2 : need, &t);
                    if (ts) {
                        finish(r, ts);
                        cv_done.notify_all();
                        continue;
                    }
                    r->target = t;
                    used += need;
                }
                batch.push_back(r);
            }
        }
        if (!batch.empty()) {
            st = launch(batch);
            if (st) {
                std::lock_guard<std::mutex> lk(m);
                for (Request* r : batch) finish(r, st);
                fatal = st;
                cv_done.notify_all();
                break;
            }
            continue;
        }
        // Nothing to admit: wait for the oldest batch when the device is full
        // or the budget is spent, else poll briefly for new arrivals.
        bool full;
        {
            std::lock_guard<std::mutex> lk(m);
            full = (int)inflight.size() >= cfg.max_inflight || !queue.empty();
        }
        if (full) {
            st = reap(true);
            if (st) {
                std::lock_guard<std::mutex> lk(m);
                fatal = st;
                break;
            }
        } else {
            std::unique_lock<std::mutex> lk(m);
            cv_work.wait_for(lk, std::chrono::microseconds(20),
                             [&] { return stop || !queue.empty(); });
        }
    }
    // Stopped: fail whatever is left and drain the device.
    (void)hipStreamSynchronize(gr_internal::context_stream(ctx));
    (void)hipStreamSynchronize(copy_stream);
    std::lock_guard<std::mutex> lk(m);
    const gr_status st = fatal ? fatal : GR_ERR_INVALID_ARGUMENT;
    const std::string why = fatal ? std::string("server stopped after a device error: ") +
                                        gr_last_error()
                                  : std::string("server destroyed before the request ran");
    for (Request* r : queue) finish(r, st, why.c_str());
    queue.clear();
    for (Batch& b : inflight)
        for (Request* r : b.reqs) finish(r, st, why.c_str());
    cv_done.notify_all();
}

gr_server::~gr_server() {
    for (Batch& b : inflight) spare.push_back(b);
    for (Batch& b : spare) {
        if (b.rendered) (void)hipEventDestroy(b.rendered);
        if (b.copied) (void)hipEventDestroy(b.copied);
        if (b.h_stats) (void)hipHostFree(b.h_stats);
        if (b.h_ctrl) (void)hipHostFree(b.h_ctrl);
    }
    for (auto& kv : free_dev) (void)hipFree(kv.second);
    for (auto& kv : free_host) (void)hipHostFree(kv.second);
    for (auto& kv : all) {
        if (kv.second->target.dev) (void)hipFree(kv.second->target.dev);
        if (kv.second->target.host) (void)hipHostFree(kv.second->target.host);
    }
    if (copy_stream) (void)hipStreamDestroy(copy_stream);
}

extern "C" {

gr_status gr_server_create(gr_context* ctx, const gr_server_config* cfg, gr_server** out) {
    if (!ctx || !cfg || !out) return fail(GR_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (cfg->target_budget < 8 || cfg->max_batch < 1 || cfg->max_inflight < 1)
        return fail(GR_ERR_INVALID_ARGUMENT, "server budget, max_batch and max_inflight must be positive");
    if ((cfg->colour_only != 0 && cfg->colour_only != 1) || cfg->reserved != 0)
        return fail(GR_ERR_INVALID_ARGUMENT, "colour_only must be 0 or 1 and reserved 0");
    if (cfg->max_queue < 0 || cfg->deadline_ms < 0)
        return fail(GR_ERR_INVALID_ARGUMENT, "max_queue and deadline_ms must not be negative");
    try {
        std::unique_ptr<gr_server> s(new gr_server);
        s->ctx = ctx;
        s->cfg = *cfg;
        if (hipSetDevice(gr_internal::context_device(ctx)) != hipSuccess ||
            hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking) != hipSuccess)
            return fail(GR_ERR_DEVICE, "server copy stream");
        gr_server* raw = s.get();
        // Consecutive batches use the context's pipelines, so a batch's
        // geometry overlaps the previous one's raster; each batch's stats and
        // control words are read on its own pipe's stream.
        s->worker = std::thread([raw] { raw->run(); });
        *out = s.release();
        return GR_OK;
    } catch (...) {
        return fail(GR_ERR_OUT_OF_MEMORY, "server allocation failed");
    }
}

void gr_server_destroy(gr_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->m);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    delete s;
}

gr_status gr_server_submit(gr_server* s, const gr_frame* frame, const gr_object* objects,
                           int32_t num_objects, int64_t* ticket) {
    if (!s || !frame || !ticket || num_objects < 0 || (num_objects > 0 && !objects))
        return fail(GR_ERR_INVALID_ARGUMENT, "null argument");
    // The checks a batch would make of this frame, so that a bad request is
    // refused alone instead of failing the batch it would share.
    const gr_status cst = gr_internal::check_request(s->ctx, frame, objects, num_objects);
    if (cst) return cst;
    if (target_bytes(*frame) > (size_t)s->cfg.target_budget)
        return fail(GR_ERR_INVALID_ARGUMENT, "frame larger than the server's target budget");
    try {
        std::unique_ptr<Request> r(new Request);
        r->frame = *frame;
        r->objects.assign(objects, objects + num_objects);
        r->result.width = frame->width;
        r->result.height = frame->height;
        std::lock_guard<std::mutex> lk(s->m);
        if (s->stop || s->fatal) return fail(s->fatal ? s->fatal : GR_ERR_INVALID_ARGUMENT,
                                             "server stopped");
        if (s->cfg.max_queue > 0 && (int64_t)s->queue.size() >= s->cfg.max_queue)
            return fail(GR_ERR_BUSY, "server queue full");
        r->ticket = s->next_ticket++;
        r->result.submit_ns = now_ns();
        *ticket = r->ticket;
        s->queue.push_back(r.get());
        s->all.emplace(r->ticket, std::move(r));
    } catch (...) {
        return fail(GR_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
    s->cv_work.notify_one();
    return GR_OK;
}

gr_status gr_server_wait(gr_server* s, int64_t ticket, int32_t timeout_ms, gr_serve_result* out) {
    if (!s || !out) return fail(GR_ERR_INVALID_ARGUMENT, "null argument");
    std::unique_lock<std::mutex> lk(s->m);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return fail(GR_ERR_INVALID_HANDLE, "unknown ticket");
    Request* r = it->second.get();
    auto ready = [&] { return r->done; };
    if (timeout_ms < 0)
        s->cv_done.wait(lk, ready);
    else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready))
        return fail(GR_ERR_TIMEOUT, "request not finished");
    *out = r->result;
    if (r->result.status) return fail((gr_status)r->result.status, r->error);
    return GR_OK;
}

gr_status gr_server_release(gr_server* s, int64_t ticket) {
    if (!s) return fail(GR_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lk(s->m);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return fail(GR_ERR_INVALID_HANDLE, "unknown ticket");
    Request* r = it->second.get();
    if (!r->done) return fail(GR_ERR_INVALID_ARGUMENT, "request still in progress");
    // The pinned half goes back to the pool (the worker reuses it by size); a
    // request that failed still holding its device half returns that too.
    if (r->target.host) s->free_host.emplace(r->target.host_bytes, r->target.host);
    r->target.host = nullptr;
    if (r->target.dev) {
        s->free_dev.emplace(r->target.bytes, r->target.dev);
        s->used -= r->target.bytes;
        r->target.dev = nullptr;
    }
    s->all.erase(it);
    return GR_OK;
}

gr_status gr_server_replay(gr_server* s, const gr_frame* frames, const gr_object* objects,
                           const int32_t* first_object, const int32_t* num_objects,
                           int32_t num_distinct, const double* arrival_s, int32_t n,
                           int64_t* latency_ns, int32_t* batch_size) {
    if (!s || !frames || !first_object || !num_objects || !arrival_s || num_distinct < 1 || n < 0)
        return fail(GR_ERR_INVALID_ARGUMENT, "null argument");
    for (int32_t d = 0; d < num_distinct; d++)
        if (num_objects[d] < 0 || (num_objects[d] > 0 && (!objects || first_object[d] < 0)))
            return fail(GR_ERR_INVALID_ARGUMENT, "request object range");
    // Collector: waits for the tickets in submission order and releases them,
    // so pinned buffers are recycled while requests keep arriving.
    std::vector<int64_t> tickets((size_t)n, 0);
    std::mutex tm;
    std::condition_variable tcv;
    int32_t submitted = 0;
    bool aborted = false;
    gr_status first_error = GR_OK;
    std::thread collector([&] {
        for (int32_t i = 0; i < n; i++) {
            int64_t t;
            {
                std::unique_lock<std::mutex> lk(tm);
                tcv.wait(lk, [&] { return submitted > i || aborted; });
                if (submitted <= i) return;
                t = tickets[(size_t)i];
            }
            if (t == 0) {  // refused at submission (queue full)
                if (latency_ns) latency_ns[i] = -1;
                if (batch_size) batch_size[i] = 0;
                continue;
            }
            gr_serve_result r;
            const gr_status st = gr_server_wait(s, t, -1, &r);
            if (latency_ns) latency_ns[i] = st == GR_ERR_BUSY ? -2 : r.done_ns - r.submit_ns;
            if (batch_size) batch_size[i] = r.batch_size;
            (void)gr_server_release(s, t);
            if (st && st != GR_ERR_BUSY) {
                std::lock_guard<std::mutex> lk(tm);
                if (!first_error) first_error = st;
            }
        }
    });
    const auto t0 = std::chrono::steady_clock::now();
    gr_status st = GR_OK;
    for (int32_t i = 0; i < n; i++) {
        const auto due = t0 + std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                                  std::chrono::duration<double>(arrival_s[i]));
        while (std::chrono::steady_clock::now() < due) {
        }
        const int32_t d = i % num_distinct;
        int64_t t;
        st = gr_server_submit(s, &frames[d], objects ? objects + first_object[d] : nullptr,
                              num_objects[d], &t);
        if (st == GR_ERR_BUSY) {  // a full queue refuses it: counted, not fatal
            t = 0;
            st = GR_OK;
        }
        std::lock_guard<std::mutex> lk(tm);
        if (st) {
            aborted = true;
            break;
        }
        tickets[(size_t)i] = t;
        submitted = i + 1;
        tcv.notify_one();
    }
    {
        std::lock_guard<std::mutex> lk(tm);
        aborted = aborted || submitted < n;
    }
    tcv.notify_one();
    collector.join();
    return st ? st : first_error;
}

}  // extern "C"
