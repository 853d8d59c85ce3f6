// serve.cpp — the serving loop behind rrxCreateServer (radeonrays_ext.h).
//
// The reference is a synchronous library with no request loop; this is the
// layer a ray-query service puts in front of rrCmdIntersect.  One host thread
// per server:
//   release   requests become visible at their release time (an open-loop
//             arrival schedule can be submitted ahead of time);
//   admit     released requests are admitted in arrival order while their
//             ray + hit bytes fit the memory budget (head-of-line order, so
//             a large request is not starved by small ones behind it);
//   dispatch  batched mode: whenever one of max_inflight streams is idle, all
//             admitted requests (up to the batch limits) go out as one
//             rrxCmdIntersectBatch, so requests that arrived while the GPU
//             was busy share the next launch and launches on different
//             streams overlap; plain mode: one rrCmdIntersect +
//             rrSumbitCommandStream + rrWaitEvent per request, in order;
//   complete  a stream's event is polled; its requests are stamped complete
//             and their bytes leave the budget.
// Requests whose rays and hits are in host memory (rrxServerSubmitHost) are
// staged through a device arena owned by the server: a request is admitted
// only when the arena has room for its rays and hits, its rays are copied in
// on the launch's stream before the launch and its hits copied out after
// it, and the room is freed at completion, so the arena bounds the device
// memory such requests hold.
// The thread talks to the GPU only through the library's public entry points
// and HIP streams/events of default priority.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <queue>
#include <thread>
#include <unordered_set>
#include <vector>

#include "../../include/radeonrays_ext.h"
#include "../../include/radeonrays_hip.h"
#include "api_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

// The thread sleeps until this long before the next release and polls from
// there (a timed sleep wakes tens of microseconds late).
constexpr uint64_t kSpinNs = 1000000;

uint32_t HitBytes(RRIntersectQueryOutput o) {
    switch (o) {
        case RR_INTERSECT_QUERY_OUTPUT_INSTANCE_ID: return 4;
        case RR_INTERSECT_QUERY_OUTPUT_FULL_HIT_T: return 32;
        default: return 16;
    }
}

// First-fit allocator over the staging arena's bytes, in 256-byte granules;
// freed ranges merge with free neighbours.
class ArenaSpace {
public:
    static constexpr uint64_t kGranule = 256;
    void Reset(uint64_t bytes) {
        free_.clear();
        if (bytes >= kGranule) free_[0] = bytes / kGranule * kGranule;
    }
    bool Alloc(uint64_t bytes, uint64_t& off) {
        bytes = (bytes + kGranule - 1) / kGranule * kGranule;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            off = it->first;
            const uint64_t rest = it->second - bytes;
            free_.erase(it);
            if (rest) free_[off + bytes] = rest;
            return true;
        }
        return false;
    }
    void Free(uint64_t off, uint64_t bytes) {
        bytes = (bytes + kGranule - 1) / kGranule * kGranule;
        auto next = free_.lower_bound(off);
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                off = prev->first;
                bytes += prev->second;
                free_.erase(prev);
            }
        }
        next = free_.lower_bound(off);
        if (next != free_.end() && off + bytes == next->first) {
            bytes += next->second;
            free_.erase(next);
        }
        free_[off] = bytes;
    }

private:
    std::map<uint64_t, uint64_t> free_;  // offset -> length
};

uint64_t StagedBytes(uint32_t rays, uint32_t hit_bytes) {
    const uint64_t g = ArenaSpace::kGranule;
    return ((uint64_t)rays * 32 + g - 1) / g * g + ((uint64_t)rays * hit_bytes + g - 1) / g * g;
}

// Host buffers of a staged request and its place in the arena.
struct HostBuffers {
    const void* rays = nullptr;
    void* hits = nullptr;
    uint64_t arena_off = 0;
    bool staged = false;  // holds arena space
};

struct Pending {
    uint64_t release;
    uint64_t ticket;
};
struct LaterFirst {
    bool operator()(const Pending& a, const Pending& b) const {
        return a.release != b.release ? a.release > b.release : a.ticket > b.ticket;
    }
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    RRCommandStream cs = nullptr;
    RRDevicePtr scratch = nullptr;
    std::vector<uint64_t> tickets;
    std::vector<RRDevicePtr> handles;  // arena handles of the launch's staged requests
    bool busy = false;
};

struct Server {
    RRContext ctx = nullptr;
    RRxServerOptions opt{};
    int device = 0;
    Clock::time_point epoch;

    std::mutex mu;
    std::condition_variable wake;     // submissions and shutdown
    std::condition_variable drained;  // completions
    std::vector<RRxIntersectRequest> reqs;  // by ticket
    std::vector<HostBuffers> host;          // by ticket (rays == nullptr: device request)
    std::vector<uint64_t> bytes;
    std::vector<RRxServedTiming> timing;
    std::vector<uint8_t> complete;
    std::priority_queue<Pending, std::vector<Pending>, LaterFirst> unreleased;
    std::deque<uint64_t> released;  // waiting for budget
    std::deque<uint64_t> admitted;  // within budget, not dispatched
    uint64_t admitted_bytes = 0, peak_bytes = 0;
    uint64_t submitted = 0, completed = 0;
    uint32_t batches = 0;
    bool stop = false;
    RRError error = RR_SUCCESS;

    std::vector<Slot> slots;
    RRCommandStream plain_cs = nullptr;
    RRDevicePtr plain_scratch = nullptr;
    RRDevicePtr arena = nullptr;  // staging arena for host requests
    uint8_t* arena_base = nullptr;
    ArenaSpace arena_space;       // guarded by mu
    std::thread worker;

    uint64_t Now() const {
        return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(Clock::now() - epoch).count();
    }

    // Move released requests into the budget (caller holds mu).
    void ReleaseAndAdmit(uint64_t now) {
        while (!unreleased.empty() && unreleased.top().release <= now) {
            released.push_back(unreleased.top().ticket);
            unreleased.pop();
        }
        while (!released.empty()) {
            const uint64_t t = released.front();
            if (admitted_bytes + bytes[t] > opt.memory_budget) break;
            HostBuffers& h = host[t];
            if (h.rays && reqs[t].ray_count) {
                // Head of line: wait for arena space rather than pass over it.
                if (!arena_space.Alloc(StagedBytes(reqs[t].ray_count, HitBytes(reqs[t].query_output)), h.arena_off))
                    break;
                h.staged = true;
            }
            admitted_bytes += bytes[t];
            peak_bytes = std::max(peak_bytes, admitted_bytes);
            admitted.push_back(t);
            released.pop_front();
        }
    }

    // Take the next batch off the admitted queue (caller holds mu).
    // (The per-ticket vectors grow under mu as requests are submitted, so
    // the batch takes copies of what the dispatch needs.)
    void TakeBatch(uint32_t max_requests, std::vector<uint64_t>& tickets, std::vector<RRxIntersectRequest>& batch,
                   std::vector<HostBuffers>& staged) {
        tickets.clear();
        batch.clear();
        staged.clear();
        uint64_t rays = 0;
        while (!admitted.empty() && tickets.size() < max_requests) {
            const uint64_t t = admitted.front();
            if (!tickets.empty() && rays + reqs[t].ray_count > opt.max_batch_rays) break;
            rays += reqs[t].ray_count;
            tickets.push_back(t);
            batch.push_back(reqs[t]);
            staged.push_back(host[t]);
            admitted.pop_front();
        }
    }

    void Finish(const std::vector<uint64_t>& tickets, uint64_t now) {
        std::lock_guard<std::mutex> lk(mu);
        for (uint64_t t : tickets) {
            timing[t].complete_ns = now;
            complete[t] = 1;
            admitted_bytes -= bytes[t];
            HostBuffers& h = host[t];
            if (h.staged) {
                arena_space.Free(h.arena_off, StagedBytes(reqs[t].ray_count, HitBytes(reqs[t].query_output)));
                h.staged = false;
            }
            ++completed;
        }
        drained.notify_all();
    }

    // Point the staged requests of a batch at their arena space (new
    // RRDevicePtr handles, appended to `handles`) and enqueue their ray
    // uploads on `stream` (synchronously when stream is null).  The caller
    // does not hold mu: `staged` holds copies taken with the batch.
    RRError StageIn(const std::vector<HostBuffers>& staged, std::vector<RRxIntersectRequest>& batch,
                    std::vector<RRDevicePtr>& handles, hipStream_t stream) {
        for (size_t k = 0; k < staged.size(); ++k) {
            const HostBuffers& h = staged[k];
            if (!h.staged) continue;
            const uint64_t ray_bytes = (uint64_t)batch[k].ray_count * 32;
            const uint64_t hits_off = h.arena_off + (ray_bytes + ArenaSpace::kGranule - 1) / ArenaSpace::kGranule *
                                                        ArenaSpace::kGranule;
            RRDevicePtr r = nullptr, o = nullptr;
            RRError e = rrGetDevicePtrFromHipPtr(ctx, arena_base, h.arena_off, &r);
            if (e == RR_SUCCESS) handles.push_back(r);
            if (e == RR_SUCCESS) e = rrGetDevicePtrFromHipPtr(ctx, arena_base, hits_off, &o);
            if (e == RR_SUCCESS) handles.push_back(o);
            if (e != RR_SUCCESS) return e;
            batch[k].rays = r;
            batch[k].hits = o;
            const hipError_t c = stream ? hipMemcpyAsync(arena_base + h.arena_off, h.rays, ray_bytes,
                                                         hipMemcpyHostToDevice, stream)
                                        : hipMemcpy(arena_base + h.arena_off, h.rays, ray_bytes, hipMemcpyHostToDevice);
            if (c != hipSuccess) return RR_ERROR_INTERNAL;
        }
        return RR_SUCCESS;
    }

    // Enqueue the staged requests' hit downloads (after the launch).
    RRError StageOut(const std::vector<HostBuffers>& staged, const std::vector<RRxIntersectRequest>& batch,
                     hipStream_t stream) {
        for (size_t k = 0; k < staged.size(); ++k) {
            const HostBuffers& h = staged[k];
            if (!h.staged) continue;
            const uint64_t ray_bytes = (uint64_t)batch[k].ray_count * 32;
            const uint64_t hits_off = h.arena_off + (ray_bytes + ArenaSpace::kGranule - 1) / ArenaSpace::kGranule *
                                                        ArenaSpace::kGranule;
            const uint64_t hit_bytes = (uint64_t)batch[k].ray_count * HitBytes(batch[k].query_output);
            const hipError_t c = stream ? hipMemcpyAsync(h.hits, arena_base + hits_off, hit_bytes,
                                                         hipMemcpyDeviceToHost, stream)
                                        : hipMemcpy(h.hits, arena_base + hits_off, hit_bytes, hipMemcpyDeviceToHost);
            if (c != hipSuccess) return RR_ERROR_INTERNAL;
        }
        return RR_SUCCESS;
    }

    void ReleaseHandles(std::vector<RRDevicePtr>& handles) {
        for (RRDevicePtr p : handles) (void)rrReleaseDevicePtr(ctx, p);
        handles.clear();
    }

    void Fail(RRError e) {
        std::lock_guard<std::mutex> lk(mu);
        if (error == RR_SUCCESS) error = e;
        stop = true;
        drained.notify_all();
    }

    void Stamp(const std::vector<uint64_t>& tickets, uint64_t now, uint32_t id) {
        std::lock_guard<std::mutex> lk(mu);
        for (uint64_t t : tickets) {
            timing[t].dispatch_ns = now;
            timing[t].batch_id = id;
            timing[t].batch_size = (uint32_t)tickets.size();
        }
    }

    void RunPlain() {
        std::vector<uint64_t> tickets;
        std::vector<RRxIntersectRequest> batch;
        std::vector<HostBuffers> staged;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu);
                for (;;) {
                    if (stop) return;
                    ReleaseAndAdmit(Now());
                    if (!admitted.empty()) break;
                    if (!unreleased.empty()) {
                        const uint64_t next = unreleased.top().release, now = Now();
                        if (next > now + kSpinNs) {
                            wake.wait_for(lk, std::chrono::nanoseconds(next - now - kSpinNs));
                        } else {
                            lk.unlock();
                            std::this_thread::yield();
                            lk.lock();
                        }
                    } else {
                        wake.wait(lk);
                    }
                }
                TakeBatch(1, tickets, batch, staged);
            }
            Stamp(tickets, Now(), batches++);
            if (batch[0].ray_count == 0) {
                Finish(tickets, Now());
                continue;
            }
            // A staged request: upload, trace, download, each waited for, as
            // a client of the per-call API does.
            std::vector<RRDevicePtr> handles;
            RRError e = StageIn(staged, batch, handles, nullptr);
            const RRxIntersectRequest& q = batch[0];
            RREvent ev = nullptr;
            if (e == RR_SUCCESS)
                e = rrCmdIntersect(ctx, q.scene_buffer, q.query, q.rays, q.ray_count, nullptr, q.query_output, q.hits,
                                   plain_scratch, plain_cs);
            if (e == RR_SUCCESS) e = rrSumbitCommandStream(ctx, plain_cs, nullptr, &ev);
            if (e == RR_SUCCESS) e = rrWaitEvent(ctx, ev);
            if (ev) (void)rrReleaseEvent(ctx, ev);
            if (e == RR_SUCCESS) e = StageOut(staged, batch, nullptr);
            ReleaseHandles(handles);
            if (e != RR_SUCCESS) return Fail(e);
            Finish(tickets, Now());
        }
    }

    void RunBatched() {
        std::vector<uint64_t> tickets;
        std::vector<RRxIntersectRequest> batch;
        std::vector<HostBuffers> staged;
        for (;;) {
            // Completions.
            for (Slot& s : slots) {
                if (!s.busy) continue;
                const hipError_t q = hipEventQuery(s.done);
                if (q == hipErrorNotReady) continue;
                if (q != hipSuccess) return Fail(RR_ERROR_INTERNAL);
                s.busy = false;
                ReleaseHandles(s.handles);
                Finish(s.tickets, Now());
            }
            Slot* free_slot = nullptr;
            bool any_busy = false;
            for (Slot& s : slots) {
                if (s.busy) any_busy = true;
                else if (!free_slot) free_slot = &s;
            }
            {
                std::unique_lock<std::mutex> lk(mu);
                if (stop) break;
                ReleaseAndAdmit(Now());
                if (free_slot && !admitted.empty()) {
                    TakeBatch(opt.max_batch_requests, tickets, batch, staged);
                } else {
                    tickets.clear();
                    if (!any_busy) {
                        // Nothing on the GPU: sleep until shortly before the
                        // next release (then poll), or until a submission.
                        if (!unreleased.empty()) {
                            const uint64_t next = unreleased.top().release, now = Now();
                            if (next > now + kSpinNs) wake.wait_for(lk, std::chrono::nanoseconds(next - now - kSpinNs));
                        } else if (released.empty()) {
                            wake.wait(lk);
                        }
                    }
                }
            }
            if (tickets.empty()) {
                std::this_thread::yield();
                continue;
            }
            Slot& s = *free_slot;
            Stamp(tickets, Now(), batches++);
            RRError e = StageIn(staged, batch, s.handles, s.stream);
            if (e == RR_SUCCESS) e = rrxCmdIntersectBatch(ctx, batch.data(), (uint32_t)batch.size(), s.scratch, s.cs);
            if (e == RR_SUCCESS) e = StageOut(staged, batch, s.stream);
            if (e == RR_SUCCESS && hipEventRecord(s.done, s.stream) != hipSuccess) e = RR_ERROR_INTERNAL;
            if (e != RR_SUCCESS) {
                (void)hipStreamSynchronize(s.stream);
                ReleaseHandles(s.handles);
                return Fail(e);
            }
            s.tickets.swap(tickets);
            s.busy = true;
        }
        // Shutdown: wait for what is still on the GPU.
        for (Slot& s : slots)
            if (s.busy) {
                (void)hipEventSynchronize(s.done);
                ReleaseHandles(s.handles);
            }
    }

    void Run() {
        (void)hipSetDevice(device);
        if (opt.mode == RRX_SERVE_PLAIN) RunPlain();
        else RunBatched();
    }
};

Server* AsServer(RRxServer s) { return reinterpret_cast<Server*>(s); }

// Live servers: a handle that was never created or is already destroyed is
// rejected with RR_ERROR_INVALID_PARAMETER rather than dereferenced.
std::mutex g_servers_mu;
std::unordered_set<const void*> g_servers;
bool LiveServer(RRxServer s) {
    std::lock_guard<std::mutex> lk(g_servers_mu);
    return s && g_servers.count(s) != 0;
}

void DestroyResources(Server* sv) {
    (void)hipSetDevice(sv->device);
    for (Slot& s : sv->slots) {
        if (s.stream) (void)hipStreamSynchronize(s.stream);
        if (s.cs) (void)rrReleaseExternalCommandStream(sv->ctx, s.cs);
        if (s.scratch) (void)rrReleaseDevicePtr(sv->ctx, s.scratch);
        if (s.done) (void)hipEventDestroy(s.done);
        if (s.stream) (void)hipStreamDestroy(s.stream);
    }
    sv->slots.clear();
    if (sv->plain_cs) (void)rrReleaseCommandStream(sv->ctx, sv->plain_cs);
    if (sv->plain_scratch) (void)rrReleaseDevicePtr(sv->ctx, sv->plain_scratch);
    if (sv->arena) (void)rrReleaseDevicePtr(sv->ctx, sv->arena);
}

}  // namespace

extern "C" {

RRError rrxCreateServer(RRContext context, const RRxServerOptions* options, RRxServer* server) {
    if (!rr::IsLiveContext(context) || !options || !server) return RR_ERROR_INVALID_PARAMETER;
    const RRxServerOptions& o = *options;
    if (rr::Raw(o.mode) != RRX_SERVE_BATCHED && rr::Raw(o.mode) != RRX_SERVE_PLAIN) return RR_ERROR_INVALID_PARAMETER;
    if (o.max_batch_rays == 0 || o.memory_budget == 0) return RR_ERROR_INVALID_PARAMETER;
    if (o.mode == RRX_SERVE_BATCHED && (o.max_inflight == 0 || o.max_inflight > 16 || o.max_batch_requests == 0))
        return RR_ERROR_INVALID_PARAMETER;
    auto* sv = new Server;
    sv->ctx = context;
    sv->opt = o;
    sv->device = rr::ContextDevice(context);
    sv->epoch = Clock::now();
    RRError e = RR_SUCCESS;
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(sv->device);
    if (o.mode == RRX_SERVE_PLAIN) {
        size_t bytes = 0;
        e = rrGetTraceMemoryRequirements(context, o.max_batch_rays, &bytes);
        if (e == RR_SUCCESS) e = rrAllocateDeviceBuffer(context, bytes, &sv->plain_scratch);
        if (e == RR_SUCCESS) e = rrAllocateCommandStream(context, &sv->plain_cs);
    } else {
        size_t bytes = 0;
        e = rrxGetIntersectBatchMemoryRequirements(context, o.max_batch_requests, o.max_batch_rays, o.max_batch_rays,
                                                    &bytes);
        sv->slots.resize(o.max_inflight);
        for (Slot& s : sv->slots) {
            if (e != RR_SUCCESS) break;
            if (hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess) {
                e = RR_ERROR_INTERNAL;
                break;
            }
            e = rrGetCommandStreamFromHipStream(context, s.stream, &s.cs);
            if (e == RR_SUCCESS) e = rrAllocateDeviceBuffer(context, bytes, &s.scratch);
        }
    }
    if (e == RR_SUCCESS && o.staging_bytes) {
        e = rrAllocateDeviceBuffer(context, o.staging_bytes, &sv->arena);
        if (e == RR_SUCCESS) {
            sv->arena_base = static_cast<uint8_t*>(rrGetHipAddress(sv->arena));
            sv->arena_space.Reset(o.staging_bytes);
        }
    }
    (void)hipSetDevice(prev);
    if (e != RR_SUCCESS) {
        DestroyResources(sv);
        delete sv;
        return e;
    }
    sv->worker = std::thread([sv] { sv->Run(); });
    {
        std::lock_guard<std::mutex> lk(g_servers_mu);
        g_servers.insert(sv);
    }
    *server = reinterpret_cast<RRxServer>(sv);
    return RR_SUCCESS;
}

RRError rrxDestroyServer(RRxServer server) {
    {
        std::lock_guard<std::mutex> lk(g_servers_mu);
        if (!server || g_servers.erase(server) == 0) return RR_ERROR_INVALID_PARAMETER;
    }
    Server* sv = AsServer(server);
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
    }
    sv->wake.notify_all();
    sv->worker.join();
    DestroyResources(sv);
    delete sv;
    return RR_SUCCESS;
}

uint64_t rrxServerNow(RRxServer server) { return LiveServer(server) ? AsServer(server)->Now() : 0; }

RRError rrxServerSubmit(RRxServer server, const RRxIntersectRequest* requests, const uint64_t* release_ns, uint32_t count,
                        uint64_t* first_ticket) {
    if (!LiveServer(server) || !first_ticket || (count && !requests)) return RR_ERROR_INVALID_PARAMETER;
    Server* sv = AsServer(server);
    for (uint32_t i = 0; i < count; ++i) {
        const RRxIntersectRequest& q = requests[i];
        if (!q.scene_buffer || (q.ray_count && (!q.rays || !q.hits))) return RR_ERROR_INVALID_PARAMETER;
        if (rr::Raw(q.query) != RR_INTERSECT_QUERY_CLOSEST && rr::Raw(q.query) != RR_INTERSECT_QUERY_ANY) return RR_ERROR_INVALID_PARAMETER;
        if (rr::Raw(q.query_output) != RR_INTERSECT_QUERY_OUTPUT_FULL_HIT &&
            rr::Raw(q.query_output) != RR_INTERSECT_QUERY_OUTPUT_INSTANCE_ID &&
            rr::Raw(q.query_output) != RR_INTERSECT_QUERY_OUTPUT_FULL_HIT_T)
            return RR_ERROR_INVALID_PARAMETER;
        if (q.ray_count > sv->opt.max_batch_rays) return RR_ERROR_INVALID_PARAMETER;
        if (uint64_t(q.ray_count) * (32 + HitBytes(q.query_output)) > sv->opt.memory_budget)
            return RR_ERROR_INVALID_PARAMETER;
    }
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        if (sv->error != RR_SUCCESS) return sv->error;
        const uint64_t now = sv->Now();
        *first_ticket = sv->submitted;
        for (uint32_t i = 0; i < count; ++i) {
            const RRxIntersectRequest& q = requests[i];
            const uint64_t t = sv->submitted++;
            const uint64_t rel = release_ns ? std::max(release_ns[i], now) : now;
            sv->reqs.push_back(q);
            sv->host.push_back(HostBuffers{});
            sv->bytes.push_back(uint64_t(q.ray_count) * (32 + HitBytes(q.query_output)));
            RRxServedTiming tm{};
            tm.arrival_ns = rel;
            sv->timing.push_back(tm);
            sv->complete.push_back(0);
            sv->unreleased.push({rel, t});
        }
    }
    sv->wake.notify_all();
    return RR_SUCCESS;
}

RRError rrxServerSubmitHost(RRxServer server, const RRxHostIntersectRequest* requests, const uint64_t* release_ns,
                            uint32_t count, uint64_t* first_ticket) {
    if (!LiveServer(server) || !first_ticket || (count && !requests)) return RR_ERROR_INVALID_PARAMETER;
    Server* sv = AsServer(server);
    for (uint32_t i = 0; i < count; ++i) {
        const RRxHostIntersectRequest& q = requests[i];
        if (!q.scene_buffer || (q.ray_count && (!q.rays || !q.hits))) return RR_ERROR_INVALID_PARAMETER;
        if (rr::Raw(q.query) != RR_INTERSECT_QUERY_CLOSEST && rr::Raw(q.query) != RR_INTERSECT_QUERY_ANY) return RR_ERROR_INVALID_PARAMETER;
        if (rr::Raw(q.query_output) != RR_INTERSECT_QUERY_OUTPUT_FULL_HIT &&
            rr::Raw(q.query_output) != RR_INTERSECT_QUERY_OUTPUT_INSTANCE_ID &&
            rr::Raw(q.query_output) != RR_INTERSECT_QUERY_OUTPUT_FULL_HIT_T)
            return RR_ERROR_INVALID_PARAMETER;
        if (q.ray_count > sv->opt.max_batch_rays) return RR_ERROR_INVALID_PARAMETER;
        if (uint64_t(q.ray_count) * (32 + HitBytes(q.query_output)) > sv->opt.memory_budget)
            return RR_ERROR_INVALID_PARAMETER;
        if (q.ray_count && (!sv->arena || StagedBytes(q.ray_count, HitBytes(q.query_output)) > sv->opt.staging_bytes))
            return RR_ERROR_INVALID_PARAMETER;
    }
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        if (sv->error != RR_SUCCESS) return sv->error;
        const uint64_t now = sv->Now();
        *first_ticket = sv->submitted;
        for (uint32_t i = 0; i < count; ++i) {
            const RRxHostIntersectRequest& q = requests[i];
            const uint64_t t = sv->submitted++;
            const uint64_t rel = release_ns ? std::max(release_ns[i], now) : now;
            RRxIntersectRequest d{};
            d.scene_buffer = q.scene_buffer;
            d.ray_count = q.ray_count;
            d.query = q.query;
            d.query_output = q.query_output;
            sv->reqs.push_back(d);
            HostBuffers h;
            h.rays = q.ray_count ? q.rays : nullptr;
            h.hits = q.hits;
            sv->host.push_back(h);
            sv->bytes.push_back(uint64_t(q.ray_count) * (32 + HitBytes(q.query_output)));
            RRxServedTiming tm{};
            tm.arrival_ns = rel;
            sv->timing.push_back(tm);
            sv->complete.push_back(0);
            sv->unreleased.push({rel, t});
        }
    }
    sv->wake.notify_all();
    return RR_SUCCESS;
}

RRError rrxServerDrain(RRxServer server) {
    if (!LiveServer(server)) return RR_ERROR_INVALID_PARAMETER;
    Server* sv = AsServer(server);
    std::unique_lock<std::mutex> lk(sv->mu);
    sv->drained.wait(lk, [sv] { return sv->completed == sv->submitted || sv->error != RR_SUCCESS; });
    return sv->error;
}

RRError rrxServerGetTimings(RRxServer server, uint64_t first_ticket, uint32_t count, RRxServedTiming* timings) {
    if (!LiveServer(server) || (count && !timings)) return RR_ERROR_INVALID_PARAMETER;
    Server* sv = AsServer(server);
    std::lock_guard<std::mutex> lk(sv->mu);
    if (first_ticket + count > sv->submitted) return RR_ERROR_INVALID_PARAMETER;
    for (uint32_t i = 0; i < count; ++i) {
        const uint64_t t = first_ticket + i;
        if (!sv->complete[t]) return RR_ERROR_INVALID_PARAMETER;
        timings[i] = sv->timing[t];
    }
    return RR_SUCCESS;
}

RRError rrxServerGetPeakAdmitted(RRxServer server, uint64_t* bytes) {
    if (!LiveServer(server) || !bytes) return RR_ERROR_INVALID_PARAMETER;
    Server* sv = AsServer(server);
    std::lock_guard<std::mutex> lk(sv->mu);
    *bytes = sv->peak_bytes;
    return RR_SUCCESS;
}

}  // extern "C"
