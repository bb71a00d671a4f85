// Serving loop of a session: continuous batching of in-memory update batches.
//
// The reference takes its batches only from the update file and runs them one after another on
// the default stream (technique_permutation.h:127-215).  Here a request loop keeps the graph,
// index and plans resident and takes batches as they arrive.  Two host threads share the work:
// the stager validates a request, computes its keys and copies it, through a pinned buffer, into
// one of two device staging slots on a copy stream; the runner waits for that copy and applies
// the batch (admission, index update, matching) on the session's streams.  So one request's
// host-to-device transfer and preparation overlap the previous request's matching, and a request
// is admitted to the device as soon as a staging slot is free; the adjacency arena admits it
// (with the blocks earlier relocations freed) or refuses it without change.  Requests apply in
// submission order, so each result equals what csm_apply_edges returns for the same sequence.
#include "engine.h"

#include <cstring>

namespace csm {

#define CSM_HIP(call)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) return std::string("HIP error: ") + hipGetErrorString(e_) + " at " #call; \
    } while (0)

std::string Engine::stage_keys(const ServeReq& r, uint32_t* kf, uint32_t* kb) const {
    for (uint64_t k = 0; k < r.n; k++) {
        const uint32_t a = r.src[k], b = r.dst[k];
        if (a >= dv_ || b >= dv_) return tagged(Code::kArgument, "batch edge endpoint out of range");
        if (a == b) return tagged(Code::kArgument, "self-loops are not supported");
        if (host_g_.vlabel[a] == kNoVertex || host_g_.vlabel[b] == kNoVertex) return tagged(Code::kArgument, "batch edge endpoint was never declared");
        if (!knows_elabel(r.elabel[k])) return tagged(Code::kArgument, "batch edge label unknown to the session");
        kf[k] = key_of_edge(b, r.elabel[k]);
        kb[k] = key_of_edge(a, r.elabel[k]);
    }
    return "";
}

std::string Engine::serve_start() {
    if (!built_) return tagged(Code::kState, "index not built");
    if (srv_) return tagged(Code::kState, "the session is already serving");
    for (auto& qs : qs_) if (!qs->planned) return tagged(Code::kState, "every query must be planned before serving");
    auto sv = std::make_unique<Server>();
    CSM_HIP(hipStreamCreateWithFlags(&sv->copy, hipStreamNonBlocking));
    for (int k = 0; k < 2; k++) {
        CSM_HIP(hipEventCreateWithFlags(&sv->ready[k], hipEventDisableTiming));
        CSM_HIP(hipHostMalloc(reinterpret_cast<void**>(&sv->pinned[k]), max_batch_ * 17ull, hipHostMallocDefault));
    }
    srv_ = std::move(sv);
    srv_->stager = std::thread([this]() { serve_stage_loop(); });
    srv_->runner = std::thread([this]() { serve_run_loop(); });
    return "";
}

std::string Engine::serve_submit(const uint32_t* src, const uint32_t* dst, const uint32_t* elabel, const uint8_t* op, uint64_t n,
                                 uint64_t& ticket) {
    if (!srv_) return tagged(Code::kState, "the session is not serving");
    if (n > max_batch_) return tagged(Code::kLimit, "batch larger than the scratch was sized for (max_batch_edges)");
    auto r = std::make_shared<ServeReq>();
    r->n = n;
    r->src.assign(src, src + n);
    r->dst.assign(dst, dst + n);
    r->elabel.assign(elabel, elabel + n);
    r->op.assign(op, op + n);
    r->t_submit = std::chrono::steady_clock::now();
    std::lock_guard<std::mutex> lk(srv_->mu);
    if (srv_->stop) return tagged(Code::kState, "the request loop is stopping");
    r->id = srv_->next_id++;
    ticket = r->id;
    srv_->open[r->id] = r;
    srv_->incoming.push_back(r);
    srv_->cv.notify_all();
    return "";
}

void Engine::serve_stage_loop() {
    Server& sv = *srv_;
    (void)hipSetDevice(opt_.device);
    for (;;) {
        std::shared_ptr<ServeReq> r;
        int slot = -1;
        {
            std::unique_lock<std::mutex> lk(sv.mu);
            sv.cv.wait(lk, [&]() { return (sv.stop && sv.incoming.empty()) || (!sv.incoming.empty() && (!sv.slot_busy[0] || !sv.slot_busy[1])); });
            if (sv.incoming.empty()) return;  // stopping, nothing left
            r = sv.incoming.front();
            sv.incoming.pop_front();
            slot = sv.slot_busy[0] ? 1 : 0;
            sv.slot_busy[slot] = true;
        }
        uint8_t* pin = sv.pinned[slot];
        uint32_t* p32 = reinterpret_cast<uint32_t*>(pin);
        const uint64_t n = r->n;
        std::string err = stage_keys(*r, p32 + 2 * n, p32 + 3 * n);
        if (err.empty() && n) {
            std::memcpy(p32, r->src.data(), n * 4);
            std::memcpy(p32 + n, r->dst.data(), n * 4);
            std::memcpy(pin + 16 * n, r->op.data(), n);
            const uint64_t at = stream_cap_ + static_cast<uint64_t>(slot) * max_batch_;
            hipError_t e = hipSuccess;
            auto cp = [&](const void* dev, const void* host, uint64_t bytes) {
                if (e == hipSuccess) e = hipMemcpyAsync(const_cast<void*>(dev), host, bytes, hipMemcpyHostToDevice, sv.copy);
            };
            cp(bin_.src + at, p32, n * 4);
            cp(bin_.dst + at, p32 + n, n * 4);
            cp(bin_.key_fwd + at, p32 + 2 * n, n * 4);
            cp(bin_.key_bwd + at, p32 + 3 * n, n * 4);
            cp(bin_.op + at, pin + 16 * n, n);
            if (e == hipSuccess) e = hipEventRecord(sv.ready[slot], sv.copy);
            // the pinned buffer is reused for the next request of this slot only after the runner
            // has consumed the slot, which it does after this copy completes
            if (e != hipSuccess) err = std::string("HIP error: ") + hipGetErrorString(e) + " staging a request";
        }
        std::lock_guard<std::mutex> lk(sv.mu);
        if (!err.empty()) {
            r->err = err;
            r->done = true;
            r->t_done = std::chrono::steady_clock::now();
            sv.slot_busy[slot] = false;
        } else {
            r->slot = slot;
            sv.staged.push_back(r);
        }
        sv.cv.notify_all();
    }
}

void Engine::serve_run_loop() {
    Server& sv = *srv_;
    (void)hipSetDevice(opt_.device);
    for (;;) {
        std::shared_ptr<ServeReq> r;
        {
            std::unique_lock<std::mutex> lk(sv.mu);
            // a request waits for the ones submitted before it (staging or failed in staging)
            sv.cv.wait(lk, [&]() {
                if (!sv.staged.empty()) return true;
                return sv.stop && sv.incoming.empty() && sv.staged.empty() && !sv.slot_busy[0] && !sv.slot_busy[1];
            });
            if (sv.staged.empty()) return;
            r = sv.staged.front();
            sv.staged.pop_front();
        }
        const int slot = r->slot;
        std::string err;
        if (hipStreamWaitEvent(s_, sv.ready[slot], 0) != hipSuccess)  // the batch build reads the staged copy
            err = tagged(Code::kDevice, "waiting for a staged request failed");
        else
            err = run_resident(stream_cap_ + static_cast<uint64_t>(slot) * max_batch_, r->n, r->src.data(), r->dst.data(), r->op.data(),
                               r->out, r->info);
        std::lock_guard<std::mutex> lk(sv.mu);
        r->err = err;
        r->done = true;
        r->t_done = std::chrono::steady_clock::now();
        sv.slot_busy[slot] = false;
        sv.cv.notify_all();
    }
}

std::string Engine::serve_wait(uint64_t ticket, std::vector<QueryResult>& out, BatchInfo& info, double& latency_ms) {
    if (!srv_) return tagged(Code::kState, "the session is not serving");
    std::shared_ptr<ServeReq> r;
    {
        std::unique_lock<std::mutex> lk(srv_->mu);
        auto it = srv_->open.find(ticket);
        if (it == srv_->open.end()) return tagged(Code::kArgument, "unknown or already collected ticket");
        r = it->second;
        srv_->cv.wait(lk, [&]() { return r->done; });
        srv_->open.erase(ticket);
    }
    latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    out = r->out;
    info = r->info;
    return r->err;
}

std::string Engine::serve_stop() {
    if (!srv_) return "";
    {
        std::lock_guard<std::mutex> lk(srv_->mu);
        srv_->stop = true;
        srv_->cv.notify_all();
    }
    if (srv_->stager.joinable()) srv_->stager.join();
    {
        std::lock_guard<std::mutex> lk(srv_->mu);
        srv_->cv.notify_all();
    }
    if (srv_->runner.joinable()) srv_->runner.join();
    (void)hipStreamSynchronize(srv_->copy);
    for (int k = 0; k < 2; k++) {
        if (srv_->ready[k]) (void)hipEventDestroy(srv_->ready[k]);
        if (srv_->pinned[k]) (void)hipHostFree(srv_->pinned[k]);
    }
    if (srv_->copy) (void)hipStreamDestroy(srv_->copy);
    srv_.reset();
    return "";
}

}  // namespace csm
