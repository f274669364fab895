// server.cpp — request loop over the batched engine (continuous batching).
//
// The reference has no serving mode: every count is one process run
// (src/main.cpp).  The server keeps engine contexts ("lanes") resident: each
// lane owns a stream, its batch buffers and captured graph, a device input area
// and a worker thread.  A free lane takes every queued request that fits its
// HBM share (in arrival order), counts them together with tg_count_many_device
// (one job per request, results identical to counting each alone) and copies
// each request's result into a pooled pinned host buffer.  Requests that arrive
// while a lane is busy are taken by the next free lane, so with two or more
// lanes the work of different batches overlaps on the device.
#include "../../include/tuna_gpu.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

void set_host_error(const std::string& m);

namespace {

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Req {
    uint64_t id = 0;
    const uint8_t* bases = nullptr;
    uint64_t n_bases = 0;
    const uint64_t* off = nullptr;
    uint64_t n_chunks = 0;
    double t_submit = 0, t_start = 0, t_done = 0;
    int status = 0;  // 0 queued, 1 running, 2 done, < 0 error code
    std::string err;
    uint8_t* res = nullptr;  // pinned: keys then counts
    size_t res_cap = 0;
    uint64_t n_out = 0;
    uint32_t lane = 0, batch_jobs = 0;
    tg_job_stats st{};
};

struct Lane {
    tg_ctx* ctx = nullptr;
    hipStream_t copy = nullptr;
    uint8_t* d_in = nullptr;
    uint64_t* d_off = nullptr;
    uint64_t in_cap = 0, off_cap = 0;
    uint64_t* h_off = nullptr;  // pinned staging of rebased offsets
    uint64_t h_off_cap = 0;
    std::thread th;
    uint64_t batches = 0, jobs = 0;
};

}  // namespace

struct tg_server {
    int device = 0;
    tg_params p{};
    tg_server_config cfg{};
    uint64_t lane_bases = 0;  // admission limit per lane batch
    uint64_t budget = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Req*> queue;
    std::unordered_map<uint64_t, std::unique_ptr<Req>> reqs;
    std::vector<std::unique_ptr<Lane>> lanes;
    bool stop = false;
    uint64_t next_id = 1;
    // pinned result buffers, reused: capacity -> free buffers
    std::multimap<size_t, uint8_t*> pool;
    size_t pool_bytes = 0;
    uint64_t completed = 0, failed = 0, rejected = 0;
};

namespace {

constexpr uint64_t BYTES_PER_BASE_NW1 = 64;  // batch class + input + result store, per base
constexpr uint64_t BYTES_PER_BASE_NW2 = 96;

uint8_t* pool_get(tg_server* s, size_t bytes, size_t* cap)
{
    {
        std::lock_guard<std::mutex> g(s->mu);
        auto it = s->pool.lower_bound(bytes);
        if (it != s->pool.end() && it->first <= 2 * bytes + (1u << 20)) {
            uint8_t* p = it->second;
            *cap = it->first;
            s->pool.erase(it);
            return p;
        }
    }
    size_t c = 1u << 16;
    while (c < bytes) c <<= 1;
    void* p = nullptr;
    if (hipHostMalloc(&p, c, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    *cap = c;
    std::lock_guard<std::mutex> g(s->mu);
    s->pool_bytes += c;
    return static_cast<uint8_t*>(p);
}

void pool_put(tg_server* s, uint8_t* p, size_t cap)
{
    if (!p) return;
    s->pool.emplace(cap, p);  // caller holds s->mu
}

int grow_dev(tg_ctx* ctx, void** p, uint64_t* cap, uint64_t need)
{
    if (*cap >= need) return TG_OK;
    if (*p) tg_dev_free(ctx, *p);
    *p = nullptr;
    *cap = 0;
    const uint64_t c = need + need / 4 + 4096;
    const int rc = tg_dev_alloc(ctx, c, p);
    if (rc == TG_OK) *cap = c;
    return rc;
}

// One lane batch: upload, count, copy results out.  Returns an error code for
// the whole batch (each request then carries it).
int run_lane_batch(tg_server* s, Lane& L, std::vector<Req*>& batch, std::string& err)
{
    const uint32_t nw = s->p.k > 32 ? 2u : 1u;
    uint64_t nb = 0, nc = 0;
    for (Req* r : batch) {
        nb += r->n_bases;
        nc += r->n_chunks;
    }
    int rc;
    if ((rc = grow_dev(L.ctx, (void**)&L.d_in, &L.in_cap, nb + 64)) ||
        (rc = grow_dev(L.ctx, (void**)&L.d_off, &L.off_cap, (nc + 1) * 8ull))) {
        err = tg_last_error();
        return rc;
    }
    if (L.h_off_cap < nc + 1) {
        if (L.h_off) hipHostFree(L.h_off);
        L.h_off = nullptr;
        const uint64_t c = (nc + 1) + (nc + 1) / 4 + 1024;
        if (hipHostMalloc((void**)&L.h_off, c * 8, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            err = "pinned staging allocation failed";
            L.h_off_cap = 0;
            return TG_ENOMEM;
        }
        L.h_off_cap = c;
    }
    std::vector<uint64_t> jc(batch.size() + 1), jb(batch.size() + 1);
    uint64_t base = 0, ch = 0;
    L.h_off[0] = 0;
    for (size_t j = 0; j < batch.size(); ++j) {
        Req* r = batch[j];
        jc[j] = ch;
        jb[j] = base;
        for (uint64_t i = 1; i <= r->n_chunks; ++i) L.h_off[ch + i] = r->off[i] + base;
        if (r->n_bases &&
            hipMemcpyAsync(L.d_in + base, r->bases, r->n_bases, hipMemcpyHostToDevice, L.copy) != hipSuccess) {
            err = "input upload failed";
            return TG_EDEVICE;
        }
        base += r->n_bases;
        ch += r->n_chunks;
    }
    jc[batch.size()] = ch;
    jb[batch.size()] = base;
    if (hipMemcpyAsync(L.d_off, L.h_off, (ch + 1) * 8ull, hipMemcpyHostToDevice, L.copy) != hipSuccess ||
        hipStreamSynchronize(L.copy) != hipSuccess) {
        err = "input upload failed";
        return TG_EDEVICE;
    }
    std::vector<tg_job_stats> st(batch.size());
    rc = tg_count_many_device(L.ctx, &s->p, L.d_in, base, L.d_off, ch, jc.data(), jb.data(),
                              (uint32_t)batch.size(), st.data(), nullptr);
    if (rc) {
        err = tg_last_error();
        return rc;
    }
    for (size_t j = 0; j < batch.size(); ++j) {
        Req* r = batch[j];
        const uint64_t* dk = nullptr;
        const uint32_t* dc = nullptr;
        uint64_t n = 0;
        if ((rc = tg_many_result_device(L.ctx, (uint32_t)j, &dk, &dc, &n))) {
            err = tg_last_error();
            return rc;
        }
        const size_t bytes = n * nw * 8ull + n * 4ull;
        r->res = pool_get(s, std::max<size_t>(bytes, 64), &r->res_cap);
        if (!r->res) {
            err = "pinned result allocation failed";
            return TG_ENOMEM;
        }
        if (n && (hipMemcpyAsync(r->res, dk, n * nw * 8ull, hipMemcpyDeviceToHost, L.copy) != hipSuccess ||
                  hipMemcpyAsync(r->res + n * nw * 8ull, dc, n * 4ull, hipMemcpyDeviceToHost, L.copy) !=
                      hipSuccess)) {
            err = "result download failed";
            return TG_EDEVICE;
        }
        r->n_out = n;
        r->st = st[j];
    }
    if (hipStreamSynchronize(L.copy) != hipSuccess) {
        err = "result download failed";
        return TG_EDEVICE;
    }
    return TG_OK;
}

void lane_main(tg_server* s, Lane* L, uint32_t lane_id)
{
    hipSetDevice(s->device);
    for (;;) {
        std::vector<Req*> batch;
        {
            std::unique_lock<std::mutex> g(s->mu);
            s->cv_work.wait(g, [&] { return s->stop || !s->queue.empty(); });
            if (s->stop && s->queue.empty()) return;
            uint64_t nb = 0;
            const double t = now_s();
            while (!s->queue.empty() && batch.size() < s->cfg.max_batch_jobs) {
                Req* r = s->queue.front();
                if (!batch.empty() && nb + r->n_bases > s->lane_bases) break;
                s->queue.pop_front();
                r->status = 1;
                r->t_start = t;
                r->lane = lane_id;
                nb += r->n_bases;
                batch.push_back(r);
            }
        }
        std::string err;
        const int rc = run_lane_batch(s, *L, batch, err);
        const double t = now_s();
        {
            std::lock_guard<std::mutex> g(s->mu);
            for (Req* r : batch) {
                r->t_done = t;
                r->batch_jobs = (uint32_t)batch.size();
                if (rc) {
                    r->status = rc;
                    r->err = err;
                    pool_put(s, r->res, r->res_cap);
                    r->res = nullptr;
                    s->failed += 1;
                } else {
                    r->status = 2;
                    s->completed += 1;
                }
            }
            L->batches += 1;
            L->jobs += batch.size();
        }
        s->cv_done.notify_all();
    }
}

Req* find_req(tg_server* s, uint64_t id)
{
    auto it = s->reqs.find(id);
    return it == s->reqs.end() ? nullptr : it->second.get();
}

}  // namespace

extern "C" {

int tg_host_alloc(uint64_t bytes, void** p)
{
    if (!p) return TG_EINVAL;
    if (hipHostMalloc(p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        set_host_error("pinned host allocation failed");
        return TG_ENOMEM;
    }
    return TG_OK;
}

int tg_host_free(void* p)
{
    if (p && hipHostFree(p) != hipSuccess) return TG_EDEVICE;
    return TG_OK;
}

int tg_server_create(int device, const tg_params* p, const tg_server_config* cfg, tg_server** out)
{
    if (!p || !out) {
        set_host_error("null argument");
        return TG_EINVAL;
    }
    if (p->k > 64 && p->k <= 256) {
        set_host_error("the server counts k <= 64; count jobs of k = " + std::to_string(p->k) +
                       " one at a time (tg_count_host, tg_count_device, tg_count_files)");
        return TG_EINVAL;
    }
    if (p->k < 2 || p->k > 64 || p->m < 1 || p->m > 32 || p->m >= p->k) {
        set_host_error("invalid k/m for the server");
        return TG_EINVAL;
    }
    auto s = std::make_unique<tg_server>();
    s->device = device;
    s->p = *p;
    s->cfg.lanes = 2;
    s->cfg.max_batch_jobs = 64;
    s->cfg.max_batch_bases = 1ull << 27;
    s->cfg.hbm_budget = 0;
    if (cfg) {
        if (cfg->lanes) s->cfg.lanes = cfg->lanes;
        if (cfg->max_batch_jobs) s->cfg.max_batch_jobs = cfg->max_batch_jobs;
        if (cfg->max_batch_bases) s->cfg.max_batch_bases = cfg->max_batch_bases;
        s->cfg.hbm_budget = cfg->hbm_budget;
    }
    if (s->cfg.lanes > 8 || s->cfg.max_batch_jobs > 4096) {
        set_host_error("server config out of range (lanes <= 8, max_batch_jobs <= 4096)");
        return TG_EINVAL;
    }
    if (hipSetDevice(device) != hipSuccess) {
        set_host_error("hipSetDevice failed");
        return TG_EDEVICE;
    }
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
        set_host_error("hipMemGetInfo failed");
        return TG_EDEVICE;
    }
    s->budget = s->cfg.hbm_budget ? s->cfg.hbm_budget : fr / 2;
    if (s->budget > fr) {
        set_host_error("HBM budget exceeds free device memory");
        return TG_ENOMEM;
    }
    const uint64_t bpb = p->k > 32 ? BYTES_PER_BASE_NW2 : BYTES_PER_BASE_NW1;
    s->lane_bases = std::min<uint64_t>(s->cfg.max_batch_bases, s->budget / s->cfg.lanes / bpb);
    if (s->lane_bases < (1u << 16)) {
        set_host_error("HBM budget too small for one lane");
        return TG_ENOMEM;
    }
    for (uint32_t i = 0; i < s->cfg.lanes; ++i) {
        auto L = std::make_unique<Lane>();
        int rc = tg_ctx_create(device, &L->ctx);
        if (rc) return rc;
        tg_batch_config bc;
        tg_get_batch_config(L->ctx, &bc);
        uint64_t bb = 1u << 16;
        while (bb < s->lane_bases && bb < (1ull << 30)) bb <<= 1;
        bc.batch_bases = std::max<uint64_t>(bb, 1u << 16);
        if ((rc = tg_set_batch_config(L->ctx, &bc))) return rc;
        if (hipStreamCreate(&L->copy) != hipSuccess) {
            set_host_error("hipStreamCreate failed");
            return TG_EDEVICE;
        }
        s->lanes.push_back(std::move(L));
    }
    tg_server* raw = s.release();
    for (uint32_t i = 0; i < raw->cfg.lanes; ++i)
        raw->lanes[i]->th = std::thread(lane_main, raw, raw->lanes[i].get(), i);
    *out = raw;
    return TG_OK;
}

void tg_server_destroy(tg_server* s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto& L : s->lanes)
        if (L->th.joinable()) L->th.join();
    hipSetDevice(s->device);
    for (auto& L : s->lanes) {
        if (L->d_in) tg_dev_free(L->ctx, L->d_in);
        if (L->d_off) tg_dev_free(L->ctx, L->d_off);
        if (L->h_off) hipHostFree(L->h_off);
        if (L->copy) hipStreamDestroy(L->copy);
        tg_ctx_destroy(L->ctx);
    }
    for (auto& kv : s->reqs)
        if (kv.second->res) hipHostFree(kv.second->res);
    for (auto& kv : s->pool) hipHostFree(kv.second);
    delete s;
}

int tg_server_submit(tg_server* s, const uint8_t* bases, uint64_t n_bases, const uint64_t* off,
                     uint64_t n_chunks, uint64_t* id)
{
    if (!s || !id || (!bases && n_bases) || !off) {
        set_host_error("null argument");
        return TG_EINVAL;
    }
    if (off[n_chunks] != n_bases || off[0] != 0) {
        set_host_error("chunk offsets must start at 0 and end at n_bases");
        return TG_EINVAL;
    }
    if (n_bases > s->lane_bases) {
        std::lock_guard<std::mutex> g(s->mu);
        s->rejected += 1;
        char b[160];
        snprintf(b, sizeof b, "request of %llu bases exceeds the per-lane HBM admission limit (%llu)",
                 (unsigned long long)n_bases, (unsigned long long)s->lane_bases);
        set_host_error(b);
        return TG_ENOMEM;
    }
    auto r = std::make_unique<Req>();
    r->bases = bases;
    r->n_bases = n_bases;
    r->off = off;
    r->n_chunks = n_chunks;
    r->t_submit = now_s();
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->stop) {
            set_host_error("server is stopping");
            return TG_EINVAL;
        }
        r->id = s->next_id++;
        *id = r->id;
        s->queue.push_back(r.get());
        s->reqs.emplace(r->id, std::move(r));
    }
    s->cv_work.notify_one();
    return TG_OK;
}

int tg_server_poll(tg_server* s, uint64_t id)
{
    if (!s) return TG_EINVAL;
    std::lock_guard<std::mutex> g(s->mu);
    Req* r = find_req(s, id);
    if (!r) {
        set_host_error("unknown request id");
        return TG_EINVAL;
    }
    if (r->status < 0) {
        set_host_error(r->err);
        return r->status;
    }
    return r->status == 2 ? 1 : 0;
}

int tg_server_wait(tg_server* s, uint64_t id, double timeout_s)
{
    if (!s) return TG_EINVAL;
    std::unique_lock<std::mutex> g(s->mu);
    Req* r = find_req(s, id);
    if (!r) {
        set_host_error("unknown request id");
        return TG_EINVAL;
    }
    auto done = [&] { return r->status == 2 || r->status < 0; };
    if (timeout_s < 0) s->cv_done.wait(g, done);
    else if (!s->cv_done.wait_for(g, std::chrono::duration<double>(timeout_s), done)) {
        set_host_error("request not finished within the timeout");
        return TG_ETIMEOUT;
    }
    if (r->status < 0) {
        set_host_error(r->err);
        return r->status;
    }
    return TG_OK;
}

int tg_server_result(tg_server* s, uint64_t id, const uint64_t** keys, const uint32_t** counts,
                     uint64_t* n, tg_job_stats* st, tg_request_times* t)
{
    if (!s) return TG_EINVAL;
    std::lock_guard<std::mutex> g(s->mu);
    Req* r = find_req(s, id);
    if (!r) {
        set_host_error("unknown request id");
        return TG_EINVAL;
    }
    if (r->status < 0) {
        set_host_error(r->err);
        return r->status;
    }
    if (r->status != 2) {
        set_host_error("request not finished");
        return TG_ETIMEOUT;
    }
    const uint32_t nw = s->p.k > 32 ? 2u : 1u;
    if (keys) *keys = reinterpret_cast<const uint64_t*>(r->res);
    if (counts) *counts = reinterpret_cast<const uint32_t*>(r->res + r->n_out * nw * 8ull);
    if (n) *n = r->n_out;
    if (st) *st = r->st;
    if (t) {
        t->submit_to_start_s = r->t_start - r->t_submit;
        t->submit_to_done_s = r->t_done - r->t_submit;
        t->lane = r->lane;
        t->batch_jobs = r->batch_jobs;
    }
    return TG_OK;
}

int tg_server_release(tg_server* s, uint64_t id)
{
    if (!s) return TG_EINVAL;
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) {
        set_host_error("unknown request id");
        return TG_EINVAL;
    }
    Req* r = it->second.get();
    if (r->status == 0 || r->status == 1) {
        set_host_error("request still queued or running");
        return TG_EINVAL;
    }
    pool_put(s, r->res, r->res_cap);
    s->reqs.erase(it);
    return TG_OK;
}

int tg_server_stats(tg_server* s, tg_server_info* out)
{
    if (!s || !out) return TG_EINVAL;
    std::lock_guard<std::mutex> g(s->mu);
    tg_server_info o{};
    o.lanes = s->cfg.lanes;
    o.lane_bases = s->lane_bases;
    o.hbm_budget = s->budget;
    o.queued = s->queue.size();
    o.completed = s->completed;
    o.failed = s->failed;
    o.rejected = s->rejected;
    o.pinned_pool_bytes = s->pool_bytes;
    for (auto& L : s->lanes) {
        o.batches += L->batches;
    }
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) == hipSuccess) o.hbm_free = fr;
    *out = o;
    return TG_OK;
}

}  // extern "C"
