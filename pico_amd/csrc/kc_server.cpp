// kc_server.cpp -- request loop for serving k-core decompositions and edge
// updates from one process on one GPU (no counterpart in the reference).
//
//  * Requests enter a FIFO.  A worker that comes free admits what the FIFO
//    head allows: a request is admitted when its HBM footprint estimate fits
//    in the budget left by the requests in flight (a request that does not
//    fit stops admission, so large requests are not starved), and updates or
//    decompositions of a resident graph wait for the previous request on
//    that graph (they apply in submission order).
//  * Continuous batching: the small fresh-graph decompositions admitted
//    together are fused into one batched launch (kc_batch_*), whatever
//    arrived while the other workers were busy.
//  * Every worker owns a HIP stream, so admitted work overlaps on the
//    device.  Results are the engines' own: the same bits alone or in a mix.
#include "kc_capi_impl.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

using namespace kci;


namespace {

struct Req {
    uint64_t id = 0;
    kc_request r{};
    std::vector<uint32_t> u, v; // update edges, copied at submit
    double t_submit = 0;
    uint64_t bytes = 0;
    // UPDATE / DECOMPOSE_GRAPH: the resident graph and its context, copied
    // under the server lock at admission (the worker runs without the lock)
    kc_graph *g = nullptr;
    kc_ctx *gctx = nullptr;
};

struct Resident {
    kc_graph *g = nullptr;
    kc_ctx *ctx = nullptr;
    bool busy = false;
};

} // namespace

struct kc_server {
    kc_server_opts o{};
    int device = 0;
    std::vector<kc_ctx *> ctxs; // one per worker
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Req *> pending;
    std::deque<kc_completion> done;
    std::deque<Resident> graphs; // read and extended only under mu
    uint64_t next_id = 1;
    uint64_t inflight_bytes = 0;
    uint32_t busy = 0;
    uint64_t outstanding = 0; // submitted, not yet completed
    bool stop = false;
    kc_server_stats st{};
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    double now() const
    {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
            .count();
    }
};

namespace {

constexpr double RMAT_A = 0.57, RMAT_B = 0.19, RMAT_C = 0.19;

uint64_t rmat_m_gen(const kc_request &r)
{
    return (uint64_t)std::llround(r.edge_factor * (double)(1ull << r.scale));
}

// HBM footprint estimate: the device build (two sort buffers of 2 keys per
// generated edge, the unsorted column bound, row offsets) plus the engine's
// per-vertex scratch.  The build's buffers are released before the peel
// runs, so its peak also covers the peel's compacted stages (two sets of at
// most 8 bytes per undirected edge, 16 per vertex).
uint64_t footprint(const kc_server *s, const kc_request &r, size_t nedges)
{
    switch (r.kind) {
    case KC_REQ_DECOMPOSE_RMAT: {
        uint64_t V = 1ull << r.scale;
        return 48 * rmat_m_gen(r) + 48 * V + (16ull << 20);
    }
    case KC_REQ_UPDATE:
        return 16 * (uint64_t)nedges + (1ull << 20);
    case KC_REQ_DECOMPOSE_GRAPH: {
        // per-vertex scratch, plus the compacted stages of a bucket peel
        const kc_graph *g = s->graphs[r.graph].g;
        return 72ull * g->V + 8ull * g->nnz + (1ull << 20);
    }
    default:
        return 0;
    }
}

bool graph_kind(const kc_request &r)
{
    return r.kind == KC_REQ_UPDATE || r.kind == KC_REQ_DECOMPOSE_GRAPH;
}

// Pick the next job under the lock: the admissible requests from the FIFO,
// fusing small fresh decompositions.  Empty when nothing can start.
std::vector<Req *> admit(kc_server *s)
{
    std::vector<Req *> job;
    std::vector<uint32_t> blocked; // resident graphs with an earlier request waiting
    uint64_t budget = s->o.hbm_budget;
    for (auto it = s->pending.begin(); it != s->pending.end();) {
        Req *q = *it;
        if (graph_kind(q->r)) {
            uint32_t gid = q->r.graph;
            bool wait = s->graphs[gid].busy ||
                        std::find(blocked.begin(), blocked.end(), gid) != blocked.end();
            if (wait || !job.empty()) {
                blocked.push_back(gid);
                ++it;
                continue;
            }
        }
        if (s->inflight_bytes + q->bytes > budget && s->inflight_bytes > 0) {
            s->st.admission_waits++;
            break; // FIFO on memory: nothing overtakes a request that does not fit
        }
        bool small = q->r.kind == KC_REQ_DECOMPOSE_RMAT && q->r.scale <= s->o.fuse_max_scale;
        if (!job.empty()) {
            // only small fresh decompositions join a fused job
            if (!small || job.size() >= s->o.max_batch)
                break;
        }
        job.push_back(q);
        s->inflight_bytes += q->bytes;
        it = s->pending.erase(it);
        if (graph_kind(q->r)) {
            Resident &rg = s->graphs[q->r.graph];
            rg.busy = true;
            q->g = rg.g;
            q->gctx = rg.ctx;
            break;
        }
        if (!small)
            break;
    }
    return job;
}

// device checksum of a core vector, on the stream that produced it
kc_status checksum_of(const uint32_t *core, uint32_t V, hipStream_t st, uint64_t &out)
{
    uint64_t *d = nullptr;
    KC_HIP(hipMallocAsync((void **)&d, 8, st), "checksum buffer");
    hipError_t e = kc::core_checksum(core, V, d, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(&out, d, 8, hipMemcpyDeviceToHost, st);
    hipError_t e2 = hipFreeAsync(d, st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    if (e == hipSuccess)
        e = e2;
    return e == hipSuccess ? KC_OK : hip_fail(e, "checksum");
}

kc_completion make_completion(const Req *q, double t_start)
{
    kc_completion c{};
    c.id = q->id;
    c.tag = q->r.tag;
    c.kind = q->r.kind;
    c.submit_ms = q->t_submit;
    c.start_ms = t_start;
    return c;
}

void run_job(kc_server *s, kc_ctx *ctx, std::vector<Req *> &job, std::vector<kc_completion> &out)
{
    double t_start = s->now();
    Req *q0 = job[0];
    if (q0->r.kind == KC_REQ_DECOMPOSE_RMAT && q0->r.scale <= s->o.fuse_max_scale) {
        // small graphs, fused: one device build and one batched launch for
        // all of them (a lone small graph takes the same path)
        uint32_t G = (uint32_t)job.size();
        std::vector<uint32_t> sc(G);
        std::vector<uint64_t> mg(G), seeds(G), sums(G);
        for (uint32_t i = 0; i < G; i++) {
            sc[i] = job[i]->r.scale;
            mg[i] = rmat_m_gen(job[i]->r);
            seeds[i] = job[i]->r.seed;
        }
        kc_batch *b = nullptr;
        std::vector<uint32_t> km(G);
        kc_batch_stats bs{};
        kc_status st = kc_batch_gen_rmat_list(ctx, G, sc.data(), mg.data(), seeds.data(), RMAT_A,
                                              RMAT_B, RMAT_C, &b);
        if (st == KC_OK)
            st = kc_batch_run(b, nullptr, nullptr, km.data(), &bs);
        if (st == KC_OK)
            st = kc_batch_checksums(b, sums.data());
        for (uint32_t i = 0; i < G && st == KC_OK; i++)
            if (job[i]->r.core_out)
                st = kc_batch_graph_cores(b, i, job[i]->r.core_out);
        kc_batch_free(b);
        double t_done = s->now();
        for (uint32_t i = 0; i < G; i++) {
            kc_completion c = make_completion(job[i], t_start);
            c.status = st;
            c.done_ms = t_done;
            c.V = 1u << sc[i];
            c.kmax = km[i];
            c.checksum = sums[i];
            c.fused = G;
            out.push_back(c);
        }
        return;
    }
    kc_completion c = make_completion(q0, t_start);
    c.fused = 1;
    kc_status st = KC_OK;
    // A large decomposition takes a share of the CUs (its grid and its
    // compacted stages' grid), so that as many as there are workers run side
    // by side instead of one after another; cores do not depend on the grid.
    kc_opts share{};
    share.grid = std::max(16u, ctx->num_cus / std::max(1u, s->o.slots));
    share.stage_grid = std::min(share.grid, std::max(1u, ctx->num_cus / 2));
    if (q0->r.kind == KC_REQ_DECOMPOSE_RMAT) {
        kc_graph *g = nullptr;
        st = kc_graph_gen_rmat(ctx, q0->r.scale, rmat_m_gen(q0->r), RMAT_A, RMAT_B, RMAT_C,
                               q0->r.seed, 4, &g);
        kc_stats ks{};
        if (st == KC_OK)
            st = kc_decompose(g, KC_PEELONE_DYN, &share, q0->r.core_out, &ks);
        if (st == KC_OK)
            st = checksum_of(g->result, g->V, ctx->stream, c.checksum);
        c.V = g ? g->V : 0;
        c.kmax = (uint32_t)ks.kmax;
        kc_graph_free(g);
    } else {
        kc_graph *g = q0->g;
        if (q0->r.kind == KC_REQ_UPDATE) {
            kc_update_stats us{};
            st = kc_update(g, (kc_update_op)q0->r.op, q0->u.data(), q0->v.data(), q0->u.size(),
                           q0->r.core_out, &us);
            c.applied = us.applied;
        } else {
            kc_stats ks{};
            st = kc_decompose(g, KC_PEELONE_DYN, &share, q0->r.core_out, &ks);
        }
        if (st == KC_OK)
            st = checksum_of(g->result, g->V, q0->gctx->stream, c.checksum);
        c.V = g->V;
        c.kmax = 0;
    }
    c.status = st;
    c.done_ms = s->now();
    out.push_back(c);
}

void worker(kc_server *s, uint32_t wi)
{
    kc_ctx *ctx = s->ctxs[wi];
    hipSetDevice(ctx->device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        std::vector<Req *> job;
        s->cv_work.wait(lk, [&] {
            if (s->stop)
                return true;
            job = admit(s);
            return !job.empty();
        });
        if (job.empty() && s->stop)
            return;
        s->busy++;
        s->st.max_busy = std::max<uint64_t>(s->st.max_busy, s->busy);
        s->st.max_inflight_bytes = std::max(s->st.max_inflight_bytes, s->inflight_bytes);
        s->st.jobs++;
        if (job.size() > 1) {
            s->st.fused_jobs++;
            s->st.fused_requests += job.size();
        }
        lk.unlock();
        std::vector<kc_completion> out;
        run_job(s, ctx, job, out);
        lk.lock();
        s->busy--;
        for (Req *q : job) {
            s->inflight_bytes -= q->bytes;
            if (graph_kind(q->r))
                s->graphs[q->r.graph].busy = false;
            delete q;
        }
        for (auto &c : out) {
            s->done.push_back(c);
            s->outstanding--;
            s->st.completed++;
            if (c.status)
                s->st.failed++;
        }
        s->cv_done.notify_all();
        s->cv_work.notify_all(); // memory and graphs came free
    }
}

} // namespace

extern "C" {

kc_status kc_server_create(int device, const kc_server_opts *opts, kc_server **out)
{
    if (!out)
        return fail(KC_EINVAL, "null argument");
    auto *s = new (std::nothrow) kc_server;
    if (!s)
        return fail(KC_ENOMEM, "host allocation");
    if (opts)
        s->o = *opts;
    s->device = device;
    if (!s->o.slots)
        s->o.slots = 4;
    if (s->o.slots > 64)
        s->o.slots = 64;
    if (!s->o.max_batch)
        s->o.max_batch = 256;
    if (!s->o.fuse_max_scale)
        s->o.fuse_max_scale = 16;
    for (uint32_t i = 0; i < s->o.slots; i++) {
        kc_ctx *c = nullptr;
        kc_status st = kc_ctx_create(device, &c);
        if (st) {
            for (kc_ctx *x : s->ctxs)
                kc_ctx_destroy(x);
            delete s;
            return st;
        }
        s->ctxs.push_back(c);
    }
    if (!s->o.hbm_budget)
        s->o.hbm_budget = s->ctxs[0]->hbm / 2;
    for (uint32_t i = 0; i < s->o.slots; i++)
        s->workers.emplace_back(worker, s, i);
    *out = s;
    return KC_OK;
}

kc_status kc_server_gen_rmat_graph(kc_server *s, uint32_t scale, double edge_factor,
                                   uint64_t seed, uint32_t *graph_id)
{
    if (!s || !graph_id)
        return fail(KC_EINVAL, "null argument");
    kc_ctx *c = nullptr;
    kc_status st = kc_ctx_create(s->device, &c);
    if (st)
        return st;
    kc_graph *g = nullptr;
    uint64_t m_gen = (uint64_t)std::llround(edge_factor * (double)(1ull << scale));
    st = kc_graph_gen_rmat(c, scale, m_gen, RMAT_A, RMAT_B, RMAT_C, seed, 4, &g);
    // resident graphs take updates: build their maintenance state now, not in
    // the first update request
    if (!st)
        st = kc_update_prepare(g);
    if (st) {
        if (g)
            kc_graph_free(g);
        kc_ctx_destroy(c);
        return st;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    s->graphs.push_back(Resident{g, c, false});
    *graph_id = (uint32_t)s->graphs.size() - 1;
    return KC_OK;
}

kc_status kc_server_graph_download(kc_server *s, uint32_t graph_id, kc_host_graph **out)
{
    if (!s)
        return fail(KC_EINVAL, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    if (graph_id >= s->graphs.size())
        return fail(KC_EINVAL, "no graph %u", graph_id);
    s->cv_done.wait(lk, [&] { return !s->graphs[graph_id].busy; });
    return kc_graph_download(s->graphs[graph_id].g, out);
}

kc_status kc_server_submit(kc_server *s, const kc_request *r, uint64_t *id)
{
    if (!s || !r)
        return fail(KC_EINVAL, "null argument");
    auto *q = new (std::nothrow) Req;
    if (!q)
        return fail(KC_ENOMEM, "host allocation");
    q->r = *r;
    switch (r->kind) {
    case KC_REQ_DECOMPOSE_RMAT:
        if (r->scale == 0 || r->scale > 31 || r->edge_factor < 0) {
            delete q;
            return fail(KC_EINVAL, "bad R-MAT request");
        }
        break;
    case KC_REQ_UPDATE:
        if (r->op != KC_INSERT && r->op != KC_DELETE) {
            delete q;
            return fail(KC_EINVAL, "bad update op");
        }
        if (r->n && (!r->u || !r->v)) {
            delete q;
            return fail(KC_EINVAL, "null edge arrays");
        }
        q->u.assign(r->u, r->u + r->n);
        q->v.assign(r->v, r->v + r->n);
        q->r.u = q->r.v = nullptr;
        [[fallthrough]];
    case KC_REQ_DECOMPOSE_GRAPH: {
        std::lock_guard<std::mutex> lk(s->mu);
        if (r->graph >= s->graphs.size()) {
            delete q;
            return fail(KC_EINVAL, "no graph %u", r->graph);
        }
        break;
    }
    default:
        delete q;
        return fail(KC_EINVAL, "unknown request kind %u", r->kind);
    }
    std::lock_guard<std::mutex> lk(s->mu);
    q->bytes = footprint(s, q->r, q->u.size());
    if (q->bytes > s->o.hbm_budget) {
        delete q;
        return fail(KC_ENOMEM, "request needs ~%llu bytes, budget is %llu",
                    (unsigned long long)q->bytes, (unsigned long long)s->o.hbm_budget);
    }
    q->id = s->next_id++;
    q->t_submit = s->now();
    if (id)
        *id = q->id;
    s->pending.push_back(q);
    s->outstanding++;
    s->st.submitted++;
    s->cv_work.notify_one();
    return KC_OK;
}

kc_status kc_server_poll(kc_server *s, kc_completion *out, uint32_t max, uint32_t *n,
                         int wait_ms)
{
    if (!s || !n || (max && !out))
        return fail(KC_EINVAL, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->done.empty() && wait_ms != 0) {
        auto pred = [&] { return !s->done.empty() || s->outstanding == 0; };
        if (wait_ms < 0)
            s->cv_done.wait(lk, pred);
        else
            s->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), pred);
    }
    uint32_t k = 0;
    while (k < max && !s->done.empty()) {
        out[k++] = s->done.front();
        s->done.pop_front();
    }
    *n = k;
    return KC_OK;
}

kc_status kc_server_drain(kc_server *s)
{
    if (!s)
        return fail(KC_EINVAL, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->outstanding == 0; });
    return KC_OK;
}

double kc_server_now_ms(const kc_server *s) { return s ? s->now() : 0.0; }

uint64_t kc_core_checksum(const uint32_t *core, uint32_t V)
{
    return core ? kch::core_checksum(core, V) : 0;
}

kc_status kc_server_get_stats(kc_server *s, kc_server_stats *out)
{
    if (!s || !out)
        return fail(KC_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->st;
    out->hbm_budget = s->o.hbm_budget;
    return KC_OK;
}

void kc_server_destroy(kc_server *s)
{
    if (!s)
        return;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_done.wait(lk, [&] { return s->outstanding == 0; });
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->workers)
        t.join();
    for (auto &rg : s->graphs) {
        kc_graph_free(rg.g);
        kc_ctx_destroy(rg.ctx);
    }
    for (kc_ctx *c : s->ctxs)
        kc_ctx_destroy(c);
    for (Req *q : s->pending)
        delete q;
    delete s;
}

} // extern "C"
