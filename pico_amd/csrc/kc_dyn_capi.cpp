// kc_dyn_capi.cpp -- kc_update: core maintenance under batched edge
// insertions and deletions on a device-resident graph (no counterpart in the
// reference, which only decomposes static graphs).
//
// Per call: the batch is validated and normalised on the host (ids < V,
// self-loops dropped, each pair once).  Insertions are split into matchings
// (no vertex twice in one matching) and applied matching by matching on the
// slack CSR, the cores maintained after every matching (raise the candidate
// region by one, lower to the fixed point).  Deletions are applied in one
// launch, one wave per edited row, then lowered from the roots of all of
// them.
#include "kc_capi_impl.hpp"

#include <algorithm>
#include <cstring>
#include <new>
#include <unordered_map>
#include <vector>

using namespace kci;

struct DynState {
    uint32_t V = 0;
    uint64_t *off = nullptr; // V+1
    uint32_t *len = nullptr, *adj = nullptr, *core = nullptr, *core_prev = nullptr;
    uint32_t *mark = nullptr, *visit = nullptr; // per-vertex stamps
    uint32_t *list[3] = {nullptr, nullptr, nullptr};
    uint32_t *cnt = nullptr; // device counters: [0] list length, [1] candidates, [2] drops, [3] changed
    uint32_t *eu = nullptr, *ev = nullptr;
    uint8_t *status = nullptr;
    uint32_t ecap = 0;
    // batched deletions: per edited row its vertex and segment start, and
    // per half-edge the neighbour to drop and the edge's index
    uint32_t *hv = nullptr, *hs = nullptr, *hn = nullptr, *he = nullptr;
    uint64_t hcap = 0;
    uint64_t slots = 0;
    uint32_t stamp = 0;
    std::vector<uint64_t> off_h;
    uint64_t bytes = 0;
};

void dyn_free(DynState *d)
{
    if (!d)
        return;
    for (void *p : {(void *)d->off, (void *)d->len, (void *)d->adj, (void *)d->core,
                    (void *)d->core_prev, (void *)d->mark, (void *)d->visit, (void *)d->list[0],
                    (void *)d->list[1], (void *)d->list[2], (void *)d->cnt, (void *)d->eu,
                    (void *)d->ev, (void *)d->status, (void *)d->hv, (void *)d->hs,
                    (void *)d->hn, (void *)d->he})
        if (p)
            hipFree(p);
    delete d;
}

namespace {

// spare slots per row: a few for every vertex, more for high degrees
inline uint64_t slack_for(uint64_t deg) { return 4 + deg / 16; }

kc_status dmalloc(DynState *d, void **p, size_t bytes, const char *what)
{
    hipError_t e = hipMalloc(p, bytes ? bytes : 4);
    if (e != hipSuccess)
        return hip_fail(e, what);
    d->bytes += bytes;
    return KC_OK;
}

// lay rows out with capacities cap[v]; copies the current rows over
kc_status relayout(DynState *d, const std::vector<uint64_t> &cap, hipStream_t st)
{
    const uint32_t V = d->V;
    std::vector<uint64_t> off(V + 1ull, 0);
    for (uint32_t v = 0; v < V; v++)
        off[v + 1] = off[v] + cap[v];
    uint64_t *noff = nullptr;
    uint32_t *nadj = nullptr;
    kc_status s;
    if ((s = dmalloc(d, (void **)&noff, (V + 1ull) * 8, "slack offsets")))
        return s;
    if ((s = dmalloc(d, (void **)&nadj, off[V] * 4, "slack adjacency"))) {
        hipFree(noff);
        return s;
    }
    hipError_t e = hipMemcpyAsync(noff, off.data(), (V + 1ull) * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = kc::dyn_copy(d->off, d->len, d->adj, noff, nadj, V, st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        hipFree(noff);
        hipFree(nadj);
        return hip_fail(e, "slack re-layout");
    }
    hipFree(d->off);
    hipFree(d->adj);
    d->bytes -= (V + 1ull) * 8 + d->slots * 4;
    d->off = noff;
    d->adj = nadj;
    d->slots = off[V];
    d->off_h.swap(off);
    return KC_OK;
}

kc_status dyn_create(kc_graph *g)
{
    kc_ctx *c = g->ctx;
    hipStream_t st = c->stream;
    auto *d = new (std::nothrow) DynState;
    if (!d)
        return fail(KC_ENOMEM, "host allocation");
    const uint32_t V = g->V;
    d->V = V;
    // current cores from the bucket peel
    kc_stats pst{};
    kc_status s = run_method(g, KC_PEELONE_DYN, kc_opts{}, pst);
    if (s) {
        delete d;
        return s;
    }
    size_t V4 = (size_t)(V ? V : 1) * 4;
    std::vector<uint64_t> rp(V + 1ull);
    hipError_t e;
    if (g->wide) {
        e = hipMemcpy(rp.data(), g->rowptr, (V + 1ull) * 8, hipMemcpyDeviceToHost);
    } else {
        std::vector<uint32_t> r32(V + 1ull);
        e = hipMemcpy(r32.data(), g->rowptr, (V + 1ull) * 4, hipMemcpyDeviceToHost);
        for (size_t v = 0; v <= V; v++)
            rp[v] = r32[v];
    }
    if (e != hipSuccess) {
        delete d;
        return hip_fail(e, "rowptr copy");
    }
    d->off_h.assign(V + 1ull, 0);
    for (uint32_t v = 0; v < V; v++) {
        uint64_t deg = rp[v + 1] - rp[v];
        d->off_h[v + 1] = d->off_h[v] + deg + slack_for(deg);
    }
    d->slots = d->off_h[V];
    if ((s = dmalloc(d, (void **)&d->off, (V + 1ull) * 8, "slack offsets")) ||
        (s = dmalloc(d, (void **)&d->adj, d->slots * 4, "slack adjacency")) ||
        (s = dmalloc(d, (void **)&d->len, V4, "row lengths")) ||
        (s = dmalloc(d, (void **)&d->core, V4, "cores")) ||
        (s = dmalloc(d, (void **)&d->core_prev, V4, "cores")) ||
        (s = dmalloc(d, (void **)&d->mark, V4, "marks")) ||
        (s = dmalloc(d, (void **)&d->visit, V4, "marks")) ||
        (s = dmalloc(d, (void **)&d->list[0], V4, "work list")) ||
        (s = dmalloc(d, (void **)&d->list[1], V4, "work list")) ||
        (s = dmalloc(d, (void **)&d->list[2], V4, "candidate list")) ||
        (s = dmalloc(d, (void **)&d->cnt, 16, "counters"))) {
        dyn_free(d);
        return s;
    }
    e = hipMemcpyAsync(d->off, d->off_h.data(), (V + 1ull) * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = kc::dyn_from_csr(g->rowptr, g->wide, g->col, d->off, d->len, d->adj, V, st);
    if (e == hipSuccess && V)
        e = hipMemcpyAsync(d->core, g->result, (size_t)V * 4, hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess)
        e = hipMemsetAsync(d->mark, 0, V4, st);
    if (e == hipSuccess)
        e = hipMemsetAsync(d->visit, 0, V4, st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        dyn_free(d);
        return hip_fail(e, "dynamic graph set-up");
    }
    g->dyn = d;
    return KC_OK;
}

uint32_t next_stamp(DynState *d, hipStream_t st)
{
    if (++d->stamp == 0) { // wrapped: clear the stamp arrays once
        size_t V4 = (size_t)(d->V ? d->V : 1) * 4;
        hipMemsetAsync(d->mark, 0, V4, st);
        hipMemsetAsync(d->visit, 0, V4, st);
        d->stamp = 1;
    }
    return d->stamp;
}

kc_status read_u32(const uint32_t *dev, uint32_t &out, hipStream_t st)
{
    KC_HIP(hipMemcpyAsync(&out, dev, 4, hipMemcpyDeviceToHost, st), "counter copy");
    KC_HIP(hipStreamSynchronize(st), "counter copy");
    return KC_OK;
}

kc::DynView view(const DynState *d)
{
    return kc::DynView{d->V, d->off, d->len, d->adj, d->core};
}

// Lower from the worklist in list[0] (n entries) to the fixed point.
kc_status lower(DynState *d, uint32_t n, kc_update_stats &us, hipStream_t st)
{
    uint32_t *in = d->list[0], *out = d->list[1];
    while (n) {
        uint32_t stamp = next_stamp(d, st);
        KC_HIP(hipMemsetAsync(d->cnt, 0, 4, st), "counter clear");
        KC_HIP(kc::dyn_round(view(d), in, n, out, d->cnt, d->mark, stamp, d->cnt + 2, d->V, st),
               "lowering round");
        us.evaluations += n;
        us.rounds++;
        kc_status s = read_u32(d->cnt, n, st);
        if (s)
            return s;
        std::swap(in, out);
    }
    return KC_OK;
}

// Apply one matching (n edges already on the device in eu/ev); grows rows
// that are full and retries those edges.  Returns the applied count.
kc_status edit_matching(kc_graph *g, DynState *d, uint32_t op, uint32_t n,
                        std::vector<uint32_t> &mu, std::vector<uint32_t> &mv,
                        std::vector<uint8_t> &stat, kc_update_stats &us, hipStream_t st)
{
    stat.assign(n, 0);
    KC_HIP(kc::dyn_edit(view(d), d->eu, d->ev, n, op, d->status, st), "edge edit");
    KC_HIP(hipMemcpyAsync(stat.data(), d->status, n, hipMemcpyDeviceToHost, st), "status copy");
    KC_HIP(hipStreamSynchronize(st), "edge edit");
    std::vector<uint32_t> ru, rv, idx;
    for (uint32_t i = 0; i < n; i++)
        if (stat[i] == 2) {
            ru.push_back(mu[i]);
            rv.push_back(mv[i]);
            idx.push_back(i);
        }
    if (!ru.empty()) {
        // grow: capacity for the current length, these edges and fresh slack
        std::vector<uint32_t> len(d->V);
        KC_HIP(hipMemcpy(len.data(), d->len, (size_t)d->V * 4, hipMemcpyDeviceToHost),
               "length copy");
        std::vector<uint64_t> cap(d->V);
        for (uint32_t v = 0; v < d->V; v++)
            cap[v] = std::max<uint64_t>(d->off_h[v + 1] - d->off_h[v],
                                        len[v] + slack_for(len[v]));
        // full rows double, so a vertex that keeps gaining edges is laid
        // out again only logarithmically often
        auto grown = [&](uint32_t x) {
            return len[x] + 1 + std::max<uint64_t>(slack_for(len[x]), len[x]);
        };
        for (size_t i = 0; i < ru.size(); i++) {
            cap[ru[i]] = std::max<uint64_t>(cap[ru[i]], grown(ru[i]));
            cap[rv[i]] = std::max<uint64_t>(cap[rv[i]], grown(rv[i]));
        }
        kc_status s = relayout(d, cap, st);
        if (s)
            return s;
        us.grows++;
        uint32_t k = (uint32_t)ru.size();
        KC_HIP(hipMemcpyAsync(d->eu, ru.data(), k * 4ull, hipMemcpyHostToDevice, st), "edges");
        KC_HIP(hipMemcpyAsync(d->ev, rv.data(), k * 4ull, hipMemcpyHostToDevice, st), "edges");
        KC_HIP(kc::dyn_edit(view(d), d->eu, d->ev, k, op, d->status, st), "edge edit");
        std::vector<uint8_t> s2(k);
        KC_HIP(hipMemcpyAsync(s2.data(), d->status, k, hipMemcpyDeviceToHost, st), "status copy");
        KC_HIP(hipStreamSynchronize(st), "edge edit");
        for (uint32_t i = 0; i < k; i++) {
            if (s2[i] == 2)
                return fail(KC_EDEVICE, "row still full after growing");
            stat[idx[i]] = s2[i];
        }
        // the device edge arrays hold the retried subset: restore the matching
        KC_HIP(hipMemcpyAsync(d->eu, mu.data(), n * 4ull, hipMemcpyHostToDevice, st), "edges");
        KC_HIP(hipMemcpyAsync(d->ev, mv.data(), n * 4ull, hipMemcpyHostToDevice, st), "edges");
        KC_HIP(hipMemcpyAsync(d->status, stat.data(), n, hipMemcpyHostToDevice, st), "status");
    }
    uint64_t applied = 0;
    for (uint8_t x : stat)
        applied += x == 1;
    us.applied += applied;
    if (op == 0)
        g->nnz += 2 * applied;
    else
        g->nnz -= 2 * applied;
    return KC_OK;
}

} // namespace

kc_status kci::dyn_materialize(kc_graph *g)
{
    DynState *d = g->dyn;
    if (!d || !g->csr_stale)
        return KC_OK;
    kc_ctx *c = g->ctx;
    hipStream_t st = c->stream;
    const uint32_t V = g->V;
    std::vector<uint32_t> len(V);
    if (V)
        KC_HIP(hipMemcpy(len.data(), d->len, (size_t)V * 4, hipMemcpyDeviceToHost),
               "length copy");
    std::vector<uint64_t> rp(V + 1ull, 0);
    for (uint32_t v = 0; v < V; v++)
        rp[v + 1] = rp[v] + len[v];
    uint64_t nnz = rp[V];
    bool wide = g->wide || nnz > 0xffffffffull;
    size_t rb = (V + 1ull) * (wide ? 8 : 4);
    void *nrp = nullptr;
    uint32_t *ncol = nullptr;
    KC_HIP(hipMalloc(&nrp, rb), "rowptr");
    hipError_t e = hipMalloc((void **)&ncol, nnz ? nnz * 4 : 4);
    if (e != hipSuccess) {
        hipFree(nrp);
        return hip_fail(e, "col");
    }
    if (wide) {
        e = hipMemcpyAsync(nrp, rp.data(), rb, hipMemcpyHostToDevice, st);
    } else {
        std::vector<uint32_t> r32(rp.begin(), rp.end());
        e = hipMemcpyAsync(nrp, r32.data(), rb, hipMemcpyHostToDevice, st);
        if (e == hipSuccess)
            e = hipStreamSynchronize(st);
    }
    if (e == hipSuccess)
        e = kc::dyn_to_csr(d->off, d->len, d->adj, nrp, wide, ncol, V, st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        hipFree(nrp);
        hipFree(ncol);
        return hip_fail(e, "CSR rebuild");
    }
    g->bytes -= (V + 1ull) * (g->wide ? 8 : 4) + (g->nnz ? g->nnz * 4 : 4);
    hipFree(g->rowptr);
    hipFree(g->col);
    g->rowptr = nrp;
    g->col = ncol;
    g->wide = wide;
    g->nnz = nnz;
    g->bytes += rb + (nnz ? nnz * 4 : 4);
    g->degsorted = false;
    g->csr_stale = false;
    return KC_OK;
}

extern "C" kc_status kc_update(kc_graph *g, kc_update_op op, const uint32_t *u,
                               const uint32_t *v, uint64_t n, uint32_t *core_out,
                               kc_update_stats *stats)
{
    if (!g)
        return fail(KC_EINVAL, "null graph");
    if (op != KC_INSERT && op != KC_DELETE)
        return fail(KC_EINVAL, "op must be KC_INSERT or KC_DELETE");
    if (n && (!u || !v))
        return fail(KC_EINVAL, "null edge arrays");
    if (!g->owns_csr)
        return fail(KC_EINVAL, "graph is a view into a batch");
    double t0 = now_ms();
    kc_update_stats us{};
    us.requested = n;
    const uint32_t V = g->V;
    // normalise: ids in range, no self-loops, each unordered pair once
    std::vector<uint64_t> keys;
    keys.reserve(n);
    for (uint64_t i = 0; i < n; i++) {
        if (u[i] >= V || v[i] >= V)
            return fail(KC_EINVAL, "edge %llu (%u, %u) out of range (V=%u)",
                        (unsigned long long)i, u[i], v[i], V);
        if (u[i] == v[i])
            continue;
        uint32_t a = std::min(u[i], v[i]), b = std::max(u[i], v[i]);
        keys.push_back(((uint64_t)a << 32) | b);
    }
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    KC_HIP(hipSetDevice(g->ctx->device), "hipSetDevice");
    hipStream_t st = g->ctx->stream;
    if (!g->dyn) {
        if (g->csr_stale)
            return fail(KC_EDEVICE, "stale CSR without dynamic state");
        kc_status s = dyn_create(g);
        if (s)
            return s;
    }
    DynState *d = g->dyn;
    // greedy split into matchings: edge -> first matching after both ends'
    // (insertions only; deletions edit every row in one launch)
    std::unordered_map<uint32_t, uint32_t> next;
    next.reserve(op == KC_INSERT ? keys.size() * 2 : 0);
    std::vector<std::vector<uint32_t>> mu, mv;
    if (op == KC_DELETE && !keys.empty()) {
        mu.emplace_back();
        mv.emplace_back();
        mu[0].reserve(keys.size());
        mv[0].reserve(keys.size());
        for (uint64_t k : keys) {
            mu[0].push_back((uint32_t)(k >> 32));
            mv[0].push_back((uint32_t)k);
        }
    }
    for (uint64_t k : op == KC_INSERT ? keys : std::vector<uint64_t>{}) {
        uint32_t a = (uint32_t)(k >> 32), b = (uint32_t)k;
        auto ia = next.find(a), ib = next.find(b);
        uint32_t m = std::max(ia == next.end() ? 0u : ia->second,
                              ib == next.end() ? 0u : ib->second);
        if (m >= mu.size()) {
            mu.emplace_back();
            mv.emplace_back();
        }
        mu[m].push_back(a);
        mv[m].push_back(b);
        next[a] = m + 1;
        next[b] = m + 1;
    }
    us.matchings = (uint32_t)mu.size();
    size_t biggest = 0;
    for (auto &x : mu)
        biggest = std::max(biggest, x.size());
    if (biggest > d->ecap) {
        hipFree(d->eu);
        hipFree(d->ev);
        hipFree(d->status);
        d->eu = d->ev = nullptr;
        d->status = nullptr;
        d->ecap = 0;
        kc_status s;
        if ((s = dmalloc(d, (void **)&d->eu, biggest * 4, "edge buffer")) ||
            (s = dmalloc(d, (void **)&d->ev, biggest * 4, "edge buffer")) ||
            (s = dmalloc(d, (void **)&d->status, biggest, "edge status")))
            return s;
        d->ecap = (uint32_t)biggest;
    }
    KC_HIP(hipEventRecord(g->ev0, st), "hipEventRecord");
    if (V)
        KC_HIP(hipMemcpyAsync(d->core_prev, d->core, (size_t)V * 4, hipMemcpyDeviceToDevice, st),
               "core copy");
    KC_HIP(hipMemsetAsync(d->cnt, 0, 16, st), "counter clear");
    std::vector<uint8_t> stat;
    kc_status s;
    if (op == KC_INSERT) {
        for (size_t m = 0; m < mu.size(); m++) {
            uint32_t k = (uint32_t)mu[m].size();
            KC_HIP(hipMemcpyAsync(d->eu, mu[m].data(), k * 4ull, hipMemcpyHostToDevice, st),
                   "edges");
            KC_HIP(hipMemcpyAsync(d->ev, mv[m].data(), k * 4ull, hipMemcpyHostToDevice, st),
                   "edges");
            if ((s = edit_matching(g, d, 0, k, mu[m], mv[m], stat, us, st)))
                return s;
            // candidate region: search from the roots, then raise it by one
            uint32_t vs = next_stamp(d, st);
            KC_HIP(hipMemsetAsync(d->cnt, 0, 8, st), "counter clear");
            KC_HIP(kc::dyn_roots(view(d), d->eu, d->ev, k, d->status, d->visit, vs, d->list[0],
                                 d->cnt, V, st),
                   "roots");
            uint32_t nf = 0;
            if ((s = read_u32(d->cnt, nf, st)))
                return s;
            uint32_t *in = d->list[0], *out = d->list[1];
            while (nf) {
                KC_HIP(hipMemsetAsync(d->cnt, 0, 4, st), "counter clear");
                KC_HIP(kc::dyn_bfs(view(d), in, nf, out, d->cnt, d->visit, vs, d->list[2],
                                   d->cnt + 1, V, st),
                       "candidate search");
                us.bfs_levels++;
                if ((s = read_u32(d->cnt, nf, st)))
                    return s;
                std::swap(in, out);
            }
            uint32_t nc = 0;
            if ((s = read_u32(d->cnt + 1, nc, st)))
                return s;
            us.candidates += nc;
            KC_HIP(kc::dyn_raise(d->core, d->list[2], nc, st), "raise");
            // lower the raised region to the fixed point
            KC_HIP(hipMemcpyAsync(d->list[0], d->list[2], (size_t)nc * 4,
                                  hipMemcpyDeviceToDevice, st),
                   "worklist");
            if ((s = lower(d, nc, us, st)))
                return s;
        }
    } else if (!mu.empty()) {
        // every deletion in one launch (rows grouped by the vertex that owns
        // them), then one lowering pass from the roots of all of them (the
        // old cores are an upper bound of the new ones)
        const uint32_t k = (uint32_t)mu[0].size();
        std::vector<std::pair<uint64_t, uint32_t>> half;
        half.reserve(2ull * k);
        for (uint32_t i = 0; i < k; i++) {
            half.push_back({(uint64_t)mu[0][i] << 32 | mv[0][i], i});
            half.push_back({(uint64_t)mv[0][i] << 32 | mu[0][i], i});
        }
        std::sort(half.begin(), half.end());
        std::vector<uint32_t> hv, hs, hn(half.size()), he(half.size());
        for (size_t j = 0; j < half.size(); j++) {
            uint32_t x = (uint32_t)(half[j].first >> 32);
            if (hv.empty() || hv.back() != x) {
                hv.push_back(x);
                hs.push_back((uint32_t)j);
            }
            hn[j] = (uint32_t)half[j].first;
            he[j] = half[j].second;
        }
        hs.push_back((uint32_t)half.size());
        if (half.size() + 1 > d->hcap) {
            for (uint32_t **p : {&d->hv, &d->hs, &d->hn, &d->he}) {
                hipFree(*p);
                *p = nullptr;
            }
            d->hcap = 0;
            uint64_t c = half.size() + 1;
            if ((s = dmalloc(d, (void **)&d->hv, c * 4, "deletion rows")) ||
                (s = dmalloc(d, (void **)&d->hs, c * 4, "deletion rows")) ||
                (s = dmalloc(d, (void **)&d->hn, c * 4, "deletion rows")) ||
                (s = dmalloc(d, (void **)&d->he, c * 4, "deletion rows")))
                return s;
            d->hcap = c;
        }
        const uint32_t nseg = (uint32_t)hv.size();
        KC_HIP(hipMemcpyAsync(d->eu, mu[0].data(), k * 4ull, hipMemcpyHostToDevice, st), "edges");
        KC_HIP(hipMemcpyAsync(d->ev, mv[0].data(), k * 4ull, hipMemcpyHostToDevice, st), "edges");
        KC_HIP(hipMemcpyAsync(d->hv, hv.data(), nseg * 4ull, hipMemcpyHostToDevice, st), "rows");
        KC_HIP(hipMemcpyAsync(d->hs, hs.data(), (nseg + 1ull) * 4, hipMemcpyHostToDevice, st),
               "rows");
        KC_HIP(hipMemcpyAsync(d->hn, hn.data(), hn.size() * 4, hipMemcpyHostToDevice, st),
               "rows");
        KC_HIP(hipMemcpyAsync(d->he, he.data(), he.size() * 4, hipMemcpyHostToDevice, st),
               "rows");
        KC_HIP(hipMemsetAsync(d->status, 0, k, st), "status clear");
        KC_HIP(kc::dyn_delete_rows(view(d), d->hv, d->hs, nseg, d->hn, d->he, d->status, st),
               "edge deletion");
        uint32_t rs = next_stamp(d, st);
        KC_HIP(hipMemsetAsync(d->cnt, 0, 4, st), "counter clear");
        KC_HIP(kc::dyn_roots(view(d), d->eu, d->ev, k, d->status, d->mark, rs, d->list[0],
                             d->cnt, V, st),
               "roots");
        stat.resize(k);
        KC_HIP(hipMemcpyAsync(stat.data(), d->status, k, hipMemcpyDeviceToHost, st),
               "status copy");
        uint32_t nr = 0;
        if ((s = read_u32(d->cnt, nr, st)))
            return s;
        uint64_t applied = 0;
        for (uint8_t x : stat)
            applied += x == 1;
        us.applied += applied;
        g->nnz -= 2 * applied;
        if ((s = lower(d, nr, us, st)))
            return s;
    }
    KC_HIP(hipMemsetAsync(d->cnt + 3, 0, 4, st), "counter clear");
    KC_HIP(kc::dyn_changed(d->core, d->core_prev, V, d->cnt + 3, st), "changed count");
    KC_HIP(hipEventRecord(g->ev1, st), "hipEventRecord");
    uint32_t changed = 0;
    if ((s = read_u32(d->cnt + 3, changed, st)))
        return s;
    float ms = 0;
    hipEventElapsedTime(&ms, g->ev0, g->ev1);
    us.changed = changed;
    us.skipped = n - us.applied;
    g->csr_stale = true;
    g->result = d->core;
    if (core_out && V)
        KC_HIP(hipMemcpy(core_out, d->core, (size_t)V * 4, hipMemcpyDeviceToHost), "core copy");
    us.device_ms = ms;
    us.total_ms = now_ms() - t0;
    if (stats)
        *stats = us;
    return KC_OK;
}

extern "C" kc_status kc_update_prepare(kc_graph *g)
{
    if (!g)
        return fail(KC_EINVAL, "null graph");
    if (!g->owns_csr)
        return fail(KC_EINVAL, "graph is a view into a batch");
    if (g->dyn)
        return KC_OK;
    if (g->csr_stale)
        return fail(KC_EDEVICE, "stale CSR without dynamic state");
    KC_HIP(hipSetDevice(g->ctx->device), "hipSetDevice");
    return dyn_create(g);
}

extern "C" kc_status kc_graph_dims(const kc_graph *g, uint32_t *V, uint64_t *nnz)
{
    if (!g)
        return fail(KC_EINVAL, "null graph");
    if (V)
        *V = g->V;
    if (nnz)
        *nnz = g->nnz;
    return KC_OK;
}
