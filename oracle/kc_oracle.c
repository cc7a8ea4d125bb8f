/*
 * kc_oracle.c -- CPU restatement of the PICO k-core methods.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
 * product path.
 *
 * Each function cites the reference file:line it restates.  The reference
 * kernels work on a degree-prefixed edge list (virtual_graph.cpp:93-157)
 * where vertex v's slot sits at nodePointer[v] = sum_{w<v}(deg(w)+1); with a
 * standard CSR that offset is rowptr[v] + v, which is what the histogram
 * layout below uses.
 */
#include "kc_oracle.h"

#include <stdlib.h>
#include <string.h>

static uint32_t deg_of(const uint64_t *rowptr, uint32_t v)
{
    return (uint32_t)(rowptr[v + 1] - rowptr[v]);
}

static void *xcalloc(size_t n, size_t sz)
{
    void *p = calloc(n ? n : 1, sz);
    if (!p)
        abort();
    return p;
}

/* ------------------------------------------------------------------------ */
/* Batagelj & Zaversnik 2003, Algorithm 1 (bin sort by degree, then peel in
 * order, moving each decremented neighbour one bin down). */
uint32_t orc_bz(uint32_t V, const uint64_t *rowptr, const uint32_t *col,
                uint32_t *core)
{
    uint32_t md = 0, kmax = 0;
    for (uint32_t v = 0; v < V; v++) {
        core[v] = deg_of(rowptr, v);
        if (core[v] > md)
            md = core[v];
    }
    uint64_t *bin = (uint64_t *)xcalloc((size_t)md + 2, sizeof(uint64_t));
    uint64_t *pos = (uint64_t *)xcalloc(V, sizeof(uint64_t));
    uint32_t *vert = (uint32_t *)xcalloc(V, sizeof(uint32_t));
    for (uint32_t v = 0; v < V; v++)
        bin[core[v]]++;
    uint64_t start = 0;
    for (uint32_t d = 0; d <= md; d++) {
        uint64_t num = bin[d];
        bin[d] = start;
        start += num;
    }
    for (uint32_t v = 0; v < V; v++) {
        pos[v] = bin[core[v]];
        vert[pos[v]] = v;
        bin[core[v]]++;
    }
    for (uint32_t d = md; d >= 1; d--)
        bin[d] = bin[d - 1];
    bin[0] = 0;
    for (uint64_t i = 0; i < V; i++) {
        uint32_t v = vert[i];
        if (core[v] > kmax)
            kmax = core[v];
        for (uint64_t e = rowptr[v]; e < rowptr[v + 1]; e++) {
            uint32_t u = col[e];
            if (core[u] > core[v]) {
                uint32_t du = core[u];
                uint64_t pu = pos[u];
                uint64_t pw = bin[du];
                uint32_t w = vert[pw];
                if (u != w) {
                    pos[u] = pw;
                    vert[pu] = w;
                    pos[w] = pu;
                    vert[pw] = u;
                }
                bin[du]++;
                core[u]--;
            }
        }
    }
    free(bin);
    free(pos);
    free(vert);
    return kmax;
}

/* ------------------------------------------------------------------------ */
/* gppeel.cu:186-228 host loop with scan_nodes (:12-43) and scatter_edges
 * (:45-104).  `core` below is the reference's d_core (live degree),
 * `fin` its d_final_core. */
void orc_gpp(uint32_t V, const uint64_t *rowptr, const uint32_t *col,
             int compat, uint32_t *out, uint64_t *iterations)
{
    uint32_t *cur = (uint32_t *)xcalloc(V, sizeof(uint32_t));
    uint8_t *deleted = (uint8_t *)xcalloc(V, 1);
    uint8_t *scatter = (uint8_t *)xcalloc(V, 1);
    for (uint32_t v = 0; v < V; v++) {
        cur[v] = deg_of(rowptr, v);
        out[v] = 0;
    }
    uint64_t itr = 0;
    uint32_t level = 0;
    int plus = 1, finished = 0;
    while (!finished) {
        itr++;
        if (plus)
            level++;
        else
            plus = 1;
        finished = 1;
        memset(scatter, 0, V);
        for (uint32_t v = 0; v < V; v++) { /* scan_nodes */
            if (deleted[v])
                continue;
            finished = 0;
            if (cur[v] <= level) {
                deleted[v] = 1;
                scatter[v] = 1;
                out[v] = level;
                plus = 0;
            }
        }
        for (uint32_t v = 0; v < V; v++) { /* scatter_edges */
            if (!scatter[v])
                continue;
            for (uint64_t e = rowptr[v]; e < rowptr[v + 1]; e++)
                if (!deleted[col[e]])
                    cur[col[e]]--;
        }
    }
    if (!compat)
        for (uint32_t v = 0; v < V; v++)
            if (rowptr[v + 1] == rowptr[v])
                out[v] = 0;
    *iterations = itr;
    free(cur);
    free(deleted);
    free(scatter);
}

/* ------------------------------------------------------------------------ */
/* peelone.cu:14-91 kernel and :154-186 host loop.  Parts are the virtual
 * nodes of virtual_graph.cpp:161-183 (ceil(deg/part_size) per vertex, in
 * vertex order); part p of a vertex with np parts walks neighbours
 * p, p+np, p+2np, ... (peelone.cu:331-347). */
void orc_peelone(uint32_t V, const uint64_t *rowptr, const uint32_t *col,
                 uint32_t part_size, uint32_t *dist, uint64_t *iterations)
{
    uint64_t P = 0;
    for (uint32_t v = 0; v < V; v++) {
        uint32_t d = deg_of(rowptr, v);
        P += (d + part_size - 1) / part_size;
        dist[v] = d;
    }
    uint8_t *sighed = (uint8_t *)xcalloc(P, 1);
    uint64_t itr = 0;
    uint32_t level = 0;
    int plus = 1, finished = 0;
    while (!finished) {
        itr++;
        if (plus)
            level++;
        else
            plus = 1;
        finished = 1;
        uint64_t p = 0;
        for (uint32_t v = 0; v < V; v++) {
            uint32_t d = deg_of(rowptr, v);
            uint32_t np = (d + part_size - 1) / part_size;
            for (uint32_t part = 0; part < np; part++, p++) {
                if (sighed[p])
                    continue;
                finished = 0;
                if (dist[v] != level)
                    continue;
                plus = 0;
                sighed[p] = 1;
                for (uint32_t i = 0; i < part_size; i++) {
                    uint64_t j = part + (uint64_t)i * np;
                    if (j >= d)
                        break;
                    uint32_t u = col[rowptr[v] + j];
                    if (dist[u] > level) {
                        uint32_t old = dist[u]--;
                        if (old == level) /* clamp, peelone.cu:342-345 */
                            dist[u] = level;
                    }
                }
            }
        }
    }
    *iterations = itr;
    free(sighed);
}

/* ------------------------------------------------------------------------ */
/* peel-dyn.cc:133-182 (peelone-dyn.cc:120-168 is the same peel with the
 * restore and enqueue statements swapped).  selectNodesAtLevel (:2-44)
 * gathers deg==level; processNodes (:47-131) drains the queue, decrementing
 * neighbours with deg>level and enqueueing those that land on level.  Run
 * sequentially no decrement ever undershoots, so the restore branch
 * (a <= level) never fires here. */
uint32_t orc_bucket_peel(uint32_t V, const uint64_t *rowptr,
                         const uint32_t *col, uint32_t *deg,
                         uint64_t *levels)
{
    uint32_t *queue = (uint32_t *)xcalloc(V, sizeof(uint32_t));
    for (uint32_t v = 0; v < V; v++)
        deg[v] = deg_of(rowptr, v);
    uint64_t count = 0;
    uint32_t level = 0;
    uint64_t trips = 0;
    while (count < V) {
        uint64_t tail = 0;
        for (uint32_t v = 0; v < V; v++)
            if (deg[v] == level)
                queue[tail++] = v;
        for (uint64_t base = 0; base < tail; base++) {
            uint32_t v = queue[base];
            for (uint64_t e = rowptr[v]; e < rowptr[v + 1]; e++) {
                uint32_t u = col[e];
                if (deg[u] > level) {
                    uint32_t a = deg[u]--;
                    if (a == level + 1)
                        queue[tail++] = u;
                }
            }
        }
        count += tail;
        level++;
        trips++;
    }
    free(queue);
    *levels = trips;
    return V ? level - 1 : 0;
}

/* ------------------------------------------------------------------------ */
/* HistoCore.cu: histo_kernel :15-75, count_kernel :77-128,
 * update_histo_kernel :130-202, host loop :242-341 (initial frontier
 * core>1 at :246-257). */
void orc_histocore(uint32_t V, const uint64_t *rowptr, const uint32_t *col,
                   uint32_t *core, uint64_t *iterations, uint32_t *hist_out)
{
    uint64_t H = rowptr[V] + V;
    uint32_t *histo = (uint32_t *)xcalloc(H, sizeof(uint32_t));
    uint8_t *sighed = (uint8_t *)xcalloc(V, 1);
    uint8_t *nbr_sighed = (uint8_t *)xcalloc(V, 1);
    for (uint32_t v = 0; v < V; v++) {
        core[v] = deg_of(rowptr, v);
        sighed[v] = core[v] > 1;
    }
    for (uint32_t v = 0; v < V; v++) { /* histo_kernel */
        uint32_t d = deg_of(rowptr, v);
        uint64_t base = rowptr[v] + v;
        if (d == 0)
            continue;
        histo[base] = d;
        for (uint64_t e = rowptr[v]; e < rowptr[v + 1]; e++) {
            uint32_t c = core[col[e]];
            histo[base + (c < d ? c : d)]++;
        }
    }
    uint64_t itr = 0;
    for (;;) {
        itr++;
        int finished = 1;
        for (uint32_t v = 0; v < V; v++) { /* count_kernel */
            if (!sighed[v])
                continue;
            uint64_t base = rowptr[v] + v;
            uint32_t old = core[v], sum = 0, k;
            for (k = old; k >= 1; k--) {
                sum += histo[base + k];
                if (sum >= k)
                    break;
            }
            if (k == old)
                continue;
            finished = 0;
            histo[base] = old;
            histo[base + k] = sum;
            core[v] = k;
            nbr_sighed[v] = 1;
        }
        memset(sighed, 0, V);
        if (finished)
            break;
        for (uint32_t v = 0; v < V; v++) { /* update_histo_kernel */
            if (!nbr_sighed[v])
                continue;
            uint64_t base = rowptr[v] + v;
            uint32_t nc = core[v], oc = histo[base];
            for (uint64_t e = rowptr[v]; e < rowptr[v + 1]; e++) {
                uint32_t u = col[e];
                uint32_t cu = core[u];
                if (nc < cu) {
                    uint64_t bu = rowptr[u] + u;
                    uint32_t sub = cu < oc ? cu : oc;
                    uint32_t cnt = histo[bu + sub]--;
                    histo[bu + nc]++;
                    if (oc >= cu && cnt == cu)
                        sighed[u] = 1;
                }
            }
        }
        memset(nbr_sighed, 0, V);
    }
    *iterations = itr;
    if (hist_out)
        memcpy(hist_out, histo, H * sizeof(uint32_t));
    free(histo);
    free(sighed);
    free(nbr_sighed);
}

/* ------------------------------------------------------------------------ */
/* CntCore.cu: core_kernel :17-80, count_kernel :82-130, nbr_kernel
 * :132-195, cnt_kernel :197-264, host loop :344-417. */
void orc_cntcore(uint32_t V, const uint64_t *rowptr, const uint32_t *col,
                 uint32_t *core, uint64_t *iterations)
{
    uint64_t H = rowptr[V] + V;
    uint32_t *histo = (uint32_t *)xcalloc(H, sizeof(uint32_t));
    uint32_t *cnt = (uint32_t *)xcalloc(V, sizeof(uint32_t));
    uint8_t *sighed = (uint8_t *)xcalloc(V, 1);
    uint8_t *sighed_nbr = (uint8_t *)xcalloc(V, 1);
    uint8_t *active = (uint8_t *)xcalloc(V, 1);
    for (uint32_t v = 0; v < V; v++)
        core[v] = deg_of(rowptr, v);
    uint64_t itr = 0;
    for (;;) {
        itr++;
        int finished = 1;
        memset(sighed, 0, V);
        memset(sighed_nbr, 0, V);
        memset(histo, 0, H * sizeof(uint32_t));
        for (uint32_t v = 0; v < V; v++) { /* core_kernel */
            active[v] = cnt[v] < core[v];
            if (!active[v])
                continue;
            uint32_t d = deg_of(rowptr, v);
            uint64_t base = rowptr[v] + v;
            for (uint64_t e = rowptr[v]; e < rowptr[v + 1]; e++) {
                uint32_t c = core[col[e]];
                histo[base + (c < d ? c : d)]++;
            }
        }
        for (uint32_t v = 0; v < V; v++) { /* count_kernel */
            if (!active[v])
                continue;
            uint64_t base = rowptr[v] + v;
            uint32_t sum = 0, k;
            for (k = deg_of(rowptr, v); k >= 1; k--) {
                sum += histo[base + k];
                if (sum >= k)
                    break;
            }
            if (k == core[v])
                continue;
            finished = 0;
            histo[base] = core[v];
            core[v] = k;
            sighed[v] = 1;
        }
        if (finished)
            break;
        for (uint32_t v = 0; v < V; v++) { /* nbr_kernel */
            if (!sighed[v])
                continue;
            for (uint64_t e = rowptr[v]; e < rowptr[v + 1]; e++) {
                cnt[col[e]] = 0;
                sighed_nbr[col[e]] = 1;
            }
        }
        for (uint32_t v = 0; v < V; v++) { /* cnt_kernel */
            if (!sighed_nbr[v])
                continue;
            for (uint64_t e = rowptr[v]; e < rowptr[v + 1]; e++)
                if (core[col[e]] >= core[v])
                    cnt[v]++;
        }
    }
    *iterations = itr;
    free(histo);
    free(cnt);
    free(sighed);
    free(sighed_nbr);
    free(active);
}

/* ------------------------------------------------------------------------ */
/* NbrCore.cu: core_kernel :12-76, count_kernel :78-113, nbr_kernel
 * :115-167, host loop :244-297 (frontier starts as all vertices, :237). */
void orc_nbrcore(uint32_t V, const uint64_t *rowptr, const uint32_t *col,
                 uint32_t *core, uint64_t *iterations)
{
    uint64_t H = rowptr[V] + V;
    uint32_t *histo = (uint32_t *)xcalloc(H, sizeof(uint32_t));
    uint8_t *sighed = (uint8_t *)xcalloc(V, 1);
    uint8_t *nbr_sighed = (uint8_t *)xcalloc(V, 1);
    for (uint32_t v = 0; v < V; v++) {
        core[v] = deg_of(rowptr, v);
        sighed[v] = 1;
    }
    uint64_t itr = 0;
    for (;;) {
        itr++;
        int finished = 1;
        memset(histo, 0, H * sizeof(uint32_t));
        for (uint32_t v = 0; v < V; v++) { /* core_kernel */
            if (!sighed[v])
                continue;
            uint32_t d = deg_of(rowptr, v);
            uint64_t base = rowptr[v] + v;
            if (d == 0)
                continue;
            histo[base] = d;
            for (uint64_t e = rowptr[v]; e < rowptr[v + 1]; e++) {
                uint32_t c = core[col[e]];
                histo[base + (c < d ? c : d)]++;
            }
        }
        for (uint32_t v = 0; v < V; v++) { /* count_kernel */
            if (!sighed[v])
                continue;
            uint64_t base = rowptr[v] + v;
            uint32_t sum = 0, k;
            for (k = deg_of(rowptr, v); k >= 1; k--) {
                sum += histo[base + k];
                if (sum >= k)
                    break;
            }
            if (k == core[v])
                continue;
            finished = 0;
            histo[base + k] = sum;
            core[v] = k;
            nbr_sighed[v] = 1;
        }
        memset(sighed, 0, V);
        if (finished)
            break;
        for (uint32_t v = 0; v < V; v++) { /* nbr_kernel */
            if (!nbr_sighed[v])
                continue;
            for (uint64_t e = rowptr[v]; e < rowptr[v + 1]; e++)
                sighed[col[e]] = 1;
        }
        memset(nbr_sighed, 0, V);
    }
    *iterations = itr;
    free(histo);
    free(sighed);
    free(nbr_sighed);
}

/* ------------------------------------------------------------------------ */
/* h-index of the multiset {min(cap, c(u)) : u in N(v)}: the largest k with
 * at least k values >= k. */
static uint32_t hindex_capped(const uint64_t *rowptr, const uint32_t *col,
                              const uint32_t *c, uint32_t v, uint32_t cap,
                              uint32_t *scratch)
{
    if (cap == 0)
        return 0;
    memset(scratch, 0, ((size_t)cap + 1) * sizeof(uint32_t));
    for (uint64_t e = rowptr[v]; e < rowptr[v + 1]; e++) {
        uint32_t x = c[col[e]];
        scratch[x < cap ? x : cap]++;
    }
    uint32_t sum = 0, k;
    for (k = cap; k >= 1; k--) {
        sum += scratch[k];
        if (sum >= k)
            break;
    }
    return k;
}

void orc_hindex_jacobi(uint32_t V, const uint64_t *rowptr, const uint32_t *col,
                       uint32_t *core, uint64_t *rounds)
{
    uint32_t md = 0;
    for (uint32_t v = 0; v < V; v++) {
        core[v] = deg_of(rowptr, v);
        if (core[v] > md)
            md = core[v];
    }
    uint32_t *next = (uint32_t *)xcalloc(V, sizeof(uint32_t));
    uint32_t *scratch = (uint32_t *)xcalloc((size_t)md + 2, sizeof(uint32_t));
    uint64_t r = 0;
    for (;;) {
        r++;
        int changed = 0;
        for (uint32_t v = 0; v < V; v++) {
            next[v] = hindex_capped(rowptr, col, core, v, core[v], scratch);
            if (next[v] != core[v])
                changed = 1;
        }
        memcpy(core, next, (size_t)V * sizeof(uint32_t));
        if (!changed)
            break;
    }
    *rounds = r;
    free(next);
    free(scratch);
}

/* ------------------------------------------------------------------------ */
void orc_bruteforce(uint32_t V, const uint64_t *rowptr, const uint32_t *col,
                    uint32_t *core)
{
    uint8_t *in = (uint8_t *)xcalloc(V, 1);
    for (uint32_t v = 0; v < V; v++)
        core[v] = 0;
    for (uint32_t k = 1;; k++) {
        uint32_t alive = 0;
        for (uint32_t v = 0; v < V; v++)
            in[v] = 1;
        int changed = 1;
        while (changed) {
            changed = 0;
            for (uint32_t v = 0; v < V; v++) {
                if (!in[v])
                    continue;
                uint32_t d = 0;
                for (uint64_t e = rowptr[v]; e < rowptr[v + 1]; e++)
                    d += in[col[e]];
                if (d < k) {
                    in[v] = 0;
                    changed = 1;
                }
            }
        }
        for (uint32_t v = 0; v < V; v++)
            if (in[v]) {
                core[v] = k;
                alive++;
            }
        if (!alive)
            break;
    }
    free(in);
}

/* ------------------------------------------------------------------------ */
uint64_t orc_check_invariants(uint32_t V, const uint64_t *rowptr,
                              const uint32_t *col, const uint32_t *core)
{
    uint32_t md = 0;
    for (uint32_t v = 0; v < V; v++)
        if (deg_of(rowptr, v) > md)
            md = deg_of(rowptr, v);
    uint32_t *scratch = (uint32_t *)xcalloc((size_t)md + 2, sizeof(uint32_t));
    uint64_t bad = 0;
    for (uint32_t v = 0; v < V && !bad; v++) {
        uint32_t d = deg_of(rowptr, v), ge = 0;
        if (core[v] > d) {
            bad = (uint64_t)v + 1;
            break;
        }
        for (uint64_t e = rowptr[v]; e < rowptr[v + 1]; e++)
            ge += core[col[e]] >= core[v];
        if (ge < core[v])
            bad = (uint64_t)v + 1;
        else if (hindex_capped(rowptr, col, core, v, d, scratch) != core[v])
            bad = (uint64_t)v + 1;
    }
    free(scratch);
    return bad;
}
