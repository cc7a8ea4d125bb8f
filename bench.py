#!/usr/bin/env python3
"""k-core decomposition benchmark on one MI355X.

Workload (the headline configuration): R-MAT scale 22, edge factor 16,
(a,b,c,d) = (.57,.19,.19,.05), symmetrized, deduplicated, self-loops dropped,
int32 CSR resident in HBM.  One step = one full decomposition of that graph
by the plain peel (peelone_dyn engine); HistoCore (h-index) is timed on the
same graph and reported beside it.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Beside it, the batched workload: 4,096 R-MAT graphs (scale uniform over the
integers 8..18, edge factor uniform in [4, 32], seed base+g) decomposed in
one persistent launch, against per-graph plain launches of the same engine
on the same resident data (graphs/s both ways, every graph checked against
the oracle).

Prints one JSON line.  value = GTEPS = undirected edges / seconds per step.
roofline.achieved = algorithmic bytes per launch / mean kernel time, with
the byte model
    B_peel = 4(V+1) + 4*2m (col) + 4*2m (degree RMW) + 4V + 4V(kmax+1)
    B_h    = 4(V+1) + 12*2m + 4(2m+V) + 4V + rounds*V + 4*changed + 16*sum deg(changed)
(2m = adjacency entries).  cpu_baseline = the oracle's sequential
Batagelj-Zaversnik peel on one host core over the same graph.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--edge-factor", type=float, default=16.0)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--method", default="peelone_dyn")
    ap.add_argument("--hindex-method", default="histocore")
    ap.add_argument("--cpu-runs", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-graphs", type=int, default=4096)
    ap.add_argument("--batch-steps", type=int, default=5)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=300)
    ap.add_argument("--serve-rate", type=float, default=40.0, help="Poisson arrivals per s")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-sweep", default="160,640",
                    help="further Poisson rates (per s) of the serving stream, comma-separated")
    ap.add_argument("--no-methods", action="store_true",
                    help="skip GPP / PeelOne / CntCore / NbrCore on the headline graph")
    ap.add_argument("--no-chunglu", action="store_true")
    ap.add_argument("--no-reader", action="store_true")
    ap.add_argument("--only", choices=["peel", "batch", "serve", "methods", "chunglu",
                                       "reader"], default=None)
    ap.add_argument("--profile", default=None,
                    help="profiling run: build the graph (--graph), run the copy probe, then "
                         "decompose twice with this one method (or 'batch': the 4,096-graph "
                         "batch twice at block 1024, one workgroup per CU); prints a short line")
    ap.add_argument("--graph", choices=["rmat", "chunglu"], default="rmat")
    ap.add_argument("--pmc-key", default=None,
                    help="with --summarize-pmc: the entry of profiles/pmc_traffic.json to "
                         "write (merged into the file)")
    ap.add_argument("--pmc-kernel", default=None,
                    help="with --summarize-pmc and --pmc-key: kernel name substring")
    ap.add_argument("--summarize-trace", metavar="KT_DB", default=None,
                    help="print the kernel table of a rocprofv3 --kernel-trace database")
    ap.add_argument("--summarize-counters", nargs="+", metavar="X", default=None,
                    help="DB KERNEL COUNTER...: per-dispatch means of counters for one kernel")
    ap.add_argument("--pmc-runs", type=int, default=2,
                    help="with --summarize-pmc and --pmc-key: decompositions in the profiled "
                         "--profile run (it runs two)")
    ap.add_argument("--pmc-stages", type=int, default=1,
                    help="with --summarize-pmc: peel_kernel launches per plain-peel step "
                         "(launches_per_step of the profiled run)")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("FETCH_DB", "WRITE_DB"),
                    help="write profiles/pmc_traffic.json from two rocprofv3 --pmc "
                         "databases (FETCH_SIZE, WRITE_SIZE) of this script's runs")
    return ap.parse_args()


def peel_bytes(V, nnz, kmax, R=4):
    return R * (V + 1) + 4 * nnz + 4 * nnz + 4 * V + 4 * V * (kmax + 1)


def hindex_bytes(V, nnz, rounds, changed, changed_edges, R=4):
    init = R * (V + 1) + 12 * nnz + 4 * (nnz + V) + 4 * V
    return init + rounds * V + R * changed + 16 * changed_edges


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def traffic_from_profile(kernel_key):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE, scaled against a copy kernel of known size); see
    profiles/pmc_traffic.json.  None when not measured."""
    p = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(p) as f:
            e = json.load(f).get(kernel_key, {})
            return e.get("bytes_per_step", e.get("bytes_per_launch"))
    except (OSError, ValueError):
        return None


def trimmed_mean(xs):
    """Mean after dropping the minimum and the maximum (of at least 3 values)."""
    xs = sorted(xs)
    if len(xs) >= 3:
        xs = xs[1:-1]
    return sum(xs) / len(xs)


def _pct(xs, q):
    import numpy as np
    return round(float(np.percentile(np.asarray(xs), q)), 3) if len(xs) else None


def bench_serving(args, rate=None, n=None):
    """Poisson request stream against kc_server: fresh R-MAT decompositions
    (scale uniform over 10..20, edge factor 16) and 10,000-edge update batches
    (uniform random insertions / deletions of present edges) on a resident
    R-MAT s20 ef16 graph.  Latency = completion - arrival (submission happens
    at the arrival instant).  Every result is then re-served alone and
    compared; the resident graph's final cores are checked against the
    oracle."""
    import threading

    import numpy as np

    import oracle
    import pico_amd as pa

    rate = float(rate or args.serve_rate)
    rng = np.random.default_rng(args.seed + 77)
    n = int(n or args.serve_requests)
    base_scale, base_ef, base_seed = 20, 16.0, args.seed + 5
    # the workload, fixed up front: kinds, arrival times, parameters
    srv = pa.Server(slots=args.serve_slots, fuse_max_scale=16)
    gid = srv.gen_rmat_graph(base_scale, base_ef, seed=base_seed)
    base = srv.download_graph(gid)
    rp, col = base.rowptr, base.col
    src = np.repeat(np.arange(base.V, dtype=np.uint32), np.diff(rp).astype(np.int64))
    up = np.flatnonzero(src < col)
    del_pool = rng.permutation(up)
    reqs, dp = [], 0
    for i in range(n):
        if rng.random() < 0.5:
            reqs.append(("rmat", int(rng.integers(10, 21)), 16.0, 10_000 + i))
        elif rng.random() < 0.5:
            u = rng.integers(0, base.V, 10_000, dtype=np.uint32)
            v = rng.integers(0, base.V, 10_000, dtype=np.uint32)
            reqs.append(("update", "insert", u, v))
        else:
            pick = del_pool[dp: dp + 10_000]
            dp += 10_000
            reqs.append(("update", "delete", src[pick].copy(), col[pick].copy()))
    gaps = rng.exponential(1.0 / rate, n)
    arrive = np.cumsum(gaps) - gaps[0]
    done = {}

    def collect(total):
        while len(done) < total:
            for c in srv.poll(wait_ms=200):
                done[c["tag"]] = c

    th = threading.Thread(target=collect, args=(n,))
    th.start()
    t0 = time.perf_counter()
    for k, r in enumerate(reqs):
        dt = arrive[k] - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        if r[0] == "rmat":
            srv.submit_rmat(r[1], r[2], seed=r[3], tag=k)
        else:
            srv.submit_update(gid, r[1], r[2], r[3], tag=k)
    th.join()
    wall = time.perf_counter() - t0
    st = srv.stats()
    final = srv.download_graph(gid)
    srv.close()
    lat = {k: c["done_ms"] - c["submit_ms"] for k, c in done.items()}
    dec = [lat[k] for k, r in enumerate(reqs) if r[0] == "rmat"]
    upd = [lat[k] for k, r in enumerate(reqs) if r[0] != "rmat"]
    first = min(c["submit_ms"] for c in done.values())
    last = max(c["done_ms"] for c in done.values())
    # every result again, alone (one slot, one request at a time)
    alone = pa.Server(slots=1, max_batch=1)
    gid2 = alone.gen_rmat_graph(base_scale, base_ef, seed=base_seed)
    same = 0
    for k, r in enumerate(reqs):
        if r[0] == "rmat":
            alone.submit_rmat(r[1], r[2], seed=r[3], tag=k)
        else:
            alone.submit_update(gid2, r[1], r[2], r[3], tag=k)
        c = alone.poll(wait_ms=-1)[0]
        same += int(c["checksum"] == done[k]["checksum"] and c["status"] == 0
                    and done[k]["status"] == 0)
    alone.close()
    last_upd = max((k for k, r in enumerate(reqs) if r[0] != "rmat"), default=None)
    final_ok = None
    if last_upd is not None:
        want = oracle.bz(final.rowptr, final.col)
        final_ok = bool(done[last_upd]["checksum"] == pa.core_checksum(want))
    return {
        "workload": "Poisson arrivals at %.1f/s: 50%% fresh R-MAT decompositions (scale uniform "
                    "10..20, edge factor 16), 50%% 10,000-edge insert/delete batches on a "
                    "resident R-MAT s20 ef16 graph" % rate,
        "requests": n, "rate_per_s": rate, "slots": args.serve_slots,
        "requests_per_s": round(n / ((last - first) * 1e-3), 2),
        "p50_ms": _pct(list(lat.values()), 50), "p99_ms": _pct(list(lat.values()), 99),
        "decompose_p50_ms": _pct(dec, 50), "decompose_p99_ms": _pct(dec, 99),
        "update_p50_ms": _pct(upd, 50), "update_p99_ms": _pct(upd, 99),
        "max_ms": round(max(lat.values()), 3), "wall_s": round(wall, 2),
        "fused_jobs": st["fused_jobs"], "fused_requests": st["fused_requests"],
        "max_busy_slots": st["max_busy"],
        "max_inflight_gb": round(st["max_inflight_bytes"] / 1e9, 3),
        "same_alone": "%d/%d" % (same, n), "final_cores_vs_oracle": final_ok,
    }


def _timed(dg, method, reps, **opts):
    """Warm run, then `reps` timed runs: (cores of the last, stats, mean device ms)."""
    import numpy as np
    dg.decompose(method, copy=False, **opts)
    ms = []
    for _ in range(reps):
        _, st = dg.decompose(method, copy=False, **opts)
        ms.append(st["device_ms"])
    core, st = dg.decompose(method, **opts)
    return core, st, float(np.mean(ms))


def _roof(b_alg, kernel_ms, traffic=None):
    a = b_alg / (kernel_ms * 1e-3) / 1e9
    out = {"bound": "hbm", "achieved": round(a, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(a / HBM_PEAK_GBS, 5), "algorithmic_bytes": int(b_alg),
           "traffic": traffic}
    if traffic:
        am = min(b_alg, traffic) / (kernel_ms * 1e-3) / 1e9
        out["frac_min_rule"] = round(am / HBM_PEAK_GBS, 5)
    return out


PEEL_METHODS = ("gpp", "peelone", "peel_dyn", "peelone_dyn")


def method_line(dg, method, V, nnz, m, R, reps, want=None, traffic_key=None, chg_deg=None,
                **opts):
    """One method on a resident graph: kernel time, GTEPS, and the roofline of
    its algorithmic bytes (the peel model for the peel family, the h-index
    model for the h-index family, both from the same cores / rounds)."""
    core, st, kms = _timed(dg, method, reps, **opts)
    if method in PEEL_METHODS:
        b = peel_bytes(V, nnz, st["kmax"], R)
        kern = "peel_kernel"
    else:
        # the three h-index methods run the same Jacobi sequence (same rounds,
        # same changed vertices); HistoCore's kernel sums the changed degrees
        cd = st["changed_degree"] if method == "histocore" or chg_deg is None else chg_deg
        b = hindex_bytes(V, nnz, st["iterations"], st["changed"], cd, R)
        kern = "hindex_kernel"
    line = {"kernel_ms": round(kms, 4), "gteps": round(m / (kms * 1e-3) / 1e9, 4),
            "iterations": st["iterations"], "kmax": st["kmax"], "stages": st["stages"],
            "grid": st["grid"], "block": st["block"], "kernel": kern,
            "roofline": _roof(b, kms, traffic_from_profile(traffic_key) if traffic_key
                              else None)}
    if method not in PEEL_METHODS:
        line["rebuild_rounds"] = st.get("rebuild_rounds")
        line["atomics"] = st["atomics"]
        line["changed"] = st["changed"]
    if want is not None:
        import numpy as np
        exp = want.copy()
        if method == "gpp" and opts.get("gpp_compat"):
            exp[exp == 0] = 1
        line["parity_vs_oracle"] = bool(np.array_equal(core, exp))
    return line


def bench_methods(dg, V, nnz, m, want, chg_deg, reps=3):
    """GPP, PeelOne, peel_dyn, CntCore and NbrCore on the headline graph (the
    bucket peel and HistoCore have lines of their own).  peel_dyn runs the
    same kernel instance and stages as peelone_dyn, whose counters it shares."""
    return {mt: method_line(dg, mt, V, nnz, m, 4, reps, want,
                            traffic_key="peel" if mt == "peel_dyn" else mt, chg_deg=chg_deg)
            for mt in ("gpp", "peelone", "peel_dyn", "cntcore", "nbrcore")}


def bench_chunglu(ctx, args, check=True):
    """The heavy-skew Chung-Lu graph (n = 2^24, exponent 2.1, mean degree 20,
    cap 4M, seed 3) with u64 row offsets: the bucket peel and HistoCore."""
    import numpy as np

    import oracle
    t0 = time.time()
    dg = ctx.gen_chunglu(1 << 24, 2.1, 20.0, 4e6, seed=3, rowptr_bytes=8)
    gen_s = time.time() - t0
    want = None
    g = dg.download()
    V, nnz = g.V, g.nnz
    deg = np.diff(g.rowptr)
    if check:
        want = oracle.bz(g.rowptr, g.col)
    out = {"workload": "Chung-Lu n=2^24, exponent 2.1, mean degree 20, max expected degree "
                       "4M, seed 3; int32 ids, int64 row offsets, built in HBM",
           "V": V, "undirected_edges": nnz // 2, "dmax": int(deg.max()),
           "hubs_over_1e5": int((deg > 100000).sum()), "graph_gen_s": round(gen_s, 2)}
    for mt in ("peelone_dyn", "histocore"):
        out[mt] = method_line(dg, mt, V, nnz, nnz // 2, 8, max(2, args.steps // 3), want,
                              traffic_key="chunglu_" + mt)
    dg.close()
    return out


def bench_reader(ctx, g, args):
    """The R-MAT s22 graph written as an edge file (example.txt dialect, one
    line per undirected edge, 1-based) and read back straight into HBM by
    kc_graph_read with both reference readers: wall time from the file (in the
    page cache) to a resident graph, and equality with the generator's graph."""
    import tempfile

    import numpy as np

    import pico_amd as pa
    d = os.environ.get("TMPDIR") or tempfile.gettempdir()
    path = os.path.join(d, "kc_bench_rmat%d_%d.txt" % (args.scale, os.getpid()))
    out = {}
    try:
        t = time.perf_counter()
        g.write_edgelist(path)
        out["write_s"] = round(time.perf_counter() - t, 3)
        out["file_bytes"] = os.path.getsize(path)
        for reader in ("dedup", "coo"):
            best = None
            for _ in range(2):
                rg, st = ctx.read(path, reader)
                if best is None or st["total_ms"] < best[1]["total_ms"]:
                    if best:
                        best[0].close()
                    best = (rg, st)
                else:
                    rg.close()
            rg, st = best
            back = rg.download()
            rg.close()
            if reader == "dedup":
                # one line per undirected edge: the generator's graph back, up
                # to trailing isolated ids (num_nodes = max id)
                same = bool(np.array_equal(back.rowptr, g.rowptr[:back.V + 1]) and
                            np.array_equal(back.col, g.col[:back.nnz]))
            else:
                # the COO reader keeps file order, which for this file (rows in
                # order, each edge once from its smaller end) is sorted order
                same = bool(np.array_equal(back.rowptr, g.rowptr[:back.V + 1]) and
                            np.array_equal(back.col, g.col[:back.nnz]))
            out[reader] = {"total_ms": round(st["total_ms"], 1),
                           "parse_ms": round(st["parse_ms"], 1),
                           "build_ms": round(st["build_ms"], 1), "threads": st["threads"],
                           "V": st["V"], "nnz": st["nnz"], "matches_graph": same,
                           "gb_per_s": round(out["file_bytes"] / (st["total_ms"] * 1e-3) / 1e9,
                                             3)}
        t = time.perf_counter()
        hg = pa.HostGraph.read_dedup(path)
        out["host_dedup_reader_ms"] = round((time.perf_counter() - t) * 1e3, 1)
        hg.close()
    finally:
        try:
            os.remove(path)
        except OSError:
            pass
    out["workload"] = ("R-MAT scale %d edge factor %g seed %d written as a 1-based edge file, "
                       "read into HBM (host threads parse, device builds the CSR)"
                       % (args.scale, args.edge_factor, args.seed))
    return out


def bench_batched(ctx, args):
    """4,096 uneven R-MAT graphs: one persistent launch vs per-graph plain
    launches, both on the same HBM-resident batch."""
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np

    import oracle

    G = args.batch_graphs
    t0 = time.time()
    db = ctx.gen_rmat_batch(G, 8, 18, 4.0, 32.0, seed=args.seed)
    gen_s = time.time() - t0
    best, tune_ms = db.autotune(reps=1)
    for _ in range(max(1, args.warmup)):
        db.run(copy=False)
    wall, dev = [], []
    for _ in range(args.batch_steps):
        _, _, st = db.run(copy=False)
        wall.append(st["total_ms"])
        dev.append(st["device_ms"])
    core, km, st = db.run()
    _, _, sp = db.run_plain("peelone_dyn", copy=False)
    # every graph against the oracle's sequential peel, on the host's cores
    hb = db.download()
    vo, eo, rp, cl = hb.arrays()

    def check(gi):
        V = int(vo[gi + 1] - vo[gi])
        want = oracle.bz(rp[vo[gi] + gi: vo[gi] + gi + V + 1].astype(np.uint64),
                         cl[eo[gi]: eo[gi + 1]])
        return bool(np.array_equal(core[vo[gi]: vo[gi] + V], want)
                    and km[gi] == (int(want.max()) if V else 0))

    threads = min(16, os.cpu_count() or 1)
    t = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        ok = sum(ex.map(check, range(G)))
    cpu_s = time.perf_counter() - t
    ms = float(np.median(wall))
    dev_med = float(np.median(dev))
    Vg = np.diff(vo).astype(np.int64)
    Eg = np.diff(eo).astype(np.int64)
    kg = km.astype(np.int64)
    b_batch = int(np.sum(4 * (Vg + 1) + 8 * Eg + 4 * Vg + 4 * Vg * (kg + 1)))
    out = {
        "workload": "%d R-MAT graphs (.57,.19,.19,.05), scale uniform over the integers 8..18, "
                    "edge factor uniform in [4,32], seed %d+g; concatenated int32 CSR in HBM"
                    % (G, args.seed),
        "vertices": int(hb.vertices), "adjacency_entries": int(hb.entries),
        "graphs_per_s": round(G / (ms * 1e-3), 1),
        "ms_per_batch": round(ms, 3), "device_ms": round(float(np.median(dev)), 3),
        "launches": 1, "steps": args.batch_steps,
        "plain_graphs_per_s": round(G / (sp["total_ms"] * 1e-3), 1),
        "plain_ms": round(sp["total_ms"], 1), "plain_device_ms": round(sp["device_ms"], 1),
        "plain_launches": sp["launches"], "plain_method": "peelone_dyn",
        "speedup_vs_plain": round(sp["total_ms"] / ms, 2),
        "autotuned": best, "levels_run": st["levels_run"], "kmax": st["kmax"],
        "roofline": dict(_roof(b_batch, dev_med, traffic_from_profile("batch")),
                         kernel="batch_kernel", kernel_ms=round(dev_med, 3),
                         accounting="sum over graphs of the peel model with each graph's kmax"),
        "parity_graphs": "%d/%d" % (ok, G),
        "cpu_bz_s": round(cpu_s, 2), "cpu_bz_threads": threads,
        "gen_s": round(gen_s, 2),
    }
    db.close()
    return out


COPY_PROBE_BYTES = 1 << 30


def _pmc_dispatches(db_path, counter):
    """Values of `counter` per dispatch, by kernel name, from a rocprofv3
    rocpd database."""
    import sqlite3
    db = sqlite3.connect(db_path)
    cur = db.cursor()
    rows = cur.execute("select name, dispatch_id, sum(counter_value) from pmc_events "
                       "where counter_name = ? group by name, dispatch_id",
                       (counter,)).fetchall()
    out = {}
    for name, _, v in rows:
        out.setdefault(name, []).append(float(v))
    return out


def _pmc_per_kernel(db_path, counter):
    """Mean value of `counter` per dispatch, by kernel name."""
    return {k: sum(v) / len(v) for k, v in _pmc_dispatches(db_path, counter).items()}


def summarize_trace(db_path, title=None):
    """Markdown table of a rocprofv3 --kernel-trace database: calls and
    mean / min / max / total duration per kernel name."""
    import sqlite3
    db = sqlite3.connect(db_path)
    cur = db.cursor()
    rows = cur.execute("select name, count(*), avg(end - start), min(end - start), "
                       "max(end - start), sum(end - start) from kernels group by name "
                       "order by sum(end - start) desc").fetchall()
    tot = sum(r[5] for r in rows) or 1
    lines = ["| kernel | calls | avg (us) | min (us) | max (us) | total % |",
             "|---|---:|---:|---:|---:|---:|"]
    for name, n, avg, mn, mx, sm in rows:
        nm = name if len(name) <= 90 else name[:87] + "..."
        lines.append("| `%s` | %d | %.1f | %.1f | %.1f | %.1f |" % (
            nm.replace("|", "\\|"), n, avg / 1e3, mn / 1e3, mx / 1e3, 100.0 * sm / tot))
    if title:
        print("### " + title)
    print("\n".join(lines))


def summarize_counters(db_path, counters, kernel):
    """Mean per dispatch and total of each counter for the kernels whose name
    contains `kernel` (the counting variant of peel_kernel excluded)."""
    out = {}
    for c in counters:
        d = _pmc_dispatches(db_path, c)
        vals = [v for k, vs in d.items() if kernel in k and "true>" not in k for v in vs]
        out[c] = {"dispatches": len(vals), "mean": (sum(vals) / len(vals)) if vals else None,
                  "total": sum(vals)}
    print(json.dumps(out, indent=1))


def profile_run(args):
    """One method (or the batch) alone in the process, for rocprofv3 counter
    passes: every dispatch of its kernel belongs to the workload named."""
    import pico_amd as pa
    ctx = pa.Context(0)
    ctx.probe_copy(COPY_PROBE_BYTES, 2)
    if args.profile == "batch":
        db = ctx.gen_rmat_batch(args.batch_graphs, 8, 18, 4.0, 32.0, seed=args.seed)
        for _ in range(2):
            _, _, st = db.run(block=1024, per_cu=1, copy=False)
        print(json.dumps({"profile": "batch", "device_ms": st["device_ms"],
                          "launches": st["launches"]}), flush=True)
        db.close()
    else:
        if args.graph == "chunglu":
            dg = ctx.gen_chunglu(1 << 24, 2.1, 20.0, 4e6, seed=3, rowptr_bytes=8)
        else:
            dg = ctx.gen_rmat(args.scale, args.edge_factor, seed=args.seed)
        for _ in range(2):
            _, st = dg.decompose(args.profile, copy=False)
        print(json.dumps({"profile": args.profile, "graph": args.graph,
                          "device_ms": st["device_ms"], "stages": st["stages"]}), flush=True)
        dg.close()
    ctx.close()


def summarize_pmc(paths, peel_stages=1, key=None, kernel=None, runs=2):
    """profiles/pmc_traffic.json: HBM bytes per launch of the dominant
    kernels = FETCH_SIZE and WRITE_SIZE (KiB per dispatch) scaled so that the
    copy probe's counters read the bytes it is known to move.  One plain-peel
    step is `peel_stages` launches of peel_kernel (one per compacted stage);
    its bytes per step are the mean per launch times that count."""
    fetch = _pmc_per_kernel(paths[0], "FETCH_SIZE")
    write = _pmc_per_kernel(paths[1], "WRITE_SIZE")

    def pick(d, key):
        # the counting variant of peel_kernel (template argument true) is
        # run once outside the timed steps: not the kernel measured
        ks = [k for k in d if key in k and "true>" not in k]
        return d[ks[0]] if ks else None

    cf, cw = pick(fetch, "copy_probe_kernel"), pick(write, "copy_probe_kernel")
    sf = COPY_PROBE_BYTES / (cf * 1024.0)
    sw = COPY_PROBE_BYTES / (cw * 1024.0)
    out = {"copy_probe": {"bytes": COPY_PROBE_BYTES, "fetch_kib": cf, "write_kib": cw,
                          "fetch_scale": sf, "write_scale": sw},
           "sources": ["FETCH_SIZE pass: " + os.path.basename(paths[0]),
                       "WRITE_SIZE pass: " + os.path.basename(paths[1])]}
    pairs = (((key, kernel),) if key else
             (("peel", "peel_kernel"), ("hindex", "hindex_kernel"), ("batch", "batch_kernel")))
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if key:  # merge one entry into the existing summary
        try:
            with open(path) as fh:
                merged = json.load(fh)
        except (OSError, ValueError):
            merged = {}
        merged.setdefault("copy_probe", out["copy_probe"])
        merged.setdefault("sources", [])
        out = merged
    fd = _pmc_dispatches(paths[0], "FETCH_SIZE")
    for k, kname in pairs:
        f, w = pick(fetch, kname), pick(write, kname)
        if f is None or w is None:
            continue
        per = int(f * 1024.0 * sf + w * 1024.0 * sw)
        if key:
            # a --profile run: its decompositions each launch the kernel the
            # same number of times (one per compacted stage)
            nd = sum(len(v) for kk, v in fd.items() if kname in kk and "true>" not in kk)
            n = max(1, round(nd / max(1, runs)))
        else:
            n = peel_stages if kname == "peel_kernel" else 1
        out[k] = {"kernel": kname, "fetch_kib": f, "write_kib": w,
                  "bytes_per_launch": per, "launches_per_step": n,
                  "bytes_per_step": per * n, "fetch_scale": sf, "write_scale": sw}
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out, indent=1))


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: the engine runs one process on one MI355X and "
              "has no multi-GPU sharding; --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    import numpy as np

    import pico_amd as pa

    if args.summarize_trace:
        summarize_trace(args.summarize_trace)
        return
    if args.summarize_counters:
        summarize_counters(args.summarize_counters[0], args.summarize_counters[2:],
                           args.summarize_counters[1])
        return
    if args.summarize_pmc:
        summarize_pmc(args.summarize_pmc, args.pmc_stages, args.pmc_key, args.pmc_kernel,
                      args.pmc_runs)
        return
    if args.profile:
        profile_run(args)
        return
    if args.only == "peel":
        args.no_batch = args.no_serve = args.no_methods = args.no_chunglu = True
        args.no_reader = True
    if args.only == "methods":
        args.no_batch = args.no_serve = args.no_chunglu = args.no_reader = True
    if args.only == "reader":
        args.no_batch = args.no_serve = args.no_chunglu = args.no_methods = True
    if args.only in ("batch", "serve", "chunglu"):  # one section alone (profiling runs)
        if args.only == "serve":
            print(json.dumps({"serving": bench_serving(args)}), flush=True)
        else:
            c = pa.Context(0)
            if args.only == "batch":
                print(json.dumps({"batched": bench_batched(c, args)}), flush=True)
            else:
                print(json.dumps({"chunglu": bench_chunglu(c, args, check=False)}), flush=True)
            c.close()
        return
    ctx = pa.Context(0)
    t0 = time.time()
    # built in HBM (the host generator's graph, edge for edge); resident
    # before timing starts
    dg = ctx.gen_rmat(args.scale, args.edge_factor, seed=args.seed)
    gen_s = time.time() - t0
    g = dg.download()  # host copy for the oracle and the CPU baseline
    V, nnz = g.V, g.nnz
    m = nnz // 2

    # streaming-copy probe: achievable HBM bandwidth on this device, and the
    # calibration kernel for the HBM byte counters
    copy_ms = ctx.probe_copy(COPY_PROBE_BYTES, 5)
    copy_gbs = 2 * COPY_PROBE_BYTES / (copy_ms * 1e-3) / 1e9

    # ---------------------------------------------------- plain peel steps
    for _ in range(args.warmup):
        dg.decompose(args.method, copy=False)
    dev_ms = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        _, st = dg.decompose(args.method, copy=False)
        dev_ms.append(st["device_ms"])
    t_total = time.perf_counter() - t_start
    ms_per_step = 1e3 * t_total / max(args.steps, 1)
    core, st = dg.decompose(args.method)
    kmax = st["kmax"]
    # one more run, untimed, of the variant that counts its atomics
    _, cst = dg.decompose(args.method, copy=False, flags=pa.kcore.KC_OPT_COUNT)
    kernel_ms = float(np.mean(dev_ms))
    b_alg = peel_bytes(V, nnz, kmax)
    # achieved: algorithmic bytes / kernel time; beside it the stricter
    # min(algorithmic, measured HBM bytes) / kernel time
    peel_traffic = traffic_from_profile("peel")
    b_used = min(b_alg, peel_traffic) if peel_traffic else b_alg
    achieved = b_alg / (kernel_ms * 1e-3) / 1e9
    achieved_min = b_used / (kernel_ms * 1e-3) / 1e9
    value = m / (ms_per_step * 1e-3) / 1e9

    # ------------------------------------------------------ h-index (same graph)
    # the first run also builds HistoCore's degree-ordered adjacency (prep_ms):
    # the cold time of one decomposition is that plus its kernel time
    _, cold = dg.decompose(args.hindex_method, copy=False)
    h_ms = []
    for _ in range(max(3, args.steps // 2)):
        _, hst = dg.decompose(args.hindex_method, copy=False)
        h_ms.append(hst["device_ms"])
    hcore, hst = dg.decompose(args.hindex_method)
    h_kernel_ms = float(np.mean(h_ms))
    bh = hindex_bytes(V, nnz, hst["iterations"], hst["changed"], hst["changed_degree"])
    h_traffic = traffic_from_profile("hindex")
    bh_used = min(bh, h_traffic) if h_traffic else bh
    h_ach = bh / (h_kernel_ms * 1e-3) / 1e9
    h_ach_min = bh_used / (h_kernel_ms * 1e-3) / 1e9

    # ------------------------------------------------------- CPU baseline
    # The reference has no CPU k-core; the baseline is the oracle's sequential
    # restatement of the method the GPU runs here (peel-dyn.cc / peelone-dyn.cc:
    # one full degree scan per level, then the level's queue), on one core,
    # over the full graph.  The textbook Batagelj-Zaversnik peel is timed beside it.
    cpu = None
    parity = None
    ref = None
    if not args.no_cpu:
        import oracle
        rp, col = g.rowptr, g.col
        times, bz_times = [], []
        for _ in range(max(1, args.cpu_runs)):
            t = time.perf_counter()
            ref = oracle.bz(rp, col)
            bz_times.append(time.perf_counter() - t)
        for _ in range(max(1, args.cpu_runs // 2)):
            t = time.perf_counter()
            pcore, _lv = oracle.bucket_peel(rp, col)
            times.append(time.perf_counter() - t)
        cpu_s = float(np.mean(times))
        bz_s = float(np.mean(bz_times))
        cpu = {"value": round(m / cpu_s / 1e9, 5), "unit": "GTEPS", "cores": 1, "kind": "port",
               "sample": "full R-MAT s%d graph (%d undirected edges), the sequential "
                         "restatement of peel-dyn.cc's bucket peel (oracle/kc_oracle.c "
                         "orc_bucket_peel), mean of %d runs = %.3f s" % (
                             args.scale, m, len(times), cpu_s),
               "nproc": os.cpu_count(), "cpu_model": cpu_model(),
               "bz_gteps": round(m / bz_s / 1e9, 5), "bz_s": round(bz_s, 3),
               "bz_runs": len(bz_times)}
        parity = bool(np.array_equal(core, ref) and np.array_equal(hcore, ref) and
                      np.array_equal(pcore, ref))

    methods = None if args.no_methods else bench_methods(dg, V, nnz, m, ref,
                                                         hst["changed_degree"])
    reader = None if args.no_reader else bench_reader(ctx, g, args)
    del g
    chunglu = None if args.no_chunglu else bench_chunglu(ctx, args, check=not args.no_cpu)
    batched = None if args.no_batch else bench_batched(ctx, args)
    serving = sweep = None
    if not args.no_serve:
        serving = bench_serving(args)
        sweep = [bench_serving(args, rate=float(r))
                 for r in args.serve_sweep.split(",") if r.strip()]

    out = {
        "metric": "k-core GTEPS + % HBM roofline (R-MAT s22); batched vs plain graphs/s; p99",
        "value": round(value, 4),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded R-MAT generator, pico_amd host C++)",
        "config": {"workload": "R-MAT scale %d edge factor %g (.57,.19,.19,.05) seed %d: "
                               "plain peel (%s) on the resident int32 CSR"
                               % (args.scale, args.edge_factor, args.seed, args.method),
                   "V": V, "undirected_edges": m, "kmax": kmax,
                   "levels_run": st["levels_run"], "grid_barriers": st["phases"],
                   "grid": st["grid"], "block": st["block"], "graph_gen_s": round(gen_s, 2)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": peel_traffic,
                     "kernel": "peel_kernel", "kernel_ms": round(kernel_ms, 4),
                     "launches_per_step": st["stages"],
                     "kernel_ms_median": round(float(np.median(dev_ms)), 4),
                     "kernel_ms_trimmed": round(trimmed_mean(dev_ms), 4),
                     "algorithmic_bytes": b_alg, "atomics": cst["atomics"],
                     "achieved_min_rule": round(achieved_min, 2),
                     "frac_min_rule": round(achieved_min / HBM_PEAK_GBS, 5),
                     "accounting": "achieved = algorithmic bytes / mean step time; "
                                   "*_min_rule = min(algorithmic, traffic) bytes / mean step time",
                     "copy_probe_gbs": round(copy_gbs, 1)},
        "hindex": {"method": args.hindex_method, "kernel_ms": round(h_kernel_ms, 4),
                   "cold_ms": round(cold["prep_ms"] + cold["device_ms"], 3),
                   "prep_ms": round(cold["prep_ms"], 3),
                   "gteps": round(m / (h_kernel_ms * 1e-3) / 1e9, 4),
                   "rounds": hst["iterations"], "changed": hst["changed"],
                   "roofline": {"bound": "hbm",
                                "achieved": round(h_ach, 2),
                                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                "frac": round(h_ach / HBM_PEAK_GBS, 5),
                                "achieved_min_rule": round(h_ach_min, 2),
                                "frac_min_rule": round(h_ach_min / HBM_PEAK_GBS, 5),
                                "traffic": h_traffic,
                                "kernel": "hindex_kernel", "algorithmic_bytes": bh,
                                "atomics": hst["atomics"]}},
        "cpu_baseline": cpu,
        "parity_vs_oracle": parity,
        "methods": methods,
        "chunglu": chunglu,
        "batched": batched,
        "serving": serving,
        "serving_sweep": sweep,
        "reader": reader,
    }
    print(json.dumps(out), flush=True)
    dg.close()
    ctx.close()


if __name__ == "__main__":
    main()
