"""Benchmark: edges/s for exact 4-node graphlet counts on one MI355X.

Workload: R-MAT/Kronecker graph, scale 22 (4.19M vertex ids), edge factor 16,
(a,b,c,d) = (0.57,0.19,0.19,0.05), symmetrised, self loops and duplicates
removed, int32 ids, uint64 counts; synthetic from a seed.

One step is the reference's timed region (graphlet_driver.cpp:69-75) plus the
preprocessing that the reference does in its constructor: neighbour ordering,
orientation and edge ids, E_ordered, the per-edge counting, the 15-sum
reduction, the micro columns in device memory and the derivation of the 17
totals.  The graph's CSR is resident in HBM before the timed region.

Prints one JSON line (rank 0).  --gpus N > 1 is refused: this build measures
one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
INT8_PEAK_TOPS = 5000.0   # dense int8 MFMA, MI355X
FP4_PEAK_TFLOPS = 10100.0  # dense fp4/fp6 MFMA peak of one MI355X


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_info() -> dict:
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count()}


def cpu_threads() -> int:
    """Host threads for the CPU baseline: every core this process may use
    (a one-GPU box grants 16, and os.cpu_count() reports that share)."""
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


def edge_walk_bytes(row: np.ndarray, col: np.ndarray, ev: np.ndarray, eu: np.ndarray, micro: bool) -> float:
    """Bytes B(G) of the min-side edge-centric walk (every edge reads both
    neighbour lists and the smaller 2-hop side): 4 sum_e [d_v + d_u +
    min(s(u)-d_v, s(v)-d_u)] + 8(n+1) + 8m + 64m[micro] + 136."""
    deg = np.diff(row).astype(np.int64)
    s = np.zeros(len(deg), dtype=np.float64)
    src = np.repeat(np.arange(len(deg)), deg)
    np.add.at(s, src, deg[col].astype(np.float64))
    dv, du = deg[ev].astype(np.float64), deg[eu].astype(np.float64)
    per = dv + du + np.minimum(s[eu] - dv, s[ev] - du)
    n, m = len(deg), len(ev)
    return 4.0 * per.sum() + 8.0 * (n + 1) + 8.0 * m + (64.0 * m if micro else 0.0) + 136.0


def measured_traffic(kernel: str, workload: str) -> float | None:
    """HBM bytes per launch of `kernel` on `workload`, from the rocprofv3
    FETCH_SIZE/WRITE_SIZE summaries committed under profiles/ (calibrated on
    a plain copy; see the summary files)."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        with open(path) as f:
            for line in f:
                p = line.split()
                if len(p) == 4 and p[0] == "traffic_bytes" and p[1] == kernel and p[2] == workload:
                    return float(p[3])
    return None


def measured_atomics(kernel: str, workload: str) -> float | None:
    """Memory-side atomic requests per launch of `kernel` on `workload`
    (rocprofv3 TCC_ATOMIC_sum, from the summaries committed under profiles/)."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        with open(path) as f:
            for line in f:
                p = line.split()
                if len(p) == 4 and p[0] == "atomic_requests" and p[1] == kernel and p[2] == workload:
                    return float(p[3])
    return None


def atomic_ceiling(ctx, words: int = 1 << 27, reps: int = 10) -> dict:
    """The device's rate of global atomic adds (pgd_atomic_rate): one atomic
    add per 32-bit counter and launch, consecutive lanes on consecutive
    counters, so a wavefront's adds form 64-byte requests of 16 counters.
    Those requests are performed past the L2 (a profile of them counts every
    one as an L2 miss and 64 bytes written per request), so this rate, not
    the HBM bandwidth, bounds a pass made of per-slot sums."""
    r, ms = ctx.atomic_rate(words, reps)
    return {"words": words, "launches": reps, "ms_per_launch": ms, "requests_per_s": r,
            "atomic_adds_per_s": r * 16.0, "method": "pgd_atomic_rate (one atomic add per int32 counter per launch)"}


def roofline(phases: list[list[dict]], workload: str = "") -> dict | None:
    """Roofline of the dominant counting kernel: its algorithmic HBM bytes
    (the engine's work model for that pass, per launch) over its average
    device time measured with HIP events on the engine's stream."""
    if not phases or not phases[0]:
        return None
    names = [p["name"] for p in phases[0]]
    avg = {nm: float(np.mean([ph[i]["ms"] for ph in phases])) for i, nm in enumerate(names)}
    dom = max(avg, key=avg.get)
    i = names.index(dom)
    by = phases[0][i]["bytes"]
    ach = by / (avg[dom] / 1e3) / 1e9
    tr = measured_traffic(dom, workload)
    return {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": ach / HBM_PEAK_GBS, "traffic": tr, "algorithmic_GB": by / 1e9, "ms": avg[dom],
            "units": phases[0][i]["units"]}


def cpu_baseline(g, target_s: float, threads: int) -> dict:
    """The oracle (the reference's OpenMP edge loop restated in C) on a
    deterministic, work-stratified sample of E_ordered, extrapolated to the
    whole graph with the reference's per-edge work model
    2 d_v + d_u + (s(u) - d_v)."""
    from tests.oracle_lib import OracleGraph
    row, col = g.csr()
    ev, eu = g.edge_arrays()
    order = g.e_ordered()
    og = OracleGraph.from_csr(row, col)
    deg = np.diff(row).astype(np.int64)
    src = np.repeat(np.arange(len(deg)), deg)
    s = np.zeros(len(deg), dtype=np.float64)
    np.add.at(s, src, deg[col].astype(np.float64))
    work = 2 * deg[ev] + deg[eu] + (s[eu] - deg[ev])        # per edge id
    w_ord = work[order]
    total_work = float(w_ord.sum())
    # calibrate the probe rate on a small stride, then size the sample
    stride = max(1, len(order) // 2000)
    samp = np.ascontiguousarray(order[::stride])
    t0 = time.perf_counter()
    og.decompose(ev, eu, samp, micro=False, threads=threads)
    dt = max(time.perf_counter() - t0, 1e-6)
    rate = float(work[samp].sum()) / dt
    want = target_s * rate
    frac = min(1.0, want / total_work)
    stride = max(1, int(round(1.0 / frac)))
    samp = np.ascontiguousarray(order[::stride])
    t0 = time.perf_counter()
    og.decompose(ev, eu, samp, micro=False, threads=threads)
    dt = time.perf_counter() - t0
    swork = float(work[samp].sum())
    est_total = dt * total_work / swork
    return {
        "value": len(order) / est_total, "unit": "edges/s", "cores": threads, "kind": "port",
        "estimated_total_s": est_total,
        "sample": f"every {stride}th edge of E_ordered ({len(samp)} edges, {swork:.3e} of {total_work:.3e} "
                  f"probes, {dt:.1f}s); total extrapolated by the probe model: {est_total:.0f}s",
        "cpu": cpu_info(),
    }


def bench_micro_writer(g) -> dict:
    """The .micro table of the last decomposition, formatted on the device
    (pgd_write_micro): to /dev/null (device formatting, pinned copies to the
    host, write calls) and to a file in the temporary directory."""
    import tempfile
    out = {}
    fd = os.open(os.devnull, os.O_WRONLY)
    try:
        g.write_micro_fd(fd)  # warm-up (buffers)
        t0 = time.perf_counter()
        nb = g.write_micro_fd(fd)
        dt = time.perf_counter() - t0
    finally:
        os.close(fd)
    out["bytes"] = nb
    out["rows"] = g.m
    out["devnull_s"] = dt
    out["devnull_GB_per_s"] = nb / dt / 1e9
    path = os.path.join(tempfile.gettempdir(), f"pgd_bench_{os.getpid()}.micro")
    try:
        t0 = time.perf_counter()
        g.write_micro_stats(path)
        dt = time.perf_counter() - t0
        out["file_s"] = dt
        out["file_GB_per_s"] = nb / dt / 1e9
    finally:
        if os.path.exists(path):
            os.remove(path)
    log(f"[bench] micro writer: {nb / 1e9:.2f} GB, {out['devnull_GB_per_s']:.1f} GB/s to /dev/null, "
        f"{out['file_GB_per_s']:.1f} GB/s to a file")
    return out


def bench_reader(pgd_amd, ctx, n, s, d, g) -> dict:
    """read_graph straight to a device CSR (pgd_graph_from_file: file text to
    CSR in reader order, as the reference's reader leaves it before
    preprocessing) on the R-MAT-22 lines written as a .mtx file, and as an
    edge list (the same text; its header line reads as a self loop), which
    takes the device renumbering by first appearance.  All host threads, and
    one thread for comparison; the file is in the page cache (just written)."""
    import shutil
    import tempfile
    out = {"threads": os.cpu_count()}
    base = os.path.join(tempfile.gettempdir(), f"pgd_bench_{os.getpid()}")
    mtx, txt = base + ".mtx", base + ".txt"
    try:
        pgd_amd.write_mtx(mtx, n, s, d)
        shutil.copyfile(mtx, txt)
        out["file_bytes"] = os.path.getsize(mtx)
        ref = pgd_amd.GraphletCore(s, d, n=n, ctx=ctx, preprocess=False)  # the host-COO route, reader order
        row0, col0 = ref.csr()
        deg0 = np.sort(np.diff(row0))
        deg0 = deg0[deg0 > 0]
        ref.free()
        for name, path in (("mtx", mtx), ("edge_list", txt)):
            res = {}
            for label, threads in (("all_threads", 0), ("one_thread", 1)):
                t0 = time.perf_counter()
                h = pgd_amd.GraphletCore.from_file(path, ctx=ctx, threads=threads, preprocess=False)
                dt = time.perf_counter() - t0
                if label == "all_threads":
                    r, c = h.csr()
                    res["edges"] = h.m
                    if name == "mtx":  # the same ids and lines: the same CSR as the host route
                        res["same_csr_as_host_route"] = bool(np.array_equal(r, row0) and np.array_equal(c, col0))
                    else:
                        # renumbered by first appearance: ids that no line names (isolated
                        # vertices of the .mtx route) are absent, and the header line adds
                        # one vertex whose self loop is dropped; the nonzero degrees agree
                        deg = np.sort(np.diff(r))
                        res["same_nonzero_degrees_as_host_route"] = bool(np.array_equal(deg[deg > 0], deg0))
                h.free()
                res[label + "_s"] = dt
                res[label + "_lines_per_s"] = len(s) / dt
            out[name] = res
            log(f"[bench] reader {name}: {res['all_threads_s']:.2f} s with all threads, "
                f"{res['one_thread_s']:.2f} s with one")
    finally:
        for p in (mtx, txt):
            if os.path.exists(p):
                os.remove(p)
    return out


def bench_dense(pgd_amd, ctx) -> dict:
    """Dense tier on G(500, 0.9) and G(2000, 0.5): device time of the int8
    MFMA engine and of the sparse oriented engine on the same graph, the
    product launches' achieved int8 rate (S = A.A, 2 N^3 operations, and the
    fused [A3 | AT], 8 N^3; the digit kernel between them is timed apart), the
    rate of the 4-clique products of the local graphs (fp4 MFMA, against the
    fp4 peak), and the tier's phases."""
    out = {}
    for n, p in ((500, 0.9), (2000, 0.5)):
        s, d = pgd_amd.gen_gnp(n, p, 1)
        g = pgd_amd.GraphletCore(s, d, n=n, ctx=ctx)
        g.sort_edges("deg")
        res = {}

        def mfma(ms_ops, peak=INT8_PEAK_TOPS, unit="TOP/s"):
            gms, ops = sorted(ms_ops)[len(ms_ops) // 2]
            return {"bound": "mfma", "achieved": ops / (gms / 1e3) / 1e12, "peak": peak,
                    "unit": unit, "frac": ops / (gms / 1e3) / 1e12 / peak, "ops": ops, "ms": gms}

        for eng in ("dense", "orient"):
            g.decompose(micro=True, engine=eng)  # warm-up (allocations, plan)
            ms, gemm, k4m, phases = [], [], [], []
            for _ in range(3):
                g.decompose(micro=True, engine=eng)
                ms.append(g.last_timing()[0])
                ph = g.last_phases()
                phases.append({q["name"]: q["ms"] for q in ph})
                mm = [q for q in ph if q["name"].startswith("k_gemm_i8")]
                if mm:  # the MFMA launches (S, then [A3 | AT]); bytes holds their operations
                    gemm.append((sum(q["ms"] for q in mm), sum(q["bytes"] for q in mm)))
                kk = [q for q in ph if q["name"] == "k_dense_k4m"]
                if kk:  # the 4-clique products of the local graphs (fp4 MFMA; operations counted on the device)
                    k4m.append((kk[0]["ms"], kk[0]["bytes"]))
            res[eng] = float(np.median(ms))
            if gemm:
                res["mfma"] = mfma(gemm)
            if k4m:  # fp4 (e2m1) operands, f32 accumulation
                res["mfma_k4"] = mfma(k4m, FP4_PEAK_TFLOPS, "TFLOP/s")
            if eng == "dense":
                res["phases_ms"] = {k: float(np.median([p_[k] for p_ in phases])) for k in phases[0]}
        out[f"gnp_{n}_{p}_seed1"] = {"m": g.m, "dense_ms": res["dense"], "orient_ms": res["orient"],
                                     "edges_per_s": g.m / (res["dense"] / 1e3), "mfma": res.get("mfma"),
                                     "mfma_k4": res.get("mfma_k4"), "phases_ms": res.get("phases_ms")}
        g.free()
        log(f"[bench] dense G({n},{p}): dense {res['dense']:.2f} ms, sparse {res['orient']:.2f} ms")
    return out


def bench_batched(pgd_amd, ctx, count: int = 8192) -> dict:
    """Batched form on the batched workload (8192 graphs): one captured
    launch sequence replayed, against plain one-graph-per-launch on a sample."""
    from pgd_amd.workloads import batched_graphs
    graphs = batched_graphs(count, seed=1)
    t0 = time.perf_counter()
    b = pgd_amd.Batch(graphs, ctx=ctx, micro=True)
    t_create = time.perf_counter() - t0
    t0 = time.perf_counter()
    s0, _ = b.run()
    t_first = time.perf_counter() - t0
    best, tune_ms = b.autotune(trials=2)
    reps = []
    for _ in range(3):
        t0 = time.perf_counter()
        s, _ = b.run()
        reps.append((time.perf_counter() - t0, b.last_ms()))
        assert np.array_equal(s, s0)
    wall, dev = sorted(reps)[1]
    idx = list(range(0, count, max(1, count // 256)))
    t0 = time.perf_counter()
    for i in idx:
        n, src, dst = graphs[i]
        g = pgd_amd.GraphletCore(src, dst, n=n, ctx=ctx)
        g.sort_edges("deg")
        sm = g.decompose(micro=True)
        assert [int(x) for x in s0[i]] == sm, i
        g.free()
    plain_gps = len(idx) / (time.perf_counter() - t0)
    M = b.m_total
    out = {"graphs": count, "edges": M, "vertices": b.n_total,
           "workload": "batched: BA(m in [2,8]) / G(n,p) avg degree [4,32], n log-uniform [20,5000], seed 1",
           "create_s": t_create, "first_run_s": t_first, "replay_ms": dev, "replay_wall_s": wall,
           "graphs_per_s": count / (t_create + t_first), "edges_per_s": M / (t_create + t_first),
           "graphs_per_s_basis": "end to end: batch creation (upload, union build, preprocessing, plan, capture) "
                                 "plus its first run",
           "one_graph_per_launch_graphs_per_s": plain_gps, "one_graph_per_launch_sample": len(idx),
           "speedup_vs_one_graph_per_launch": (count / (t_create + t_first)) / plain_gps,
           "replay_graphs_per_s": count / wall, "replay_speedup_vs_one_graph_per_launch": (count / wall) / plain_gps,
           "tuning": best.as_dict(), "autotuned_ms": tune_ms}
    b.free()
    log(f"[bench] batched: {out['graphs_per_s']:.0f} graphs/s end to end ({out['replay_graphs_per_s']:.0f} replayed) "
        f"vs {plain_gps:.0f} one-graph-per-launch")
    return out


def bench_served(pgd_amd, ctx, rate: float = 2000.0, seconds: float = 4.0, large_frac: float = 0.01) -> dict:
    """Served form: Poisson arrivals of batched-workload graphs mixed with
    R-MAT scale-18 graphs, per-edge results kept, counted by the server."""
    from pgd_amd.workloads import batched_graph
    pool = [batched_graph(i, seed=9)[:3] for i in range(1500)]
    large = []
    for sd in (1, 2):
        s, d = pgd_amd.gen_rmat(18, 16, 0.57, 0.19, 0.19, sd)
        large.append((int(max(s.max(), d.max())) + 1, s, d))
    srv = pgd_amd.Server(ctx=ctx, hbm_budget=96 << 30)
    rng = np.random.default_rng(1)
    t0 = time.perf_counter()
    t, k, nl = 0.0, 0, 0
    while t < seconds:
        t += rng.exponential(1.0 / rate)
        now = time.perf_counter() - t0
        if t > now:
            time.sleep(t - now)
        if rng.random() < large_frac:
            g = large[nl % 2]
            nl += 1
        else:
            g = pool[k % len(pool)]
        srv.submit(*g, micro=True)
        k += 1
        if k % 256 == 0:
            for rid in srv.poll():
                srv.release(rid)
    t_sub = time.perf_counter() - t0
    srv.drain(600000)
    for rid in srv.poll():
        srv.release(rid)
    st = srv.stats()
    srv.free()
    out = {"offered_req_per_s": rate, "large_fraction": large_frac, "submitted": k, "large": nl,
           "submit_s": t_sub, "graphs_per_s": st["finished"] / st["elapsed_s"],
           "edges_per_s": st["edges"] / st["elapsed_s"], "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"],
           "batches": st["batches"], "elapsed_s": st["elapsed_s"]}
    log(f"[bench] served: {out['graphs_per_s']:.0f} graphs/s p50 {out['p50_ms']:.1f} ms p99 {out['p99_ms']:.1f} ms")
    # capacity: the same request mix submitted as fast as the host can, so the
    # queue never runs dry; graphs/s is then what the server can finish
    srv = pgd_amd.Server(ctx=ctx, hbm_budget=96 << 30)
    rng = np.random.default_rng(3)
    ncap, nl2 = 16000, 0
    for i in range(ncap):
        if rng.random() < large_frac:
            g = large[nl2 % 2]
            nl2 += 1
        else:
            g = pool[i % len(pool)]
        srv.submit(*g, micro=True)
        if i % 256 == 255:
            for rid in srv.poll():
                srv.release(rid)
    srv.drain(600000)
    for rid in srv.poll():
        srv.release(rid)
    st = srv.stats()
    srv.free()
    out["capacity"] = {"submitted": ncap, "large": nl2, "graphs_per_s": st["finished"] / st["elapsed_s"],
                       "edges_per_s": st["edges"] / st["elapsed_s"], "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"],
                       "batches": st["batches"], "elapsed_s": st["elapsed_s"]}
    log(f"[bench] served capacity: {out['capacity']['graphs_per_s']:.0f} graphs/s")
    # one graph per launch on the same mix: the plain path per request (build,
    # preprocessing, E_ordered, counting with micro columns), timed on 256 pool
    # graphs and on each large graph twice, weighted 99:1
    def plain(n, s, d):
        t0 = time.perf_counter()
        g = pgd_amd.GraphletCore(s, d, n=n, ctx=ctx)
        g.sort_edges("deg")
        g.decompose(micro=True)
        g.free()
        return time.perf_counter() - t0
    plain(*pool[0])
    t_small = float(np.mean([plain(*pool[i]) for i in range(0, len(pool), max(1, len(pool) // 256))]))
    t_large = float(np.mean([plain(*large[j % 2]) for j in range(4)]))
    mix = 1.0 / ((1 - large_frac) * t_small + large_frac * t_large)
    out["one_graph_per_launch_same_mix_graphs_per_s"] = mix
    out["one_graph_per_launch_small_ms"] = t_small * 1e3
    out["one_graph_per_launch_large_ms"] = t_large * 1e3
    out["capacity_vs_one_graph_per_launch_same_mix"] = out["capacity"]["graphs_per_s"] / mix
    log(f"[bench] one graph per launch on the served mix: {mix:.0f} graphs/s "
        f"(small {t_small * 1e3:.2f} ms, R-MAT-18 {t_large * 1e3:.1f} ms)")
    return out


def served_fresh_process() -> dict:
    """bench_served in a child process of its own (a fresh HIP context)."""
    try:
        r = subprocess.run([sys.executable, "-u", os.path.abspath(__file__), "--served-only"], capture_output=True,
                           text=True, timeout=600, cwd=ROOT)
        for line in r.stderr.splitlines():
            log(line)
        if r.returncode != 0:
            return {"error": f"served child exited {r.returncode}: {r.stderr[-400:]}"}
        res = json.loads(r.stdout.strip().splitlines()[-1])
        res["process"] = "fresh child process, measured before the other sections"
        return res
    except Exception as e:  # the section must not hide the headline
        return {"error": repr(e)}


def validate_cpu_baseline(pgd_amd, ctx, scales, target_s: float, threads: int) -> list:
    """The sampled CPU baseline's extrapolation against full runs of the
    oracle over every edge of E_ordered, on smaller R-MAT graphs."""
    from tests.oracle_lib import OracleGraph
    res = []
    for scale in scales:
        s, d = pgd_amd.gen_rmat(scale, 16, 0.57, 0.19, 0.19, 1)
        n = int(max(s.max(), d.max())) + 1
        g = pgd_amd.GraphletCore(s, d, n=n, ctx=ctx)
        g.sort_edges("deg")
        est = cpu_baseline(g, target_s, threads)
        row, col = g.csr()
        ev, eu = g.edge_arrays()
        order = g.e_ordered()
        g.free()
        og = OracleGraph.from_csr(row, col)
        log(f"[bench] full oracle run on R-MAT-{scale} ({len(order)} edges, {threads} threads)")
        # in 32 consecutive ranges of E_ordered, each timed (the edges are
        # independent: the sum is the full run), with a line after each
        full = 0.0
        cuts = np.linspace(0, len(order), 33).astype(np.int64)
        for k in range(32):
            part = np.ascontiguousarray(order[cuts[k]:cuts[k + 1]])
            t0 = time.perf_counter()
            og.decompose(ev, eu, part, micro=False, threads=threads)
            full += time.perf_counter() - t0
            log(f"[bench]   range {k + 1}/32 done, {full:.1f} s so far")
        res.append({"scale": scale, "edges": int(len(order)), "full_s": full,
                    "estimated_s": est["estimated_total_s"], "sample": est["sample"],
                    "relative_error": est["estimated_total_s"] / full - 1.0})
        log(f"[bench] R-MAT-{scale}: full {full:.1f} s, estimated {est['estimated_total_s']:.1f} s")
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--ef", type=int, default=16)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--sections", default="io,dense,batched,served",
                    help="extra measurements after the headline: any of io,dense,batched,served (or none)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy three times (scale for the PMC byte counters)")
    ap.add_argument("--atomic-ceiling", action="store_true",
                    help="only measure the device's global atomic-add request rate (pgd_atomic_rate) and print it")
    ap.add_argument("--served-only", action="store_true",
                    help="only the served measurement, in this (fresh) process; prints its JSON")
    ap.add_argument("--validate-cpu-baseline", default="",
                    help="comma-separated R-MAT scales: compare the sampled CPU baseline's extrapolation with full "
                         "oracle runs there, print that JSON and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU: run with --gpus 1"}))
        return 2
    if args.validate_cpu_baseline:
        import pgd_amd
        ctx = pgd_amd.Context(0)
        scales = [int(x) for x in args.validate_cpu_baseline.split(",")]
        print(json.dumps({"cpu_baseline_validation": validate_cpu_baseline(
            pgd_amd, ctx, scales, args.cpu_seconds, cpu_threads())}), flush=True)
        ctx.close()
        return 0
    if args.served_only:
        import pgd_amd
        ctx = pgd_amd.Context(0)
        print(json.dumps(bench_served(pgd_amd, ctx)), flush=True)
        ctx.close()
        return 0
    served = None
    if "served" in args.sections.split(","):
        # the served form is measured first, in a fresh child process, so that
        # nothing this process allocates or runs beforehand affects it
        served = served_fresh_process()
    if args.copy_calibration:
        import torch
        a = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration": {"bytes_read": a.numel() * 4, "bytes_written": a.numel() * 4,
                                               "launches": 3}}))
        return 0
    if args.atomic_ceiling:
        import pgd_amd
        ctx = pgd_amd.Context(0)
        print(json.dumps({"atomic_ceiling": atomic_ceiling(ctx)}))
        ctx.close()
        return 0
    import pgd_amd
    import torch

    t0 = time.perf_counter()
    s, d = pgd_amd.gen_rmat(args.scale, args.ef, 0.57, 0.19, 0.19, args.seed)
    n = int(max(s.max(), d.max())) + 1          # the .mtx reader's n
    log(f"[bench] generated R-MAT scale {args.scale}: n={n} m={len(s)} in {time.perf_counter() - t0:.1f}s")
    ctx = pgd_amd.Context(0)
    g = pgd_amd.GraphletCore(s, d, n=n, ctx=ctx, preprocess=False)
    log(f"[bench] device CSR built; max degree {g.max_degree}")
    # host edge list -> device CSR (upload over PCIe, dedup, CSR assembly), timed
    # on a second construction so the first-use costs are out of it
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g2 = pgd_amd.GraphletCore(s, d, n=n, ctx=ctx, preprocess=False)
    torch.cuda.synchronize()
    t_upload = time.perf_counter() - t0
    g2.free()

    def step():
        g.preprocess("deg", False)
        g.sort_edges("deg", False)
        sums = g.decompose(micro=True)
        return pgd_amd.derive(g.n, g.m, sums)

    for i in range(args.warmup):
        step()
        log(f"[bench] warmup {i} done ({g.last_engine()})")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dom = []
    tot = []
    phases = []
    for i in range(args.steps):
        totals = step()
        a, b, _ = g.last_timing()
        tot.append(a)
        dom.append(b)
        phases.append(g.last_phases())
        log(f"[bench] step {i}: counting {a:.1f} ms (dominant kernel {b:.1f} ms)")
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    m = g.m
    row, col = g.csr()
    ev, eu = g.edge_arrays()
    B = edge_walk_bytes(row, col, ev, eu, micro=True)
    out = {
        "metric": "edges/s for exact 4-node graphlet counts (1 GPU)",
        "value": m / dt, "unit": "edges/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64", "data": "synthetic R-MAT from a seed",
        "config": {"workload": f"rmat_scale{args.scale}_ef{args.ef}_a0.57_b0.19_c0.19_seed{args.seed}",
                   "n": g.n, "m": m, "max_degree": g.max_degree, "micro": True, "engine": g.last_engine()},
        "counting_ms": float(np.mean(tot)),
        "upload_and_csr_ms": t_upload * 1e3,
        "pcie_inclusive_edges_per_s": m / (dt + t_upload),
        "roofline": roofline(phases, f"rmat_scale{args.scale}_ef{args.ef}_a0.57_b0.19_c0.19_seed{args.seed}"),
        "phases": [{"name": p["name"], "ms": round(float(np.mean([ph[j]["ms"] for ph in phases])), 3),
                    "GB": round(p["bytes"] / 1e9, 3), "units": p["units"]} for j, p in enumerate(phases[0])],
        "edge_walk_model": {"bytes": B, "achieved_GBs": B / dt / 1e9, "frac_of_hbm_peak": B / dt / 1e9 / HBM_PEAK_GBS},
        "totals": dict(zip(pgd_amd.TOTAL_NAMES, totals)),
    }
    # the bound the dominant kernel meets first: its per-slot sums are global
    # atomics, performed past the L2 at the device's atomic-request rate
    rl = out["roofline"]
    if rl:
        try:
            ceil = atomic_ceiling(ctx)
            req = measured_atomics(rl["kernel"], out["config"]["workload"])
            rate = req / (rl["ms"] / 1e3) if req else None
            out["atomics"] = {"kernel": rl["kernel"], "requests_per_launch": req, "requests_per_s": rate,
                              "ceiling": ceil, "frac_of_ceiling": rate / ceil["requests_per_s"] if rate else None}
        except Exception as e:  # a section must not hide the headline
            out["atomics"] = {"error": repr(e)}
    if "io" in args.sections.split(","):
        try:
            out["micro_writer"] = bench_micro_writer(g)
        except Exception as e:  # a section must not hide the headline
            out["micro_writer"] = {"error": repr(e)}
        try:
            out["reader"] = bench_reader(pgd_amd, ctx, n, s, d, g)
        except Exception as e:  # a section must not hide the headline
            out["reader"] = {"error": repr(e)}
    for sec, fn in (("dense", bench_dense), ("batched", bench_batched)):
        if sec in args.sections.split(","):
            try:
                out[sec] = fn(pgd_amd, ctx)
            except Exception as e:  # a section must not hide the headline
                out[sec] = {"error": repr(e)}
    if served is not None:
        out["served"] = served
        plain = out.get("batched", {}).get("one_graph_per_launch_graphs_per_s")
        if plain and "capacity" in served:
            served["capacity_vs_one_graph_per_launch"] = served["capacity"]["graphs_per_s"] / plain
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(g, args.cpu_seconds, cpu_threads())
            out["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
        except Exception as e:  # the baseline must not hide the GPU number
            out["cpu_baseline"] = {"error": repr(e)}
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
