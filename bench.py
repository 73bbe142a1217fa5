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


def survey_bytes(row: np.ndarray, col: np.ndarray, ev: np.ndarray, eu: np.ndarray, micro: bool) -> float:
    """B(G) = 4 sum_e [d_v + d_u + min(s(u)-d_v, s(v)-d_u)] + 8(n+1) + 8m + 64m[micro] + 136."""
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
        "sample": f"every {stride}th edge of E_ordered ({len(samp)} edges, {swork:.3e} of {total_work:.3e} "
                  f"probes, {dt:.1f}s); total extrapolated by the probe model: {est_total:.0f}s",
        "cpu": cpu_info(),
    }


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
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy three times (scale for the PMC byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU: run with --gpus 1"}))
        return 2
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
    import pgd_amd
    import torch

    t0 = time.perf_counter()
    s, d = pgd_amd.gen_rmat(args.scale, args.ef, 0.57, 0.19, 0.19, args.seed)
    n = int(max(s.max(), d.max())) + 1          # the .mtx reader's n
    log(f"[bench] generated R-MAT scale {args.scale}: n={n} m={len(s)} in {time.perf_counter() - t0:.1f}s")
    ctx = pgd_amd.Context(0)
    g = pgd_amd.GraphletCore(s, d, n=n, ctx=ctx, preprocess=False)
    log(f"[bench] device CSR built; max degree {g.max_degree}")

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
    B = survey_bytes(row, col, ev, eu, micro=True)
    out = {
        "metric": "edges/s for exact 4-node graphlet counts (1 GPU)",
        "value": m / dt, "unit": "edges/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64", "data": "synthetic R-MAT from a seed",
        "config": {"workload": f"rmat_scale{args.scale}_ef{args.ef}_a0.57_b0.19_c0.19_seed{args.seed}",
                   "n": g.n, "m": m, "max_degree": g.max_degree, "micro": True, "engine": g.last_engine()},
        "counting_ms": float(np.mean(tot)),
        "roofline": roofline(phases, f"rmat_scale{args.scale}_ef{args.ef}_a0.57_b0.19_c0.19_seed{args.seed}"),
        "phases": [{"name": p["name"], "ms": round(float(np.mean([ph[j]["ms"] for ph in phases])), 3),
                    "GB": round(p["bytes"] / 1e9, 3), "units": p["units"]} for j, p in enumerate(phases[0])],
        "survey_model": {"bytes": B, "achieved_GBs": B / dt / 1e9, "frac_of_hbm_peak": B / dt / 1e9 / HBM_PEAK_GBS},
        "totals": dict(zip(pgd_amd.TOTAL_NAMES, totals)),
    }
    if not args.no_cpu_baseline:
        try:
            threads = min(os.cpu_count() or 1, 16)
            out["cpu_baseline"] = cpu_baseline(g, args.cpu_seconds, threads)
            out["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
        except Exception as e:  # the baseline must not hide the GPU number
            out["cpu_baseline"] = {"error": repr(e)}
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
