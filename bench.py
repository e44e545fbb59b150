#!/usr/bin/env python3
"""Benchmark of the Poisson (MG-PCG) path on one MI355X.

Workload: 1024 x 1024 RGB Neumann Poisson problems (three independent channels
solved together by one batched plan); each channel is the seeded sum of 8 random
Gaussians plus N(0, 0.01) noise, its right-hand side ``div(grad target)``.
One step = one complete solve of the RGB image to a true FP64 relative residual
<= 1e-8 per channel, from device-resident right-hand sides to device-resident
solutions.

``value`` = solved MPix/s of the mixed solver (FP32 V-cycle, FP64 outer PCG);
the same line carries the plain FP64 MG-PCG rate, their ratio, request latency
percentiles, the max-abs error against the exact solution and the FP64 solve,
the live per-kernel roofline of the dominant kernel and the CPU oracle baseline.
"""

from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for _p in (ROOT, ROOT / "visual-computing_amd"):
    if str(_p) not in sys.path:
        sys.path.insert(0, str(_p))

METRIC = "solved MPix/s @1e-8 residual, mixed vs FP64 MG-PCG; p50/p99 request latency"
HBM_PEAK_GBS = 8000.0
SHAPE = (1024, 1024)
CHANNELS = 3


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_inputs():
    from gdpoisson import workloads as W

    # seeds 2000 + channel (configuration 2, problem 0)
    return W.rgb_neumann_problem(SHAPE, config=2, problem=0)


def time_solves(plan, bs, steps, warmup):
    import torch

    outs = [torch.empty_like(b) for b in bs]
    for _ in range(warmup):
        plan.mgcg(bs, outs)
    torch.cuda.synchronize()
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        s = time.perf_counter()
        _, reps = plan.mgcg(bs, outs)  # returns once its stream has finished the solve
        lat.append(time.perf_counter() - s)
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    return total, lat, outs, reps


def pct(xs, q):
    xs = sorted(xs)
    k = min(len(xs) - 1, max(0, int(round(q / 100.0 * (len(xs) - 1)))))
    return xs[k]


def _cpu_solves(bs, runs: int, limit):
    """Median seconds of ``runs`` oracle RGB solves after one warm-up, with the thread
    pools of numpy / scipy limited to ``limit`` (None: the libraries' defaults)."""
    from threadpoolctl import threadpool_limits

    from oracle import solvers as OS

    times, its = [], []
    with threadpool_limits(limits=limit):
        for k in range(runs + 1):
            t0 = time.perf_counter()
            for b in bs:
                _, it, _ = OS.mgcg(b, bc="neumann")
                its.append(it)
            if k:
                times.append(time.perf_counter() - t0)
    return statistics.median(times), max(its)


def cpu_baseline(runs: int = 5):
    """The oracle's FP64 MG-PCG (Neumann) on the whole RGB workload: one core (the
    reference is effectively single-threaded), and once more with the libraries' default
    thread pools.  Median of ``runs`` timed solves after one untimed warm-up (about 2 s
    each here)."""
    from threadpoolctl import threadpool_info

    from oracle import generators as G

    bs = [G.neumann_problem(SHAPE, seed=2000 + c)[0] for c in range(CHANNELS)]
    dt, its = _cpu_solves(bs, runs, 1)
    dt_def, _ = _cpu_solves(bs, max(1, runs // 2), None)
    pools = sorted({f"{i.get('internal_api')}:{i.get('num_threads')}" for i in threadpool_info()})
    return {
        "value": round(SHAPE[0] * SHAPE[1] / 1e6 / dt, 4),
        "unit": "MPix/s",
        "cores": 1,
        "kind": "port",
        "sample": f"the whole 1024x1024 RGB workload with the numpy FP64 MG-PCG oracle "
        f"({its} iterations per channel), median of {runs} solves after one warm-up, "
        f"{dt:.2f} s per RGB solve, {os.cpu_count()} host cores visible, 1 used",
        "default_threads": {
            "value": round(SHAPE[0] * SHAPE[1] / 1e6 / dt_def, 4),
            "seconds_per_solve": round(dt_def, 3),
            "thread_pools": pools,
            "note": "the same solves with numpy/scipy's default thread pools; elementwise "
                    "numpy, scipy CSR products and pocketfft run one thread either way",
        },
    }


COPY_BYTES = 64 << 23  # the --profile-run calibration copy: 512 MiB read + 512 MiB written


def calibration_copy():
    """A plain device copy of COPY_BYTES (read and written once each): the reference
    dispatch that scales the FETCH_SIZE / WRITE_SIZE counters of a profile run."""
    import torch

    src = torch.empty(COPY_BYTES // 8, dtype=torch.float64, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()


SERVE_MIX = ("70% RGB seamless compositing (squares of sides 512..2048 in steps of 256, "
             "blob masks, rectangle route), 30% tone mapping 3840x2160 (log-normal radiance, "
             "single-scale attenuation (|grad|/alpha)^(beta-1), exact DCT integration)")


def serve_stream(count: int, qps_points, slots: int) -> dict:
    """The mixed request stream at each arrival rate of ``qps_points`` (Poisson arrivals),
    served by one continuous-batching Server with ``slots`` streams: its plans stay
    resident across the runs, their PCG iterations replay captured hipGraphs, and the tile
    shapes of every plane shape seen at warm-up come from the autotuner.  Returns latency
    and batching statistics per rate."""
    import numpy as np

    from gdpoisson import workloads as W
    from gdpoisson.serving import Server, latency_stats

    srv = Server(slots=slots, keep_outputs=False)
    warm = W.serving_stream(count=48, qps=1e4, seed=4999)
    log("serving: tuning tiles of the warm-up shapes")
    srv.tune(warm, solves=1)
    log("serving: warm-up stream")
    srv.run(warm, realtime=False)  # first-use costs (plans, graphs) outside the timed streams
    points = []
    for k, qps in enumerate(qps_points):
        log(f"serving stream: {count} requests at {qps:g}/s, {slots} slots")
        reqs = W.serving_stream(count=count, qps=qps, seed=5000 + k)
        nb0, hits0, builds0 = len(srv.batches), srv.plan_hits, srv.plan_builds
        done = srv.run(reqs)
        st = latency_stats(done)
        lat = {}
        for c in done:
            lat.setdefault(c.kind, []).append(c.latency * 1e3)
        batches = srv.batches[nb0:]
        points.append({
            "qps": qps,
            **{k2: (round(v, 3) if isinstance(v, float) else v) for k2, v in st.items()},
            "p50_ms_by_kind": {k2: round(float(np.percentile(v, 50)), 3) for k2, v in lat.items()},
            "p99_ms_by_kind": {k2: round(float(np.percentile(v, 99)), 3) for k2, v in lat.items()},
            "batches": len(batches),
            "mean_requests_per_batch": round(sum(n for _, n, _ in batches) / max(len(batches), 1), 2),
            "max_requests_per_batch": max((n for _, n, _ in batches), default=0),
            "plan_hits": srv.plan_hits - hits0,
            "plan_builds": srv.plan_builds - builds0,
            "all_converged": all(r.converged and r.relative_residual <= 1e-8
                                 for c in done for r in c.reports),
        })
    st = srv.stats()
    srv.close()
    return {
        "requests_per_point": count,
        "slots": slots,
        "mix": SERVE_MIX,
        "method": srv.method,
        "tonemap_method": srv.tonemap_method,
        "graphs": st["graphs"],
        "tuned_shapes": st["tuned_shapes"],
        "resident_plan_peak_GB": st["resident_plan_peak_GB"],
        "budget_GB": round(srv.total_budget / 1e9, 1),
        "peak_admitted_GB": round(srv.peak_bytes / 1e9, 2),
        "points": points,
    }


def composite_leg(steps: int) -> dict:
    """2048 x 2048 RGB seamless compositing of a blob-masked source (about 30% of the
    pixels selected), mixed MG-PCG to 1e-8, by both routes: the rectangle (Dirichlet over
    the mask's bounding box, poisson.py:337-347) and the mask domain (the unknowns are the
    selected pixels, poisson.py:154-244).  Time per composite from device-resident images
    to the device-resident result, right-hand side and plan set-up included."""
    import torch

    import gdpoisson
    from gdpoisson import workloads as W

    side = 2048
    src, tgt, mask = W.compositing_images((side, side))
    s, t, mk = (torch.from_numpy(x).cuda() for x in (src, tgt, mask))
    out = {"image": [side, side, 3], "selected": round(float(mask.mean()), 4), "steps": steps}
    for domain in ("rectangle", "mask"):
        log(f"compositing ({domain} route)")
        gdpoisson.seamless_clone(s, t, mk, (0, 0), domain=domain, method="mgcg_mixed")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            _, reps = gdpoisson.seamless_clone(s, t, mk, (0, 0), domain=domain,
                                               method="mgcg_mixed")
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        out[domain] = {
            "ms_per_composite": round(dt * 1e3, 2),
            "MPix_per_s": round(side * side / 1e6 / dt, 1),
            "iterations": [int(r.iterations) for r in reps],
            "max_relative_residual": float(f"{max(r.relative_residual for r in reps):.3e}"),
            "converged": all(r.converged and r.relative_residual <= 1e-8 for r in reps),
        }
    return out


def _family(base: str) -> str | None:
    """The level-kernel family of a dispatch: the sweep-order, streamed and block-tiled down
    kernels are one family (they run the same levels), ranked together by level area, and
    so are the strip and block-tiled up kernels."""
    if "k_sdown<" in base or "k_sdown_stream<" in base or "k_sdown_blk<" in base:
        return "k_sdown<"
    if "k_sup_blk<" in base:
        return "k_sup<"
    for key in ("k_sup<", "k_down<", "k_up<"):
        if key in base:
            return key
    return None


def _level_area(name: str, grid: int) -> int:
    """Grid size x rows per strip (the third template argument of the strip kernels; the
    block-tiled kernels' is rows per block, four strips' worth): proportional to the
    level's area whatever tile shape the level was given."""
    base = name.split("(")[0]
    if _family(base) in ("k_sdown<", "k_sup<"):
        args = base[base.index("<") + 1:].split(",")
        rows = int(args[2].strip(" >"))
        return grid * rows // 4 if "_blk<" in base else grid * rows
    return grid


def kernel_kind(name: str, grid: int, grids: dict) -> str | None:
    """The plan profiler's kind name of a rocprofv3 dispatch of this workload.

    Levels are ranked per kernel family by grid size x rows per strip (largest = finest).
    """
    base = name.split("(")[0]
    fam = _family(base)
    if fam is not None:
        kind = "down" if fam in ("k_sdown<", "k_down<") else "up"
        # k_sup<T, nu, rows, dots, masked>; k_sdown_stream<T, nu, rows, upd> and
        # k_sdown_blk<T, nu, rows, upd> (upd: the finest level carrying the PCG update)
        args = [a.strip(" >") for a in base[base.index("<") + 1:].split(",")]
        dots = kind == "up" and len(args) > 3 and args[3] == "true"
        upd = (("k_sdown_stream<" in base or "k_sdown_blk<" in base) and len(args) > 3
               and args[3] == "true")
        lv = sorted(grids[fam], reverse=True).index(_level_area(name, grid))
        return f"{kind} L{lv}" + ("+dots" if dots else "") + ("+update" if upd else "")
    if "k_update<" in base:
        # k_update<T, TH, TW, defer_x>: the x half left to the tail launch or the next apply
        return "update (x deferred)" if base.rstrip(" >").endswith("true") else "update"
    for key, kind in (("k_stail<", "tail"), ("k_tail<", "tail"),
                      ("k_apply_s<", "apply"), ("k_init<", "init"),
                      ("k_residual<", "residual"), ("k_finalize<", "finalize"),
                      ("k_init_stats<", "init_stats")):
        if key in base:
            return kind
    return None


def reduce_traffic(fetch_csv: str, write_csv: str) -> dict:
    """Per-kernel HBM bytes per launch from FETCH_SIZE / WRITE_SIZE (KiB per dispatch).

    Each counter is scaled by the ratio of the known bytes of the plain device copy in
    the same run (__amd_rocclr_copyBuffer of 512 MiB) to what the counter reads for it.
    Dispatches that exited early (converged problems) are left out: the mean is over
    dispatches whose count is at least half the kernel's maximum.
    """
    import csv
    import collections

    out = {"calibration": {}, "kernels": {}}
    per = collections.defaultdict(dict)
    for counter, path in (("FETCH_SIZE", fetch_csv), ("WRITE_SIZE", write_csv)):
        rows = [r for r in csv.DictReader(open(path)) if r["Counter_Name"] == counter]
        copy = [float(r["Counter_Value"]) * 1024 for r in rows
                if "copyBuffer" in r["Kernel_Name"] and int(r["Grid_Size"]) >= 1 << 16]
        if not copy:
            raise SystemExit(f"{path}: no calibration copy dispatch found")
        scale = COPY_BYTES / max(copy)
        out["calibration"][counter] = {"copy_bytes": COPY_BYTES, "counter_bytes": max(copy),
                                       "scale": round(scale, 4)}
        grids = collections.defaultdict(set)
        for r in rows:
            fam = _family(r["Kernel_Name"].split("(")[0])
            if fam is not None:
                grids[fam].add(_level_area(r["Kernel_Name"], int(r["Grid_Size"])))
        # grouped by instantiation and launch shape: a kind with two of them (the tail
        # with and without the deferred x update) averages over both, as its algorithmic
        # bytes per launch do; the early-exit filter then compares each dispatch only with
        # its own variant
        vals = collections.defaultdict(lambda: collections.defaultdict(list))
        for r in rows:
            kind = kernel_kind(r["Kernel_Name"], int(r["Grid_Size"]), grids)
            if kind:
                variant = (r["Kernel_Name"].split("(")[0], int(r["Grid_Size"]))
                vals[kind][variant].append(float(r["Counter_Value"]) * 1024 * scale)
        for kind, by_grid in vals.items():
            full = []
            for v in by_grid.values():
                full += [x for x in v if x >= 0.5 * max(v)] or v
            per[kind][counter] = sum(full) / len(full)
    for kind, d in per.items():
        f, w = d.get("FETCH_SIZE", 0.0), d.get("WRITE_SIZE", 0.0)
        out["kernels"][kind] = {"fetch_bytes": round(f), "write_bytes": round(w),
                                "hbm_bytes_per_launch": round(f + w)}
    return out


def _short(name: str) -> str:
    import re

    return re.sub(r"\(.*", "", name).replace("void ", "").replace("gdp::", "")


def trace_summary(kt_dir: str) -> str:
    """Markdown of a rocprofv3 --kernel-trace --stats run of --profile-run: the per-kernel
    statistics and the timeline of one solve (the sixth of seven)."""
    import csv

    d = Path(kt_dir)
    stats = list(csv.DictReader(open(next(d.glob("*kernel_stats.csv")))))
    trace = sorted(csv.DictReader(open(next(d.glob("*kernel_trace.csv")))),
                   key=lambda r: int(r["Start_Timestamp"]))
    out = ["## Kernel statistics (rocprofv3 --stats)", "",
           "| kernel | calls | total us | avg us | min us | max us | % |",
           "|---|---|---|---|---|---|---|"]
    for r in stats:
        out.append(f"| `{_short(r['Name'])[:72]}` | {r['Calls']} | "
                   f"{float(r['TotalDurationNs']) / 1e3:.1f} | {float(r['AverageNs']) / 1e3:.2f} | "
                   f"{float(r['MinNs']) / 1e3:.2f} | {float(r['MaxNs']) / 1e3:.2f} | "
                   f"{float(r['Percentage']):.2f} |")
    starts = [i for i, r in enumerate(trace) if "k_init<" in r["Kernel_Name"]]
    if len(starts) >= 2:
        a, b = starts[-2], starts[-1]
        t0 = int(trace[a]["Start_Timestamp"])
        out += ["", "## One solve from the kernel trace: start offset, duration, grid", "",
                "| start us | dur us | grid | kernel |", "|---|---|---|---|"]
        for r in trace[a:b]:
            st = int(r["Start_Timestamp"])
            out.append(f"| {(st - t0) / 1e3:.1f} | {(int(r['End_Timestamp']) - st) / 1e3:.2f} | "
                       f"{r['Grid_Size_X']} | `{_short(r['Kernel_Name'])[:64]}` |")
    return "\n".join(out) + "\n"


def traffic_md(fetch_csv: str, write_csv: str, bench_json: str, key: str = "kernels") -> str:
    """Markdown table: algorithmic bytes per launch (the bench line's kernel table at
    ``key``, e.g. "batched.kernels" for the 64-problem leg) against the HBM bytes the two
    counter passes measured."""
    t = reduce_traffic(fetch_csv, write_csv)
    k = json.loads(Path(bench_json).read_text())
    for part in key.split("."):
        k = k[part]
    cal = t["calibration"]
    out = [f"Calibration (the 512 MiB device copy at the end of `--profile-run`): FETCH_SIZE "
           f"scale {cal['FETCH_SIZE']['scale']}, WRITE_SIZE scale {cal['WRITE_SIZE']['scale']}.",
           "", "| kernel | algorithmic B/launch | measured B/launch (fetch + write) | "
           "measured / algorithmic | avg us (events) |", "|---|---|---|---|---|"]
    for name, v in t["kernels"].items():
        if name not in k or not k[name]["GBps"]:
            continue
        alg = k[name]["GBps"] * 1e9 * k[name]["avg_us"] * 1e-6
        out.append(f"| {name} | {alg / 1e6:.2f} MB | {v['hbm_bytes_per_launch'] / 1e6:.2f} MB | "
                   f"{v['hbm_bytes_per_launch'] / alg:.2f} | {k[name]['avg_us']} |")
    return "\n".join(out) + "\n"


def batch_inputs():
    """The batched workload: 64 RGB Neumann problems, sides log-uniform in [128, 2048],
    aspect U[0.5, 2] (192 planes); returns (plane shapes, device right-hand sides, exact
    zero-mean solutions on the host)."""
    from gdpoisson import workloads as W

    shapes = W.batch_shapes()
    plane_shapes, bs, exact = [], [], []
    for p, s in enumerate(shapes):
        for c in range(3):
            t = W.blob_image(s, 4000 + 10 * p + c)
            bs.append(W.neumann_rhs(t))
            exact.append(t - t.mean())
            plane_shapes.append(s)
    return shapes, plane_shapes, bs, exact


def batch_leg(steps: int, tune: bool = True) -> dict:
    """The batched configuration (64 RGB problems of uneven sizes, 192 planes, one mixed
    plan): solved MPix/s, and the per-kernel HBM rates of one profiled solve over a working
    set far larger than the 256 MB Infinity Cache, so the rates are HBM rates."""
    import numpy as np
    import torch

    import gdpoisson

    log("batched 64-problem workload: building inputs")
    shapes, plane_shapes, bs, exact = batch_inputs()
    plan = gdpoisson.Plan(plane_shapes, boundary="neumann",
                          options=gdpoisson.PlanOptions(precision="mixed"))
    tiles = None
    if tune:
        log("batched 64-problem workload: autotuning tile shapes")
        kinds = ["down", "up", "krylov"]
        ch = gdpoisson.autotune(plan, bs, solves=1)
        tiles = {f"{kinds[k]}{l}": v for (k, l), v in sorted(ch.items())}
    outs = [torch.empty_like(b) for b in bs]
    log("batched 64-problem workload: timing")
    plan.mgcg(bs, outs)  # first use: graph capture
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        _, reps = plan.mgcg(bs, outs)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    err = max(float(np.abs(u.cpu().numpy() - x).max()) for u, x in zip(outs, exact))
    mpix = sum(a * b for a, b in shapes) / 1e6
    points = sum(a * b for a, b in plane_shapes)
    plan.profile(True)
    plan.mgcg(bs, outs)
    prof = plan.profile_report()
    plan.profile(False)
    total_ms = sum(v["ms"] for v in prof.values())
    kernels = {
        k: {"avg_us": round(v["ms"] / v["launches"] * 1e3, 2), "launches": v["launches"],
            "GBps": round(v["bytes"] / (v["ms"] / 1e3) / 1e9, 1) if v["ms"] > 0 else None,
            "frac": round(v["bytes"] / (v["ms"] / 1e3) / 1e9 / HBM_PEAK_GBS, 3)
            if v["ms"] > 0 else None,
            "share": round(v["ms"] / total_ms, 3)}
        for k, v in sorted(prof.items(), key=lambda kv: -kv[1]["ms"])[:12]
    }
    # the last update of a batch of iterations runs after planes that converged at fewer
    # iterations have finished (they exit at once), so its algorithmic bytes overstate
    # what it moves: no rate for it
    if "update" in kernels:
        kernels["update"].update(GBps=None, frac=None,
                                 note="last update of a batch of iterations: planes that "
                                      "finished earlier exit at once")
    # HBM bytes per launch measured by the FETCH_SIZE / WRITE_SIZE passes of this workload
    tfile = ROOT / "profiles" / "traffic_batch64_rgb_mixed.json"
    if tfile.exists():
        meas = json.loads(tfile.read_text()).get("kernels", {})
        for k, v in kernels.items():
            if k in meas and v["GBps"]:
                alg = v["GBps"] * 1e9 * v["avg_us"] * 1e-6
                v["traffic"] = meas[k]["hbm_bytes_per_launch"]
                v["traffic_vs_algorithmic"] = round(meas[k]["hbm_bytes_per_launch"] / alg, 3)
    ws = plan.workspace_bytes + 2 * 8 * points
    return {
        "workload": "64 RGB Neumann problems, sides log-uniform in [128, 2048], aspect "
                    "U[0.5, 2], 192 planes in one mixed plan, to 1e-8 true residual",
        "MPix": round(mpix, 2),
        "unknowns": points,
        "ms_per_solve": round(dt * 1e3, 2),
        "MPix_per_s": round(mpix / dt, 1),
        "iterations": [min(r.iterations for r in reps), max(r.iterations for r in reps)],
        "max_relative_residual": float(f"{max(r.relative_residual for r in reps):.3e}"),
        "max_abs_err_vs_exact": float(f"{err:.3e}"),
        "converged": all(r.converged and r.relative_residual <= 1e-8 for r in reps),
        "tiles": tiles,
        "working_set_bytes": int(ws),
        "working_set_vs_infinity_cache": round(ws / (256 << 20), 1),
        "kernels": kernels,
    }


def cold_first_solve(nu: int) -> dict:
    """Latency of the first solve of a new image: a new plan (created, workspace
    allocated and bound, graphs captured on first use) solving right-hand sides it has
    not seen, against the same plan's warm re-solve of other new inputs."""
    import torch

    import gdpoisson
    from gdpoisson import workloads as W

    bs1, _ = W.rgb_neumann_problem(SHAPE, config=2, problem=7)
    bs2, _ = W.rgb_neumann_problem(SHAPE, config=2, problem=8)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan = gdpoisson.Plan([SHAPE] * CHANNELS, boundary="neumann",
                          options=gdpoisson.PlanOptions(precision="mixed", pre_smoothing=nu,
                                                        post_smoothing=nu))
    plan.mgcg(bs1)
    torch.cuda.synchronize()
    cold = time.perf_counter() - t0
    t0 = time.perf_counter()
    plan.mgcg(bs2)
    torch.cuda.synchronize()
    new_rhs = time.perf_counter() - t0
    return {"cold_first_solve_ms": round(cold * 1e3, 3),
            "first_solve_of_new_inputs_on_a_warm_plan_ms": round(new_rhs * 1e3, 3),
            "of": "1024x1024 RGB mixed: plan creation + workspace + bind + graph capture + "
                  "solve, then a warm plan's first solve of other inputs"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-tune", action="store_true", help="keep the default tile shapes")
    ap.add_argument("--tiles", metavar="BENCH_JSON",
                    help="reuse the tile shapes recorded in a bench line (config.tiles) instead "
                    "of tuning, so a profiled run has exactly the benched configuration")
    ap.add_argument("--trace-summary", metavar="KT_DIR",
                    help="markdown summary of a rocprofv3 --kernel-trace --stats output folder")
    ap.add_argument("--traffic-md", nargs=3, metavar=("FETCH_CSV", "WRITE_CSV", "BENCH_JSON"),
                    help="markdown table of measured vs algorithmic bytes per kernel")
    ap.add_argument("--serve-requests", type=int, default=120,
                    help="requests of the mixed serving stream (0: skip it)")
    ap.add_argument("--composite-steps", type=int, default=3,
                    help="2048x2048 RGB composites timed per route (0: skip them)")
    ap.add_argument("--serve-qps", type=float, nargs="+", default=[100.0, 200.0, 400.0, 800.0],
                    help="arrival rates of the stream (the p50/p99 of the line is the 200/s point when "
                    "listed, else the first)")
    ap.add_argument("--serve-slots", type=int, default=4,
                    help="serving workers (HIP streams; four hardware queues per process)")
    ap.add_argument("--nu", type=int, default=2,
                    help="pre/post smoothing sweeps of the mixed solver's V-cycle")
    ap.add_argument("--traffic-key", default="kernels",
                    help="where --traffic-md finds the kernel table in BENCH_JSON "
                    "(batched.kernels for the 64-problem leg)")
    ap.add_argument("--traffic", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="reduce two rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE counter_collection "
                    "CSVs of a --profile-run into per-kernel HBM bytes per launch (JSON on stdout)")
    ap.add_argument("--profile-run", nargs="?", const="rgb", choices=("rgb", "batch"),
                    help="only run a few mixed solves of the 1024x1024 RGB workload (rgb) or "
                    "of the 64-problem batch (batch) and a plain device copy of 512 MiB (the "
                    "FETCH_SIZE / WRITE_SIZE calibration) for rocprofv3; no JSON line")
    ap.add_argument("--batch-steps", type=int, default=3,
                    help="timed solves of the 64-problem batched workload (0: skip)")
    args = ap.parse_args()
    if args.trace_summary:
        print(trace_summary(args.trace_summary), end="")
        return
    if args.traffic_md:
        print(traffic_md(*args.traffic_md, key=args.traffic_key), end="")
        return
    if args.traffic:
        print(json.dumps(reduce_traffic(*args.traffic), indent=1))
        return
    if args.gpus != 1:
        print(
            f"bench.py measures one GPU only (got --gpus {args.gpus}): the Poisson path has "
            "no multi-GPU code in this build; problems shard by whole images with no exchange",
            file=sys.stderr,
        )
        sys.exit(2)

    import numpy as np
    import torch

    import gdpoisson

    torch.cuda.set_device(0)
    if args.profile_run == "batch":
        _, plane_shapes, bs, _ = batch_inputs()
        plan = gdpoisson.Plan(plane_shapes, boundary="neumann",
                              options=gdpoisson.PlanOptions(precision="mixed"))
        if args.tiles:  # the tiles the bench line's batched leg tuned
            rec = json.loads(Path(args.tiles).read_text()).get("batched", {}).get("tiles") or {}
            kinds = ["down", "up", "krylov"]
            plan.set_tiles({(kinds.index(k.rstrip("0123456789")),
                             int(k[len(k.rstrip("0123456789")):])): v for k, v in rec.items()})
        for _ in range(1 + args.steps):
            plan.mgcg(bs)
        calibration_copy()
        log("profile run done")
        return
    log("building inputs")
    bs, exact = make_inputs()
    mixed = gdpoisson.Plan([SHAPE] * CHANNELS, boundary="neumann",
                           options=gdpoisson.PlanOptions(precision="mixed",
                                                         pre_smoothing=args.nu,
                                                         post_smoothing=args.nu))
    # plain FP64 MG-PCG: the reference's mgcg cycle (nu = 1/1, solvers.py:137) in FP64
    fp64 = gdpoisson.Plan([SHAPE] * CHANNELS, boundary="neumann",
                          options=gdpoisson.PlanOptions(precision="fp64"))
    # the same cycle as the mixed solver, all in FP64
    fp64_same = gdpoisson.Plan([SHAPE] * CHANNELS, boundary="neumann",
                               options=gdpoisson.PlanOptions(precision="fp64",
                                                             pre_smoothing=args.nu,
                                                             post_smoothing=args.nu))

    # tile shapes picked by the autotuner for exactly this workload (outside the timed
    # region; cached per plan configuration)
    tiles = {}
    plans = (("mixed", mixed), ("fp64", fp64), ("fp64_same_cycle", fp64_same))
    kinds = ["down", "up", "krylov"]
    if args.tiles:
        recorded = json.loads(Path(args.tiles).read_text())["config"]["tiles"]
        for name, plan in plans:
            sel = {(kinds.index(k.rstrip("0123456789")), int(k[len(k.rstrip("0123456789")):])): v
                   for k, v in recorded.get(name, {}).items()}
            plan.set_tiles(sel)
            tiles[name] = recorded.get(name, {})
    elif not args.no_tune:
        log("autotuning tile shapes")
        for name, plan in plans:
            ch = gdpoisson.autotune(plan, bs)
            tiles[name] = {f"{kinds[k]}{l}": v for (k, l), v in sorted(ch.items())}

    if args.profile_run:
        for _ in range(2 + args.steps):
            mixed.mgcg(bs)
        calibration_copy()
        log("profile run done")
        return

    log("timing mixed")
    t_mix, lat_mix, u_mix, rep_mix = time_solves(mixed, bs, args.steps, args.warmup)
    log("timing fp64")
    fsteps = max(3, args.steps // 3)
    t_64, lat_64, u_64, rep_64 = time_solves(fp64, bs, fsteps, max(1, args.warmup))
    t_64s, _, _, rep_64s = time_solves(fp64_same, bs, fsteps, max(1, args.warmup))

    mpix = SHAPE[0] * SHAPE[1] / 1e6
    v_mix = mpix * args.steps / t_mix
    v_64 = mpix * fsteps / t_64
    v_64s = mpix * fsteps / t_64s

    # accuracy of every channel
    err_exact, err_fp64, res = [], [], []
    for c in range(CHANNELS):
        um = u_mix[c].cpu().numpy()
        err_exact.append(float(np.abs(um - exact[c]).max()))
        err_fp64.append(float(np.abs(um - u_64[c].cpu().numpy()).max()))
        res.append(rep_mix[c].relative_residual)
    ok = all(r.converged for r in rep_mix) and max(res) <= 1e-8

    # live per-kernel timing (HIP events on the launch stream) of the same solve
    log("profiling kernels")
    mixed.profile(True)
    for _ in range(3):
        mixed.mgcg(bs, u_mix)
    prof = mixed.profile_report()
    mixed.profile(False)
    dom_name, dom = max(prof.items(), key=lambda kv: kv[1]["ms"])
    avg_s = dom["ms"] / dom["launches"] / 1e3
    bytes_per_launch = dom["bytes"] / dom["launches"]
    achieved = bytes_per_launch / avg_s / 1e9
    kernels = {
        k: {
            "avg_us": round(v["ms"] / v["launches"] * 1e3, 2),
            "launches_per_solve": round(v["launches"] / 3.0, 2),
            "GBps": round(v["bytes"] / (v["ms"] / 1e3) / 1e9, 1) if v["ms"] > 0 else None,
            "share": round(v["ms"] / sum(x["ms"] for x in prof.values()), 3),
        }
        for k, v in sorted(prof.items(), key=lambda kv: -kv[1]["ms"])
    }

    traffic = None
    tfile = ROOT / "profiles" / "traffic_1024x1024_rgb_mixed.json"
    if tfile.exists():
        t = json.loads(tfile.read_text()).get("kernels", {}).get(dom_name)
        if t:
            traffic = t["hbm_bytes_per_launch"]

    line = {
        "metric": METRIC,
        "value": round(v_mix, 2),
        "unit": "MPix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_mix / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 V-cycle + f64 PCG",
        "data": "synthetic (seeded Gaussian blobs + N(0,0.01) noise per channel)",
        "config": {
            "workload": "1024x1024 RGB Neumann Poisson, mixed FP32-MG / FP64-refined MG-PCG "
            "to 1e-8 true relative residual, 3 channels batched in one plan",
            "shape": list(SHAPE),
            "channels": CHANNELS,
            "levels": mixed.levels(0),
            "parallelism": "single GPU, channels batched per launch",
            "tiles": tiles,
        },
        # the denominator is the fastest FP64 GPU configuration measured (the same
        # kernels with the reference's nu = 1/1 cycle or the mixed solver's cycle)
        "fp64_value": round(max(v_64, v_64s), 2),
        "fp64_config": "fp64_same_cycle" if v_64s >= v_64 else "fp64_reference_cycle",
        "mixed_vs_fp64": round(v_mix / max(v_64, v_64s), 3),
        "fp64_reference_cycle_value": round(v_64, 2),
        "fp64_reference_cycle_ms_per_step": round(t_64 / fsteps * 1e3, 4),
        "mixed_vs_fp64_reference_cycle": round(v_mix / v_64, 3),
        "fp64_same_cycle_value": round(v_64s, 2),
        "mixed_vs_fp64_same_cycle": round(v_mix / v_64s, 3),
        "smoothing": {"mixed": [args.nu, args.nu], "fp64_reference_cycle": [1, 1],
                      "fp64_same_cycle": [args.nu, args.nu]},
        "iterations": {"mixed": [r.iterations for r in rep_mix],
                       "fp64_reference_cycle": [r.iterations for r in rep_64],
                       "fp64_same_cycle": [r.iterations for r in rep_64s]},
        "relative_residual": [float(f"{r:.3e}") for r in res],
        "converged": ok,
        "max_abs_err_vs_exact": float(f"{max(err_exact):.3e}"),
        "max_abs_err_vs_fp64": float(f"{max(err_fp64):.3e}"),
        "latency_ms": {"p50": round(pct(lat_mix, 50) * 1e3, 4),
                       "p99": round(pct(lat_mix, 99) * 1e3, 4),
                       "fp64_p50": round(pct(lat_64, 50) * 1e3, 4)},
        "roofline": {
            "bound": "hbm",
            "kernel": dom_name,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "bytes_per_launch": int(bytes_per_launch),
            "avg_launch_us": round(avg_s * 1e6, 2),
        },
        "kernels": kernels,
    }
    log("cold first solve")
    line["cold_latency"] = cold_first_solve(args.nu)
    if args.composite_steps > 0:
        line["compositing"] = composite_leg(args.composite_steps)
    if args.batch_steps > 0:
        line["batched"] = batch_leg(args.batch_steps, tune=not args.no_tune)
    if args.serve_requests > 0:
        line["serving"] = serve_stream(args.serve_requests, args.serve_qps, args.serve_slots)
        head_qps = 200.0 if 200.0 in args.serve_qps else args.serve_qps[0]
        head = next(p for p in line["serving"]["points"] if p["qps"] == head_qps)
        line["latency_ms"] = {
            "p50": head["p50_ms"], "p99": head["p99_ms"],
            "of": f"requests of the mixed serving stream at {head['qps']:g}/s "
                  "(arrival to completion)",
            "solve_p50": line["latency_ms"]["p50"], "solve_p99": line["latency_ms"]["p99"],
            "solve_fp64_p50": line["latency_ms"]["fp64_p50"]}
    if not args.no_cpu:
        log("CPU oracle baseline")
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
