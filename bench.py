"""Benchmark: temporal set inversion on one MI355X.

One step = one frame of the plain per-frame path on the 64-capsule workload (seeded
smooth-min union of 64 capsules translating linearly in t under a twist warp of angle
2 t z; localsubdiv 9 = 512^3 cells; 1024x1024 frame; delta 0.01; Lipschitz correction on):
set inversion (tsi_set_inversion) + trace of the frame (tsi_trace) + root re-enqueue,
at t_k = k/60.  The first W frames are untimed warm-up (frame 0 evaluates every cell);
the next K frames are timed with the device synchronised on both sides.

Prints one JSON line.  `value` = interval cell-evals per second over the timed frames;
`roofline` reports the tape-evaluation kernels against the FP32 VALU peak; `cpu_baseline`
times the oracle (the C restatement of the reference) on a bounded sample of the same tape.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X FP32 vector issue: 256 CU x 4 SIMD x 16 lanes x 4 (wave64 over 4 cycles) ... expressed as
# lane-results: 157.3 TFLOP/s counts an FMA as 2 and assumes packed dual-FP32 issue, i.e.
# 78.6 T f32 results/s; without packed issue 39.3 T/s.
PEAK_VALU_TOPS = 78.6
PEAK_VALU_UNPACKED_TOPS = 39.3
PEAK_HBM_GBS = 8000.0

WORKLOADS = {
    # plain method on one MI355X: 512^3, 1024x1024, 64 capsules
    "capsules64_L9_1024": dict(scene="capsules", seed=1, localsubdiv=9, width=1024, height=1024,
                               lipschitz_corrected=True, swept_volume=False, point_sample=True),
    # plain method, large: 1024^3, 1920x1080, 256-primitive CSG, swept, point-sampled
    "csg256_L10_1920x1080": dict(scene="csg", seed=2, localsubdiv=10, width=1920, height=1080,
                                 lipschitz_corrected=False, swept_volume=True, point_sample=True),
    # the 8-moving-spheres field at 64^3
    "spheres8_L6_256": dict(scene="spheres", seed=0, localsubdiv=6, width=256, height=256,
                            lipschitz_corrected=False, swept_volume=False, point_sample=True),
}


def profiled_traffic(workload: str, engine: str):
    """HBM bytes per eval launch from the committed rocprofv3 summary of this workload, if any
    (FETCH_SIZE/WRITE_SIZE passes scaled by the copy calibration; see profiles/)."""
    path = os.path.join(ROOT, "profiles", f"tape_eval_{workload.split('_')[0]}_{workload.split('_', 1)[1]}.txt")
    if engine != "compiled" or not os.path.exists(path):
        return None, None, None
    rows, times = {}, {}
    in_traffic = False
    for line in open(path):
        parts = line.split()
        in_traffic = in_traffic or "hbm_bytes/disp" in line
        if in_traffic and len(parts) == 5 and parts[0].startswith("tsi_tape_eval_"):
            rows[parts[0]] = (int(parts[1]), float(parts[4]))
        elif not in_traffic and len(parts) == 5 and parts[0].startswith("tsi_tape_eval_"):
            times[parts[0]] = (int(parts[1]), float(parts[3]))   # dispatches, avg_us
    if not rows:
        return None, None, None
    n = sum(k for k, _ in rows.values())
    m = sum(k for k, _ in times.values())
    avg_us = sum(k * t for k, t in times.values()) / m if m else None
    return sum(k * b for k, b in rows.values()) / n, os.path.relpath(path, ROOT), avg_us


def _short(name: str) -> str:
    return name.split("(")[0]


def summarize_profile(args) -> None:
    """Prints the committed profile summary (profiles/tape_eval_<workload>.txt) from the
    rocprofv3 databases of the kernel-trace run, the FETCH_SIZE / WRITE_SIZE runs of the same
    workload and the same two counters over a 256 MiB device copy (the scale that turns the
    counters into bytes on this machine)."""
    import sqlite3

    def per_kernel(db, sql):
        out = {}
        for name, n, v in sqlite3.connect(db).execute(sql):
            k = _short(name)
            m, t = out.get(k, (0, 0.0))
            out[k] = (m + n, t + v)
        return out

    kt = per_kernel(args.summarize, "select name, count(*), sum(duration) from kernels group by name")
    total = sum(t for _, t in kt.values())
    frames = kt.get("tsi::k_frame_begin", (0, 0))[0]
    W, R = args.workload, "$R"
    tag = {"fetch": args.pmc_fetch, "write": args.pmc_write}
    lines = [f"# Kernel time and HBM traffic: {W}, compiled tape engine, bench defaults",
             "#",
             "# commands (run from /tmp on the GPU host, repo at $R; <d> = an output directory):",
             f"#   rocprofv3 --kernel-trace --stats -d <d>/kt -o kt -- python3 {R}/bench.py --steps 5 --warmup 2 --no-cpu --workload {W}"]
    for c in ("FETCH_SIZE", "WRITE_SIZE"):
        lines.append(f"#   rocprofv3 --pmc {c} -d <d>/{c} -o pmc -- python3 {R}/bench.py --steps 2 --warmup 1 --no-cpu --workload {W}")
        lines.append(f"#   rocprofv3 --pmc {c} -d <d>/copy_{c} -o pmc -- python3 {R}/bench.py --calibrate-copy-mib {args.copy_mib}")
    lines.append(f"#   python3 bench.py --workload {W} --summarize <d>/kt/kt_results.db --pmc-fetch <d>/FETCH_SIZE/pmc_results.db "
                 "--pmc-write <d>/WRITE_SIZE/pmc_results.db --copy-fetch <d>/copy_FETCH_SIZE/pmc_results.db "
                 "--copy-write <d>/copy_WRITE_SIZE/pmc_results.db")
    lines += ["#", f"# {frames} frames (warm-up + timed); one level kernel per octree level running the forward-only",
              "# and the reverse list side by side (tsi_tape_eval_level), one snapshot + trace per frame on",
              "# a trace stream.", ""]
    lines.append(f"{'kernel':<44} {'dispatches':>10} {'total_ms':>10} {'avg_us':>9} {'share':>7}")
    for k, (n, t) in sorted(kt.items(), key=lambda x: -x[1][1]):
        lines.append(f"{k[:44]:<44} {n:>10} {t / 1e6:>10.3f} {t / n / 1e3:>9.2f} {100 * t / total:>6.1f}%")
    if args.pmc_fetch and args.pmc_write and args.copy_fetch and args.copy_write:
        q = "select name, count(*), sum(counter_value) from pmc_events group by name"
        kib = args.copy_mib * 1024.0

        def scale(db):   # KiB the counter reports per copy dispatch -> bytes
            n, v = per_kernel(db, q)["__amd_rocclr_copyBuffer"]
            return kib / (v / n)

        sf, sw = scale(args.copy_fetch), scale(args.copy_write)
        f, w = per_kernel(args.pmc_fetch, q), per_kernel(args.pmc_write, q)
        lines += ["", f"copy calibration ({args.copy_mib} MiB device-to-device): FETCH_SIZE scale {sf:.3f}, "
                      f"WRITE_SIZE scale {sw:.3f} (bytes = KiB x scale x 1024)",
                  f"{'kernel':<24} {'dispatches':>10} {'fetch_KiB/disp':>15} {'write_KiB/disp':>15} {'hbm_bytes/disp':>15}"]
        for k in ("tsi_tape_eval_level", "tsi::k_trace", "tsi::k_snapshot"):
            if k in f and k in w:
                nf, vf = f[k]
                nw, vw = w[k]
                b = (vf / nf * sf + vw / nw * sw) * 1024.0
                lines.append(f"{k:<24} {nf:>10} {vf / nf:>15.1f} {vw / nw:>15.1f} {b:>15.0f}")
    print("\n".join(lines))


def _cpu_worker(job):
    """One oracle frame sequence (frames t_k = (k + phase)/60) until the time budget is spent."""
    import oracle

    ll, w, L, phase, budget_s = job
    tape = oracle.parse_ll(ll)
    fr = oracle.Frame(tape, localsubdiv=L, swept_volume=w["swept_volume"],
                      lipschitz_corrected=w["lipschitz_corrected"], point_sample=w["point_sample"],
                      voxel_capacity=1 << 22)
    cells = frames = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s and frames < 240:
        st = fr.set_inversion(float(np.float32((frames + phase) / 60.0)))
        cells += st.cell_evals
        frames += 1
        fr.enqueue_root()
    dt = time.perf_counter() - t0
    fr.close()
    return cells, frames, dt


def cpu_baseline(ll, w, budget_s: float):
    """The oracle (C restatement of compute.wgsl `main`) on the host cores: one independent frame
    sequence per worker process (phases 7 frames apart), same tape and parameters at localsubdiv
    L-3, each until the time budget is spent.  Runs before the GPU is initialised."""
    import multiprocessing as mp

    L = max(3, w["localsubdiv"] - 3)
    workers = max(1, min(16, os.cpu_count() or 1))
    ctx = mp.get_context("fork")
    t0 = time.perf_counter()
    with ctx.Pool(workers) as pool:
        res = pool.map(_cpu_worker, [(ll, w, L, 7 * i, budget_s) for i in range(workers)])
    wall = time.perf_counter() - t0
    cells = sum(r[0] for r in res)
    frames = sum(r[1] for r in res)
    return {"value": cells / wall, "unit": "cell-evals/s", "cores": workers, "kind": "port",
            "sample": f"oracle (C restatement) on {workers} host cores, one frame sequence per core, same tape "
                      f"and parameters at localsubdiv {L}, {frames} frames, {cells} cell-evals in {wall:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="capsules64_L9_1024", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds per CPU-baseline worker")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--engine", default="compiled", choices=["compiled", "interpreter"],
                    help="tape evaluation: compiled straight-line kernels (default) or the interpreter")
    ap.add_argument("--sync-trace", action="store_true",
                    help="trace each frame synchronously instead of overlapping it with the next set inversion")
    ap.add_argument("--prune", default="auto", choices=["auto", "on", "off", "exact"],
                    help="liveness-pruned evaluation: on, off (full tape per cell), exact (only for "
                         "tapes whose pruned results are provably identical), auto (default): off for "
                         "these workloads, where it does not pay (capsules: the reverse-list cells "
                         "cannot use their parents' records; csg: the swept windows keep ~80 %% of "
                         "the records live)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="instead of the benchmark: copy a buffer of this many MiB device-to-device 5 times "
                         "(the reference kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
    ap.add_argument("--summarize", metavar="KERNEL_TRACE_DB",
                    help="instead of the benchmark: summarise a rocprofv3 --kernel-trace --stats database")
    ap.add_argument("--pmc-fetch", metavar="DB")
    ap.add_argument("--pmc-write", metavar="DB")
    ap.add_argument("--copy-fetch", metavar="DB")
    ap.add_argument("--copy-write", metavar="DB")
    ap.add_argument("--copy-mib", type=int, default=256)
    args = ap.parse_args()
    if args.summarize:
        summarize_profile(args)
        return
    if args.calibrate_copy_mib:
        import torch

        n = args.calibrate_copy_mib << 20
        src = torch.ones(n // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n, "launches": 5}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: this build runs the "
              f"set-inversion path on a single MI355X (it would shard by octree subtree)", file=sys.stderr)
        sys.exit(2)

    from temporalsetinversion_amd import scenes

    w = WORKLOADS[args.workload]
    ll = scenes.SCENES[w["scene"]](w["seed"])
    cpu = None if args.no_cpu else cpu_baseline(ll, w, args.cpu_budget)

    import temporalsetinversion_amd as T

    tape = T.tape_from_ll(ll)
    prune = args.prune if args.prune != "auto" else "off"
    st = T.State(tape, width=w["width"], height=w["height"], localsubdiv=w["localsubdiv"], delta=0.01,
                 point_sample=w["point_sample"], swept_volume=w["swept_volume"],
                 lipschitz_corrected=w["lipschitz_corrected"], voxel_capacity=1 << 27,
                 prune={"on": True, "off": False, "exact": "exact"}[prune], engine=args.engine)
    info = st.tape_info()
    st.reset()
    frame = 0

    def step():
        """Queues frame k (set inversion, trace, root re-enqueue) and returns frame k-1's result:
        the host never waits for the frame it has just queued, so the device always has the
        next frame's work; the trace of frame k runs on a trace stream beside the set
        inversion of frame k+1."""
        nonlocal frame, inflight
        st.update(float(np.float32(frame / 60.0)))
        frame += 1
        if args.sync_trace:
            r = st.set_inversion()
            st.render(pixels=False)
            return [r]
        st.set_inversion_async()
        st.render_async()
        inflight += 1
        if inflight < 2:
            return []
        inflight -= 1
        return [st.set_inversion_result()]

    def drain():
        nonlocal inflight
        out = []
        while inflight:
            out.append(st.set_inversion_result())
            inflight -= 1
        return out

    inflight = 0
    for _ in range(args.warmup):
        step()
    drain()
    st.trace_wait(pixels=False)
    frame_results = []
    ms0, visits0, n0 = st.trace_totals()
    tot = dict(cells=0, rev=0, fwd_ops=0, rev_ops=0, gpu_ms=0.0, eval_ms=0.0, launches=0, pruned=0)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        frame_results += step()
    frame_results += drain()
    st.trace_wait(pixels=False)   # the last frame's trace is part of the timed work
    wall = time.perf_counter() - t0
    for r in frame_results:
        s = r.stats
        tot["cells"] += s["cell_evals"]
        tot["rev"] += s["rev_evals"]
        tot["fwd_ops"] += s["fwd_ops"]
        tot["rev_ops"] += s["rev_ops"]
        tot["gpu_ms"] += s["gpu_ms"]
        tot["eval_ms"] += s["eval_ms"]
        tot["launches"] += s["eval_launches"]
        tot["pruned"] += s["pruned_instr"]
    assert len(frame_results) == args.steps
    ms1, visits1, n1 = st.trace_totals()
    assert n1 - n0 == args.steps
    trace = dict(ms=ms1 - ms0, visits=visits1 - visits0)
    st.close()

    ops = tot["fwd_ops"] + tot["rev_ops"]
    achieved = ops / (tot["eval_ms"] * 1e-3) / 1e12 if tot["eval_ms"] > 0 else 0.0
    launches = max(tot["launches"], 1)
    traffic, traffic_src, prof_avg_us = profiled_traffic(args.workload, info["engine"])
    # trace: SURVEY-style algorithmic bytes, 40 B per node visit + 20 B per pixel (colour + depth)
    px = w["width"] * w["height"]
    trace_bytes = trace["visits"] * 40 + args.steps * px * 20
    trace_gbs = trace_bytes / (trace["ms"] * 1e-3) / 1e9 if trace["ms"] > 0 else 0.0
    out = {
        "metric": "interval cell-evals/s",
        "value": tot["cells"] / wall,
        "unit": "cell-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.workload, "tape_instructions": int(len(tape)), "localsubdiv": w["localsubdiv"],
                   "frame": f"{w['width']}x{w['height']}", "delta": 0.01,
                   "lipschitz_corrected": w["lipschitz_corrected"], "swept_volume": w["swept_volume"],
                   "point_sample": w["point_sample"], "times": "t_k = k/60", "engine": info["engine"],
                   "trace": "synchronous" if args.sync_trace else "overlapped with the next set inversion",
                   "prune": prune, "frames_in_flight": 1 if args.sync_trace else 2},
        "tape_compile_s": info["compile_ms"] / 1e3, "tape_compile_cached": bool(info["cache_hit"]),
        "records_run_per_cell": (len(tape) - 4) - tot["pruned"] / max(tot["cells"], 1),
        "set_inversion_ms_per_frame": tot["gpu_ms"] / args.steps,
        "cell_evals_per_frame": tot["cells"] / args.steps,
        "rev_evals_per_frame": tot["rev"] / args.steps,
        "frames_per_s": args.steps / wall,
        "roofline": {"bound": "valu",
                     "kernel": ("tsi_tape_eval_level (one launch per octree level: forward-only and reverse lists)"
                                if info["engine"] == "compiled" else
                                "k_eval<false>/<true> (one forward-only and one reverse launch per octree level)"),
                     "achieved": achieved, "peak": PEAK_VALU_TOPS, "unit": "T f32-ops/s",
                     "frac": achieved / PEAK_VALU_TOPS, "traffic": traffic, "traffic_unit": "HBM bytes per launch",
                     "traffic_source": traffic_src,
                     "algorithmic_ops_per_launch": ops / launches,
                     # HIP events around the level loop: kernels plus the dispatch gaps between them
                     "avg_launch_us": tot["eval_ms"] * 1e3 / launches,
                     # the same kernels' average duration in the committed rocprofv3 summary
                     "profile_avg_kernel_us": prof_avg_us,
                     "peak_unpacked": PEAK_VALU_UNPACKED_TOPS,
                     "eval_ms_per_frame": tot["eval_ms"] / args.steps,
                     "ops_per_frame": ops / args.steps},
        "roofline_trace": {"bound": "hbm", "kernel": "k_trace", "achieved": trace_gbs, "peak": PEAK_HBM_GBS,
                           "unit": "GB/s", "frac": trace_gbs / PEAK_HBM_GBS,
                           "trace_ms_per_frame": trace["ms"] / args.steps,
                           "node_visits_per_frame": trace["visits"] / args.steps},
    }
    if cpu is not None:
        out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
