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


def cpu_baseline(tape, w, budget_s: float):
    """Oracle (single-threaded C restatement) on a bounded sample: the same tape and
    parameters at a reduced depth, frames t_k = k/60 until the time budget is spent."""
    import oracle

    L = max(3, w["localsubdiv"] - 3)
    fr = oracle.Frame(tape, localsubdiv=L, swept_volume=w["swept_volume"],
                      lipschitz_corrected=w["lipschitz_corrected"], point_sample=w["point_sample"],
                      voxel_capacity=1 << 22)
    cells = 0
    frames = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s and frames < 120:
        st = fr.set_inversion(float(np.float32(frames / 60.0)))
        cells += st.cell_evals
        frames += 1
        fr.enqueue_root()
    dt = time.perf_counter() - t0
    fr.close()
    return {"value": cells / dt, "unit": "cell-evals/s", "cores": 1, "kind": "port",
            "sample": f"oracle, same tape at localsubdiv {L}, {frames} frames t=k/60, {cells} cell-evals, "
                      f"{dt:.1f} s"}


def _rocpd_kernels(db):
    """(name, dispatches, total ns) per kernel of a rocprofv3 --kernel-trace database."""
    import sqlite3

    c = sqlite3.connect(db)
    agg = {}
    for name, d in c.execute("select name, duration from kernels"):   # ns per dispatch
        a = agg.setdefault(name, [0, 0.0])
        a[0] += 1
        a[1] += float(d)
    return sorted(((n, k, t) for n, (k, t) in agg.items()), key=lambda r: -r[2])


def _rocpd_counter(db):
    """kernel short name -> (dispatches, summed counter value in KiB) of a --pmc run."""
    import sqlite3

    c = sqlite3.connect(db)
    out = {}
    for name, v in c.execute("select name, counter_value from pmc_events"):
        k = name.split("(")[0].replace("void ", "")
        a = out.setdefault(k, [0, 0.0])
        a[0] += 1
        a[1] += v
    return out


def summarize_profile(args):
    """Text summary of a rocprofv3 --kernel-trace --stats run and, when given, of the
    FETCH_SIZE / WRITE_SIZE passes of the same command and of the copy calibration."""
    lines = ["kernel                                                   dispatches   total_ms    avg_us   share"]
    top = _rocpd_kernels(args.summarize)
    total = sum(t for _, _, t in top) or 1.0
    for n, k, t in top[:14]:
        lines.append(f"{n.split('(')[0][:56]:56s} {k:10d} {t / 1e6:10.3f} {t / k / 1e3:9.2f} {100 * t / total:6.1f}%")
    if args.pmc_fetch and args.pmc_write and args.copy_fetch and args.copy_write:
        mib = args.copy_mib
        cf = _rocpd_counter(args.copy_fetch)["__amd_rocclr_copyBuffer"]
        cw = _rocpd_counter(args.copy_write)["__amd_rocclr_copyBuffer"]
        f_scale = (mib * 1024.0) / (cf[1] / cf[0])
        w_scale = (mib * 1024.0) / (cw[1] / cw[0])
        lines.append("")
        lines.append(f"copy calibration ({mib} MiB device-to-device): FETCH_SIZE reads {cf[1] / cf[0]:.0f} KiB "
                     f"-> scale {f_scale:.3f}; WRITE_SIZE reads {cw[1] / cw[0]:.0f} KiB -> scale {w_scale:.3f}")
        fe, wr = _rocpd_counter(args.pmc_fetch), _rocpd_counter(args.pmc_write)
        lines.append("kernel                                 dispatches  fetch_KiB/disp  write_KiB/disp  hbm_bytes/disp")
        for k in sorted(fe, key=lambda k: -fe[k][1])[:8]:
            n, fkib = fe[k]
            wkib = wr.get(k, [n, 0.0])[1] / max(wr.get(k, [n, 0.0])[0], 1)
            hbm = (fkib / n * f_scale + wkib * w_scale) * 1024.0
            lines.append(f"{k[:38]:38s} {n:10d} {fkib / n * f_scale:15.1f} {wkib * w_scale:15.1f} {hbm:15.0f}")
    print("\n".join(lines))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="capsules64_L9_1024", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--engine", default="compiled", choices=["compiled", "interpreter"],
                    help="tape evaluation: compiled straight-line kernels (default) or the interpreter")
    ap.add_argument("--prune", default="on", choices=["on", "off", "exact"],
                    help="pruned per-group programs: on (default), off (full tape per cell), exact "
                         "(only for tapes whose pruned results are provably identical)")
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

    import temporalsetinversion_amd as T
    from temporalsetinversion_amd import scenes

    w = WORKLOADS[args.workload]
    ll = scenes.SCENES[w["scene"]](w["seed"])
    tape = T.tape_from_ll(ll)
    st = T.State(tape, width=w["width"], height=w["height"], localsubdiv=w["localsubdiv"], delta=0.01,
                 point_sample=w["point_sample"], swept_volume=w["swept_volume"],
                 lipschitz_corrected=w["lipschitz_corrected"], voxel_capacity=1 << 27,
                 prune={"on": True, "off": False, "exact": "exact"}[args.prune], engine=args.engine)
    info = st.tape_info()
    st.reset()
    frame = 0

    def step():
        nonlocal frame
        st.update(float(np.float32(frame / 60.0)))
        r = st.set_inversion()
        st.render(pixels=False)
        frame += 1
        return r

    for _ in range(args.warmup):
        step()
    tot = dict(cells=0, rev=0, fwd_ops=0, rev_ops=0, gpu_ms=0.0, eval_ms=0.0, launches=0, pruned=0)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = step()
        s = r.stats
        tot["cells"] += s["cell_evals"]
        tot["rev"] += s["rev_evals"]
        tot["fwd_ops"] += s["fwd_ops"]
        tot["rev_ops"] += s["rev_ops"]
        tot["gpu_ms"] += s["gpu_ms"]
        tot["eval_ms"] += s["eval_ms"]
        tot["launches"] += s["eval_launches"]
        tot["pruned"] += s["pruned_instr"]
    wall = time.perf_counter() - t0
    st.close()

    ops = tot["fwd_ops"] + tot["rev_ops"]
    achieved = ops / (tot["eval_ms"] * 1e-3) / 1e12 if tot["eval_ms"] > 0 else 0.0
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
                   "prune": args.prune if info["engine"] == "interpreter" else "n/a"},
        "tape_compile_s": info["compile_ms"] / 1e3, "tape_compile_cached": bool(info["cache_hit"]),
        "records_run_per_cell": (len(tape) - 4) - tot["pruned"] / max(tot["cells"], 1),
        "set_inversion_ms_per_frame": tot["gpu_ms"] / args.steps,
        "cell_evals_per_frame": tot["cells"] / args.steps,
        "rev_evals_per_frame": tot["rev"] / args.steps,
        "frames_per_s": args.steps / wall,
        "roofline": {"bound": "valu", "kernel": "k_eval (tape evaluation, forward + reverse)",
                     "achieved": achieved, "peak": PEAK_VALU_TOPS, "unit": "T f32-ops/s",
                     "frac": achieved / PEAK_VALU_TOPS, "traffic": None,
                     "peak_unpacked": PEAK_VALU_UNPACKED_TOPS,
                     "eval_ms_per_frame": tot["eval_ms"] / args.steps,
                     "ops_per_frame": ops / args.steps},
    }
    if not args.no_cpu:
        import oracle

        out["cpu_baseline"] = cpu_baseline(oracle.parse_ll(ll), w, args.cpu_budget)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
