"""Benchmark: temporal set inversion on one MI355X.

One step = one frame of the plain per-frame path on the 1024^3 workload (seeded straight-line
CSG of 256 spheres / boxes / tori under min / max / smooth-min with sin/cos domain warps in x, y,
z and t; localsubdiv 10 = 1024^3 cells; 1920x1080 frame; delta 0.01; swept over each frame's
window [t - 1/60, t]; point-sampled leaves): set inversion + trace of the frame + root
re-enqueue, at t_k = k/60.  The first W frames are untimed warm-up (frame 0 evaluates every
cell); the next K frames are timed with the device synchronised on both sides.

Prints one JSON line.  `value` = interval cell-evals per second over the timed frames;
`roofline` reports the tape-evaluation kernel against the FP32 VALU peak (its device time from
HIP events right around each level launch); `temporal_vs_full` compares the set inversion of
the temporal frames with full re-evaluation (every frame from an empty octree), the quantity
the reference publishes (~3x at 1024^3, delta 0.01); `cpu_baseline` times the oracle (the C
restatement of the reference) on a bounded sample of the same workload.
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

# MI355X FP32 vector peak as published: 157.3 TFLOP/s (256 CUs x 2.4 GHz x 256 flops per clock:
# an FMA counted as 2 and packed dual-FP32 issue).  The tape kernels count one op per f32 result
# and cannot use FMA (results must be bit-identical to the unfused rules) or packed FP32 (off on
# this machine); their attainable rate is one result per lane per clock, 39.3 T/s.
PEAK_VALU_TFLOPS = 157.3
PEAK_VALU_UNPACKED_TOPS = 39.3
PEAK_HBM_GBS = 8000.0

WORKLOADS = {
    # plain method on one MI355X: 512^3, 1024x1024, 64 capsules (512^3 cells)
    "capsules64_L9_1024": dict(scene="capsules", seed=1, localsubdiv=9, width=1024, height=1024,
                               lipschitz_corrected=True, swept_volume=False, point_sample=True),
    # plain method, large: 1024^3, 1920x1080, 256-primitive CSG, swept, point-sampled
    "csg256_L10_1920x1080": dict(scene="csg", seed=2, localsubdiv=10, width=1920, height=1080,
                                 lipschitz_corrected=False, swept_volume=True, point_sample=True),
    # the 8-moving-spheres field at 64^3 (the reference's CPU configuration)
    "spheres8_L6_256": dict(scene="spheres", seed=0, localsubdiv=6, width=256, height=256,
                            lipschitz_corrected=False, swept_volume=False, point_sample=True),
    # the capsules at 1024^3 (an animated scene at the resolution of the reference's published
    # temporal-vs-full figure)
    "capsules64_L10_1024": dict(scene="capsules", seed=1, localsubdiv=10, width=1024, height=1024,
                                lipschitz_corrected=True, swept_volume=False, point_sample=True),
}
DEFAULT_WORKLOAD = "csg256_L10_1920x1080"


# the bench options of a profiled run: the headline workload's frames only
PROFILE_ONLY = "--no-cpu --no-full --no-c1 --no-batched --no-served --no-cinematic"


def _profile_path(workload: str) -> str:
    return os.path.join(ROOT, "profiles", f"tape_eval_{workload}.txt")


def profiled_traffic(workload: str, engine: str):
    """(HBM bytes per eval launch, summary path, average kernel us) from the committed rocprofv3
    summary of this workload (timed frames only), if any."""
    path = _profile_path(workload)
    if engine != "compiled" or not os.path.exists(path):
        return None, None, None
    traffic = avg = None
    for line in open(path):
        parts = line.split()
        # "tsi_tape_eval_level timed <n> dispatches avg_us <us>" and
        # "tsi_tape_eval_level hbm <n> <fetch KiB> <write KiB> <bytes per dispatch>"
        if len(parts) >= 6 and parts[0] == "tsi_tape_eval_level" and parts[1] == "timed":
            avg = float(parts[5])
        if len(parts) >= 6 and parts[0] == "tsi_tape_eval_level" and parts[1] == "hbm":
            traffic = float(parts[5])
    return traffic, os.path.relpath(path, ROOT), avg


def _csv_rows(path):
    import csv

    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def _kernel(name: str) -> str:
    return name.split("(")[0].strip()


def summarize_profile(args) -> None:
    """Prints the committed profile summary (profiles/tape_eval_<workload>.txt) from rocprofv3 CSV
    output: the kernel trace of a bench run, the counter-collection runs of the same command
    (FETCH_SIZE, WRITE_SIZE, SQ counters) and FETCH_SIZE / WRITE_SIZE over a 256 MiB device copy
    (the scale that turns those counters into bytes on this machine).  Dispatches of the first
    --skip-frames frames (the warm-up: frame 0 evaluates every cell) are left out of the averages."""
    W, skip = args.workload, args.skip_frames
    L = WORKLOADS[W]["localsubdiv"]
    rows = sorted(_csv_rows(args.summarize), key=lambda r: int(r["Start_Timestamp"]))
    # frames are delimited by k_frame_begin
    frame_of, k = [], -1
    for r in rows:
        if _kernel(r["Kernel_Name"]) == "tsi::k_frame_begin":
            k += 1
        frame_of.append(k)
    frames = k + 1
    per, timed = {}, {}
    for r, f in zip(rows, frame_of):
        name = _kernel(r["Kernel_Name"])
        d = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3   # us
        n, t = per.get(name, (0, 0.0))
        per[name] = (n + 1, t + d)
        if f >= skip:
            n, t = timed.get(name, (0, 0.0))
            timed[name] = (n + 1, t + d)
    total = sum(t for _, t in per.values())
    R = "$R"
    lines = [f"# Kernel time, HBM traffic and SQ counters: {W}, compiled tape engine, bench defaults",
             "#",
             "# commands (from /tmp on the GPU host, repo at $R, <d> = an output directory):",
             f"#   rocprofv3 --kernel-trace --stats -f csv -d <d>/kt -o kt -- python3 {R}/bench.py --steps 5 --warmup 2 "
             f"{PROFILE_ONLY} --workload {W}"]
    for c in ("FETCH_SIZE", "WRITE_SIZE"):
        lines.append(f"#   rocprofv3 --pmc {c} -f csv -d <d>/{c} -o pmc -- python3 {R}/bench.py --steps 2 --warmup 1 "
                     f"{PROFILE_ONLY} --workload {W}")
        lines.append(f"#   rocprofv3 --pmc {c} -f csv -d <d>/copy_{c} -o pmc -- python3 {R}/bench.py --calibrate-copy-mib "
                     f"{args.copy_mib}")
    if args.pmc_sq:
        lines.append(f"#   rocprofv3 --pmc <SQ counters of each pass below> -f csv -d <d>/SQ<n> -o pmc -- python3 {R}/bench.py "
                     f"--steps 2 --warmup 1 {PROFILE_ONLY} --workload {W}")
    lines.append(f"#   python3 bench.py --workload {W} --skip-frames {skip} --summarize <d>/kt/..._kernel_trace.csv "
                 "--pmc-fetch <FETCH_SIZE counter_collection.csv> --pmc-write <WRITE_SIZE ...> --copy-fetch <...> "
                 "--copy-write <...> --pmc-sq <SQ pass csv files>")
    lines += ["#", f"# {frames} frames ({skip} warm-up, {frames - skip} timed); one tape-evaluation launch per octree "
              f"level ({L} per frame) running the", "# forward-only and the reverse list side by side, one more "
              "launch per frame (the speculative pass of the next", "# frame's shallow levels, on its own stream), one "
              "snapshot + trace per frame on a trace stream.", ""]
    lines.append(f"{'kernel':<40} {'dispatches':>10} {'total_ms':>10} {'avg_us':>9} {'share':>7}")
    for name, (n, t) in sorted(per.items(), key=lambda x: -x[1][1]):
        lines.append(f"{name[:40]:<40} {n:>10} {t / 1e3:>10.3f} {t / n:>9.2f} {100 * t / total:>6.1f}%")
    # per octree level: the k-th evaluation launch of a frame on the frame's queue runs level k;
    # the speculative pass is the evaluation launch that follows k_spec_gather on its own queue
    # (it runs beside the previous frame's last levels)
    spec_rows, pending = set(), set()
    for i, r in enumerate(rows):
        k, q = _kernel(r["Kernel_Name"]), r["Queue_Id"]
        if k == "tsi::k_spec_gather":
            pending.add(q)
        elif k == "tsi_tape_eval_level" and q in pending:
            pending.discard(q)
            spec_rows.add(i)
    evals, per_level = {}, {}
    for i, (r, f) in enumerate(zip(rows, frame_of)):
        if _kernel(r["Kernel_Name"]) == "tsi_tape_eval_level" and f >= skip:
            d = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
            if i in spec_rows:
                per_level.setdefault("spec", []).append(d)
            else:
                evals.setdefault(f, []).append(d)
    for f, ds in evals.items():
        if len(ds) != L:
            continue   # another workload's frame
        for lab, d in zip(range(L), ds):
            per_level.setdefault(lab, []).append(d)
    if per_level:
        lines += ["", "tsi_tape_eval_level per octree level, timed frames (us): level  mean  min  max",
                  "(spec: the speculative pass over the octree's nodes at depths < 7 for the next frame, on a "
                  "stream of its own beside levels 8-9, after its one-block node list k_spec_gather)"]
        for d in sorted(per_level, key=lambda x: -1 if x == "spec" else x):
            v = per_level[d]
            lines.append(f"  level {d:>4} {sum(v) / len(v):>9.1f} {min(v):>9.1f} {max(v):>9.1f}")
    lines.append("")
    lines.append("timed frames only (the figures bench.py reads):")
    for name in ("tsi_tape_eval_level", "void tsi::k_trace<false>", "void tsi::k_trace_quads<false>"):
        if name in timed:
            n, t = timed[name]
            lines.append(f"{name} timed {n} dispatches avg_us {t / n:.2f}")
    if args.pmc_fetch and args.pmc_write and args.copy_fetch and args.copy_write:
        def per_kernel(path, counter, skip_frames=0):
            """{kernel: [values per dispatch in dispatch order]} of one counter, leaving out the
            dispatches of the first skip_frames frames (frames start at k_frame_begin)."""
            out, frame = {}, -1
            for r in sorted(_csv_rows(path), key=lambda r: int(r["Dispatch_Id"])):
                if r["Counter_Name"] != counter:
                    continue
                k = _kernel(r["Kernel_Name"])
                if k == "tsi::k_frame_begin":
                    frame += 1
                if frame >= skip_frames or "copyBuffer" in k:
                    out.setdefault(k, []).append(float(r["Counter_Value"]))
            return out

        kib = args.copy_mib * 1024.0

        def scale(path, counter):   # KiB the counter reports per copy dispatch -> bytes
            v = [x for k, xs in per_kernel(path, counter).items() if "copyBuffer" in k for x in xs]
            return kib / (sum(v) / len(v))

        sf, sw = scale(args.copy_fetch, "FETCH_SIZE"), scale(args.copy_write, "WRITE_SIZE")
        # the counter runs have 1 warm-up frame: its dispatches are left out
        f, w = per_kernel(args.pmc_fetch, "FETCH_SIZE", 1), per_kernel(args.pmc_write, "WRITE_SIZE", 1)
        lines += ["", f"copy calibration ({args.copy_mib} MiB device-to-device): FETCH_SIZE scale {sf:.3f}, "
                      f"WRITE_SIZE scale {sw:.3f} (bytes = KiB x scale x 1024); level launches of the timed frames",
                  f"{'kernel':<24} {'':>4} {'dispatches':>10} {'fetch_KiB/disp':>15} {'write_KiB/disp':>15} {'bytes/disp':>15}"]
        for k in ("tsi_tape_eval_level", "void tsi::k_trace<false>", "void tsi::k_trace_quads<false>", "tsi::k_snapshot"):
            if k in f and k in w:
                vf, vw = f[k], w[k]
                if not vf or not vw:
                    continue
                b = (sum(vf) / len(vf) * sf + sum(vw) / len(vw) * sw) * 1024.0
                lines.append(f"{k:<24} {'hbm':>4} {len(vf):>10} {sum(vf) / len(vf):>15.1f} {sum(vw) / len(vw):>15.1f} "
                             f"{b:>15.0f}")
    for path in args.pmc_sq:
        _valu_issue_per_level(path, L, lines)
        rows = _csv_rows(path)
        counters = sorted({r["Counter_Name"] for r in rows})
        vals = {}
        frame = -1
        for r in sorted(rows, key=lambda r: (int(r["Dispatch_Id"]), r["Counter_Name"])):
            k = _kernel(r["Kernel_Name"])
            if k == "tsi::k_frame_begin" and r["Counter_Name"] == counters[0]:
                frame += 1
            if k == "tsi_tape_eval_level" and frame >= 1:
                vals.setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
        lines += ["", f"SQ counters per tsi_tape_eval_level dispatch (timed frames), {os.path.basename(path)}:"]
        for c in counters:
            v = vals.get(c, [])
            if v:
                lines.append(f"  {c:<28} {sum(v) / len(v):>16.1f}")
    print("\n".join(lines))


# MI355X: 256 CUs x 4 SIMDs; a wave64 VALU instruction occupies a SIMD for 4 cycles at 2.4 GHz
VALU_ISSUE_PER_US = 256 * 4 * 2400 / 4


def _valu_issue_per_level(path, L, lines):
    """SQ_INSTS_VALU per level against the issue capacity of the same dispatch's duration (counter
    run, frames after the first): how close each level runs to one VALU instruction per SIMD
    every 4 cycles."""
    rows = sorted(_csv_rows(path), key=lambda r: int(r["Dispatch_Id"]))
    if not any(r["Counter_Name"] == "SQ_INSTS_VALU" for r in rows):
        return
    frame, spec_q, seen, per, pos = -1, set(), set(), {}, {}
    for r in rows:
        k, q, i = _kernel(r["Kernel_Name"]), r["Queue_Id"], int(r["Dispatch_Id"])
        if i in seen:
            continue
        seen.add(i)
        if k == "tsi::k_frame_begin":
            frame += 1
            pos[frame] = 0
        if k == "tsi::k_spec_gather":
            spec_q.add(q)
        if k != "tsi_tape_eval_level" or frame < 1:
            continue
        v = [float(x["Counter_Value"]) for x in rows if int(x["Dispatch_Id"]) == i and x["Counter_Name"] == "SQ_INSTS_VALU"]
        us = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
        if q in spec_q:
            lab = "spec"
        else:
            lab = pos[frame]
            pos[frame] += 1
        per.setdefault(lab, []).append((v[0] if v else 0.0, us))
    lines += ["", "VALU issue per level (counter run, frames after the first): SQ_INSTS_VALU per dispatch and its share "
              "of the issue", f"capacity of the dispatch's duration ({VALU_ISSUE_PER_US / 1e6:.2f} M wave64 instructions "
              "per us: 1024 SIMDs, one every 4 cycles, 2.4 GHz)"]
    for lab in sorted(per, key=lambda x: -1 if x == "spec" else x):
        if lab != "spec" and lab >= L:
            continue
        insts = sum(x for x, _ in per[lab]) / len(per[lab])
        us = sum(u for _, u in per[lab]) / len(per[lab])
        lines.append(f"  level {lab:>4} {insts / 1e6:>10.2f} M instructions in {us:>8.1f} us  "
                     f"{100 * insts / (VALU_ISSUE_PER_US * us):>6.1f} %")


def summarize_kernels(args) -> None:
    """Per-kernel dispatch counts and durations of a rocprofv3 kernel trace (the batched form's
    profile summary: the persistent set-inversion launch and the trace launch of each submit)."""
    rows = _csv_rows(args.summarize_kernels)
    per = {}
    for r in rows:
        name = _kernel(r["Kernel_Name"])
        if args.split_grid:   # one line per launch shape (threads per grid dimension)
            name = name + " " + r["Grid_Size_X"] + "x" + r["Grid_Size_Y"]
        d = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
        per.setdefault(name, []).append(d)
    total = sum(sum(v) for v in per.values())
    R = "$R"
    lines = [f"# Kernel time: {args.summary_title}", "#",
             "# command (from /tmp on the GPU host, repo at $R, <d> = an output directory):",
             f"#   rocprofv3 --kernel-trace --stats -f csv -d <d> -o kt -- python3 {R}/bench.py {args.summary_cmd}",
             f"#   python3 bench.py --summarize-kernels <d>/kt_kernel_trace.csv --summary-title ... --summary-cmd ..."
             + (" --split-grid" if args.split_grid else ""),
             "", f"{'kernel':<44} {'dispatches':>10} {'total_ms':>10} {'avg_us':>11} {'min_us':>11} {'max_us':>11} {'share':>7}"]
    for name, v in sorted(per.items(), key=lambda x: -sum(x[1])):
        lines.append(f"{name[:44]:<44} {len(v):>10} {sum(v) / 1e3:>10.3f} {sum(v) / len(v):>11.1f} {min(v):>11.1f} "
                     f"{max(v):>11.1f} {100 * sum(v) / total:>6.1f}%")
    print("\n".join(lines))


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cpu_worker(job):
    """Oracle frames t_k = k/60 over one sub-box of the root until the time budget is spent."""
    import oracle

    ll, w, L, box, budget_s = job
    tape = oracle.parse_ll(ll)
    fr = oracle.Frame(tape, localsubdiv=L, swept_volume=w["swept_volume"], lipschitz_corrected=w["lipschitz_corrected"],
                      point_sample=w["point_sample"], voxel_capacity=1 << 22, root_box=box)
    cells = frames = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s and frames < 240:
        st = fr.set_inversion(float(np.float32(frames / 60.0)))
        cells += st.cell_evals
        frames += 1
        fr.enqueue_root()
    dt = time.perf_counter() - t0
    fr.close()
    return cells, frames, dt


def _sub_boxes(ll, w, depth, count, seed=0):
    """Boxes of `count` depth-`depth` octree cells the surface passes through (seeded choice),
    from the oracle's first frame at that depth."""
    import oracle

    fr = oracle.Frame(oracle.parse_ll(ll), localsubdiv=depth, swept_volume=w["swept_volume"],
                      lipschitz_corrected=w["lipschitz_corrected"], point_sample=False, voxel_capacity=1 << 20)
    fr.set_inversion(0.0)
    n = fr.export()
    fr.close()
    cand = n[(n["depth"] == depth) & (n["f"] == 0.0)]
    pick = cand[np.random.default_rng(seed).choice(len(cand), min(count, len(cand)), replace=False)]
    size = 10.0 / (1 << depth)
    out = []
    for c in pick:
        lo = [-5.0 + size * int(c[a]) for a in ("i", "j", "k")]
        out.append(tuple(lo) + tuple(x + size for x in lo))
    return out


def cpu_baseline(ll, w, budget_s: float, workers: int):
    """The oracle (C restatement of compute.wgsl `main`) on the host cores over a bounded sample of
    the same workload: `workers` of the 512 depth-3 subtrees of the octree that the surface passes
    through, one per worker process, each an octree of its own over that sub-box with the same
    tape, parameters and epsilon (so the same cells as that subtree of the full octree), frames
    t_k = k/60 until the time budget is spent.  Runs before the GPU is initialised."""
    import multiprocessing as mp

    boxes = _sub_boxes(ll, w, 3, workers)
    ctx = mp.get_context("fork")
    t0 = time.perf_counter()
    with ctx.Pool(len(boxes)) as pool:
        res = pool.map(_cpu_worker, [(ll, w, w["localsubdiv"], b, budget_s) for b in boxes])
    wall = time.perf_counter() - t0
    cells = sum(r[0] for r in res)
    frames = sum(r[1] for r in res)
    return {"value": cells / wall, "unit": "cell-evals/s", "cores": len(boxes), "kind": "port",
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "sample": f"oracle (CPU restatement, not lavapipe: no Vulkan driver or cargo on the host) on "
                      f"{len(boxes)} host cores (the host's CPU share for one GPU), one process per depth-3 "
                      f"subtree the surface passes through ({len(boxes)} of 512), same tape, localsubdiv, "
                      f"delta, swept window and point sampling; {frames} frames, {cells} cell-evals in {wall:.1f} s"}


def cpu_baseline_c1(budget_s: float, workers: int):
    """The reference's CPU configuration (8 moving spheres, localsubdiv 6, 256x256 frame): one frame
    sequence per worker (phases 7 frames apart), set inversion and trace per frame."""
    import multiprocessing as mp

    ctx = mp.get_context("fork")
    t0 = time.perf_counter()
    with ctx.Pool(workers) as pool:
        res = pool.map(_c1_worker, [(7 * i, budget_s) for i in range(workers)])
    wall = time.perf_counter() - t0
    frames = sum(r[0] for r in res)
    cells = sum(r[1] for r in res)
    return {"frames_per_s": frames / wall, "cell_evals_per_s": cells / wall, "cores": workers, "kind": "port",
            "config": "spheres8_L6_256 (8 spheres, c(t) = c0 + 0.2 sin(2 pi t + phi), localsubdiv 6, 256x256, "
                      "delta 0.01, t_k = k/60)", "frames": frames, "seconds": wall}


def _c1_worker(job):
    import oracle
    import temporalsetinversion_amd as T
    from temporalsetinversion_amd import scenes

    phase, budget_s = job
    fr = oracle.Frame(oracle.parse_ll(scenes.spheres(0)), localsubdiv=6, voxel_capacity=1 << 20)
    view = T.diagonal_camera()
    zd = float(np.float32(256.0) / np.float32(np.tan(np.float32(np.pi / 4))))   # compute.wgsl:4592
    frames = cells = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s and frames < 600:
        t = float(np.float32((frames + phase) / 60.0))
        cells += fr.set_inversion(t).cell_evals
        fr.trace(view, 256, 256, t, zd)
        fr.enqueue_root()
        frames += 1
    fr.close()
    return frames, cells


def gpu_c1(T, frames: int = 120) -> dict:
    """The reference's CPU configuration on the plain GPU path, as cpu_baseline_c1 runs it: 8 moving
    spheres, localsubdiv 6, 256x256, a blocking set inversion and trace per frame at t_k = k/60."""
    from temporalsetinversion_amd import scenes

    st = T.State(T.tape_from_ll(scenes.spheres(0)), width=256, height=256, localsubdiv=6, voxel_capacity=1 << 20)
    view = T.diagonal_camera()
    st.reset()
    for k in range(10):   # warm-up
        st.update(float(np.float32(k / 60.0)), view)
        st.set_inversion()
        st.trace(pixels=False)
        st.enqueue_root()
    cells = 0
    t0 = time.perf_counter()
    for k in range(10, 10 + frames):
        st.update(float(np.float32(k / 60.0)), view)
        cells += st.set_inversion().stats["cell_evals"]
        st.trace(pixels=False)
        st.enqueue_root()
    wall = time.perf_counter() - t0
    st.close()
    return {"frames_per_s": frames / wall, "cell_evals_per_s": cells / wall, "frames": frames,
            "config": "spheres8_L6_256 (as cpu_baseline_c1), blocking set inversion + trace per frame"}


def cpu_cinematic(budget_s: float) -> dict:
    """The oracle's cinematic trace (the path tracer of get_color_supersample / get_color_single)
    on one host core over a bounded sample: rows of the spheres8_L6_256 frame at t = 10/60, 8 rows
    at a time, until the budget is spent."""
    import oracle
    import temporalsetinversion_amd as T
    from temporalsetinversion_amd import scenes

    fr = oracle.Frame(oracle.parse_ll(scenes.spheres(0)), localsubdiv=6, voxel_capacity=1 << 20,
                      shading=T._lib.TSI_SHADE_CINEMATIC)
    for k in range(11):
        fr.set_inversion(float(np.float32(k / 60.0)))
        if k < 10:
            fr.enqueue_root()
    view = np.array(T.diagonal_camera(), np.float32)
    zd = float(T.camera_zdist(256))
    rows = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s and rows < 256:
        fr.trace(view, 256, 256, float(np.float32(10 / 60.0)), zd, row0=rows, rows=8)
        rows += 8
    wall = time.perf_counter() - t0
    fr.close()
    return {"value": rows * 256 / wall, "unit": "pixels/s", "cores": 1, "kind": "port",
            "sample": f"{rows} rows of the spheres8_L6_256 frame at t = 10/60 (8 jittered samples, 2 bounces)",
            "seconds": wall}


def cinematic_form(T, args) -> dict:
    """The cinematic path tracer (TSI_SHADE_CINEMATIC) after the same set inversion as the plain
    frames: the reference's CPU configuration (spheres8_L6_256) and the headline workload.  Device
    time of the trace launch, node visits (octree child reads) per frame, and the algorithmic bytes
    of the trace (40 B per node visit + 20 B per pixel) over its time."""
    from temporalsetinversion_amd import scenes

    out = {}
    cases = [("spheres8_L6_256", WORKLOADS["spheres8_L6_256"], T.diagonal_camera(), 20),
             (args.workload, WORKLOADS[args.workload], None, 3)]
    for name, w, view, frames in cases:
        tape = T.tape_from_ll(scenes.SCENES[w["scene"]](w["seed"]))
        st = T.State(tape, width=w["width"], height=w["height"], localsubdiv=w["localsubdiv"], delta=0.01,
                     point_sample=w["point_sample"], swept_volume=w["swept_volume"],
                     lipschitz_corrected=w["lipschitz_corrected"],
                     voxel_capacity=1 << (20 if w["localsubdiv"] <= 6 else 27), prune=False,
                     shading=T._lib.TSI_SHADE_CINEMATIC)
        st.reset()
        ms = visits = 0.0
        for k in range(3 + frames):
            st.update(float(np.float32(k / 60.0)), view)
            st.set_inversion()
            st.trace(pixels=False)
            st.enqueue_root()
            if k >= 3:
                m, v = st.last_trace()
                ms += m
                visits += v
        st.close()
        px = w["width"] * w["height"]
        ms /= frames
        visits /= frames
        gbs = (visits * 40 + px * 20) / (ms * 1e-3) / 1e9
        out[name] = {"trace_ms_per_frame": ms, "pixels_per_s": px / (ms * 1e-3), "node_visits_per_frame": visits,
                     "node_visits_per_pixel": visits / px, "frame": f"{w['width']}x{w['height']}",
                     "roofline": {"bound": "hbm", "kernel": "k_trace<true>", "achieved": gbs, "peak": PEAK_HBM_GBS,
                                  "unit": "GB/s", "frac": gbs / PEAK_HBM_GBS}, "frames": frames}
    return out


def full_reevaluation(st, frames: int, t0_frame: int):
    """Set inversion of `frames` frames each from an empty octree (tsi_reset before the frame: every
    cell new), device ms per frame; the temporal method's comparison of the reference."""
    ms = cells = 0.0
    for k in range(frames):
        st.update(float(np.float32((t0_frame + k) / 60.0)))
        st.reset()
        r = st.set_inversion()
        ms += r.gpu_ms
        cells += r.stats["cell_evals"]
    return ms / frames, cells / frames


def temporal_frames(st, frames: int, warmup: int):
    """Set inversion of `frames` consecutive temporal frames after `warmup` untimed ones, from a
    reset; device ms and cell-evals per frame."""
    st.reset()
    ms = cells = 0.0
    for k in range(warmup + frames):
        st.update(float(np.float32(k / 60.0)))
        r = st.set_inversion()
        st.enqueue_root()
        if k >= warmup:
            ms += r.gpu_ms
            cells += r.stats["cell_evals"]
    return ms / frames, cells / frames


def make_state(T, w, tape, prune, engine):
    return T.State(tape, width=w["width"], height=w["height"], localsubdiv=w["localsubdiv"], delta=0.01,
                   point_sample=w["point_sample"], swept_volume=w["swept_volume"],
                   lipschitz_corrected=w["lipschitz_corrected"], voxel_capacity=1 << 27,
                   prune={"on": True, "off": False, "exact": "exact"}[prune], engine=engine)


# ---------------------------------------------------------------- batched and served forms

# node capacity per session by octree depth (generous: a request that still exhausts its pool is
# re-run with 4x the nodes during calibration, before anything is timed)
BATCH_CAPACITY = {6: 1 << 17, 7: 1 << 18, 8: 1 << 20, 9: 1 << 21, 10: 1 << 23}
SERVED_CAPACITY = {7: 1 << 20, 8: 1 << 21, 9: 1 << 22, 10: 1 << 24}


def _same_nodes(a, b) -> bool:
    if len(a) != len(b):
        return False
    a = a[np.lexsort((a["k"], a["j"], a["i"], a["depth"]))]
    b = b[np.lexsort((b["k"], b["j"], b["i"], b["depth"]))]
    for f in a.dtype.names:
        x, y = (a[f].view(np.uint32), b[f].view(np.uint32)) if a[f].dtype.kind == "f" else (a[f], b[f])
        if not np.array_equal(x, y):
            return False
    return True


def _plain_request(T, states, tapes, r, view, cap):
    """One batched request on the plain path: the tape's context with the request's depth and
    window, reset, one blocking set inversion and trace."""
    st = states.get(r["tape"])
    if st is None:
        st = states[r["tape"]] = T.State(tapes[r["tape"]], width=1920, height=1080, localsubdiv=10,
                                         voxel_capacity=cap, prune=False)
    st.set_params(localsubdiv=r["localsubdiv"], swept_volume=True, swept_window=r["window"])
    st.uniforms.resolution[0], st.uniforms.resolution[1] = r["width"], r["height"]
    st.width, st.height = r["width"], r["height"]
    st.reset()
    st.update(r["time"], view)
    res = st.set_inversion()
    st.trace(pixels=False)
    return st, res


def batched_form(T, args) -> dict:
    """The batched workload: `--batch-requests` independent requests (16 seeded sphere / capsule
    tapes of 100-2000 records, localsubdiv 6..10, swept t-windows of width 1/30..1/2, frames
    512x512..1920x1080), all in one submit (one persistent set-inversion launch + one trace
    launch), against the same requests one at a time on the plain path."""
    from temporalsetinversion_amd import scenes, tuning
    from temporalsetinversion_amd.batch import Batch

    tapes = [T.tape_from_ll(x) for x in scenes.batch_tapes()]
    reqs = scenes.batch_requests(args.batch_requests, len(tapes))
    view = T.diagonal_camera()
    caps = [BATCH_CAPACITY[r["localsubdiv"]] for r in reqs]
    b = Batch(tapes, arena_bytes=args.batch_arena_gib << 30, max_sessions=len(reqs))
    compile_s = b.info()["compile_ms"] / 1e3

    def open_all():
        return [b.open(tape=r["tape"], localsubdiv=r["localsubdiv"], swept_volume=True, swept_window=r["window"],
                       voxel_capacity=c, max_width=r["width"], max_height=r["height"]) for r, c in zip(reqs, caps)]

    sids = open_all()
    frames = [(s, r["time"], r["width"], r["height"], view, True) for s, r in zip(sids, reqs)]
    # calibration (untimed): requests whose pool is too small get 4x the nodes
    for _ in range(4):
        res, _, _ = b.run(frames)
        bad = [i for i, x in enumerate(res) if x["status"] != 0]
        if not bad:
            break
        for i in bad:
            caps[i] *= 4
        for s in sids:
            b.close_session(s)
        sids = open_all()
        frames = [(s, r["time"], r["width"], r["height"], view, True) for s, r in zip(sids, reqs)]
    else:
        raise RuntimeError("batched requests still fail after recalibration")
    tune = tuning.autotune(b, frames, repeats=1) if not args.no_tune else None
    # timed: whole submits of all requests
    walls, sets, traces = [], [], []
    for _ in range(args.batch_steps):
        t0 = time.perf_counter()
        res, s_ms, t_ms = b.run(frames)
        walls.append(time.perf_counter() - t0)
        sets.append(s_ms)
        traces.append(t_ms)
        assert all(x["status"] == 0 for x in res)
    cells = sum(x["cell_evals"] for x in res)
    rev_ops = sum(x["rev_ops"] + x["fwd_ops"] for x in res)
    wall = float(np.median(walls))
    # the same requests one at a time on the plain path (contexts per tape made beforehand)
    states = {}
    cap_of = {}
    for r, c in zip(reqs, caps):
        cap_of[r["tape"]] = max(cap_of.get(r["tape"], 0), c)
    for i, r in enumerate(reqs):   # untimed pass: contexts, trace buffers at their largest
        _plain_request(T, states, tapes, r, view, cap_of[r["tape"]])
    t0 = time.perf_counter()
    for r in reqs:
        _plain_request(T, states, tapes, r, view, cap_of[r["tape"]])
    plain_wall = time.perf_counter() - t0
    # identical leaf sets and frames on a seeded sample of requests
    rng = np.random.default_rng(11)
    sample = sorted(set(rng.choice(len(reqs), min(args.batch_check, len(reqs)), replace=False).tolist()) |
                    {int(np.argmax([r["localsubdiv"] * 10 + r["window"] for r in reqs]))})
    same = 0
    for i in sample:
        r = reqs[i]
        st, _ = _plain_request(T, states, tapes, r, view, cap_of[r["tape"]])
        ok = _same_nodes(st.read_nodes(), b.read_nodes(sids[i]))
        rgba, depth = st.trace()
        brgba, bdepth = b.frame(sids[i], r["width"], r["height"])
        ok = ok and np.array_equal(rgba.view(np.uint32), brgba.view(np.uint32)) and \
            np.array_equal(depth.view(np.uint32), bdepth.view(np.uint32))
        same += ok
    for st in states.values():
        st.close()
    info = b.info()
    b.close()
    L = np.bincount([r["localsubdiv"] for r in reqs], minlength=11)[6:].tolist()
    return {"workload": f"batched{len(reqs)}", "requests": len(reqs), "tapes": len(tapes),
            "requests_per_localsubdiv_6_to_10": L,
            "frames_per_s": len(reqs) / wall, "wall_ms_per_submit": wall * 1e3,
            "set_inversion_ms": float(np.median(sets)), "trace_ms": float(np.median(traces)),
            "cell_evals_per_submit": cells, "cell_evals_per_s": cells / wall, "ops_per_submit": rev_ops,
            "plain_frames_per_s": len(reqs) / plain_wall, "plain_wall_ms": plain_wall * 1e3,
            "speedup_vs_plain": plain_wall / wall,
            "identical_to_plain": {"checked": len(sample), "identical": same},
            "grid": {k: info[k] for k in ("grid_blocks", "waves_per_simd", "claim_units", "stack_rows")},
            "module_compile_s": compile_s, "module_cached": bool(info["cache_hit"]),
            "tuning": ({"best": tune["best"], "seconds": tune["seconds"], "trials": tune["trials"]} if tune else None)}


def served_form(T, args) -> dict:
    """The served workload: 32 seeded capsule scenes (200-2000 records, localsubdiv 7..10,
    Lipschitz correction, frames 512x512..1920x1080), requests arriving as a Poisson process;
    each request is its scene's next 1/60 window.  Continuous batching (serve.Server) against the
    plain path serving the same schedule one request at a time (serve.PlainServer)."""
    from temporalsetinversion_amd import scenes, serve
    from temporalsetinversion_amd.batch import Batch

    tapes = [T.tape_from_ll(x) for x in scenes.served_tapes()]
    rng = np.random.default_rng(12)
    view = T.diagonal_camera()
    sc = []
    for i in range(len(tapes)):
        L = int(rng.integers(7, 11))
        W, H = scenes.FRAME_SIZES[int(rng.integers(0, len(scenes.FRAME_SIZES)))]
        sc.append(serve.Scene(tape=i, localsubdiv=L, width=W, height=H, view=view,
                              params=dict(lipschitz_corrected=True), voxel_capacity=SERVED_CAPACITY[L]))
    reqs = serve.poisson_requests(len(sc), args.served_rate, args.served_seconds, seed=13)
    b = Batch(tapes, arena_bytes=args.served_budget_gib << 30, max_sessions=len(sc))
    b.tune(waves_per_simd=b.info()["max_waves_per_simd"])
    srv = serve.Server(b, sc)
    served = srv.run(reqs)
    plain = serve.PlainServer(tapes, sc)
    pl = plain.run(reqs)
    # saturation: an offered load above what either path sustains (requests continue each scene's
    # frame sequence); frames/s is then each path's capacity
    sat = None
    if args.served_saturation_rate > 0:
        more = serve.poisson_requests(len(sc), args.served_saturation_rate, args.served_saturation_seconds, seed=14)
        done = [max([r.frame for r in reqs if r.scene == s], default=-1) + 1 for s in range(len(sc))]
        for r in more:
            r.frame += done[r.scene]
        s_sat = srv.run(more)
        p_sat = plain.run(more)
        reqs = reqs + more
        sat = {"rate_req_per_s": args.served_saturation_rate, "seconds": args.served_saturation_seconds,
               "requests": len(more), "frames_per_s": s_sat["frames_per_s"], "p50_ms": s_sat["p50_ms"],
               "p99_ms": s_sat["p99_ms"], "mean_batch": s_sat["mean_batch"], "failed": s_sat["failed"],
               "plain_frames_per_s": p_sat["frames_per_s"], "plain_p50_ms": p_sat["p50_ms"],
               "plain_p99_ms": p_sat["p99_ms"], "plain_failed": p_sat["failed"],
               "speedup_vs_plain_frames_per_s": s_sat["frames_per_s"] / p_sat["frames_per_s"]}
    # each scene's octree after the same windows, alone (plain) and in the mix (served)
    same = sum(_same_nodes(plain.states[s].read_nodes(), b.read_nodes(srv.sid[s])) for s in srv.sid)
    # solo latency: each scene's next frame alone on an idle device
    solo = []
    for s in sorted(srv.sid):
        k = max([r.frame for r in reqs if r.scene == s], default=-1) + 1
        t0 = time.perf_counter()
        b.run([(srv.sid[s], serve.frame_time(k), sc[s].width, sc[s].height, view)])
        solo.append(time.perf_counter() - t0)
    info = b.info()
    plain.close()
    b.close()
    solo_ms = np.asarray(solo) * 1e3
    drop = ("latencies",)
    return {"workload": f"served{len(sc)}", "rate_req_per_s": args.served_rate, "seconds": args.served_seconds,
            "requests": len(reqs), "budget_gib": args.served_budget_gib,
            "arena_used_gib": info["arena_used"] / 2**30, "sessions": len(srv.sid),
            **{k: v for k, v in served.items() if k not in drop},
            "solo_p50_ms": float(np.median(solo_ms)), "solo_p99_ms": float(np.percentile(solo_ms, 99)),
            "p99_over_solo_p99": served["p99_ms"] / float(np.percentile(solo_ms, 99)),
            "p99_over_solo_p50": served["p99_ms"] / float(np.median(solo_ms)),
            "plain": {k: v for k, v in pl.items() if k not in drop},
            "speedup_vs_plain_frames_per_s": served["frames_per_s"] / pl["frames_per_s"] if pl["frames_per_s"] else None,
            "sessions_identical_to_plain": {"checked": len(srv.sid), "identical": int(same)},
            "saturation": sat,
            "levels": [s.localsubdiv for s in sc]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default=DEFAULT_WORKLOAD, choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds per CPU-baseline worker")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-full", action="store_true", help="skip the temporal-vs-full re-evaluation comparison")
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
    ap.add_argument("--skip-frames", type=int, default=0,
                    help="with --summarize: leave out the first N frames' dispatches (the warm-up frames)")
    ap.add_argument("--pmc-fetch", metavar="DB")
    ap.add_argument("--pmc-write", metavar="DB")
    ap.add_argument("--pmc-sq", metavar="DB", nargs="*", default=[])
    ap.add_argument("--copy-fetch", metavar="DB")
    ap.add_argument("--copy-write", metavar="DB")
    ap.add_argument("--copy-mib", type=int, default=256)
    ap.add_argument("--summarize-kernels", metavar="KERNEL_TRACE_CSV",
                    help="instead of the benchmark: per-kernel summary of a rocprofv3 kernel trace")
    ap.add_argument("--summary-title", default="")
    ap.add_argument("--summary-cmd", default="")
    ap.add_argument("--split-grid", action="store_true", help="with --summarize-kernels: a line per launch shape")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched form")
    ap.add_argument("--no-served", action="store_true", help="skip the served form")
    ap.add_argument("--no-cinematic", action="store_true", help="skip the cinematic path tracer")
    ap.add_argument("--no-c1", action="store_true", help="skip the reference's CPU configuration on the GPU")
    ap.add_argument("--no-tune", action="store_true", help="batched form: keep the default launch shape")
    ap.add_argument("--batch-requests", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--batch-check", type=int, default=8, help="batched requests compared with the plain path")
    ap.add_argument("--batch-arena-gib", type=int, default=96)
    ap.add_argument("--served-rate", type=float, default=400.0, help="requests per second")
    ap.add_argument("--served-seconds", type=float, default=5.0)
    ap.add_argument("--served-saturation-rate", type=float, default=3000.0,
                    help="offered load above either path's capacity (0: skip)")
    ap.add_argument("--served-saturation-seconds", type=float, default=1.5)
    ap.add_argument("--served-budget-gib", type=int, default=200, help="HBM admission budget of the served form")
    args = ap.parse_args()
    if args.summarize:
        summarize_profile(args)
        return
    if args.summarize_kernels:
        summarize_kernels(args)
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
    cpu = cpu_c1 = cpu_cin = None
    if not args.no_cpu:   # before the GPU is touched (worker processes are forked)
        workers = max(1, min(16, os.cpu_count() or 1))
        print(f"cpu baseline: {workers} workers x {args.cpu_budget:.0f} s (same workload), then the spheres "
              f"configuration", file=sys.stderr, flush=True)
        cpu = cpu_baseline(ll, w, args.cpu_budget, workers)
        cpu_c1 = cpu_baseline_c1(min(args.cpu_budget, 6.0), workers)
        cpu_cin = None if args.no_cinematic else cpu_cinematic(min(args.cpu_budget, 6.0))

    import temporalsetinversion_amd as T

    tape = T.tape_from_ll(ll)
    prune = args.prune if args.prune != "auto" else "off"
    st = make_state(T, w, tape, prune, args.engine)
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
    print(f"{args.workload}: {args.warmup} warm-up + {args.steps} timed frames", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        step()
    drain()
    st.trace_wait(pixels=False)
    frame_results = []
    ms0, visits0, n0 = st.trace_totals()
    tot = dict(cells=0, rev=0, fwd_ops=0, rev_ops=0, gpu_ms=0.0, eval_ms=0.0, kernel_ms=0.0, launches=0, pruned=0)
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
        tot["kernel_ms"] += s["eval_kernel_ms"]
        tot["launches"] += s["eval_launches"]
        tot["pruned"] += s["pruned_instr"]
    assert len(frame_results) == args.steps
    ms1, visits1, n1 = st.trace_totals()
    assert n1 - n0 == args.steps
    trace = dict(ms=ms1 - ms0, visits=visits1 - visits0)
    print(f"timed: {wall * 1e3 / args.steps:.3f} ms per frame", file=sys.stderr, flush=True)

    # the reference's published comparison: temporal frames against full re-evaluation
    tvf = None
    if not args.no_full:
        si_ms = tot["gpu_ms"] / args.steps
        full_ms, full_cells = full_reevaluation(st, 5, frame)
        tvf = {"workload": args.workload, "temporal_set_inversion_ms": si_ms, "full_set_inversion_ms": full_ms,
               "speedup": full_ms / si_ms, "temporal_cell_evals": tot["cells"] / args.steps,
               "full_cell_evals": full_cells}
        if args.workload != "capsules64_L10_1024":
            w2 = WORKLOADS["capsules64_L10_1024"]
            st2 = make_state(T, w2, T.tape_from_ll(scenes.SCENES[w2["scene"]](w2["seed"])), "off", args.engine)
            t_ms, t_cells = temporal_frames(st2, 10, 3)
            f_ms, f_cells = full_reevaluation(st2, 5, 13)
            st2.close()
            tvf["capsules64_L10_1024"] = {"temporal_set_inversion_ms": t_ms, "full_set_inversion_ms": f_ms,
                                          "speedup": f_ms / t_ms, "temporal_cell_evals": t_cells,
                                          "full_cell_evals": f_cells}
    st.close()

    ops = tot["fwd_ops"] + tot["rev_ops"]
    kernel_s = tot["kernel_ms"] * 1e-3
    achieved = ops / kernel_s / 1e12 if kernel_s > 0 else 0.0
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
                     "achieved": achieved, "peak": PEAK_VALU_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_VALU_TFLOPS,
                     "frac_vs_unpacked_issue": achieved / PEAK_VALU_UNPACKED_TOPS,
                     "peak_unpacked_issue": PEAK_VALU_UNPACKED_TOPS,
                     "traffic": traffic, "traffic_unit": "HBM bytes per launch", "traffic_source": traffic_src,
                     "algorithmic_ops_per_launch": ops / launches,
                     "ops_per_cell_eval": ops / max(tot["cells"], 1),
                     # HIP events right before and after each level launch, over the timed frames
                     "avg_launch_us": tot["kernel_ms"] * 1e3 / launches,
                     # the same kernel's average over the timed frames of the committed rocprofv3 summary
                     "profile_avg_kernel_us": prof_avg_us,
                     "eval_ms_per_frame": tot["kernel_ms"] / args.steps,
                     "eval_span_ms_per_frame": tot["eval_ms"] / args.steps,
                     "ops_per_frame": ops / args.steps},
        "roofline_trace": {"bound": "hbm", "kernel": "k_trace / k_trace_quads (the form the runtime picks per frame)",
                           "achieved": trace_gbs, "peak": PEAK_HBM_GBS,
                           "unit": "GB/s", "frac": trace_gbs / PEAK_HBM_GBS,
                           "trace_ms_per_frame": trace["ms"] / args.steps,
                           "node_visits_per_frame": trace["visits"] / args.steps},
    }
    if tvf is not None:
        out["temporal_vs_full"] = tvf
    if not args.no_c1:
        out["gpu_c1"] = gpu_c1(T)
    if not args.no_batched:
        print("batched form", file=sys.stderr, flush=True)
        out["batched"] = batched_form(T, args)
    if not args.no_served:
        print("served form", file=sys.stderr, flush=True)
        out["served"] = served_form(T, args)
    if not args.no_cinematic:
        print("cinematic path tracer", file=sys.stderr, flush=True)
        out["cinematic"] = cinematic_form(T, args)
    if cpu is not None:
        out["cpu_baseline"] = cpu
        out["cpu_baseline_c1"] = cpu_c1
        if cpu_cin is not None:
            out["cpu_baseline_cinematic"] = cpu_cin
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
