"""Frame loop of src/main.rs (1056-1087) without a window: update -> compute -> render per frame,
at t_k = k/60 (main.rs:1062), with the reference's optional outputs -- images/{k}.png
(main.rs:693-715) and timing.txt (main.rs:952-953, 1071-1073).

    python -m temporalsetinversion_amd.run scene.ll --frames 60 --localsubdiv 8 --size 1024 \\
        --images images --timing timing.txt
"""
from __future__ import annotations

import argparse
import os
import time

import numpy as np

from . import State, diagonal_camera, timing_line
from . import scenes


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("scene", help="a .ll file, or the name of a generated scene: " + ", ".join(sorted(scenes.SCENES)))
    ap.add_argument("--frames", type=int, default=60)
    ap.add_argument("--localsubdiv", type=int, default=8)
    ap.add_argument("--size", type=int, nargs="+", default=[1024], help="W [H]")
    ap.add_argument("--delta", type=float, default=0.01)
    ap.add_argument("--swept", action="store_true")
    ap.add_argument("--lipschitz", action="store_true")
    ap.add_argument("--images", help="directory for {frame}.png")
    ap.add_argument("--timing", help="timing.txt path")
    ap.add_argument("--engine", default="compiled", choices=["compiled", "interpreter"])
    args = ap.parse_args(argv)
    text = open(args.scene).read() if os.path.exists(args.scene) else scenes.SCENES[args.scene]()
    W = args.size[0]
    H = args.size[1] if len(args.size) > 1 else W
    if args.images:
        os.makedirs(args.images, exist_ok=True)
    timing = open(args.timing, "w") if args.timing else None
    accum = 0.0
    with State(text, width=W, height=H, localsubdiv=args.localsubdiv, delta=args.delta,
               swept_volume=args.swept, lipschitz_corrected=args.lipschitz, engine=args.engine) as st:
        last = time.perf_counter()
        for k in range(args.frames):
            now = time.perf_counter()
            dt_ms = int((now - last) * 1e6) / 1000.0   # dt.as_micros() as f64 / 1000.0 (main.rs:1072)
            last = now
            st.update(float(np.float32(k * np.float32(1.0 / 60.0))), diagonal_camera())
            accum += st.compute()
            if timing:
                timing.write(timing_line(k, accum, dt_ms))
            st.render(pixels=False)
            if args.images:
                st.save_png(os.path.join(args.images, f"{k}.png"))
    if timing:
        timing.close()
    print(f"{args.frames} frames, set inversion {accum:.3f} ms in total")


if __name__ == "__main__":
    main()
