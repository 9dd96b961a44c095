"""Autotuner of the micrograph batch (gtomx_mpick): for each image side it times
the 2D search's orientation loop (gtomx_mpick_tune: one warm-up, then timed
runs with HIP events) at every candidate batch that fits, with each row-pass
kernel (GTOMX_ROWS_PLAIN / GTOMX_ROWS_PREFETCH / GTOMX_ROWS_ONE at 2048, 3072, 4096), and writes the fastest
(batch, variant) per side to gtom_amd/tuning/micrograph_batches.json, which
MicrographBatch users (bench.py's micrograph row) read.

    python -m gtom_amd.tune_micrographs [--sides 1024 2048 3072 4096] [--out PATH]

The projector is the 64^3 random-blob reference's (p = 128); the timings do not
depend on the data, only on the sizes.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning", "micrograph_batches.json")
CANDIDATES = {1024: [16, 32, 64, 128, 240], 2048: [8, 16, 32, 64], 3072: [4, 8, 16, 32], 4096: [4, 8, 16, 24]}


def load_table(path=TABLE):
    """{side: (batch, rows variant)} from the committed table, or {} when there is none."""
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        js = json.load(f)
    return {int(k): (int(v["batch"]), int(v.get("variant", 0))) for k, v in js["sides"].items()}


def tune(sides, box=64, orientations=64, repeats=2, log=print):
    import torch
    import gtom_amd as gx
    from gtom_amd import synth
    dev = torch.device("cuda:0")
    p = 2 * box
    t = synth.blob_template(box, np.random.default_rng(0), radius=min(14.0, box / 2 - 6))
    projft = gx.prepare_projector(torch.from_numpy(t).to(dev), p)
    angles = gx.fibonacci_angles(orientations // 8, 8)
    out = {}
    for n in sides:
        batches = CANDIDATES.get(n, [8, 16, 32, 64])
        variants = [gx.ROWS_PLAIN, gx.ROWS_PREFETCH] if n > 256 else [gx.ROWS_PLAIN]
        if n in (2048, 3072, 4096):
            variants.append(gx.ROWS_ONE)
        cands = [(b, v) for v in variants for b in batches]
        t0 = time.time()
        ms = gx.mpick_tune(p, n, projft, [b for b, _ in cands], angles, repeats=repeats,
                           variants=[v for _, v in cands])
        best = int(np.argmin(ms))
        out[str(n)] = {"batch": int(cands[best][0]), "variant": int(cands[best][1]),
                       "ms_per_orientation": {f"{b}/{v}": (None if not np.isfinite(m) else round(m, 5))
                                              for (b, v), m in zip(cands, ms)}}
        log(f"[tune] {n}^2: {dict(zip([f'{b}/{v}' for b, v in cands], [round(m, 4) for m in ms]))} "
            f"ms/orientation (batch/variant) -> {cands[best]} ({time.time() - t0:.1f} s)")
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--sides", type=int, nargs="+", default=[1024, 2048, 3072, 4096])
    ap.add_argument("--out", default=TABLE)
    args = ap.parse_args(argv)
    sides = tune(args.sides, log=lambda m: print(m, file=sys.stderr, flush=True))
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    with open(args.out, "w") as f:
        json.dump({"projector": 128, "orientations_timed": 64, "sides": sides}, f, indent=1)
    print(json.dumps(sides))
    return 0


if __name__ == "__main__":
    sys.exit(main())
