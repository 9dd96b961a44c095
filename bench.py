#!/usr/bin/env python3
"""Throughput of the template-matching path on one MI355X.

Workload (the 256^3 configuration): a 512x512x256 tomogram processed as two
256^3 blocks with 40 planted copies of a 48^3 random-blob template (SNR 0.05),
a per-block CTF (defocus U(2,5) um, 10 A/px, 300 kV), and 9,000 orientations
(500 Fibonacci-sphere directions x 18 in-plane steps); projector box p = 96
(oversampling 0.375), local-std radius 24. One step is one full
d_PickSubTomograms call over both blocks and all orientations, with the block
spectra, CTFs and projector already resident in HBM.

Prints one JSON line. `value` is voxel-orientations per second (2 * 256^3 *
9000 per step); `roofline` is the search pipeline's: the algorithmic bytes of one
batch of orientations over one block divided by the summed time of the batch's
kernels, from HIP events recorded around every launch on the stream the search
runs on, with each kernel's own figures as a breakdown;
`cpu_baseline` times the oracle's float32 restatement on this host's cores over
a bounded sample (one block, the first K orientations).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

METRIC = "voxel-orientations/s, 256³ tomo template match; learned vs plain at ≥99% recall"
HBM_PEAK_GBS = 8000.0
BF16_MFMA_PEAK_TFLOPS = 2500.0  # dense
STAGES = ("k_pick_project_z", "k_pick_y", "k_pick_x_max")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=0, help="orientations per pass (0: library default)")
    ap.add_argument("--angles", type=int, default=0,
                    help="search only the first N orientations (profiling runs; not a bench line)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target length of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB device copy (calibrates rocprof byte counters)")
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--search", choices=["plain", "learned"], default="plain",
                    help="learned: the learned-proposal search against the exhaustive one on held-out blocks of the "
                         "learned configuration (one JSON line of its own)")
    ap.add_argument("--learned-blocks", type=int, default=10,
                    help="held-out blocks timed by the learned comparison")
    ap.add_argument("--row", choices=["prob_picker", "picker", "pick2d", "diff2", "local_peaks", "fft_micrograph",
                                      "prob_picker_micrograph", "pick2d_micrograph", "micrographs", "split", "pnet",
                                      "pick_large", "tomo_picker", "pick_generic"],
                    default=None,
                    help="measure a row next to the search instead (one JSON line of its own)")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def copy_probe():
    import torch
    n = 1 << 28  # 1 GiB of float32
    a = torch.ones(n, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    log(f"copy probe: 5 x {4 * n} bytes read and written")


def disc_bins(n):
    """Half-spectrum bins in the columns (x, y) with x^2 + y^2 <= (n/2)^2: the
    only columns that hold bins of the sphere |k| <= n/2, and so the only ones
    pass 1 reads E and the CTF for and writes, and pass 2 reads."""
    x = np.arange(n // 2 + 1)[None, :]
    y = np.fft.fftfreq(n, 1.0 / n)[:, None]
    return int(np.count_nonzero(x * x + y * y <= (n // 2) ** 2)) * n


def algorithmic_bytes(stage, n, nb):
    """Minimum HBM bytes of one launch of a stage for nb orientations of one n^3 volume."""
    H = (n // 2 + 1) * n * n
    Hd = disc_bins(n)
    V = n ** 3
    if stage == 0:  # read E (8 B/bin) and the CTF (4 B/bin) once, write nb column-transformed spectra
        return 12 * Hd + 8 * Hd * nb
    if stage == 1:  # read the disc columns, write the full y-transformed spectra
        return 8 * Hd * nb + 8 * H * nb
    return 8 * H * nb + 16 * V  # read nb spectra, read+write best and argmax


def b_algo_batch(n, b):
    """Algorithmic HBM bytes of one batch of b orientations over one n^3 volume, the whole
    search pipeline: n^3 * (16h*b + 8h + 20) with h = (n/2+1)/n -- per orientation the
    product spectrum written and read once (16h per voxel), per batch the block's
    half spectrum read once (8h) and the epilogue (sigma^-1 4, best 8, argmax 8)."""
    h = (n // 2 + 1) / n
    return n ** 3 * (16 * h * b + 8 * h + 20)


def measured_traffic(kernel, n, batch):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary
    (profiles/pick_cfg2_traffic.json, calibrated against a plain copy), when it was taken at the
    same launch shape; None otherwise."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "pick_cfg2_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        js = json.load(f)
    if js.get("block") != n or js.get("batch") != batch:
        return None
    recs = [v for k, v in js["kernels"].items() if k.startswith(f"gx::{kernel}<{n}")]
    return int(recs[0]["traffic_bytes"]) if recs else None


def pipeline_roofline(n, nvol, nang, batch, steps, ms, cnt, ms_per_step):
    """Roofline of the search pipeline: one 'launch' is one batch of orientations of one
    volume through its kernels (pass 1, plane correction + pass 2, pass 3); its algorithmic
    bytes are the single-round-trip figure b_algo_batch, its time the sum of the stages'
    live HIP-event times. Per-stage figures on each kernel's own I/O follow as a breakdown."""
    sizes = [min(batch, nang - b0) for b0 in range(0, nang, batch)]
    launches = len(sizes) * nvol * steps
    assert all(c == launches for c in cnt), (cnt, launches)
    total_ms = sum(ms)
    algo = sum(b_algo_batch(n, b) for b in sizes) * nvol * steps
    per_launch_ms = total_ms / launches
    achieved = algo / launches / (per_launch_ms / 1e3) / 1e9
    traffic = [measured_traffic(k, n, batch) for k in ("k_pick_project_z", "k_pick_herm_norm", "k_pick_y",
                                                       "k_pick_x_max")]
    stages = {}
    for k, name in enumerate(STAGES):
        own = sum(algorithmic_bytes(k, n, b) for b in sizes) * nvol * steps / launches
        avg = ms[k] / launches
        gbs = own / (avg / 1e3) / 1e9
        t = measured_traffic(name, n, batch)
        if k == 1:
            name = "k_pick_herm_norm + k_pick_y"
            t = None if t is None or traffic[1] is None else t + traffic[1]
        stages[name] = {"avg_launch_ms": round(avg, 4), "own_algorithmic_bytes": int(own),
                        "achieved_GBs": round(gbs, 1), "frac_own": round(gbs / HBM_PEAK_GBS, 4), "traffic": t,
                        "ms_per_step": round(ms[k] / steps, 2)}
    vo_step = nvol * n ** 3 * nang
    e2e = b_algo_batch(n, batch) / (n ** 3 * batch) * vo_step / (ms_per_step / 1e3) / 1e9
    return {"bound": "hbm", "kernel": "search pipeline per batch (k_pick_project_z, k_pick_herm_norm, k_pick_y, "
                                      "k_pick_x_max)",
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None if None in traffic else int(sum(traffic)),
            "avg_launch_ms": round(per_launch_ms, 4), "algorithmic_bytes_per_launch": int(algo / launches),
            "algorithmic_bytes_per_vo": round(algo / (vo_step * steps), 3),
            "end_to_end_frac": round(e2e / HBM_PEAK_GBS, 4), "stages": stages}


def cpu_baseline(w, target_seconds):
    from oracle import pyoracle as po
    import torch
    n, p = w["n"], w["p"]
    # Same float32 inputs as the GPU run, block 0.
    E = w["exp_ft"][:1].cpu().numpy()
    P = w["projft"].cpu().numpy()
    C = w["ctf"][0].cpu().numpy()
    angles = w["angles"].astype(np.float64)
    threads = po.num_threads()
    probe = max(1, min(4, threads // 4))
    r = po.pick(P, p, w["oversample"], E, C, n, angles[:probe], w["maskradius"], dtype=np.float32)
    per = r["loop_seconds"] / probe
    k = int(max(2, min(len(angles), math.ceil(target_seconds / max(per, 1e-6)))))
    log(f"cpu baseline: {per:.3f} s/orientation on {threads} threads, timing {k} orientations")
    r = po.pick(P, p, w["oversample"], E, C, n, angles[:k], w["maskradius"], dtype=np.float32)
    del torch
    return {"value": n ** 3 * k / r["loop_seconds"], "unit": "voxel-orientations/s", "cores": threads,
            "kind": "port",
            "sample": f"1 block 256^3 x first {k} orientations, float32 oracle restatement, "
                      f"orientation loop {r['loop_seconds']:.2f} s"}


def learned_comparison(nblocks):
    """The learned search (gtomx_pick_subtomograms_learned with the committed trained network)
    against the exhaustive search on held-out blocks of the learned configuration (256^3, 48^3
    template, 9000 orientations, SNR 0.3): per block both are timed end to end on the host clock
    (the learned time includes the network and its one synchronisation), and the exhaustive
    search's peaks (d_LocalPeaks above the workload's threshold) are recalled when the learned
    map holds the same value and orientation there. Two operating points: the shortlist point
    (top-16 directions, fallback below confidence 0.5) and the recall point (a tile is skipped only
    when the network gives "no peak" at least the recall confidence, every other tile is searched
    over every direction); the metric's ratio is the recall point's. None without trained weights."""
    import torch
    from gtom_amd import proposal as pr
    from gtom_amd import train_proposal as tp
    if not os.path.exists(tp.WEIGHTS):
        log("no trained proposal network committed; learned comparison skipped")
        return None
    w = tp.Workload()
    net = pr.ProposalNet(max_tiles=512, params=np.load(tp.WEIGHTS)["params"])
    seeds = [tp.HELDOUT_SEEDS[0] + k for k in range(nblocks + 1)]
    tp.recall(w, net, seeds[:1])  # warm-up (plans, kernels)
    points = {}
    for name, (topk, conf) in (("shortlist", (tp.SHORTLIST_TOPK, tp.SHORTLIST_CONFIDENCE)),
                               ("recall", (tp.RECALL_TOPK, tp.RECALL_CONFIDENCE))):
        res = tp.recall(w, net, seeds[1:], topk, conf)
        ex = sum(r["exhaustive_s"] for r in res)
        le = sum(r["learned_s"] for r in res)
        peaks = sum(r["peaks"] for r in res)
        found = sum(r["found"] for r in res)
        vo = len(res) * tp.N ** 3 * tp.NDIRS * tp.NPSI
        points[name] = {"topk": topk, "confidence": conf, "exhaustive_vo_per_s": vo / ex,
                        "learned_exhaustive_equivalent_vo_per_s": vo / le, "learned_over_plain": ex / le,
                        "recall": found / max(peaks, 1), "peaks": peaks, "recalled": found,
                        "tiles": {k: sum(r["stats"][k] for r in res) for k in res[0]["stats"]},
                        "exhaustive_s": round(ex, 3), "learned_s": round(le, 3)}
        log(f"learned ({name}): {points[name]}")
    net.close()
    torch.cuda.synchronize()
    head = points["recall"]
    return {"workload": f"{nblocks} held-out 256^3 blocks (seeds {seeds[1]}..{seeds[-1]}), 20 copies of the "
                        f"48^3 template at SNR {tp.SNR}, 9000 orientations, proposals per 32^3 tile",
            "exhaustive_vo_per_s": head["exhaustive_vo_per_s"],
            "learned_exhaustive_equivalent_vo_per_s": head["learned_exhaustive_equivalent_vo_per_s"],
            "learned_over_plain": head["learned_over_plain"], "target_ratio": 5.0, "recall": head["recall"],
            "peaks": head["peaks"], "recalled": head["recalled"], "exhaustive_s": head["exhaustive_s"],
            "learned_s": head["learned_s"], "operating_points": points}


def row_learned(args):
    cmp_ = learned_comparison(args.learned_blocks)
    if cmp_ is None:
        return 1
    out = {"row": "learned", "metric": METRIC, "value": cmp_["learned_exhaustive_equivalent_vo_per_s"],
           "unit": "voxel-orientations/s (exhaustive-equivalent)", "n_gpus": 1, "steps": args.learned_blocks,
           "warmup": 1, "ms_per_step": cmp_["learned_s"] * 1e3 / args.learned_blocks, "higher_is_better": True,
           "dtype": "f32 search, bf16 network", "data": "synthetic", "config": {"workload": cmp_["workload"]},
           "learned": cmp_}
    print(json.dumps(out), flush=True)
    return 0


def _blob(box, rng, n=10):
    g = np.arange(box) - box // 2
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    t = np.zeros((box, box, box))
    for _ in range(n):
        c = rng.uniform(-box / 4, box / 4, 3)
        s = rng.uniform(1.0, box / 10, 3)
        t += np.exp(-(((x - c[0]) / s[0]) ** 2 + ((y - c[1]) / s[1]) ** 2 + ((z - c[2]) / s[2]) ** 2) / 2)
    return t.astype(np.float32)


def _timed_steps(args, step):
    import torch
    for i in range(args.warmup):
        step()
        torch.cuda.synchronize()
        log(f"warmup step {i + 1}/{args.warmup} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
        torch.cuda.synchronize()
        log(f"step {i + 1}/{args.steps} done")
    return (time.perf_counter() - t0) / args.steps


def _row_line(args, row, metric, unit, work, sec, config, cpu, extra=None):
    out = {"row": row, "metric": metric, "value": work / sec, "unit": unit, "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": sec * 1e3, "higher_is_better": True, "dtype": "f32",
           "data": "synthetic", "config": config}
    if extra:
        out.update(extra)
    if cpu is not None and not args.no_cpu_baseline:
        from oracle import pyoracle as po
        cw, ct, sample = cpu()
        out["cpu_baseline"] = {"value": cw / ct, "unit": unit, "cores": po.num_threads(), "kind": "port",
                               "sample": sample}
    print(json.dumps(out), flush=True)
    return 0


def row_picker(args):
    """The Picker class, 3D: a 128^3 image, a 32^3 reference, anglestep 0.5 rad
    (648 orientations). One step is PerformCorrelation over all of them."""
    import torch
    import gtom_amd as gx
    rng = np.random.default_rng(args.seed)
    nr, di, step = 32, (128, 128, 128), 0.5
    ref = _blob(nr, rng)
    image = rng.standard_normal(di).astype(np.float32)
    dev = torch.device("cuda:0")
    ctf = np.ones((nr, nr, nr // 2 + 1), np.float32)
    pk = gx.Picker(torch.from_numpy(ref).to(dev), di)
    dimg = torch.from_numpy(image.copy()).to(dev)
    dctf = torch.from_numpy(ctf).to(dev)
    pk.set_image(dimg, dctf)
    best = torch.full(di, -1e30, device=dev)
    ang = torch.zeros(di + (3,), device=dev)
    angles = gx.equal_angular_spacing((0, 2 * np.pi), (0, np.pi), (0, 2 * np.pi), step)
    sec = _timed_steps(args, lambda: pk.perform_correlation(best, ang, anglestep=step))
    pk.close()
    vox = int(np.prod(di))

    def cpu():
        from oracle import pyoracle as po
        k = 40
        t0 = time.perf_counter()
        po.picker(ref.astype(np.float64), image.astype(np.float64), ctf.astype(np.float64), angles[:k])
        ct = time.perf_counter() - t0
        return vox * k, ct, f"same image and reference, the first {k} orientations, FP64 oracle, {ct:.1f} s"
    return _row_line(args, "picker", "image voxel-orientations/s, Picker 3D", "voxel-orientations/s",
                     vox * len(angles), sec, {"workload": "Picker.PerformCorrelation", "image": list(di),
                                              "ref": [nr] * 3, "anglestep": step, "orientations": len(angles)}, cpu)


def row_tomo_picker(args):
    """The TomoPicker class: 41 tilt images of 512^2 (tilt -60 ... 60 degrees, weights 0.5 ... 1), a
    32^3 blob reference with a spherical mask of radius 14 (circular), a 512 x 512 x 128 correlation
    volume, the 588 orientations of anglestep 30 degrees. One step is PerformCorrelation over all of
    them, writing the maximum, its angle and index and the two sums."""
    import torch
    import gtom_amd as gx
    rng = np.random.default_rng(args.seed)
    nr, nx, ny, nimg, dc = 32, 512, 512, 41, (512, 512, 128)
    g = np.arange(nr) - nr // 2
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    ref = _blob(nr, rng)
    mask = (x * x + y * y + z * z <= 14 * 14).astype(np.float32)
    images = rng.standard_normal((nimg, ny, nx)).astype(np.float32)
    tilts = np.array([[0.0, np.deg2rad(t), 0.0] for t in np.linspace(-60.0, 60.0, nimg)], np.float32)
    w = rng.uniform(0.5, 1.0, nimg).astype(np.float32)
    angles = gx.equal_angular_spacing((0, 2 * np.pi), (0, np.pi), (0, 2 * np.pi), np.deg2rad(30.0))
    dev = torch.device("cuda:0")
    tp = gx.TomoPicker(torch.from_numpy(ref).to(dev), torch.from_numpy(mask).to(dev), (ny, nx), nimg)
    tp.set_image(torch.from_numpy(images).to(dev), tilts, w)
    shape = (dc[2], dc[1], dc[0])
    corr = torch.empty(shape, device=dev)
    ang = torch.empty(shape + (3,), device=dev)
    idx = torch.empty(shape, dtype=torch.int32, device=dev)
    s1, s2 = torch.empty_like(corr), torch.empty_like(corr)

    def step():
        corr.fill_(-1e30)
        tp.perform_correlation(corr, ang, angles=angles, index=idx, sum1=s1, sum2=s2)
    sec = _timed_steps(args, step)
    tp.close()
    vox = int(np.prod(dc))
    # the update's HBM bytes per voxel and orientation: maximum, sums read and written (20 B); the
    # angle and index are written only where the maximum rises, and the 41 maps (43 MB) stay in cache
    achieved = vox * len(angles) * 20 / sec / 1e9

    def cpu():
        from oracle import pyoracle as po
        k = 2
        t0 = time.perf_counter()
        po.tomo_picker(ref.astype(np.float64), mask.astype(np.float64), True, images.astype(np.float64),
                       tilts.astype(np.float64), w.astype(np.float64), angles[:k].astype(np.float64), dc)
        ct = time.perf_counter() - t0
        return vox * k, ct, (f"same images, reference and volume, the first {k} orientations, FP64 oracle, "
                             f"{ct:.1f} s including its set-up")
    return _row_line(args, "tomo_picker", "correlation-volume voxel-orientations/s, TomoPicker",
                     "voxel-orientations/s", vox * len(angles), sec,
                     {"workload": "TomoPicker.PerformCorrelation", "tilt_images": [nimg, ny, nx], "ref": [nr] * 3,
                      "volume": list(dc), "anglestep_deg": 30, "orientations": len(angles)}, cpu,
                     extra={"roofline": {"bound": "hbm", "kernel": "whole step over the update's bytes (20 B per "
                                                                   "voxel-orientation)",
                                         "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                         "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None}})


def row_pick_large(args):
    """d_PickSubTomograms on a 512^3 volume (the generic-transform form above 256): one volume of
    noise with 8 planted copies of a 48^3 blob template (projector 96^3, oversample 96/512), a CTF
    (defocus 3 um, 10 A/px), the first 1,152 orientations of the 500-direction x 18 set (batch 15),
    local-std radius 24."""
    import torch
    import gtom_amd as gx
    from gtom_amd import synth
    rng = np.random.default_rng(args.seed)
    n, box = 512, 48
    p = 2 * box
    dev = torch.device("cuda:0")
    template = synth.blob_template(box, rng, radius=14.0)
    angles = gx.fibonacci_angles(500, 18)[:1152]
    pos = [tuple(int(v) for v in rng.integers(box, n - box, 3)) for _ in range(8)]
    ori = [tuple(float(a) for a in angles[int(k)]) for k in rng.integers(0, len(angles), 8)]
    tomo = (synth.plant((n, n, n), template, pos, ori, 0.3) + rng.standard_normal((n, n, n), dtype=np.float32))
    E = gx.d_FFTR2C(torch.from_numpy(tomo.astype(np.float32)[None]).to(dev))
    del tomo
    ctf = gx.ctf_3d([gx.CtfParams.make(pixelsize=10e-10, Cs=2.7e-3, voltage=300e3, defocus=-3e-6, amplitude=0.07)],
                    n, device=dev)
    projft = gx.prepare_projector(torch.from_numpy(template).to(dev), p)
    plan = gx.PickPlan(p, n)
    best = torch.empty((1, n, n, n), device=dev)
    idx = torch.empty((1, n, n, n), dtype=torch.int32, device=dev)
    sec = _timed_steps(args, lambda: gx.d_PickSubTomograms(plan, projft, p / n, E, ctf, angles, 24.0, best=best,
                                                            bestidx=idx))
    batch = plan.batch
    plan.close()
    work = n ** 3 * len(angles)
    return _row_line(args, "pick_large", "voxel-orientations/s, d_PickSubTomograms on a 512^3 volume",
                     "voxel-orientations/s", work, sec,
                     {"workload": "one 512^3 volume, 48^3 template, 1152 orientations", "volume": [n] * 3,
                      "orientations": len(angles), "batch": batch}, None)


def row_pick_generic(args):
    """d_PickSubTomograms on a 200^3 volume, a side without a fused search (the long-row form on
    chirp-z transforms): one volume of noise with 4 planted copies of a 32^3 blob template
    (projector 64^3, oversample 64/200), a CTF (defocus 3 um, 10 A/px), the first 576 orientations of
    the 500-direction x 18 set, local-std radius 16."""
    import torch
    import gtom_amd as gx
    from gtom_amd import synth
    rng = np.random.default_rng(args.seed)
    n, box = 200, 32
    p = 2 * box
    dev = torch.device("cuda:0")
    template = synth.blob_template(box, rng, radius=10.0)
    angles = gx.fibonacci_angles(500, 18)[:576]
    pos = [tuple(int(v) for v in rng.integers(box, n - box, 3)) for _ in range(4)]
    ori = [tuple(float(a) for a in angles[int(k)]) for k in rng.integers(0, len(angles), 4)]
    tomo = (synth.plant((n, n, n), template, pos, ori, 0.3) + rng.standard_normal((n, n, n), dtype=np.float32))
    tomo = tomo.astype(np.float32)
    E = gx.d_FFTR2C(torch.from_numpy(tomo[None]).to(dev))
    ctf = gx.ctf_3d([gx.CtfParams.make(pixelsize=10e-10, Cs=2.7e-3, voltage=300e3, defocus=-3e-6, amplitude=0.07)],
                    n, device=dev)
    projft = gx.prepare_projector(torch.from_numpy(template).to(dev), p)
    plan = gx.PickPlan(p, n)
    best = torch.empty((1, n, n, n), device=dev)
    idx = torch.empty((1, n, n, n), dtype=torch.int32, device=dev)
    sec = _timed_steps(args, lambda: gx.d_PickSubTomograms(plan, projft, p / n, E, ctf, angles, 16.0, best=best,
                                                            bestidx=idx))
    batch = plan.batch
    plan.close()
    work = n ** 3 * len(angles)

    def cpu():
        from oracle import pyoracle as po
        k = 2
        t0 = time.perf_counter()
        po.pick(projft.cpu().numpy().astype(np.complex128), p, p / n, E.cpu().numpy().astype(np.complex128),
                ctf.cpu().numpy()[0].astype(np.float64), n, angles[:k].astype(np.float64), 16.0)
        ct = time.perf_counter() - t0
        return n ** 3 * k, ct, f"same volume, the first {k} orientations, FP64 oracle, {ct:.1f} s"
    return _row_line(args, "pick_generic", "voxel-orientations/s, d_PickSubTomograms on a 200^3 volume (chirp-z form)",
                     "voxel-orientations/s", work, sec,
                     {"workload": "one 200^3 volume, 32^3 template, 576 orientations", "volume": [n] * 3,
                      "orientations": len(angles), "batch": batch}, cpu)


def row_pick2d(args):
    """The 2D branch of d_PickSubTomograms: 16 images of 256^2, a 64^3 template
    (projector 128^3, oversample 0.5), 2,000 orientations (250 Fibonacci
    directions x 8 in-plane steps), one CTF, local-std radius 32."""
    import torch
    import gtom_amd as gx
    from oracle import pyoracle as po
    rng = np.random.default_rng(args.seed)
    n, box, nimg = 256, 64, 16
    p = 2 * box
    dev = torch.device("cuda:0")
    angles = gx.fibonacci_angles(250, 8)
    projft = gx.prepare_projector(torch.from_numpy(_blob(box, rng)).to(dev), p)
    imgs = rng.standard_normal((nimg, n, n)).astype(np.float32)
    E = gx.d_FFTR2C(torch.from_numpy(imgs).to(dev), ndims=2)
    ctf = gx.ctf_3d([gx.CtfParams.make(pixelsize=5e-10, Cs=2.7e-3, voltage=300e3, defocus=-2e-6, amplitude=0.07)],
                    n, device=dev, ndims=2)
    from gtom_amd import tune_micrographs as tm
    batch, variant = tm.load_table().get(n, (0, gx.ROWS_PLAIN))  # the autotuner's choice for this side
    plan = gx.PickPlan(p, n, batch, ndims=2)
    plan.set_rows_variant(variant)
    best = torch.empty((nimg, n, n), device=dev)
    idx = torch.empty((nimg, n, n), dtype=torch.int32, device=dev)
    sec = _timed_steps(args, lambda: gx.d_PickSubTomograms(plan, projft, p / n, E, ctf, angles, n / 8, best=best,
                                                            bestidx=idx))
    batch = plan.batch
    plan.close()
    work = nimg * n * n * len(angles)

    def cpu():
        k = 200
        P = projft.cpu().numpy()
        E0 = E[:1].cpu().numpy()
        C = ctf[0].cpu().numpy()
        r = po.pick(P, p, p / n, E0, C, n, angles[:k].astype(np.float64), n / 8, ndims=2)
        return n * n * k, r["loop_seconds"], (f"1 image, the first {k} orientations, FP64 oracle, orientation loop "
                                              f"{r['loop_seconds']:.1f} s")
    return _row_line(args, "pick2d", "image pixel-orientations/s, d_PickSubTomograms 2D branch",
                     "pixel-orientations/s", work, sec,
                     {"workload": "d_PickSubTomograms (2D)", "images": nimg, "image": [n, n], "template": [box] * 3,
                      "orientations": len(angles), "batch": batch, "rows_variant": variant}, cpu)


def row_pick2d_micrograph(args):
    """The 2D branch of d_PickSubTomograms at micrograph size: one 4096^2
    micrograph, a 64^3 template (projector 128^3, oversample 1/32), 480
    orientations (60 Fibonacci directions x 8 in-plane steps), one CTF at
    1 A/px, local-std radius 512."""
    import torch
    import gtom_amd as gx
    from oracle import pyoracle as po
    rng = np.random.default_rng(args.seed)
    n, box, nimg = 4096, 64, 1
    p = 2 * box
    dev = torch.device("cuda:0")
    angles = gx.fibonacci_angles(60, 8)
    projft = gx.prepare_projector(torch.from_numpy(_blob(box, rng)).to(dev), p)
    imgs = rng.standard_normal((nimg, n, n)).astype(np.float32)
    E = gx.d_FFTR2C(torch.from_numpy(imgs).to(dev), ndims=2)
    ctf = gx.ctf_3d([gx.CtfParams.make(pixelsize=1e-10, Cs=2.7e-3, voltage=300e3, defocus=-2e-6, amplitude=0.07)],
                    n, device=dev, ndims=2)
    from gtom_amd import tune_micrographs as tm
    batch, variant = tm.load_table().get(n, (0, gx.ROWS_PLAIN))  # the autotuner's choice for this side
    plan = gx.PickPlan(p, n, batch, ndims=2)
    plan.set_rows_variant(variant)
    best = torch.empty((nimg, n, n), device=dev)
    idx = torch.empty((nimg, n, n), dtype=torch.int32, device=dev)
    sec = _timed_steps(args, lambda: gx.d_PickSubTomograms(plan, projft, p / n, E, ctf, angles, n / 8, best=best,
                                                            bestidx=idx))
    batch = plan.batch
    plan.close()
    work = nimg * n * n * len(angles)

    def cpu():
        k = 16
        P = projft.cpu().numpy()
        E0 = E[:1].cpu().numpy()
        C = ctf[0].cpu().numpy()
        r = po.pick(P, p, p / n, E0, C, n, angles[:k].astype(np.float64), n / 8, ndims=2)
        return n * n * k, r["loop_seconds"], (f"the micrograph, the first {k} orientations, FP64 oracle, "
                                              f"orientation loop {r['loop_seconds']:.1f} s")
    return _row_line(args, "pick2d_micrograph", "micrograph pixel-orientations/s, d_PickSubTomograms 2D branch",
                     "pixel-orientations/s", work, sec,
                     {"workload": "d_PickSubTomograms (2D)", "images": nimg, "image": [n, n], "template": [box] * 3,
                      "orientations": len(angles), "batch": batch, "rows_variant": variant}, cpu)


def micrograph_workload(seed, nmic=64, sides=(1024, 2048, 3072, 4096), box=64, copies=200, snr=0.02):
    """The uneven micrograph batch: nmic micrographs of sides drawn uniformly from `sides`, each
    with its own astigmatic CTF (defocus U(0.8, 3.5) um, astigmatism U(0, 100) nm at a random
    angle, 1.0 A/px, 300 kV, Cs 2.7 mm) and `copies` planted projections of a box^3 random-blob
    reference (projections at random orientations through the reference's own projector), the
    signal scaled for `snr` (signal variance inside the support of the projection over the noise
    variance, before the CTF) and CTF-filtered, in unit Gaussian noise. Returns (projector, p, image dicts, sides)."""
    import torch
    import gtom_amd as gx
    from gtom_amd import synth
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(seed)
    p = 2 * box
    t = synth.blob_template(box, np.random.default_rng(seed + 1), radius=min(14.0, box / 2 - 6))
    projft = gx.prepare_projector(torch.from_numpy(t).to(dev), p)
    amp = synth.snr_amplitude(t.sum(axis=0) / np.sqrt(box), snr)
    ns = [int(rng.choice(sides)) for _ in range(nmic)]
    images = []
    gen = torch.Generator(device=dev)
    for k, n in enumerate(ns):
        angs = np.array([synth.random_orientation(rng) for _ in range(copies)], np.float32)
        sec = gx.d_rlnProject(projft, p, p, angles=angs, ndims=2)  # (copies, p, p/2+1) central sections
        proj = torch.fft.fftshift(torch.fft.irfft2(sec, s=(p, p)), dim=(-2, -1)) / np.sqrt(box)
        sig = torch.zeros((n, n), device=dev)
        pos = rng.integers(0, n - p, (copies, 2))
        for c in range(copies):
            sig[pos[c, 1]:pos[c, 1] + p, pos[c, 0]:pos[c, 0] + p] += proj[c]
        par = gx.CtfParamsFull.make(pixelsize=1e-10, Cs=2.7e-3, voltage=300e3,
                                    defocus=-float(rng.uniform(0.8e-6, 3.5e-6)),
                                    defocusdelta=float(rng.uniform(0, 100e-9)),
                                    astigmatismangle=float(rng.uniform(0, np.pi)), amplitude=0.07)
        ctf = gx.ctf_2d([par], n, device=dev)[0].contiguous()
        sig = torch.fft.irfft2(torch.fft.rfft2(sig) * ctf, s=(n, n)) * amp
        gen.manual_seed(seed * 1000 + k)
        img = sig + torch.randn((n, n), generator=gen, device=dev)
        E = gx.d_FFTR2C(img[None], ndims=2)[0].contiguous()
        del img, sig
        images.append({"spectrum": E, "ctf": ctf, "maskradius": n / 8, "best": torch.empty((n, n), device=dev),
                       "index": torch.empty((n, n), dtype=torch.int32, device=dev)})
    return projft, p, images, ns


def row_micrographs(args):
    """Batching: 64 micrographs of uneven sides (1024^2 .. 4096^2), each with its own astigmatic
    CTF, 200 planted projections of a 64^3 random-blob reference at SNR 0.02, searched over 4,000
    orientations (500 Fibonacci directions x 8 in-plane steps) in one gtomx_mpick_run: one plan per
    side with the autotuner's batch and row kernel (gtom_amd/tuning/micrograph_batches.json), each side's
    orientation loop captured once as a hipGraph and replayed per micrograph."""
    import torch
    import gtom_amd as gx
    from gtom_amd import tune_micrographs as tm
    from oracle import pyoracle as po
    t0 = time.time()
    projft, p, images, ns = micrograph_workload(args.seed)
    log(f"micrograph batch built in {time.time() - t0:.1f} s: sides {sorted(set(ns))}, "
        f"{sum(n * n for n in ns):.3g} pixels")
    table = tm.load_table()
    sides = sorted(set(ns))
    batches = [table.get(n, (0, 0))[0] for n in sides]
    variants = [table.get(n, (0, 0))[1] for n in sides]
    mb = gx.MicrographBatch(p, sides, batches, variants)
    angles = gx.fibonacci_angles(500, 8)
    stream = torch.cuda.current_stream()
    sec = _timed_steps(args, lambda: mb.run(projft, images, angles, use_graphs=True, stream=stream))
    info = mb.info()
    # graphs against eager launches on the smallest side, where launch overhead weighs most
    small = [im for im, n in zip(images, ns) if n == sides[0]][:4]
    mb.run(projft, small, angles, use_graphs=False, stream=stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    mb.run(projft, small, angles, use_graphs=False, stream=stream)
    torch.cuda.synchronize()
    eager = time.perf_counter() - t1
    t1 = time.perf_counter()
    mb.run(projft, small, angles, use_graphs=True, stream=stream)
    torch.cuda.synchronize()
    graphed = time.perf_counter() - t1
    mb.close()
    work = sum(n * n for n in ns) * len(angles)
    # algorithmic bytes per pixel-orientation of the 2D search at batch B (the 2D analogue of
    # 16h + (8h + 20)/B, h = (n/2+1)/n): summed over the micrographs at their batches
    algo = 0.0
    for n in ns:
        h = (n // 2 + 1) / n
        b = info["batches"][n]
        algo += n * n * len(angles) * (16 * h + (8 * h + 20) / b)
    achieved = algo / sec / 1e9

    def cpu():
        k = 4
        im = images[ns.index(sides[0])]
        n = sides[0]
        r = po.pick(projft.cpu().numpy(), p, p / n, im["spectrum"][None].cpu().numpy(), im["ctf"].cpu().numpy(), n,
                    angles[:k].astype(np.float64), n / 8, ndims=2)
        return n * n * k, r["loop_seconds"], (f"one {n}^2 micrograph, the first {k} orientations, FP64 oracle, "
                                              f"orientation loop {r['loop_seconds']:.1f} s")
    out_cfg = {"workload": "64 uneven micrographs (1024^2..4096^2, per-micrograph astigmatic CTF, 200 planted "
                           "projections of a 64^3 reference at SNR 0.02) x 4,000 orientations in one batched call",
               "micrographs": len(ns), "sides": {str(n): ns.count(n) for n in sides}, "orientations": len(angles),
               "batches": {str(k): v for k, v in info["batches"].items()},
               "row_variants": {str(k): v for k, v in info["variants"].items()}, "graphs_captured": info["captures"],
               "graph_replays": info["replays"],
               "small_side_check": {"side": sides[0], "micrographs": len(small), "eager_s": round(eager, 3),
                                    "graphs_s": round(graphed, 3)}}
    rc = _row_line(args, "micrographs", "micrograph pixel-orientations/s, batched uneven micrographs",
                   "pixel-orientations/s", work, sec, out_cfg, cpu, extra={
                       "roofline": {"bound": "hbm", "kernel": "whole batched search (every kernel of every batch)",
                                    "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                                    "algorithmic_bytes_per_po": round(algo / work, 3)}})
    return rc


def row_split(args):
    """Precision: the 64^3 configuration (3 planted copies, 588 orientations, CTF) searched with
    every per-orientation inverse transform's radix-16 stage on the matrix cores, fp16 and bf16
    hi/lo split (and hi alone), beside the fp32 search. The CPU baseline leg runs the FP64 oracle on
    the whole workload; its result is also the checker of every precision's error (none without it)."""
    import torch
    import gtom_amd as gx
    from gtom_amd import synth
    dev = torch.device("cuda:0")
    w = synth.cfg1(seed=1)
    n, p, s = w["n"], w["p"], w["oversample"]
    projft = gx.prepare_projector(torch.from_numpy(w["template"]).to(dev), p)
    E = gx.d_FFTR2C(torch.from_numpy(w["tomogram"][None]).to(dev))
    ctf = gx.ctf_3d([gx.CtfParams.make(pixelsize=10e-10, Cs=2.7e-3, voltage=300e3, defocus=-3e-6, amplitude=0.07)],
                    n, device=dev)
    angles = w["angles"]
    ref = cpu = None
    if not args.no_cpu_baseline:
        from oracle import pyoracle as po
        t0 = time.perf_counter()
        ref = po.pick(projft.cpu().numpy().astype(np.complex128), p, s, E.cpu().numpy().astype(np.complex128),
                      ctf.cpu().numpy()[0].astype(np.float64), n, angles.astype(np.float64), w["maskradius"])
        ct = time.perf_counter() - t0
        cpu = {"value": n ** 3 * len(angles) / ct, "unit": "voxel-orientations/s", "cores": po.num_threads(),
               "kind": "port", "sample": f"the whole workload, FP64 oracle, {ct:.1f} s"}
        scale = np.abs(ref["best"]).max()
        sure = (ref["top1"] - ref["top2"]) > 1e-4 * np.abs(ref["top1"]).max()
    # 64 orientations per batch: the split pipeline's fastest (5.66 ms per call against 5.78 at 256);
    # the fp32 search is timed at that batch and, as "fp32-default", at the library's default (256)
    plan = gx.PickPlan(p, n, args.batch or 64)
    plan_default = gx.PickPlan(p, n)
    results = {}
    for name in ("fp32", "fp32-default", "fp16", "bf16", "fp16-hi", "bf16-hi"):
        if name.startswith("fp32"):
            pl = plan if name == "fp32" else plan_default
            step = lambda pl=pl: gx.d_PickSubTomograms(pl, projft, s, E, ctf, angles, w["maskradius"])
        else:
            prec = gx.PRECISIONS[name]
            step = lambda prec=prec: gx.d_PickSubTomogramsSplit(plan, prec, projft, s, E, ctf, angles,
                                                                w["maskradius"])
        best, idx = step()
        torch.cuda.synchronize()
        b, i = best.cpu().numpy(), idx.cpu().numpy()
        sec = _timed_steps(args, step)
        results[name] = {"ms_per_step": round(sec * 1e3, 3), "vo_per_s": n ** 3 * len(angles) / sec}
        if ref is not None:
            results[name].update({"max_cc_error_vs_fp64": float(np.abs(b - ref["best"]).max() / scale),
                                  "argmax_agreement": float((i[sure] == ref["idx"][sure]).mean()),
                                  "argmax_voxels": int(sure.sum())})
        log(f"{name}: {results[name]}")
    batch = plan.batch
    plan.close()
    plan_default.close()
    h = (n // 2 + 1) / n
    algo = 16 * h + (8 * h + 20) / batch
    f16 = results["fp16"]
    out = {"row": "split", "metric": "voxel-orientations/s, 64^3 template match, split-precision matrix-core "
                                     "transforms", "value": f16["vo_per_s"], "unit": "voxel-orientations/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": f16["ms_per_step"],
           "higher_is_better": True, "dtype": "fp16 hi/lo MFMA (radix-16), fp32 accumulation and VALU stages",
           "data": "synthetic",
           "config": {"workload": "64^3 tomogram, 3 planted copies of a 32^3 blob template (projector 64^3), 588 "
                                  "orientations (30-degree GetEqualAngularSpacing), CTF 3 um", "batch": batch},
           "precisions": results,
           "roofline": {"bound": "hbm", "kernel": "split-precision search pipeline",
                        "achieved": round(f16["vo_per_s"] * algo / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(f16["vo_per_s"] * algo / 1e9 / HBM_PEAK_GBS, 4), "traffic": None,
                        "algorithmic_bytes_per_vo": round(algo, 3)}}
    if cpu is not None:
        out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)
    return 0


def pnet_flops_per_tile():
    """Multiply-adds x 2 of one 32^3 tile through the proposal network (forward): four 3^3
    convolutions (1->16 channels at 32^3, 16->16 at 16^3, 8^3, 8^3) and the 1024 x 501 head."""
    conv = lambda side, ci: side ** 3 * 16 * ci * 27 * 2
    fwd = conv(32, 1) + conv(16, 16) + conv(8, 16) + conv(8, 16) + 1024 * 501 * 2
    # training: forward, data gradients of every layer but the first, weight gradients of all
    train = 2 * fwd + (fwd - conv(32, 1))
    return fwd, train


def row_pnet(args):
    """The orientation-proposal network's kernels (bf16 matrix-core implicit GEMMs): inference on
    every 32^3 tile of a 256^3 block (standardisation, forward, top-16 proposals; 512 tiles a step)
    and Adam steps on batches of 256 tiles. Matrix-core FLOP rates are reported against the dense
    bf16 peak; the network is small enough that activations' memory traffic and the launches of its
    15 kernels, not the matrix cores, set its speed."""
    import torch
    from gtom_amd import proposal as pr
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(args.seed)
    net = pr.ProposalNet(max_tiles=512, seed=1)
    vol = torch.from_numpy(rng.standard_normal((256, 256, 256), dtype=np.float32)).to(dev)
    x = torch.from_numpy(rng.standard_normal((256, 32, 32, 32), dtype=np.float32)).to(dev)
    y = torch.from_numpy(rng.integers(0, pr.NCLS, 256).astype(np.int32)).to(dev)
    stream = torch.cuda.current_stream()

    def infer():
        k = net.load_volume(vol, stream=stream)
        net.forward(k, stream=stream)
        net.propose(k, 16, stream=stream, full=True)

    def train():
        net.load_tiles(x, stream=stream)
        net.train_step(y, 1e-4, stream=stream)

    reps = 20
    sec_i = _timed_steps(args, lambda: [infer() for _ in range(reps)]) / reps
    sec_t = _timed_steps(args, lambda: [train() for _ in range(reps)]) / reps
    net.close()
    fwd, trn = pnet_flops_per_tile()
    tf_i = 512 * fwd / sec_i / 1e12
    tf_t = 256 * trn / sec_t / 1e12
    out = {"row": "pnet", "metric": "32^3 tiles/s through the orientation-proposal network (inference)",
           "value": 512 / sec_i, "unit": "tiles/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": sec_i * 1e3, "higher_is_better": True,
           "dtype": "bf16 matrix-core operands, fp32 accumulation", "data": "synthetic (random tiles and labels)",
           "config": {"workload": "inference: every 32^3 tile of one 256^3 block (512 tiles) with top-16 "
                                  "proposals; training: Adam steps on 256-tile batches",
                      "flops_per_tile_forward": fwd, "flops_per_tile_training": trn},
           "training": {"tiles_per_s": 256 / sec_t, "ms_per_step": round(sec_t * 1e3, 3),
                        "tflops": round(tf_t, 2)},
           "roofline": {"bound": "mfma", "kernel": "whole inference pass", "achieved": round(tf_i, 2),
                        "peak": BF16_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": round(tf_i / BF16_MFMA_PEAK_TFLOPS, 5), "traffic": None}}
    print(json.dumps(out), flush=True)
    return 0


def row_diff2(args):
    """d_PickSubTomogramsDiff2: 8 volumes of 64^3, a 32^3 template (projector 64^3,
    oversample 1), relevant shifts 12^3, 500 orientations (50 x 10)."""
    import torch
    import gtom_amd as gx
    from oracle import pyoracle as po
    rng = np.random.default_rng(args.seed)
    n, box, nvol, rel = 64, 32, 8, (12, 12, 12)
    p = 2 * box
    dev = torch.device("cuda:0")
    angles = gx.fibonacci_angles(50, 10)
    projft = gx.prepare_projector(torch.from_numpy(_blob(box, rng)).to(dev), p)
    E = gx.d_FFTR2C(torch.from_numpy(rng.standard_normal((nvol, n, n, n)).astype(np.float32)).to(dev))
    ctf = torch.from_numpy(rng.uniform(-1, 1, (nvol, n, n, n // 2 + 1)).astype(np.float32)).to(dev)
    sec = _timed_steps(args, lambda: gx.d_PickSubTomogramsDiff2(projft, p, p / n, E, ctf, rel, angles))
    shifts = int(np.prod(rel))
    work = nvol * shifts * len(angles)

    def cpu():
        k = 4
        P = projft.cpu().numpy().astype(np.complex128)
        E0 = E[0].cpu().numpy().astype(np.complex128)
        C0 = ctf[0].cpu().numpy().astype(np.float64)
        t0 = time.perf_counter()
        po.diff2(P, p, p / n, E0, C0, n, rel, angles[:k].astype(np.float64))
        ct = time.perf_counter() - t0
        return shifts * k, ct, f"1 volume, the first {k} orientations, FP64 oracle, {ct:.1f} s"
    return _row_line(args, "diff2", "shift-orientations/s, d_PickSubTomogramsDiff2", "shift-orientations/s", work,
                     sec, {"workload": "d_PickSubTomogramsDiff2", "volumes": nvol, "volume": [n] * 3,
                           "relevant": list(rel), "orientations": len(angles)}, cpu)


def row_local_peaks(args):
    """d_LocalPeaks on a 256^3 smooth random map (noise low-passed to about 4
    voxels, normalised), local extent 8, threshold 3."""
    import torch
    import gtom_amd as gx
    from oracle import pyoracle as po
    rng = np.random.default_rng(args.seed)
    n, extent, thr = 256, 8, 3.0
    f = np.fft.rfftn(rng.standard_normal((n, n, n)).astype(np.float32))
    k = np.sqrt(np.fft.fftfreq(n)[:, None, None] ** 2 + np.fft.fftfreq(n)[None, :, None] ** 2 +
                np.fft.rfftfreq(n)[None, None, :] ** 2)
    vol = np.fft.irfftn(f * np.exp(-(k * 16) ** 2), s=(n, n, n)).astype(np.float32)
    vol = ((vol - vol.mean()) / vol.std()).astype(np.float32)
    dev = torch.device("cuda:0")
    dv = torch.from_numpy(vol).to(dev)
    found = []
    sec = _timed_steps(args, lambda: found.append(len(gx.d_LocalPeaks(dv, extent, thr)[0])))
    vox = n ** 3

    def cpu():
        t0 = time.perf_counter()
        po.local_peaks(vol, extent, thr)
        ct = time.perf_counter() - t0
        return vox, ct, f"the same map, FP32 oracle, {ct:.1f} s"
    return _row_line(args, "local_peaks", "voxels/s, d_LocalPeaks", "voxels/s", vox, sec,
                     {"workload": "d_LocalPeaks", "volume": [n] * 3, "localextent": extent, "threshold": thr,
                      "peaks": found[-1]}, cpu)


def row_prob_picker(args):
    """ProbabilisticPicker, 3D: a 128^3 volume (mean 0 / std 1 noise with three
    planted copies), references of 32^3 with a circular mask of radius 14,
    low-pass 24 (rotation at 48^3), anglestep 0.35 rad (1,819 orientations of
    GetEqualAngularSpacing). One step is PerformCorrelation for one reference:
    all orientations, host rand() integers included (the call is synchronous)."""
    import torch
    import gtom_amd as gx
    from oracle import pyoracle as po
    rng = np.random.default_rng(args.seed)
    di, dr, lp, step = (128, 128, 128), (32, 32, 32), 24, 0.35
    g = np.meshgrid(*[np.arange(d) - d // 2 for d in dr], indexing="ij")
    mask = (np.sqrt(sum(a.astype(float) ** 2 for a in g)) < 14).astype(np.float32)
    refs = (rng.standard_normal((2,) + dr).astype(np.float32) * mask)
    image = rng.standard_normal(di).astype(np.float32)
    for c in ((40, 50, 60), (90, 30, 70), (64, 100, 20)):
        image[tuple(slice(x - 16, x + 16) for x in c)] += refs[0][::-1, ::-1, ::-1]
    dev = torch.device("cuda:0")
    pp = gx.ProbabilisticPicker(torch.from_numpy(refs).to(dev), torch.from_numpy(np.stack([mask] * 2)).to(dev), di,
                                lp)
    pp.set_image(torch.from_numpy(image).to(dev))
    nang = pp.count_angles(step)
    best = torch.full(di, -1e30, device=dev)
    ang = torch.zeros(di + (3,), device=dev)
    ref = torch.zeros(di, dtype=torch.int32, device=dev)
    for i in range(args.warmup):
        pp.perform_correlation(i % 2, step, best, ang, ref)
        log(f"warmup step {i + 1}/{args.warmup} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        pp.perform_correlation(i % 2, step, best, ang, ref)
        log(f"step {i + 1}/{args.steps} done")
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / args.steps
    pp.close()
    vox = int(np.prod(di))
    value = vox * nang / sec
    out = {"row": "prob_picker", "metric": "image voxel-orientations/s, ProbabilisticPicker 3D", "value": value,
           "unit": "voxel-orientations/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": sec * 1e3, "higher_is_better": True, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "ProbabilisticPicker.PerformCorrelation", "image": list(di), "ref": list(dr),
                      "lowpassfreq": lp, "anglestep": step, "orientations": nang}}
    if not args.no_cpu_baseline:
        cstep = 1.2  # a bounded sample: the 50 orientations of anglestep 1.2 rad
        ncpu = len(po.equal_angular_spacing((0, 2 * np.pi), (0, np.pi), (0, 2 * np.pi), cstep))
        po.gaussrand(0, reset=True)
        t0 = time.perf_counter()
        po.prob_picker(refs[0].astype(np.float64), mask.astype(np.float64), image.astype(np.float64), lp, cstep)
        ct = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": vox * ncpu / ct, "unit": "voxel-orientations/s", "cores": po.num_threads(),
                               "kind": "port",
                               "sample": f"same volume and reference, the {ncpu} orientations of anglestep {cstep}, "
                                         f"FP64 oracle restatement, {ct:.1f} s including its set-up"}
    print(json.dumps(out), flush=True)
    return 0


def row_fft_micrograph(args):
    """The generic transforms at micrograph size: 8 images of 4096^2, R2C then
    C2R (renormalised) per step; both axes go through the two-pass path."""
    import torch
    import gtom_amd as gx
    from oracle import pyoracle as po
    n, nb = 4096, 8
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(args.seed)
    x = torch.randn((nb, n, n), device=dev, generator=g)
    X = torch.empty((nb, n, n // 2 + 1), dtype=torch.complex64, device=dev)
    y = torch.empty_like(x)
    L = gx.lib()

    def step():
        gx.check(L.gtomx_fft_r2c(gx._ptr(x), gx._ptr(X), 2, gx.int3(n, n, 1), nb, None), "gtomx_fft_r2c")
        gx.check(L.gtomx_ifft_c2r(gx._ptr(X), gx._ptr(y), 2, gx.int3(n, n, 1), nb, 1, None), "gtomx_ifft_c2r")
    sec = _timed_steps(args, step)
    pts = n * n * nb

    def cpu():
        xs = np.random.default_rng(args.seed).standard_normal((1, 1, n, n))
        k, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < 10.0:  # a bounded sample: whole images for about 10 s
            po.c2r(po.r2c(xs), n)
            k += 1
        ct = time.perf_counter() - t0
        return n * n * k, ct, f"{k} 4096^2 images one at a time, R2C + C2R with the FP64 oracle's own FFT, {ct:.1f} s"
    return _row_line(args, "fft_micrograph", "transformed points/s (R2C + C2R), 4096^2 images", "points/s",
                     pts, sec, {"workload": "d_FFTR2C + d_IFFTC2R", "image": [n, n], "batch": nb}, cpu)


def row_prob_picker_micrograph(args):
    """ProbabilisticPicker, 2D, at micrograph size: one 4096^2 micrograph
    (noise with planted copies), a 64^2 reference with a circular mask of
    radius 28, low-pass 32, anglestep 0.05 rad (126 in-plane orientations).
    One step is PerformCorrelation for one reference."""
    import torch
    import gtom_amd as gx
    from oracle import pyoracle as po
    rng = np.random.default_rng(args.seed)
    di, dr, lp, step = (4096, 4096), (64, 64), 32, 0.05
    g = np.meshgrid(*[np.arange(d) - d // 2 for d in dr], indexing="ij")
    mask = (np.sqrt(sum(a.astype(float) ** 2 for a in g)) < 28).astype(np.float32)
    refs = rng.standard_normal((1,) + dr).astype(np.float32) * mask
    image = rng.standard_normal(di).astype(np.float32)
    for _ in range(200):
        cy, cx = rng.integers(32, 4096 - 32, 2)
        image[cy - 32:cy + 32, cx - 32:cx + 32] += 0.5 * refs[0]
    dev = torch.device("cuda:0")
    pp = gx.ProbabilisticPicker(torch.from_numpy(refs).to(dev), torch.from_numpy(mask[None]).to(dev), di, lp)
    pp.set_image(torch.from_numpy(image).to(dev))
    nang = pp.count_angles(step)
    best = torch.full(di, -1e30, device=dev)
    ang = torch.zeros(di + (3,), device=dev)
    ref = torch.zeros(di, dtype=torch.int32, device=dev)
    sec = _timed_steps(args, lambda: pp.perform_correlation(0, step, best, ang, ref))
    cstep = 0.4  # the CPU sample: the orientations of anglestep 0.4 rad
    ncpu = pp.count_angles(cstep)
    pp.close()
    px = int(np.prod(di))

    def cpu():
        po.gaussrand(0, reset=True)
        t0 = time.perf_counter()
        po.prob_picker(refs[0].astype(np.float64), mask.astype(np.float64), image.astype(np.float64), lp, cstep)
        ct = time.perf_counter() - t0
        return px * ncpu, ct, (f"same micrograph and reference, the {ncpu} orientations of anglestep {cstep}, FP64 "
                               f"oracle restatement, {ct:.1f} s including its set-up")
    return _row_line(args, "prob_picker_micrograph", "micrograph pixel-orientations/s, ProbabilisticPicker 2D",
                     "pixel-orientations/s", px * nang, sec,
                     {"workload": "ProbabilisticPicker.PerformCorrelation", "image": list(di), "ref": list(dr),
                      "lowpassfreq": lp, "anglestep": step, "orientations": nang}, cpu)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (orientation sharding across "
              "GPUs is a design note, not built)", file=sys.stderr)
        return 2
    if args.copy_probe:
        copy_probe()
        return 0
    if args.search == "learned":
        return row_learned(args)
    if args.row is not None:
        return {"prob_picker": row_prob_picker, "picker": row_picker, "pick2d": row_pick2d, "diff2": row_diff2,
                "local_peaks": row_local_peaks, "fft_micrograph": row_fft_micrograph,
                "prob_picker_micrograph": row_prob_picker_micrograph,
                "pick2d_micrograph": row_pick2d_micrograph, "micrographs": row_micrographs,
                "split": row_split, "pnet": row_pnet, "pick_large": row_pick_large,
                "tomo_picker": row_tomo_picker, "pick_generic": row_pick_generic}[args.row](args)
    import torch
    import gtom_amd as gx
    from gtom_amd import synth

    t0 = time.time()
    w = synth.cfg2(seed=args.seed)
    log(f"synthetic workload built in {time.time() - t0:.1f} s")
    dev = torch.device("cuda:0")
    n, p = w["n"], w["p"]
    nvol = w["blocks"].shape[0]
    angles = w["angles"] if args.angles <= 0 else w["angles"][:args.angles]
    H = (n // 2 + 1) * n * n
    blocks = torch.from_numpy(w["blocks"]).to(dev)
    exp_ft = gx.d_FFTR2C(blocks)
    del blocks
    projft = gx.prepare_projector(torch.from_numpy(w["template"]).to(dev), p)
    ctf = gx.ctf_3d([gx.CtfParams.make(pixelsize=w["pixelsize"], Cs=2.7e-3, voltage=300e3, defocus=-float(d),
                                       amplitude=0.07) for d in w["defoci"]], n, device=dev)
    plan = gx.PickPlan(p, n, args.batch)
    best = torch.empty((nvol, n, n, n), dtype=torch.float32, device=dev)
    idx = torch.empty((nvol, n, n, n), dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()

    def step():
        gx.d_PickSubTomograms(plan, projft, w["oversample"], exp_ft, ctf, angles, w["maskradius"],
                              ctf_volume_stride=H, best=best, bestidx=idx, stream=stream)

    for i in range(args.warmup):
        step()
        torch.cuda.synchronize()
        log(f"warmup step {i + 1}/{args.warmup} done")
    plan.enable_timing(True)
    plan.timings(reset=True)
    torch.cuda.synchronize()
    # one event per step boundary on the search stream: the mean is the metric, the median is kept beside it
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    evs[0].record(stream)
    for i in range(args.steps):
        step()
        evs[i + 1].record(stream)
    torch.cuda.synchronize()
    step_ms = [evs[i].elapsed_time(evs[i + 1]) for i in range(args.steps)]
    total_ms = evs[0].elapsed_time(evs[-1])
    ms, cnt = plan.timings(reset=True)
    ms_per_step = total_ms / args.steps
    work = nvol * n ** 3 * len(angles)
    value = work / (ms_per_step / 1000.0)
    log(f"{args.steps} steps, {ms_per_step:.1f} ms/step, {value:.4g} vo/s; stage ms {ms}, launches {cnt}")

    roof = pipeline_roofline(n, nvol, len(angles), plan.batch, args.steps, ms, cnt, ms_per_step)

    cpu = None
    if not args.no_cpu_baseline:
        w2 = dict(n=n, p=p, oversample=w["oversample"], maskradius=w["maskradius"], angles=angles,
                  exp_ft=exp_ft, projft=projft, ctf=ctf)
        cpu = cpu_baseline(w2, args.cpu_seconds)

    out = {"metric": METRIC, "value": value, "unit": "voxel-orientations/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded noise + planted blob templates)",
           "config": {"workload": "plain exhaustive search: 512x512x256 tomogram as 2 blocks of 256^3, 48^3 "
                                  "template (projector 96^3, oversample 0.375), 9000 orientations (500 Fibonacci "
                                  "directions x 18 psi), per-block 3D CTF, local-std radius 24",
                      "blocks": nvol, "block": n, "orientations": len(angles), "batch": plan.batch,
                      "search": "plain"},
           "median_ms_per_step": float(np.median(step_ms)), "roofline": roof, "cpu_baseline": cpu}
    plan.close()
    # the learned-vs-plain comparison of the metric, on its own workload (see learned_comparison)
    out["learned"] = learned_comparison(args.learned_blocks) if args.learned_blocks > 0 else None
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
