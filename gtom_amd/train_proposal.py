"""Training and evaluation of the orientation-proposal network on the
exhaustive search's own results.

Workload (the learned-proposal configuration): 256^3 blocks of Gaussian noise
with 20 copies of the fixed 48^3 random-blob template each (per-block CTF with
defocus in [2, 5] um), searched over 500 Fibonacci directions x 18 in-plane
steps with the 96^3 projector (oversample 0.375) and local-std radius 24 -- the
256^3 template-matching configuration, one block per seed, at SNR 0.3. At the
plain configuration's SNR 0.05 the planted copies score 3.8-6.6 on the
normalised map and the noise alone reaches 6.0-6.5 in every block (measured on
four blocks), so the exhaustive search's peaks there are mostly noise that no
predictor can anticipate; at SNR 0.3 the copies score 6.1-12.6 against a noise
maximum of 6.1-6.2, and a threshold of 7.5 keeps the peaks that are particles.

Labels: the exhaustive search's normalised map is reduced to peaks with
d_LocalPeaks (src/Correlation/LocalPeaks.cu:24-62; extent PEAK_EXTENT, values
above PEAK_THRESHOLD). A 32^3 tile is labelled with the direction (argmax
orientation index // 18, src/Correlation/SubTomograms.cu:138-158) of the
strongest peak within LABEL_MARGIN voxels of its box, or "no peak".

    python -m gtom_amd.train_proposal label --first S --count N --out labels.npz
    python -m gtom_amd.train_proposal train --labels a.npz [b.npz ...] --out weights.npz
    python -m gtom_amd.train_proposal eval --weights weights.npz --first S --count N
    python -m gtom_amd.train_proposal scores --weights weights.npz --first S --count N

Peak recall (eval): an exhaustive peak is recalled when the learned search's
map holds the exhaustive value at that voxel (the argmax orientation was
searched there) and d_LocalPeaks finds it on the learned map too.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

N = 256
BOX = 48
P = 2 * BOX
OVERSAMPLE = P / N
NDIRS = 500
NPSI = 18
MASK_RADIUS = 24.0
NCOPIES = 20
SNR = 0.3
PEAK_EXTENT = 12
PEAK_THRESHOLD = 7.5
LABEL_MARGIN = 4
# Operating points of the learned search. Shortlist: a tile takes its top-16 directions or "no peak",
# whichever the network makes more probable, and every direction below confidence 0.5. Recall: with
# top-k = all 500 directions the summed direction probability is 1 - P(no peak), so a tile is
# skipped only when P(no peak) reaches the confidence and every other tile is searched over every
# direction; the confidence is the lowest at which every exhaustive peak of the calibration
# blocks (CALIBRATION_SEEDS, 418 peaks) lies in a searched tile (0.86 loses one).
SHORTLIST_TOPK, SHORTLIST_CONFIDENCE = 16, 0.5
RECALL_TOPK, RECALL_CONFIDENCE = 500, 0.87
CALIBRATION_SEEDS = (300_000, 20)
TRAIN_SEEDS = (100_000, 2000)
HELDOUT_SEEDS = (200_000, 50)
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "proposal_weights.npz")


def log(msg):
    print(f"[proposal] {msg}", file=sys.stderr, flush=True)


class Workload:
    """Device-resident pieces shared by every block: template, projector, orientations, plan."""

    def __init__(self, batch=0):
        import torch
        import gtom_amd as gx
        from . import synth
        self.gx = gx
        self.dev = torch.device("cuda:0")
        t = synth.proposal_template(BOX)
        self.amplitude = synth.snr_amplitude(t, SNR)
        self.template = torch.from_numpy(t).to(self.dev)
        self.projft = gx.prepare_projector(self.template, P)
        self.angles = gx.fibonacci_angles(NDIRS, NPSI)
        self.plan = gx.PickPlan(P, N, batch)

    def block(self, seed):
        from . import synth
        return synth.tomogram_block_gpu(seed, self.template, self.amplitude, n=N, ncopies=NCOPIES)

    def spectra(self, blocks, defoci):
        gx = self.gx
        E = gx.d_FFTR2C(blocks)
        ctf = gx.ctf_3d([gx.CtfParams.make(pixelsize=10e-10, Cs=2.7e-3, voltage=300e3, defocus=-float(d),
                                           amplitude=0.07) for d in defoci], N, device=self.dev)
        return E, ctf

    def exhaustive(self, E, ctf):
        H = (N // 2 + 1) * N * N
        return self.gx.d_PickSubTomograms(self.plan, self.projft, OVERSAMPLE, E, ctf, self.angles, MASK_RADIUS,
                                          ctf_volume_stride=H)

    def peaks(self, best):
        """(k, 3) (x, y, z) peaks of each map above PEAK_THRESHOLD."""
        return self.gx.d_LocalPeaks(best, PEAK_EXTENT, PEAK_THRESHOLD)


def tile_labels(peaks, values, dirs, n=N, margin=LABEL_MARGIN):
    """Class per 32^3 tile (tz, ty, tx order): the direction of the strongest peak within
    `margin` voxels of the tile's box, else 500 (no peak)."""
    nt = n // 32
    labels = np.full(nt ** 3, NDIRS, np.int32)
    bestv = np.full(nt ** 3, -np.inf)
    for (x, y, z), v, d in zip(peaks, values, dirs):
        lo = [max(0, (c - margin) // 32) for c in (z, y, x)]
        hi = [min(nt - 1, (c + margin) // 32) for c in (z, y, x)]
        for tz in range(lo[0], hi[0] + 1):
            for ty in range(lo[1], hi[1] + 1):
                for tx in range(lo[2], hi[2] + 1):
                    t = (tz * nt + ty) * nt + tx
                    if v > bestv[t]:
                        bestv[t] = v
                        labels[t] = d
    return labels


def label(args):
    import torch
    w = Workload()
    out = {"seeds": [], "labels": [], "peaks": [], "values": [], "index": []}
    t0 = time.time()
    for k in range(args.count):
        seed = args.first + k
        blk, pos, ori, df = w.block(seed)
        E, ctf = w.spectra(blk[None], [df])
        best, idx = w.exhaustive(E, ctf)
        pk = w.peaks(best[0])[0]
        b = best[0]
        iv = idx[0]
        sel = torch.from_numpy(pk).to(w.dev)
        vals = b[sel[:, 2], sel[:, 1], sel[:, 0]].cpu().numpy() if len(pk) else np.zeros(0, np.float32)
        inds = iv[sel[:, 2], sel[:, 1], sel[:, 0]].cpu().numpy() if len(pk) else np.zeros(0, np.int32)
        out["seeds"].append(seed)
        out["labels"].append(tile_labels(pk, vals, inds // NPSI))
        out["peaks"].append(pk.astype(np.int16))
        out["values"].append(vals.astype(np.float32))
        out["index"].append(inds.astype(np.int32))
        if (k + 1) % 25 == 0 or k + 1 == args.count:
            log(f"labelled {k + 1}/{args.count} blocks ({time.time() - t0:.0f} s); last: {len(pk)} peaks")
    npk = np.array([len(p) for p in out["peaks"]])
    np.savez_compressed(args.out, seeds=np.array(out["seeds"]), labels=np.stack(out["labels"]),
                        npeaks=npk, peaks=np.concatenate(out["peaks"]) if npk.sum() else np.zeros((0, 3), np.int16),
                        values=np.concatenate(out["values"]), index=np.concatenate(out["index"]))
    log(f"wrote {args.out}: {len(out['seeds'])} blocks, {int(npk.sum())} peaks")


def _standardised_tiles(blk):
    """(512, 32, 32, 32) tiles of a 256^3 block, standardised like gtomx_pnet_load_volume."""
    import torch
    m = blk.double().mean()
    sd = blk.double().std(unbiased=False)
    b = ((blk - m.float()) / sd.float())
    nt = N // 32
    return b.reshape(nt, 32, nt, 32, nt, 32).permute(0, 2, 4, 1, 3, 5).reshape(nt ** 3, 32, 32, 32)


def train(args):
    import torch
    from . import proposal as pr
    files = [np.load(f) for f in args.labels]
    seeds = np.concatenate([f["seeds"] for f in files])
    labels = np.concatenate([f["labels"] for f in files])
    w = Workload()
    rng = np.random.default_rng(args.seed)
    nval = max(1, int(len(seeds) * 0.05))
    order = rng.permutation(len(seeds))
    val_ids, train_ids = order[:nval], order[nval:]
    # dataset: every tile with a direction label, and as many "no peak" tiles per block (at least 16)
    xs, ys = [], []
    vx, vy = [], []
    t0 = time.time()
    for k, i in enumerate(order):
        blk, _, _, _ = w.block(int(seeds[i]))
        tiles = _standardised_tiles(blk)
        lab = labels[i]
        pos = np.flatnonzero(lab < pr.NO_PEAK)
        neg = np.flatnonzero(lab == pr.NO_PEAK)
        take = np.concatenate([pos, rng.choice(neg, size=min(len(neg), max(16, len(pos))), replace=False)])
        sel = torch.from_numpy(take).to(w.dev)
        if i in set(val_ids.tolist()):
            vx.append(tiles.clone())  # validation keeps every tile of its blocks
            vy.append(torch.from_numpy(lab).to(w.dev))
        else:
            xs.append(tiles[sel].clone())
            ys.append(torch.from_numpy(lab[take]).to(w.dev))
        if (k + 1) % 200 == 0:
            log(f"dataset: {k + 1}/{len(order)} blocks ({time.time() - t0:.0f} s)")
    X = torch.cat(xs)
    Y = torch.cat(ys).int()
    VX = torch.cat(vx)
    VY = torch.cat(vy).int()
    del xs, ys, vx, vy
    log(f"dataset: {len(X)} training tiles ({int((Y < pr.NO_PEAK).sum())} with a direction), "
        f"{len(VX)} validation tiles")
    bs = args.batch
    net = pr.ProposalNet(max_tiles=max(bs, 512), seed=args.seed)
    steps_per_epoch = len(X) // bs
    total = steps_per_epoch * args.epochs
    step = 0
    history = []
    for ep in range(args.epochs):
        perm = torch.randperm(len(X), device=w.dev)
        tl = 0.0
        for b in range(steps_per_epoch):
            ids = perm[b * bs:(b + 1) * bs]
            lr = args.lr * 0.5 * (1 + np.cos(np.pi * step / total))
            net.load_tiles(X[ids].contiguous())
            loss = net.train_step(Y[ids].contiguous(), float(lr), args.wd)
            tl += float(loss.mean()) if b % 50 == 0 else 0.0
            step += 1
        metrics = evaluate_tiles(net, VX, VY)
        metrics.update(epoch=ep + 1, seconds=round(time.time() - t0, 1), train_loss=round(tl / max(1, b // 50 + 1), 4))
        history.append(metrics)
        log(json.dumps(metrics))
        # written after every epoch, so a run cut short still leaves its last network
        params = net.get_params()
        np.savez_compressed(args.out, params=params, history=json.dumps(history),
                            config=json.dumps(dict(snr=SNR, peak_threshold=PEAK_THRESHOLD, peak_extent=PEAK_EXTENT,
                                                   label_margin=LABEL_MARGIN, blocks=len(seeds), epochs=ep + 1,
                                                   batch=bs, lr=args.lr, wd=args.wd)))
        log(f"wrote {args.out} (epoch {ep + 1})")
        if args.max_minutes > 0 and time.time() - t0 > 60 * args.max_minutes:
            log(f"time budget of {args.max_minutes} min reached after epoch {ep + 1}")
            break


def evaluate_tiles(net, VX, VY, topk=16, conf=0.5):
    """Validation: loss, direction top-k accuracy on tiles with a direction, and the decisions."""
    import torch
    from . import proposal as pr
    n = len(VX)
    hits = tot = 0
    dec = {"skip": 0, "short": 0, "exhaustive": 0}
    neg_skip = neg = 0
    pos_short_hit = pos = 0
    for b0 in range(0, n, 512):
        x = VX[b0:b0 + 512].contiguous()
        y = VY[b0:b0 + 512].cpu().numpy()
        k = net.load_tiles(x)
        net.forward(k)
        dirs, _, _, mass, pnp = net.propose(k, topk, full=True)
        dirs, mass, pnp = dirs.cpu().numpy(), mass.cpu().numpy(), pnp.cpu().numpy()
        for i in range(k):
            # the decision of gtomx_pick_subtomograms_learned: "no peak" or the top-k directions,
            # whichever is more probable, exhaustive when that probability is below conf
            none = pnp[i] >= mass[i]
            if max(pnp[i], mass[i]) < conf:
                d = "exhaustive"
            elif none:
                d = "skip"
            else:
                d = "short"
            dec[d] += 1
            if y[i] == pr.NO_PEAK:
                neg += 1
                neg_skip += d == "skip"
            else:
                pos += 1
                ok = y[i] in dirs[i]
                hits += ok
                pos_short_hit += (d == "exhaustive") or (d == "short" and ok)
    return {"val_tiles": n, "topk_acc": round(hits / max(pos, 1), 4),
            "pos_covered": round(pos_short_hit / max(pos, 1), 4), "neg_skipped": round(neg_skip / max(neg, 1), 4),
            "decisions": dec}


def recall(w, net, seeds, topk=16, conf=0.5):
    """Exhaustive vs learned search on each seed's block: peak recall, timings, decisions."""
    import torch
    gx = w.gx
    H = (N // 2 + 1) * N * N
    res = []
    for seed in seeds:
        blk, pos, ori, df = w.block(seed)
        E, ctf = w.spectra(blk[None], [df])
        torch.cuda.synchronize()
        t0 = time.time()
        bx, ix = w.exhaustive(E, ctf)
        torch.cuda.synchronize()
        t1 = time.time()
        bl, il, st = gx.d_PickSubTomogramsLearned(w.plan, net, w.projft, OVERSAMPLE, E, ctf, w.angles, MASK_RADIUS,
                                                  npsi=NPSI, topk=topk, confidence=conf, ctf_volume_stride=H)
        torch.cuda.synchronize()
        t2 = time.time()
        pk = w.peaks(bx[0])[0]
        pl = {tuple(p) for p in w.peaks(bl[0])[0]}
        bxn, bln, ixn, iln = bx[0].cpu().numpy(), bl[0].cpu().numpy(), ix[0].cpu().numpy(), il[0].cpu().numpy()
        same = [bool(bln[z, y, x] == bxn[z, y, x] and iln[z, y, x] == ixn[z, y, x]) for x, y, z in pk]
        found = [tuple(p) in pl for p in pk]
        res.append(dict(seed=int(seed), peaks=len(pk), same=int(sum(same)), found=int(sum(s and f for s, f in
                                                                                        zip(same, found))),
                        exhaustive_s=round(t1 - t0, 4), learned_s=round(t2 - t1, 4), stats=st))
    return res


def block_scores(w, net, seed, topk=16):
    """One block: the exhaustive search's peaks [(x, y, z, value, direction)] and, per 32^3 tile
    (tz, ty, tx order), the network's P(no peak) and top-k directions."""
    blk, _, _, df = w.block(seed)
    E, ctf = w.spectra(blk[None], [df])
    bx, ix = w.exhaustive(E, ctf)
    pk = w.peaks(bx[0])[0]
    b, iv = bx[0].cpu().numpy(), ix[0].cpu().numpy()
    k = net.load_tiles(_standardised_tiles(blk).contiguous())
    net.forward(k)
    dirs, _, _, _, pnp = net.propose(k, topk, full=True)
    peaks = [(int(x), int(y), int(z), float(b[z, y, x]), int(iv[z, y, x]) // NPSI) for x, y, z in pk]
    return peaks, pnp.cpu().numpy(), dirs.cpu().numpy()


def tile_of(x, y, z, n=N):
    nt = n // 32
    return ((z // 32) * nt + y // 32) * nt + x // 32


def detection_auc(blocks):
    """Area under the ROC of 1 - P(no peak) separating the tiles that hold an exhaustive peak from
    the tiles with no peak within LABEL_MARGIN voxels (0.5 for a scorer without information)."""
    pos, neg = [], []
    for peaks, pnp, _ in blocks:
        hold = np.zeros(len(pnp), bool)
        near = np.zeros(len(pnp), bool)
        for x, y, z, _, _ in peaks:
            hold[tile_of(x, y, z)] = True
            for dz in (-LABEL_MARGIN, 0, LABEL_MARGIN):
                for dy in (-LABEL_MARGIN, 0, LABEL_MARGIN):
                    for dx in (-LABEL_MARGIN, 0, LABEL_MARGIN):
                        xx, yy, zz = x + dx, y + dy, z + dz
                        if 0 <= xx < N and 0 <= yy < N and 0 <= zz < N:
                            near[tile_of(xx, yy, zz)] = True
        pos.append(1 - pnp[hold])
        neg.append(1 - pnp[~near])
    pos, neg = np.concatenate(pos), np.concatenate(neg)
    return float((pos[:, None] > neg[None, :]).mean() + 0.5 * (pos[:, None] == neg[None, :]).mean())


def scores(args):
    """Per block: the exhaustive search's peaks and the network's per-tile P(no peak) and top-k
    directions -- what the operating points are chosen from (every decision of
    gtomx_pick_subtomograms_learned is a function of these)."""
    from . import proposal as pr
    w = Workload()
    net = pr.ProposalNet(max_tiles=512, params=np.load(args.weights)["params"])
    out = []
    for seed in range(args.first, args.first + args.count):
        peaks, pnp, dirs = block_scores(w, net, seed, args.topk)
        out.append({"seed": seed, "peaks": [list(p) for p in peaks], "p_no_peak": [round(float(v), 5) for v in pnp],
                    "top_dirs": dirs.tolist()})
        log(f"scored block {seed}: {len(peaks)} peaks")
    net.close()
    print(json.dumps(out), flush=True)


def evaluate(args):
    from . import proposal as pr
    w = Workload()
    params = np.load(args.weights)["params"]
    net = pr.ProposalNet(max_tiles=512, params=params)
    seeds = list(range(args.first, args.first + args.count))
    res = recall(w, net, seeds, args.topk, args.confidence)
    tot = sum(r["peaks"] for r in res)
    summary = {"blocks": len(res), "peaks": tot, "recall": sum(r["found"] for r in res) / max(tot, 1),
               "value_recall": sum(r["same"] for r in res) / max(tot, 1),
               "exhaustive_s": sum(r["exhaustive_s"] for r in res), "learned_s": sum(r["learned_s"] for r in res)}
    summary["speedup"] = summary["exhaustive_s"] / max(summary["learned_s"], 1e-9)
    print(json.dumps({"summary": summary, "blocks": res}), flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("label")
    a.add_argument("--first", type=int, default=TRAIN_SEEDS[0])
    a.add_argument("--count", type=int, default=TRAIN_SEEDS[1])
    a.add_argument("--out", required=True)
    t = sub.add_parser("train")
    t.add_argument("--labels", nargs="+", required=True)
    t.add_argument("--out", default=WEIGHTS)
    t.add_argument("--epochs", type=int, default=12)
    t.add_argument("--batch", type=int, default=256)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--wd", type=float, default=1e-4)
    t.add_argument("--seed", type=int, default=0)
    t.add_argument("--max-minutes", type=float, default=0.0, help="stop after the epoch that passes this budget")
    e = sub.add_parser("eval")
    e.add_argument("--weights", default=WEIGHTS)
    e.add_argument("--first", type=int, default=HELDOUT_SEEDS[0])
    e.add_argument("--count", type=int, default=HELDOUT_SEEDS[1])
    e.add_argument("--topk", type=int, default=16)
    e.add_argument("--confidence", type=float, default=0.5)
    c = sub.add_parser("scores")
    c.add_argument("--weights", default=WEIGHTS)
    c.add_argument("--first", type=int, default=HELDOUT_SEEDS[0])
    c.add_argument("--count", type=int, default=HELDOUT_SEEDS[1])
    c.add_argument("--topk", type=int, default=16)
    args = ap.parse_args(argv)
    return {"label": label, "train": train, "eval": evaluate, "scores": scores}[args.cmd](args)


if __name__ == "__main__":
    sys.exit(main())
