"""Learned warm start: training driver and evaluation against random starts.

The network (cdx_unet_*, BF16 MFMA kernels in csrc/unet.hip) is trained on
the GPU on pairs from the forward model: random ellipse / polygon supports of
radius 20-60 px, smoothed-noise amplitude in [0.2, 1] and phase in [-pi, pi],
photons log-uniform in a range (PHOTONS, 1e5-1e8, by default; the shipped
weights were trained on 1e5-1e10), Poisson intensities.  Training seeds start
at 1,000,000, held-out seeds at 2,000,000.

Evaluation runs the same HIO schedule from two starts on held-out frames,
the frames advancing together in batched reconstructions:
  random  - uniform random phase on sqrt(I), the reference's rectangular
            support (createSupport, oversampling 2, src/post_gpu/cdi.cc:240-251)
  learned - cdx_unet_warm_start: predicted object and support
The schedule is blocks of 9 HIO + 1 ER (relaxation 1, i.e. the support
projection), so the R-factor recorded after each ER step is the R-factor of
the support-constrained estimate; shrink-wrap every other block.
Random-start HIO stays the fallback and the judge of the learned start
(reconstruct_guarded; the "guarded" summary of evaluate): both starts run
and the lower final R is kept, the pattern of util/cdi.cpp:36-45.

Command line (on the GPU):
  python -m cdixx_amd.warmstart train --steps 30000 --photons 1e5 1e10 --out cdixx_amd/models/unet256.bin
  python -m cdixx_amd.warmstart eval --weights cdixx_amd/models/unet256.bin --frames 1024 --blocks 100
"""
import argparse
import json
import os
import sys
import time

import numpy as np

TRAIN_SEED0 = 1_000_000
HELDOUT_SEED0 = 2_000_000
PHOTONS = (1e5, 1e8)
SIZE, LEVELS, BASE = 256, 4, 32
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "models", "unet256.bin")


def train(net, steps, batch=16, lr=1e-3, seed0=TRAIN_SEED0, chunk=50, photons=PHOTONS, log=print):
    """steps of Adam on fresh forward-model batches; cosine learning-rate
    decay to lr/20.  Returns the per-step losses."""
    losses = []
    t0 = time.time()
    done = 0
    while done < steps:
        k = min(chunk, steps - done)
        lr_now = lr * (0.05 + 0.95 * 0.5 * (1 + np.cos(np.pi * done / max(steps, 1))))
        losses.extend(net.train(k, batch, seed0 + done * batch, lr_now, photons))
        done += k
        log(f"step {done}/{steps} lr {lr_now:.2e} loss {np.mean(losses[-k:]):.4f} ({time.time() - t0:.1f} s)")
    return np.array(losses)


def _iters_to(trace, target, every):
    """first recorded iteration (a multiple of `every`) whose R-factor is <= target"""
    for j in range(every, len(trace), every):
        if trace[j] <= target:
            return j
    return None


def _eval_chunk(cdx, net, nframes, seed0, photons, blocks, beta, support_threshold, shrink_threshold, arms):
    """one batched run of every arm on frames seed0 .. seed0 + nframes - 1:
    per-arm R-factor traces (every projection), the frames' noise floors and
    the predicted supports' IoU"""
    n = net.size
    b = cdx.Batch([n] * nframes)
    truth, r_true = [], []
    for i in range(nframes):
        f = b.frame(i)
        ctx = b.context(i)
        net.synth_frame(seed0 + i, f["g"], f["I"], photons)
        f["bs"].zero()
        ctx.fft(f["g"], f["F"], cdx.FORWARD)
        # the true object's transform scaled to the measured total (Parseval):
        # its R-factor is the frame's noise floor
        Ft, Ih = b.get(i, "F"), b.get(i, "I")
        scale = np.sqrt(Ih.sum() / max((np.abs(Ft.astype(np.complex128)) ** 2).sum(), 1e-300))
        a = np.sqrt(np.maximum(Ih, 0))
        r_true.append(float(np.abs(np.abs(Ft) * scale - a).sum() / a.sum()))
        truth.append(b.get(i, "g"))
    b.sync()
    masks = [np.abs(t) > 0 for t in truth]
    alg = f"{blocks // 2}*(9*HIO+1*ER+9*HIO+1*ER+1*shrinkWrap)"
    nproj = (blocks // 2) * 20
    p = cdx.Params.default(beta=beta, relax=1.0, recordR=1, shrinkThreshold=shrink_threshold)
    traces, seconds, iou = {}, {}, []
    for arm in arms:
        for i in range(nframes):
            f = b.frame(i)
            ctx = b.context(i)
            if arm == "learned":
                net.warm_start(ctx, f["I"], f["F"], f["g"], f["spt"], threshold=support_threshold)
                pred = b.get(i, "spt") > 0.5
                iou.append((pred & masks[i]).sum() / max((pred | masks[i]).sum(), 1))
                continue
            if arm == "learned_support":
                net.warm_start(ctx, f["I"], f["F"], f["g"], f["spt"], threshold=support_threshold)
            ctx.random_start(f["I"], None, f["F"], seed=seed0 + i, mode=1)
            ctx.fft(f["F"], f["g"], cdx.INVERSE)
            ctx.apply_norm(f["g"], 1.0 / n)
            if arm == "random":
                ctx.create_support(f["spt"], 2.0)
            elif arm == "true_support":
                b.put(i, "spt", masks[i].astype(np.float32))
        b.sync()
        b.set_threshold(0.5)
        t0 = time.perf_counter()
        _, tr = b.phase_retrieve(alg, p, rtrace_cap=nproj)
        b.sync()
        seconds[arm] = time.perf_counter() - t0
        traces[arm] = tr
    b.close()
    return traces, seconds, r_true, iou, nproj


def evaluate(cdx, net, nframes=64, seed0=HELDOUT_SEED0, photons=PHOTONS, blocks=60, beta=0.9, target=0.05,
             support_threshold=0.5, shrink_threshold=0.02, arms=("random", "learned"), chunk=256):
    """Iterations to R <= target (and to R <= max(target, 1.1 x the true
    object's R, the frame's noise floor)) and final R per start on `nframes`
    held-out frames (batched runs of at most `chunk` frames).  Arms: "random"
    (random phase, rectangular support), "learned" (warm start),
    "learned_support" (random phase, predicted support), "true_support"
    (random phase, the true support: a bound on what a support estimate
    alone can give).  Also reported: the R-factor of the true object (the
    noise floor of each frame), the overlap of the predicted support with the
    true one, and per frame the iterations each arm needs to reach the random
    start's final R.  Returns (per-frame dict, summary)."""
    traces = {arm: [] for arm in arms}
    seconds = {arm: 0.0 for arm in arms}
    r_true, iou, nproj = [], [], 0
    for k in range(0, nframes, chunk):
        tr, sec, rt, io, nproj = _eval_chunk(cdx, net, min(chunk, nframes - k), seed0 + k, photons, blocks, beta,
                                             support_threshold, shrink_threshold, arms)
        for arm in arms:
            traces[arm].append(tr[arm])
            seconds[arm] += sec[arm]
        r_true += rt
        iou += io
    out = {}
    floor = [max(target, 1.1 * r) for r in r_true]
    for arm in arms:
        tr = np.concatenate(traces[arm], axis=0)
        out[arm] = dict(trace=tr[:, ::10], iters=[_iters_to(tr[i], target, 10) for i in range(nframes)],
                        iters_floor=[_iters_to(tr[i], floor[i], 10) for i in range(nframes)],
                        final_R=np.array([tr[i][nproj - 10] for i in range(nframes)]), seconds=seconds[arm])
    summary = {"frames": nframes, "iterations": nproj, "target_R": target, "photons": list(photons),
               "median_R_true_object": float(np.median(r_true)),
               "frames_with_R_true_below_target": int(np.sum(np.array(r_true) <= target))}
    if iou:
        summary["median_support_iou"] = float(np.median(iou))
    if "random" in out and "learned" in out:
        # the learned start judged by the random start (util/cdi.cpp:36-45:
        # keep the run with the lower residual): per frame the lower final R
        rr, rl = out["random"]["final_R"], out["learned"]["final_R"]
        guarded = np.minimum(rr, rl)
        summary["guarded_median_final_R"] = float(np.median(guarded))
        summary["guarded_frames_keeping_learned"] = int(np.sum(rl <= rr))
        summary["guarded_frames_above_random"] = int(np.sum(guarded > rr))
    ref_final = out["random"]["final_R"] if "random" in out else None
    for arm, o in out.items():
        reached = [x for x in o["iters"] if x is not None]
        summary[f"{arm}_reached_target"] = len(reached)
        summary[f"{arm}_median_iters_to_target"] = float(np.median(reached)) if reached else None
        fl = [x for x in o["iters_floor"] if x is not None]
        summary[f"{arm}_reached_floor"] = len(fl)
        summary[f"{arm}_median_iters_to_floor"] = float(np.median(fl)) if fl else None
        summary[f"{arm}_median_final_R"] = float(np.median(o["final_R"]))
        summary[f"{arm}_median_R_at_100"] = float(np.median(o["trace"][:, 10]))
        if ref_final is not None:
            # probe index (x10 iterations) where the arm first reaches the random start's final R
            its = []
            for i in range(nframes):
                hit = np.nonzero(o["trace"][i, 1:] <= ref_final[i])[0]
                its.append(10 * (hit[0] + 1) if len(hit) else None)
            got = [x for x in its if x is not None]
            summary[f"{arm}_frames_reaching_random_final_R"] = len(got)
            summary[f"{arm}_median_iters_to_random_final_R"] = float(np.median(got)) if got else None
            o["iters_to_random_final"] = its
        summary[f"{arm}_seconds"] = round(o["seconds"], 3)
    return out, summary


def reconstruct_guarded(cdx, net, ctx_small, ctx, I, bs, algorithm, params, seed, support_radius,
                        threshold=0.5, dilate=0):
    """The learned start judged by random-start phase retrieval, the pattern
    of util/cdi.cpp:36-45 (rerun from a new start, keep the result with the
    lowest residual): the schedule runs from the network's scaled warm start
    (cdx_unet_warm_start_scaled, phases only, disk support of
    `support_radius`) and from a random-phase start (same support, Philox
    seed `seed`), and the run whose final F = FFT(g) has the lower R-factor
    against the measurement is kept.  The guarded result's R therefore never
    exceeds the random start's.  The schedule should end on ER steps, so the
    final object is support-projected.  Returns a dict with both runs' final
    R, rtraces, seconds, which start was kept ("learned"/"random") and its
    device buffers F, g, spt (the other run's are freed)."""
    n = ctx.rows
    runs = {}
    for arm in ("learned", "random"):
        F, g, spt = ctx.cplx(), ctx.cplx(), ctx.real()
        if arm == "learned":
            net.warm_start_scaled(ctx_small, ctx, I, bs, F, g, spt, threshold=threshold, dilate=dilate, seed=seed)
        else:
            ctx.random_start(I, bs, F, seed=seed, mode=1)
            ctx.fft(F, g, cdx.INVERSE)
            ctx.apply_norm(g, 1.0 / n)
        ctx.create_circle(spt, support_radius)
        ctx.set_threshold(0.5)
        ctx.sync()
        t0 = time.perf_counter()
        nit, rt = ctx.phase_retrieve(algorithm, params, F, g, I, spt, bs=bs, rtrace_cap=nit_cap(cdx, algorithm))
        ctx.sync()
        sec = time.perf_counter() - t0
        runs[arm] = dict(F=F, g=g, spt=spt, R=ctx.rfactor(F, I, bs), rtrace=rt, seconds=sec)
    keep = "learned" if runs["learned"]["R"] <= runs["random"]["R"] else "random"
    drop = "random" if keep == "learned" else "learned"
    for k in ("F", "g", "spt"):
        runs[drop][k].free()
    return dict(kept=keep, R_learned=runs["learned"]["R"], R_random=runs["random"]["R"],
                R=runs[keep]["R"], rtrace_learned=runs["learned"]["rtrace"], rtrace_random=runs["random"]["rtrace"],
                seconds=runs["learned"]["seconds"] + runs["random"]["seconds"],
                F=runs[keep]["F"], g=runs[keep]["g"], spt=runs[keep]["spt"])


def nit_cap(cdx, algorithm):
    """projection steps of a schedule (its R-factor trace length)"""
    return sum(1 for t in cdx.expand_schedule(algorithm) if t != cdx.SHRINKWRAP)


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--steps", type=int, default=3000)
    t.add_argument("--batch", type=int, default=16)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--out", default=DEFAULT_WEIGHTS)
    t.add_argument("--init", default=None, help="continue from these weights")
    t.add_argument("--photons", type=float, nargs=2, default=list(PHOTONS))
    e = sub.add_parser("eval")
    e.add_argument("--weights", default=DEFAULT_WEIGHTS)
    e.add_argument("--frames", type=int, default=64)
    e.add_argument("--blocks", type=int, default=60)
    e.add_argument("--photons", type=float, nargs=2, default=list(PHOTONS))
    e.add_argument("--target", type=float, default=0.05)
    e.add_argument("--support-threshold", type=float, default=0.5)
    e.add_argument("--arms", default="random,learned")
    args = ap.parse_args(argv)
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import cdixx_amd as cdx
    if args.cmd == "train":
        net = cdx.UNet(SIZE, LEVELS, BASE, max_batch=args.batch, seed=1)
        if args.init:
            net.load(args.init)
        losses = train(net, args.steps, args.batch, args.lr, photons=tuple(args.photons),
                       log=lambda s: print(s, flush=True))
        os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
        net.save(args.out)
        print(json.dumps({"steps": args.steps, "batch": args.batch, "first_loss": float(np.mean(losses[:20])),
                          "last_loss": float(np.mean(losses[-20:])), "weights": args.out}), flush=True)
    else:
        net = cdx.UNet(SIZE, LEVELS, BASE, max_batch=1, seed=1)
        if args.weights and os.path.exists(args.weights):
            net.load(args.weights)
        _, summary = evaluate(cdx, net, args.frames, photons=tuple(args.photons), blocks=args.blocks,
                              target=args.target, support_threshold=args.support_threshold,
                              arms=tuple(args.arms.split(",")))
        print(json.dumps(summary), flush=True)


if __name__ == "__main__":
    main()
