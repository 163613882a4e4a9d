"""Trains the U-Net denoiser on pairs rendered by this path tracer.

Each training pair is a 256 x 256 crop (a camera aimed at a random window of
a 1920 x 1080 view, so objects keep their full-resolution pixel footprint):
the input is the accumulation after a few samples per pixel (4 by default, or
drawn per pair from a list such as 4,8,16,32,64) with its albedo / normal
AOVs, the target the same crop at 4096 spp (log(1 + RGB)).  Scenes come from this package's
generators with seeds other than the benchmark scenes': env-lit fields of
displaced spheres (the 1080p workload's generator: sphere layout, metal and
glass materials, sky noise and sun position redrawn) and terrain rooms with
64 random Disney materials and 4 quad lights.  Everything stays on the GPU:
captures go straight from the accumulation buffers into the device-side
dataset, and AdamW steps run on device.

    python -m bluetruthengine_amd.train_denoiser --pairs 256 --steps 4000

The shipped weights (bluetruthengine_amd/weights/unet_c4.npy) come from
    python -m bluetruthengine_amd.train_denoiser --pairs 512 --steps 12000 \
        --input-spp 4,8,16,32,64 --envlit-frac 0.85 --seed 3
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import denoiser, renderer, scenes

CROP = 256
FULL_W, FULL_H = 1920, 1080


def log(msg):
    print(msg, flush=True)


def make_scene(kind, seed, rng):
    cx = int(rng.integers(CROP // 2, FULL_W - CROP // 2))
    cy = int(rng.integers(CROP // 2, FULL_H - CROP // 2))
    if kind == "envlit":
        eye = scenes.C3_EYE
        cam = scenes.crop_camera(eye, scenes.C3_TARGET, (0, 1, 0), scenes.C3_FOV, FULL_W, FULL_H, cx, cy, CROP)
        n = int(rng.integers(150, 500))
        return scenes.envlit_c3(CROP, CROP, spp=1, max_depth=12, n_spheres=n, slices=24, bands=12,
                                env_size=(1024, 512), seed=seed, camera=cam)
    cam = scenes.crop_camera((0, 1, 3.6), (0, 1, 0), (0, 1, 0), 40.0, FULL_W, FULL_H, cx, cy, CROP)
    return scenes.terrain_c2(CROP, CROP, spp=1, max_depth=10, freq=40, seed=seed, camera=cam)


def render_pairs(pairs, input_spp, target_spp, envlit_frac, seed):
    """input_spp: a list; each pair's input takes one of its values at random."""
    rng = np.random.default_rng(seed)
    denoiser.dataset_reserve(pairs, CROP, CROP)
    t0 = time.perf_counter()
    for i in range(pairs):
        kind = "envlit" if rng.uniform() < envlit_frac else "terrain"
        sd = make_scene(kind, 10_000 + i, rng)
        spp = int(rng.choice(input_spp)) if len(input_spp) > 1 else input_spp[0]
        r = renderer.Renderer(sd)
        r.enable_aov(True)
        r.build_accel()
        r.render(spp)
        denoiser.capture(i, 0)
        left = target_spp - spp
        while left > 0:
            n = min(left, 1024)
            r.render(n)
            left -= n
        denoiser.capture(i, 1)
        if i % 8 == 7 or i == pairs - 1:
            log(f"pair {i + 1}/{pairs} ({kind}, {sd.triangles.shape[0]} tris) {time.perf_counter() - t0:.0f} s")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=256)
    ap.add_argument("--input-spp", default="4",
                    help="input samples per pixel, or a comma-separated list to draw each pair's from")
    ap.add_argument("--target-spp", type=int, default=4096)
    ap.add_argument("--envlit-frac", type=float, default=0.75)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--weight-decay", type=float, default=1e-4)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--loss", choices=["l1", "relmse"], default="l1",
                    help="l1: L1 of log(1 + RGB) (default); relmse: relative MSE of the linear radiance")
    ap.add_argument("--init", default="", help="start from these weights (.npy) instead of a random init")
    ap.add_argument("--finetune-steps", type=int, default=0,
                    help="after the main run, this many more steps with the relMSE loss at lr/3 (saved to --finetune-out)")
    ap.add_argument("--finetune-out", default="")
    ap.add_argument("--swa-from", type=float, default=0.0,
                    help="average the weights of every 250-step snapshot after this fraction of the steps (0 = off)")
    ap.add_argument("--out", default=denoiser.WEIGHTS_PATH)
    ap.add_argument("--log-json", default="")
    args = ap.parse_args(argv)

    renderer.lib()
    denoiser.init(args.seed)
    t0 = time.perf_counter()
    input_spp = [int(x) for x in str(args.input_spp).split(",") if x]
    render_pairs(args.pairs, input_spp, args.target_spp, args.envlit_frac, args.seed)
    t_render = time.perf_counter() - t0
    log(f"rendered {args.pairs} pairs in {t_render:.0f} s")
    if args.init:
        denoiser.load(args.init)          # after the renders: Renderer() keeps the weights
        log(f"starting from {args.init}")
    denoiser.set_loss(denoiser.LOSS_RELMSE if args.loss == "relmse" else denoiser.LOSS_L1_LOG)
    # AdamW at lr, then two tenfold decays over the last 30 % of the steps
    phases = [(0.7, args.lr), (0.2, args.lr * 0.3), (0.1, args.lr * 0.1)]
    history = []
    t1 = time.perf_counter()
    done = 0
    swa_sum, swa_n = None, 0
    for frac, lr in phases:
        steps = max(1, int(round(args.steps * frac)))
        chunk = 250
        while steps > 0:
            n = min(chunk, steps)
            losses = denoiser.train(n, args.batch, lr=lr, weight_decay=args.weight_decay, seed=args.seed * 7919 + done)
            history.extend(float(x) for x in losses)
            done += n
            steps -= n
            log(f"step {done}: lr {lr:.1e}, mean {args.loss} loss of the last {n}: {losses.mean():.5f}, "
                f"{time.perf_counter() - t1:.0f} s")
            if args.swa_from > 0 and done >= args.swa_from * args.steps:
                w = denoiser.get_weights().astype(np.float64)
                swa_sum = w if swa_sum is None else swa_sum + w
                swa_n += 1
    t_train = time.perf_counter() - t1
    if swa_n:
        # weight averaging over the snapshots of the last part of the run
        denoiser.save(args.out.replace(".npy", "_last.npy"))
        denoiser.set_weights((swa_sum / swa_n).astype(np.float32))
        log(f"averaged {swa_n} weight snapshots from step {int(args.swa_from * args.steps)} on")
    denoiser.save(args.out)
    log(f"saved {args.out} ({denoiser.param_count()} parameters); training {t_train:.0f} s")
    ft_history = []
    if args.finetune_steps > 0:
        denoiser.set_loss(denoiser.LOSS_RELMSE)
        left = args.finetune_steps
        while left > 0:
            n = min(250, left)
            losses = denoiser.train(n, args.batch, lr=args.lr / 3, weight_decay=args.weight_decay,
                                    seed=args.seed * 104729 + left)
            ft_history.extend(float(x) for x in losses)
            left -= n
            log(f"fine-tune (relMSE) {args.finetune_steps - left}/{args.finetune_steps}: mean loss of the last {n}: "
                f"{losses.mean():.5f}")
        denoiser.save(args.finetune_out or args.out.replace(".npy", "_relmse.npy"))
        denoiser.set_loss(denoiser.LOSS_L1_LOG)
    if args.log_json:
        with open(args.log_json, "w") as fh:
            json.dump({"pairs": args.pairs, "input_spp": input_spp, "target_spp": args.target_spp,
                       "steps": done, "batch": args.batch, "lr": args.lr, "weight_decay": args.weight_decay,
                       "render_s": t_render, "train_s": t_train,
                       "loss": args.loss, "init": args.init, "swa_snapshots": swa_n,
                       "loss_first_100": float(np.mean(history[:100])),
                       "loss_last_100": float(np.mean(history[-100:])),
                       "finetune_relmse_steps": args.finetune_steps,
                       "finetune_loss_last_100": float(np.mean(ft_history[-100:])) if ft_history else None}, fh)
    return 0


if __name__ == "__main__":
    sys.exit(main())
