#!/usr/bin/env python3
"""Benchmark: fwd+bwd iterations/s of the tile rasterizer on one MI355X.

Workload (``scale``): 1024x1024 RGB canvas on a white background, 8192
primitives over four 128x128 fp32 alpha bitmaps (soft disk, five-point star,
ring, blurred thresholded noise), 16x16 tiles, target a smooth random field.
One step = GPU binning + forward + background composite + MSE + backward +
Adam (lr 1e-2) + parameter clamp -- nothing in the optimisation step is
skipped.  ``value`` is steps/s over the K timed steps (wall clock between two
device synchronisations); per-step hipEvent times of the fwd+bwd segment
(binning, forward, loss, backward) are reported as p50/p99 beside it.
Before any timed run a throwaway copy of the workload is fitted, untimed, for
0.5 s so that short timed regions do not start on idle GPU clocks; the timed
jobs start from the original parameters.

``roofline``: the dominant kernel (backward raster) -- its compulsory HBM
bytes per launch divided by its mean duration, measured live with hipEvents
on the launch stream, against 8 TB/s.  ``issue_roofline`` and
``lds_roofline``: the same kernel's wave64 VALU instructions and LDS bytes per
launch (from the committed PMC passes under ``profiles/``) over the same live
duration, against 614 G instructions/s and 78.6 TB/s.  ``cpu_baseline``: the reference
algorithm in fp32 on the host cores (OpenMP build of the oracle): full
optimisation steps of the same workload (median of 3), and the plain
workload's 100-step fit (``plain_fit``, also run on the GPU).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK = 8.0e12
# wave64 VALU instructions per second: 256 CUs x 2.4 GHz x one instruction per
# CU per cycle (4 SIMDs of 16 lanes, 4 cycles per wave64 instruction, no
# packed fp32)
VALU_ISSUE_PEAK = 256 * 2.4e9
# LDS bytes per second: 128 B per CU per cycle (32 banks x 4 B), 256 CUs, 2.4 GHz
LDS_PEAK = 128 * 256 * 2.4e9


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--precision", default=None, choices=["fp32", "fp16", "bf16"],
                    help="parameter / bitmap storage (arithmetic is fp32); default fp32, fp16 for "
                         "--workload stress")
    ap.add_argument("--eps-t", type=float, default=None)
    ap.add_argument("--mode", default="fused", choices=["fused", "autograd"],
                    help="fused: the device-resident FitJob step (binning, forward, loss, backward, "
                         "Adam as HIP kernels); autograd: the reference-compatible TileRasterizer "
                         "behind torch autograd and torch.optim.Adam")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=3,
                    help="full host steps of the scale workload timed for the CPU baseline")
    ap.add_argument("--psnr-steps", type=int, default=0,
                    help="also fit this many Adam steps at fp32 and at --precision from the same "
                         "start and report both final PSNRs")
    ap.add_argument("--workload", default="scale",
                    choices=["scale", "stress", "serving", "export", "lanes"],
                    help="scale: the headline fit step; stress: the 2048^2 / 32768-primitive "
                         "RGBA colour-mapped fit step; serving: a Poisson stream of fit jobs "
                         "through the batched multi-job engine (each its own JSON line)")
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the fp32 / fp16 / mixed-storage rate and PSNR comparison")
    ap.add_argument("--no-serving", action="store_true",
                    help="leave the bounded serving run out of the headline line")
    ap.add_argument("--serve-jobs", type=int, default=48)
    ap.add_argument("--serve-rate", type=float, default=0.0,
                    help="arrival rate in jobs/s (0: every job arrives at t = 0, i.e. capacity)")
    ap.add_argument("--serve-steps", type=int, default=200)
    ap.add_argument("--serve-max-jobs", type=int, default=32)
    ap.add_argument("--serve-budget-gib", type=float, default=64.0)
    ap.add_argument("--serve-no-graph", action="store_true")
    ap.add_argument("--serve-no-share", action="store_true",
                    help="every job holds the texel records of its own bitmaps")
    ap.add_argument("--prewarm", type=float, default=0.5,
                    help="seconds of untimed fitting of a throwaway copy of the workload before the "
                         "timed runs (0: none, e.g. under a profiler, so that its per-kernel averages "
                         "cover the timed steps)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="run one device-to-device copy of this many MiB first (a known-size "
                         "reference kernel for scaling HBM byte counters)")
    args = ap.parse_args()
    if args.precision is None:
        args.precision = "fp16" if args.workload == "stress" else "fp32"
    return args


def algorithmic_bytes(H, W, N, T, Ht, Wt, P, tiles, b_p, b_t):
    """Compulsory HBM bytes per step of binning, forward and backward: every
    input read once, every output written once (primitive parameters
    8 values, template texels, the P tile-list entries, per-pixel output and
    saved state, per-primitive fp32 gradients)."""
    b_bin = N * (8 * b_p + 4) + 4 * P + 4 * (tiles + 1)
    b_fwd = N * (8 * b_p + 4) + T * Ht * Wt * b_t + 4 * P + H * W * 16 + 8 * H * W
    b_bwd = H * W * 16 + 12 * H * W + 8 * H * W + N * (8 * b_p + 4) + T * Ht * Wt * b_t + 4 * P + 32 * N
    return b_bin, b_fwd, b_bwd


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


class _CpuFit:
    """The reference's optimisation step on the host: binning, forward and
    backward of the fp32 port of the reference algorithm (per-pixel caches,
    O(K^2) suffix loop, all host cores), background composite + MSE and
    torch.optim.Adam's update with the reference's clamp in numpy."""

    def __init__(self, scene, target, lr=1e-2, betas=(0.9, 0.999), eps=1e-8):
        import oracle as O
        self.O, self.sc, self.target = O, scene, target.astype(np.float32)
        self.p = [np.array(a, np.float32) for a in (scene.means2D, scene.radii, scene.rotations,
                                                     scene.opacities, scene.colors)]
        self.m = [np.zeros_like(a) for a in self.p]
        self.v = [np.zeros_like(a) for a in self.p]
        self.lr, self.b1, self.b2, self.eps, self.t = lr, betas[0], betas[1], eps, 0
        self.ext = O.template_extents_np(scene.templates)

    def step(self) -> float:
        O, sc = self.O, self.sc
        m2, r, th, v, c = self.p
        offs, idx = O.bin_lists_c(m2, r, th, sc.bmp_sel, self.ext, sc.H, sc.W, sc.tile_size)
        osc = O.OracleScene(sc.H, sc.W, m2, r, th, v, c, sc.templates, sc.bmp_sel, offs, idx,
                            ts=sc.tile_size, alpha_ub=sc.alpha_ub)
        col, al, _ = O.forward(osc, "f32")
        res = col + (1.0 - al)[..., None] - self.target
        loss = float((res.astype(np.float64) ** 2).mean())
        gC = (2.0 / res.size) * res
        gA = -gC.sum(-1)
        g = O.backward(osc, gC, gA, "f32")
        self.t += 1
        bc1, bc2 = 1 - self.b1 ** self.t, 1 - self.b2 ** self.t
        for i, k in enumerate(("means2D", "radii", "rotations", "opacities", "colors")):
            gi = g[k].astype(np.float32)
            self.m[i] = self.b1 * self.m[i] + (1 - self.b1) * gi
            self.v[i] = self.b2 * self.v[i] + (1 - self.b2) * gi * gi
            self.p[i] = (self.p[i] - (self.lr / bc1) * self.m[i] /
                         (np.sqrt(self.v[i]) / math.sqrt(bc2) + self.eps)).astype(np.float32)
        self.p[0][:, 0] = np.clip(self.p[0][:, 0], 0, sc.W - 1)
        self.p[0][:, 1] = np.clip(self.p[0][:, 1], 0, sc.H - 1)
        self.p[1] = np.clip(self.p[1], 2, min(sc.H, sc.W) // 4).astype(np.float32)
        return loss


def cpu_baseline(scene, scale_steps=3, c0_steps=100):
    """The reference algorithm in fp32 on the host's cores: ``scale_steps``
    full optimisation steps of the scale workload (median step time is the
    value), and the plain workload's ``c0_steps``-step fit."""
    import oracle as O
    from diffbmp_amd import scenes as S
    cores = int(O.lib("f32").oracle_num_threads())
    fit = _CpuFit(scene, scene.target)
    times = []
    for _ in range(scale_steps):
        t0 = time.perf_counter()
        fit.step()
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    # plain workload: 100 Adam steps to the second random render (the target
    # rendered by the same port)
    c0 = S.make_scene(0, with_target=False)
    tp = S.make_target_params(0, c0.H, c0.W, c0.N)
    offs, idx = O.bin_lists_c(tp[0], tp[1], tp[2], c0.bmp_sel, O.template_extents_np(c0.templates),
                              c0.H, c0.W, c0.tile_size)
    col, al, _ = O.forward(O.OracleScene(c0.H, c0.W, *tp, c0.templates, c0.bmp_sel, offs, idx), "f32")
    f0 = _CpuFit(c0, col + (1.0 - al)[..., None])
    t0 = time.perf_counter()
    losses = [f0.step() for _ in range(c0_steps)]
    c0_s = time.perf_counter() - t0
    return dict(value=round(1.0 / med, 5), unit="iters/s", cores=cores, kind="port",
                cpu_model=_cpu_model(),
                sample=f"scale workload: {scale_steps} full optimisation steps (binning, forward, "
                       f"composite/MSE, backward, Adam; fp32 port of the reference algorithm), "
                       f"median {med * 1e3:.0f} ms per step",
                step_ms=[round(x * 1e3, 1) for x in times],
                plain_fit={"steps": c0_steps, "steps_per_s": round(c0_steps / c0_s, 3),
                           "loss_first": round(losses[0], 6), "loss_last": round(losses[-1], 6)})


def gpu_plain_fit(dev, steps=100):
    """The plain workload end to end on the GPU: FitJob, ``steps`` Adam steps
    (lr 1e-2) of MSE to the second random render (rendered by the drop-in
    rasterizer)."""
    import torch
    from diffbmp_amd import scenes as S
    from diffbmp_amd.fit import FitJob
    from diffbmp_amd.tile_rasterizer import TileRasterizer
    sc = S.make_scene(0, with_target=False)
    tp = S.make_target_params(0, sc.H, sc.W, sc.N)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    rast = TileRasterizer(sc.H, sc.W, 16, 0.0, 1.0, 151, sc.N)
    with torch.no_grad():
        c, a = rast(*[t(x) for x in tp], torch.empty(0, device=dev), t(sc.templates), t(sc.bmp_sel),
                    0.0, torch.zeros(7, device=dev))
    target = (c + (1 - a).unsqueeze(-1)).contiguous()
    init = [t(x) for x in (sc.means2D, sc.radii, sc.rotations, sc.opacities, sc.colors)]
    job = FitJob(*init, t(sc.templates), t(sc.bmp_sel), target, lr=1e-2, max_steps=steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        job.step()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    L = job.loss_history().cpu().numpy()
    return {"steps": steps, "steps_per_s": round(steps / el, 2), "loss_first": round(float(L[0]), 6),
            "loss_last": round(float(L[-1]), 6)}


def make_problem(sc, dev, sdt):
    import torch
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    return dict(params=[t(a) for a in (sc.means2D, sc.radii, sc.rotations, sc.opacities, sc.colors)],
                templates=t(sc.templates).to(sdt), sel=t(sc.bmp_sel), target=t(sc.target))


def run_fused(args, sc, dev, sdt):
    """Device-resident FitJob: every stage of the step is a HIP kernel."""
    import torch
    from diffbmp_amd.fit import FitJob
    pr = make_problem(sc, dev, sdt)
    job = FitJob(*pr["params"], pr["templates"], pr["sel"], pr["target"], lr=1e-2, store_dtype=sdt,
                 max_steps=args.warmup + 2 * args.steps + 16)
    if args.eps_t is not None:
        job.eps_T = args.eps_t

    def fwd_bwd():
        job.fwd_bwd()

    def reset():
        job.grad_rows.zero_()

    return job, job.step, job.handle, fwd_bwd, reset


def run_autograd(args, sc, dev, sdt, fused_loss=False, calling_thread=False, fused_adam=False):
    """The reference-compatible TileRasterizer behind torch autograd, with the
    background composite + MSE in torch as the reference's renderer writes it,
    or (`fused_loss`) through the one-kernel composite_mse_loss op.
    `calling_thread`: the backward runs on the thread that calls it
    (torch.autograd.set_multithreading_enabled(False)) instead of being handed
    to torch's per-device autograd thread.  `fused_adam`: the step's Adam and
    clamp are diffbmp_amd.optim.FusedAdam (one kernel) instead of
    torch.optim.Adam and three clamp_ calls."""
    import torch
    import torch.nn.functional as F
    from diffbmp_amd.tile_rasterizer import TileRasterizer
    pr = make_problem(sc, dev, sdt)
    leaves = [p.clone().requires_grad_() for p in pr["params"]]
    H, W, N = sc.H, sc.W, sc.N
    rast = TileRasterizer(H, W, 16, 0.0, 1.0, 151, N, use_fp16=(sdt != torch.float32), storage_dtype=sdt)
    if args.eps_t is not None:
        rast.eps_T = args.eps_t
    rmax = float(min(H, W) // 4)
    if fused_adam:
        from diffbmp_amd.optim import FusedAdam
        opt = FusedAdam([{"params": [leaves[0]], "clamp": ((0.0, 0.0), (float(W - 1), float(H - 1)))},
                         {"params": [leaves[1]], "clamp": (2.0, rmax)},
                         {"params": leaves[2:]}], lr=1e-2, handle=rast.handle)
    else:
        opt = torch.optim.Adam(leaves, lr=1e-2)
    bg = torch.ones((H, W, 3), device=dev)
    empty = torch.empty(0, device=dev)
    lr_conf = torch.tensor([1e-2, 1, 1, 1, 1, 1, 1], device=dev)
    templates, sel, target = pr["templates"], pr["sel"], pr["target"]

    def fwd_bwd():
        opt.zero_grad(set_to_none=True)
        color, alpha = rast(*leaves, empty, templates, sel, 0.0, lr_conf, None)
        if fused_loss:
            loss = rast.composite_mse_loss(color, alpha, target)
        else:
            loss = F.mse_loss(color.float() + (1.0 - alpha.float()).unsqueeze(-1) * bg, target)
        if calling_thread:
            with torch.autograd.set_multithreading_enabled(False):
                loss.backward()
        else:
            loss.backward()

    def step():
        fwd_bwd()
        opt.step()
        if not fused_adam:
            with torch.no_grad():
                leaves[0][:, 0].clamp_(0.0, float(W - 1))
                leaves[0][:, 1].clamp_(0.0, float(H - 1))
                leaves[1].clamp_(2.0, rmax)

    return rast, step, rast.handle, fwd_bwd, (lambda: None)


def gpu_prewarm(sc, dev, sdt, seconds=0.5):
    """Untimed: fit steps of a throwaway copy of the workload until the GPU has
    been busy for ``seconds``, so that the timed runs below start on a GPU
    whose clocks have ramped up.  The timed jobs are built afterwards from the
    original parameters, so they time the same steps as without it."""
    import torch
    from diffbmp_amd.fit import FitJob
    pr = make_problem(sc, dev, sdt)
    job = FitJob(*pr["params"], pr["templates"], pr["sel"], pr["target"], lr=1e-2, store_dtype=sdt,
                 max_steps=4096)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds and job.t + 16 <= 4096:
        for _ in range(16):
            job.step()
        torch.cuda.synchronize()
    del job, pr


def time_mode(args, runner, sc, dev, sdt):
    import torch
    obj, step, handle, fwd_bwd, reset = runner(args, sc, dev, sdt)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    handle.timing_reset()
    handle.timing_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    handle.timing_enable(False)
    f_ms, f_n, b_ms, b_n = handle.timing_stats()
    stream = torch.cuda.current_stream(dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    lat = []
    for _ in range(min(args.steps, 100)):
        e0.record(stream)
        step()
        e1.record(stream)
        e1.synchronize()
        lat.append(e0.elapsed_time(e1))
    # the hot path alone (binning + forward + backward, with the loss that
    # feeds it), hipEvents on the stream around K passes
    e0.record(stream)
    for _ in range(args.steps):
        fwd_bwd()
    e1.record(stream)
    e1.synchronize()
    fb_s = e0.elapsed_time(e1) * 1e-3 / args.steps
    reset()
    return dict(obj=obj, handle=handle, total=t1 - t0, lat=np.array(lat),
                fwd_s=f_ms / max(f_n, 1) * 1e-3, bwd_s=b_ms / max(b_n, 1) * 1e-3, fwd_bwd_s=fb_s)


def psnr_compare(args, sc, dev, sdt):
    """Fit `psnr_steps` Adam steps from the same start at fp32 and at the chosen
    storage precision; PSNR (dB) of the final render against the target."""
    import torch
    from diffbmp_amd.fit import FitJob
    res = {}
    for name, dt in (("fp32", torch.float32), (args.precision, sdt)):
        if name in res:
            continue
        pr = make_problem(sc, dev, dt)
        job = FitJob(*pr["params"], pr["templates"], pr["sel"], pr["target"], lr=1e-2, store_dtype=dt,
                     max_steps=args.psnr_steps + 1)
        for _ in range(args.psnr_steps):
            job.step()
        img = job.render()
        mse = float(((img - pr["target"]) ** 2).mean().item())
        res[name] = round(FitJob.psnr(mse), 4)
    res["steps"] = args.psnr_steps
    return res


def list_stats(handle, sc, dev, sdt, eps_T):
    """Pair count and list-walk depth at the initial parameters."""
    import torch
    from diffbmp_amd.tile_rasterizer import _Run
    pr = make_problem(sc, dev, sdt)
    run = _Run(handle, sdt, *pr["params"], None, pr["templates"], pr["sel"], 0.0, None, eps_T)
    _, _, _, last = run.forward()
    torch.cuda.synchronize()
    offs = run.lists.offsets.long()
    lens = (offs[1:] - offs[:-1]).float()
    walked = (last & 0x3FFFFFFF).float().mean().item()
    return int(offs[-1].item()), float(lens.mean().item()), float(walked)


def run_stress(args, dev):
    """Fused fit step on the stress workload: 2048^2 canvas, 32768 primitives,
    16 RGBA 256^2 bitmaps (alpha template + per-template colour map) blended
    with c_blend 0.5, storage in --precision (fp16 by default here)."""
    import torch
    from diffbmp_amd import scenes as S
    from diffbmp_amd.fit import FitJob

    sc = S.make_scene(3)
    sdt = {"fp32": torch.float32, "fp16": torch.float16, "bf16": torch.bfloat16}[args.precision]
    pr = make_problem(sc, dev, sdt)
    co = torch.from_numpy(np.ascontiguousarray(sc.colors_orig_tmpl)).to(dev)
    job = FitJob(*pr["params"], pr["templates"], pr["sel"], pr["target"], lr=1e-2, store_dtype=sdt,
                 max_steps=args.warmup + args.steps + 8, colors_orig=co, c_blend=sc.c_blend,
                 colors_orig_per_template=True)
    for _ in range(args.warmup):
        job.step()
    torch.cuda.synchronize()
    job.handle.timing_reset()
    job.handle.timing_enable(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        job.step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    job.handle.timing_enable(False)
    f_ms, f_n, b_ms, b_n = job.handle.timing_stats()
    T, Ht, Wt = sc.templates.shape
    out = {
        "metric": "fit steps/s at 2048^2, 32768 prims, 16 RGBA 256^2 bitmaps",
        "value": round(args.steps / (t1 - t0), 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round((t1 - t0) / args.steps * 1e3, 4),
        "higher_is_better": True,
        "dtype": {"fp32": "f32", "fp16": "f16-store/f32-acc", "bf16": "bf16-store/f32-acc"}[args.precision],
        "data": "synthetic",
        "config": {"workload": "stress: 2048x2048 RGB, 32768 primitives, 16x256^2 RGBA bitmaps "
                               "(per-template colour maps, c_blend 0.5), Adam lr 1e-2 on MSE",
                   "canvas": [sc.H, sc.W], "primitives": sc.N, "bitmaps": [T, Ht, Wt],
                   "precision": args.precision},
        "raster_fwd_ms": round(f_ms / max(f_n, 1), 4),
        "raster_bwd_ms": round(b_ms / max(b_n, 1), 4),
        "final_loss": float(job.losses[job.t - 1].item()),
    }
    print(json.dumps(out))


def stress_precision_summary(dev, steps=40, warmup=10, psnr_steps=200):
    """The stress workload (2048^2, 32768 primitives, 16 RGBA 256^2 bitmaps,
    c_blend 0.5) with fp32, fp16 and mixed storage: step rate, its ratio to
    fp32, and the PSNR (dB) after `psnr_steps` Adam steps from one start.  Its
    RGBA records (64 B per texel in fp32, 32 B in fp16) do not fit the L2, so
    the storage width shows in the step time here."""
    import torch
    from diffbmp_amd import scenes as S
    from diffbmp_amd.fit import FitJob
    sc = S.make_scene(3)
    co = torch.from_numpy(np.ascontiguousarray(sc.colors_orig_tmpl)).to(dev)
    out = {}
    for name, dt, geom32 in (("fp32", torch.float32, False), ("fp16", torch.float16, False),
                             ("fp16-app", torch.float16, True)):
        def new_job(n):
            pr = make_problem(sc, dev, dt)
            return pr, FitJob(*pr["params"], pr["templates"], pr["sel"], pr["target"], lr=1e-2,
                              store_dtype=dt, fp32_geometry=geom32, max_steps=n + 1, colors_orig=co,
                              c_blend=sc.c_blend, colors_orig_per_template=True)
        _, job = new_job(warmup + steps)
        for _ in range(warmup):
            job.step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            job.step()
        torch.cuda.synchronize()
        rate = steps / (time.perf_counter() - t0)
        del job
        pr, job = new_job(psnr_steps)
        for _ in range(psnr_steps):
            job.step()
        img = job.render()
        mse = float(((img - pr["target"]) ** 2).mean().item())
        out[name] = {"steps_per_s": round(rate, 2), "psnr_db": round(FitJob.psnr(mse), 4)}
        del job, pr
    for name in out:
        out[name]["speed_vs_fp32"] = round(out[name]["steps_per_s"] / out["fp32"]["steps_per_s"], 4)
        out[name]["psnr_minus_fp32_db"] = round(out[name]["psnr_db"] - out["fp32"]["psnr_db"], 4)
    out["psnr_steps"] = psnr_steps
    out["workload"] = ("stress: 2048x2048, 32768 primitives, 16x256^2 RGBA bitmaps (per-template "
                       "colour maps, c_blend 0.5), Adam lr 1e-2 on MSE")
    return out


def serving_summary(dev, jobs=40, rate=10.0, steps=200, seed=2025):
    """A bounded serving run for the headline line: ``jobs`` fit jobs of the
    serving mix arriving as a Poisson stream of ``rate`` jobs/s."""
    import torch
    from diffbmp_amd import serve as SV
    SV.template_library(256)
    rng = np.random.default_rng(seed)
    specs = [SV.random_spec(rng, steps=steps) for _ in range(jobs)]
    for s, a in zip(specs, SV.poisson_arrivals(rate, jobs, rng)):
        s.arrival = float(a)
    warm = SV.JobSpec(256, 256, 512, (0,), steps=5, seed=1, name="warm")
    SV.Server(SV.BatchEngine(store_dtype=torch.float16, device=dev), SV.HbmBudget(1 << 34)).run([warm])
    eng = SV.BatchEngine(store_dtype=torch.float16, device=dev)
    stats = SV.Server(eng, SV.HbmBudget(64 << 30), max_jobs=32).run(specs)
    lat = np.array([r.latency for r in stats.results])
    return {"jobs": jobs, "arrival_rate_jobs_per_s": rate, "steps_per_job": steps,
            "jobs_per_s": round(len(stats.results) / stats.wall, 3),
            "latency_p50_s": round(float(np.percentile(lat, 50)), 4),
            "latency_p99_s": round(float(np.percentile(lat, 99)), 4),
            "workload": "serving mix: canvas side U{256,512,1024,2048}, primitives log-U[512,32768], "
                        "1-16 RGBA 256^2 bitmaps, fp16 storage, 200 Adam steps per job"}


def run_serving(args, dev):
    """Poisson stream of fit jobs (canvas side uniform over {256, 512, 1024,
    2048}, primitives log-uniform in [512, 32768], 1-16 of the stress
    workload's 256^2 RGBA bitmaps, fp16 storage / fp32 arithmetic, 200 Adam
    steps each to a smooth random field) through the batched engine."""
    import torch
    from diffbmp_amd import serve as SV

    SV.template_library(256)          # host-side bitmap generation, outside the timed run
    rng = np.random.default_rng(2024)
    specs = [SV.random_spec(rng, steps=args.serve_steps) for _ in range(args.serve_jobs)]
    if args.serve_rate > 0:
        t = SV.poisson_arrivals(args.serve_rate, len(specs), rng)
        for s, a in zip(specs, t):
            s.arrival = float(a)
    for i, s in enumerate(specs):
        s.name = f"job{i}"
    # warm-up: one small job end to end (kernel loading, graph capture path)
    warm = SV.JobSpec(256, 256, 512, (0,), steps=5, seed=1, name="warm")
    eng = SV.BatchEngine(store_dtype=torch.float16, use_graph=not args.serve_no_graph, device=dev)
    SV.Server(eng, SV.HbmBudget(1 << 34)).run([warm])
    eng = SV.BatchEngine(store_dtype=torch.float16, use_graph=not args.serve_no_graph, device=dev,
                         share_bitmaps=not args.serve_no_share)
    budget = SV.HbmBudget(int(args.serve_budget_gib * (1 << 30)))
    srv = SV.Server(eng, budget, max_jobs=args.serve_max_jobs)
    stats = srv.run(specs)
    summ = stats.summary()
    lat = np.array([r.latency for r in stats.results])
    q = np.array([r.queue for r in stats.results])
    prims = sum(s.num_primitives for s in specs)
    px = sum(s.height * s.width for s in specs)
    out = {
        "metric": "serving jobs/s (200 Adam steps per job)",
        "value": round(summ["jobs_per_s"], 3),
        "unit": "jobs/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "f16-store/f32-acc",
        "data": "synthetic",
        "config": {"workload": "serving: Poisson arrivals of fit jobs, canvas side U{256,512,1024,2048}, "
                               "primitives log-U[512,32768], 1-16 RGBA 256^2 bitmaps, c_blend 0.5, "
                               "fp16 storage, Adam lr 1e-2 on MSE to a smooth random field",
                   "jobs": len(specs), "steps_per_job": args.serve_steps,
                   "arrival_rate": args.serve_rate or "all at t=0",
                   "max_jobs": args.serve_max_jobs, "hbm_budget_gib": args.serve_budget_gib,
                   "graph": not args.serve_no_graph, "shared_bitmaps": not args.serve_no_share},
        "latency_p50_s": round(float(np.percentile(lat, 50)), 4),
        "latency_p99_s": round(float(np.percentile(lat, 99)), 4),
        "queue_p50_s": round(float(np.percentile(q, 50)), 4),
        "wall_s": round(summ["wall_s"], 3),
        "batched_steps": summ["batched_steps"],
        "ms_per_batched_step": round(summ["wall_s"] / max(summ["batched_steps"], 1) * 1e3, 4),
        "max_active": summ["max_active"],
        "total_primitives": prims,
        "total_megapixels": round(px / 1e6, 2),
        "mean_final_over_first_loss": round(float(np.mean([r.final_loss / r.first_loss
                                                           for r in stats.results])), 4),
    }
    print(json.dumps(out))


def run_lanes(args, dev):
    """Lane utilisation of the raster kernels, from the profiling build
    (libdiffbmp_amd_lanestats.so, loaded through DBMP_LIBRARY): per (wave,
    list entry) evaluation of the forward and backward, the share with a
    pushed lane and the pushed lanes per evaluation, over fused fit steps of
    the scale and stress workloads."""
    import ctypes
    import torch
    from diffbmp_amd import _native
    from diffbmp_amd import scenes as S
    from diffbmp_amd.fit import FitJob
    if "lanestats" not in _native.LIB_PATH:
        raise SystemExit("--workload lanes needs DBMP_LIBRARY=diffbmp_amd/libdiffbmp_amd_lanestats.so "
                         "(python -m diffbmp_amd.build --lane-stats)")
    L = _native.lib()
    out = {"metric": "raster lane utilisation", "library": os.path.basename(_native.LIB_PATH),
           "steps": args.steps}
    for name, cfg, sdt, co in (("scale", 1, torch.float32, False), ("stress", 3, torch.float16, True)):
        sc = S.make_scene(cfg)
        pr = make_problem(sc, dev, sdt)
        kw = {}
        if co:
            kw = dict(colors_orig=torch.from_numpy(np.ascontiguousarray(sc.colors_orig_tmpl)).to(dev),
                      c_blend=sc.c_blend, colors_orig_per_template=True)
        job = FitJob(*pr["params"], pr["templates"], pr["sel"], pr["target"], lr=1e-2, store_dtype=sdt,
                     max_steps=args.warmup + args.steps + 4, **kw)
        for _ in range(args.warmup):
            job.step()
        st = (ctypes.c_uint64 * 8)()
        _native.check(L.dbmp_lane_stats(st, 1))
        for _ in range(args.steps):
            job.step()
        _native.check(L.dbmp_lane_stats(st, 1))
        res = {}
        for k, b in (("forward", 0), ("backward", 3)):
            ev, hit, lanes = st[b], st[b + 1], st[b + 2]
            res[k] = {"evaluations_per_step": ev // args.steps,
                      "share_with_a_pushed_lane": round(hit / max(ev, 1), 4),
                      "pushed_lanes_per_evaluation": round(lanes / max(ev, 1), 2),
                      "pushed_lanes_per_pushing_evaluation": round(lanes / max(hit, 1), 2),
                      "lane_utilisation": round(lanes / max(64 * ev, 1), 4)}
        out[name] = res
    print(json.dumps(out))


def run_export(args, dev):
    """Layer export of the scale scene: the uint8 RGBA crop (1.5 r square) of
    every primitive into one packed buffer; HBM-write bound."""
    import ctypes
    import torch
    from diffbmp_amd import _native
    from diffbmp_amd import scenes as S
    from diffbmp_amd.psd_export_uint8 import layer_bounds

    sc = S.make_scene(1)
    N = sc.N
    rng = np.random.default_rng(5)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    md, rd, thd, seld, tmd = t(sc.means2D), t(sc.radii), t(sc.rotations), t(sc.bmp_sel), t(sc.templates)
    cd = t(rng.uniform(0, 1, (N, 3)).astype(np.float32))
    vd = t(rng.uniform(0, 1, N).astype(np.float32))
    boxes, sizes, offs, total = layer_bounds(md, rd, sc.H, sc.W, 1.0)
    buf = torch.empty(total, dtype=torch.uint8, device=dev)
    amax = torch.empty(N, dtype=torch.int32, device=dev)
    p = _native.LayerPrims(md.data_ptr(), rd.data_ptr(), thd.data_ptr(), cd.data_ptr(), vd.data_ptr(),
                           seld.data_ptr(), N)
    T, Ht, Wt = sc.templates.shape
    st = torch.cuda.current_stream(dev)
    call = lambda: _native.check(_native.lib().dbmp_psd_layers(
        ctypes.byref(p), tmd.data_ptr(), T, Ht, Wt, None, 0, 0.0, sc.H, sc.W, 1.0, 1.0,
        boxes.data_ptr(), offs.data_ptr(), total, buf.data_ptr(), amax.data_ptr(), _native._stream(dev)))
    for _ in range(args.warmup):
        call()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(args.steps):
        call()
    e1.record(st)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    px = total // 4
    # algorithmic bytes: the crops written once, the per-layer constants and
    # the templates read once
    algo = total + N * 40 + T * Ht * Wt * 4
    out = {
        "metric": "layer export GB/s (uint8 RGBA crops, 1024^2, 8192 prims)",
        "value": round(total / ms / 1e6, 2), "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": "layer export: scale scene (1024x1024, 8192 primitives, 4x128^2 bitmaps), "
                               "1.5 r crops, scale 1, no colour maps", "layers": N, "pixels": int(px)},
        "roofline": {"bound": "hbm", "achieved": round(algo / ms / 1e6, 2), "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": round(algo / ms / 1e6 / (HBM_PEAK / 1e9), 5),
                     "traffic": None, "kernel": "k_psd_layers", "algorithmic_bytes": int(algo)},
    }
    print(json.dumps(out))


def precision_summary(sc, dev, steps=200, warmup=20, psnr_steps=200):
    """Step rate of the fused step with fp32, fp16 and mixed storage (fp32
    positions, radii and rotations; fp16 opacities, colours and bitmaps), each
    as a ratio to fp32, and the PSNR (dB) each reaches after `psnr_steps` Adam
    steps from the same start."""
    import torch
    from diffbmp_amd.fit import FitJob
    kinds = (("fp32", torch.float32, False), ("fp16", torch.float16, False),
             ("fp16-app", torch.float16, True))
    out = {}
    for name, dt, geom32 in kinds:
        pr = make_problem(sc, dev, dt)
        job = FitJob(*pr["params"], pr["templates"], pr["sel"], pr["target"], lr=1e-2, store_dtype=dt,
                     fp32_geometry=geom32, max_steps=warmup + steps + 1)
        for _ in range(warmup):
            job.step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            job.step()
        torch.cuda.synchronize()
        rate = steps / (time.perf_counter() - t0)
        del job
        pr = make_problem(sc, dev, dt)
        job = FitJob(*pr["params"], pr["templates"], pr["sel"], pr["target"], lr=1e-2, store_dtype=dt,
                     fp32_geometry=geom32, max_steps=psnr_steps + 1)
        for _ in range(psnr_steps):
            job.step()
        img = job.render()
        mse = float(((img - pr["target"]) ** 2).mean().item())
        out[name] = {"steps_per_s": round(rate, 2), "psnr_db": round(FitJob.psnr(mse), 4)}
        del job
    for name in out:
        out[name]["speed_vs_fp32"] = round(out[name]["steps_per_s"] / out["fp32"]["steps_per_s"], 4)
        out[name]["psnr_minus_fp32_db"] = round(out[name]["psnr_db"] - out["fp32"]["psnr_db"], 4)
    out["psnr_steps"] = psnr_steps
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the rasterizer "
              "path has no multi-GPU mode.", file=sys.stderr)
        sys.exit(2)
    import torch
    from diffbmp_amd import scenes as S

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if args.workload == "serving":
        run_serving(args, dev)
        return
    if args.workload == "stress":
        run_stress(args, dev)
        return
    if args.workload == "export":
        run_export(args, dev)
        return
    if args.workload == "lanes":
        run_lanes(args, dev)
        return
    if args.copy_probe:
        nb = args.copy_probe << 20
        src = torch.ones(nb // 4, device=dev)
        dst = torch.empty_like(src)
        torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
    sc = S.make_scene(1)
    H, W, N = sc.H, sc.W, sc.N
    T, Ht, Wt = sc.templates.shape
    sdt = {"fp32": torch.float32, "fp16": torch.float16, "bf16": torch.bfloat16}[args.precision]
    runners = {"fused": run_fused, "autograd": run_autograd}
    if args.prewarm > 0:
        gpu_prewarm(sc, dev, sdt, args.prewarm)
    res = time_mode(args, runners[args.mode], sc, dev, sdt)
    other = "autograd" if args.mode == "fused" else "fused"
    res_other = time_mode(args, runners[other], sc, dev, sdt)
    res_fl = time_mode(args, lambda a, s, d, t: run_autograd(a, s, d, t, fused_loss=True), sc, dev, sdt)
    res_ct = time_mode(args, lambda a, s, d, t: run_autograd(a, s, d, t, fused_loss=True,
                                                             calling_thread=True), sc, dev, sdt)
    res_fa = time_mode(args, lambda a, s, d, t: run_autograd(a, s, d, t, fused_loss=True,
                                                             fused_adam=True), sc, dev, sdt)
    eps_T = args.eps_t if args.eps_t is not None else 1e-6
    P, list_mean, walked = list_stats(res["handle"], sc, dev, sdt, eps_T)

    traffic = None
    # the dominant kernel: the backward (with the fused composite/MSE prologue
    # in the fused step)
    bwd_kernel = "k_raster_bwd_mse" if args.mode == "fused" else "k_raster_bwd"
    prof = os.path.join(HERE, "profiles", "pmc_traffic_scale_fp32.json")
    if args.precision == "fp32" and os.path.exists(prof):
        # HBM bytes per backward launch from the committed FETCH_SIZE / WRITE_SIZE
        # passes of this workload, scaled by the copy-probe factors recorded there
        with open(prof) as f:
            traffic = json.load(f)["kernels"].get(bwd_kernel, {}).get("traffic_bytes")
    b_p = torch.tensor([], dtype=sdt).element_size()
    b_bin, b_fwd, b_bwd = algorithmic_bytes(H, W, N, T, Ht, Wt, P, res["handle"].tiles, b_p, b_p)
    achieved = b_bwd / res["bwd_s"]
    lat = res["lat"]
    out = {
        "metric": "fwd+bwd iters/s at 1024^2, 8192 prims",
        "value": round(args.steps / res["total"], 3),
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["total"] / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp32": "f32", "fp16": "f16-store/f32-acc", "bf16": "bf16-store/f32-acc"}[args.precision],
        "data": "synthetic",
        "config": {"workload": "scale: 1024x1024 RGB, 8192 primitives, 4x128^2 bitmaps, 16x16 tiles, "
                               "Adam lr 1e-2 on MSE to a smooth random field; one step = binning + "
                               "forward + composite/MSE + backward + Adam + clamp",
                   "canvas": [H, W], "primitives": N, "bitmaps": [T, Ht, Wt], "tile": 16,
                   "eps_T": eps_T, "precision": args.precision, "mode": args.mode},
        "step_ms_p50": round(float(np.percentile(lat, 50)), 4),
        "step_ms_p99": round(float(np.percentile(lat, 99)), 4),
        "raster_fwd_ms": round(res["fwd_s"] * 1e3, 4),
        "raster_bwd_ms": round(res["bwd_s"] * 1e3, 4),
        f"{other}_iters_per_s": round(args.steps / res_other["total"], 3),
        # binning + forward + backward alone (the step above adds the Adam
        # update and clamp)
        "fwd_bwd_iters_per_s": round(1.0 / res["fwd_bwd_s"], 3),
        f"{other}_fwd_bwd_iters_per_s": round(1.0 / res_other["fwd_bwd_s"], 3),
        # the drop-in autograd path with the composite + MSE in one kernel
        # (composite_mse_loss) instead of torch's elementwise and reduction ops
        "autograd_fused_loss_iters_per_s": round(args.steps / res_fl["total"], 3),
        "autograd_fused_loss_fwd_bwd_iters_per_s": round(1.0 / res_fl["fwd_bwd_s"], 3),
        # the same with the backward on the calling thread: the default hands
        # each backward to torch's autograd device thread, whose hand-off costs
        # host time that this launch-heavy pass cannot hide
        "autograd_fused_loss_calling_thread_iters_per_s": round(args.steps / res_ct["total"], 3),
        "autograd_fused_loss_calling_thread_fwd_bwd_iters_per_s": round(1.0 / res_ct["fwd_bwd_s"], 3),
        # the drop-in step with diffbmp_amd.optim.FusedAdam (Adam + clamp in
        # one kernel) in place of torch.optim.Adam and the clamp_ calls
        "autograd_fused_loss_fused_adam_iters_per_s": round(args.steps / res_fa["total"], 3),
        "samples_per_s_fwd": float(f"{walked * H * W / res['fwd_s']:.4g}"),
        "pairs": P,
        "list_len_mean": round(list_mean, 2),
        "entries_walked_per_pixel": round(walked, 2),
        "roofline": {"bound": "hbm", "achieved": round(achieved / 1e9, 2), "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 5), "traffic": traffic,
                     "traffic_source": "profiles/pmc_traffic_scale_fp32.json" if traffic else None,
                     "kernel": bwd_kernel, "algorithmic_bytes": int(b_bwd),
                     "step_algorithmic_bytes": int(b_bin + b_fwd + b_bwd)},
    }
    # the bound the raster kernels actually sit at: VALU instruction issue
    # (instruction count per launch from the committed SQ_INSTS_VALU pass of
    # this workload, divided by the live hipEvent duration)
    vprof = os.path.join(HERE, "profiles", "pmc_valu_scale_fp32.json")
    if args.precision == "fp32" and os.path.exists(vprof):
        with open(vprof) as f:
            vk = json.load(f)["kernels"].get(bwd_kernel)
        if vk:
            ach = vk["SQ_INSTS_VALU"] / res["bwd_s"]
            out["issue_roofline"] = {"bound": "valu_issue", "achieved": round(ach / 1e9, 2),
                                     "peak": VALU_ISSUE_PEAK / 1e9, "unit": "G wave64 VALU instr/s",
                                     "frac": round(ach / VALU_ISSUE_PEAK, 4), "kernel": bwd_kernel,
                                     "valu_instructions": vk["SQ_INSTS_VALU"],
                                     "source": "profiles/pmc_valu_scale_fp32.json"}
    # LDS: bytes the backward moves through LDS per launch (committed
    # SQ_LDS_IDX_ACTIVE pass: 128 B per conflict-free active cycle) over the
    # live duration, against 128 B/clk per CU
    lprof = os.path.join(HERE, "profiles", "pmc_lds_scale_fp32.json")
    if args.precision == "fp32" and os.path.exists(lprof):
        with open(lprof) as f:
            lk = json.load(f)["kernels"].get(bwd_kernel)
        if lk:
            ach = lk["lds_bytes"] / res["bwd_s"]
            out["lds_roofline"] = {"bound": "lds", "achieved": round(ach / 1e9, 1),
                                   "peak": LDS_PEAK / 1e9, "unit": "GB/s",
                                   "frac": round(ach / LDS_PEAK, 4), "kernel": bwd_kernel,
                                   "lds_bytes": lk["lds_bytes"],
                                   "source": "profiles/pmc_lds_scale_fp32.json"}
    if args.psnr_steps:
        out["psnr"] = psnr_compare(args, sc, dev, sdt)
    if not args.no_precision:
        out["storage_precision"] = precision_summary(sc, dev)
        out["stress_storage_precision"] = stress_precision_summary(dev)
    out["plain_fit"] = gpu_plain_fit(dev)
    if not args.no_serving:
        out["serving"] = serving_summary(dev)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(sc, args.cpu_steps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
