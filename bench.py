"""Training-step throughput of the CVQPG hot path on one MI355X.

Workload (the metric's configuration): RGB 1920x1080, 150,000 Gaussians, FP32 /
complex64 field, 3 wavelengths (639/515/473 nm), 8 um pitch, 2x padded BLASM
(2816 x 4480 FFT grid) to 4 planes at 0, 2, 4, 6 mm; one step = rasterize -> pad ->
glue -> BLASM -> multi-plane L2 + SSIM loss -> full adjoint -> Adan.

Prints one JSON line (rank 0).  ``time_to_psnr`` is a reduced learned-warm-start study
(8 training pairs from 400-iteration 1080p fits, 800 render-loss fine-tuning steps, 2
held-out 1080p scenes; ``--workload warmstart`` runs the full one): speed-up of the warm
start to the plain fit's final per-plane PSNR.  ``rooflines`` has one entry per hot kernel
(raster adjoint and forward, the two BLASM column passes and the two per-plane row passes,
the loss), each timed with HIP events on the replay stream inside the timed steps (its stage
is captured as a graph of its own), with its algorithmic rate against the published peak
and its measured HBM bytes; ``roofline`` is the entry of the longest kernel.
``cpu_baseline`` times the reference's PyTorch path on the host CPU on a bounded sample of
the same workload (see ``cpu_baseline``).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
METRIC = "train step/s @1080p RGB 150k Gaussians; time-to-PSNR warm-start vs plain"
WORKLOAD = "rgb1080p_150k"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--path", default="fused", choices=["fused", "autograd", "torch"],
                    help="fused: all-HIP graph-captured step; autograd: reference-structured "
                         "step (HIP raster, PyTorch elsewhere); torch: the reference's PyTorch "
                         "path on the GPU (3-sigma / 64-px tile renderer, torch.fft BLASM, "
                         "torch loss and Adan), the denominator of the speed-up target")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-full-step", action="store_true",
                    help="also time one whole step of the reference's PyTorch path at full "
                         "size on the host CPU (about a minute) and report it as the CPU "
                         "baseline, next to the sampled estimate")
    ap.add_argument("--no-warmstart", action="store_true",
                    help="skip the reduced warm-start time-to-PSNR study of the default run")
    ap.add_argument("--workload", default=WORKLOAD,
                    help="rgb1080p_150k (the metric's workload), gray256_10k, rgb4k_600k, "
                         "batch16 (16 uneven scenes, batched vs sequential) or warmstart "
                         "(full learned warm-start time-to-PSNR study)")
    ap.add_argument("--warm-pairs", type=int, default=16,
                    help="--workload warmstart: training pairs (1080p plain fits)")
    ap.add_argument("--warm-train-steps", type=int, default=600,
                    help="--workload warmstart: supervised Adam steps on the pairs")
    ap.add_argument("--warm-eval", type=int, default=4,
                    help="--workload warmstart: held-out 1920x1080 scenes (2000-iteration "
                         "plain fits); the 512x512 set keeps 4 scenes")
    ap.add_argument("--warm-out", default=None, metavar="PATH",
                    help="--workload warmstart: also write the whole result (every held-out "
                         "scene's row) as JSON to PATH")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="run one device-to-device copy of MIB MiB after the timed steps (a "
                         "plain copy kernel of known size for scaling HBM byte counters)")
    return ap.parse_args()


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hot_kernels_traffic_rgb1080p_150k.txt")
VALU_FILE = os.path.join(ROOT, "profiles", "hot_kernels_valu_rgb1080p_150k.txt")


def read_kv(path):
    """key=value lines of a committed profile summary ('#' lines are comments)."""
    vals = {}
    if os.path.exists(path):
        with open(path) as f:
            for line in f:
                if "=" in line and not line.lstrip().startswith("#"):
                    k, v = line.split("=", 1)
                    vals[k.strip()] = v.strip()
    return vals


# Hot kernels of the step: (stage of FusedTrainer(fine=True), kernel, bound).  Their
# algorithmic work per launch follows the unit definitions of the hot-path measurement
# plan:
#   raster: unit = reference pixel-pair = keys x 256; per pair (45 + 14 C) FP32 VALU ops +
#           (1 + 2 C) transcendentals in the adjoint, (20 + 8 C) + (1 + 2 C) in the forward
#           (all pairs of the reference's tile lists, including those the kernels skip);
#   BLASM:  unit = grid point of a one-dimensional pass over the 2x grid, 16 B (8 B read +
#           8 B written, complex64); the column pass of a direction does 1 + D transforms
#           of every column of the 2x grid per channel, the per-plane row pass D per
#           channel, of the rows that survive the crop to the image (the other rows of
#           the 2x grid are cropped away before that pass, so it never transforms them);
#   loss:   16 B per pixel and plane (read intensity, target, mask; write the gradient).
HOT = (("render_bwd", "k_render_bwd", "valu"), ("blasm_bwd_cols", "k_a2_cols", "hbm"),
       ("blasm_fwd_cols", "k_f2_cols", "hbm"), ("render_fwd", "k_render_fwd", "valu"),
       ("loss", "k_loss_stream", "hbm"), ("blasm_bwd_rows", "k_a3_rows", "hbm"),
       ("blasm_fwd_out", "k_f3_rows", "hbm"))


def raster_ops_per_pair(C, adjoint):
    """(VALU ops, transcendentals) per reference pixel-pair."""
    return (45 + 14 * C, 1 + 2 * C) if adjoint else (20 + 8 * C, 1 + 2 * C)


def kernel_model(stage, spec, pairs):
    """Algorithmic work of one launch of a hot kernel: (amount, unit of the rate)."""
    W, H = spec["W"], spec["H"]
    C = len(spec["wavelengths"])
    sys.path.insert(0, ROOT)
    from cvqpg_amd import synthetic
    D = len(synthetic.distances_for(spec))
    pH, pW = synthetic.pad_size_for(W, H)
    grid = 4 * pH * pW
    if stage in ("render_bwd", "render_fwd"):
        v, t = raster_ops_per_pair(C, stage == "render_bwd")
        return pairs * (v + t), "TOP/s"
    if stage in ("blasm_bwd_cols", "blasm_fwd_cols"):
        return 16.0 * grid * C * (1 + D), "GB/s"
    if stage in ("blasm_bwd_rows", "blasm_fwd_out"):
        return 16.0 * H * (2 * pW) * C * D, "GB/s"
    if stage == "loss":
        return 16.0 * D * C * H * W, "GB/s"
    raise KeyError(stage)


def rooflines(kernel_ms, spec, pairs, workload):
    """One entry per timed hot kernel (average of HIP events bracketing its own graph on
    the replay stream, inside the timed steps), with its algorithmic rate against the
    published peak, the HBM bytes it moves per launch (committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary) and, for the VALU-bound raster, the share of cycles the SIMDs spend
    issuing VALU instructions (committed SQ counter summary)."""
    traffic = read_kv(TRAFFIC_FILE) if workload == WORKLOAD else {}
    valu = read_kv(VALU_FILE) if workload == WORKLOAD else {}
    out = []
    for stage, kern, bound in HOT:
        ms = kernel_ms.get(stage)
        if not ms:
            continue
        amount, unit = kernel_model(stage, spec, pairs)
        if unit == "TOP/s":
            achieved = amount / (ms * 1e-3) / 1e12
            peak = FP32_PEAK_TFLOPS
        else:
            achieved = amount / (ms * 1e-3) / 1e9
            peak = HBM_PEAK_GBS
        e = {"kernel": kern, "stage": stage, "bound": bound, "achieved": achieved,
             "peak": peak, "unit": unit, "frac": achieved / peak, "avg_ms": ms,
             "algorithmic": amount,
             "traffic": float(traffic[kern]) if kern in traffic else None}
        if e["traffic"]:
            e["traffic_gbs"] = e["traffic"] / (ms * 1e-3) / 1e9
            e["traffic_frac"] = e["traffic_gbs"] / HBM_PEAK_GBS
        if kern + ".valu_busy" in valu:
            e["valu_busy"] = float(valu[kern + ".valu_busy"])
        if unit == "TOP/s":
            e["pairs"] = pairs
            e["ops_per_pair"] = raster_ops_per_pair(len(spec["wavelengths"]),
                                                    stage == "render_bwd")
        out.append(e)
    return out


def cpu_threads():
    """Host threads for the CPU baseline: the host's CPUs, capped by OMP_NUM_THREADS (a
    GPU box exports its share of the host there)."""
    n = os.cpu_count() or 1
    try:
        n = min(n, int(os.environ.get("OMP_NUM_THREADS", n)))
    except ValueError:
        pass
    return max(1, n)


def cpu_baseline(spec, seed=0):
    """The reference's PyTorch path (3-sigma / 64-px tile renderer, torch.fft BLASM, torch
    loss + SSIM, Adan; the structure of train_2d_gaussian_rgb.py:187-211) timed on the host
    CPU, on a bounded sample of the workload, scaled to step/s.

    A whole 1920 x 1080 step of that path keeps every tile batch for autograd (more than
    45 GB of host memory), so the sample times, at full size, one complete instance of each
    part of the step and scales it by its multiplicity:
      (a) raster forward + autograd backward of the tile renderer on a 256 x 256 window (the
          Gaussians that reach it), x frame area / window area;
      (b) guard-band pad + glue + BLASM of one (plane, channel) on the full 2816 x 4480 grid,
          forward + autograd backward, x planes x channels;
      (c) one plane's 3-term L2 + SSIM loss at 1920 x 1080 x 3, forward + backward, x planes;
      (d) the Adan step over all Gaussians, x 1.
    The whole step of the small configuration (256 x 256 gray, 10k Gaussians) is timed as
    well (median of 3 after one warm-up), as a check of the sampling."""
    import torch

    sys.path.insert(0, ROOT)
    from oracle import model as omodel
    import __graft_entry__ as ge
    ge.load_pkg()
    from cvqpg_amd import loss as ploss, model as pmodel, synthetic as syn, train
    from cvqpg_amd.optimizer import Adan
    from cvqpg_amd.propagator import propagator
    threads = cpu_threads()
    torch.set_num_threads(threads)
    W, H, N = spec["W"], spec["H"], spec["N"]
    C = len(spec["wavelengths"])
    dist = syn.distances_for(spec)
    D = len(dist)
    pH, pW = syn.pad_size_for(W, H)
    parts = {}
    # (a) raster window
    raw = syn.make_scene(N, W, H, spec["wavelengths"], spec["pitch"], spec["kappa"], seed)
    win, x0, y0, m = 256, 832, 412, 24
    tr = {k: torch.as_tensor(v) for k, v in raw.items()}
    mt = torch.tanh(tr["means_2d"])
    cx, cy = (mt[:, 0] + 1) * 0.5 * W, (mt[:, 1] + 1) * 0.5 * H
    sel = (cx > x0 - m) & (cx < x0 + win + m) & (cy > y0 - m) & (cy < y0 + win + m)
    sub = {k: v[sel].clone().requires_grad_() for k, v in tr.items()}
    t0 = time.perf_counter()
    ri = omodel.raster_inputs(sub, (W, H), spec["wavelengths"], spec["pitch"])
    ri["means"] = ri["means"] - torch.tensor([[float(x0), float(y0)]])
    out = omodel.render_tile_torch(ri, win, win)
    (out.abs() ** 2).sum().backward()
    parts["raster"] = (time.perf_counter() - t0, W * H / (win * win))
    # (b) one (plane, channel) of pad + glue + BLASM
    g = torch.Generator().manual_seed(seed)
    h = torch.complex(torch.randn(1, H, W, generator=g), torch.randn(1, H, W, generator=g))
    h = (0.3 * h).requires_grad_()
    prop = propagator(resolution=(pH, pW), wavelengths=spec["wavelengths"][:1],
                      pixel_pitch=spec["pitch"], number_of_frames=1, distances=dist[:1],
                      aperture_size=10 * max(W, H), backend="torch")
    t0 = time.perf_counter()
    padded = pmodel.soft_window_pad(h, (pH, pW))
    ph = torch.angle(padded) % (2 * torch.pi)
    amp = torch.clamp(torch.abs(padded), 0.0, 1.0)
    inten = prop.reconstruct(ph, amplitude=amp, no_grad=False)[0, 0, 0]
    crop = inten[pH // 2 - H // 2:pH // 2 + H // 2, pW // 2 - W // 2:pW // 2 + W // 2]
    crop.sum().backward()
    parts["blasm"] = (time.perf_counter() - t0, D * C)
    # (c) one plane's loss
    # (the loss's cost does not depend on the target's content: random target and depth,
    # no defocus synthesis)
    lf = ploss.multiplane_loss_odak(torch.rand(C, H, W, generator=g), torch.rand(1, H, W, generator=g),
                                    number_of_planes=D, weights=[1.0, 1.0, 1.0, 0.0],
                                    scheme="none", backend="torch")
    pred = torch.rand(C, H, W, generator=g).requires_grad_()
    t0 = time.perf_counter()
    Lp = lf(pred, lf.targets[0], 0, inject_noise=True)
    Lp = Lp + ploss.GaussianLoss(pred, lf.targets[0], backend="torch")
    Lp.backward()
    parts["loss"] = (time.perf_counter() - t0, D)
    # (d) Adan over all Gaussians
    params = [torch.nn.Parameter(torch.as_tensor(raw[k]).clone()) for k in raw]
    for q in params:
        q.grad = torch.randn_like(q)
    opt = Adan([{"params": params, "lr": 1e-3}], lr=0, eps=1e-8, foreach=False)
    t0 = time.perf_counter()
    torch.nn.utils.clip_grad_norm_(params, 1.0)
    opt.step()
    parts["adan"] = (time.perf_counter() - t0, 1)
    t_step = sum(t * k for t, k in parts.values())
    # the small configuration's whole step
    wl0 = train.Workload("gray256_10k", device="cpu", seed=seed, renderer="torch")
    at = train.AutogradTrainer(wl0, backend="torch")
    at.step()
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        at.step()
        ts.append(time.perf_counter() - t0)
    ts.sort()
    sample = "; ".join(f"{k} {t:.2f}s x{m:.1f}" for k, (t, m) in parts.items())
    return {"value": 1.0 / t_step, "unit": "step/s", "cores": threads, "kind": "port",
            "host_cpus": os.cpu_count(),
            "sample": (f"reference PyTorch path on the host CPU, FP32, {threads} threads: one "
                       f"full-size instance of each part of the {W}x{H} step, scaled by its "
                       f"multiplicity ({sample}; {int(sel.sum())} Gaussians reach the raster "
                       f"window); whole 256x256 gray 10k step: {ts[1]:.2f}s"),
            "parts_s": {k: t for k, (t, _) in parts.items()},
            "gray256_10k_step_s": ts[1]}


def cpu_full_step(name, seed=0):
    """One whole training step of the reference's PyTorch path (``--renderer torch`` with
    ``grad_ckpt``, torch.fft BLASM, torch loss + SSIM, Adan) at the workload's full size
    on the host CPU: the tile loop keeps one tile's output per tile for autograd instead of
    its Gaussian batches, which bounds host memory.  Returns (seconds, threads)."""
    import torch
    from cvqpg_amd import train
    threads = cpu_threads()
    torch.set_num_threads(threads)
    wl = train.Workload(name, device="cpu", seed=seed, renderer="torch")
    wl.args_prop.grad_ckpt = True
    at = train.AutogradTrainer(wl, backend="torch")
    print(f"[bench] CPU full step of {name} ({threads} threads)", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    at.step()
    return time.perf_counter() - t0, threads


def stage_report(trainer, steps=3):
    """Eager per-stage GPU times (HIP events between the stages), for diagnostics."""
    import torch
    ev = {}
    for _ in range(steps):
        trainer.step_timed(ev)
    torch.cuda.synchronize()
    return {k: sum(a.elapsed_time(b) for a, b in v) / len(v) for k, v in ev.items()}


BATCH_METRIC = "aggregate train step/s, 16 uneven RGB scenes in one captured step"


def bench_batch(args, dev):
    """16 independent scenes (512^2 / 1024x768 / 1280x720 / 1920x1080, N = 0.07 x pixels
    clamped to 20k-150k, 3 planes at seeded z in 0-6 mm): aggregate step/s of the
    batched graph against the same 16 scenes stepped one after another, each from its
    own captured graph."""
    import torch
    from cvqpg_amd import batch, synthetic
    specs = synthetic.batch_specs(16, seed=0)
    print("[bench] building 16 scenes", file=sys.stderr, flush=True)
    bt = batch.make_batch(specs, device=dev, warmup_iters=400)
    log = lambda m: print(f"[bench] autotune {m}", file=sys.stderr, flush=True)
    chosen = bt.autotune(log=log)
    print(f"[bench] autotune choice {chosen}", file=sys.stderr, flush=True)
    bt.capture()
    bt.capture_sequential()
    for _ in range(args.warmup):
        bt.run()
        bt.run_sequential()
    ms_seq = batch.timed_steps(bt.run_sequential, args.steps, dev)
    ms = batch.timed_steps(bt.run, args.steps, dev)
    if not bt.check_capacity():
        raise RuntimeError("key capacity exceeded during the timed steps")
    k = {}
    for name in ("render_fwd", "render_bwd"):
        k[name] = bt._time(lambda: bt._batched(name), 10)
    keys_ref = sum(t.reference_keys() for t in bt.trainers)
    pairs = keys_ref * 256
    v, t = raster_ops_per_pair(3, True)
    achieved = pairs * (v + t) / (k["render_bwd"] * 1e-3) / 1e12
    roof = {"kernel": "k_render_bwd_batch (persistent tile-queue raster adjoint, all scenes)",
            "bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None, "avg_ms": k["render_bwd"],
            "pairs": pairs, "ops_per_pair": [v, t], "fwd_avg_ms": k["render_fwd"]}
    n = len(bt.trainers)
    out = {"metric": BATCH_METRIC, "value": n * 1e3 / ms, "unit": "scene-step/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic",
           "config": {"workload": "batch16_uneven_rgb", "scenes": [
                          [t.W, t.H, t.N] for t in bt.trainers],
                      "planes": 3, "keys": keys_ref,
                      "keys_live": sum(t.last_pairs for t in bt.trainers),
                      "sequential_value": n * 1e3 / ms_seq, "sequential_ms_per_round": ms_seq,
                      "speedup_vs_sequential": ms_seq / ms,
                      "autotune": {"render_fwd": {"grid": bt.grid["render_fwd"],
                                                  "chunk": bt.chunk["render_fwd"]},
                                   "render_bwd": {"grid": bt.grid["render_bwd"],
                                                  "chunk": bt.chunk["render_bwd"]},
                                   "streams": bt.n_streams},
                      "peak_hbm_gb": torch.cuda.max_memory_allocated(dev) / 1e9},
           "roofline": roof, "cpu_baseline": None}
    print(json.dumps(out), flush=True)


WARM_METRIC = "time-to-PSNR warm-start vs plain (speed-up at equal per-plane PSNR)"


def warm_summary(res):
    """Compact form of warmstart.experiment's result for the JSON line."""
    out = {"network": res["network"], "evals": []}
    for e in res["evals"]:
        d = {k: e[k] for k in ("frame", "gaussians", "plain_iters")}
        d["net"] = e["net"]
        d["mean_start"] = e["mean_start"]
        d["speedup_median"] = (1.0 / e["net"]["median_time_ratio"]
                               if e["net"]["median_time_ratio"] else None)
        d["psnr_plain_final"] = [r["plain_final_psnr"] for r in e["rows"]]
        d["warm_iters"] = [r["warm_iters"] for r in e["rows"]]
        out["evals"].append(d)
    return out


def bench_warmstart(args, dev, quick=False):
    """Learned warm start against the plain start: wall-clock (GPU time of the fit steps,
    plus the network forward and decode for the warm start) to reach the plain fit's
    final per-plane PSNR within 0.1 dB, on held-out seeded scenes."""
    from cvqpg_amd import warmstart
    log = lambda m: print(f"[bench] warmstart {m}", file=sys.stderr, flush=True)
    if quick:
        res = warmstart.experiment(dev, n_train=8, fit_iters=400, train_steps=300,
                                   ft_scenes=8, ft_steps=800,
                                   eval_sets=(((1920, 1080), 2, 1000),), log=log,
                                   with_mean=False)
    else:
        res = warmstart.experiment(dev, n_train=args.warm_pairs,
                                   train_steps=args.warm_train_steps,
                                   eval_sets=(((1920, 1080), args.warm_eval, 2000),
                                              ((512, 512), 4, 1000)), log=log)
        if args.warm_out:
            rows = {"network": res["network"],
                    "evals": [{k: v for k, v in e.items()} for e in res["evals"]]}
            with open(args.warm_out, "w") as f:
                json.dump(rows, f, indent=1, default=float)
    return warm_summary(res)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the library "
              "has no multi-GPU path", file=sys.stderr)
        sys.exit(2)
    import torch
    sys.path.insert(0, ROOT)
    import __graft_entry__ as ge
    ge.load_pkg()
    from cvqpg_amd import fused, rasterizer, synthetic, train

    if args.workload == "batch16":
        return bench_batch(args, torch.device("cuda:0"))
    if args.workload == "warmstart":
        dev = torch.device("cuda:0")
        t0 = time.perf_counter()
        res = bench_warmstart(args, dev)
        hd = [e for e in res["evals"] if e["frame"] == [1920, 1080]][0]
        out = {"metric": WARM_METRIC, "value": hd["speedup_median"], "unit": "x",
               "n_gpus": 1, "steps": None, "warmup": None,
               "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": "f32 fit, bf16/f32-accumulate network",
               "data": "synthetic", "config": {"workload": "warmstart", **res,
                                               "wall_s": time.perf_counter() - t0},
               "roofline": None, "cpu_baseline": None}
        print(json.dumps(out), flush=True)
        return

    dev = torch.device("cuda:0")
    spec = synthetic.WORKLOADS[args.workload]
    W, H, N = spec["W"], spec["H"], spec["N"]
    C = len(spec["wavelengths"])
    print(f"[bench] building workload {args.workload}", file=sys.stderr, flush=True)
    wl = train.Workload(args.workload, device=dev, seed=0,
                        renderer="torch" if args.path == "torch" else "hip")
    # the step's peak counts the workload's resident tensors, not the temporaries of its setup
    torch.cuda.reset_peak_memory_stats(dev)
    pH, pW = synthetic.pad_size_for(W, H)
    one_graph_ms = None   # the kernel-timing loop's step time (fused path)
    if args.path in ("autograd", "torch"):
        trainer = train.AutogradTrainer(wl, backend="torch")
        path = ("autograd: HIP raster + torch.fft BLASM + torch loss/Adan" if args.path == "autograd"
                else "torch: PyTorch tile renderer + torch.fft BLASM + torch loss/Adan")
        for _ in range(args.warmup):
            trainer.step()
        torch.cuda.synchronize()
        rasterizer.timer.enabled = True
        rasterizer.timer.reset()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            trainer.step()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        rasterizer.timer.enabled = False
        P = P_ref = rasterizer._capacity_hint.get((N, W, H), 0)   # the reference's lists
        kernel_ms = {"render_fwd": rasterizer.timer.mean_ms("k_render_fwd"),
                     "render_bwd": rasterizer.timer.mean_ms("k_render_bwd")}
        stages = kernel_ms
    else:
        trainer = fused.FusedTrainer(wl, warmup_iters=400, seed=0, fine=True)
        path = ("fused HIP step, hipGraph replay (prologue, bin, sort, raster, BLASM, loss, "
                "adjoints, Adan)")
        trainer.step()
        stages = stage_report(trainer)
        print(f"[bench] eager stage ms: {json.dumps(stages)}", file=sys.stderr, flush=True)
        # (1) every hot kernel's stage captured as a graph of its own, so HIP events on the
        # replay stream bracket each of them (the events between graphs cost ~2 % of the
        # step, so this loop only measures the kernels)
        trainer.capture(split=[h[0] for h in HOT])
        for _ in range(args.warmup):
            trainer.run()
        ev = {}
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            trainer.run(split_events=ev)
        torch.cuda.synchronize()
        split_ms = 1e3 * (time.perf_counter() - t0) / args.steps
        kernel_ms = {k: sum(a.elapsed_time(b) for a, b in v) / len(v) for k, v in ev.items()}
        # (2) the timed steps: the step as three graphs, the longest kernel's stage alone in
        # the middle one, bracketed by HIP events on the replay stream
        dominant = max(kernel_ms, key=kernel_ms.get)
        trainer.capture(split=dominant)
        for _ in range(args.warmup):
            trainer.run()
        ev = {}
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            trainer.run(split_events=ev)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        kernel_ms[dominant] = sum(a.elapsed_time(b) for a, b in ev[dominant]) / len(ev[dominant])
        if not trainer.check_capacity():
            raise RuntimeError("key capacity exceeded during the timed steps")
        P = trainer.last_pairs
        one_graph_ms = split_ms
        # the algorithmic unit counts the reference's tile lists (the step bins live lists)
        P_ref = trainer.reference_keys()
    ms = 1e3 * dt / args.steps
    pairs = P_ref * 256
    roofs = rooflines(kernel_ms, spec, pairs, args.workload)
    roof = max(roofs, key=lambda e: e["avg_ms"]) if roofs else None
    peak_gb = torch.cuda.max_memory_allocated(dev) / 1e9   # the step's own peak
    if args.calib_copy > 0:
        n = args.calib_copy * (1 << 20) // 4
        src = torch.ones(n, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"[bench] calibration copy of {args.calib_copy} MiB done", file=sys.stderr, flush=True)
    warm = None
    if not args.no_warmstart and args.workload == WORKLOAD:
        try:
            warm = bench_warmstart(args, dev, quick=True)
        except Exception as e:  # the step/s line must still be printed
            print(f"[bench] warm-start study failed: {e!r}", file=sys.stderr, flush=True)
            warm = {"error": repr(e)}
    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] CPU baseline sample", file=sys.stderr, flush=True)
        cpu = cpu_baseline(spec)
        if args.cpu_full_step:
            t_full, threads = cpu_full_step(args.workload)
            cpu["sampled_value"] = cpu["value"]
            cpu["value"] = 1.0 / t_full
            cpu["full_step_s"] = t_full
            cpu["sample"] = (f"one whole {W}x{H} step of the reference PyTorch path on the host "
                             f"CPU, FP32, {threads} threads, gradient checkpointing of the tile "
                             f"loop: {t_full:.1f}s (the sampled estimate: "
                             f"{1.0 / cpu['sampled_value']:.1f}s)")
    out = {"metric": METRIC, "value": 1e3 / ms, "unit": "step/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic",
           "config": {"workload": args.workload, "width": W, "height": H, "gaussians": N,
                      "channels": C, "planes": len(synthetic.distances_for(spec)),
                      "fft_grid": [2 * pH, 2 * pW], "keys": P_ref, "keys_live": P,
                      "path": path,
                      "peak_hbm_gb": peak_gb, "stage_ms_eager": stages,
                      "ms_per_step_all_kernels_split": one_graph_ms},
           "roofline": roof, "rooflines": roofs,
           "cpu_baseline": cpu, "time_to_psnr": warm}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
