"""Training-step throughput of the CVQPG hot path on one MI355X.

Workload (the metric's configuration): RGB 1920x1080, 150,000 Gaussians, FP32 /
complex64 field, 3 wavelengths (639/515/473 nm), 8 um pitch, 2x padded BLASM
(2816 x 4480 FFT grid) to 4 planes at 0, 2, 4, 6 mm; one step = rasterize -> pad ->
glue -> BLASM -> multi-plane L2 + SSIM loss -> full adjoint -> Adan.

Prints one JSON line (rank 0).  ``time_to_psnr`` is a reduced learned-warm-start study
(8 training pairs from 400-iteration 1080p fits, 800 render-loss fine-tuning steps, 2
held-out 1080p scenes; ``--workload warmstart`` runs the full one):
speed-up of the warm start to the plain fit's final per-plane PSNR.  ``roofline_blasm`` gives
the achieved HBM rate of the FFT passes.  ``roofline`` describes the dominant kernel, timed
with HIP events on its launch stream inside the timed region.  ``cpu_baseline``
times the CPU oracle on a bounded sample of the same workload (see
``cpu_baseline``).
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
    ap.add_argument("--no-warmstart", action="store_true",
                    help="skip the reduced warm-start time-to-PSNR study of the default run")
    ap.add_argument("--workload", default=WORKLOAD,
                    help="rgb1080p_150k (the metric's workload), gray256_10k, rgb4k_600k, "
                         "batch16 (16 uneven scenes, batched vs sequential) or warmstart "
                         "(full learned warm-start time-to-PSNR study)")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="run one device-to-device copy of MIB MiB after the timed steps (a "
                         "plain copy kernel of known size for scaling HBM byte counters)")
    return ap.parse_args()


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "render_bwd_traffic_rgb1080p_150k.txt")


BLASM_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "blasm_traffic_rgb1080p_150k.txt")


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


def blasm_roofline(stages, workload):
    """HBM roofline of the BLASM passes (forward f1-f3 + adjoint a3-a1) from the eager
    per-stage HIP-event times: the measured bytes of the committed rocprofv3 summary and
    the algorithmic model of the unpruned schedule (30 two-dimensional FFTs of the 2x grid
    at 32 B per point), both per step."""
    vals = read_kv(BLASM_TRAFFIC_FILE)
    if workload != WORKLOAD or "blasm_bytes_per_step" not in vals or not stages:
        return None
    ms = stages.get("blasm_fwd", 0.0) + stages.get("blasm_bwd", 0.0)
    if ms <= 0:
        return None
    meas = float(vals["blasm_bytes_per_step"])
    model = float(vals["blasm_model_bytes_per_step"])
    # achieved: the bytes the kernels move (rocprofv3) over their time; model_gbs: the
    # unpruned schedule's algorithmic bytes over the same time (above the HBM peak once the
    # pruning removes more than the time the passes lose to latency)
    return {"kernels": "k_f1/f2/f3 + k_a3/a2/a1 (BLASM forward and adjoint)", "bound": "hbm",
            "achieved": meas / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": meas / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": meas, "ms": ms,
            "model_bytes": model, "model_gbs": model / (ms * 1e-3) / 1e9,
            "traffic_source": os.path.relpath(BLASM_TRAFFIC_FILE, ROOT)}


def measured_traffic():
    """HBM bytes per k_render_bwd launch on the default workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (counter values scaled by the same
    counters' reading on a plain copy kernel of known size); None if absent."""
    try:
        return float(read_kv(TRAFFIC_FILE)["hbm_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def raster_flops_per_pair(C):
    """Algorithmic FP32 operations per (pixel, Gaussian) pair of the reference math
    (cuda/forward.cu:131-151, cuda/backward.cu:101-148), counting each add, mul,
    compare and transcendental as one operation and an FMA as two."""
    fwd = 17 + 9 * C          # offsets 2, mahal 9, clamp 2, exp 1, alpha 1, base 1 ; per channel 9
    bwd = 46 + 19 * C
    return fwd, bwd


def cpu_baseline(spec, seed=0):
    """CPU oracle on a bounded sample of the workload, scaled to step/s.

    Sample: (a) the reference's PyTorch tile renderer (3 sigma, 64-px tiles,
    oracle.model.render_tile_torch) forward + autograd backward in FP32 on a
    512 x 512 window holding the Gaussians whose centres fall inside it, scaled by
    the frame/window area ratio; (b) one forward + one inverse complex64 2D FFT of
    the full 2816 x 4480 grid (torch.fft / pocketfft), scaled by the 4 D C FFTs of
    the reference step (forward and autograd backward).  Elementwise work of the
    step is not counted, so the baseline is optimistic for the CPU."""
    import torch

    sys.path.insert(0, ROOT)
    from oracle import model as omodel
    import __graft_entry__ as ge
    pkg = ge.load_pkg()
    from cvqpg_amd import synthetic as syn
    W, H, N = spec["W"], spec["H"], spec["N"]
    C = len(spec["wavelengths"])
    raw = syn.make_scene(N, W, H, spec["wavelengths"], spec["pitch"], spec["kappa"], seed)
    win = 512
    tr = {k: torch.as_tensor(v) for k, v in raw.items()}
    mt = torch.tanh(tr["means_2d"])
    cx, cy = (mt[:, 0] + 1) * 0.5 * W, (mt[:, 1] + 1) * 0.5 * H
    sel = (cx < win) & (cy < win)
    sub = {k: v[sel].clone().requires_grad_() for k, v in tr.items()}
    threads = torch.get_num_threads()
    t0 = time.perf_counter()
    ri = omodel.raster_inputs(sub, (W, H), spec["wavelengths"], spec["pitch"])
    out = omodel.render_tile_torch(ri, win, win)
    (out.abs() ** 2).sum().backward()
    t_r = time.perf_counter() - t0
    pH, pW = syn.pad_size_for(W, H)
    u = torch.randn(2 * pH, 2 * pW, dtype=torch.complex64)
    t0 = time.perf_counter()
    torch.fft.ifft2(torch.fft.fft2(u))
    t_fft = time.perf_counter() - t0
    D = len(syn.distances_for(spec))
    t_step = t_r * (W * H) / (win * win) + t_fft * 2 * D * C
    return {"value": 1.0 / t_step, "unit": "step/s", "cores": threads, "kind": "port",
            "sample": (f"oracle PyTorch tile renderer fwd+bwd FP32 on a {win}x{win} window "
                       f"({int(sel.sum())} Gaussians) scaled x{W * H / win / win:.1f}; one "
                       f"fft2+ifft2 pair of {2 * pH}x{2 * pW} scaled x{2 * D * C}; "
                       f"elementwise work not counted; {t_r:.1f}s + {t_fft:.1f}s measured")}


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
    pairs = sum(t.last_pairs for t in bt.trainers) * 256
    f_fwd, f_bwd = raster_flops_per_pair(3)
    achieved = pairs * f_bwd / (k["render_bwd"] * 1e-3) / 1e12
    roof = {"kernel": "k_render_bwd_batch (persistent tile-queue raster adjoint, all scenes)",
            "bound": "mfma", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None, "avg_ms": k["render_bwd"],
            "pairs": pairs, "flop_per_pair": f_bwd, "fwd_avg_ms": k["render_fwd"]}
    n = len(bt.trainers)
    out = {"metric": BATCH_METRIC, "value": n * 1e3 / ms, "unit": "scene-step/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic",
           "config": {"workload": "batch16_uneven_rgb", "scenes": [
                          [t.W, t.H, t.N] for t in bt.trainers],
                      "planes": 3, "keys": sum(t.last_pairs for t in bt.trainers),
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
        res = warmstart.experiment(dev, log=log)
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
    f_fwd, f_bwd = raster_flops_per_pair(C)
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
        P = rasterizer._capacity_hint.get((N, W, H), 0)
        k_ms = rasterizer.timer.mean_ms("k_render_bwd")
        stages = {"k_render_fwd": rasterizer.timer.mean_ms("k_render_fwd"), "k_render_bwd": k_ms}
    else:
        trainer = fused.FusedTrainer(wl, warmup_iters=400, seed=0)
        path = "fused HIP step, hipGraph replay (prologue, bin, sort, raster, BLASM, loss, adjoints, Adan)"
        trainer.step()
        stages = stage_report(trainer)
        print(f"[bench] eager stage ms: {json.dumps(stages)}", file=sys.stderr, flush=True)
        trainer.capture(split="render_bwd")
        for _ in range(args.warmup):
            trainer.run()
        torch.cuda.synchronize()
        ev = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            trainer.run(split_events=ev)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        k_ms = sum(a.elapsed_time(b) for a, b in ev) / len(ev)
        if not trainer.check_capacity():
            raise RuntimeError("key capacity exceeded during the timed steps")
        P = trainer.last_pairs
    ms = 1e3 * dt / args.steps
    pairs = P * 256
    traffic = measured_traffic() if args.workload == WORKLOAD else None
    roof = None   # the PyTorch path has no HIP raster kernel to report
    if k_ms:
        achieved = pairs * f_bwd / (k_ms * 1e-3) / 1e12
        roof = {"kernel": "k_render_bwd (raster adjoint tile kernel)",
                "bound": "mfma",
                "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP32_PEAK_TFLOPS, "traffic": traffic, "avg_ms": k_ms,
                "pairs": pairs, "flop_per_pair": f_bwd}
        if traffic is not None:
            roof["traffic_source"] = os.path.relpath(TRAFFIC_FILE, ROOT)
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
    out = {"metric": METRIC, "value": 1e3 / ms, "unit": "step/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic",
           "config": {"workload": args.workload, "width": W, "height": H, "gaussians": N,
                      "channels": C, "planes": len(synthetic.distances_for(spec)),
                      "fft_grid": [2 * pH, 2 * pW], "keys": P, "path": path,
                      "peak_hbm_gb": peak_gb, "stage_ms_eager": stages},
           "roofline": roof, "roofline_blasm": blasm_roofline(stages, args.workload),
           "cpu_baseline": cpu, "time_to_psnr": warm}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
