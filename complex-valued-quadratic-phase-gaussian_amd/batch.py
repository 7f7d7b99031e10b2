"""Many independent hologram fits advanced by one captured step.

The reference fits one scene per process (train_2d_gaussian_rgb.py:187-211, one
``qpf_render`` per iteration).  ``BatchTrainer`` advances S scenes of uneven resolution
and Gaussian count together; one step is

    per scene   prologue, bin_count, bin_sort           (fork over side streams)
    all scenes  cvqpg_render_forward_batch               (one persistent launch)
    per scene   blasm_fwd, loss, blasm_bwd               (fork over side streams)
    all scenes  cvqpg_render_backward_batch              (one persistent launch)
    per scene   reduce, adan                             (fork over side streams)

and the whole step is recorded into ONE HIP graph.  The raster, the expensive part,
runs from a persistent grid that drains a queue holding every scene's 16x16 tiles
(heaviest scene first), so small scenes fill the tail of large ones instead of leaving
the GPU idle between launches.  The per-scene stages keep their single-scene kernels
(their transform plans are compiled per frame size) and overlap across default-priority
side streams joined back into the capture stream.

Every scene's arithmetic is exactly that of its own ``FusedTrainer`` step (same
kernels, same per-tile order), so a batched step is bit-identical to running the
scenes one after another.

``autotune`` picks the persistent grid and the queue chunk (tiles per fetch) of the two
batched raster kernels by timing them in place, and the number of side streams by
timing captured steps (parameters and optimiser state are restored afterwards).
"""
import ctypes
import time

import torch

from . import _lib
from .fused import FusedTrainer

PHASES = (("scene", ("prologue", "bin_count", "bin_sort")),
          ("batch", "render_fwd"),
          ("scene", ("blasm_fwd", "loss", "blasm_bwd")),
          ("batch", "render_bwd"),
          ("scene", ("reduce", "adan")))


class BatchTrainer:
    def __init__(self, workloads, streams=4, grid=None, chunk=1, trainers=None, **trainer_kw):
        self.trainers = trainers if trainers is not None else [
            FusedTrainer(wl, **trainer_kw) for wl in workloads]
        t0 = self.trainers[0]
        self.dev = t0.dev
        self.C = t0.C
        for t in self.trainers:
            if t.C != self.C or t.dev != self.dev:
                raise ValueError("every scene of a batch needs the same channel count and device")
        self.n_cu = torch.cuda.get_device_properties(self.dev).multi_processor_count
        self.grid = {"render_fwd": grid or 8 * self.n_cu, "render_bwd": grid or 5 * self.n_cu}
        self.chunk = {"render_fwd": chunk, "render_bwd": chunk}
        self.graph = None
        self._stages = [dict(t._stages()) for t in self.trainers]
        self._build_table()
        self.set_streams(streams)

    # ------------------------------------------------------------------ set-up
    def cost(self, i):
        """Relative cost of scene i: its last key count (pixel-pairs / 256)."""
        return max(1, int(self.trainers[i].last_pairs))

    def _build_table(self):
        """Device table of the batched raster: scenes in descending cost, so the queue
        serves the heaviest scene's tiles first."""
        L = _lib.lib()
        self.order = sorted(range(len(self.trainers)), key=lambda i: -self.cost(i))
        scenes = (_lib.RasterScene * len(self.order))()
        for j, i in enumerate(self.order):
            t = self.trainers[i]
            sc = scenes[j]
            sc.n, sc.width, sc.height, sc.channels = t.N, t.W, t.H, t.C
            sc.ws, sc.ws_bytes, sc.capacity = t.raster_ws.data_ptr(), t.raster_ws.numel(), t.capacity
            sc.point_list, sc.ranges = t.point_list.data_ptr(), t.ranges.data_ptr()
            sc.field = torch.view_as_real(t.field).data_ptr()
            sc.grad_field = torch.view_as_real(t.grad_field).data_ptr()
            sc.k = t.k.data_ptr()
        nb = int(L.cvqpg_raster_batch_table_size(len(self.order)))
        host = (ctypes.c_uint8 * nb)()
        _lib.check(L.cvqpg_raster_batch_table(scenes, len(self.order), host, nb))
        img = torch.frombuffer(bytearray(host), dtype=torch.uint8)
        self.table = img.to(self.dev)
        torch.cuda.synchronize(self.dev)
        self.tiles = sum(t.T for t in self.trainers)

    def set_streams(self, n):
        """Spread the scenes over n default-priority side streams, longest-processing-time
        first; n = 0 runs every per-scene stage on the caller's stream."""
        self.n_streams = int(n)
        self.streams = [torch.cuda.Stream(self.dev) for _ in range(self.n_streams)]
        load = [0] * max(1, self.n_streams)
        self.assign = {}
        for i in self.order:
            s = min(range(len(load)), key=lambda k: load[k])
            self.assign[i] = s
            load[s] += self.cost(i)
        self.graph = None

    # ------------------------------------------------------------------ step
    def _batched(self, name):
        L = _lib.lib()
        fn = L.cvqpg_render_forward_batch if name == "render_fwd" else L.cvqpg_render_backward_batch
        _lib.check(fn(_lib.ptr(self.table), self.C, self.grid[name], self.chunk[name],
                      _lib.stream_handle(self.dev)))

    def _per_scene(self, names):
        if self.n_streams == 0:
            for i in self.order:
                for n in names:
                    self._stages[i][n]()
            return
        main = torch.cuda.current_stream(self.dev)
        for s in self.streams:
            s.wait_stream(main)
        for i in self.order:
            side = self.streams[self.assign[i]]
            with torch.cuda.stream(side):
                for n in names:
                    self._stages[i][n]()
                    # every per-scene stage enqueues on the current stream only; a stage that
                    # forks onto streams of its own inside these side streams is not supported
                    # (the capture of such a nesting has crashed in hipStreamEndCapture)
                    if torch.cuda.current_stream(self.dev) != side:
                        raise RuntimeError(f"stage {n!r} of scene {i} changed the current "
                                           "stream; per-scene stages must not fork streams")
        for s in self.streams:
            main.wait_stream(s)

    def step(self):
        """One step of every scene, stream-ordered on the current stream (eager)."""
        for kind, what in PHASES:
            if kind == "scene":
                self._per_scene(what)
            else:
                self._batched(what)

    def capture(self):
        """Record one batched step into a single HIP graph."""
        torch.cuda.synchronize(self.dev)
        g = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream(self.dev)
        with torch.cuda.stream(side):
            with torch.cuda.graph(g, stream=side):
                self.step()
        torch.cuda.synchronize(self.dev)
        self.graph = g
        return g

    def run(self):
        if self.graph is None:
            self.step()
        else:
            self.graph.replay()

    # ------------------------------------------------------------------ capacity
    def check_capacity(self):
        """One host read per scene; True if every scene's keys still fit."""
        return all(t.check_capacity() for t in self.trainers)

    def grow(self, headroom=1.25):
        """Re-size the key buffers of every scene that overflowed, rebuild the table and
        drop the captured graph (capture again before the next ``run``)."""
        for t in self.trainers:
            if not t.check_capacity():
                t.grow(headroom)
        self._stages = [dict(t._stages()) for t in self.trainers]
        self._build_table()
        self.set_streams(self.n_streams)

    # ------------------------------------------------------------------ sequential baseline
    def capture_sequential(self):
        """One HIP graph per scene (each scene's own fused step)."""
        for t in self.trainers:
            t.capture()

    def run_sequential(self):
        for i in self.order:
            self.trainers[i].run()

    # ------------------------------------------------------------------ autotuning
    def _time(self, fn, reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        fn()
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        b.synchronize()
        return a.elapsed_time(b) / reps

    def snapshot(self):
        return [([p.clone() for p in t.params], t.state.clone(), t.iteration.clone())
                for t in self.trainers]

    def restore(self, snap):
        for t, (ps, st, it) in zip(self.trainers, snap):
            for p, q in zip(t.params, ps):
                p.copy_(q)
            t.state.copy_(st)
            t.iteration.copy_(it)
        torch.cuda.synchronize(self.dev)

    def autotune(self, grids=(2, 4, 5, 8, 16), chunks=(1, 2, 4), streams=(0, 2, 4, 8),
                 reps=5, log=None):
        """Choose grid (per-CU multiples), chunk and stream count by measurement.
        Returns {name: (ms, setting)} of the choices."""
        snap = self.snapshot()
        self.graph = None
        self.step()  # keys, fields and pixel gradients of the current parameters
        torch.cuda.synchronize(self.dev)
        chosen = {}
        for name in ("render_fwd", "render_bwd"):
            best = None
            for m in grids:
                for c in chunks:
                    self.grid[name], self.chunk[name] = m * self.n_cu, c
                    ms = self._time(lambda: self._batched(name), reps)
                    if log:
                        log(f"{name} grid {m}x{self.n_cu} chunk {c}: {ms:.3f} ms")
                    if best is None or ms < best[0]:
                        best = (ms, (m * self.n_cu, c))
            self.grid[name], self.chunk[name] = best[1]
            chosen[name] = best
        best = None
        for n in streams:
            self.set_streams(n)
            self.capture()
            ms = self._time(self.run, reps)
            if log:
                log(f"streams {n}: {ms:.3f} ms per batched step")
            if best is None or ms < best[0]:
                best = (ms, n)
        self.set_streams(best[1])
        chosen["streams"] = best
        self.restore(snap)
        return chosen

    def losses(self):
        return torch.cat([t.loss for t in self.trainers])


def make_batch(specs, device="cuda", renderer="hip", **kw):
    """BatchTrainer over synthetic scenes built from ``synthetic.batch_specs``."""
    from .train import Workload
    wls = [Workload(f"scene{i}", device=device, seed=s.get("seed", i), renderer=renderer, spec=s)
           for i, s in enumerate(specs)]
    return BatchTrainer(wls, **kw)


def timed_steps(fn, steps, device):
    """Wall-clock ms per call of fn over ``steps`` calls, synchronised on both sides."""
    torch.cuda.synchronize(device)
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize(device)
    return (time.perf_counter() - t0) * 1e3 / steps
