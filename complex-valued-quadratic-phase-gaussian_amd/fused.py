"""All-HIP training step for one scene, graph-capturable.

One iteration of train_2d_gaussian_rgb.py:187-211 as a fixed sequence of libcvqpg calls
on the current stream:

    prologue      cvqpg_prologue_forward_live  activations, cov, inverse, radii, packing, live count
    binning       cvqpg_bin_scan, cvqpg_bin_sort
    raster        cvqpg_render_forward_packed  -> field [C,H,W]
    optics        cvqpg_blasm_forward          -> plane fields + intensities [D,C,H,W]
    loss          cvqpg_loss_forward_backward_stats  -> dL/dI, loss (target statistics
                  from cvqpg_loss_target_stats, once per fit)
    optics adj.   cvqpg_blasm_backward         -> dL/dfield
    raster adj.   cvqpg_render_backward_tiles + cvqpg_raster_backward_raw -> raw grads
    optimiser     cvqpg_adan_step              clip, Adan, schedule, iteration += 1

Nothing in ``step`` allocates or synchronises, so ``capture`` can record it into a HIP
graph (torch.cuda.CUDAGraph drives hipStreamBeginCapture/EndCapture on ROCm).  The key
buffers are sized once from an eager count with headroom.  A step whose keys do not fit
renders nothing, and its Adan update and iteration advance are skipped on the device
(the raster workspace's overflow flag is the optimiser's skip flag), so a replayed graph
never applies an update computed from an empty field; the workspace also counts such
steps.  ``check_capacity`` reads the key count and that count back (one host read) so a
caller can re-size with ``grow``.
"""
import ctypes
import math

import torch

from . import _lib
from .model import PARAM_NAMES
from .optics import HologramOptics
from .rasterizer import workspace_bytes

TILE = 16
TABLES, TABLES_KEPT = 8, 16   # CVQPG_BLASM_TABLES, CVQPG_BLASM_TABLES_KEPT
ADAN_BETAS = (0.98, 0.92, 0.99)
GROUP_LR = {"means_2d": 0.01, "pre_act_scales": 0.005, "pre_act_rotation": 0.001,
            "colours": 0.0025, "pre_act_phase": 0.0025, "pre_act_opacities": 0.025,
            "curvature": 0.001}


class FusedTrainer:
    def __init__(self, wl, num_itrs=2001, lr=0.01, warmup_iters=400, primitive="curv",
                 seed=0, inject_noise=True, headroom=1.25, fine=False):
        """``fine`` runs the BLASM directions pass by pass (stages blasm_fwd_rows / _cols /
        _out and blasm_bwd_rows / _cols / _in instead of blasm_fwd / blasm_bwd), so a
        split capture can time the column passes on their own."""
        L = _lib.lib()
        self.fine = bool(fine)
        g = wl.gaussians
        self.wl = wl
        self.dev = g.means_2d.device
        dev = self.dev
        self.params = [getattr(g, n).data for n in PARAM_NAMES]
        for p in self.params:
            assert p.is_contiguous() and p.dtype == torch.float32
        self.N = g.means_2d.shape[0]
        self.C = g.channels
        self.W, self.H = wl.img_size
        ap = wl.args_prop
        self.wavelengths = list(ap.wavelengths)
        self.k = (2 * math.pi / torch.tensor(self.wavelengths, dtype=torch.float64)).float().to(dev)
        self.dxy2 = float(ap.pixel_pitch) ** 2
        self.curv_lambda = float(self.wavelengths[0] if self.C == 1 else min(self.wavelengths))
        self.curv_mode = 0 if primitive == "flat" else 2
        self.curv_nyquist = float(getattr(ap, "curv_nyquist", 0.9))
        self.warmup = int(warmup_iters) if primitive != "flat" else 0
        self.num_itrs, self.lr0 = int(num_itrs), float(lr)
        self.seed = int(seed)
        self.T = math.ceil(self.W / TILE) * math.ceil(self.H / TILE)
        self.optics = HologramOptics(wl.img_size, ap.pad_size, self.wavelengths, ap.pixel_pitch,
                                     ap.distances, aperture_size=ap.aperture_size)
        D, C, H, W, N = self.optics.D, self.C, self.H, self.W, self.N
        self.D = D
        # per-step buffers
        self.means = torch.empty(N, 2, device=dev)
        self.radii = torch.empty(N, dtype=torch.int32, device=dev)
        self.num_pairs = torch.zeros(1, dtype=torch.int32, device=dev)
        self.ranges = torch.empty(self.T, 2, dtype=torch.int32, device=dev)
        self.field = torch.empty(C, H, W, dtype=torch.complex64, device=dev)
        self.inten = torch.empty(D, C, H, W, device=dev)
        self.plane_field = torch.empty(D, C, H, W, dtype=torch.complex64, device=dev)
        self.grad_inten = torch.empty(D, C, H, W, device=dev)
        self.grad_field = torch.empty(C, H, W, dtype=torch.complex64, device=dev)
        self.loss = torch.zeros(1, device=dev)
        nl = L.cvqpg_loss_workspace_size(C, D, H, W)
        self.loss_ws = torch.empty(int(nl), dtype=torch.uint8, device=dev)
        self.targets = wl.targets.contiguous().float()
        self.masks = wl.masks.contiguous().float()
        # the targets are fixed for the fit: their SSIM window statistics are filtered once
        nts = int(L.cvqpg_loss_target_stats_size(C, D, H, W))
        self.target_stats = torch.empty(nts // 4, device=dev)
        _lib.check(L.cvqpg_loss_target_stats(_lib.ptr(self.targets), C, D, H, W,
                                             _lib.ptr(self.target_stats), nts,
                                             _lib.stream_handle()))
        if inject_noise:
            rng = (self.targets.flatten(1).max(1).values - self.targets.flatten(1).min(1).values)
            self.noise_scale = [1e-3 * float(v) for v in rng.cpu()]
        else:
            self.noise_scale = None
        self.sizes = [p.numel() for p in self.params]
        self.total = sum(self.sizes)
        self.grads = torch.empty(self.total, device=dev)
        self.n_partials = math.ceil(N / 256)
        self.sumsq = torch.empty(self.n_partials, device=dev)
        self.state = torch.zeros(4 * self.total, device=dev)
        self.iteration = torch.zeros(1, dtype=torch.int32, device=dev)
        self.group_lr = [GROUP_LR[n] for n in PARAM_NAMES]
        self.graph = None
        self.capacity = 0
        # the BLASM transform tables of these sizes, once per fit (the step keeps them)
        ows = self.optics._workspace(dev)
        _lib.check(L.cvqpg_blasm_forward_passes(
            TABLES, None, *self.optics._common(), _lib.ptr(ows), ows.numel(), None, None,
            _lib.stream_handle(dev)))
        self._size_keys(headroom)

    # ------------------------------------------------------------------ sizing
    def _alloc_raster(self, capacity):
        L = _lib.lib()
        self.capacity = int(capacity)
        nb = workspace_bytes(self.N, self.C, self.W, self.H, self.capacity)
        self.raster_ws = torch.empty(nb, dtype=torch.uint8, device=self.dev)
        self.point_list = torch.empty(max(self.capacity, 1), dtype=torch.int32, device=self.dev)
        flags = ctypes.c_void_p()
        _lib.check(L.cvqpg_raster_overflow_flags(_lib.ptr(self.raster_ws), nb, self.N,
                                                 self.capacity, ctypes.byref(flags)))
        self.overflow_flags = flags
        off = flags.value - self.raster_ws.data_ptr()
        # [key count, overflow of the last count, overflowed counts since the clear]
        self._misc = self.raster_ws[off - 4:off + 8].view(torch.int32)
        _lib.check(L.cvqpg_raster_clear_overflow(_lib.ptr(self.raster_ws), nb, self.N,
                                                 self.capacity, self._s()))
        self.overflowed_steps = 0

    def _size_keys(self, headroom):
        """Eager prologue + count with a provisional capacity, then fix the capacity."""
        self._alloc_raster(max(1024, 16 * self.N))
        self._prologue()
        self._count()
        P = int(self.num_pairs.item())
        self.last_pairs = P
        # sized on the reference's key count, which bounds every live count of the same
        # parameters: while a fit sharpens its Gaussians the live lists grow much faster
        # than the rectangles (+17 % against +4 % over 400 steps of a 256 x 256 fit)
        self._alloc_raster(max(1024, int(max(P, self.reference_keys()) * headroom)))

    def reference_keys(self):
        """Number of keys of the reference's tile lists (every tile of each 5-sigma
        rectangle, cuda/forward.cu:15-26) for the Gaussians of the last prologue, counted
        with cvqpg_bin_count into a scratch workspace (the step's own buffers and overflow
        flags are untouched).  One host read."""
        L = _lib.lib()
        nb = workspace_bytes(self.N, self.C, self.W, self.H, 1)
        ws = torch.empty(nb, dtype=torch.uint8, device=self.dev)
        out = torch.zeros(1, dtype=torch.int32, device=self.dev)
        _lib.check(L.cvqpg_bin_count(_lib.ptr(self.means), _lib.ptr(self.radii), self.N, self.W,
                                     self.H, _lib.ptr(ws), nb, 1, _lib.ptr(out), self._s()))
        return int(out.item())

    def check_capacity(self):
        """One host read of the last key count and of the number of overflowed steps since
        the buffers were sized; True if every step fitted (``overflowed_steps`` holds the
        count: those steps changed neither the parameters nor the iteration)."""
        P, _, over = (int(v) for v in self._misc.cpu())
        self.last_pairs = P
        self.overflowed_steps = over
        return P <= self.capacity and over == 0

    def grow(self, headroom=1.25):
        """Re-size the key buffers for the current parameters (optimiser state and the
        iteration counter are kept) and re-capture the graph if one was captured."""
        split = getattr(self, "_split", None)
        had_graph = self.graph is not None
        self.graph = None
        self._size_keys(headroom)
        if had_graph:
            self.capture(split=split)

    # ------------------------------------------------------------------ stages
    def _s(self):
        return _lib.stream_handle(self.dev)

    def _ws(self):
        return _lib.ptr(self.raster_ws), self.raster_ws.numel(), self.capacity

    def _prologue(self):
        # the prologue fused with the live count (the raster kernels' own tile lists: keys
        # below their skip threshold left out; results those of the full lists bit for bit)
        ws, nb, cap = self._ws()
        _lib.check(_lib.lib().cvqpg_prologue_forward_live(
            _lib.ptr_array(self.params), self.N, self.C, self.W, self.H, _lib.ptr(self.k),
            self.dxy2, self.curv_lambda, self.curv_mode, self.curv_nyquist, ws, nb, cap,
            _lib.ptr(self.means), _lib.ptr(self.radii), self._s()))

    def _count(self):
        # the scan of the counts the prologue wrote
        ws, nb, cap = self._ws()
        _lib.check(_lib.lib().cvqpg_bin_scan(self.N, ws, nb, cap, _lib.ptr(self.num_pairs),
                                             self._s()))

    def _stages(self):
        L = _lib.lib()
        S = self._s  # resolved at launch time: the capture stream during capture
        ws, nb, cap = self._ws()
        N, C, W, H = self.N, self.C, self.W, self.H
        o = self.optics
        ows = o._workspace(self.dev)
        ns = None if self.noise_scale is None else _lib.floats(self.noise_scale)
        b1, b2, b3 = ADAN_BETAS
        fr = lambda t: _lib.ptr(torch.view_as_real(t))
        stages = [
            ("prologue", self._prologue),
            ("bin_count", self._count),
            ("bin_sort", lambda: _lib.check(L.cvqpg_bin_sort(
                _lib.ptr(self.means), _lib.ptr(self.radii), N, W, H, ws, nb, cap,
                _lib.ptr(self.point_list), _lib.ptr(self.ranges), S()))),
            ("render_fwd", lambda: _lib.check(L.cvqpg_render_forward_packed(
                N, C, W, H, ws, nb, cap, _lib.ptr(self.point_list), _lib.ptr(self.ranges),
                fr(self.field), S()))),
        ]
        # the transform tables were built once in __init__ (TABLES_KEPT: no launches for
        # them in the step)
        blasm_fwd = lambda m: lambda: _lib.check(L.cvqpg_blasm_forward_passes(
            m | TABLES_KEPT, fr(self.field), *o._common(), _lib.ptr(ows), ows.numel(), fr(self.plane_field),
            _lib.ptr(self.inten), S()))
        blasm_bwd = lambda m: lambda: _lib.check(L.cvqpg_blasm_backward_passes(
            m | TABLES_KEPT, fr(self.field), _lib.ptr(self.grad_inten), fr(self.plane_field), *o._common(),
            _lib.ptr(ows), ows.numel(), fr(self.grad_field), S()))
        if self.fine:
            stages += [("blasm_fwd_rows", blasm_fwd(1)), ("blasm_fwd_cols", blasm_fwd(2)),
                       ("blasm_fwd_out", blasm_fwd(4))]
        else:
            stages.append(("blasm_fwd", blasm_fwd(7)))
        stages += [
            ("loss", lambda: _lib.check(L.cvqpg_loss_forward_backward_stats(
                _lib.ptr(self.inten), _lib.ptr(self.targets), _lib.ptr(self.target_stats),
                _lib.ptr(self.masks), None, ns,
                self.seed, _lib.ptr(self.iteration), C, self.D, H, W, _lib.ptr(self.loss_ws),
                self.loss_ws.numel(), _lib.ptr(self.grad_inten), _lib.ptr(self.loss), S()))),
        ]
        if self.fine:
            stages += [("blasm_bwd_rows", blasm_bwd(1)), ("blasm_bwd_cols", blasm_bwd(2)),
                       ("blasm_bwd_in", blasm_bwd(4))]
        else:
            stages.append(("blasm_bwd", blasm_bwd(7)))
        stages += [
            ("render_bwd", lambda: _lib.check(L.cvqpg_render_backward_tiles(
                fr(self.grad_field), _lib.ptr(self.k), N, C, W, H, ws, nb, cap,
                _lib.ptr(self.point_list), _lib.ptr(self.ranges), S()))),
            ("reduce", lambda: _lib.check(L.cvqpg_raster_backward_raw(
                _lib.ptr_array(self.params), N, C, W, H, self.dxy2, self.curv_lambda,
                self.curv_mode, self.curv_nyquist, _lib.ptr(self.iteration), self.warmup, ws, nb,
                cap, _lib.ptr(self.grads), _lib.ptr(self.sumsq), S()))),
            ("adan", lambda: _lib.check(L.cvqpg_adan_step(
                _lib.ptr_array(self.params), _lib.int64s(self.sizes), _lib.ptr(self.grads),
                _lib.ptr(self.state), _lib.ptr(self.sumsq), self.n_partials,
                _lib.floats(self.group_lr), self.lr0, 1e-3, self.num_itrs, b1, b2, b3, 1e-8,
                1.0, self.warmup, 1 if self.curv_mode != 0 else 0, _lib.ptr(self.iteration),
                self.overflow_flags, S()))),
        ]
        return stages

    STAGES = ("prologue", "bin_count", "bin_sort", "render_fwd", "blasm_fwd", "loss",
              "blasm_bwd", "render_bwd", "reduce", "adan")

    def step(self):
        for _, fn in self._stages():
            fn()

    def loss_and_grads(self):
        """The step without the optimiser: loss into ``self.loss`` and the raw gradients
        into ``self.grads`` (eager, stream-ordered, no host synchronisation)."""
        for name, fn in self._stages():
            if name != "adan":
                fn()

    def step_timed(self, events):
        """Eager step with a HIP event after every stage (events: dict name -> list)."""
        e0 = torch.cuda.Event(enable_timing=True)
        e0.record()
        prev = e0
        for name, fn in self._stages():
            fn()
            e = torch.cuda.Event(enable_timing=True)
            e.record()
            events.setdefault(name, []).append((prev, e))
            prev = e

    # ------------------------------------------------------------------ graphs
    def capture(self, split=None):
        """Record one step into HIP graphs.  ``split`` (a stage name or a list of them)
        records each named stage as a graph of its own, between graphs of the stages in
        between, so ``run`` can bracket the named stages with events on the replay
        stream."""
        torch.cuda.synchronize(self.dev)
        names_split = [] if split is None else ([split] if isinstance(split, str) else list(split))
        self._split = split
        stages = self._stages()
        groups, cur = [], []
        for name, fn in stages:
            if name in names_split:
                if cur:
                    groups.append((None, cur))
                groups.append((name, [(name, fn)]))
                cur = []
            else:
                cur.append((name, fn))
        if cur:
            groups.append((None, cur))
        missing = set(names_split) - {n for n, _ in groups}
        if missing:
            raise ValueError(f"unknown stages {sorted(missing)}")
        graphs = []
        side = torch.cuda.Stream(self.dev)
        for name, grp in groups:
            g = torch.cuda.CUDAGraph()
            with torch.cuda.stream(side):
                with torch.cuda.graph(g, stream=side):
                    for _, fn in grp:
                        fn()
            graphs.append((name, g))
        torch.cuda.synchronize(self.dev)
        self.graphs = graphs
        self.graph = graphs
        return graphs

    def run(self, split_events=None):
        """One step: graph replay if captured, else eager.  With a split capture and
        ``split_events`` (a dict), an event pair bracketing each named stage's graph is
        appended to ``split_events[name]``."""
        if self.graph is None:
            self.step()
            return
        for name, g in self.graphs:
            if name is not None and split_events is not None:
                a = torch.cuda.Event(enable_timing=True)
                b = torch.cuda.Event(enable_timing=True)
                a.record()
                g.replay()
                b.record()
                split_events.setdefault(name, []).append((a, b))
            else:
                g.replay()

    def grad_views(self):
        """Raw gradients as tensors shaped like the parameters."""
        out, off = {}, 0
        for n, p in zip(PARAM_NAMES, self.params):
            out[n] = self.grads[off:off + p.numel()].view_as(p)
            off += p.numel()
        return out
