"""Batched fits of independent holograms of uneven size (one captured step).

The batched step runs every scene's own kernels, except that the raster forward and
adjoint of all scenes come from one persistent launch draining a shared tile queue.
Each tile is computed by the same per-tile code in the same order as in the scene's own
launch, so the bar is bit-identity with the scenes' single-scene fused steps (whose
parity with the FP64 oracle is tested in test_gpu_step.py).
"""
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda"
RGB = [639e-9, 515e-9, 473e-9]
SMALL = [dict(W=64, H=48, N=300, wavelengths=RGB, pitch=8e-6, distances=[0.0, 3e-3, 6e-3],
              kappa=0.05, seed=0),
         dict(W=96, H=80, N=700, wavelengths=RGB, pitch=8e-6, distances=[1e-3, 2e-3, 5e-3],
              kappa=0.02, seed=1),
         dict(W=48, H=48, N=200, wavelengths=RGB, pitch=8e-6, distances=[0.5e-3, 4e-3, 4.5e-3],
              kappa=0.02, seed=2)]


def scenes(specs):
    from cvqpg_amd import train
    return [train.Workload(f"s{i}", device=DEV, seed=s["seed"], spec=s) for i, s in enumerate(specs)]


def trainers(specs):
    from cvqpg_amd import fused
    return [fused.FusedTrainer(wl, num_itrs=100, warmup_iters=2) for wl in scenes(specs)]


def same(a, b):
    return bool(torch.equal(a, b))


def test_table_order_and_tiles(pkg):
    from cvqpg_amd.batch import BatchTrainer
    bt = BatchTrainer(None, trainers=trainers(SMALL), streams=2)
    assert bt.tiles == sum(t.T for t in bt.trainers)
    costs = [bt.cost(i) for i in bt.order]
    assert costs == sorted(costs, reverse=True)
    hdr = bt.table[:64].cpu().view(torch.int32)
    assert int(hdr[4]) == len(SMALL) and int(hdr[5]) == bt.tiles
    assert int(hdr[0]) == 0 and int(hdr[1]) == 0


@pytest.mark.parametrize("streams", [0, 2])
def test_batched_steps_equal_single_scene_steps(pkg, streams):
    from cvqpg_amd.batch import BatchTrainer
    ref = trainers(SMALL)
    bt = BatchTrainer(None, trainers=trainers(SMALL), streams=streams, chunk=2)
    for it in range(5):
        if it == 2:
            bt.capture()
            for t in ref:
                t.capture()
        bt.run()
        for t in ref:
            t.run()
        torch.cuda.synchronize()
        for i, (a, b) in enumerate(zip(bt.trainers, ref)):
            assert same(a.loss, b.loss), (it, i, float(a.loss), float(b.loss))
            assert same(a.field, b.field), (it, i)
            for p, q in zip(a.params, b.params):
                assert same(p, q), (it, i)
        print(f"step {it}: losses " + " ".join(f"{float(t.loss):.6f}" for t in bt.trainers),
              flush=True)


@pytest.mark.parametrize("grid,chunk", [(1, 1), (7, 3), (4096, 1)])
def test_queue_shapes_and_counter_reset(pkg, grid, chunk):
    """Any grid (fewer or more workgroups than tiles) and chunk gives the same fields and
    gradients; consecutive launches reuse the queue counters."""
    from cvqpg_amd.batch import BatchTrainer
    bt = BatchTrainer(None, trainers=trainers(SMALL), streams=0)
    snap = bt.snapshot()
    bt.step()
    torch.cuda.synchronize()
    fields = [t.field.clone() for t in bt.trainers]
    grads = [t.grads.clone() for t in bt.trainers]
    for rep in range(2):
        bt.restore(snap)
        # the step up to the optimiser again, from the restored parameters
        for i in bt.order:
            for n in ("prologue", "bin_count", "bin_sort"):
                bt._stages[i][n]()
        bt.grid["render_fwd"], bt.chunk["render_fwd"] = grid, chunk
        bt.grid["render_bwd"], bt.chunk["render_bwd"] = grid, chunk
        bt._batched("render_fwd")
        for i in bt.order:
            for n in ("blasm_fwd", "loss", "blasm_bwd"):
                bt._stages[i][n]()
        bt._batched("render_bwd")
        for i in bt.order:
            bt._stages[i]["reduce"]()
        torch.cuda.synchronize()
        hdr = bt.table[:16].cpu().view(torch.int32)
        assert int(hdr[0]) == 0 and int(hdr[1]) == 0 and int(hdr[2]) == 0 and int(hdr[3]) == 0
        for t, f, g in zip(bt.trainers, fields, grads):
            assert same(t.field, f)
            # the restored iteration counter replays the same step: same noise, same grads
            assert same(t.grads, g)


def test_full_size_scenes_bit_identical(pkg):
    """Two scenes at their batched sizes (512^2 / 20k and 1024x768 / 55k Gaussians)."""
    from cvqpg_amd import synthetic
    from cvqpg_amd.batch import BatchTrainer
    specs = [s for s in synthetic.batch_specs(16, seed=0)][:2]
    assert {(s["W"], s["H"]) for s in specs} == {(512, 512), (1024, 768)}
    ref = trainers(specs)
    bt = BatchTrainer(None, trainers=trainers(specs), streams=2)
    for it in range(2):
        bt.run()
        for t in ref:
            t.run()
    torch.cuda.synchronize()
    for a, b in zip(bt.trainers, ref):
        print(f"{a.W}x{a.H} N={a.N}: loss {float(a.loss):.6f}", flush=True)
        assert same(a.loss, b.loss)
        for p, q in zip(a.params, b.params):
            assert same(p, q)


def test_autotune_restores_state(pkg):
    from cvqpg_amd.batch import BatchTrainer
    bt = BatchTrainer(None, trainers=trainers(SMALL), streams=2)
    bt.step()
    torch.cuda.synchronize()
    before = bt.snapshot()
    chosen = bt.autotune(grids=(1, 4), chunks=(1, 2), streams=(0, 2), reps=2)
    print(chosen, flush=True)
    for t, (ps, st, it) in zip(bt.trainers, before):
        for p, q in zip(t.params, ps):
            assert same(p, q)
        assert same(t.state, st) and same(t.iteration, it)
    assert bt.n_streams in (0, 2)
    bt.capture()
    bt.run()
    torch.cuda.synchronize()
    assert all(torch.isfinite(t.loss).all() for t in bt.trainers)
