"""CPU oracle for the CVQPG hologram hot path (test infrastructure only).

This package restates, on the CPU, the reference algorithm of every function on
the hot path so that the HIP kernels can be checked against it.  It is imported
only by ``tests/``, by ``__graft_entry__.smoke()`` (as the checker) and by the
``cpu_baseline`` leg of ``bench.py``.  The shipped package never imports it and
never falls back to it.

Modules and the reference code they restate:

* ``binning``   -- 5-sigma tile binning, key emission, key sort and tile ranges
                   (``cuda/forward.cu:8-60``, ``cuda/raster.cu:30-105``).  numpy,
                   float32 rect arithmetic, bit-exact.
* ``raster``    -- dense FP64 forward of the quadratic-phase Gaussian sum
                   (``cuda/rasterizer.py:74-90``), the binned forward
                   (``cuda/forward.cu:62-162``) and the closed-form adjoint
                   (``cuda/backward.cu:11-168``).
* ``model``     -- activations, covariance, inverse, radii, soft_window_pad and
                   the reconstruct glue (``model_2d_gaussian_rgb.py:83-153,
                   186-241, 424-439``; ``train_2d_gaussian_rgb.py:89-95``), plus
                   the 3-sigma / 64-px torch tile renderer (``:269-422``) that
                   serves as the CPU baseline.
* ``optics``    -- band-limited angular spectrum propagation with an FP64
                   transfer-function phase (``utils/propagator.py:41-64,
                   202-290`` and odak's BLASM kernel it calls).
* ``loss``      -- the three-term masked L2, SSIM term and target synthesis
                   (``utils/propagator.py:361-449``, ``utils/data_utils.py:22-82``).
* ``export``    -- DPAC phase-only encoding with 8-bit quantisation and the per-plane
                   PSNR / SSIM evaluation (``export_poh.py:110-119,173-178``,
                   ``train_2d_gaussian_rgb.py:98-112``).
* ``adan``      -- Adan optimiser step, nan_to_num, global clip and the means
                   cosine schedule (``utils/optimizer.py:105-262``,
                   ``train_2d_gaussian_rgb.py:58-86,205-211``).
* ``step``      -- one full training step in torch FP64/FP32 on the CPU
                   (``train_2d_gaussian_rgb.py:179-211``).

How each function is pinned (golden fixtures, closed forms, brute force,
finite differences) is stated in its docstring and exercised by
``tests/test_oracle_*.py``.
"""
