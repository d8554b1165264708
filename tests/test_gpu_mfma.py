"""Matrix-core DFT-as-matrix variant (complex64 only, opt-in by plan flag).

Each axis pass is a real (2N x 2N) x (2N x 16) product per 16 lines on
v_mfma_f32_16x16x{16,32} instructions, both operands split into high + low half-precision
parts, hh + hl + lh accumulated in FP32, every tile scaled by a power of two before an fp16
split.  Checked against the exact oracle; the bars below are the variants' stated accuracy,
looser than the FFT path's 2e-6, and hold at any input scale."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

SHAPES = [(6, 6, 6), (8, 8, 8), (12, 12, 12), (17, 17, 17), (8, 8, 16)]
BARS = {"bf16": 2e-5, "fp16": 6e-7}  # observed 4-7e-6 and 1.1-1.9e-7


@pytest.mark.parametrize("split", ["bf16", "fp16"])
def test_mfma_plans_match_oracle(gpu, split):
    f = gpu.fft3d
    flag = f.MFMA_BF16 if split == "bf16" else f.MFMA_FP16
    for shape in SHAPES:
        for direction in (f.FORWARD, f.BACKWARD):
            for batch in (1, 33):
                x = oracle.make_input(shape, batch, seed=7 + batch, dist="uniform",
                                      dtype=np.complex64)
                plan = f.Plan(shape, batch, direction, f.C64, flags=flag)
                name = plan.kernel
                assert name.startswith("mfma_") and name.endswith(split + "x3")
                y = plan.run(x)
                plan.close()
                err = oracle.rel_l2(y, oracle.fftn_numpy(x.astype(np.complex128), sign=direction))
                print(f"{name} dir={direction} B={batch} rel_l2={err:.2e}", flush=True)
                assert np.all(np.isfinite(y))
                assert err <= BARS[split], (shape, direction, batch, err)


@pytest.mark.parametrize("split", ["bf16", "fp16"])
@pytest.mark.parametrize("scale,dc", [(1e-3, 0.0), (1e-6, 0.0), (300.0, 0.0), (1e4, 0.0),
                                      (1.0, 300.0), (1e-30, 0.0)])
def test_mfma_accuracy_at_any_input_scale(gpu, split, scale, dc):
    """Small, large and DC-heavy inputs: finite results at the variant's stated bar (fp16
    parts hold only |x| < 65504 and lose bits below 6e-5 without the per-line scaling)."""
    f = gpu.fft3d
    flag = f.MFMA_BF16 if split == "bf16" else f.MFMA_FP16
    for shape in [(17, 17, 17), (8, 8, 8), (12, 12, 12)]:
        x = (oracle.make_input(shape, 5, seed=19, dtype=np.complex128) * scale + dc)
        x = x.astype(np.complex64)
        plan = f.Plan(shape, 5, f.FORWARD, f.C64, flags=flag)
        name = plan.kernel
        y = plan.run(x)
        plan.close()
        err = oracle.rel_l2(y, oracle.fftn_numpy(x.astype(np.complex128)))
        print(f"{name} scale={scale} dc={dc} rel_l2={err:.2e}", flush=True)
        assert np.all(np.isfinite(y))
        assert err <= BARS[split], (shape, scale, dc, err)


def test_every_matrix_core_kernel_forced_by_name(gpu):
    """Every registered matrix-core form (tile sizes, K widths, the 17-point fold, the x pass
    through LDS or straight to HBM, persistent and pipelined), forced by name, both
    directions, at a batch that leaves a partial last block, against the oracle."""
    f = gpu.fft3d
    names = [(n, sh, d) for n, sh, prec, d in f.kernels() if n.startswith("mfma_")]
    assert any("fold" in n for n, _, _ in names)
    for name, shape, d in names:
        split = "bf16" if name.endswith("bf16x3") else "fp16"
        x = oracle.make_input(shape, 37, seed=23, dtype=np.complex64)
        plan = f.Plan(shape, 37, d, f.C64, kernel=name)
        y = plan.run(x)
        plan.close()
        err = oracle.rel_l2(y, oracle.fftn_numpy(x.astype(np.complex128), sign=d))
        print(f"{name} B=37 rel_l2={err:.2e}", flush=True)
        assert np.all(np.isfinite(y)), name
        assert err <= BARS[split], (name, err)


def test_mfma_normalized_round_trip(gpu):
    f = gpu.fft3d
    x = oracle.make_input((17, 17, 17), 8, seed=3, dtype=np.complex64)
    fwd = f.Plan((17, 17, 17), 8, f.FORWARD, f.C64, flags=f.MFMA_FP16)
    bwd = f.Plan((17, 17, 17), 8, f.BACKWARD, f.C64, flags=f.MFMA_FP16 | f.NORMALIZE)
    back = bwd.run(fwd.run(x))
    assert oracle.rel_l2(back, x) <= 2 * BARS["fp16"]
    fwd.close()
    bwd.close()


def test_mfma_flags_rejected_where_unsupported(gpu):
    f = gpu.fft3d
    with pytest.raises(ValueError):
        f.Plan((17, 17, 17), 1, f.FORWARD, f.C128, flags=f.MFMA_BF16)
    with pytest.raises(ValueError):
        f.Plan((36, 36, 36), 1, f.FORWARD, f.C64, flags=f.MFMA_BF16)
    with pytest.raises(ValueError):
        f.Plan((8, 8, 8), 1, f.FORWARD, f.C64, flags=f.MFMA_BF16 | f.MFMA_FP16)
    # without the flag a complex64 plan never picks a matrix-core kernel
    p = f.Plan((8, 8, 8), 64, f.FORWARD, f.C64)
    assert not p.kernel.startswith("mfma_")
    p.close()


@pytest.mark.parametrize("split", ["bf16", "fp16"])
def test_mfma_in_one_mixed_launch(gpu, split):
    """Uneven batches of every fused geometry, both directions, one matrix-core launch (one
    workgroup per task); each item against the oracle at the variant's bar, and equal to the
    same item through the single-geometry matrix-core plan to within rounding of the same
    products."""
    f = gpu.fft3d
    flag = f.MFMA_BF16 if split == "bf16" else f.MFMA_FP16
    items = [((6, 6, 6), 37, -1), ((8, 8, 8), 19, 1), ((12, 12, 12), 5, -1),
             ((17, 17, 17), 3, 1), ((8, 8, 16), 9, -1), ((6, 6, 6), 1, 1)]
    mp = f.MultiPlan(items, f.C64, flags=flag)
    name, ntasks, grid = mp.info()
    assert split in name
    assert grid == ntasks > len(items), (grid, ntasks)
    xs = [oracle.make_input(sh, b, seed=31 + k, dtype=np.complex64)
          for k, (sh, b, d) in enumerate(items)]
    ys = mp.run(xs)
    mp.close()
    for (sh, b, d), x, y in zip(items, xs, ys):
        err = oracle.rel_l2(y, oracle.fftn_numpy(x.astype(np.complex128), sign=d))
        print(f"mixed {split} {sh} B={b} dir={d} rel_l2={err:.2e}", flush=True)
        assert np.all(np.isfinite(y))
        assert err <= BARS[split], (sh, b, d, err)
    with pytest.raises(ValueError):
        f.MultiPlan(items, f.C128, flags=flag)
    with pytest.raises(ValueError):
        f.MultiPlan(items, f.C64, flags=f.MFMA_BF16 | f.MFMA_FP16)
