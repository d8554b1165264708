"""GPU: the mixed-batch persistent launch (fft3d_mx_multi_*) against the oracle, and each
item's result identical whether it runs alone or inside the mix."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

MIX = [((6, 6, 6), 300, -1), ((8, 8, 8), 130, 1), ((12, 12, 12), 17, -1),
       ((17, 17, 17), 9, 1), ((8, 8, 16), 21, -1), ((8, 8, 8), 1, -1)]


@pytest.mark.parametrize("prec", [0, 1])
def test_mixed_batch_matches_oracle(gpu, prec):
    f = gpu.fft3d
    dt = np.complex128 if prec == 0 else np.complex64
    tol = 1e-13 if prec == 0 else 2e-6
    xs = [oracle.make_input(s, b, seed=11 + k, dtype=dt) for k, (s, b, d) in enumerate(MIX)]
    m = f.MultiPlan(MIX, precision=prec)
    name, ntasks, grid = m.info()
    assert ntasks > len(MIX)
    # persistent: at most 8 workgroups per CU (MI355X: 256 CUs), never more than the tasks
    assert 0 < grid <= min(ntasks, 8 * 256), (grid, ntasks)
    ys = m.run(xs)
    for x, y, (s, b, d) in zip(xs, ys, MIX):
        err = oracle.rel_l2(y, oracle.fftn_numpy(x, sign=d))
        print(f"{name} {s} B={b} dir={d}: {err:.2e}", flush=True)
        assert err <= tol, (s, d, err)
    m.close()


def test_mixed_equals_alone(gpu):
    """Each item alone (a one-item mixed launch) and inside the mix give identical bits."""
    f = gpu.fft3d
    xs = [oracle.make_input(s, b, seed=5 + k) for k, (s, b, d) in enumerate(MIX)]
    m = f.MultiPlan(MIX, flags=f.NO_TUNE)
    mixed = m.run(xs)
    m.close()
    for x, y, item in zip(xs, mixed, MIX):
        one = f.MultiPlan([item], flags=f.NO_TUNE)
        alone = one.run([x])[0]
        one.close()
        assert np.array_equal(alone.view(np.uint8), y.view(np.uint8)), item


def test_mixed_rejects_bad_items(gpu):
    f = gpu.fft3d
    with pytest.raises(ValueError):
        f.MultiPlan([((36, 36, 36), 2, -1)])
    with pytest.raises(ValueError):
        f.MultiPlan([((8, 8, 8), 2, 0)])
    with pytest.raises(ValueError):
        f.MultiPlan([((8, 8, 8), 1, -1)] * 17)
