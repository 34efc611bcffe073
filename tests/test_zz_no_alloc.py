"""lbm_mlp_loss_grad_noalloc (the loss gradient with its scratch in a fixed
device array instead of allocations) against lbm_mlp_loss_grad, through the
C ABI: the same kernels and launch shape, so the same bits, on two calls in
a row (the second with fewer samples, so the array holds the first call's
larger contents), and sizes beyond its limit are rejected before any
device work."""
import numpy as np
import pytest

from simuverse_amd import lbm as L

pytestmark = pytest.mark.gpu

NOALLOC_MAX = 16384


def test_noalloc_gradient_equals_allocating_gradient():
    assert L.device_count() >= 1, "no gfx950 device visible"
    rng = np.random.default_rng(31)
    net = L.Network(seed=13)
    net.set_params((rng.standard_normal(5385) * 0.2).astype(np.float32))
    for n in (NOALLOC_MAX, 3001):
        print(f"case: loss_grad_noalloc n={n}", flush=True)
        x = rng.standard_normal((n, 9)).astype(np.float32)
        t = rng.standard_normal((n, 9)).astype(np.float32)
        l_ref, g_ref = net.loss_grad(x, t)
        l_na, g_na = net.loss_grad_noalloc(x, t)
        assert l_na == l_ref
        assert np.array_equal(g_na, g_ref)
        assert np.all(np.isfinite(g_na)) and np.any(g_na != 0)
    # a second network on the same device shares the array
    other = L.Network(seed=14)
    x = rng.standard_normal((700, 9)).astype(np.float32)
    t = rng.standard_normal((700, 9)).astype(np.float32)
    assert np.array_equal(other.loss_grad_noalloc(x, t)[1], other.loss_grad(x, t)[1])
    print("case: loss_grad_noalloc beyond its limit", flush=True)
    big = np.zeros((NOALLOC_MAX + 1, 9), np.float32)
    with pytest.raises(L.LbmError) as ei:
        net.loss_grad_noalloc(big, big)
    assert ei.value.code == 1 and "16384" in str(ei.value)
    other.close()
    net.close()
