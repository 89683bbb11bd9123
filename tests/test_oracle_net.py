"""The network oracle's gradient against central finite differences (float64),
and its bfloat16 mode against float64 (the rounding is small)."""
import numpy as np

from bamf_amd.proposal import init_params, param_layout, training_batch
from oracle import net as onet


def small_case(seed=3, n=3, side=6):
    rng = np.random.default_rng(seed)
    img, occ, logi = training_batch(rng, n, side, density=(20.0, 40.0))
    p = init_params(seed).astype(np.float64)
    p += np.random.default_rng(seed + 1).normal(0, 0.02, p.size)  # break the zero biases
    return p, img, occ, logi


def test_layout_matches_device_layout():
    lay = param_layout()
    olay, tot = onet.layout()
    assert tot == lay["total"]
    for k, (o, s) in olay.items():
        assert lay[k] == (o, s)


def test_gradient_matches_finite_differences():
    p, img, occ, logi = small_case()
    loss, g, _ = onet.loss_and_grad(p, img, occ, logi, bf16_mode=False)
    rng = np.random.default_rng(0)
    lay, _ = onet.layout()
    for name in ("w1", "b1", "w2", "w3", "b3", "w4", "b4", "wh", "bh"):
        o, s = lay[name]
        for i in rng.choice(int(np.prod(s)), size=min(4, int(np.prod(s))), replace=False):
            h = 1e-6
            pp, pm = p.copy(), p.copy()
            pp[o + i] += h
            pm[o + i] -= h
            fd = (onet.loss_and_grad(pp, img, occ, logi, False)[0].sum()
                  - onet.loss_and_grad(pm, img, occ, logi, False)[0].sum()) / (2 * h)
            assert abs(fd - g[o + i]) <= 1e-6 + 1e-4 * abs(fd), (name, i, fd, g[o + i])


def test_bf16_mode_close_to_float64():
    p, img, occ, logi = small_case(seed=5, n=4, side=8)
    l64, g64, _ = onet.loss_and_grad(p, img, occ, logi, bf16_mode=False)
    l16, g16, _ = onet.loss_and_grad(p.astype(np.float32), img, occ, logi, bf16_mode=True)
    assert np.allclose(l16, l64, rtol=2e-2, atol=1e-4)
    assert np.linalg.norm(g16 - g64) < 5e-2 * np.linalg.norm(g64)
