"""Proposal network on the matrix cores against the numpy oracle."""
import numpy as np
import pytest

from bamf_amd import proposal
from bamf_amd.rjmcmc import RoiSet
from oracle import net as onet

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("batch", [8, 96])
def test_gradients_match_oracle(batch):
    """batch 96: the weight-gradient work-groups each take more than one image."""
    rng = np.random.default_rng(7)
    img, occ, logi = proposal.training_batch(rng, batch, 16)
    p = proposal.init_params(3)
    p += np.random.default_rng(4).normal(0, 0.02, p.size).astype(np.float32)
    net = proposal.Network(p, batch, 16, 16)
    g, loss = net.gradients(img, occ, logi)
    lo, go, h = onet.loss_and_grad(p, img, occ, logi, bf16_mode=True)
    out = net.forward(img)
    net.close()
    print("loss", loss, lo, flush=True)
    assert np.allclose(loss, lo, rtol=1e-3, atol=1e-6)
    assert rel(out, h) < 1e-3
    lay, _ = onet.layout()
    for name, (o, s) in lay.items():
        n = int(np.prod(s))
        e = rel(g[o:o + n], go[o:o + n])
        print(f"{name}: rel err {e:.2e}", flush=True)
        assert e < 2e-2, name


def test_training_lowers_loss():
    p, hist = proposal.train(300, batch=128, lr=3e-3, seed=1, log_every=50)
    first, last = hist[0][1] + hist[0][2], hist[-1][1] + hist[-1][2]
    print("loss", first, "->", last, flush=True)
    assert last < 0.6 * first


def test_training_with_chain_labels():
    """Device-generated batches, every second one labelled by plain chains'
    final states: the labelled targets follow the truth rule on those states
    and training still lowers the loss."""
    p, hist = proposal.train(120, batch=64, lr=3e-3, seed=3, log_every=20, chain_labels=0.5, chain_jumps=1000)
    first, last = hist[0][1] + hist[0][2], hist[-1][1] + hist[-1][2]
    print("loss", first, "->", last, flush=True)
    assert np.isfinite(p).all() and last < 0.8 * first


def test_proposal_maps_equal_per_roi_network():
    """Maps from the 32 x 32 canvas with masking equal the network run on each
    ROI at its own size (zero padding at the ROI border)."""
    rng = np.random.default_rng(2)
    p = proposal.init_params(5)
    sides = [(8, 8), (12, 16), (16, 12), (32, 32), (5, 9)]
    imgs, offs = [], [0]
    for (x, y) in sides:
        im, _, _ = proposal.training_batch(rng, 1, max(x, y))
        imgs.append(im[0, :x, :y].copy())
        offs.append(offs[-1] + x * y)
    data = np.concatenate([im.ravel() for im in imgs])  # [x][y] = data[y + x*Y]
    rs = RoiSet(data, np.array(offs), [s[0] for s in sides], [s[1] for s in sides])
    eps = 0.05
    dq, dmu = proposal.proposal_maps(p, rs, eps=eps)
    q, mu = dq.download(np.float32), dmu.download(np.float32)
    for r, (x, y) in enumerate(sides):
        qr, mr = q[offs[r]:offs[r + 1]], mu[offs[r]:offs[r + 1]]
        assert abs(qr.sum() - 1) < 1e-4 and np.all(qr >= eps / (x * y) * 0.999)
        _, _, h = onet.loss_and_grad(p, imgs[r][None], np.zeros((1, x, y), np.float32),
                                     np.zeros((1, x, y), np.float32), bf16_mode=True)
        s = 1 / (1 + np.exp(-h[0, ..., 0].astype(np.float64)))
        qo = (1 - eps) * s / s.sum() + eps / (x * y)
        assert rel(qr, qo.ravel()) < 2e-3, r
        assert rel(mr, h[0, ..., 1].ravel()) < 2e-3, r
