"""Posterior agreement with long reference chains (the fourth correctness
check): plain and learned-proposal GPU chains against long oracle chains on
ten ROIs of the batched uneven workload, dense 32 x 32 ROIs included.

Reference: tests/golden/posterior_batched_uneven.npz, written by
tests/golden/make_posterior_fixture.py: per ROI 128 independent oracle
chains of 20,000 + 980,000 jumps; each chain's final state is one posterior
sample (emitter count, localization errors against the truth).

GPU: per ROI 1,024 independent chains (distinct stream keys) of the same
length and parameters; each chain's final state is one posterior sample.
(At 100,000 jumps the emitter count of the 30-emitter 32 x 32 ROI has not
forgotten the empty start yet, so the reference chains run 1,000,000; two
halves of the oracle's own samples agree to KS p 0.55 and 1.2 % in RMSE.)

Checks, per proposal (bamf_amd/posterior.py):
  - emitter-count histograms: two-sample KS of the GPU samples' N against
    the oracle's, p > 0.01 on the pooled samples of all ROIs and p > 0.001 on
    each ROI (ten ROIs: 1 % family-wise);
  - localization: the posterior localization RMSE against the truth within
    2 % of the oracle's (its Monte Carlo error at these sample sizes is below
    1 %).  MAP-N and its localizations are printed beside it.
"""
import os

import numpy as np
import pytest

from bamf_amd import _hip, posterior
from bamf_amd.rjmcmc import LEARNED, PLAIN, run_batch

pytestmark = pytest.mark.gpu

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "posterior_batched_uneven.npz")
K = 1024
CAP = 96


@pytest.fixture(scope="module")
def ref():
    return posterior.load_reference(FIX)


def _samples(ref, plan, rois, mode):
    res = run_batch(plan, rois, seed=4242, mode=mode, final_cap=CAP, hist_len=128)
    assert np.all(res.status == 0), np.unique(res.status)
    assert res.final[:, 0].max() < CAP
    return posterior.final_samples(res.final, len(ref["side"]), K, CAP)


def _check_agreement(ref, ns, pos, label):
    a = posterior.agreement(ref, ns, pos)
    for r in range(len(ns)):
        print(f"[{label}] roi {r} side {ref['side'][r]}: N gpu {ns[r].mean():.2f} +- {ns[r].std():.2f}, "
              f"oracle {ref['oracle_n_roi'][r].mean():.2f} +- {ref['oracle_n_roi'][r].std():.2f}, "
              f"KS p {a['ks_p'][r]:.3g}", flush=True)
    print(f"[{label}] pooled KS p {a['ks_p_pooled']:.3g}; posterior localization RMSE gpu {a['rmse_nm']:.3f} nm, "
          f"oracle {a['oracle_rmse_nm']:.3f} nm, ratio {a['rmse_ratio']:.4f}; MAP-N gpu {a['map_n']}, oracle "
          f"{a['oracle_map_n']}; MAP-N localization RMSE gpu {a['map_rmse_nm']:.2f} nm, oracle "
          f"{a['oracle_map_rmse_nm']:.2f} nm", flush=True)
    assert a["ks_p_pooled"] > 0.01, a["ks_p_pooled"]
    assert min(a["ks_p"]) > 0.001, a["ks_p"]
    assert abs(a["rmse_ratio"] - 1) <= 0.02, (a["rmse_nm"], a["oracle_rmse_nm"])


def test_plain_chain_posterior_matches_long_oracle_chains(ref):
    plan = posterior.reference_plan(ref)
    rois = posterior.replicated_batch(ref, K)
    ns, pos = _samples(ref, plan, rois, PLAIN)
    _check_agreement(ref, ns, pos, "plain")


def test_learned_chain_posterior_matches_long_oracle_chains(ref):
    from bamf_amd import proposal

    plan = posterior.reference_plan(ref)
    rois = posterior.replicated_batch(ref, K)
    prm, _ = proposal.train(400, batch=256, lr=3e-3, seed=5)
    mm = proposal.MapMaker(prm)
    dq, dmu = _hip.DeviceBuffer(rois.total_pixels * 4), _hip.DeviceBuffer(rois.total_pixels * 4)
    mm.run(rois, dq, dmu, 0.05)
    _hip.synchronize()
    mm.close()
    plan.set_proposal(dq, dmu, alpha=0.2, halfwidth=0.35, beta=1.0)
    plan.set_intensity_proposal("fit", 1.3, 0.5)
    ns, pos = _samples(ref, plan, rois, LEARNED)
    _check_agreement(ref, ns, pos, "learned")
