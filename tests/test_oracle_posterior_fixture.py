"""The posterior reference fixture (tests/golden/posterior_batched_uneven.npz)
is what the current oracle produces: a few of its chains, run again, end in
the recorded states, and the stored per-ROI summaries match the recorded
samples."""
import os

import numpy as np

from bamf_amd import posterior
from bamf_amd.rjmcmc import make_params

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "posterior_batched_uneven.npz")


def test_fixture_chains_replay_in_the_oracle(oracle):
    ref = posterior.load_reference(FIX)
    p = make_params(n_burnin=int(ref["n_burnin"]), n_trials=int(ref["n_trials"]), rho=float(ref["rho"]),
                    p_burnin=tuple(ref["p_burnin"]), p_trials=tuple(ref["p_trials"]))
    op = oracle.McmcParams.from_buffer_copy(p)
    o = ref["offset"]
    back = np.zeros_like(ref["sig_pdf"])
    for r in (0, 2):  # an 8 x 8 and a 12 x 12 ROI (fast chains)
        s = int(ref["side"][r])
        for c in range(2):
            res = oracle.rjmcmc(ref["data"][o[r]:o[r + 1]], s, s, op, ref["off_pdf"], ref["sig_pdf"], back,
                                bg0=float(ref["bg0"]), seed=int(ref["chain_seed"]), roi=r * 1000 + c)
            assert res["status"] == 0
            assert res["n"] == ref["oracle_n_roi"][r][c]


def test_fixture_summaries_are_consistent():
    ref = posterior.load_reference(FIX)
    n_roi = len(ref["side"])
    assert len(ref["oracle_n_roi"]) == n_roi and all(len(v) == int(ref["chains"]) for v in ref["oracle_n_roi"])
    assert abs(posterior.posterior_rmse_nm(ref["oracle_err_roi"]) - float(ref["oracle_post_rmse_nm"])) < 1e-9
    # the dense 32 x 32 ROIs are in it
    assert int(ref["side"].max()) == 32 and max(len(t) for t in ref["truth_roi"]) >= 40
