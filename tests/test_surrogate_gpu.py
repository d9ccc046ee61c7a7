"""Inverse design end to end on the GPU at a small scale (3 levels, 64^3): training pairs from
batched ensembles, the network trained by the library's MFMA kernels, and what it gains over the
method without it: its parameter errors against the prior mean, and the runs it needs to bring a
hidden target's 32 log shells within +-5% against the plain sweep's."""
import numpy as np
import pytest

from fastweavturb_amd import surrogate as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def trained(tmp_path_factory):
    bank = S.CenterlineBank(4321, nlev=3)
    bp = str(tmp_path_factory.mktemp("bank") / "bank.npz")
    bank.save(bp)
    jobs = S.family_jobs(1024, 5, 64, nlev=3)
    X, Y, M, Ns = S.generate_pairs(jobs, ensemble=64, workers=4, bank_path=bp)
    assert np.all(Ns == 64) and np.all(np.isfinite(X))
    ntr = 896
    sur = S.Surrogate(seed=3)
    hist = sur.fit(X[:ntr], Y[:ntr], M[:ntr], epochs=300)
    assert hist[-1] < 0.1 * hist[0]
    return bank, sur, X[ntr:], Y[ntr:], M[ntr:]


def test_network_beats_prior_mean(trained):
    _, sur, X, Y, M = trained
    P = sur.predict(X)
    e = S.masked_rmse(P, Y, M)
    base = S.masked_rmse(np.broadcast_to(sur.ym, P.shape), Y, M)
    a_slots = [4, 8]                       # log2 circulation ratios of levels 1 and 2
    assert np.all(e[a_slots] < 0.6 * base[a_slots])
    assert e[a_slots].sum() < 0.4 * base[a_slots].sum()
    b = np.arange(3) * 4 + 1
    acc = np.mean([abs(S.snap_density(p[b]) - y[1]) < 1e-9 for p, y in zip(P, Y)])
    assert acc >= 0.95


def test_surrogate_reaches_targets_in_fewer_runs_than_the_sweep(trained):
    bank, sur, *_ = trained
    ver = S.Verifier(bank, 64)
    try:
        s_runs, w_runs = [], []
        for d in S.target_designs(2, seed=9100, nlev=3):
            tgt = ver.run(d)
            r = S.surrogate_design(sur, tgt, ver, max_runs=6)
            assert r["found"], r["errors"]
            sw = S.sweep_design(tgt, ver)
            s_runs.append(r["runs"])
            w_runs.append(sw["runs"])
        assert sum(s_runs) * 5 <= sum(w_runs)
    finally:
        ver.close()
