"""The deposition's curved, twist and varying-core terms (tube.f90:844-892) against closed forms.

Straight tubes with twist and with a varying core (straight branch, tube.f90:854-860) and circles
with and without twist and varying core (Frenet branch, tube.f90:868-892) have fields that follow
from the geometry alone (tests/closed_forms.py). The FP64 oracle is checked against them on the
CPU; the GPU path (FP64 and FP32 contexts) against them and against the oracle.
"""
import numpy as np
import pytest

import closed_forms as CF

CASES = CF.cases()
IDS = [c.name for c in CASES]


def oracle_field(O, case):
    p, t = case.params, case.tubes
    P = O.Params(p.N, p.n_d, p.line_s, p.line_e, p.sigma_ratio, p.eta_t, p.ntf_times, p.ntf_pre,
                 p.Rtube_ratio, p.dzeta_sup)
    geo = O.Geometry(P, t.npoint, t.cx, t.cy, t.cz)
    return geo.deposit(t.sigma, t.gamma, t.n_T)


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_oracle_matches_closed_form(oracle, case):
    om, hits = oracle_field(oracle, case)
    assert hits > 0
    CF.check(case, om)


def _mistakes(case):
    m = []
    if case.params.eta_t != 0:
        m.append("flip_twist")
    if case.params.sigma_ratio != 1:
        m.append("drop_core")
    if case.name.startswith("ring") and (case.params.eta_t != 0 or case.params.sigma_ratio != 1):
        m.append("drop_D")
    return m


PLANTED = [(c, m) for c in CASES for m in _mistakes(c)]


@pytest.mark.parametrize("case,mistake", PLANTED, ids=[f"{c.name}-{m}" for c, m in PLANTED])
def test_closed_forms_reject_planted_mistakes(oracle, case, mistake):
    """Each bound is tight enough to catch a twist term of the wrong sign, a dropped varying-core
    term or a dropped 1/D: the oracle's field is far outside the bound of the altered expectation."""
    om, _ = oracle_field(oracle, case)
    alt = case.with_mistake(mistake)
    d = np.linalg.norm(om - alt.expect) / np.linalg.norm(alt.expect)
    assert d > 20 * case.rtol_l2, (case.name, mistake, d)


def _gpu_field(case, precision):
    from fastweavturb_amd.api import WovenGPU
    with WovenGPU(case.params.N, precision, 0) as g:
        g.set_params(case.params)
        g.set_tubes(case.tubes)
        g.build_segments()
        g.bin()
        st = g.deposit()
        return g.get_fields("omega"), st["hits"]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_gpu_matches_closed_form_and_oracle(oracle, case, precision):
    print(f"closed form {case.name} FP{precision}", flush=True)
    om, hits = _gpu_field(case, precision)
    ref, ref_hits = oracle_field(oracle, case)
    assert hits == ref_hits
    CF.check(case, om)
    rel = np.linalg.norm(om - ref) / np.linalg.norm(ref)
    assert rel <= (1e-12 if precision == 64 else 1e-5), rel
