"""The 1024^3 workload at its full size, through properties that do not depend on the size:
structure functions of orders 2-6 at every separation 1..512 against their closed form for
u_x = sin x, and one 8-level woven field (reduced density) generated with fields resident, whose
normalised velocity carries u' = 1."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 1024


def test_structure_functions_all_separations_closed_form():
    from fastweavturb_amd.api import WovenGPU
    g = WovenGPU(N, 32)
    x = (2.0 * np.pi / N) * np.arange(N)
    ux = np.broadcast_to(np.sin(x).astype(np.float32)[None, None, :], (N, N, N))
    g.set_field("ux", np.ascontiguousarray(ux))
    seps = np.arange(1, 513, dtype=np.int32)
    S = g.structure_functions(1, 2, 6, seps)
    s = np.sin(np.pi * seps / N)                 # sin(a/2), a = 2 pi r / N
    exact = {2: 2 * s ** 2, 4: 6 * s ** 4, 6: 20 * s ** 6}
    for p, ref in exact.items():
        np.testing.assert_allclose(S[p - 2], ref, rtol=2e-4)
    for p in (3, 5):                             # odd orders vanish
        assert np.max(np.abs(S[p - 2])) <= 1e-5 * np.max(exact[p + 1])
    g.close()


def test_eight_level_field_is_normalised():
    from fastweavturb_amd import fbb
    from fastweavturb_amd.api import WovenGPU
    p, t = fbb.config_c3(density=0.125)
    g = WovenGPU(N, 32)
    g.set_params(p)
    g.set_tubes(t)
    etot, up, ens = g.generate()
    assert up > 0 and math.isfinite(ens) and ens > 0
    _, up2, _ = g.scalars()
    assert up2 == up
    E = g.spectrum(1)
    # u' = 1: the shells hold 3/2 minus the corner modes beyond k = nek
    assert 1.4 < E.sum() <= 1.5 + 1e-6
    dep = g.deposit_stats()
    assert dep["overflow"] == 0 and dep["hits"] > 0
    S = g.structure_functions(1, 2, 6, np.array([1, 8, 64, 512], dtype=np.int32))
    assert np.all(S[0] > 0) and np.all(S[2] > 0)
    g.close()
