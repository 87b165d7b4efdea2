"""The coupled advection-diffusion transport (advection_diffusion_source,
heat_transport.cpp:774-816; solve_implicit_advection_diffusion, :1083-1318):
the oracle's restatements against the compiled reference, bit for bit, and
the source's conservation (sum A H = 0 to rounding, with latent heat of zero
area sum)."""
import numpy as np
import pytest

import support as S

needs_ref = pytest.mark.skipif(not S.have_ref(), reason="oracle/_ref not built (reference sources absent)")


def source(lib, m, case, mean, energy, latent=None):
    eddy, adv = np.zeros(m.n), np.zeros(m.n)
    lat = None if latent is None else np.ascontiguousarray(latent)
    lib.check(lib.advection_source(m.handle, case["layers"], case["conductance"], case["outflow"],
                                   None if lat is None else lat.ctypes.data, mean, energy, eddy, adv))
    return eddy, adv


@pytest.mark.parametrize("level,layers", [(3, 1), (4, 3)])
def test_oracle_advection_source(level, layers):
    m = S.mesh("oracle", level)
    case = S.advection_case(m, layers, 5)
    theta, _, energy, _ = case["response"](np.zeros(m.n))
    latent = np.sin(np.arange(m.n)) * 3.0
    latent -= np.sum(latent * m.area) / np.sum(m.area) * np.ones(m.n)
    eddy, adv = source(S.oracle(), m, case, theta, energy, latent)
    total = m.area * (eddy + adv)
    assert abs(total.sum()) <= 1e-11 * np.abs(total).sum()
    if S.have_ref():
        r_eddy, r_adv = source(S.ref(), S.mesh("ref", level), case, theta, energy, latent)
        assert np.array_equal(eddy.view(np.uint64), r_eddy.view(np.uint64))
        assert np.array_equal(adv.view(np.uint64), r_adv.view(np.uint64))


@needs_ref
@pytest.mark.parametrize("level,layers,quartic", [(3, 2, False), (4, 3, True)])
def test_oracle_advection_diffusion_matches_reference(level, layers, quartic):
    m = S.mesh("oracle", level)
    case = S.advection_case(m, layers, 9 + level, quartic)
    a = S.solve_advection(S.oracle(), m.handle, m.n, case)
    b = S.solve_advection(S.ref(), S.mesh("ref", level).handle, m.n, case)
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x.view(np.uint64), y.view(np.uint64))
    fa = [getattr(a[3], f) for f, _ in S.AdvectionResult._fields_]
    fb = [getattr(b[3], f) for f, _ in S.AdvectionResult._fields_]
    assert fa == fb
    assert a[3].consistency_residual <= 1e-6 and a[3].newton_iterations >= 1
