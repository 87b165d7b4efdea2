"""The coupled advection-diffusion transport on the device, against the
oracle (pinned bit for bit to the compiled reference in
test_advection_oracle.py):

- advection_diffusion_source: bit-identical;
- solve_implicit_advection_diffusion: the device's inner BiCGSTAB (Jacobi-
  smoothed V-cycle, block ILU(0) bit-identical to the reference's, dot
  products in another fixed order) reaches the same Newton solution: both
  end with a consistency residual <= 1e-6 W/m^2 and |dH| <= 1e-5 W/m^2,
  |dtheta| <= 1e-6 K; the source conserves energy (sum A H = 0 to 1e-12 of
  sum A |H|).
"""
import numpy as np
import pytest

import planet_game_amd as P
import support as S
from test_gpu_parity import announce, bits_equal, product_mesh

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("level,layers", [(4, 1), (5, 3)])
def test_advection_source_bit_exact(ctx, level, layers):
    announce("advection source", level, layers)
    om = S.mesh("oracle", level)
    case = S.advection_case(om, layers, 5)
    theta, _, energy, _ = case["response"](np.linspace(-50.0, 80.0, om.n))
    latent = 2.0 * np.cos(np.arange(om.n))
    eddy, adv = P.advection_source(P.Graph(ctx, product_mesh(level)), layers, case["conductance"], case["outflow"],
                                   theta, energy, latent)
    r_eddy, r_adv = np.zeros(om.n), np.zeros(om.n)
    S.oracle().check(S.oracle().advection_source(om.handle, layers, case["conductance"], case["outflow"],
                                                 latent.ctypes.data, theta, energy, r_eddy, r_adv))
    assert bits_equal(eddy, r_eddy) and bits_equal(adv, r_adv)


@pytest.mark.parametrize("level,layers,quartic", [(3, 2, False), (4, 3, True), (5, 2, True)])
def test_advection_diffusion_matches_oracle(ctx, level, layers, quartic):
    announce("advection diffusion", level, layers, quartic)
    om = S.mesh("oracle", level)
    case = S.advection_case(om, layers, 9 + level, quartic)
    sd, md, ed, rd = P.implicit_advection_diffusion(P.Graph(ctx, product_mesh(level)), layers, case["conductance"],
                                                    case["outflow"], case["floor"], case["response"])
    sr, mr, er, rr = S.solve_advection(S.oracle(), om.handle, om.n, case)
    print("level", level, "layers", layers, "newton", rd.newton_iterations, rr.newton_iterations, "linear",
          rd.linear_iterations, rr.linear_iterations, "residual", rd.consistency_residual_W_m2,
          rr.consistency_residual, "dH", float(np.max(np.abs(sd - sr))), "dT", float(np.max(np.abs(md - mr))),
          flush=True)
    assert rd.consistency_residual_W_m2 <= 1e-6
    assert abs(rd.sum_W) <= 1e-12 * rd.absolute_sum_W
    assert np.max(np.abs(sd - sr)) <= 1e-5
    assert np.max(np.abs(md - mr)) <= 1e-6
    assert rd.advective_absolute_sum_W > 0.0 and rd.eddy_absolute_sum_W > 0.0


def test_advection_diffusion_rejects_negative_conductance(ctx):
    announce("advection diffusion errors")
    om = S.mesh("oracle", 3)
    case = S.advection_case(om, 1, 2)
    bad = case["conductance"].copy()
    bad[3] = -1.0
    with pytest.raises(P.PsimError, match="eddy conductances must be finite and non-negative"):
        P.implicit_advection_diffusion(P.Graph(ctx, product_mesh(3)), 1, bad, case["outflow"], case["floor"],
                                       case["response"])
