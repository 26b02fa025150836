"""The configurations' physical set-up (turbulentbl_amd/cases.py), pinned by
what fixes it: the Blasius solution's textbook constants, the reference's
own momentum-thickness definition Re_theta = theta Re Ue (dnssetbl.f90:
1151-1153) on the inflow profile, the friction-velocity units (mean wall
shear 1 for both inflow profiles), the reference's tanh map with the stated
stretch, and the CFL rule."""
import math

import numpy as np
import pytest

from oracle import orc
from tests.case_util import make_config_case
from turbulentbl_amd import cases, inflow


def test_blasius_constants_from_the_ode():
    # f''(0) = 0.332057 makes f' -> 1; f' = 0.99 at eta = 4.910; theta = 0.6641 (textbook values)
    eta = np.linspace(0.0, 12.0, 120001)
    fp = cases.blasius_fprime(eta)
    assert abs(fp[-1] - 1.0) < 1e-5
    assert eta[np.argmax(fp >= 0.99)] == pytest.approx(cases.BLASIUS_ETA99, abs=2e-3)
    assert np.trapezoid(fp * (1 - fp), eta) == pytest.approx(cases.BLASIUS_THETA_ETA, rel=1e-4)
    # slope at the wall
    assert (fp[1] - fp[0]) / (eta[1] - eta[0]) == pytest.approx(cases.BLASIUS_FPP0, rel=1e-3)


@pytest.mark.parametrize("case", ["c1", "c2", "c3", "c4", "c5a"])
def test_flow_parameters_reproduce_re_theta_and_unit_wall_shear(case):
    Re, Ue = cases.flow_parameters(case)
    z = np.linspace(0.0, 1.5, 300001)
    U = (cases.blasius_profile(z, Ue) if cases.CASES[case]["ic"] == "blasius"
         else inflow.mean_profile(z, Re, Ue))
    # the reference's Re_theta = theta Re Ue (dnssetbl.f90:1151-1153), theta in delta99 units
    theta = cases.theta_of(U, z, Ue)
    assert theta * Re * Ue == pytest.approx(cases.CASES[case]["re_theta"], rel=2e-4)
    # velocities in u_tau: the wall shear dU/dz / Re of the inflow profile is 1
    assert (U[1] - U[0]) / (z[1] - z[0]) / Re == pytest.approx(1.0, rel=2e-3)
    assert U[-1] == pytest.approx(Ue, rel=1e-4)
    if cases.CASES[case]["ic"] != "blasius":
        # Coles-Fernholz: U_inf/u_tau = ln(Re_theta)/0.384 + 4.127
        assert Ue == pytest.approx(math.log(cases.CASES[case]["re_theta"]) / 0.384 + 4.127)


def test_stretched_map_is_the_references_with_stretch_2p5():
    g, s = make_config_case("c3", (16, 8, 192))
    x3 = g.mesh["x3"]
    eta = np.linspace(0.0, 1.0, g.nZ)
    want = 3.0 * (1.0 - np.tanh(2.5 * (1.0 - eta)) / np.tanh(2.5))
    assert np.allclose(x3, want, rtol=1e-9, atol=1e-12)
    assert x3[0] == pytest.approx(0.0, abs=1e-15) and x3[-1] == pytest.approx(3.0)
    assert (x3[1] - x3[0]) == pytest.approx(1.072e-3, rel=1e-3)


def test_cfl_rule_and_domain():
    g, s = make_config_case("c2", (32, 16, 24))
    m = g.mesh
    assert m["x1"][-1] == pytest.approx(60.0) and m["x2"][-1] == pytest.approx(6.0)
    dx, dy = m["x1"][1] - m["x1"][0], m["x2"][1] - m["x2"][0]
    dz = 0.5 * (m["x3"][2:] - m["x3"][:-2])
    U = s.Ut[1:-1, 1:-1, 1:-1]
    c = g.dt * (np.abs(U[..., 0]) / dx + np.abs(U[..., 1]) / dy + np.abs(U[..., 2]) / dz[None, None, :])
    assert c.max() == pytest.approx(0.3, rel=1e-12)


def test_blasius_case_initial_field():
    g, s = make_config_case("c1")
    U = s.Ut
    assert s.nstep == 0 and g.i_rec == 51
    assert not U[:, :, 0, :].any()                                     # no slip
    assert np.array_equal(U[:, 0], U[:, g.nY - 2]) and np.array_equal(U[:, -1], U[:, 1])
    mean = U[1:-1, 1:-1, 1:-1, 0].mean(axis=(0, 1))
    want = cases.blasius_profile(g.mesh["x3"][1:-1], g.Ue)
    assert np.abs(mean - want).max() <= 0.05 * g.Ue * 4 / math.sqrt((g.nX - 2) * (g.nY - 2))
    noise = U[1:-1, 1:-1, 1:-1, 1]
    assert np.abs(noise).max() <= 0.05 * g.Ue and np.abs(noise).max() >= 0.045 * g.Ue
