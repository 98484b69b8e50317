"""The oracle's integrands of two trial arguments (DependsOn<0, 1>): the solid
materials with their acceleration (MaterialStressFunctor,
phy/solid_mechanics.hpp:834-852: source density * a) and the heat integrand
with the temperature rate (ThermalMaterialIntegrand,
phy/heat_transfer.hpp:380-406: source rho c_p * dT/dt), pinned on the
perturbed unit cube, whose trilinear cells still tile [0,1]^3:

- the rows of a residual sum to the source integrated over the cube (the test
  functions sum to one, so their gradients sum to zero): density * a0 per
  component for a constant acceleration with the solid at rest, and
  rho c_p r0 - s for a constant temperature rate, whatever the temperature;
- the second argument enters linearly: R(U, a) - R(U, 0) is the residual of
  the mass integrand {density u, 0} at a (the oracle's mass-diffusion
  qfunction, an independent path), the same for every U, and R(U, 0) is the
  one-argument residual bit for bit;
- the stored derivative in the first argument does not see the second: its
  J.v equals the one-argument form's bit for bit."""
import numpy as np
import pytest

import oracle as O

CASES = {
    "neohookean": (O.QF_NEOHOOKEAN, (1.3, 10.0, 0.25), 3, 1.3),
    "linear_elastic": (O.QF_LINEAR_ELASTIC, (1.2, 0.8, 0.7), 3, 0.7),
    "j2": (O.QF_J2_LINEAR, (200.0, 0.3, 0.5, 2.0, 0.0, 1.9, 1.0), 3, 1.9),
    "heat": (O.QF_HEAT, (1.1, 0.4, 2.5), 1, 2.5),
}


def forms(case, p=2, n=3):
    kind, params, nc, coef = CASES[case]
    m = O.Mesh.cuboid(n, n, n)
    m.perturb(42, 0.2)
    s = O.Space(m, p, nc)
    f2, f1 = O.Form(s), O.Form(s)
    if kind == O.QF_J2_LINEAR:
        f2.add_domain_integral2(kind, params, state_init=np.zeros(10))
        f1.add_domain_integral_state(kind, params, np.zeros(10))
    else:
        f2.add_domain_integral2(kind, params)
        f1.add_domain_integral(kind, params)
    return s, f2, f1, coef


def field(s, seed, amp):
    return np.random.default_rng(seed).uniform(-amp, amp, s.ndof)


@pytest.mark.parametrize("case", ["neohookean", "linear_elastic", "j2"])
def test_constant_acceleration_integrates_to_density_times_a(case):
    s, f2, _, rho = forms(case)
    a0 = np.array([0.3, -1.1, 2.0])
    R = f2.residual2(np.zeros(s.ndof), np.tile(a0, s.nnodes)).reshape(-1, 3)
    assert np.abs(R.sum(axis=0) - rho * a0).max() < 1e-13


@pytest.mark.parametrize("p", [1, 2, 3])
def test_constant_rate_integrates_to_source(p):
    s, f2, _, c = forms("heat", p=p)
    r0, src = 0.7, CASES["heat"][1][1]
    R = f2.residual2(field(s, 1, 1.0), np.full(s.ndof, r0))
    assert abs(R.sum() - (c * r0 - src)) < 1e-13


@pytest.mark.parametrize("case", list(CASES))
def test_second_argument_enters_as_a_mass_term(case):
    s, f2, f1, coef = forms(case)
    nc = CASES[case][2]
    mass = O.Form(s)
    mass.add_domain_integral(O.QF_MASS_DIFFUSION, (coef, 0.0))
    amp = 1.0 if case == "heat" else 0.01
    a = field(s, 7, 1.0)
    M_a = mass.residual(a)
    for seed in (3, 4):
        U = field(s, seed, amp)
        r0 = f2.residual2(U, np.zeros(s.ndof))
        assert np.array_equal(r0, f1.residual(U))
        d = f2.residual2(U, a) - r0
        assert np.abs(d - M_a).max() <= 1e-13 * max(np.abs(M_a).max(), np.abs(r0).max())
    assert nc in (1, 3)


@pytest.mark.parametrize("case", list(CASES))
def test_first_argument_derivative_ignores_the_second(case):
    s, f2, f1, _ = forms(case)
    amp = 1.0 if case == "heat" else 0.01
    U, a, dU = field(s, 3, amp), field(s, 5, 1.0), field(s, 6, 1.0)
    f2.linearize2(U, a)
    f1.linearize(U)
    assert np.array_equal(f2.jvp(dU), f1.jvp(dU))
