"""Oracle pinning for J2SmallStrain<LinearHardening> with quadrature-point state.

Known answers restated from the reference's own material tests
(physics/materials/tests/J2_material.cpp): the uniaxial-stress closed form
(:147-176, tolerance 1e-10), the consistency check (:121-145) and the
Brannon-Leelavanichkul verification path (:23-114, tolerance 5e-2).  The
kinematic-hardening and cyclic-shear cases are closed forms of the same model
(linear hardening makes the radial return exact in one increment); the tangent
is checked against central differences of the point function.
"""
import numpy as np
import pytest

import oracle as O

# p = {E, nu, sigma_y, Hi, Hk, density, geom_nonlin}
STATE0 = np.zeros(10)


def stress(params, H, state):
    """Cauchy stress of the material alone (geometric nonlinearity off: flux = sigma)."""
    p = list(params[:6]) + [0.0]
    _, f, st = O.qf_eval_state(O.QF_J2_LINEAR, p, 3, np.zeros(3), np.asarray(H, float).ravel(), state)
    return f, st


def tangent(params, H, state):
    p = list(params[:6]) + [0.0]
    return O.qf_tangent(O.QF_J2_LINEAR, p, 3, np.zeros(3), np.asarray(H, float).ravel(), state)


def uniaxial_stress_history(params, t_max, num_steps, eps_xx):
    """uniaxial_stress_test (physics/materials/material_verification_tools.hpp:47-87):
    lateral strains solved by Newton so that sigma_yy = sigma_zz = 0, state committed per step."""
    state = STATE0.copy()
    dudx = np.zeros((3, 3))
    t, dt, out = 0.0, t_max / (num_steps - 1), []
    for _ in range(num_steps):
        x = np.array([dudx[1, 1], dudx[2, 2]])
        for _ in range(20):
            H = np.diag([eps_xx(t), x[0], x[1]])
            s, _ = stress(params, H, state)
            r = np.array([s[1, 1], s[2, 2]])
            if np.abs(r).max() < 1e-15:
                break
            D = tangent(params, H, state)
            Jm = np.array([[D[4, 4], D[4, 8]], [D[8, 4], D[8, 8]]])
            x = x - np.linalg.solve(Jm, r)
        dudx = np.diag([eps_xx(t), x[0], x[1]])
        s, state = stress(params, dudx, state)
        out.append((t, dudx.copy(), s.copy(), state.copy()))
        t += dt
    return out


@pytest.mark.parametrize("Hk", [0.0, 0.004])
def test_uniaxial_closed_form(Hk):
    # J2_material.cpp:147-176 (Hk = 0); with kinematic hardening the uniaxial
    # plastic modulus is Hi + Hk
    E, nu, sy = 1.0, 0.25, 0.01
    Hi = E / 100.0
    Ht = Hi + Hk
    params = [E, nu, sy, Hi, Hk, 1.0]
    hist = uniaxial_stress_history(params, 2.0, 4, lambda t: sy / E * t)
    exact_s = lambda e: E * e if e < sy / E else E / (E + Ht) * (sy + Ht * e)
    exact_p = lambda e: 0.0 if e < sy / E else (E * e - sy) / (E + Ht)
    plastic_steps = 0
    for t, dudx, s, st in hist:
        e = dudx[0, 0]
        assert abs(s[0, 0] - exact_s(e)) <= 1e-10 * abs(exact_s(e))
        assert abs(st[0] - exact_p(e)) <= 1e-10 * abs(exact_p(e))
        assert abs(s[1, 1]) < 1e-14 and abs(s[2, 2]) < 1e-14
        plastic_steps += exact_p(e) > 0
    assert plastic_steps == 2


def test_consistency_after_return():
    # J2_material.cpp:121-145 with linear hardening: after the return map the
    # Mises stress equals the flow stress and dev(sigma) = 2G dev(sym(H) - eps_p)
    H = np.array([[0.7551559, 0.3129729, 0.12388372], [0.548188, 0.8851279, 0.30576992],
                  [0.82008433, 0.95633745, 0.3566252]])
    E, nu, sy, Hi = 1.0, 0.25, 0.1, 0.05
    s, st = stress([E, nu, sy, Hi, 0.0, 1.0], H, STATE0)
    dev = lambda A: A - np.trace(A) / 3.0 * np.eye(3)
    mises = np.sqrt(1.5) * np.linalg.norm(dev(s))
    assert st[9] > 0
    assert abs(mises - (sy + Hi * st[9])) <= 1e-9 * mises
    ep = st[:9].reshape(3, 3)
    s_el = E / (1 + nu) * dev(0.5 * (H + H.T) - ep)
    assert np.linalg.norm(s_el - dev(s)) / np.linalg.norm(s_el) < 1e-9
    # plastic strain is deviatoric and symmetric
    assert abs(np.trace(ep)) < 1e-14 and np.array_equal(ep, ep.T)


def brannon_analytic(t):
    # J2_material.cpp:35-70 (coefficients of the published solution as the test restates them)
    a = np.exp(12.33 * t)
    if t <= 0.201:
        s = [-474.0 * t, -474.0 * t, 948.0 * t]
    elif t <= 1.0:
        s = [-95.26, -95.26, 190.50]
    else:
        c = np.array([[189.40, 0.1704, -0.003242, 0.00001712], [-76.87, -1.4430, 0.001316, 0.00001712],
                      [-112.50, 1.2720, 0.001926, 0.00001712]])
        s = [(c[i, 0] + c[i, 1] * np.sqrt(a) + c[i, 2] * a) / (1.0 + c[i, 3] * a) for i in range(3)]
    return np.diag(s)


def test_brannon_verification_path():
    # J2_material.cpp:72-114: single_quadrature_point_test, 64 steps to t = 2
    G = 79000.0
    K = 10 * G
    E = 9 * K * G / (3 * K + G)
    nu = (3 * K - 2 * G) / (2 * (3 * K + G))
    sy = 165 * np.sqrt(3.0)
    params = [E, nu, sy, 0.0, 0.0, 1.0]
    eps = [np.array([-0.0030000, -0.003, 0.0060000]), np.array([-0.0103923, 0.000, 0.0103923])]
    state = STATE0.copy()
    t, nsteps = 0.0, 64
    dt = 2.0 / (nsteps - 1)
    for _ in range(nsteps):
        a0 = t if t < 1 else 2.0 - t
        a1 = 0.0 if t < 1 else t - 1.0
        H = np.diag(a0 * eps[0] + a1 * eps[1])
        s, state = stress(params, H, state)
        if t > 0:
            assert np.linalg.norm(s - brannon_analytic(t)) / np.linalg.norm(s) < 5e-2
        t += dt


def test_cyclic_shear_closed_form():
    # simple shear eps_12 = gamma/2 with isotropic hardening: equivalent plastic
    # strain = |gamma_p| / sqrt(3); each yielding increment solves
    # sqrt(3) |tau| = sigma_y + Hi eqps exactly
    E, nu, sy, Hi = 200.0, 0.3, 0.5, 2.0
    G = E / (2 * (1 + nu))
    params = [E, nu, sy, Hi, 0.0, 1.0]
    path = np.concatenate([np.linspace(0, 0.03, 7)[1:], np.linspace(0.03, -0.03, 13)[1:],
                           np.linspace(-0.03, 0.03, 13)[1:]])
    state = STATE0.copy()
    gp, eq = 0.0, 0.0  # closed-form plastic shear and accumulated equivalent strain
    for g in path:
        H = np.zeros((3, 3))
        H[0, 1] = g
        s, state = stress(params, H, state)
        tau_tr = G * (g - gp)
        f = np.sqrt(3.0) * abs(tau_tr) - (sy + Hi * eq)
        if f > 1e-10 * sy:
            dq = f / (3 * G + Hi)  # increment of equivalent plastic strain
            gp += np.sign(tau_tr) * np.sqrt(3.0) * dq
            eq += dq
        tau = G * (g - gp)
        assert abs(s[0, 1] - tau) <= 1e-12 * max(abs(tau), sy)
        assert abs(state[9] - eq) <= 1e-12 * max(eq, 1e-3)
        assert abs(2 * state[1] - gp) <= 1e-12 * max(abs(gp), 1e-3)
    assert eq > 0.05  # the path yielded in both directions


@pytest.mark.parametrize("geom", [0.0, 1.0])
@pytest.mark.parametrize("plastic", [False, True])
def test_tangent_matches_central_differences(geom, plastic):
    rng = np.random.default_rng(7)
    E, nu, sy, Hi, Hk = 200.0, 0.3, 0.5, 2.0, 1.0
    p = [E, nu, sy, Hi, Hk, 1.0, geom]
    st = STATE0.copy()
    st[:9] = (lambda A: 0.5 * (A + A.T) - np.trace(A) / 3 * np.eye(3))(rng.uniform(-1e-3, 1e-3, (3, 3))).ravel()
    st[9] = 2e-3
    H = rng.uniform(-1, 1, (3, 3)) * (0.02 if plastic else 1e-4)
    _, f0, st1 = O.qf_eval_state(O.QF_J2_LINEAR, p, 3, np.zeros(3), H.ravel(), st)
    assert (st1[9] > st[9]) == plastic
    D = O.qf_tangent(O.QF_J2_LINEAR, p, 3, np.zeros(3), H.ravel(), st)
    h = 1e-7
    Dfd = np.zeros((9, 9))
    for b in range(9):
        Hp, Hm = H.ravel().copy(), H.ravel().copy()
        Hp[b] += h
        Hm[b] -= h
        _, fp, _ = O.qf_eval_state(O.QF_J2_LINEAR, p, 3, np.zeros(3), Hp, st)
        _, fm, _ = O.qf_eval_state(O.QF_J2_LINEAR, p, 3, np.zeros(3), Hm, st)
        Dfd[:, b] = (fp - fm).ravel() / (2 * h)
    assert np.abs(D - Dfd).max() <= 1e-6 * np.abs(D).max()
    if not geom:
        # small strain: the consistent tangent of associative J2 with linear
        # hardening is major-symmetric and has minor symmetry in the strain
        assert np.abs(D - D.T).max() <= 1e-12 * np.abs(D).max()


def elastic_params(E, nu):
    K = E / (3.0 * (1.0 - 2.0 * nu))
    G = 0.5 * E / (1.0 + nu)
    return K - 2.0 / 3.0 * G, G


def j2_form(n=2, p=2, params=None, perturb=0.1):
    m = O.Mesh.cuboid(n, n, n)
    if perturb:
        m.perturb(42, perturb)
    sp = O.Space(m, p, 3, 1)
    f = O.Form(sp)
    f.add_domain_integral_state(O.QF_J2_LINEAR, params, STATE0)
    return m, sp, f


def test_elastic_range_equals_linear_elasticity():
    E, nu = 200.0, 0.3
    params = [E, nu, 0.5, 2.0, 0.0, 1.0, 0.0]
    m, sp, f = j2_form(params=params)
    lam, mu = elastic_params(E, nu)
    le = O.Form(sp)
    le.add_domain_integral(O.QF_LINEAR_ELASTIC, [lam, mu])
    xyz = sp.node_coords()
    U = np.zeros(sp.ndof)
    U[0::3] = 1e-5 * np.sin(3 * xyz[:, 1])
    U[1::3] = 2e-5 * xyz[:, 0] * xyz[:, 2]
    R = f.residual(U)
    R2 = le.residual(U)
    assert np.abs(R - R2).max() <= 1e-13 * np.abs(R2).max()
    assert np.abs(f.qdata()).max() == 0.0  # elastic: nothing to commit


def test_state_commit_semantics():
    params = [200.0, 0.3, 0.5, 2.0, 1.0, 1.0, 1.0]
    m, sp, f = j2_form(params=params)
    xyz = sp.node_coords()
    U = np.zeros(sp.ndof)
    U[0::3] = 0.03 * xyz[:, 1]  # shear beyond yield
    R0 = f.residual(U)
    assert np.abs(f.qdata()).max() == 0.0  # residual without update leaves the state
    f.linearize(U)
    assert np.abs(f.qdata()).max() == 0.0
    f.update_qdata(True)
    R1 = f.residual(U)
    f.update_qdata(False)
    st1 = f.qdata()
    assert np.array_equal(R0, R1)  # the committing evaluation returns the same residual
    assert st1[..., 9].min() > 0
    # re-evaluating at the committed state does not enter the return map again
    f.update_qdata(True)
    f.residual(U)
    f.update_qdata(False)
    assert np.array_equal(f.qdata(), st1)
    # set/get round trip
    st2 = st1 * 0.5
    f.set_qdata(st2)
    assert np.array_equal(f.qdata(), st2)
