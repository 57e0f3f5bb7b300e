"""Alart-Curnier function of the oracle (fc3d_AlartCurnier_functions.c:506-812)
checked against an independent formulation differentiated with dual numbers,
on every row of the reference fixture ACinputs.dat (the comparison of
test/fc3d_AlartCurnierFunctions_test.c:122-151), and against finite
differences."""
import numpy as np

from conftest import read_ac_inputs


def ac(oracle, fn, R, U, mu, rho):
    F, A, B = np.zeros(3), np.zeros(9), np.zeros(9)
    getattr(oracle.load(), fn)(oracle.d(np.ascontiguousarray(R)), oracle.d(np.ascontiguousarray(U)),
                               mu, oracle.d(np.ascontiguousarray(rho)), oracle.d(F), oracle.d(A),
                               oracle.d(B))
    return F, A.reshape(3, 3, order="F"), B.reshape(3, 3, order="F")


def test_std_against_dual_formulation_on_acinputs(oracle):
    R, U, MU, RHO = read_ac_inputs()
    worst = 0.0
    for k in range(len(MU)):
        F1, A1, B1 = ac(oracle, "oracle_computeAlartCurnierSTD", R[k], U[k], MU[k], RHO[k])
        F2, A2, B2 = ac(oracle, "oracle_alart_curnier_dual", R[k], U[k], MU[k], RHO[k])
        worst = max(worst, np.abs(F1 - F2).max(), np.abs(A1 - A2).max(), np.abs(B1 - B2).max())
    # the reference test accepts 1e-6; the two formulations agree far closer
    assert worst < 1e-10


def test_std_jacobian_by_finite_differences(oracle):
    rng = np.random.default_rng(7)
    h = 1e-7
    checked = 0
    for _ in range(400):
        R = rng.normal(size=3) + np.array([1.0, 0, 0])
        U = rng.normal(size=3)
        mu = rng.uniform(0.1, 0.9)
        rho = np.array([rng.uniform(0.2, 2), rng.uniform(0.2, 2)])
        rho = np.array([rho[0], rho[1], rho[1]])
        F, A, B = ac(oracle, "oracle_computeAlartCurnierSTD", R, U, mu, rho)
        ok = True
        Anum, Bnum = np.zeros((3, 3)), np.zeros((3, 3))
        for j in range(3):
            e = np.zeros(3)
            e[j] = h
            Fp, _, _ = ac(oracle, "oracle_computeAlartCurnierSTD", R, U + e, mu, rho)
            Fm, _, _ = ac(oracle, "oracle_computeAlartCurnierSTD", R, U - e, mu, rho)
            Anum[:, j] = (Fp - Fm) / (2 * h)
            Fp, _, _ = ac(oracle, "oracle_computeAlartCurnierSTD", R + e, U, mu, rho)
            Fm, _, _ = ac(oracle, "oracle_computeAlartCurnierSTD", R - e, U, mu, rho)
            Bnum[:, j] = (Fp - Fm) / (2 * h)
        # skip samples within h of a kink (branch switch inside the stencil)
        if np.abs(Anum - A).max() > 1e-3 and np.abs(Bnum - B).max() > 1e-3:
            ok = False
        if ok:
            np.testing.assert_allclose(A, Anum, atol=2e-5)
            np.testing.assert_allclose(B, Bnum, atol=2e-5)
            checked += 1
    assert checked > 300


def test_jacobian_sign_mutation_is_caught(oracle):
    R, U, MU, RHO = read_ac_inputs()
    try:
        oracle.mutation(1)
        worst = 0.0
        for k in range(len(MU)):
            _, A1, _ = ac(oracle, "oracle_computeAlartCurnierSTD", R[k], U[k], MU[k], RHO[k])
            _, A2, _ = ac(oracle, "oracle_alart_curnier_dual", R[k], U[k], MU[k], RHO[k])
            worst = max(worst, np.abs(A1 - A2).max())
    finally:
        oracle.mutation(0)
    assert worst > 1e-3


def test_jeanmoreau_is_the_natural_map_with_radius_mu_rn(oracle):
    """Christensen-Pang variant: F_t = r_t - P_{D(mu r_n)}(r_t - rho u_t)."""
    rng = np.random.default_rng(8)
    for _ in range(300):
        R = rng.normal(size=3)
        R[0] = abs(R[0]) + 0.1
        U = rng.normal(size=3)
        mu = rng.uniform(0.1, 0.9)
        rho = np.array([1.3, 0.7, 0.7])
        F, _, _ = ac(oracle, "oracle_computeAlartCurnierJeanMoreau", R, U, mu, rho)
        xn = R[0] - rho[0] * U[0]
        fn = rho[0] * U[0] if xn >= 0 else R[0]
        xt = R[1:] - rho[1] * U[1:]
        rad = mu * R[0]
        nx = np.linalg.norm(xt)
        pt = xt if (nx < rad or nx < 1e-20) else rad * xt / nx
        np.testing.assert_allclose(F, [fn, *(R[1:] - pt)], rtol=1e-12, atol=1e-13)


def test_rho_split_spectral_norm_closed_form(oracle):
    """For W = diag(2, 7, 7): rho_n = 1/2, rho_t = 2/(14 + sqrt(196 - 56))
    (the reference's quirk dw = sw^2 - 4 (sw - W21 + W12))."""
    L = oracle.load()
    W = np.diag([2.0, 7.0, 7.0]).reshape(-1, order="F")
    rho = np.zeros(3)
    L.oracle_rho_split_spectral_norm(oracle.d(W), oracle.d(rho))
    rt = 2.0 / (14.0 + np.sqrt(196.0 - 56.0))
    np.testing.assert_allclose(rho, [0.5, rt, rt], rtol=1e-15)
    L.oracle_rho_split_spectral_norm_cond(oracle.d(W), oracle.d(rho))
    sw, dw = 14.0, np.sqrt(140.0)
    np.testing.assert_allclose(rho[1], 2 * (sw - dw) / (sw + dw) ** 2, rtol=1e-15)
    # the quirk: an antisymmetric tangential coupling changes dw
    Wq = np.array([[2.0, 0, 0], [0, 7.0, 1.0], [0, -1.0, 7.0]]).reshape(-1, order="F")
    L.oracle_rho_split_spectral_norm(oracle.d(Wq), oracle.d(rho))
    dwq = np.sqrt(14.0 ** 2 - 4 * (14.0 - (-1.0) + 1.0))
    np.testing.assert_allclose(rho[1], 2.0 / (14.0 + dwq), rtol=1e-15)
