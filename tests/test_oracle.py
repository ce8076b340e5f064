"""Pins the CPU oracle (oracle/hes_oracle.cpp) without the reference.

The reference (Julia) cannot run in this image, so every function of the oracle
is checked against something its own code does not share: numpy/LAPACK for the
3x3 algebra and eigenvalues, closed forms derived independently of the dual
number path, finite differences, exact physical identities and the reference's
own known answers (uniform states, identical phases at a = 1/2).
"""
import math

import numpy as np
import pytest

from oracle_bind import rel_err_per_component

MAT = np.array([8.93, 4.6, 3.9e-4, 300.0, 2.1, 1.0, 3.0, 2.0])       # Barton2009() (EquationsOfState.jl:90)
MAT_ALT = np.array([8.93, 6.22, 9.0e-4, 300.0, 3.16, 1.0, 3.577, 2.088])  # main.jl:133 alternative
MAT2 = np.concatenate([MAT, [2.7, 4.6, 3.9e-4, 300.0, 3.1, 1.0, 3.0, 2.0]])


def vec(M):
    return np.asarray(M, dtype=float).reshape(-1, order="F").copy()


def mat(v):
    return np.asarray(v, dtype=float).reshape(3, 3, order="F")


def rand_A(rng, scale=0.05):
    while True:
        A = np.eye(3) + scale * rng.standard_normal((3, 3))
        if np.linalg.det(A) > 0.3:
            return A


def rot(rng):
    q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    return q * np.sign(np.linalg.det(q))


# --------------------------------------------------------------------------- SimpleLA / Strains
def test_inv_det_against_numpy(ora):
    rng = np.random.default_rng(0)
    for _ in range(50):
        M = rng.standard_normal((3, 3))
        np.testing.assert_allclose(mat(ora.inv(vec(M))), np.linalg.inv(M), rtol=1e-12, atol=1e-12)
        assert abs(ora.det(vec(M)) - np.linalg.det(M)) <= 1e-13 * max(1, abs(np.linalg.det(M)))
        assert abs(ora.det_lu(vec(M)) - np.linalg.det(M)) <= 1e-13 * max(1, abs(np.linalg.det(M)))


def test_finger_and_invariants(ora):
    rng = np.random.default_rng(1)
    for _ in range(30):
        A = rand_A(rng, 0.2)
        G = mat(ora.finger(vec(A)))
        np.testing.assert_allclose(G, np.linalg.inv(A @ A.T), rtol=1e-12, atol=1e-13)
        lam = np.linalg.eigvalsh(G)  # invariants are the elementary symmetric functions
        i = ora.invariants(vec(G))
        np.testing.assert_allclose(i, [lam.sum(), lam[0] * lam[1] + lam[1] * lam[2] + lam[0] * lam[2],
                                       lam.prod()], rtol=1e-12)


# --------------------------------------------------------------------------- energy / entropy
def _energy_eig(m, S, G):
    """Energy written through the eigenvalues of G: I1^2/3 - I2 = 1/6 sum_{i<j}(l_i - l_j)^2."""
    rho0, c0, cv, t0, b0, al, be, ga = m
    k0 = c0 ** 2 - 4.0 / 3.0 * b0 ** 2
    lam = np.linalg.eigvalsh(G)
    i3 = lam.prod()
    dev = ((lam[0] - lam[1]) ** 2 + (lam[1] - lam[2]) ** 2 + (lam[0] - lam[2]) ** 2) / 6.0
    return (0.5 * k0 / al ** 2 * (i3 ** (al / 2) - 1) ** 2 + cv * t0 * i3 ** (ga / 2) * (math.exp(S / cv) - 1)
            + 0.5 * b0 ** 2 * i3 ** (be / 2) * dev)


@pytest.mark.parametrize("m", [MAT, MAT_ALT])
def test_energy_known_values(ora, m):
    assert ora.energy(m, 0.0, vec(np.eye(3))) == 0.0  # undeformed, S = 0
    rng = np.random.default_rng(2)
    for _ in range(30):
        A = rand_A(rng, 0.1)
        S = rng.uniform(-1e-3, 2e-3)
        G = np.linalg.inv(A @ A.T)
        e = ora.energy(m, S, vec(G))
        assert abs(e - _energy_eig(m, S, G)) <= 1e-11 * max(1.0, abs(e))
        # rotation invariance of an isotropic energy
        R = rot(rng)
        assert abs(ora.energy(m, S, vec(R @ G @ R.T)) - e) <= 1e-11 * max(1.0, abs(e))


def test_entropy_inverts_energy_and_clamps(ora):
    rng = np.random.default_rng(3)
    for _ in range(30):
        G = np.linalg.inv(rand_A(rng, 0.1) @ rand_A(rng, 0.0).T)
        S = rng.uniform(-1e-3, 3e-3)
        e = ora.energy(MAT, S, vec(G))
        assert abs(ora.entropy(MAT, e, vec(G)) - S) <= 1e-9 * max(1e-3, abs(S))
    # far below the cold curve the logarithm's argument is clamped at 1e-6
    G = vec(np.eye(3))
    assert ora.entropy(MAT, -1.0, G) == pytest.approx(MAT[2] * math.log(1e-6), rel=1e-15)


def test_temperature_is_de_dS(ora):
    rng = np.random.default_rng(4)
    for _ in range(10):
        G = vec(np.linalg.inv(rand_A(rng) @ rand_A(rng, 0.0).T))
        S, h = 1e-3, 1e-8
        fd = (ora.energy(MAT, S + h, G) - ora.energy(MAT, S - h, G)) / (2 * h)
        assert abs(ora.temperature(MAT, S, G) - fd) <= 1e-6 * abs(fd)


# --------------------------------------------------------------------------- stress
def _stress_closed(m, S, A):
    """sigma = -2 rho [kappa (G^2 - I1/3 G) + c3 I], derived by hand from the energy."""
    rho0, c0, cv, t0, b0, al, be, ga = m
    k0 = c0 ** 2 - 4.0 / 3.0 * b0 ** 2
    rho = rho0 / np.linalg.det(A)
    G = np.linalg.inv(A @ A.T)
    i1 = np.trace(G)
    i2 = 0.5 * (i1 ** 2 - np.trace(G @ G))
    i3 = np.linalg.det(G)
    sa, sb, sg = i3 ** (al / 2), i3 ** (be / 2), i3 ** (ga / 2)
    W = 0.5 * b0 ** 2 * sb * (i1 ** 2 / 3 - i2)
    Eth = cv * t0 * sg * (math.exp(S / cv) - 1)
    c3 = k0 / (2 * al) * (sa - 1) * sa + ga / 2 * Eth + be / 2 * W
    return -2 * rho * (0.5 * b0 ** 2 * sb * (G @ G - i1 / 3 * G) + c3 * np.eye(3))


@pytest.mark.parametrize("m", [MAT, MAT_ALT])
def test_stress(ora, m):
    assert np.all(ora.stress(m, 0.0, vec(np.eye(3))) == 0.0)  # sigma(I, 0) = 0
    rng = np.random.default_rng(5)
    for _ in range(20):
        A = rand_A(rng)
        S = rng.uniform(0, 2e-3)
        sig = mat(ora.stress(m, S, vec(A)))
        ref = _stress_closed(m, S, A)
        np.testing.assert_allclose(sig, ref, rtol=1e-11, atol=1e-11 * np.abs(ref).max())
        np.testing.assert_allclose(sig, sig.T, atol=1e-12 * np.abs(sig).max())
        R = rot(rng)  # objectivity: sigma(R A) = R sigma(A) R^T
        np.testing.assert_allclose(mat(ora.stress(m, S, vec(R @ A))), R @ sig @ R.T,
                                   atol=1e-11 * np.abs(sig).max())


def test_stress_is_gradient_of_energy(ora):
    """-2 rho G dE/dG with dE/dG by central differences of the oracle's energy."""
    rng = np.random.default_rng(6)
    A = rand_A(rng)
    S = 1e-3
    G = np.linalg.inv(A @ A.T)
    h = 1e-6
    D = np.zeros((3, 3))
    for k in range(3):
        for l in range(3):
            E = np.zeros((3, 3))
            E[k, l] = h
            D[k, l] = (ora.energy(MAT, S, vec(G + E)) - ora.energy(MAT, S, vec(G - E))) / (2 * h)
    ref = -2 * (MAT[0] / np.linalg.det(A)) * G @ D
    np.testing.assert_allclose(mat(ora.stress(MAT, S, vec(A))), ref, atol=1e-7 * np.abs(ref).max())


def test_stress_den_form(ora):
    """stress(eos, den, e_int, F) recomputes S from e_int and uses the given density."""
    rng = np.random.default_rng(7)
    A = rand_A(rng)
    S = 1.5e-3
    G = vec(np.linalg.inv(A @ A.T))
    e = ora.energy(MAT, S, G)
    den = 3.3
    s1 = ora.stress_den(MAT, den, e, vec(A))
    s2 = ora.stress(MAT, S, vec(A)) * (den / (MAT[0] / np.linalg.det(A)))
    np.testing.assert_allclose(s1, s2, rtol=1e-9, atol=1e-12)


# --------------------------------------------------------------------------- acoustic tensor / eigenvalues
def test_acoustic_at_identity(ora):
    ac = ora.acoustic(MAT, 0.0, vec(np.eye(3)))
    np.testing.assert_allclose(ac, np.diag([MAT[1] ** 2, MAT[4] ** 2, MAT[4] ** 2]), atol=1e-12)


@pytest.mark.parametrize("n", [(1.0, 0.0, 0.0), (0.3, -0.5, 0.81)])
def test_acoustic_is_stress_derivative(ora, n):
    """Omega_ij = (1/rho) sum_m n_m d sigma(m,i) / dA along dA = e_j (A^T n)^T, by central differences."""
    rng = np.random.default_rng(8)
    n = np.asarray(n)
    for _ in range(5):
        A = rand_A(rng)
        S = 1e-3
        rho = MAT[0] / np.linalg.det(A)
        h = 1e-6
        ref = np.zeros((3, 3))
        for j in range(3):
            dA = np.zeros((3, 3))
            dA[j, :] = A.T @ n
            sp = mat(ora.stress(MAT, S, vec(A + h * dA)))
            sm = mat(ora.stress(MAT, S, vec(A - h * dA)))
            ref[:, j] = n @ (sp - sm) / (2 * h) / rho
        ac = ora.acoustic(MAT, S, vec(A), n)
        np.testing.assert_allclose(ac, ref, atol=2e-7 * np.abs(ref).max())
        np.testing.assert_allclose(ac, ac.T, atol=1e-12 * np.abs(ac).max())


def test_acoustic_prim_is_fixed_energy_derivative(ora):
    """acoustic(eos, P, n): derivative of stress(eos, den, e_int, F) at fixed den and e_int."""
    rng = np.random.default_rng(9)
    A = rand_A(rng)
    S = 1e-3
    G = vec(np.linalg.inv(A @ A.T))
    e = ora.energy(MAT, S, G)
    den = 0.7 * 9.1
    P = np.concatenate([[0.7, 9.1, 0.1, 0.2, 0.3, S], vec(A)])
    h = 1e-6
    ref = np.zeros((3, 3))
    for j in range(3):
        dA = np.zeros((3, 3))
        dA[j, :] = A[0, :]
        sp = mat(ora.stress_den(MAT, den, e, vec(A + h * dA)))
        sm = mat(ora.stress_den(MAT, den, e, vec(A - h * dA)))
        ref[:, j] = (sp - sm)[0, :] / (2 * h) / den
    np.testing.assert_allclose(ora.acoustic_prim(MAT, P), ref, atol=2e-7 * np.abs(ref).max())


def test_eigvals_against_lapack(ora):
    rng = np.random.default_rng(10)
    for _ in range(50):
        M = rng.standard_normal((3, 3))
        M = M + M.T
        np.testing.assert_allclose(np.sort(ora.eigvals3(M)), np.sort(np.linalg.eigvals(M).real),
                                   rtol=1e-12, atol=1e-12 * np.abs(M).max())
    for _ in range(20):  # the acoustic tensors themselves (non-symmetric at round-off)
        ac = ora.acoustic(MAT, 1e-3, vec(rand_A(rng, 0.1)))
        np.testing.assert_allclose(np.sort(ora.eigvals3(ac)), np.sort(np.linalg.eigvals(ac).real),
                                   rtol=1e-12)


# --------------------------------------------------------------------------- model functions
def _rand_prims(rng, nph, mats):
    out = []
    for p in range(nph):
        A = rand_A(rng)
        blk = [mats[8 * p] / np.linalg.det(A), *rng.uniform(-0.5, 0.5, 3), rng.uniform(0, 2e-3), *vec(A)]
        out += ([rng.uniform(0.05, 0.95)] if nph == 2 else []) + blk
    if nph == 2:
        out[15] = 1 - out[0]
    return np.array(out)


@pytest.mark.parametrize("nph", [1, 2])
def test_prim_cons_roundtrip_and_rho_ignored(ora, nph):
    rng = np.random.default_rng(11)
    mats = MAT if nph == 1 else MAT2
    for _ in range(20):
        P = _rand_prims(rng, nph, mats)
        Q = ora.prim2cons(nph, mats, P)
        np.testing.assert_allclose(ora.cons2prim(nph, mats, Q), P, rtol=1e-10, atol=1e-13)
        # the conserved a*rho is never read (HyperelasticityMPh.jl:110-114)
        Q2 = Q.copy()
        Q2[0 if nph == 1 else 1] *= 1.7
        assert np.array_equal(ora.cons2prim(nph, mats, Q2)[1:], ora.cons2prim(nph, mats, Q)[1:])
        assert np.array_equal(ora.flux(nph, mats, Q2), ora.flux(nph, mats, Q))


def test_flux_at_rest_is_minus_stress(ora):
    rng = np.random.default_rng(12)
    A = rand_A(rng)
    P = np.concatenate([[MAT[0] / np.linalg.det(A), 0, 0, 0, 1e-3], vec(A)])
    F = ora.flux(1, MAT, ora.prim2cons(1, MAT, P))
    sig = mat(ora.stress(MAT, 1e-3, vec(A)))
    np.testing.assert_allclose(F[1:4], -sig[0, :], rtol=1e-10)
    assert F[0] == 0 and F[4] == 0 and np.all(F[5:] == 0)


def test_flux_jacobian_spectrum_is_the_characteristic_speeds(ora):
    """The 1D system Q_t + F(Q)_x = 0 is hyperbolic with characteristic speeds
    u1 +- c_k (c_k^2 the eigenvalues of the acoustic tensor), u1 for the
    advected components and 0 for the three first-row entries of rho A, whose
    flux vanishes identically.  So the eigenvalues of dF/dQ (central
    differences of flux, HyperelasticityMPh.jl:146-175) must be exactly the
    six values get_eigvals returns (HyperelasticityMPh.jl:252-266), u1 four
    times and 0 four times: a dropped or misplaced flux term moves them."""
    rng = np.random.default_rng(31)
    for _ in range(4):
        Q = ora.prim2cons(1, MAT, _rand_prims(rng, 1, MAT))
        n = Q.size
        J = np.zeros((n, n))
        for k in range(n):
            h = 1e-6 * max(abs(Q[k]), 1e-3)
            Qp, Qm = Q.copy(), Q.copy()
            Qp[k] += h
            Qm[k] -= h
            J[:, k] = (ora.flux(1, MAT, Qp) - ora.flux(1, MAT, Qm)) / (2 * h)
        got = np.sort(np.linalg.eigvals(J).real)
        u1 = ora.cons2prim(1, MAT, Q)[1]
        want = np.sort(np.concatenate([ora.get_eigvals(1, MAT, Q), [u1] * 4, [0.0] * 4]))
        assert np.abs(got - want).max() <= 1e-7 * np.abs(want).max()


def test_interface_exchange_of_one_rusanov_step(ora):
    """Two pieces whose phases differ only in volume fraction: along the path
    psi(s) = Q_l (1-s) + Q_r s every primitive is constant, so B(psi) is, and
    one Rusanov step (D- = D+ = 1/2 int B dpsi, main.jl:59) changes the sums
    over all cells by exactly -dt/dx [F(Q_r) - F(Q_l) + B da].  B's column
    (HyperelasticityMPh.jl:204-226, omega = 0, k = 1/2, beta = 0) is built here
    from the separately pinned stress and temperature: u_I = (u1 + u2)/2,
    sigma_I = (th2 sigma1 + th1 sigma2)/(th1 + th2), sigma_I(:,1).u_I, and the
    distortion rows rho A_1c u + rho (A^T (u_I - u))_c.  Swapping the
    temperature weights, or taking u_I = u1 or sigma_I = sigma1, misses by a
    factor of 2 or more."""
    rng = np.random.default_rng(41)
    P = _rand_prims(rng, 2, MAT2)
    PL, PR = P.copy(), P.copy()
    PL[0], PL[15] = 0.3, 0.7
    PR[0], PR[15] = 0.8, 0.2
    QL, QR = ora.prim2cons(2, MAT2, PL), ora.prim2cons(2, MAT2, PR)
    nx = 20
    Q0 = np.asfortranarray(np.stack([QL if i < nx // 2 else QR for i in range(nx)], axis=1))
    Q1, t, s, dt = ora.run(2, MAT2, 0, Q0, 0.6, 1.0 / nx, steps=1)
    dtdx = dt[0] * nx
    pr = ora.cons2prim(2, MAT2, QL)
    sig, th, u, rho, A = [], [], [], [], []
    for p in range(2):
        m = MAT2[8 * p:8 * p + 8]
        b = pr[15 * p:15 * p + 15]
        sig.append(mat(ora.stress(m, b[5], b[6:15])))
        th.append(ora.temperature(m, b[5], ora.finger(b[6:15])))
        u.append(b[2:5])
        rho.append(b[1])
        A.append(mat(b[6:15]))
    uI = 0.5 * u[0] + 0.5 * u[1]
    sI = (th[1] * sig[0] + th[0] * sig[1]) / (th[0] + th[1])
    FL, FR = ora.flux(2, MAT2, QL), ora.flux(2, MAT2, QR)
    dS = Q1.sum(axis=1) - Q0.sum(axis=1)
    for p in range(2):
        da = PR[15 * p] - PL[15 * p]
        Bc = np.zeros(15)
        Bc[0] = uI[0]
        Bc[2:5] = sI[:, 0]
        Bc[5] = sI[:, 0] @ uI
        atdv = A[p].T @ (uI - u[p])
        for c in range(3):
            Bc[6 + 3 * c:9 + 3 * c] = rho[p] * A[p][0, c] * u[p]
            Bc[6 + 3 * c] += rho[p] * atdv[c]
        o = 15 * p
        pred = -dtdx * (FR[o:o + 15] - FL[o:o + 15] + Bc * da)
        scale = np.abs(pred) + 1e-12 * np.abs(pred).max()
        assert np.all(np.abs(dS[o:o + 15] - pred) <= 1e-11 * scale), (p, dS[o:o + 15] - pred)


def test_noncons_structure(ora):
    rng = np.random.default_rng(13)
    Q = ora.prim2cons(2, MAT2, _rand_prims(rng, 2, MAT2))
    B = ora.noncons_flux(MAT2, Q)
    cols = [c for c in range(30) if np.any(B[:, c] != 0)]
    assert cols == [0, 15]  # one non-zero column per phase block (the a_p column)
    assert np.all(B[15:, 0] == 0) and np.all(B[:15, 15] == 0)
    assert B[1, 0] == 0 and B[16, 15] == 0
    # interface velocity/stress rows are common to both phases, so with
    # da_2 = -da_1 the mixture momentum and energy contributions cancel
    np.testing.assert_array_equal(B[[0, 2, 3, 4, 5], 0], B[[15, 17, 18, 19, 20], 15])


def test_fluxes_of_equal_states(ora):
    rng = np.random.default_rng(14)
    for nph, mats in ((1, MAT), (2, MAT2)):
        Q = ora.prim2cons(nph, mats, _rand_prims(rng, nph, mats))
        c, dm, dp = ora.lxf(nph, mats, Q, Q, 7.0)
        np.testing.assert_array_equal(c, ora.flux(nph, mats, Q))
        assert np.all(dm == 0) and np.all(dp == 0)
        e = ora.get_eigvals(nph, mats, Q)
        c, dm, dp = ora.hll(nph, mats, Q, Q, e, e)
        assert np.all(c == 0) and np.all(dm == 0) and np.all(dp == 0)


def test_hll_reduces_to_textbook_hll(ora):
    """With no non-conservative term, D- = F_hll - F_l and D+ = F_r - F_hll with the
    Harten-Lax-van Leer flux F_hll = (s_r F_l - s_l F_r + s_l s_r (Q_r - Q_l)) / (s_r - s_l)."""
    rng = np.random.default_rng(15)
    for _ in range(10):
        Ql = ora.prim2cons(1, MAT, _rand_prims(rng, 1, MAT))
        Qr = ora.prim2cons(1, MAT, _rand_prims(rng, 1, MAT))
        el, er = ora.get_eigvals(1, MAT, Ql), ora.get_eigvals(1, MAT, Qr)
        em = ora.get_eigvals(1, MAT, 0.5 * (Ql + Qr))
        sl = min(0.0, em.min(), el.min())
        sr = max(0.0, em.max(), er.max())
        Fl, Fr = ora.flux(1, MAT, Ql), ora.flux(1, MAT, Qr)
        Fh = (sr * Fl - sl * Fr + sl * sr * (Qr - Ql)) / (sr - sl)
        _, dm, dp = ora.hll(1, MAT, Ql, Qr, el, er)
        scale = np.abs(Fl).max() + np.abs(Fr).max()
        np.testing.assert_allclose(dm, Fh - Fl, atol=1e-13 * scale)
        np.testing.assert_allclose(dp, Fr - Fh, atol=1e-13 * scale)


def _bcols(ora, mats, Q):
    """The two non-zero columns of noncons_flux(Q) (HyperelasticityMPh.jl:178-231,
    omega = 0, k = 1/2, beta = 0), as a (2, 15) array, built from the separately
    pinned cons2prim, stress and temperature: u_I = (u1 + u2)/2,
    sigma_I = (th2 sigma1 + th1 sigma2)/(th1 + th2), sigma_I(:,1).u_I and the
    distortion rows rho A_1c u + rho (A^T (u_I - u))_c."""
    pr = ora.cons2prim(2, mats, Q)
    sig, th, u, rho, A = [], [], [], [], []
    for p in range(2):
        m = mats[8 * p:8 * p + 8]
        b = pr[15 * p:15 * p + 15]
        sig.append(mat(ora.stress(m, b[5], b[6:15])))
        th.append(ora.temperature(m, b[5], ora.finger(b[6:15])))
        u.append(b[2:5])
        rho.append(b[1])
        A.append(mat(b[6:15]))
    uI = 0.5 * u[0] + 0.5 * u[1]
    sI = (th[1] * sig[0] + th[0] * sig[1]) / (th[0] + th[1])
    cols = np.zeros((2, 15))
    for p in range(2):
        cols[p, 0] = uI[0]
        cols[p, 2:5] = sI[:, 0]
        cols[p, 5] = sI[:, 0] @ uI
        atdv = A[p].T @ (uI - u[p])
        for c in range(3):
            cols[p, 6 + 3 * c:9 + 3 * c] = rho[p] * A[p][0, c] * u[p]
            cols[p, 6 + 3 * c] += rho[p] * atdv[c]
    return cols


def _b_int(ora, mats, Qa, Qb):
    """Gauss-Legendre-6 integral of B(psi) dpsi along psi(s) = Qa (1 - s) + Qb s on
    [0, 1] (NumFluxes.jl:94-107), nodes and weights from numpy's leggauss."""
    x, w = np.polynomial.legendre.leggauss(6)
    out = np.zeros(30)
    for xi, wi in zip((x + 1.0) / 2.0, w / 2.0):
        cols = _bcols(ora, mats, Qa * (1.0 - xi) + Qb * xi)
        for p in range(2):
            out[15 * p:15 * p + 15] += wi * cols[p] * (Qb[15 * p] - Qa[15 * p])
    return out


def test_two_phase_hll_across_a_volume_fraction_jump(ora):
    """Two-phase HLL where the volume fractions jump and every other primitive
    differs across the face too, so B varies along the path (NumFluxes.jl:85-132,
    the flux main.jl:225 runs on test case 6).  D-, D+ are rebuilt here from
    pinned pieces: the flux, B's column from the pinned stress and temperature,
    Gauss-Legendre nodes from numpy, the speed bounds from the pinned
    eigenvalues, Q_hll = (Q_r s_r - Q_l s_l - (B_int(Q_l, Q_r) + F_r - F_l)) /
    (s_r - s_l) (:111), and the two-segment integral
    B_int(Q_l, Q_hll) + B_int(Q_hll, Q_r) (:128-129).  The oracle must agree to
    1e-12, and D- + D+ must equal F_r - F_l plus the integral along
    l -> hll -> r.  The faces are chosen so that the two-segment integral differs
    from the straight one by more than 1e-3 relative; the test shows that the
    plausible mistakes -- B_int(Q_l, Q_r) in place of the two-segment sum,
    s_l and s_r swapped in Q_hll, Q_hll without its path term -- each miss the
    oracle by far more than the tolerance."""
    rng = np.random.default_rng(61)
    for case in range(6):
        PL, PR = _rand_prims(rng, 2, MAT2), _rand_prims(rng, 2, MAT2)
        PL[0] = rng.uniform(0.15, 0.45)
        PR[0] = rng.uniform(0.55, 0.85)
        PL[15], PR[15] = 1.0 - PL[0], 1.0 - PR[0]
        QL, QR = ora.prim2cons(2, MAT2, PL), ora.prim2cons(2, MAT2, PR)
        el, er = ora.get_eigvals(2, MAT2, QL), ora.get_eigvals(2, MAT2, QR)
        em = ora.get_eigvals(2, MAT2, 0.5 * (QL + QR))
        sl = min(0.0, em.min(), el.min())
        sr = max(0.0, em.max(), er.max())
        FL, FR = ora.flux(2, MAT2, QL), ora.flux(2, MAT2, QR)

        def build(sl_h, sr_h, with_path=True, two_segments=True):
            pint = _b_int(ora, MAT2, QL, QR) if with_path else 0.0
            Qh = (QR * sr_h - QL * sl_h - (pint + FR - FL)) / (sr_h - sl_h)
            if two_segments:
                I = _b_int(ora, MAT2, QL, Qh) + _b_int(ora, MAT2, Qh, QR)
            else:
                I = _b_int(ora, MAT2, QL, QR)
            J = FR - FL + I
            d = sr - sl
            dm = -sl / d * J + sl * sr / d * (QR - QL)
            dp = sr / d * J - sl * sr / d * (QR - QL)
            return dm, dp, J, I

        dm, dp, J, I2 = build(sl, sr)
        cons, odm, odp = ora.hll(2, MAT2, QL, QR, el, er)
        assert np.all(cons == 0.0)  # the conservative part is zeros(30) (NumFluxes.jl:82)
        scale = max(np.abs(dm).max(), np.abs(dp).max())
        tol = 1e-12 * scale
        assert np.abs(odm - dm).max() <= tol, (case, np.abs(odm - dm).max() / scale)
        assert np.abs(odp - dp).max() <= tol, (case, np.abs(odp - dp).max() / scale)
        assert np.abs((odm + odp) - J).max() <= tol, case
        # the two-segment path matters at these faces
        I1 = _b_int(ora, MAT2, QL, QR)
        assert np.abs(I2 - I1).max() >= 1e-3 * np.abs(I2).max(), case
        # planted mistakes miss by far more than the tolerance
        for wrong in (build(sl, sr, two_segments=False), build(sr, sl), build(sl, sr, with_path=False)):
            miss = max(np.abs(odm - wrong[0]).max(), np.abs(odp - wrong[1]).max())
            assert miss >= 1e3 * tol, (case, miss / scale)


def test_quadrature_rules(ora):
    x, w = ora.quadrature(0)
    xr, wr = np.polynomial.legendre.leggauss(6)
    np.testing.assert_allclose(x, xr, atol=2e-16)
    np.testing.assert_allclose(w, wr, atol=2e-16)
    import mpmath
    mpmath.mp.dps = 40
    s7 = mpmath.sqrt(7)
    lo = [-1, -mpmath.sqrt(mpmath.mpf(1) / 3 + 2 * s7 / 21), -mpmath.sqrt(mpmath.mpf(1) / 3 - 2 * s7 / 21)]
    lo = lo + [-v for v in reversed(lo)]
    lw = [mpmath.mpf(1) / 15, (14 - s7) / 30, (14 + s7) / 30]
    lw = lw + list(reversed(lw))
    x, w = ora.quadrature(1)
    np.testing.assert_allclose(x, [float(v) for v in lo], atol=2e-16)
    np.testing.assert_allclose(w, [float(v) for v in lw], atol=2e-16)
    # Lobatto-6 integrates polynomials up to degree 9 exactly
    assert abs(sum(wi * xi ** 8 for xi, wi in zip(x, w)) - 2 / 9) < 1e-15


# --------------------------------------------------------------------------- time loop
def _riemann(ora, nph, mats, nx, seed):
    rng = np.random.default_rng(seed)
    Ql = ora.prim2cons(nph, mats, _rand_prims(rng, nph, mats))
    Qr = ora.prim2cons(nph, mats, _rand_prims(rng, nph, mats))
    return np.asfortranarray(np.stack([Ql if i < nx / 2 else Qr for i in range(nx)], axis=1))


@pytest.mark.parametrize("nph,flux", [(1, 0), (1, 1), (1, 2), (2, 0), (2, 1), (2, 2)])
def test_uniform_state_stays_bit_identical(ora, nph, flux):
    mats = MAT if nph == 1 else MAT2
    rng = np.random.default_rng(16)
    Q = ora.prim2cons(nph, mats, _rand_prims(rng, nph, mats))
    Q0 = np.asfortranarray(np.repeat(Q[:, None], 20, axis=1))
    Q1, t, s, dt = ora.run(nph, mats, flux, Q0, 0.6, 0.05, steps=5)
    assert s == 5 and np.array_equal(Q1, Q0)


@pytest.mark.parametrize("flux", [0, 1, 2])
def test_identical_phases_at_half_are_half_the_single_phase_run(ora, flux):
    """HyperelasticityMPh test cases 4/5 replicate the single-phase cases with two
    identical phases at a = 1/2: every non-conservative term is multiplied by a
    zero volume-fraction jump and everything else scales by 1/2 exactly."""
    nx = 40
    Q1 = _riemann(ora, 1, MAT, nx, seed=17)
    Q2 = np.zeros((30, nx), order="F")
    for p in range(2):
        Q2[15 * p] = 0.5
        Q2[15 * p + 1:15 * p + 15] = 0.5 * Q1
    mats2 = np.concatenate([MAT, MAT])
    R1, t1, s1, d1 = ora.run(1, MAT, flux, Q1, 0.6, 1.0 / nx, steps=8)
    R2, t2, s2, d2 = ora.run(2, mats2, flux, Q2, 0.6, 1.0 / nx, steps=8)
    assert np.array_equal(d1, d2)
    for p in range(2):
        assert np.all(R2[15 * p] == 0.5)
        assert np.array_equal(R2[15 * p + 1:15 * p + 15], 0.5 * R1)


def test_lax_friedrichs_is_the_textbook_scheme(ora):
    """With lambda = dx / dt on every face (main.jl:30-41, 50) and no path
    term, the update is the Lax-Friedrichs scheme in its textbook form
    Q_i <- (Q_i-1 + Q_i+1) / 2 - dt / (2 dx) (F_i+1 - F_i-1)."""
    nx = 32
    Q0 = _riemann(ora, 1, MAT, nx, seed=21)
    Q0[:, 9:13] = ora.prim2cons(1, MAT, _rand_prims(np.random.default_rng(22), 1, MAT))[:, None]
    Q1, t, s, dt = ora.run(1, MAT, 2, Q0, 0.6, 1.0 / nx, steps=1)
    dtdx = dt[0] * nx
    F = ora.flux(1, MAT, Q0)
    ref = 0.5 * (Q0[:, :-2] + Q0[:, 2:]) - 0.5 * dtdx * (F[:, 2:] - F[:, :-2])
    scale = np.abs(Q0).max(axis=1) + dtdx * np.abs(F).max(axis=1)
    assert np.all(np.abs(Q1[:, 1:-1] - ref) <= 1e-14 * scale[:, None])
    assert np.array_equal(Q1[:, [0, -1]], Q0[:, [0, -1]])  # frozen ends (main.jl:219-220)
    # and it differs from the Rusanov step (a face speed below dx / dt)
    Qr, _, _, dtr = ora.run(1, MAT, 0, Q0, 0.6, 1.0 / nx, steps=1)
    assert dtr[0] == dt[0] and not np.array_equal(Qr, Q1)


def test_lax_friedrichs_and_rusanov_differ_by_the_face_speed(ora):
    """Two phases: the path integrals do not depend on lambda, so one step of
    each differs exactly by the viscosity term with the two face speeds,
    dt/dx [(L - lam_r) (Q_i+1 - Q_i) - (L - lam_l) (Q_i - Q_i-1)] / 2 with
    L = dx / dt and lam the max |eigenvalue| over the face's two cells."""
    nx = 24
    Q0 = _riemann(ora, 2, MAT2, nx, seed=23)
    Q0[:, 5:8] = ora.prim2cons(2, MAT2, _rand_prims(np.random.default_rng(24), 2, MAT2))[:, None]
    QL, _, _, dt = ora.run(2, MAT2, 2, Q0, 0.6, 1.0 / nx, steps=1)
    QR, _, _, dtr = ora.run(2, MAT2, 0, Q0, 0.6, 1.0 / nx, steps=1)
    assert dt[0] == dtr[0]
    dtdx = dt[0] * nx
    lam = np.abs(ora.get_eigvals(2, MAT2, Q0)).max(axis=0)
    lf = np.maximum(lam[:-1], lam[1:])              # face i+1/2, i = 0..nx-2
    dQ = Q0[:, 1:] - Q0[:, :-1]
    L = 1.0 / dtdx
    vis = (L - lf) * dQ                              # per face
    ref = 0.5 * dtdx * (vis[:, 1:] - vis[:, :-1])    # interior cells 1..nx-2
    scale = np.abs(Q0).max(axis=1)
    assert np.all(np.abs((QL - QR)[:, 1:-1] - ref) <= 1e-13 * scale[:, None])


def test_conservation_and_t_end(ora):
    nx = 60
    Q0 = _riemann(ora, 1, MAT, nx, seed=18)
    Q0[:, -5:] = Q0[:, :5]  # same state at both frozen ends
    Q1, t, s, dt = ora.run(1, MAT, 1, Q0, 0.6, 1.0 / nx, t_end=0.002)
    assert t >= 0.002 and t - dt[-1] < 0.002  # `while t < T` (main.jl:202)
    assert t == pytest.approx(dt.sum(), rel=1e-14)
    err = np.abs(Q1.sum(axis=1) - Q0.sum(axis=1)) / np.abs(Q0).sum(axis=1)
    assert np.all(err < 1e-14)


def test_oracle_reproduces_golden_fixtures(ora):
    """The committed fixtures were written by tests/golden/make_golden.py."""
    import os
    here = os.path.join(os.path.dirname(__file__), "golden")
    g = np.load(os.path.join(here, "testcase6_hll.npz"))
    Q, t, s, dt = ora.run(2, g["mat"], 1, g["Q0"], 0.6, 1.0 / g["Q0"].shape[1], steps=int(g["steps"]))
    assert np.array_equal(Q, g["Q"]) and t == float(g["t"])
    g = np.load(os.path.join(here, "config1_rusanov.npz"))
    Q, t, s, dt = ora.run(1, g["mat"], 0, g["Q0"], 0.6, 1.0 / g["Q0"].shape[1], t_end=0.06)
    assert s == int(g["steps"]) and np.array_equal(Q, g["Q"])


def _dual_case(ora, n, steps, seed=13):
    """Pieces, constants and a target for the dual checks (the target: the
    oracle's run with b0 x 1.05 and c0 x 0.97)."""
    from golden.make_adjoint_dual import oracle_profiles, pieces
    mats, prims = pieces(seed)
    mt = mats.copy()
    mt[4] *= 1.05
    mt[1] *= 0.97
    rho_t, u_t = oracle_profiles(ora, mt, prims, n, steps)
    return mats, prims, rho_t, u_t


def _theta_loss(ora, theta, n, steps, rho_t, u_t):
    """The misfit as a function of the 28 scalars, through the double oracle."""
    from golden.make_adjoint_dual import oracle_profiles
    prims = np.asfortranarray(np.stack([theta[:13], theta[13:26]], axis=1))
    mats = MAT.copy()
    mats[4], mats[1] = theta[26], theta[27]
    rho, u1 = oracle_profiles(ora, mats, prims, n, steps)
    return float(np.sum((rho - rho_t) ** 2) + np.sum((u1 - u_t) ** 2))


def test_dual_gradient_matches_central_differences(ora):
    """The 28 partials that the duals carry through the time loop (64 cells,
    10 Rusanov steps) against central differences of the double oracle,
    h = 1e-6 |theta_k|.  The dual's value is the double run's misfit."""
    n, steps = 64, 10
    mats, prims, rho_t, u_t = _dual_case(ora, n, steps)
    # a target away from the run itself, so that every partial is sizeable
    rho_t = rho_t * (1.0 + 0.01 * np.sin(np.arange(n)))
    L, g = ora.adjoint_dual(mats, prims, 0.5, n, steps, 0.6, rho_t, u_t)
    theta = np.concatenate([prims[:, 0], prims[:, 1], [mats[4], mats[1]]])
    assert L == pytest.approx(_theta_loss(ora, theta, n, steps, rho_t, u_t), rel=1e-12)
    fd = np.empty(28)
    for k in range(28):
        h = 1e-6 * abs(theta[k])
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        fd[k] = (_theta_loss(ora, tp, n, steps, rho_t, u_t) -
                 _theta_loss(ora, tm, n, steps, rho_t, u_t)) / (2 * h)
    # the differences' rounding error is about eps L / h: 2e-8 of the largest
    # partial for the off-diagonal entries of A (|theta| = 0.05)
    scale = np.abs(g).max()
    for k in range(28):
        assert abs(g[k] - fd[k]) <= 5e-6 * abs(g[k]) + 2e-8 * scale, (k, g[k], fd[k])


def test_dual_eigenvalue_partials_match_differences(ora):
    """The dual eigenvalues' partials (v^T dA v with v from the Jacobi
    rotations) against central differences of the Jacobi eigenvalues, on
    acoustic tensors of random states moved along random directions."""
    import ctypes
    from oracle_bind import P as ptr
    rng = np.random.default_rng(3)
    for _ in range(20):
        A = rand_A(rng)
        ac = ora.acoustic(MAT, 1e-3, A.reshape(-1, order="F")).reshape(-1, order="F")
        d = rng.standard_normal((3, 3))
        d = (d + d.T).reshape(-1, order="F")
        w, dw = np.zeros(3), np.zeros(3)
        ora.L.ora_eigvals3_dual(ptr(ac), ptr(d), ptr(w), ptr(dw))
        assert np.array_equal(w, ora.eigvals3(ac))
        h = 1e-6 * np.abs(ac).max()
        fd = (ora.eigvals3(ac + h * d) - ora.eigvals3(ac - h * d)) / (2 * h)
        # the Jacobi order of the eigenvalues is the same at +-h (they are well apart)
        np.testing.assert_allclose(dw, fd, rtol=1e-6, atol=1e-8 * np.abs(d).max())
        # first-order perturbation of a symmetric matrix's eigenvalues, from
        # LAPACK's eigenvectors (eigh: ascending eigenvalues)
        ev, V = np.linalg.eigh(0.5 * (ac + ac.reshape(3, 3).reshape(-1, order="F")).reshape(3, 3, order="F"))
        Dm = d.reshape(3, 3, order="F")
        ref = np.array([V[:, i] @ Dm @ V[:, i] for i in range(3)])
        np.testing.assert_allclose(dw[np.argsort(w)], ref, rtol=1e-9, atol=1e-12 * np.abs(d).max())


@pytest.mark.slow
def test_adjoint_dual_fixture_reproduces(ora):
    """tests/golden/adjoint_dual_512.npz (512 cells, 120 steps; written by
    tests/golden/make_adjoint_dual.py) is what the duals give now: the GPU
    adjoint's gradient is checked against it."""
    import os
    from golden.make_adjoint_dual import FIXTURE, make
    g = np.load(FIXTURE)
    d = make(ora)
    assert np.array_equal(d["rho_t"], g["rho_t"]) and np.array_equal(d["u_t"], g["u_t"])
    assert d["loss"] == float(g["loss"])
    assert np.array_equal(d["grad"], g["grad"])
    assert os.path.basename(FIXTURE) == "adjoint_dual_512.npz"
