"""Pins of the oracle's restatement of the engine's Anderson update
(ccd_oracle.CCD.solve_anderson), which the GPU trajectory is checked against:

* the coefficient solve against the constrained least-squares problem it
  states, solved independently through its KKT system;
* on a linear fixed point x = M x + b it reaches the known solution;
* on the CCD equations it stops on the fixed point of the reference update
  (the oracle's diis_t1t2 restatement, pinned by the reference's fixtures).
"""
import numpy as np
import pytest

import ccd_oracle as co


def test_coefficients_solve_the_constrained_least_squares():
    rng = np.random.default_rng(4)
    for m in (1, 2, 3, 5, 8):
        F = rng.normal(size=(40, m)) + 1j * rng.normal(size=(40, m))
        G = np.real(F.conj().T @ F)
        c = co.CCD.anderson_coefficients(G)
        kkt = np.zeros((m + 1, m + 1))
        kkt[:m, :m] = 2 * G
        kkt[:m, m] = kkt[m, :m] = 1.0
        rhs = np.zeros(m + 1)
        rhs[m] = 1.0
        want = np.linalg.solve(kkt, rhs)[:m]
        assert abs(c.sum() - 1.0) < 1e-14
        assert np.allclose(c, want, rtol=1e-10, atol=1e-12)


def test_coefficients_reject_parallel_vectors():
    f = np.array([1.0, 2.0, -1.0])
    F = np.stack([f, 2.0 * f], 1)
    assert co.CCD.anderson_coefficients(F.T @ F) is None


def test_linear_fixed_point_known_answer():
    """x = M x + b with a contraction M (dim 12, spectral radius 0.95, so plain
    iteration needs ~600 steps for 1e-13): Anderson with depth 8 and beta = 1,
    written against the same coefficient solve (dropping the oldest pair while
    the system is singular, as solve_anderson does), reaches the solution of
    (1 - M) x = b to 1e-11 within 100 steps."""
    rng = np.random.default_rng(9)
    n = 12
    Q = np.linalg.qr(rng.normal(size=(n, n)))[0]
    M = Q @ np.diag(np.linspace(-0.95, 0.95, n)) @ Q.T
    b = rng.normal(size=n)
    xs = np.linalg.solve(np.eye(n) - M, b)
    x = np.zeros(n)
    ys, fs = [], []
    for _ in range(100):
        f = M @ x + b - x
        if np.linalg.norm(f) < 1e-13:
            break
        if len(ys) == 8:
            ys.pop(0), fs.pop(0)
        ys.append(x + f)
        fs.append(f)
        while True:
            c = co.CCD.anderson_coefficients(np.array([[fj @ fk for fk in fs] for fj in fs]))
            if c is not None or len(ys) == 1:
                c = np.ones(1) if c is None else c
                break
            ys.pop(0), fs.pop(0)
        x = sum(cj * yj for cj, yj in zip(c, ys))
    assert np.linalg.norm(x - xs) < 1e-11 * np.linalg.norm(xs)


@pytest.mark.parametrize("beta,depth", [(1.0, 8), (0.7, 4)])
def test_same_fixed_point_as_reference_update(beta, depth):
    lat = co.Lattice.build(14, "pnm", 0.16, 1)
    cc = co.CCD(lat, co.MinnesotaInteraction(lat))
    e_ref = cc.solve(6, 6, 1e-11, 200)
    ecc_ref = cc.ecc
    e_aa = cc.solve_anderson(depth=depth, beta=beta, tol=1e-11, maxit=200)
    assert abs(cc.ecc - ecc_ref) < 1e-9
    assert len(e_aa) < len(e_ref)
