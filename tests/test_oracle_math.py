"""Pins of the oracle by mathematics alone (no reference output involved).

* The channel-blocked, ordered-pair residual of the oracle (the reference's
  factorisation into I6/I7/I9/I4 intermediates with dynamic denominators,
  ccm_t2_eqn.f90:4-1518) against the textbook CCD amplitude equation written
  densely over all orbital indices (antisymmetrised form with a general Fock
  matrix, as in Crawford and Schaefer, Rev. Comput. Chem. 14 (2000) 33, with
  the T1 terms dropped), on the tiny PNM A=14 |n_i|<=1 lattice: R - D T must
  equal it for arbitrary amplitudes; the constant, linear and quadratic parts
  are checked separately so that a dropped or doubled term cannot hide.
* The oracle's Pulay DIIS (ccm_diis.f90:74-249) on a linear fixed-point problem
  of dimension 3: with 6 stored iterates the first extrapolation must land on
  the exact solution (the error vectors of a linear iteration are an affine image
  of the iterates, so a vanishing combination of 5 of them in 3 dimensions exists).
* The MBPT2 energy against the closed form 1/4 sum |<ij||ab>|^2 / D.
"""
import numpy as np
import pytest

import ccd_oracle as co
from conftest import fixture_interaction, fixture_lattice, golden


@pytest.fixture(scope="module", params=["minnesota", "pnm14_n1_nnlo"])
def tiny(request):
    """PNM A=14 |n_i|<=1 with the closed-form Minnesota interaction, or with the
    reference's chiral NNLO elements of the same lattice (a fixture): the chiral
    tensor force does not conserve Sz, so the spin-partner (off-diagonal) one-body
    terms are exercised too."""
    if request.param == "minnesota":
        lat = co.Lattice.build(14, "pnm", 0.16, 1)
        vf = co.MinnesotaInteraction(lat)
        cc = co.CCD(lat, vf)
    else:
        g = golden(request.param)
        lat = fixture_lattice(g)
        vf = fixture_interaction(g, lat)
        cc = co.CCD(lat, vf, fock=g["fock"], e0=g["e0"])
    n = lat.nsp
    p = np.arange(n)
    V = vf(p[:, None, None, None], p[None, :, None, None], p[None, None, :, None], p[None, None, None, :])
    return lat, cc, np.asarray(V, np.complex128)


def dense_t(cc, T):
    """Ordered dense t[a,b,i,j] (particles and holes 0-based within their sets)."""
    lat, ch = cc.lat, cc.ch
    A, nsp = lat.A, lat.nsp
    a = np.arange(A, nsp)[:, None, None, None]
    b = np.arange(A, nsp)[None, :, None, None]
    i = np.arange(A)[None, None, :, None]
    j = np.arange(A)[None, None, None, :]
    idx = ch.flat_index(*np.broadcast_arrays(a, b, i, j))
    x = cc.flat(T)
    return np.where(idx >= 0, x[np.maximum(idx, 0)], 0.0)


def textbook_ccd(V, f, t, A):
    """Dense CCD amplitude equation (everything on the right of 0 = ...)."""
    h, p = slice(0, A), slice(A, None)
    Vppph = V[p, p, h, h]
    out = Vppph.copy()
    fpp, fhh = f[p, p], f[h, h]
    Pab = lambda x: x - x.transpose(1, 0, 2, 3)
    Pij = lambda x: x - x.transpose(0, 1, 3, 2)
    out += Pab(np.einsum("bc,acij->abij", fpp, t))
    out -= Pij(np.einsum("kj,abik->abij", fhh, t))
    out += 0.5 * np.einsum("abcd,cdij->abij", V[p, p, p, p], t, optimize=True)
    out += 0.5 * np.einsum("klij,abkl->abij", V[h, h, h, h], t, optimize=True)
    out += Pab(Pij(np.einsum("kbcj,acik->abij", V[h, p, p, h], t, optimize=True)))
    Vhhpp = V[h, h, p, p]
    out += 0.25 * np.einsum("klcd,cdij,abkl->abij", Vhhpp, t, t, optimize=True)
    # the summand is symmetric under the joint swap (a<->b, i<->j), so this is
    # the same as P(ij) alone
    out += 0.5 * Pij(Pab(np.einsum("klcd,acik,bdjl->abij", Vhhpp, t, t, optimize=True)))
    out -= 0.5 * Pij(np.einsum("klcd,dcik,ablj->abij", Vhhpp, t, t, optimize=True))
    out -= 0.5 * Pab(np.einsum("klcd,aclk,dbij->abij", Vhhpp, t, t, optimize=True))
    return out


def test_interaction_antisymmetry_and_hermiticity(tiny):
    lat, cc, V = tiny
    s = np.max(np.abs(V))
    assert np.max(np.abs(V + V.transpose(1, 0, 2, 3))) < 1e-13 * s
    assert np.max(np.abs(V + V.transpose(0, 1, 3, 2))) < 1e-13 * s
    assert np.max(np.abs(V - np.conj(V.transpose(2, 3, 0, 1)))) < 1e-12 * s


def test_residual_is_the_textbook_ccd_equation(tiny):
    """R(T) - D(T) T of the oracle equals the dense textbook CCD equation for a
    generic (non-converged, non-MBPT2) amplitude set."""
    lat, cc, V = tiny
    A = lat.A
    rng = np.random.default_rng(7)
    T0 = cc.mbpt2_start()
    # generic amplitudes with the right symmetry: scale the MBPT2 amplitudes by
    # a random factor symmetric under a<->b and i<->j
    T = []
    for c, t in enumerate(T0):
        pp, hh = cc.ch.pp[c], cc.ch.hh[c]
        g = rng.uniform(0.5, 1.5, (lat.nsp, lat.nsp))
        g = g + g.T
        fac = g[pp[:, 0], pp[:, 1]][:, None] * g[hh[:, 0], hh[:, 1]][None, :]
        T.append(t * fac)
    t = dense_t(cc, T)
    assert np.max(np.abs(t + t.transpose(1, 0, 2, 3))) < 1e-15

    def oracle(lam):
        Tl = [lam * x for x in T]
        R, d_h, d_p = cc.residual(Tl)
        D = cc.denominators(d_h, d_p)
        return dense_t(cc, [R[c] - D[c] * Tl[c] for c in range(len(T))])

    # parts of a quadratic in lam from lam = 0, 1, 2
    def parts(f0, f1, f2):
        return f0, (4 * f1 - f2 - 3 * f0) / 2, (f2 - 2 * f1 + f0) / 2

    mine = parts(oracle(0.0), oracle(1.0), oracle(2.0))
    book = parts(*(textbook_ccd(V, cc.fock, lam * t, A) for lam in (0.0, 1.0, 2.0)))
    for m, b in zip(mine, book):
        scale = np.max(np.abs(b))
        assert scale > 0.1
        assert np.max(np.abs(m - b)) < 1e-12 * scale


def test_energy_and_mbpt2_closed_forms(tiny):
    lat, cc, V = tiny
    A = lat.A
    f = np.real(np.diag(cc.fock))
    T0 = cc.mbpt2_start()
    t = dense_t(cc, T0)
    Vhhpp = V[:A, :A, A:, A:]
    den = (f[:A][:, None, None, None] + f[:A][None, :, None, None]
           - f[A:][None, None, :, None] - f[A:][None, None, None, :])
    with np.errstate(divide="ignore", invalid="ignore"):
        e2 = 0.25 * np.sum(np.where(den != 0, np.abs(Vhhpp) ** 2 / den, 0.0))
    assert abs(cc.energy(T0) - e2) < 1e-12 * abs(e2)
    assert abs(cc.energy(T0) - 0.25 * np.einsum("ijab,abij->", Vhhpp, t)) < 1e-12 * abs(e2)
    # E0 = kinetic + 1/2 sum <ij||ij>
    e0 = np.sum(lat.e[:A]) + 0.5 * np.einsum("ijij->", V[:A, :A, :A, :A])
    assert abs(cc.e0 - e0) < 1e-10


class _LinearToy(co.CCD):
    """x = M x + b in 3 complex dimensions, posed as the CCD fixed point T = R/D
    with D = 1; the 'energy' is a fixed linear functional of x."""

    def __init__(self, M, b, w, x0):
        self.M, self.b, self.w, self.x0 = M, b, w, x0
        self.e0 = 0.0

    def mbpt2_start(self):
        return [self.x0.copy()]

    def flat(self, T):
        return T[0]

    def unflat(self, x):
        return [x]

    def energy(self, T):
        return np.sum(self.w * T[0])

    def residual(self, T):
        return [self.M @ T[0] + self.b], None, None

    def denominators(self, d_h, d_p):
        return [1.0]


def test_diis_lands_on_linear_fixed_point():
    rng = np.random.default_rng(3)
    M = 0.6 * (rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))) / 3
    b = rng.standard_normal(3) + 1j * rng.standard_normal(3)
    w = rng.standard_normal(3)
    xs = np.linalg.solve(np.eye(3) - M, b)
    toy = _LinearToy(M, b, w, np.zeros(3, np.complex128))
    en = toy.solve(subspace=6, diis_step=6, tol=-1.0, maxit=6)
    # the 6th update is the first extrapolation
    assert np.max(np.abs(toy.T[0] - xs)) < 1e-10 * np.max(np.abs(xs))
    assert abs(en[-1] - np.sum(w * xs)) < 1e-10 * np.max(np.abs(xs))
    # without extrapolation the 0.75/0.25 mixing is still far away after 6 steps
    plain = _LinearToy(M, b, w, np.zeros(3, np.complex128))
    plain.solve(subspace=6, diis_step=1000, tol=-1.0, maxit=6)
    assert np.max(np.abs(plain.T[0] - xs)) > 1e-3
