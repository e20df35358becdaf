"""Pins the oracle (oracle/ccd_oracle.py) to the compiled reference: every fixture
in tests/golden/ was written by the reference's own routines
(oracle/ref_driver.f90 + make_golden.py)."""
import os

import numpy as np
import pytest

import ccd_oracle as co
from conftest import GOLDEN, fixture_interaction, fixture_lattice, golden

# fixtures that carry the reference's orbital table
DUMPED = sorted(f[:-4] for f in os.listdir(GOLDEN)
                if f.endswith(".npz") and "orb_int" in np.load(os.path.join(GOLDEN, f)).files)


@pytest.mark.parametrize("name", DUMPED)
def test_oracle_basis_equals_reference_orbital_table(name):
    """The oracle's basis, built from the deck parameters alone, is the
    reference's orbital table state for state: degenerate momenta in the order
    its quicksort leaves them (so open shells get the same holes), momenta and
    energies to rounding."""
    g = golden(name)
    A, nmax = int(g["params"][0]), int(g["params"][1])
    lat = co.Lattice.build(A, str(g["matter"]), float(g["rho"]), nmax, twist=tuple(g["twist"]))
    oi, od = g["orb_int"], g["orb_dbl"]
    assert np.array_equal(lat.n.T, oi[:3])
    assert np.array_equal(lat.sz, oi[3]) and np.array_equal(lat.tz, oi[4])
    assert np.allclose(lat.k.T, od[:3], rtol=0, atol=1e-15)
    assert np.allclose(lat.e, od[3], rtol=1e-15, atol=0)


def test_reference_index_order_sorts_with_ties():
    """The restated quicksort returns a permutation that sorts its keys, also
    with many equal keys and on the short arrays its insertion pass handles."""
    rng = np.random.default_rng(3)
    for n in (1, 2, 6, 7, 8, 30, 125, 343, 1000):
        for levels in (3, 50, 10 ** 9):
            arr = list(rng.integers(0, levels, n).astype(float))
            order = co.reference_index_order(arr)
            assert sorted(order) == list(range(n))
            keys = [arr[i] for i in order]
            assert keys == sorted(arr)


@pytest.mark.parametrize("name", ["pnm14_n1_nnlo", "snm4_n1_nnlo", "pnm14_n1_minn", "snm4_n1_minn",
                                  "pnm14_n1_minn_twist"])
def test_oracle_fock_e0_against_reference(name):
    g = golden(name)
    lat = fixture_lattice(g)
    cc = co.CCD(lat, fixture_interaction(g, lat))
    assert np.max(np.abs(cc.fock - g["fock"])) < 1e-11
    assert abs(cc.e0 - g["e0"]) < 1e-10


@pytest.mark.parametrize("name", ["pnm14_n1_nnlo", "snm4_n1_nnlo", "pnm14_n1_minn", "snm4_n1_minn",
                                  "pnm14_n1_minn_twist"])
def test_oracle_ccd_iterations_against_reference(name):
    """Same V and Fock as the reference: every iteration's correlation energy.
    Before the first DIIS step the agreement is at rounding level; the DIIS
    solve is ill-conditioned (its coefficients reach 1e2 with alternating signs),
    so rounding-level differences in the iterates grow to ~1e-7 relative after a
    DIIS step and shrink again towards the fixed point."""
    g = golden(name)
    lat = fixture_lattice(g)
    cc = co.CCD(lat, fixture_interaction(g, lat), fock=g["fock"], e0=g["e0"])
    A, nmax, sub, step, maxit, order = (int(x) for x in g["params"])
    en = np.array(cc.solve(sub, step, float(g["tol"]), maxit))
    ref = g["energies"]
    assert len(en) == len(ref)
    k = min(step, len(ref))
    assert np.max(np.abs(en[:k] - ref[:k])) < 1e-11 * np.max(np.abs(ref))
    assert np.max(np.abs(en - ref)) < 1e-6 * np.max(np.abs(ref))
    # the loop stops at |dE| <= tol, so the two fixed-point approaches may differ
    # by up to tol
    assert abs(cc.ecc - g["eccd"]) < 1e-9 + float(g["tol"])


def test_oracle_minnesota_matches_reference_function():
    """The oracle's own Minnesota restatement against the reference's
    vmom_minnesota elements (stored V records of a Minnesota fixture)."""
    for name in ("pnm14_n1_minn", "snm4_n1_minn", "pnm14_n1_minn_twist"):
        g = golden(name)
        lat = fixture_lattice(g)
        mn = co.MinnesotaInteraction(lat)
        idx = g["v_idx"].astype(np.int64)
        v = mn(idx[:, 1] - 1, idx[:, 2] - 1, idx[:, 3] - 1, idx[:, 4] - 1)
        assert np.max(np.abs(v - g["v_val"])) < 1e-14 * max(1.0, np.max(np.abs(g["v_val"])))


def test_oracle_converged_t2_against_reference():
    g = golden("pnm14_n1_minn")
    lat = fixture_lattice(g)
    cc = co.CCD(lat, fixture_interaction(g, lat), fock=g["fock"], e0=g["e0"])
    cc.solve(8, 8, float(g["tol"]), 100)
    t = g["t2_idx"].astype(np.int64) - 1
    idx = cc.ch.flat_index(t[:, 0], t[:, 1], t[:, 2], t[:, 3])
    mine = cc.flat(cc.T)[idx]
    assert np.max(np.abs(mine - g["t2_val"])) < 1e-9 * np.max(np.abs(g["t2_val"]))


def t_from_fixture(cc, g):
    """Ordered amplitudes rebuilt from the reference's converged canonical T2."""
    x = np.zeros(cc.ch.size, np.complex128)
    t = g["t2_idx"].astype(np.int64) - 1
    a, b, i, j = t[:, 0], t[:, 1], t[:, 2], t[:, 3]
    for (p, q, r, s_, sg) in ((a, b, i, j, 1), (b, a, i, j, -1), (a, b, j, i, -1), (b, a, j, i, 1)):
        idx = cc.ch.flat_index(p, q, r, s_)
        x[idx[idx >= 0]] = sg * g["t2_val"][idx >= 0]
    return cc.unflat(x)


@pytest.mark.parametrize("name", ["snm4_n1_nnlo", "snm4_n1_minn", "pnm14_n1_minn_twist"])
def test_oracle_triples_against_reference(name):
    """(T) of the oracle on the reference's own converged T2 and V."""
    g = golden(name)
    lat = fixture_lattice(g)
    vi = fixture_interaction(g, lat)
    cc = co.CCD(lat, vi, fock=g["fock"], e0=g["e0"])
    T = t_from_fixture(cc, g)
    h = g["hppp_idx"].astype(np.int64) - 1
    # vhppp records hold <cb||kd> keyed (k,d,c,b)
    hv = co.TableInteraction(lat.nsp, [(h[:, 2], h[:, 3], h[:, 0], h[:, 1], g["hppp_val"])])
    et = cc.triples(T, vhppp=hv)
    ref = g["eccdt"] - g["eccd"]
    assert abs(et - ref) < 1e-12 * max(1.0, abs(ref))


def test_oracle_ccdt1_energies_against_reference_program():
    """The CCDT-1 triples energy (setup_t3, t3_intermediate, t3_energy) of the
    oracle on the reference's own V, Fock and <cb||kd> blocks (SNM A=4,
    |n_i|<=1, chiral NNLO with the normal-ordered NNLO 3NF) against the
    'CCDT-1 energy' lines the reference program prints: the start T3 = W3/D,
    and the first iteration's T3 = (W3 + W(T2))/D from the MBPT2 amplitudes
    (this pins the sign of the cross term W conj(W3))."""
    import chiral_oracle as ch
    g = golden("snm4_n1_nnlo_3nf")
    p = golden("snm4_n1_nnlo_3nf_ccdt1")
    assert abs(p["result_line"][4] * 4 - g["e0"].real) < 4e-5
    assert abs(p["ccd_energies"][0] - (g["e0"] + g["energies"][0])) < 1e-10 * abs(g["e0"])
    lat = fixture_lattice(g)
    cc = co.CCD(lat, fixture_interaction(g, lat), fock=g["fock"], e0=g["e0"])
    T = cc.mbpt2_start()
    h = g["hppp_idx"].astype(np.int64) - 1
    hv = co.TableInteraction(lat.nsp, [(h[:, 2], h[:, 3], h[:, 0], h[:, 1], g["hppp_val"])])
    c3 = ch.ThreeBodyConstants(ch.ChiralConstants(ch.NNLO, 450.0), 0.017, 0.79, 450.0, 3)
    w3 = lambda a, b, c, i, j, k: ch.chiral3_asym(lat, c3, a, b, c, i, j, k)
    e30 = cc.triples(T, vhppp=hv, mode="w3", w3=w3)
    assert abs(e30 - p["e3"][0]) < 1e-9 * abs(p["e3"][0]), (e30, p["e3"][0])
    e31 = cc.triples(T, vhppp=hv, mode="ccdt1", w3=w3)
    assert abs(e31 - p["e3"][1]) < 1e-9 * abs(p["e3"][1]), (e31, p["e3"][1])


def test_oracle_triples_with_residual_3nf_against_reference():
    """(T) at tnf_approx 3 (pert_triples2's et_5: the antisymmetrized 3NF
    element added to the triples numerator) on the reference's converged T2,
    V and <cb||kd> blocks (SNM A=4, |n_i|<=1, NNLO with the normal-ordered
    3NF)."""
    import chiral_oracle as ch
    g = golden("snm4_n1_nnlo_3nf3")
    lat = fixture_lattice(g)
    cc = co.CCD(lat, fixture_interaction(g, lat), fock=g["fock"], e0=g["e0"])
    T = t_from_fixture(cc, g)
    h = g["hppp_idx"].astype(np.int64) - 1
    hv = co.TableInteraction(lat.nsp, [(h[:, 2], h[:, 3], h[:, 0], h[:, 1], g["hppp_val"])])
    c3 = ch.ThreeBodyConstants(ch.ChiralConstants(ch.NNLO, 450.0), 0.017, 0.79, 450.0, 3)
    w3 = lambda a, b, c, i, j, k: ch.chiral3_asym(lat, c3, a, b, c, i, j, k)
    et5 = cc.triples(T, vhppp=hv, mode="td", w3=w3)
    ref = g["eccdt"] - g["eccd"]
    assert abs(et5 - ref) < 1e-10 * abs(ref), (et5, ref)
