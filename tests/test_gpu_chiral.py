"""GPU parity of the built-in chiral NN generator: elements, Fock/E0, the CCD
trajectory and (T) against the compiled reference's fixtures, and elements
against the oracle on lattices larger than the fixtures."""
import numpy as np
import pytest

import ccd_oracle as co
import chiral_oracle as ch
from conftest import golden
from test_gpu_parity import check_trajectory, params, set_reference_t2, step_loop

pytestmark = pytest.mark.gpu

# (fixture, chiral order, cutoff)
CASES = [("pnm14_n1_nnlo", 3, 450.0), ("snm4_n1_nnlo", 3, 450.0), ("snm4_n1_lo450", 0, 450.0),
         ("snm4_n1_nlo450", 2, 450.0), ("snm4_n1_nnlo394", 3, 394.0), ("snm4_n1_nnlo500", 3, 500.0),
         ("pnm14_n1_nnlo_twist", 3, 450.0), ("pnm66_n2_nnlo", 3, 450.0)]


def chiral_solver(ccnm, g, order, cutoff):
    A, nmax, sub, step, maxit = params(g)
    matter = ccnm.PNM if str(g["matter"]) == "pnm" else ccnm.SNM
    s = ccnm.Solver(A=A, matter=matter, rho=float(g["rho"]), nmax=nmax, twist=tuple(g["twist"]))
    s.set_chiral(order, cutoff)
    s.generate_interaction(ccnm.INT_CHIRAL)
    return s


def label_map(s, g):
    """Reference orbital index -> engine orbital index, through (n, sz, tz)."""
    mine = s.orbitals()
    lab = {(mine["nx"][x], mine["ny"][x], mine["nz"][x], mine["sz"][x], mine["tz"][x]): x
           for x in range(len(mine["e"]))}
    oi = g["orb_int"]
    return np.array([lab[(oi[0][x], oi[1][x], oi[2][x], oi[3][x], oi[4][x])] for x in range(oi.shape[1])])


@pytest.mark.parametrize("name,order,cutoff", CASES)
def test_native_chiral_against_reference(ccnm, name, order, cutoff):
    """Lattice, chiral V, Fock, E0, MBPT2 and the CCD trajectory built by the
    engine versus the reference run (chiral_NN_with_delta) on the same input."""
    g = golden(name)
    A, nmax, sub, step, maxit = params(g)
    s = chiral_solver(ccnm, g, order, cutoff)
    assert abs(s.e0() - g["e0"]) < 1e-10 * abs(g["e0"])
    emp2 = s.setup_t_amplitudes()
    assert abs(emp2 - g["energies"][0]) < 1e-11 * abs(g["energies"][0])
    en = step_loop(s, sub, step, float(g["tol"]), maxit)
    check_trajectory(en, g["energies"], step, float(g["tol"]))


@pytest.mark.parametrize("name,order,cutoff", [c for c in CASES if c[0] != "pnm66_n2_nnlo"])
def test_chiral_elements_against_reference(ccnm, name, order, cutoff):
    """Every stored <pq||rs> record of the reference's blocks, evaluated by the
    device generator on the engine's own lattice (orbitals matched by labels)."""
    g = golden(name)
    s = chiral_solver(ccnm, g, order, cutoff)
    m = label_map(s, g)
    idx = g["v_idx"].astype(np.int64) - 1
    v = s.vnn_elements(m[idx[:, 1]], m[idx[:, 2]], m[idx[:, 3]], m[idx[:, 4]])
    ref = g["v_val"]
    assert np.max(np.abs(v - ref)) < 1e-12 * np.max(np.abs(ref))


def _conserving_quadruples(o, n, seed):
    rng = np.random.default_rng(seed)
    nsp = len(o["e"])
    key = {}
    for x in range(nsp):
        key.setdefault((o["nx"][x], o["ny"][x], o["nz"][x], o["tz"][x]), []).append(x)
    out = []
    while len(out) < n:
        p, q, r = (int(v) for v in rng.integers(0, nsp, 3))
        k = tuple(o[c][p] + o[c][q] - o[c][r] for c in ("nx", "ny", "nz", "tz"))
        if k in key:
            out.append((p, q, r, key[k][rng.integers(0, len(key[k]))]))
    return np.array(out, np.int64).T


@pytest.mark.parametrize("A,matter,rho,nmax,twist,ci", [
    (66, "pnm", 0.16, 3, (0.0, 0.0, 0.0), None),
    (66, "pnm", 0.12, 4, (0.5, 2.5, 1.5), None),
    (28, "snm", 0.16, 2, (0.0, 0.0, 0.0), None),
    (66, "pnm", 0.16, 2, (0.0, 0.0, 0.0), (-0.74 * 1.2, -0.49, -0.65 * 0.8, 0.96)),
])
def test_chiral_elements_against_oracle(ccnm, A, matter, rho, nmax, twist, ci):
    """Random conserving quadruples on lattices beyond the fixtures, and with
    the c1/c3 LECs moved by +-20%, against the oracle's chiral_NN_with_delta."""
    s = ccnm.Solver(A=A, matter=ccnm.PNM if matter == "pnm" else ccnm.SNM, rho=rho, nmax=nmax, twist=twist)
    s.set_chiral(3, 450.0, ci)
    s.generate_interaction(ccnm.INT_CHIRAL)
    o = s.orbitals()
    lat = co.Lattice.build(A, matter, rho, nmax, twist=twist)
    lab = {(tuple(lat.n[x]), int(lat.sz[x]), int(lat.tz[x])): x for x in range(lat.nsp)}
    m = np.array([lab[((o["nx"][x], o["ny"][x], o["nz"][x]), o["sz"][x], o["tz"][x])] for x in range(len(o["e"]))])
    p, q, r, t = _conserving_quadruples(o, 4000, 11)
    v = s.vnn_elements(p, q, r, t)
    it = ch.ChiralInteraction(lat, 3, 450.0, c1=None if ci is None else ci[0], c3=None if ci is None else ci[2])
    want = it(m[p], m[q], m[r], m[t])
    assert np.max(np.abs(want)) > 0.1
    assert np.max(np.abs(v - want)) < 1e-12 * np.max(np.abs(want))


# (fixture, chiral order, 3NF approximation)
TNF_CASES = [("pnm14_n1_nnlo_3nf", 3, 2), ("snm4_n1_nnlo_3nf", 3, 2), ("snm4_n1_nlo450_3nf", 2, 2),
             ("pnm14_n1_nnlo_3nf1", 3, 1)]


def chiral_3nf_solver(ccnm, g, order, approx, reference_orbitals=True):
    """The chiral NN plus the normal-ordered 3NF generated on the GPU, on the
    reference's own orbital table (its numbering) or on the engine's lattice."""
    A, nmax, sub, step, maxit = params(g)
    if reference_orbitals:
        oi, od = g["orb_int"], g["orb_dbl"]
        s = ccnm.Solver(orbitals={"nx": oi[0], "ny": oi[1], "nz": oi[2], "sz": oi[3], "tz": oi[4],
                                  "kx": od[0], "ky": od[1], "kz": od[2], "e": od[3], "A": A,
                                  "box": float(g["box"]), "volume": float(g["volume"])})
    else:
        matter = ccnm.PNM if str(g["matter"]) == "pnm" else ccnm.SNM
        s = ccnm.Solver(A=A, matter=matter, rho=float(g["rho"]), nmax=nmax, twist=tuple(g["twist"]))
    s.set_chiral(order, 450.0)
    s.set_3nf(approx, 0.017, 0.79, 450.0, 3)
    s.generate_interaction(ccnm.INT_CHIRAL)
    return s


@pytest.mark.parametrize("name,order,approx", TNF_CASES)
def test_native_chiral_3nf_against_reference(ccnm, name, order, approx):
    """Chiral NN plus the NNLO three-nucleon force normal-ordered on the GPU
    (E0; and Fock; and the two-body blocks), on the reference's orbital table,
    against the reference run: E0 and the Fock matrix to rounding.  The
    reference fills the non-canonical orderings of its hhpp block with their
    own 3NF sums, which the float32 2/3 of the 3NF regulator makes differ from
    the antisymmetric partner by ~1e-7 of the 3NF part, so MBPT2 and the
    iterations agree to 1e-7 relative."""
    g = golden(name)
    A, nmax, sub, step, maxit = params(g)
    s = chiral_3nf_solver(ccnm, g, order, approx)
    assert abs(s.e0() - g["e0"]) < 1e-11 * abs(g["e0"])
    fd = s.fock_diag()
    assert np.max(np.abs(fd - np.diag(g["fock"]))) < 1e-11 * np.max(np.abs(fd))
    emp2 = s.setup_t_amplitudes()
    scale = np.max(np.abs(g["energies"]))
    assert abs(emp2 - g["energies"][0]) < 1e-7 * scale
    en = step_loop(s, sub, step, float(g["tol"]), maxit)
    k = min(step, len(en), len(g["energies"]))
    assert np.max(np.abs(en[:k] - g["energies"][:k])) < 1e-7 * scale
    assert abs(en[-1] - g["energies"][-1]) < 1e-7 * scale + 20 * float(g["tol"])


@pytest.mark.parametrize("name,order,approx", TNF_CASES[:2])
def test_native_lattice_3nf_against_oracle(ccnm, name, order, approx):
    """On the engine's own lattice (built from the deck parameters; its
    numbering is the reference's) the normal-ordered E0 and Fock matrix equal
    the oracle's and the reference's."""
    import ccd_oracle as co
    import chiral_oracle as ch
    g = golden(name)
    A, nmax, sub, step, maxit = params(g)
    s = chiral_3nf_solver(ccnm, g, order, approx, reference_orbitals=False)
    lat = co.Lattice.build(A, str(g["matter"]), float(g["rho"]), nmax, twist=tuple(g["twist"]))
    mine = s.orbitals()
    assert np.array_equal(np.asarray(mine["nx"]), lat.n[:, 0]) and np.array_equal(np.asarray(mine["sz"]), lat.sz)
    nn = ch.ChiralInteraction(lat, order, 450.0)
    no = ch.NormalOrdered3NF(lat, nn.cc)
    i, j = (x.ravel() for x in np.meshgrid(np.arange(A), np.arange(A), indexing="ij"))
    e0 = np.sum(lat.e[:A]) + 0.5 * np.sum(nn(i, j, i, j)) + no.e0()
    assert abs(s.e0() - e0) < 1e-12 * abs(e0)
    assert abs(s.e0() - g["e0"]) < 1e-12 * abs(g["e0"])
    pd = np.arange(lat.nsp)
    f = lat.e.astype(np.complex128) + no.fock(pd, pd)
    for h in range(A):
        f = f + nn(pd, np.full(lat.nsp, h), pd, np.full(lat.nsp, h))
    assert np.max(np.abs(s.fock_diag() - f)) < 1e-12 * np.max(np.abs(f))


@pytest.mark.parametrize("name", ["pnm14_n1_nnlo_3nf", "snm4_n1_nnlo_3nf"])
def test_triples_3nf_against_reference(ccnm, name):
    """(T) with <cb||kd> carrying the normal-ordered 3NF (and <jk||lc> the NN
    alone, as the reference's vnn_hhhp), on the reference's converged T2."""
    g = golden(name)
    s = chiral_3nf_solver(ccnm, g, int(g["params"][5]), int(g["tnf_approx"]))
    s.setup_t_amplitudes()
    set_reference_t2(ccnm, s, g)
    et = s.pert_triples2()
    ref = g["eccdt"] - g["eccd"]
    assert abs(et - ref) < 1e-11 * max(1.0, abs(ref)), (et, ref)


@pytest.mark.parametrize("name", ["pnm14_n1_nnlo", "snm4_n1_nnlo", "pnm14_n1_nnlo_twist"])
def test_triples_native_chiral_against_reference(ccnm, name):
    """(T) with the <cb||kd> and <jk||lc> elements from the device chiral
    generator, on the reference's converged T2, versus pert_triples2."""
    g = golden(name)
    s = chiral_solver(ccnm, g, int(g["params"][5]), float(g.get("cutoff", 450.0)))
    s.setup_t_amplitudes()
    set_reference_t2(ccnm, s, g)
    et = s.pert_triples2()
    ref = g["eccdt"] - g["eccd"]
    assert abs(et - ref) < 1e-11 * max(1.0, abs(ref)), (et, ref)


@pytest.mark.parametrize("name,tnf", [("pnm14_n1_nnlo", "F, 0"), ("pnm14_n1_nnlo_3nf", "T, 2")])
def test_fortran_host_reference_deck_chiral(ccnm, tmp_path, name, tnf):
    """The unmodified reference deck (no engine line: chiral NNLO at 450 MeV,
    DIIS 6/6, tolerance 1e-6, CCD(T); without and with the normal-ordered 3NF)
    through the Fortran driver ccm_gpu; its result block matches the
    reference run of the same deck."""
    import os
    import subprocess
    from conftest import ROOT
    exe = os.path.join(ROOT, "coupled-cluster-nuclearmatter_amd", "host", "ccm_gpu")
    if not os.path.exists(exe):
        pytest.skip("Fortran host not built (flang absent)")
    g = golden(name)
    deck = tmp_path / "ccm_in"
    deck.write_text("""!Chiral order for Deltas(LO = 0,NLO=2,NNLO=3,N3LO=4) and cutoff
3, 450
! cE and cD 3nf parameters:
 0.017, 0.79
! number of particles
14
! specify: pnm/snm, input type: density/kfermi
pnm, density
! specify boundary conditions (PBC/TABC/TABCsp)
PBC
! dens/kf, ntwist,  nmax, my_twists
0.16,1,1,1
! specify cluster approximation: CCD, CCD(T)
CCD(T)
! tnf switch (T/F) and specify 3nf approximation: 0=tnf0b, 1=tnf1b, 2=tnf2b
TNF
! 3nf cutoff(MeV),non-local reg. exp
450, 3
""".replace("TNF", tnf))
    res = subprocess.run([exe, str(deck)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    line = [l for l in res.stdout.splitlines() if l.startswith("     0.01700")][-1]
    vals = [float(x) for x in line.split()]
    A = 14
    ref = [0.017, 0.79, float(g["kf"]), 0.16, g["e0"].real / A, (g["e0"] + g["energies"][0]).real / A,
           g["eccd"].real / A, g["eccdt"].real / A]
    assert np.allclose(vals, ref, atol=2e-5), (vals, ref)
