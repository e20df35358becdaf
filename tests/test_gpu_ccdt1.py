"""CCDT-1 and the (T) with the residual three-nucleon element (tnf_approx 3).

CCDT-1 is the reference's CCDT1 branch of ccd_iter (ccm_iter.f90:26-92): T2
iterates as in CCD, and every energy adds sum T3 conj(W3) over a<b<c, i<j<k
of T3 = (W3 + W(T2))/D (t3_intermediate from the T2 before the update,
t3_energy after it), with W3 = <abc||ijk> of the NNLO three-nucleon force and
W the 18-term (T) numerator.  The fixtures *_ccdt1*.npz hold the 'CCD energy'
and 'CCDT-1 energy' lines the reference program itself prints for the deck.
The normal-ordered 3NF makes the engine's and the reference's hhpp blocks
differ by ~1e-7 of the 3NF part (float32 2/3 of the regulator), so the
trajectories agree to ~1e-7 relative.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, golden
from test_gpu_parity import params, set_reference_t2

pytestmark = pytest.mark.gpu

TOL = float(np.float32(1e-6))


def native_3nf(ccnm, A, matter, rho, nmax, twist, approx):
    s = ccnm.Solver(A=A, matter=ccnm.PNM if matter == "pnm" else ccnm.SNM, rho=rho, nmax=nmax, twist=twist)
    s.set_chiral(ccnm.NNLO, 450.0)
    s.set_3nf(approx, 0.017, 0.79, 450.0, 3)
    s.generate_interaction(ccnm.INT_CHIRAL)
    return s


CCDT1 = [("snm4_n1_nnlo_3nf_ccdt1", 4, "snm", None), ("pnm14_n1_nnlo_3nf_ccdt1", 14, "pnm", None),
         ("pnm14_n1_nnlo_3nf_ccdt1_tabc2_3", 14, "pnm", "pnm14_n1_nnlo_tabc2_3")]


@pytest.mark.parametrize("name,A,matter,twist_from", CCDT1)
def test_ccdt1_step_api_against_reference_program(ccnm, name, A, matter, twist_from):
    """The CCDT1 loop of ccd_iter written against the step API on the
    engine's own lattice (the twisted case at the reference's Gauss-Legendre
    twist): every printed CCD and CCDT-1 energy of the reference program, and
    its iteration count."""
    p = golden(name)
    twist = tuple(golden(twist_from)["twist"]) if twist_from else (0.0, 0.0, 0.0)
    s = native_3nf(ccnm, A, matter, float(p["rho"]), 1, twist, 2)
    e0 = s.e0()
    e1 = s.setup_t_amplitudes()
    ccd, e3s = [e0 + e1], []
    e3 = s.setup_t3()
    e3s.append(e3)
    e1 = e1 + e3
    count = nstep = 0
    dener = 4.0
    for it in range(1, 101):
        if not abs(dener) > TOL:
            break
        s.t3_intermediate()
        s.t2_intermediate()
        count, nstep = s.diis_t1t2(count + 1, nstep + 1, 6, 6)
        e2 = s.ccd_energy_save()
        ccd.append(e0 + e2)
        e3 = s.t3_energy()
        e3s.append(e3)
        e2 = e2 + e3
        dener = e2 - e1
        e1 = e2
    ccd, e3s = np.array(ccd), np.array(e3s)
    assert len(e3s) == len(p["e3"]), (len(e3s), len(p["e3"]))
    scale = abs(p["ccd_energies"][0])
    # up to the first DIIS extrapolation (iteration 6) the trajectories agree
    # to the 1e-7 of the hhpp blocks; after it the ill-conditioned DIIS system
    # amplifies that, and the converged energies agree within the tolerance
    k = 6
    assert np.max(np.abs(ccd[:k] - p["ccd_energies"][:k])) < 1e-7 * scale
    assert np.max(np.abs(e3s[:k] - p["e3"][:k])) < 1e-7 * np.max(np.abs(p["e3"][:k])) + 1e-10
    total = ccd[-1] + e3s[-1]
    ref = p["ccd_energies"][-1] + p["e3"][-1]
    assert abs(total - ref) < 1e-7 * scale + 20 * TOL, (total, ref)
    assert abs(e3s[-1] - p["e3"][-1]) < 1e-4 * abs(p["e3"][-1]) + 1e-8
    if str(p["boundary"]) == "PBC":
        assert abs(total.real / A - p["result_line"][6]) < 2e-5


def test_fortran_host_ccdt1_deck(ccnm, tmp_path):
    """The unmodified reference deck with cc_approx CCDT1 (SNM A=4, |n_i|<=1,
    NNLO with the normal-ordered 3NF) through the Fortran driver: its result
    line equals the reference program's to the printed precision."""
    exe = os.path.join(ROOT, "coupled-cluster-nuclearmatter_amd", "host", "ccm_gpu")
    if not os.path.exists(exe):
        pytest.skip("Fortran host not built (flang absent)")
    p = golden("snm4_n1_nnlo_3nf_ccdt1")
    deck = tmp_path / "ccm_in"
    deck.write_text("""!Chiral order for Deltas(LO = 0,NLO=2,NNLO=3,N3LO=4) and cutoff
3, 450
! cE and cD 3nf parameters:
 0.017, 0.79
! number of particles
4
! specify: pnm/snm, input type: density/kfermi
snm, density
! specify boundary conditions (PBC/TABC/TABCsp)
PBC
! dens/kf, ntwist,  nmax, my_twists
0.16,1,1,1
! specify cluster approximation: CCD, CCD(T)
CCDT1
! tnf switch (T/F) and specify 3nf approximation: 0=tnf0b, 1=tnf1b, 2=tnf2b
T, 2
! 3nf cutoff(MeV),non-local reg. exp
450, 3
""")
    res = subprocess.run([exe, str(deck)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr
    line = [l for l in res.stdout.splitlines() if l.startswith("     0.01700")][-1]
    vals = [float(x) for x in line.split()]
    assert np.allclose(vals, p["result_line"], atol=2e-5), (vals, p["result_line"])
    assert sum("CCDT-1 energy" in l for l in res.stdout.splitlines()) == len(p["e3"])


def test_triples_residual_3nf_on_reference_t2(ccnm):
    """(T) at tnf_approx 3 on the reference's converged T2 (SNM A=4): the
    energy added to CCD is et_5 = sum |W + W3|^2/D, and et_4 = sum |W|^2/D
    equals the tnf_approx 2 (T) of the same amplitudes."""
    g = golden("snm4_n1_nnlo_3nf3")
    g2 = golden("snm4_n1_nnlo_3nf")
    A, nmax, sub, step, maxit = params(g)
    s = native_3nf(ccnm, A, "snm", 0.16, nmax, (0.0, 0.0, 0.0), 3)
    s.setup_t_amplitudes()
    set_reference_t2(ccnm, s, g)
    et = s.pert_triples2()
    et4, et5 = s.triples_parts()
    ref5 = g["eccdt"] - g["eccd"]
    ref4 = g2["eccdt"] - g2["eccd"]
    assert abs(et - et5) == 0
    assert abs(et5 - ref5) < 1e-7 * abs(ref5), (et5, ref5)
    assert abs(et4 - ref4) < 1e-7 * abs(ref4), (et4, ref4)


def test_triples_residual_3nf_native_twisted(ccnm):
    """The whole tnf_approx 3 CCD(T) on the engine's twisted lattice (PNM
    A=14) against the reference run: CCD and the (T) with the residual 3NF."""
    g = golden("pnm14_n1_nnlo_3nf3")
    A, nmax, sub, step, maxit = params(g)
    s = native_3nf(ccnm, A, "pnm", 0.16, nmax, tuple(g["twist"]), 3)
    s.setup_t_amplitudes()
    r = s.ccd_iter(subspace=sub, diis_step=step, tol=float(g["tol"]), maxit=maxit)
    assert abs(r["eccd"] - g["eccd"]) < 1e-7 * abs(g["eccd"])
    et = s.pert_triples2()
    ref = g["eccdt"] - g["eccd"]
    assert abs(et - ref) < 1e-6 * abs(ref), (et, ref)


def test_ccdt1_inside_the_device_loop(ccnm):
    """ccgpu_solve with ccdt1 = 1 (the whole CCDT1 loop on the device side of
    the C ABI) against the reference program: iteration count and final
    energy (PNM A=14, |n_i|<=1, NNLO with the normal-ordered 3NF)."""
    p = golden("pnm14_n1_nnlo_3nf_ccdt1")
    s = native_3nf(ccnm, 14, "pnm", 0.16, 1, (0.0, 0.0, 0.0), 2)
    s.setup_t_amplitudes()
    r = s.ccd_iter(subspace=6, diis_step=6, tol=TOL, maxit=100, ccdt1=1)
    ref = p["ccd_energies"][-1] + p["e3"][-1]
    assert r["iterations"] == len(p["e3"]) - 1, (r["iterations"], len(p["e3"]) - 1)
    assert abs(r["eccd"] - ref) < 1e-7 * abs(p["ccd_energies"][0]) + 20 * TOL, (r["eccd"], ref)
    assert abs(r["embpt2"] - p["ccd_energies"][0]) < 1e-7 * abs(p["ccd_energies"][0])


def test_fortran_host_tabc_all_twists_ccdt1(tmp_path):
    """my_twists = 0 with CCDT1: the four twist combinations of the ntwist=2
    deck in one batched CCDT-1 solve; the printed twist-averaged CCD(T-1)
    energy is the weighted sum of the reference program's per-twist results
    and each twist takes the program's iteration count."""
    import itertools
    exe = os.path.join(ROOT, "coupled-cluster-nuclearmatter_amd", "host", "ccm_gpu")
    if not os.path.exists(exe):
        pytest.skip("Fortran host not built (flang absent)")
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
TABC
! dens/kf, ntwist,  nmax, my_twists
0.16,2,1,0
! specify cluster approximation: CCD, CCD(T)
CCDT1
! tnf switch (T/F) and specify 3nf approximation: 0=tnf0b, 1=tnf1b, 2=tnf2b
T, 2
! 3nf cutoff(MeV),non-local reg. exp
450, 3
""")
    res = subprocess.run([exe, str(deck)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr
    x, w = np.polynomial.legendre.leggauss(2)
    w = w * np.pi / 2
    weights = [len(set(itertools.permutations((a, b, c)))) * w[a] * w[b] * w[c] / np.pi ** 3
               for a in range(2) for b in range(a, 2) for c in range(b, 2)]
    ps = [golden(f"pnm14_n1_nnlo_3nf_ccdt1_tabc2_{k}") for k in range(1, 5)]
    cc = sum(wt * (p["ccd_energies"][-1] + p["e3"][-1]) for wt, p in zip(weights, ps)) / 14
    line = [l for l in res.stdout.splitlines() if l.strip().startswith("Twist averaged CCSD")][-1]
    got = complex(*(float(v) for v in line.split("(")[1].split(")")[0].split(",")))
    assert abs(got - cc) < 1e-7 * abs(cc) + 20 * TOL / 14, (got, cc)
    rows = [l.split() for l in res.stdout.splitlines() if l.startswith("twist")]
    assert [int(r[-1]) for r in rows] == [len(p["e3"]) - 1 for p in ps]
