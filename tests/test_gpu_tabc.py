"""Twist-averaged boundary conditions through the Fortran driver ccm_gpu.

The reference's TABC branch (ccm_main.f90:161-279) enumerates the
(nx <= ny <= nz) combinations of ntwist Gauss-Legendre twists on [0, pi]
(ccm_main.f90:909-921), solves the combination my_twists and accumulates
factor * w_x w_y w_z / pi^3 * E0 (factor 1, 3 or 6 for the permutations;
the MBPT2, CCD and CCD(T) sums are commented out there and print as zero).
The fixtures tests/golden/pnm14_n1_nnlo_tabc*_k.npz hold, per combination k,
the reference's own twist (its gauss_legendre), the E0 that the reference
program ccm.exe prints for my_twists = k, and the reference's CCD trajectory
and (T) at that twist.  The weights are checked against numpy's Gauss-Legendre
rule, independent of both.
"""
import itertools
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, golden

pytestmark = pytest.mark.gpu

A = 14
EXE = os.path.join(ROOT, "coupled-cluster-nuclearmatter_amd", "host", "ccm_gpu")

DECK = """!Chiral order for Deltas(LO = 0,NLO=2,NNLO=3,N3LO=4) and cutoff
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
0.16,{n},1,{k}
! specify cluster approximation: CCD, CCD(T)
{cc}
! tnf switch (T/F) and specify 3nf approximation: 0=tnf0b, 1=tnf1b, 2=tnf2b
F, 0
! 3nf cutoff(MeV),non-local reg. exp
450, 3
"""


def combos(n):
    """(indices, weight) of the TABC combinations from numpy's rule."""
    x, w = np.polynomial.legendre.leggauss(n)
    w = w * np.pi / 2
    out = []
    for a in range(n):
        for b in range(a, n):
            for c in range(b, n):
                factor = len(set(itertools.permutations((a, b, c))))
                out.append(((a + 1, b + 1, c + 1), factor * w[a] * w[b] * w[c] / np.pi ** 3))
    return out


def run_host(tmp_path, n, k, cc):
    if not os.path.exists(EXE):
        pytest.skip("Fortran host not built (flang absent)")
    deck = tmp_path / "ccm_in"
    deck.write_text(DECK.format(n=n, k=k, cc=cc))
    res = subprocess.run([EXE, str(deck)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr
    return res.stdout


def complex_line(out, label):
    line = [l for l in out.splitlines() if l.strip().startswith(label)][-1]
    re_, im = re.search(r"\(([^,]+),([^)]+)\)", line).groups()
    return complex(float(re_), float(im))


@pytest.mark.parametrize("n,k", [(2, 1), (2, 2), (2, 3), (2, 4), (3, 5)])
def test_fortran_host_tabc_one_twist(tmp_path, n, k):
    """my_twists = k: the driver solves the k-th combination at the
    reference's twist (bit for bit) and prints the weighted E0 per particle
    that the reference program prints; MBPT2 and CCD print as zero."""
    out = run_host(tmp_path, n, k, "CCD")
    g = golden(f"pnm14_n1_nnlo_tabc{n}_{k}")
    idx, wt = combos(n)[k - 1]
    assert abs(float(g["tabc_weight"]) - wt) < 1e-14
    line = [l for l in out.splitlines() if "my twists are:" in l][-1]
    assert tuple(int(v) for v in line.split()[-3:]) == idx
    line = [l for l in out.splitlines() if "twist angles" in l][-1]
    assert [float(v) for v in line.split()[-3:]] == list(g["twist"])
    e0 = complex_line(out, "Twist averaged E0/N")
    assert abs(e0 - g["tabc_e0_per_n"]) < 1e-13 * abs(g["tabc_e0_per_n"]), (e0, g["tabc_e0_per_n"])
    assert complex_line(out, "Twist averaged mbpt2") == 0
    assert complex_line(out, "Twist averaged CCSD") == 0


def test_fortran_host_tabc_all_twists(tmp_path):
    """my_twists = 0: every combination of the deck in one batched solve,
    then (T) per twist: per twist the reference's iteration count, and the
    printed twist averages of E0, MBPT2, CCD and CCD(T) equal the weighted
    sums of the reference's per-twist results."""
    out = run_host(tmp_path, 2, 0, "CCD(T)")
    cs = combos(2)
    gs = [golden(f"pnm14_n1_nnlo_tabc2_{k}") for k in range(1, len(cs) + 1)]
    e0 = sum(w * g["e0"] for (_, w), g in zip(cs, gs)) / A
    mp = sum(w * (g["e0"] + g["energies"][0]) for (_, w), g in zip(cs, gs)) / A
    cc = sum(w * g["eccd"] for (_, w), g in zip(cs, gs)) / A
    ct = sum(w * g["eccdt"] for (_, w), g in zip(cs, gs)) / A
    assert abs(sum(g["tabc_e0_per_n"] for g in gs) - e0) < 1e-13 * abs(e0)
    assert abs(complex_line(out, "Twist averaged E0/N") - e0) < 1e-12 * abs(e0)
    assert abs(complex_line(out, "Twist averaged mbpt2") - mp) < 1e-12 * abs(mp)
    assert abs(complex_line(out, "Twist averaged CCSD") - cc) < 1e-9 * abs(cc)
    rows = [l.split() for l in out.splitlines() if l.startswith("twist")]
    assert len(rows) == len(cs)
    for row, (_, w), g in zip(rows, cs, gs):
        assert [float(v) for v in row[2:5]] == list(g["twist"])
        assert abs(float(row[5]) - w) < 1e-14
        assert int(row[-1]) == len(g["energies"]) - 1, (row, len(g["energies"]) - 1)
    line = [l for l in out.splitlines() if l.startswith("     0.01700")][-1]
    vals = [float(x) for x in line.split()]
    assert np.allclose(vals[4:], [e0.real, mp.real, cc.real, ct.real], atol=2e-5), (vals, ct)
