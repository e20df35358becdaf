"""The engine at the reference's production sizes, against the reference's own
runs (fixtures written by the compiled reference, tests/golden/make_golden.py):

* PNM A=66 |n_i|<=4 (1458 states), chiral NNLO(450) NN, DIIS 8/8, tol 1e-8 MeV,
  at rho = 0.08, 0.12, 0.16, 0.20 fm^-3 (pnm66_n4_nnlo_r*.npz: E0, the MBPT2
  start and every iteration's correlation energy of the reference's ccd_iter,
  ccm_iter.f90:1-170): the native solve, and every plan of the fast path
  (learned start, one batch, hipGraph, FP64 / V32 / FP32 contractions, the
  reference update or the Anderson update) within 1e-5 MeV per nucleon;
* the reference's own sample deck (deltafull-CCMkspace/ccm_in: SNM A=28,
  rho=0.165, |n_i|<=2, chiral NNLO(450) plus the 3NF normal-ordered to
  two-body, CCD, DIIS 6/6, tol 1e-6) through the Fortran host, against the
  reference program's result block (ccm_main.f90:322-348, ccm_in_sample.npz);
* one residual R(T) of PNM A=66 |n_i|<=2 (chiral NNLO) element by element and
  term by term against the oracle's restatement of t2_intermediate
  (ccm_t2_eqn.f90:4-1304).
"""
import os
import subprocess

import numpy as np
import pytest

import ccd_oracle as co
import chiral_oracle as cho
from conftest import ROOT, golden

pytestmark = pytest.mark.gpu

A = 66
RHO = (0.08, 0.12, 0.16, 0.20)
NAMES = [f"pnm66_n4_nnlo_r{int(round(r * 100)):03d}" for r in RHO]


@pytest.fixture(scope="module")
def lattices(ccnm):
    out = {}
    for rho, nm in zip(RHO, NAMES):
        s = ccnm.Solver(A=A, matter=ccnm.PNM, rho=rho, nmax=4)
        s.set_chiral(ccnm.NNLO, 450.0)
        s.generate_interaction(ccnm.INT_CHIRAL)
        out[nm] = s
    yield out
    for s in out.values():
        s.close()


@pytest.mark.parametrize("name", NAMES)
def test_native_nmax4_against_reference(ccnm, lattices, name):
    """E0 and the MBPT2 start to 1e-10 relative, the iterations before the
    first DIIS extrapolation to 1e-11 relative, the converged energy within the
    stopping tolerance (after an extrapolation of the nearly singular depth-8
    DIIS system any two correct implementations take different paths to the
    same fixed point)."""
    g = golden(name)
    s = lattices[name]
    assert abs(g["box"] - s.dims["box"]) < 1e-12 * g["box"]
    assert abs(s.e0() - g["e0"]) <= 1e-10 * abs(g["e0"]), (s.e0(), g["e0"])
    ref = g["energies"]
    emb = s.setup_t_amplitudes()
    assert abs(emb - ref[0]) <= 1e-10 * abs(ref[0]), (emb, ref[0])
    tol = float(g["tol"])
    r = s.ccd_iter(subspace=8, diis_step=8, tol=tol, maxit=100, use_graph=1)
    assert r["converged"] and r["status"] == 0, r
    en = s.solve_energies()
    k = min(8, len(ref), len(en))
    assert np.max(np.abs(en[:k] - ref[:k])) <= 1e-11 * np.max(np.abs(ref)), (en[:k], ref[:k])
    # the converged energy: its real part (the printed E/N, ccm_main.f90:333-339)
    # within the stopping rule; the imaginary part vanishes at the fixed point
    # and is checked as an invariant (after an extrapolation of the reference's
    # complex, un-conjugated DIIS it decays at the mixing rate)
    assert abs(en[-1].real - ref[-1].real) < 1e-9 + 20 * tol, (en[-1], ref[-1])
    assert abs(en[-1].imag) < 1e-8 * abs(en[-1].real)
    assert abs(r["eccd"].real / A - g["eccd"].real / A) < 1e-8


@pytest.mark.parametrize("name", NAMES[2:3])
def test_anderson_nmax4_reaches_reference_fixed_point(ccnm, lattices, name):
    """The Anderson update lands on the reference's converged energy (same
    fixed point of the FP64 residual) in fewer iterations."""
    g = golden(name)
    s = lattices[name]
    s.setup_t_amplitudes()
    r = s.ccd_iter(subspace=8, diis_step=8, tol=1e-8, maxit=100, use_graph=1, accel=ccnm.ACCEL_ANDERSON, mix=1.0)
    assert r["converged"] and r["status"] == 0, r
    assert abs(r["eccd"] - g["eccd"]) < 1e-9 + 20 * 1e-8, (r["eccd"], g["eccd"])
    assert r["iterations"] < len(g["energies"]) - 1


@pytest.fixture(scope="module")
def warm_net(ccnm):
    """A warm-start network trained on converged |n_i|<=2 solves at densities
    and twists other than those of the test lattices."""
    net = ccnm.WarmStart(seed=5)
    rng = np.random.default_rng(17)
    for rho in (0.07, 0.10, 0.14, 0.18, 0.22):
        s = ccnm.Solver(A=A, matter=ccnm.PNM, rho=rho, nmax=2, twist=tuple(rng.uniform(0.0, 3.0, 3)))
        s.set_chiral(ccnm.NNLO, 450.0)
        s.generate_interaction(ccnm.INT_CHIRAL)
        s.setup_t_amplitudes()
        s.ccd_iter(subspace=8, diis_step=8, tol=1e-8, maxit=100)
        net.add_solve(s)
        s.close()
    net.train(800, batch=8192, lr=2e-3, seed=3)
    return net


@pytest.mark.parametrize("accel", [0, 1])
@pytest.mark.parametrize("prec", ["fp64", "v32", "fp32", "bf16x3", "bf16x3/v32"])
def test_fast_arm_nmax4_within_1e5_per_nucleon(ccnm, lattices, warm_net, prec, accel):
    """Every plan of the fast path (learned start, the four densities in one
    batch replayed from a hipGraph, FP64 / V32 / FP32 / BF16x3 contractions
    with the stopping test in FP64, or in V32 after BF16x3, the reference
    update or Anderson) lands within 1e-5 MeV per nucleon of the reference's
    converged E/N."""
    sols = [lattices[n] for n in NAMES]
    for s in sols:
        warm_net.warm_start(s)
    names = {"fp64": ccnm.PREC_FP64, "v32": ccnm.PREC_V32, "fp32": ccnm.PREC_FP32, "bf16x3": ccnm.PREC_BF16X3}
    low, _, tail = prec.partition("/")
    res = ccnm.batch_solve(sols, subspace=8, diis_step=8, tol=1e-8, maxit=100, precision=names[low],
                           fp32_switch=1.0 if low in ("fp32", "bf16x3") else 0.0, use_graph=1, accel=accel,
                           mix=1.0, tail=names[tail or "fp64"])
    for r, nm in zip(res, NAMES):
        g = golden(nm)
        assert r["converged"], (nm, r)
        assert r["iterations"] > r["fp32_iterations"]   # the stopping test is an FP64 iteration
        assert abs(r["eccd"].real - g["eccd"].real) / A <= 1e-5, (nm, r["eccd"], g["eccd"])
        if prec != "fp64":
            assert r["fp32_iterations"] >= 1


def test_fortran_host_reference_sample_deck(ccnm, tmp_path):
    """The reference's own input deck (ccm_in of deltafull-CCMkspace, the one
    deck it ships) through the Fortran driver ccm_gpu: the printed result line
    (cE, cD, kF, rho, E0/N, MBPT2/N, CCD/N, CCD(T)/N) equals the reference
    program's at its printed precision."""
    exe = os.path.join(ROOT, "coupled-cluster-nuclearmatter_amd", "host", "ccm_gpu")
    if not os.path.exists(exe):
        pytest.skip("Fortran host not built (flang absent)")
    g = golden("ccm_in_sample")
    deck = tmp_path / "ccm_in"
    # the deck's values, in its layout (positional reads, ccm_main.f90:31-37, 893-900)
    deck.write_text("""!Chiral order for Deltas(LO = 0,NLO=2,NNLO=3,N3LO=4) and cutoff
3, 450
! cE and cD 3nf parameters:
 0.0170000000000000012, 0.790000000000000036
! number of particles
28
! specify: pnm/snm, input type: density/kfermi
snm, density
! specify boundary conditions (PBC/TABC/TABCsp)
PBC
! dens/kf, ntwist,  nmax
0.165,1,2,1
! specify cluster approximation: CCD, CCDT
CCD
! tnf switch (T/F) and specify 3nf approximation: 0=tnf0b, 1=tnf1b, 2=tnf2b
T, 2
! 3nf cutoff(MeV),non-local reg. exp
450, 3
""")
    res = subprocess.run([exe, str(deck)], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-2000:]
    line = [x for x in res.stdout.splitlines() if x.startswith("     0.01700")][-1]
    vals = np.array([float(x) for x in line.split()])
    ref = g["result_line"]
    assert np.allclose(vals, ref, atol=1.5e-5), (vals, ref)


def test_native_sample_deck_trajectory(ccnm):
    """The same deck through the C ABI: E0, the MBPT2 start and every iteration
    of the reference program's 'CCD energy' lines (E0 plus the correlation
    energy) up to the first DIIS extrapolation, and the converged energy."""
    g = golden("ccm_in_sample")
    s = ccnm.Solver(A=28, matter=ccnm.SNM, rho=0.165, nmax=2)
    s.set_chiral(ccnm.NNLO, 450.0)
    s.set_3nf(2, 0.0170000000000000012, 0.790000000000000036, 450.0, 3)
    s.generate_interaction(ccnm.INT_CHIRAL)
    e0 = s.e0()
    # E0 with the normal-ordered 3NF: the 'Twist averaged E0/N' line (the
    # 'Vacuum expectation value' line is printed before the 3NF is added)
    assert abs(e0 / 28 - g["e0_per_n"]) < 1e-10 * abs(g["e0_per_n"])
    emb = s.setup_t_amplitudes()
    ref = g["ccd_energies"]           # E0 + correlation energy: start, then per iteration
    # the float32 2/3 of the 3NF regulator makes the normal-ordered blocks
    # symmetric only to ~1e-7 (the reference fills one triangle, the engine the
    # same one): the start and early iterations agree to that level
    assert abs(e0 + emb - ref[0]) < 2e-7 * abs(ref[0])
    r = s.ccd_iter(subspace=6, diis_step=6, tol=float(np.float32(1e-6)), maxit=100, use_graph=1)
    assert r["converged"]
    en = e0 + s.solve_energies()
    k = min(6, len(ref), len(en))
    assert np.max(np.abs(en[:k] - ref[:k])) < 2e-7 * np.max(np.abs(ref)), (en[:k], ref[:k])
    tol = float(np.float32(1e-6))
    assert abs(r["eccd"] - ref[-1]) < 1e-9 + 20 * tol, (r["eccd"], ref[-1])
    assert abs(r["eccd"].real / 28 - g["eccd_per_n"].real) < (1e-9 + 20 * tol) / 28
    s.close()


@pytest.fixture(scope="module")
def oracle_n2():
    lat = co.Lattice.build(A, "pnm", 0.16, 2)
    return co.CCD(lat, cho.ChiralInteraction(lat, cho.NNLO, 450.0))


@pytest.mark.parametrize("terms", [0, 1, 2, 4, 8, 15])
@pytest.mark.parametrize("start", ["mbpt2", "iterated"])
def test_residual_terms_elementwise_nmax2(ccnm, oracle_n2, monkeypatch, terms, start):
    """One residual R(T) of PNM A=66 |n_i|<=2 (chiral NNLO(450)), the lattice of
    the CPU reference configuration, element by element against the oracle:
    the conj(V) start with the one-body terms (1), the pp ladder (2), the hh
    ladder with I9 (4), the ring (8) and all of them, plus the dynamic
    denominators, at the MBPT2 start and at amplitudes after 12 Anderson
    iterations.  Bound: 1e-12 of the largest element."""
    cc = oracle_n2
    s = ccnm.Solver(A=A, matter=ccnm.PNM, rho=0.16, nmax=2)
    s.set_chiral(ccnm.NNLO, 450.0)
    s.generate_interaction(ccnm.INT_CHIRAL)
    s.setup_t_amplitudes()
    if start == "iterated":
        s.ccd_iter(subspace=8, diis_step=8, tol=0.0, maxit=12, accel=ccnm.ACCEL_ANDERSON, mix=1.0)
    t = s.t2_unique()
    lab = s.t2_labels()
    a, b, i, j = lab[:, 0], lab[:, 1], lab[:, 2], lab[:, 3]
    # the oracle's ordered amplitudes from the unique ones (antisymmetry)
    x = np.zeros(cc.ch.size, np.complex128)
    for (p, q, r_, u), sg in (((a, b, i, j), 1.0), ((b, a, i, j), -1.0), ((a, b, j, i), -1.0), ((b, a, j, i), 1.0)):
        idx = cc.ch.flat_index(p, q, r_, u)
        assert np.all(idx >= 0)
        x[idx] = sg * t
    R, dh, dp = cc.residual(cc.unflat(x), terms)
    ref = cc.flat(R)[cc.ch.flat_index(a, b, i, j)]
    monkeypatch.setenv("CCGPU_TERMS", str(terms))
    s.set_t2_unique(t)
    s.t2_intermediate()
    got = s.debug_fetch("R", len(t))
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(got - ref)) <= 1e-12 * scale, np.max(np.abs(got - ref)) / scale
    gdh = s.debug_fetch("dh", A)
    gdp = s.debug_fetch("dp", s.dims["nparticles"])
    assert np.max(np.abs(gdh - dh)) <= 1e-12 * np.max(np.abs(dh))
    assert np.max(np.abs(gdp - dp[A:])) <= 1e-12 * np.max(np.abs(dp[A:]))
    s.close()


def _abs_oracle(cc):
    """The oracle with every interaction block replaced by its modulus: its
    ladder terms of |T| bound, element by element, the sums of |products| that
    the ladder contractions of the true residual add up."""
    import copy
    ca = copy.copy(cc)
    for nm in ("v_hhpp", "v_pppp", "v_hhhh", "vtil", "w0"):
        setattr(ca, nm, [np.abs(x) for x in getattr(cc, nm)])
    return ca


def _ring_abs(cc, xabs):
    """Sum of |products| behind every element of the ring term (the oracle's
    Pandya chain T~ -> I4 = 1/2 V~ T~ + W0 -> R~ = T~ I4 -> four scattered
    adds, ccm_t2_eqn.f90:657-1185) with moduli throughout and the four adds
    taken with sign +1, so cancellations between them do not hide the error
    each one carries."""
    ch = cc.ch
    out = np.zeros(ch.size)
    for K in range(len(ch.ph_keys)):
        ph, hp = ch.ph[K], ch.hp[K]
        a, i = ph[:, 0][:, None], ph[:, 1][:, None]
        j, b = hp[:, 0][None, :], hp[:, 1][None, :]
        idx = ch.flat_index(a, b, i, j)
        tt = np.where(idx >= 0, xabs[np.maximum(idx, 0)], 0.0)
        rt = tt @ (0.5 * (np.abs(cc.vtil[K]) @ tt) + np.abs(cc.w0[K]))
        for aa, bb, ii, jj in ((a, b, i, j), (b, a, i, j), (a, b, j, i), (b, a, j, i)):
            idx2 = ch.flat_index(aa, bb, ii, jj)
            m = idx2 >= 0
            np.add.at(out, idx2[m], rt[m])
    return out


@pytest.mark.parametrize("terms", [2, 4, 8])
def test_bf16x3_residual_within_split_product_bound(ccnm, oracle_n2, monkeypatch, terms):
    """The BF16x3 residual (every contraction on split bf16 products with FP32
    sums; the pp ladder on the FP32 copy of V, the ring on FP32 Pandya blocks)
    against the FP64 residual, element by element, on PNM A=66 |n_i|<=2
    (chiral NNLO(450)) at amplitudes after 12 Anderson iterations.

    Bound: a real product x*y summed as xh*yh + xh*yl + xl*yh, with
    xh = bf16(x) and xl = bf16(x - xh), is within 3 * 2^-16 |x y| of x*y
    (2^-16 |x| from each split, 2^-16 |x y| for the dropped xl*yl); an FP32
    sum of K products adds at most K * 2^-24 of the sum of their moduli, and
    the FP64 -> FP32 conversion 2^-24 per operand.  The hh ladder and the ring
    are two GEMMs in a row (I9 or I4, then R), which doubles the bound, so
        |R_bf16x3 - R_fp64| <= 2 eps |R|(|V|, |T|),
        eps = 3 * 2^-16 + (K + 2) * 2^-24,
    where |R|(|V|, |T|) is the oracle's residual with every block and
    amplitude replaced by its modulus and K the longest contraction.  The pp
    ladder is one GEMM on FP32-rounded V: eps = 3 * 2^-16 + (K + 3) * 2^-24.
    The BF16x3 kernels must have run, and the error must be far below that of
    plain bf16 products (2^-8)."""
    cc = oracle_n2
    s = ccnm.Solver(A=A, matter=ccnm.PNM, rho=0.16, nmax=2)
    s.set_chiral(ccnm.NNLO, 450.0)
    s.generate_interaction(ccnm.INT_CHIRAL)
    s.setup_t_amplitudes()
    s.ccd_iter(subspace=8, diis_step=8, tol=0.0, maxit=12, accel=ccnm.ACCEL_ANDERSON, mix=1.0)
    t = s.t2_unique()
    lab = s.t2_labels()
    a, b, i, j = lab[:, 0], lab[:, 1], lab[:, 2], lab[:, 3]
    x = np.zeros(cc.ch.size, np.complex128)
    for (p, q, r_, u), sg in (((a, b, i, j), 1.0), ((b, a, i, j), -1.0), ((a, b, j, i), -1.0), ((b, a, j, i), 1.0)):
        x[cc.ch.flat_index(p, q, r_, u)] = sg * t
    if terms == 8:
        bound_abs = _ring_abs(cc, np.abs(x))[cc.ch.flat_index(a, b, i, j)]
    else:
        ca = _abs_oracle(cc)
        Rabs, _, _ = ca.residual(ca.unflat(np.abs(x)), terms)
        bound_abs = np.real(ca.flat(Rabs)[cc.ch.flat_index(a, b, i, j)])
    monkeypatch.setenv("CCGPU_TERMS", str(terms))
    s.set_t2_unique(t)
    s.t2_intermediate(ccnm.PREC_FP64)
    r64 = s.debug_fetch("R", len(t))
    monkeypatch.setenv("CCGPU_TIMING", "2")   # names every launched GEMM class with its kernel kind
    s.t2_intermediate(ccnm.PREC_BF16X3)
    rbf = s.debug_fetch("R", len(t))
    launched = list(ccnm.kernel_stats())
    monkeypatch.delenv("CCGPU_TIMING")
    if terms == 2:
        eps = 3 * 2.0**-16 + (s.dims["max_npp"] + 3) * 2.0**-24
    else:
        # longest contraction: a pp block (X = V_hhpp T), or a particle-hole
        # channel (both ring GEMMs run over ph or hp pairs of one channel)
        kring = max(max(len(x) for x in cc.ch.ph), max(len(x) for x in cc.ch.hp))
        kmax = max(s.dims["max_npp"], kring)
        eps = 2 * (3 * 2.0**-16 + (kmax + 2) * 2.0**-24)
        assert any(".k6." in k for k in launched), launched
    err = np.abs(rbf - r64)
    assert np.all(np.isfinite(rbf))
    assert np.all(err <= eps * bound_abs + 1e-12 * np.abs(r64).max()), (err / np.maximum(bound_abs, 1e-300)).max()
    assert err.max() < 2.0**-12 * np.abs(r64).max()
    s.close()



def test_bf16x3_ladder_split_copy_of_t_gives_the_same_bits(ccnm, monkeypatch):
    """The BF16x3 pp ladder reading T split once per residual (k_t_split, kernel
    kind 10) against the ladder that splits T itself (kind 9,
    CCGPU_PPL_TSPLIT=0): the same residual bit for bit, whole residual and
    ladder alone, and each run launched the form it names."""
    s = ccnm.Solver(A=A, matter=ccnm.PNM, rho=0.16, nmax=2)
    s.set_chiral(ccnm.NNLO, 450.0)
    s.generate_interaction(ccnm.INT_CHIRAL)
    s.setup_t_amplitudes()
    s.ccd_iter(subspace=8, diis_step=8, tol=0.0, maxit=5, accel=ccnm.ACCEL_ANDERSON, mix=1.0)
    t = s.t2_unique()
    monkeypatch.setenv("CCGPU_TIMING", "2")
    for terms in ("2", "15"):
        monkeypatch.setenv("CCGPU_TERMS", terms)
        out = {}
        for split in ("0", "1"):
            monkeypatch.setenv("CCGPU_PPL_TSPLIT", split)
            s.set_t2_unique(t)
            s.t2_intermediate(ccnm.PREC_BF16X3)
            out[split] = s.debug_fetch("R", len(t))
            launched = list(ccnm.kernel_stats())
            assert any((".k10." if split == "1" else ".k9.") in k for k in launched), launched
        assert np.array_equal(out["0"], out["1"])
        assert np.any(out["1"] != 0)
    s.close()
