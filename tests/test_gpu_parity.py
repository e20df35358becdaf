"""GPU parity: the HIP engine (through its C ABI) against the compiled reference's
golden fixtures and against the oracle, on identical inputs."""
import os

import numpy as np
import pytest

import ccd_oracle as co
from conftest import GOLDEN, fixture_interaction, fixture_lattice, golden

pytestmark = pytest.mark.gpu

MINN_CASES = ["pnm14_n1_minn", "snm4_n1_minn", "pnm14_n1_minn_twist", "pnm66_n2_minn",
              "pnm66_n2_minn_ref", "pnm66_n2_minn_twist", "snm28_n2_minn"]


def params(g):
    A, nmax, sub, step, maxit, order = (int(x) for x in g["params"])
    return A, nmax, sub, step, maxit


def step_loop(s, sub, step, tol, maxit, precision=0):
    """ccd_iter (ccm_iter.f90:60-96) written against the step API."""
    e1 = s.ccd_energy_save()
    out = [e1]
    count = nstep = 0
    dener = 4.0
    for it in range(1, maxit + 1):
        if not abs(dener) > tol:
            continue
        s.t2_intermediate(precision)
        count += 1
        nstep += 1
        count, nstep = s.diis_t1t2(count, nstep, sub, step)
        e2 = s.ccd_energy_save()
        dener = e2 - e1
        e1 = e2
        out.append(e2)
    return np.array(out)


def check_trajectory(en, ref, step, tol, strict=False):
    """Iterations before the first DIIS extrapolation agree at rounding level.
    The reference's DIIS Gram matrix is nearly singular (condition ~1e16 at
    depth 8: geometric error vectors), so after a DIIS step any two
    implementations follow different but equally valid paths to the same fixed
    point; there only the converged energy is compared, at the scale set by the
    stopping tolerance.  strict=True (well-conditioned DIIS) compares every
    iteration."""
    scale = np.max(np.abs(ref))
    k = min(step, len(ref), len(en))
    assert np.max(np.abs(en[:k] - ref[:k])) < 1e-11 * scale
    if strict:
        assert len(en) == len(ref), (len(en), len(ref))
        assert np.max(np.abs(en - ref)) < 1e-9 * scale
    assert abs(en[-1] - ref[-1]) < 1e-9 + 20 * tol, (en[-1], ref[-1])


@pytest.mark.parametrize("name", MINN_CASES)
def test_native_minnesota_against_reference(ccnm, name):
    """Lattice, Minnesota V, Fock, E0, MBPT2 and the whole CCD trajectory built by
    the engine itself versus the reference run on the same physical input."""
    g = golden(name)
    A, nmax, sub, step, maxit = params(g)
    matter = ccnm.PNM if str(g["matter"]) == "pnm" else ccnm.SNM
    s = ccnm.Solver(A=A, matter=matter, rho=float(g["rho"]), nmax=nmax, twist=tuple(g["twist"]))
    assert abs(s.dims["box"] - float(g["box"])) < 1e-12
    s.generate_interaction(ccnm.INT_MINNESOTA)
    assert abs(s.e0() - g["e0"]) < 1e-10 * abs(g["e0"])
    emp2 = s.setup_t_amplitudes()
    assert abs(emp2 - g["energies"][0]) < 1e-11 * abs(g["energies"][0])
    en = step_loop(s, sub, step, float(g["tol"]), maxit)
    check_trajectory(en, g["energies"], step, float(g["tol"]))


def hand_over(s, kind, rows, vals):
    """One reference-layout block from records rows[:, 0:4] (1-based orbitals of
    the bra pair and the ket pair) and their values."""
    bra = np.unique(rows[:, 0:2], axis=0)
    ket = np.unique(rows[:, 2:4], axis=0)
    bpos = {tuple(x): n for n, x in enumerate(bra)}
    kpos = {tuple(x): n for n, x in enumerate(ket)}
    val = np.zeros((len(bra), len(ket)), np.complex128)
    for row, v in zip(rows, vals):
        val[bpos[(row[0], row[1])], kpos[(row[2], row[3])]] = v
    s.set_block(kind, bra[:, 0], bra[:, 1], ket[:, 0], ket[:, 1], val)


def hand_over_blocks(ccnm, s, g, triples=False):
    idx = g["v_idx"].astype(np.int64)
    kinds = {1: ccnm.BLOCK_HHHH, 3: ccnm.BLOCK_HHPP, 4: ccnm.BLOCK_HPHP, 6: ccnm.BLOCK_PPPP}
    if triples:
        kinds[2] = ccnm.BLOCK_HHHP
    for k, kind in kinds.items():
        sel = idx[:, 0] == k
        hand_over(s, kind, idx[sel][:, 1:5], g["v_val"][sel])
    if triples:
        # vnn_hppp records (k, d, c, b) hold <cb||kd>: rows (k,d), columns (c,b)
        hand_over(s, ccnm.BLOCK_HPPP, g["hppp_idx"].astype(np.int64), g["hppp_val"])


DUMPED = sorted(f[:-4] for f in os.listdir(GOLDEN)
                if f.endswith(".npz") and "orb_int" in np.load(os.path.join(GOLDEN, f)).files)


@pytest.mark.parametrize("name", DUMPED)
def test_native_basis_equals_reference_orbital_table(ccnm, name):
    """The engine's own basis (ccgpu_create from the deck parameters) is the
    reference's orbital table state for state, degenerate momenta in the
    order of its quicksort."""
    g = golden(name)
    A, nmax = int(g["params"][0]), int(g["params"][1])
    s = ccnm.Solver(A=A, matter=ccnm.PNM if str(g["matter"]) == "pnm" else ccnm.SNM,
                    rho=float(g["rho"]), nmax=nmax, twist=tuple(g["twist"]))
    o = s.orbitals()
    oi, od = g["orb_int"], g["orb_dbl"]
    for r, k in enumerate(("nx", "ny", "nz", "sz", "tz")):
        assert np.array_equal(o[k], oi[r]), k
    assert np.allclose(o["e"], od[3], rtol=1e-15, atol=0)


@pytest.mark.parametrize("name", ["pnm14_n1_nnlo", "snm4_n1_nnlo", "pnm14_n1_minn"])
def test_dropin_reference_blocks(ccnm, name):
    """Drop-in path: the reference's own orbitals, V blocks (chiral NNLO here) and
    Fock matrix handed over through ccgpu_set_block / ccgpu_set_fock."""
    g = golden(name)
    A, nmax, sub, step, maxit = params(g)
    oi, od = g["orb_int"], g["orb_dbl"]
    orbs = {"nx": oi[0], "ny": oi[1], "nz": oi[2], "sz": oi[3], "tz": oi[4],
            "kx": od[0], "ky": od[1], "kz": od[2], "e": od[3], "A": A,
            "box": float(g["box"]), "volume": float(g["volume"])}
    s = ccnm.Solver(orbitals=orbs)
    hand_over_blocks(ccnm, s, g)
    s.set_fock(g["fock"], g["e0"])
    s.finalize_setup()
    emp2 = s.setup_t_amplitudes()
    assert abs(emp2 - g["energies"][0]) < 1e-12 * abs(g["energies"][0])
    en = step_loop(s, sub, step, float(g["tol"]), maxit)
    check_trajectory(en, g["energies"], step, float(g["tol"]))
    # converged amplitudes by orbital labels
    t = g["t2_idx"].astype(np.int64) - 1
    mine = s.get_t2(t[:, 0], t[:, 1], t[:, 2], t[:, 3])
    assert np.max(np.abs(mine - g["t2_val"])) < 1e-6 * np.max(np.abs(g["t2_val"])) + float(g["tol"])


def test_solve_equals_step_api_and_oracle(ccnm):
    """ccgpu_solve (library-side loop) and the step API give the same energies,
    and both agree with the oracle's ordered-pair restatement."""
    s = ccnm.Solver(A=14, matter=ccnm.PNM, rho=0.16, nmax=1)
    s.generate_interaction()
    s.setup_t_amplitudes()
    en = step_loop(s, 6, 6, 1e-8, 100)
    s2 = ccnm.Solver(A=14, matter=ccnm.PNM, rho=0.16, nmax=1)
    s2.generate_interaction()
    s2.setup_t_amplitudes()
    r = s2.ccd_iter(subspace=6, diis_step=6, tol=1e-8, maxit=100)
    # same kernels, same reductions: the library loop reproduces the step API exactly
    assert r["iterations"] == len(en) - 1
    assert r["ecorr"] == en[-1]
    lat = co.Lattice.build(14, "pnm", 0.16, 1)
    cc = co.CCD(lat, co.MinnesotaInteraction(lat))
    ref = np.array(cc.solve(6, 6, 1e-8, 100))
    check_trajectory(en, ref, 6, 1e-8)


def test_diis_trajectory_matches_oracle_when_well_conditioned(ccnm):
    """With a two-vector DIIS (subspace 3, every 3rd iteration) the Gram matrix
    is well conditioned and the whole trajectory, DIIS steps included, must
    match the oracle iteration by iteration."""
    for A, matter, nmax in ((14, "pnm", 1), (4, "snm", 1)):
        s = ccnm.Solver(A=A, matter=ccnm.PNM if matter == "pnm" else ccnm.SNM, rho=0.16, nmax=nmax)
        s.generate_interaction()
        s.setup_t_amplitudes()
        en = step_loop(s, 3, 3, 1e-9, 60)
        lat = co.Lattice.build(A, matter, 0.16, nmax)
        cc = co.CCD(lat, co.MinnesotaInteraction(lat))
        ref = np.array(cc.solve(3, 3, 1e-9, 60))
        check_trajectory(en, ref, 3, 1e-9, strict=True)
    # amplitudes by labels against the oracle's converged T2
    lab = s.t2_labels()
    mine = s.t2_unique()
    idx = cc.ch.flat_index(lab[:, 0], lab[:, 1], lab[:, 2], lab[:, 3])
    assert np.all(idx >= 0)
    theirs = cc.flat(cc.T)[idx]
    assert np.max(np.abs(mine - theirs)) < 1e-8 * np.max(np.abs(theirs))


@pytest.mark.parametrize("split", ["default", "all"])
def test_mfma_gemm_matches_valu_gemm(ccnm, monkeypatch, split):
    """The MFMA grouped GEMM against the plain VALU GEMM (independent lane
    layout, whole k range per tile) on the PNM A=66, nmax=2 lattice.  With
    split="all", on the nmax=3 lattice (pp pairs up to 620 per channel), every
    group with k >= 512 outside the pp ladder (which never splits) is split 8
    ways (split quantum lowered to 64 k), so the split-K partials and their
    fixed-order sum are checked against the VALU result as well."""
    nmax = 2
    if split == "all":
        monkeypatch.setenv("CCGPU_SPLIT_TILES", "1000000")
        monkeypatch.setenv("CCGPU_SPLIT_K", "64")
        nmax = 3
    res = []
    for mode in ("mfma", "valu"):
        if mode == "valu":
            os.environ["CCGPU_GEMM"] = "valu"
        try:
            s = ccnm.Solver(A=66, matter=ccnm.PNM, rho=0.16, nmax=nmax)
            s.generate_interaction()
            s.setup_t_amplitudes()
            res.append(step_loop(s, 8, 8, 1e-8, 100))
            s.close()
        finally:
            os.environ.pop("CCGPU_GEMM", None)
    a, b = res
    k = 8
    assert np.max(np.abs(a[:k] - b[:k])) < 1e-12 * np.max(np.abs(a))
    assert abs(a[-1] - b[-1]) < 1e-9 + 20 * 1e-8


def test_graph_replay_matches_eager(ccnm):
    out = []
    for graph in (0, 1):
        s = ccnm.Solver(A=66, matter=ccnm.PNM, rho=0.16, nmax=2)
        s.generate_interaction()
        s.setup_t_amplitudes()
        out.append(s.ccd_iter(subspace=8, diis_step=8, tol=1e-8, maxit=100, use_graph=graph))
    assert out[0]["iterations"] == out[1]["iterations"]
    assert out[0]["eccd"] == out[1]["eccd"]


@pytest.mark.parametrize("share", ["0", "1"])
def test_batch_matches_sequential(ccnm, monkeypatch, share):
    """Three twists in one batched launch sequence versus three separate solves.
    With one pp ladder per twist (CCGPU_SHARE_V=0) the kernels and reduction
    orders per solve are those of the separate solves: bitwise equal.  With
    the shared ladder (one V_pppp over the union of the twists' pairs) the
    ladder sums the same products in another order: equal to 1e-12."""
    monkeypatch.setenv("CCGPU_SHARE_V", share)
    twists = [(0.0, 0.0, 0.0), (0.7, 1.9, 2.6), (2.1, 0.4, 1.3)]
    seq = []
    for tw in twists:
        s = ccnm.Solver(A=66, matter=ccnm.PNM, rho=0.16, nmax=2, twist=tw)
        s.generate_interaction()
        s.setup_t_amplitudes()
        seq.append(s.ccd_iter(subspace=8, diis_step=8, tol=1e-8, maxit=100))
    ss = []
    for tw in twists:
        s = ccnm.Solver(A=66, matter=ccnm.PNM, rho=0.16, nmax=2, twist=tw)
        s.generate_interaction()
        s.setup_t_amplitudes()
        ss.append(s)
    bat = ccnm.batch_solve(ss, subspace=8, diis_step=8, tol=1e-8, maxit=100)
    for a, b in zip(seq, bat):
        assert a["iterations"] == b["iterations"]
        if share == "0":
            assert a["eccd"] == b["eccd"]
        else:
            assert abs(a["eccd"] - b["eccd"]) <= 1e-12 * abs(a["eccd"])
    g = golden("pnm66_n2_minn_twist")
    assert abs(bat[1]["eccd"] - g["eccd"]) < 1e-9 + 20 * float(g["tol"])


def test_fp32_mixed_precision_error(ccnm):
    """FP32 contractions for the first DIIS cycle, FP64 afterwards: the solve
    converges under the FP64 test and E/N lies within 1e-5 MeV of plain FP64
    (the tolerance of the mixed-precision arm; in practice ~1e-10)."""
    out = {}
    for prec in (ccnm.PREC_FP64, ccnm.PREC_FP32):
        s = ccnm.Solver(A=66, matter=ccnm.PNM, rho=0.16, nmax=2)
        s.generate_interaction()
        s.setup_t_amplitudes()
        out[prec] = s.ccd_iter(subspace=8, diis_step=8, tol=1e-8, maxit=100, precision=prec, fp32_switch=1.0)
    d = abs(out[0]["eccd"] - out[1]["eccd"]) / 66
    assert out[1]["converged"]
    assert out[1]["fp32_iterations"] >= 8, out[1]
    assert out[0]["fp32_iterations"] == 0
    assert out[1]["iterations"] - out[1]["fp32_iterations"] >= 1
    assert d < 1e-5, d


def set_reference_t2(ccnm, s, g):
    """Load the reference's converged T2 (its own orbital numbering) into the
    engine's solver by matching orbitals through their labels (n, sz, tz)."""
    mine = s.orbitals()
    lab = {(mine["nx"][x], mine["ny"][x], mine["nz"][x], mine["sz"][x], mine["tz"][x]): x
           for x in range(len(mine["e"]))}
    oi = g["orb_int"]
    ref2mine = np.array([lab[(oi[0][x], oi[1][x], oi[2][x], oi[3][x], oi[4][x])] for x in range(oi.shape[1])])
    t = g["t2_idx"].astype(np.int64) - 1
    s.set_t2(ref2mine[t[:, 0]], ref2mine[t[:, 1]], ref2mine[t[:, 2]], ref2mine[t[:, 3]], g["t2_val"])


@pytest.mark.parametrize("name", ["pnm14_n1_minn", "snm4_n1_minn", "pnm14_n1_minn_twist"])
def test_triples_against_reference(ccnm, name):
    """(T) of the GPU kernel on the reference's converged T2 versus the
    reference's pert_triples2 energy; and on the engine's own converged T2."""
    g = golden(name)
    A, nmax, sub, step, maxit = params(g)
    matter = ccnm.PNM if str(g["matter"]) == "pnm" else ccnm.SNM
    s = ccnm.Solver(A=A, matter=matter, rho=float(g["rho"]), nmax=nmax, twist=tuple(g["twist"]))
    s.generate_interaction()
    s.setup_t_amplitudes()
    set_reference_t2(ccnm, s, g)
    et = s.pert_triples2()
    ref = g["eccdt"] - g["eccd"]
    assert abs(et - ref) < 1e-12 * max(1.0, abs(ref)), (et, ref)
    # element count: brute force over (a<b<c, i<j<k) with equal total momentum and isospin
    o = s.orbitals()
    A_, nsp = s.dims["nholes"], s.dims["nsp"]
    key = {}
    for x in range(nsp):
        key[x] = (o["nx"][x], o["ny"][x], o["nz"][x], o["tz"][x])
    import itertools
    hk = {}
    for i, j, k in itertools.combinations(range(A_), 3):
        kk = tuple(key[i][q] + key[j][q] + key[k][q] for q in range(4))
        hk[kk] = hk.get(kk, 0) + 1
    want = 0
    for a, b, c in itertools.combinations(range(A_, nsp), 3):
        want += hk.get(tuple(key[a][q] + key[b][q] + key[c][q] for q in range(4)), 0)
    assert s.triples_stats()[0] == want
    # own converged amplitudes: agreement at the stopping-tolerance scale
    s.setup_t_amplitudes()
    r = s.ccd_iter(subspace=sub, diis_step=step, tol=float(g["tol"]), maxit=maxit)
    et2 = s.pert_triples2()
    assert abs(et2 - ref) < 1e-6 * abs(ref) + float(g["tol"])


@pytest.mark.parametrize("name", ["snm4_n1_nnlo", "snm4_n1_minn", "pnm14_n1_minn_twist"])
def test_triples_dropin_blocks(ccnm, name):
    """(T) from handed-over blocks (the reference's chiral NNLO or Minnesota
    hhhp and hppp elements) on the reference's converged T2, against the
    reference's pert_triples2 energy."""
    g = golden(name)
    A = int(g["params"][0])
    oi, od = g["orb_int"], g["orb_dbl"]
    orbs = {"nx": oi[0], "ny": oi[1], "nz": oi[2], "sz": oi[3], "tz": oi[4],
            "kx": od[0], "ky": od[1], "kz": od[2], "e": od[3], "A": A,
            "box": float(g["box"]), "volume": float(g["volume"])}
    s = ccnm.Solver(orbitals=orbs)
    hand_over_blocks(ccnm, s, g, triples=True)
    s.set_fock(g["fock"], g["e0"])
    s.finalize_setup()
    s.setup_t_amplitudes()
    set_reference_t2(ccnm, s, g)
    et = s.pert_triples2()
    ref = g["eccdt"] - g["eccd"]
    assert abs(et - ref) < 1e-11 * max(1.0, abs(ref)), (et, ref)


def test_fp32_until_tolerance_with_fp64_confirmation(ccnm):
    """fp32_switch <= 0: FP32 contractions until |dE| <= tol, then FP64
    iterations until an FP64 |dE| <= tol; E/N within 1e-5 MeV of plain FP64."""
    ref = None
    for prec, tol in ((ccnm.PREC_FP64, 1e-8), (ccnm.PREC_FP32, 1e-5)):
        s = ccnm.Solver(A=66, matter=ccnm.PNM, rho=0.16, nmax=2)
        s.generate_interaction()
        s.setup_t_amplitudes()
        out = s.ccd_iter(subspace=8, diis_step=8, tol=tol, maxit=100, precision=prec, fp32_switch=0.0)
        if ref is None:
            ref = out
            continue
    assert out["converged"]
    assert out["fp32_iterations"] >= 8
    assert out["iterations"] > out["fp32_iterations"]
    assert abs(out["eccd"] - ref["eccd"]) / 66 < 1e-5


def test_step_api_fp32_residual(ccnm):
    """t2_intermediate in FP32 through the step API: the residual agrees with
    the FP64 one to FP32 rounding (relative 1e-5) and is not bit-identical."""
    s = ccnm.Solver(A=66, matter=ccnm.PNM, rho=0.16, nmax=2)
    s.generate_interaction()
    s.setup_t_amplitudes()
    n = s.dims["t2_unique"]
    s.t2_intermediate(ccnm.PREC_FP64)
    r64 = s.debug_fetch("R", n)
    s.t2_intermediate(ccnm.PREC_FP32)
    r32 = s.debug_fetch("R", n)
    rel = np.abs(r32 - r64).max() / np.abs(r64).max()
    assert np.all(np.isfinite(r32))
    assert 0 < rel < 1e-5, rel


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_fortran_host_drop_in(ccnm, tmp_path, prec):
    """The Fortran driver ccm_gpu reads a reference ccm_in deck (plus the engine
    option line) and prints the reference's result block; E/N values match the
    reference run of the same physics (PNM A=14, nmax=1, Minnesota), in plain
    FP64 and with the FP32 first DIIS cycle."""
    import subprocess
    from conftest import ROOT
    exe = os.path.join(ROOT, "coupled-cluster-nuclearmatter_amd", "host", "ccm_gpu")
    if not os.path.exists(exe):
        pytest.skip("Fortran host not built (flang absent)")
    g = golden("pnm14_n1_minn")
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
F, 0
! 3nf cutoff(MeV),non-local reg. exp
450, 3
! interaction, DIIS subspace, DIIS step, tolerance, max iterations, precision, fp32 switch
minnesota, 8, 8, 1e-8, 100, PREC, 1.0
""".replace("PREC", prec))
    res = subprocess.run([exe, str(deck)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    line = [l for l in res.stdout.splitlines() if l.startswith("     0.01700")][-1]
    vals = [float(x) for x in line.split()]
    A = 14
    ref = [0.017, 0.79, float(g["kf"]), 0.16, g["e0"].real / A, (g["e0"] + g["energies"][0]).real / A,
           g["eccd"].real / A, g["eccdt"].real / A]
    assert np.allclose(vals, ref, atol=2e-5), (vals, ref)


def test_triples_parity_table_equals_full_table(ccnm, monkeypatch):
    """The <cb||kd> table filled over one element of each parity pair (no
    twist: n -> -n maps the lattice and the Fermi sea onto themselves) against
    the table with every element evaluated, through (T) with the normal-ordered
    3NF on SNM A=28 |n_i|<=2. The image's 3NF sum over holes runs in another
    order, so the two agree to rounding, not bit for bit."""
    out = {}
    for par in ("0", "1"):
        monkeypatch.setenv("CCGPU_T3_PARITY", par)
        s = ccnm.Solver(A=28, matter=ccnm.SNM, rho=0.16, nmax=2)
        s.set_chiral(ccnm.NNLO, 450.0)
        s.set_3nf(2, 0.017, 0.79, 450.0, 3)
        s.generate_interaction(ccnm.INT_CHIRAL)
        s.setup_t_amplitudes()
        assert s.ccd_iter(subspace=8, diis_step=8, tol=1e-8, maxit=100)["converged"]
        out[par] = s.pert_triples2()
        s.close()
    assert out["0"] != 0
    assert abs(out["1"] - out["0"]) <= 1e-13 * abs(out["0"]), out


def test_3nf_vpppp_parity_fill_equals_full_fill(ccnm, monkeypatch):
    """V_pppp's normal-ordered 3NF filled over one element of each parity
    pair against every element evaluated (SNM A=28 |n_i|<=2, no twist). The
    image's sum over holes runs in another order, so the two agree to
    rounding: the MBPT2 energy and the iterations before the first DIIS step
    to 1e-12, and the converged energy within the stopping tolerance (the
    DIIS system at depth 8 amplifies rounding in between)."""
    out = {}
    for par in ("0", "1"):
        monkeypatch.setenv("CCGPU_3NF_PARITY", par)
        s = ccnm.Solver(A=28, matter=ccnm.SNM, rho=0.16, nmax=2)
        s.set_chiral(ccnm.NNLO, 450.0)
        s.set_3nf(2, 0.017, 0.79, 450.0, 3)
        s.generate_interaction(ccnm.INT_CHIRAL)
        e2 = s.setup_t_amplitudes()
        r = s.ccd_iter(subspace=8, diis_step=8, tol=1e-9, maxit=100)
        assert r["converged"]
        out[par] = (e2, r["iterations"], s.solve_energies(), r["eccd"])
        s.close()
    (m0, i0, t0, e0), (m1, i1, t1, e1) = out["0"], out["1"]
    assert abs(m1 - m0) <= 1e-12 * abs(m0)
    assert i1 == i0
    assert np.max(np.abs(t1[:8] - t0[:8])) <= 1e-12 * np.max(np.abs(t0))
    assert abs(e1 - e0) <= 1e-8


@pytest.mark.parametrize("matter", ["pnm", "snm"])
def test_triples_staged_kernel_equals_table_kernel(ccnm, monkeypatch, matter):
    """(T) of the staged kernel (F-term amplitudes from a per-call table, H-term
    rows of T in LDS) against the table kernel that gathers every amplitude:
    the same 18 terms in the same order per element, so the same bits."""
    A = 66 if matter == "pnm" else 28
    s = ccnm.Solver(A=A, matter=ccnm.PNM if matter == "pnm" else ccnm.SNM, rho=0.16, nmax=2)
    s.set_chiral(ccnm.NNLO, 450.0)
    s.generate_interaction(ccnm.INT_CHIRAL)
    s.setup_t_amplitudes()
    assert s.ccd_iter(subspace=8, diis_step=8, tol=1e-8, maxit=100)["converged"]
    out = {}
    for variant in ("3", "1"):
        monkeypatch.setenv("CCGPU_T3_KERNEL", variant)
        out[variant] = s.pert_triples2()
    s.close()
    assert out["1"] == out["3"], out
    assert out["1"] != 0


def test_hbm_read_probe(ccnm):
    """The HBM streaming-read probe returns a rate between a quarter of and
    the published 8 TB/s peak, and rejects a buffer under 1 MiB."""
    gbs = ccnm.hbm_read_probe(1 << 28, 3)
    assert 2000.0 < gbs < 8000.0, gbs
    with pytest.raises(RuntimeError, match="hbm_read_probe"):
        ccnm.hbm_read_probe(1000, 1)
