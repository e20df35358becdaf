"""The shared pp ladder of a batch of twists (one V_pppp over the union of the
twists' particle pairs, one wide GEMM per (K, Tz) channel whose columns are the
hole pairs of every twist) against one ladder per twist.

The NN element <ab||cd> depends on the lattice labels of the four states only
(the twist cancels in every momentum difference), which is what lets the
twists of a twist-averaged run (ccm_main.f90:163-279, one process per twist in
the reference) share V_pppp.  The pp ladder is ccm_t2_eqn.f90:574-607."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TWISTS = [(0.3, 1.1, 2.0), (2.9, 0.4, 1.4), (1.7, 2.2, 0.1), (0.0, 0.0, 0.0)]


def solvers(ccnm, nmax, interaction, rho=0.16, twists=TWISTS):
    out = []
    for tw in twists:
        s = ccnm.Solver(A=66, matter=ccnm.PNM, rho=rho, nmax=nmax, twist=tw)
        if interaction == ccnm.INT_CHIRAL:
            s.set_chiral(ccnm.NNLO, 450.0)
        s.generate_interaction(interaction)
        s.setup_t_amplitudes()
        out.append(s)
    return out


def run(ccnm, monkeypatch, share, interaction, nmax=2, **kw):
    monkeypatch.setenv("CCGPU_SHARE_V", "1" if share else "0")
    monkeypatch.setenv("CCGPU_TIMING", "1")   # per-step timers: name the launched steps
    ss = solvers(ccnm, nmax, interaction)
    try:
        res = ccnm.batch_solve(ss, **kw)
        stats = ccnm.kernel_stats()
        t2 = [s.t2_unique() for s in ss]
        return res, t2, stats
    finally:
        for s in ss:
            s.close()
        monkeypatch.delenv("CCGPU_TIMING")
        monkeypatch.delenv("CCGPU_SHARE_V")


@pytest.mark.parametrize("interaction", ["chiral", "minnesota"])
def test_shared_ladder_equals_per_twist_ladder_fp64(ccnm, monkeypatch, interaction):
    """Three FP64 iterations of four twists (PNM A=66 |n_i|<=2): every twist's
    amplitudes with the shared ladder equal those with its own ladder to 1e-12
    of the largest (the two sum the same products in another order)."""
    kind = ccnm.INT_CHIRAL if interaction == "chiral" else ccnm.INT_MINNESOTA
    kw = dict(subspace=8, diis_step=8, tol=0.0, maxit=3)
    r0, t0, st0 = run(ccnm, monkeypatch, False, kind, **kw)
    r1, t1, st1 = run(ccnm, monkeypatch, True, kind, **kw)
    assert "pp_share_gather" in st1 and "pp_share_gather" not in st0
    for a, b in zip(t0, t1):
        assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(a)), np.max(np.abs(a - b)) / np.max(np.abs(a))
    for a, b in zip(r0, r1):
        assert abs(a["eccd"] - b["eccd"]) <= 1e-12 * abs(a["eccd"])


@pytest.mark.parametrize("prec", ["bf16x3", "fp32"])
def test_shared_ladder_low_precision_solve(ccnm, monkeypatch, prec):
    """The fast plan on four twists at |n_i|<=3 (chiral NNLO): low-precision
    contractions with the shared ladder on BF16x3 products until |dE| <= 1e-2
    MeV, then FP64 to |dE| <= 1e-6 MeV twice (Anderson): every twist's E/N
    within 1e-5 MeV of its FP64 solve with one ladder per twist."""
    p = ccnm.PREC_BF16X3 if prec == "bf16x3" else ccnm.PREC_FP32
    ref, _, _ = run(ccnm, monkeypatch, False, ccnm.INT_CHIRAL, nmax=3, subspace=8, diis_step=8, tol=1e-8, maxit=100)
    got, _, st = run(ccnm, monkeypatch, True, ccnm.INT_CHIRAL, nmax=3, subspace=8, diis_step=8, tol=1e-6,
                     maxit=100, precision=p, fp32_switch=1e-2, accel=ccnm.ACCEL_ANDERSON, mix=1.0)
    assert "pp_share_gather" in st
    for a, b in zip(ref, got):
        assert a["converged"] and b["converged"] and b["fp32_iterations"] >= 1
        assert abs(a["eccd"].real - b["eccd"].real) / 66 <= 1e-5, (a["eccd"], b["eccd"])


def test_no_sharing_across_densities(ccnm, monkeypatch):
    """Lattices of different densities have different V: no shared ladder."""
    monkeypatch.setenv("CCGPU_TIMING", "1")
    ss = []
    for rho in (0.12, 0.16):
        s = ccnm.Solver(A=66, matter=ccnm.PNM, rho=rho, nmax=2)
        s.generate_interaction(ccnm.INT_MINNESOTA)
        s.setup_t_amplitudes()
        ss.append(s)
    try:
        ccnm.batch_solve(ss, subspace=8, diis_step=8, tol=0.0, maxit=2)
        assert "pp_share_gather" not in ccnm.kernel_stats()
    finally:
        for s in ss:
            s.close()
        monkeypatch.delenv("CCGPU_TIMING")
