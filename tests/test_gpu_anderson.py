"""The Anderson update (ccgpu_solve_opts.accel = CCGPU_ACCEL_ANDERSON) against the
oracle's restatement (ccd_oracle.CCD.solve_anderson), and the failure handling
of the solve loop with the failures forced from the host: a singular
DIIS/Anderson system in one solve of a batch, non-finite amplitudes in one
solve of a batch, the iteration cap.  The reference has none of this: its
diis_t1t2 uses ZGESV's solution whatever its info (ccm_diis.f90:173-176), and
ccd_iter stops silently at the cap (ccm_iter.f90:71-74)."""
import numpy as np
import pytest

import ccd_oracle as co
import chiral_oracle as cho
from conftest import fixture_interaction, fixture_lattice, golden

pytestmark = pytest.mark.gpu


def native(ccnm, A, matter, rho, nmax, twist=(0.0, 0.0, 0.0), chiral=True):
    s = ccnm.Solver(A=A, matter=ccnm.PNM if matter == "pnm" else ccnm.SNM, rho=rho, nmax=nmax, twist=twist)
    if chiral:
        s.set_chiral(ccnm.NNLO, 450.0)
        s.generate_interaction(ccnm.INT_CHIRAL)
    else:
        s.generate_interaction(ccnm.INT_MINNESOTA)
    return s


@pytest.mark.parametrize("case", [("pnm", 14, 1, True, (0.0, 0.0, 0.0), 8, 1.0),
                                  ("snm", 4, 1, True, (0.0, 0.0, 0.0), 4, 0.7),
                                  ("pnm", 14, 1, False, (0.3, 1.1, 2.0), 6, 1.0)])
def test_anderson_trajectory_matches_oracle(ccnm, case):
    """Every iteration's energy of the GPU Anderson update equals the oracle's
    (1e-10 of the largest), with the same iteration count: the residual
    vectors, their FP32 Gram matrix, the scaled coefficient solve and the
    combination of the mixed iterates."""
    matter, A, nmax, chiral, twist, depth, beta = case
    s = native(ccnm, A, matter, 0.16, nmax, twist, chiral)
    lat = co.Lattice.build(A, matter, 0.16, nmax, twist)
    vf = cho.ChiralInteraction(lat, cho.NNLO, 450.0) if chiral else co.MinnesotaInteraction(lat)
    cc = co.CCD(lat, vf)
    ref = np.array(cc.solve_anderson(depth=depth, beta=beta, tol=1e-10, maxit=60))
    s.setup_t_amplitudes()
    r = s.ccd_iter(subspace=depth, diis_step=depth, tol=1e-10, maxit=60, accel=ccnm.ACCEL_ANDERSON, mix=beta)
    en = s.solve_energies()
    assert r["converged"] and r["status"] == 0
    assert len(en) == len(ref), (len(en), len(ref))
    assert np.max(np.abs(en - ref)) <= 1e-10 * np.max(np.abs(ref)), np.max(np.abs(en - ref))
    assert abs(r["eccd"] - cc.ecc) < 1e-9


def test_anderson_and_reference_update_share_the_fixed_point(ccnm):
    """Both updates stop on the same FP64 residual fixed point."""
    s = native(ccnm, 66, "pnm", 0.16, 2)
    s.setup_t_amplitudes()
    r0 = s.ccd_iter(subspace=8, diis_step=8, tol=1e-9, maxit=100)
    s.setup_t_amplitudes()
    r1 = s.ccd_iter(subspace=8, diis_step=8, tol=1e-9, maxit=100, accel=ccnm.ACCEL_ANDERSON, mix=1.0)
    assert r0["converged"] and r1["converged"]
    assert abs(r0["eccd"] - r1["eccd"]) < 1e-7
    assert r1["iterations"] < r0["iterations"]


@pytest.mark.parametrize("accel", [0, 1])
def test_singular_system_in_one_solve_keeps_the_batch_going(ccnm, monkeypatch, accel):
    """A singular DIIS (reference update) or Anderson system in solve 1 of a
    batch of two: that solve takes its plain update for the step and reports
    CCGPU_STATUS_DIIS_FALLBACK; solve 0 is unaffected, both converge to the
    energies of undisturbed solves."""
    sols = [native(ccnm, 14, "pnm", rho, 1, chiral=False) for rho in (0.12, 0.16)]
    want = []
    for s in sols:
        s.setup_t_amplitudes()
        want.append(s.ccd_iter(subspace=6, diis_step=6, tol=1e-10, maxit=200, accel=accel, mix=1.0))
    monkeypatch.setenv("CCGPU_FORCE_SINGULAR", "1")
    for s in sols:
        s.setup_t_amplitudes()
    got = ccnm.batch_solve(sols, subspace=6, diis_step=6, tol=1e-10, maxit=200, accel=accel, mix=1.0)
    assert got[0]["status"] == 0
    assert got[1]["status"] & ccnm.STATUS_DIIS_FALLBACK
    assert got[0]["iterations"] == want[0]["iterations"]
    assert got[0]["eccd"] == want[0]["eccd"]
    assert got[1]["converged"]
    assert abs(got[1]["eccd"] - want[1]["eccd"]) < 1e-8


@pytest.mark.parametrize("accel", [0, 1])
def test_nonfinite_solve_is_stopped_alone(ccnm, accel):
    """NaN amplitudes handed to solve 0 of a batch: that solve stops with
    CCGPU_STATUS_NONFINITE (not converged); solve 1 converges as alone."""
    sols = [native(ccnm, 14, "pnm", rho, 1, chiral=False) for rho in (0.12, 0.16)]
    sols[1].setup_t_amplitudes()
    want = sols[1].ccd_iter(subspace=6, diis_step=6, tol=1e-10, maxit=200, accel=accel, mix=1.0)
    for s in sols:
        s.setup_t_amplitudes()
    t = sols[0].t2_unique()
    t[len(t) // 2] = np.nan
    sols[0].set_t2_unique(t)
    got = ccnm.batch_solve(sols, subspace=6, diis_step=6, tol=1e-10, maxit=200, accel=accel, mix=1.0)
    assert got[0]["status"] & ccnm.STATUS_NONFINITE
    assert not got[0]["converged"]
    assert got[1]["status"] == 0 and got[1]["converged"]
    assert got[1]["iterations"] == want["iterations"]
    assert abs(got[1]["eccd"] - want["eccd"]) < 1e-12


@pytest.mark.parametrize("accel", [0, 1])
def test_iteration_cap_is_reported(ccnm, accel):
    s = native(ccnm, 14, "pnm", 0.16, 1, chiral=False)
    s.setup_t_amplitudes()
    r = s.ccd_iter(subspace=6, diis_step=6, tol=1e-12, maxit=3, accel=accel, mix=1.0)
    assert r["iterations"] == 3
    assert not r["converged"]
    assert r["status"] == ccnm.STATUS_MAXIT
    assert len(s.solve_energies()) == 4


def test_refine_every_trajectory(ccnm):
    """FP32 contractions with an FP64 residual every 3rd iteration
    (refine_every), then FP64 to the stopping test: the same fixed point as
    FP64 throughout, with FP64 iterations at the stated places."""
    s = native(ccnm, 66, "pnm", 0.16, 2)
    s.setup_t_amplitudes()
    r64 = s.ccd_iter(subspace=8, diis_step=8, tol=1e-9, maxit=100)
    s.setup_t_amplitudes()
    r = s.ccd_iter(subspace=8, diis_step=8, tol=1e-9, maxit=100, precision=ccnm.PREC_FP32, refine_every=3,
                   fp32_switch=1e-4)
    assert r["converged"]
    # iterations 3, 6, ... of the low-precision phase ran in FP64
    assert r["fp32_iterations"] >= 4
    assert r["iterations"] - r["fp32_iterations"] >= r["fp32_iterations"] // 3
    assert abs(r["eccd"] - r64["eccd"]) / 66 < 1e-9


def test_confirm_counts_consecutive_small_changes(ccnm):
    """ccgpu_solve_opts.confirm: a solve ends at the first iteration that
    completes `confirm` consecutive |dE| <= tol; the iterates do not depend on
    it, so the shorter trajectories are prefixes of the longest. 0 is the
    default (two for the Anderson update, one for the reference's)."""
    s = native(ccnm, 66, "pnm", 0.16, 2)
    tol = 1e-4
    traj = {}
    for c in (0, 1, 2, 3):
        s.setup_t_amplitudes()
        r = s.ccd_iter(subspace=8, diis_step=8, tol=tol, maxit=100, accel=ccnm.ACCEL_ANDERSON, mix=1.0,
                       confirm=c)
        assert r["converged"]
        traj[c] = s.solve_energies()
    longest = traj[3]
    for c in (1, 2, 3):
        en = traj[c]
        assert np.array_equal(en, longest[:len(en)])
        small = np.abs(np.diff(en)) <= tol
        run = 0
        for it, sm in enumerate(small, start=1):
            run = run + 1 if sm else 0
            if run == c:
                break
        assert len(en) - 1 == it, (c, len(en) - 1, it)
    assert np.array_equal(traj[0], traj[2])
    with pytest.raises(RuntimeError, match="confirm"):
        s.ccd_iter(subspace=8, diis_step=8, tol=tol, maxit=100, accel=ccnm.ACCEL_ANDERSON, confirm=5)
    s.close()
