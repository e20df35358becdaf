"""Time to converged E/N for PNM CCD, A=66, |n_i|<=4 (1458 states), densities
0.08/0.12/0.16/0.20 fm^-3, Delta-full chiral NNLO NN at 450 MeV (the reference's
chiral_NN_with_delta at its default LECs, generated on the GPU), on one MI355X.

Two arms on the same device-resident lattices:
  plain : MBPT2 start, FP64 contractions, one density after another (the plain
          GPU solver)
  fast  : learned warm start, then all four densities in one batched launch
          sequence replayed from a hipGraph, with one of three precision plans
          (FP64; FP32-stored V_pppp then FP64; FP32 MFMA for the first DIIS
          cycle then FP64), the fastest one that lands within 1e-5 MeV/N
Both arms stop at the reference setting |dE| <= 1e-8 MeV (the plain arm's
energies are the FP64 answer), with DIIS depth 8 and extrapolation every 8
iterations.  Every precision plan ends with FP64 iterations, so the stopping
test is always an FP64 one.  The fast arm's E/N must agree with the plain FP64
E/N within 1e-5 MeV/nucleon for every density; a plan that misses is not
chosen, and the run fails if none lands.  The JSON also reports both arms at
1e-5 MeV with the chosen plan.

The warm-start network is trained before the timed region on converged solves
of |n_i|<=3 lattices (the solver's own output): random densities, twists and
pi-N LECs c1, c3 within +-20% of their defaults, and the four bench densities
in that smaller basis; its inference is inside the timed region.  A "step" is one whole job
of the fast arm (four converged solves).  The timed region starts with the
interaction, Fock matrix and tables already in HBM.  Prints one JSON line.
Beside the headline the JSON carries (T) on the symmetric-matter deck (SNM
A=132 |n_i|<=3 with the normal-ordered 3NF) and CCDT-1 on PNM A=66 |n_i|<=2.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
DENSITIES = [0.08, 0.12, 0.16, 0.20]
A, NMAX = 66, 4
TOL, DEPTH, MAXIT = 1e-8, 8, 100
TOL_FAST = 1e-5      # MeV: the looser comparison (FP32 until here, FP64 confirmation)
FP32_SWITCH = 1.0    # at tol 1e-8: FP32 up to the first DIIS extrapolation
CHIRAL_ORDER, CUTOFF = 3, 450.0   # NNLO, 450 MeV: the first line of the reference's ccm_in
CI_DEFAULT = (-0.74, -0.49, -0.65, 0.96)   # c1..c4 (GeV^-1) of that order
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6
FP32_MFMA_PEAK_TFS = 157.3


def load_ccnm():
    spec = importlib.util.spec_from_file_location(
        "ccnm", os.path.join(ROOT, "coupled-cluster-nuclearmatter_amd", "ccnm.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["ccnm"] = mod
    spec.loader.exec_module(mod)
    return mod


def lattice(ccnm, rho, nmax, twist=(0.0, 0.0, 0.0), ci=None):
    """PNM A=66 lattice with the chiral NNLO(450) NN generated on the GPU."""
    s = ccnm.Solver(A=A, matter=ccnm.PNM, rho=rho, nmax=nmax, twist=tuple(twist))
    s.set_chiral(CHIRAL_ORDER, CUTOFF, ci)
    s.generate_interaction(ccnm.INT_CHIRAL)
    return s


def run_plain(ccnm, solvers, tol=TOL):
    out = []
    for s in solvers:
        s.setup_t_amplitudes()
        out.append(s.ccd_iter(subspace=DEPTH, diis_step=DEPTH, tol=tol, maxit=MAXIT,
                              precision=ccnm.PREC_FP64, use_graph=1))
    return out


# Precision plans of the fast arm (all with the learned start and the batch):
#   fp64      FP64 contractions throughout
#   v32       pp ladder on the complex64 copy of V_pppp (FP64 arithmetic) until
#             every |dE| <= tol, then FP64 iterations to the FP64 stopping test
#   fp32      FP32 MFMA contractions up to the first DIIS extrapolation (or, at
#             the loose tolerance, until every |dE| <= tol), FP64 after
FAST_PLANS = ("fp64", "v32", "fp32")


def run_fast(ccnm, solvers, net, tol=TOL, plan="fp64"):
    for s in solvers:
        net.warm_start(s)
    prec = {"fp64": ccnm.PREC_FP64, "v32": ccnm.PREC_V32, "fp32": ccnm.PREC_FP32}[plan]
    switch = FP32_SWITCH if (plan == "fp32" and tol < TOL_FAST) else 0.0
    return ccnm.batch_solve(solvers, subspace=DEPTH, diis_step=DEPTH, tol=tol, maxit=MAXIT,
                            precision=prec, fp32_switch=switch, use_graph=1)


def timed(fn, reps):
    t = time.perf_counter()
    for _ in range(reps):
        r = fn()
    return (time.perf_counter() - t) / reps, r


def train_warm_start(ccnm, nsolves=24, seed=2024):
    """The warm-start network trained on converged |n_i|<=3 solves: 24 at
    random densities in [0.06, 0.22] fm^-3, random twists and c1, c3 within
    +-20% of their defaults, and the four bench densities at twist zero (the
    same systems in the smaller basis, as an emulator is trained on cheap
    calculations)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    net = ccnm.WarmStart(seed=11)

    def add(s):
        s.setup_t_amplitudes()
        s.ccd_iter(subspace=DEPTH, diis_step=DEPTH, tol=TOL, maxit=MAXIT, use_graph=1)
        return net.add_solve(s)
    for _ in range(nsolves):
        ci = list(CI_DEFAULT)
        ci[0] *= rng.uniform(0.8, 1.2)
        ci[2] *= rng.uniform(0.8, 1.2)
        n = add(lattice(ccnm, rng.uniform(0.06, 0.22), 3, tuple(rng.uniform(0.0, 3.14159265358979, 3)), ci))
    for rho in DENSITIES:
        n = add(lattice(ccnm, rho, 3))
    losses = []
    for lr, steps in ((2e-3, 3000), (5e-4, 1500), (1e-4, 1000)):
        losses.append(net.train(steps, batch=16384, lr=lr, seed=steps))
    return net, n, losses


def twist_batch(ccnm, ntwist=32, seed=20241202):
    """Twist averaging: 32 twists theta ~ U[0, pi]^3 (seeded) of PNM A=66
    |n_i|<=3 at rho=0.16, every twist its own lattice and channel sizes.
    Sequential plain FP64 solves against one batched solve of all twists (FP64,
    and FP32 for the first DIIS cycle), all to |dE| <= 1e-8 MeV."""
    import numpy as np
    rng = np.random.default_rng(seed)
    twists = rng.uniform(0.0, 3.14159265358979, (ntwist, 3))
    t = time.perf_counter()
    sols = []
    for tw in twists:
        sols.append(lattice(ccnm, 0.16, 3, tw))
    setup = time.perf_counter() - t

    def seq():
        out = []
        for s in sols:
            s.setup_t_amplitudes()
            out.append(s.ccd_iter(subspace=DEPTH, diis_step=DEPTH, tol=TOL, maxit=MAXIT, use_graph=1))
        return out

    def bat(prec):
        for s in sols:
            s.setup_t_amplitudes()
        return ccnm.batch_solve(sols, subspace=DEPTH, diis_step=DEPTH, tol=TOL, maxit=MAXIT, precision=prec,
                                fp32_switch=FP32_SWITCH, use_graph=1)

    seq()
    seq_s, r_seq = timed(seq, 1)
    bat(ccnm.PREC_FP64)
    bat64_s, r64 = timed(lambda: bat(ccnm.PREC_FP64), 2)
    bat(ccnm.PREC_FP32)
    bat32_s, r32 = timed(lambda: bat(ccnm.PREC_FP32), 2)
    e_seq = np.mean([r["eccd"].real for r in r_seq]) / A
    e32 = np.mean([r["eccd"].real for r in r32]) / A
    res = {"workload": f"{ntwist} twists ~ U[0,pi]^3 (seed {seed}), PNM A=66 |n_i|<=3, rho 0.16",
           "setup_s": round(setup, 3), "sequential_fp64_s": round(seq_s, 4), "batched_fp64_s": round(bat64_s, 4),
           "batched_mixed_s": round(bat32_s, 4), "speedup_batched_fp64": round(seq_s / bat64_s, 3),
           "speedup_batched_mixed": round(seq_s / bat32_s, 3),
           "twist_averaged_e_per_n": round(float(e_seq), 8),
           "mixed_minus_fp64_e_per_n": float(e32 - e_seq),
           "iterations_sequential": [r["iterations"] for r in r_seq],
           "iterations_batched_mixed": [r["iterations"] for r in r32]}
    del sols
    return res


def triples_rate(ccnm):
    """The (T) workload of the symmetric-matter configuration: SNM A=132
    |n_i|<=3 (1372 states), rho=0.16, chiral NNLO(450) NN plus the NNLO 3NF
    normal-ordered to two-body (cE, cD, c_i at the deck's defaults), CCD to
    1e-8 MeV, then (T) in FP64 on its converged amplitudes: the first call
    builds the element tables, the second is the timed kernel."""
    t = time.perf_counter()
    s = ccnm.Solver(A=2 * A, matter=ccnm.SNM, rho=0.16, nmax=3)
    s.set_chiral(CHIRAL_ORDER, CUTOFF)
    s.set_3nf(2, 0.017, 0.79, 450.0, 3)
    s.generate_interaction(ccnm.INT_CHIRAL)
    setup = time.perf_counter() - t
    s.setup_t_amplitudes()
    r = s.ccd_iter(subspace=DEPTH, diis_step=DEPTH, tol=TOL, maxit=MAXIT, use_graph=1)
    t = time.perf_counter()
    s.pert_triples2()
    first = time.perf_counter() - t
    et = s.pert_triples2()
    n, sec = s.triples_stats()
    return {"workload": "SNM A=132 |n_i|<=3 (1372 states), rho 0.16, chiral NNLO(450) + NNLO 3NF normal-ordered "
                        "to two-body, CCD to 1e-8 MeV then (T)",
            "setup_s": round(setup, 3), "ccd_iterations": r["iterations"],
            "e_ccd_per_n": round(r["eccd"].real / (2 * A), 8), "e_t_per_n": round(et.real / (2 * A), 8),
            "elements": n, "seconds": round(sec, 4), "tables_s": round(first - sec, 3),
            "elements_per_s": round(n / sec, 1)}


def ccdt1_rate(ccnm):
    """CCDT-1 (the reference's CCDT1 loop) on PNM A=66 |n_i|<=2, rho=0.16,
    chiral NNLO(450) plus the NNLO 3NF normal-ordered to two-body: one T3
    energy sum (the (T) numerator and the three-body element per (abc, ijk)
    element) timed alone, then the whole solve to |dE| <= 1e-6 MeV with the
    reference's DIIS 6/6."""
    import numpy as np
    s = ccnm.Solver(A=A, matter=ccnm.PNM, rho=0.16, nmax=2)
    s.set_chiral(CHIRAL_ORDER, CUTOFF)
    s.set_3nf(2, 0.017, 0.79, 450.0, 3)
    s.generate_interaction(ccnm.INT_CHIRAL)
    s.setup_t_amplitudes()
    s.setup_t3()
    n, _ = s.triples_stats()
    t = time.perf_counter()
    s.t3_intermediate()
    sec = time.perf_counter() - t
    s.setup_t_amplitudes()
    t = time.perf_counter()
    r = s.ccd_iter(subspace=6, diis_step=6, tol=float(np.float32(1e-6)), maxit=MAXIT, ccdt1=1)
    solve = time.perf_counter() - t
    return {"workload": "PNM A=66 |n_i|<=2, rho 0.16, chiral NNLO(450) + NNLO 3NF normal-ordered to two-body, "
                        "CCDT-1 to 1e-6 MeV (DIIS 6/6)",
            "elements": n, "t3_sum_s": round(sec, 4), "elements_per_s": round(n / sec, 1),
            "solve_s": round(solve, 3), "iterations": r["iterations"], "e_per_n": round(r["eccd"].real / A, 8)}


def learned_study(ccnm, solvers, net):
    """Start error and iterations of the learned start against MBPT2, under the
    bench schedule (DIIS 8/8) and the reference's default (6/6)."""
    import numpy as np
    out = []
    for s in solvers:
        n = s.dims["t2_unique"]
        s.setup_t_amplitudes()
        t_m = s.debug_fetch("T", n)
        net.warm_start(s)
        t_l = s.debug_fetch("T", n)
        row = {}
        for sub, step in ((8, 8), (6, 6)):
            s.setup_t_amplitudes()
            rm = s.ccd_iter(subspace=sub, diis_step=step, tol=TOL, maxit=MAXIT, use_graph=1)
            t_c = s.debug_fetch("T", n)
            net.warm_start(s)
            rl = s.ccd_iter(subspace=sub, diis_step=step, tol=TOL, maxit=MAXIT, use_graph=1)
            row[f"iters_{sub}{step}"] = [rm["iterations"], rl["iterations"]]
            row[f"ms_{sub}{step}"] = [round(rm["seconds"] * 1e3, 2), round(rl["seconds"] * 1e3, 2)]
        nc = np.linalg.norm(t_c)
        row["start_err"] = [round(float(np.linalg.norm(t_m - t_c) / nc), 4),
                            round(float(np.linalg.norm(t_l - t_c) / nc), 4)]
        out.append(row)
    return out


def ladder_roofline(ccnm, s, precision):
    """Average duration of the pp-ladder GEMM launch from HIP events recorded on
    the stream it runs on, and its algorithmic bytes and flops per launch."""
    os.environ["CCGPU_TIMING"] = "1"
    try:
        s.setup_t_amplitudes()
        s.ccd_iter(subspace=DEPTH, diis_step=DEPTH, tol=TOL, precision=precision, fp32_switch=0.0,
                   maxit=MAXIT if precision == ccnm.PREC_FP64 else DEPTH)
        st = ccnm.kernel_stats()
    finally:
        os.environ.pop("CCGPU_TIMING", None)
    d = s.dims
    ms, n = st["gemm_pp_ladder"]   # FP32 runs stop at DEPTH iterations: all in FP32
    vbytes = 16.0 if precision == ccnm.PREC_FP64 else 8.0
    # V read once, T read, R read and written once (complex16 each)
    alg_bytes = vbytes * d["vpppp_unique"] + 3 * 16.0 * d["t2_unique"]
    flops = d["pp_ladder_flops"]
    return ms, n, alg_bytes, flops, st


def profiled_traffic():
    """HBM bytes per pp-ladder pass from the committed rocprofv3 FETCH_SIZE and
    WRITE_SIZE summary (profiles/pp_ladder_pnm66_nmax4_fp64.md: the line
    "HBM traffic per pass = ... = **X GB**"), or None."""
    import re
    path = os.path.join(ROOT, "profiles", "pp_ladder_pnm66_nmax4_fp64.md")
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(r"HBM traffic per pass[^*]*= \*\*([0-9.]+) GB\*\*", text, re.S)
    return float(m.group(1)) * 1e9 if m else None


def cpu_baseline(plain_iters):
    """The compiled reference (oracle/_ref/ref_driver, the reference's own
    t2_intermediate/diis_t1t2/ccd_energy_save) on this host's cores: a bounded
    sample of 2 CCD iterations at rho=0.16 on the same A=66, nmax=4 lattice,
    timed by the reference's MPI_WTIME around the loop, scaled to the plain
    arm's iteration counts.  The deck asks for LO: the reference evaluates
    every V element with its own loop-function quadrature, and at LO that setup
    takes seconds instead of minutes; a CCD iteration does the same work for
    any interaction.  Returns None where the binary or mpiexec is absent."""
    drv = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    mpiexec = "/opt/conda/bin/mpiexec"
    if not (os.path.exists(drv) and os.path.exists(mpiexec)):
        return None
    cores = min(16, os.cpu_count() or 1)
    deck = """!order, cutoff
0, 450
! cE cD
 0.0, 0.0
! A
66
! matter, input
pnm, density
! BC
PBC
! dens, ntwist, nmax, my_twists
0.16,1,4,1
! cc
CCD
! tnf
F, 0
! 3nf
450, 3
"""
    tmp = tempfile.mkdtemp(prefix="ccref_")
    try:
        with open(os.path.join(tmp, "ccm_in"), "w") as fh:
            fh.write(deck)
        env = dict(os.environ, OMP_NUM_THREADS="1", MKL_NUM_THREADS="1")
        cmd = [mpiexec, "-n", str(cores), drv, os.path.join(tmp, "ccm_in"), tmp, "chiral",
               str(DEPTH), str(DEPTH), "1e-8", "2", "0", "0", "0", "0", "0"]
        res = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=300)
        if res.returncode != 0:
            return {"error": "reference run failed: " + res.stderr[-300:]}
        loop = None
        for line in open(os.path.join(tmp, "energies.txt")):
            f = line.split()
            if f[0] == "looptime":
                loop = float(f[1])
        per_it = loop / 2.0
        total = per_it * sum(plain_iters)
        return {"value": total, "unit": "s", "cores": cores, "kind": "reference",
                "sample": f"2 CCD iterations of PNM A=66 nmax=4 rho=0.16 (reference ccd_iter "
                          f"routines, MPI x{cores}, MKL sequential; LO chiral elements shorten its V "
                          f"setup, the iteration's work does not depend on the interaction): "
                          f"{per_it:.3f} s/iteration, scaled to the plain arm's {sum(plain_iters)} "
                          f"iterations over 4 densities"}
    except Exception as e:  # noqa: BLE001
        return {"error": str(e)[:300]}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-triples", action="store_true",
                    help="short run for counter collection: the (T) workload (SNM A=132 |n_i|<=3, chiral NNLO + "
                         "normal-ordered 3NF), its tables and one (T) kernel; no JSON line")
    ap.add_argument("--profile-run", action="store_true",
                    help="short run for counter collection: 6 FP64 iterations of the rho=0.16 lattice, "
                         "then 6 FP32 iterations of the four-density batch; no JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus above 1 is not supported "
                                   "(the path shards by twist/channel, see the design notes)"}))
        sys.exit(2)

    ccnm = load_ccnm()
    if args.profile_triples:
        t3 = triples_rate(ccnm)
        print(f"# profile (T): {t3['elements']:.3e} elements in {t3['seconds']:.3f} s", file=sys.stderr, flush=True)
        return
    t0 = time.perf_counter()
    solvers = []
    for rho in DENSITIES:
        solvers.append(lattice(ccnm, rho, NMAX))
    setup_s = time.perf_counter() - t0
    print(f"# setup (lattice, V, Fock) for 4 densities: {setup_s:.2f} s", file=sys.stderr, flush=True)
    if args.profile_run:
        solvers[2].setup_t_amplitudes()
        r64 = solvers[2].ccd_iter(subspace=DEPTH, diis_step=DEPTH, tol=1e-30, maxit=6, use_graph=1)
        for s in solvers:
            s.setup_t_amplitudes()
        r32 = ccnm.batch_solve(solvers, subspace=DEPTH, diis_step=DEPTH, tol=1e-30, maxit=6,
                               precision=ccnm.PREC_FP32, fp32_switch=0.0, use_graph=1)
        print(f"# profile run: FP64 {r64['seconds']*1e3:.1f} ms / 6 iterations, "
              f"FP32 batch {r32[0]['seconds']*1e3:.1f} ms / 6 iterations", file=sys.stderr, flush=True)
        return

    # plain FP64 arm (reference for the error check and the speedup)
    plain = run_plain(ccnm, solvers)   # warm
    shapes = ccnm.autotune_report()   # the headline lattice's lists (later studies tune their own)
    t = time.perf_counter()
    plain = run_plain(ccnm, solvers)
    plain_s = time.perf_counter() - t
    print(f"# plain FP64: {plain_s:.3f} s, iterations {[r['iterations'] for r in plain]}",
          file=sys.stderr, flush=True)

    t = time.perf_counter()
    net, nsamples, losses = train_warm_start(ccnm)
    train_s = time.perf_counter() - t
    print(f"# warm start trained on {nsamples} amplitudes in {train_s:.2f} s, losses {losses}",
          file=sys.stderr, flush=True)

    # choose the precision plan: each runs once warm and once timed, and must
    # land within 1e-5 MeV/N of the plain FP64 energies
    plans = {}
    for plan in FAST_PLANS:
        run_fast(ccnm, solvers, net, TOL, plan)
        sec, r = timed(lambda: run_fast(ccnm, solvers, net, TOL, plan), 1)
        err = max(abs(f["eccd"].real - p["eccd"].real) / A for f, p in zip(r, plain))
        ok = all(f["converged"] for f in r) and err <= 1e-5
        plans[plan] = {"s": round(sec, 4), "iterations": [x["iterations"] for x in r],
                       "low_precision_iterations": [x["fp32_iterations"] for x in r],
                       "max_err_per_n": err, "ok": ok}
        print(f"# fast plan {plan}: {sec * 1e3:.1f} ms, iterations {plans[plan]['iterations']}, err {err:.2e}",
              file=sys.stderr, flush=True)
    best = min((p for p in FAST_PLANS if plans[p]["ok"]), key=lambda p: plans[p]["s"], default=None)
    if best is None:
        print(json.dumps({"error": "no fast plan landed within 1e-5 MeV/N of FP64", "plans": plans}))
        sys.exit(1)
    for _ in range(args.warmup):
        run_fast(ccnm, solvers, net, TOL, best)
    times = []
    fast = None
    for _ in range(args.steps):
        t = time.perf_counter()
        fast = run_fast(ccnm, solvers, net, TOL, best)
        times.append(time.perf_counter() - t)
    ms_per_step = 1e3 * sorted(times)[len(times) // 2]   # median step
    errs = [abs(f["eccd"].real - p["eccd"].real) / A for f, p in zip(fast, plain)]
    # the same comparison at the looser tolerance
    plain_t5_s, plain_t5 = timed(lambda: run_plain(ccnm, solvers, TOL_FAST), 2)
    fast_t5_s, fast_t5 = timed(lambda: run_fast(ccnm, solvers, net, TOL_FAST, best), 2)
    errs5 = [abs(f["eccd"].real - p["eccd"].real) / A for f, p in zip(fast_t5, plain)]
    print(f"# fast: {ms_per_step:.1f} ms/job, iterations {[r['iterations'] for r in fast]}, "
          f"max |dE/N| vs FP64 {max(errs):.2e}", file=sys.stderr, flush=True)
    if not all(f["converged"] for f in fast) or max(errs) > 1e-5:
        print(json.dumps({"error": "fast arm missed 1e-5 MeV/N of FP64", "errs": errs}))
        sys.exit(1)

    # roofline of the dominant kernel: the pp ladder (FP64 in the plain arm)
    ms64, n64, bytes64, flops64, st64 = ladder_roofline(ccnm, solvers[2], ccnm.PREC_FP64)
    ms32, n32, bytes32, flops32, _ = ladder_roofline(ccnm, solvers[2], ccnm.PREC_FP32)
    gbs64 = bytes64 / (ms64 * 1e-3) / 1e9
    tfs64 = flops64 / (ms64 * 1e-3) / 1e12
    ai64 = flops64 / bytes64
    bound = "hbm" if ai64 < FP64_PEAK_TFS * 1e12 / (HBM_PEAK_GBS * 1e9) else "mfma"
    traffic = profiled_traffic()
    if bound == "hbm":
        roof = {"bound": "hbm", "achieved": round(gbs64, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(gbs64 / HBM_PEAK_GBS, 4), "traffic": traffic}
    else:
        roof = {"bound": "mfma", "achieved": round(tfs64, 2), "peak": FP64_PEAK_TFS, "unit": "TFLOP/s",
                "frac": round(tfs64 / FP64_PEAK_TFS, 4), "traffic": traffic}
    roof.update({"kernel": "k_zgemm_f64 (pp ladder, FP64)", "avg_ms": round(ms64, 4), "launches": n64,
                 "alg_bytes_per_launch": bytes64, "alg_flops_per_launch": flops64,
                 "tflops": round(tfs64, 2), "arith_intensity": round(ai64, 2)})
    roof32 = {"kernel": "k_zgemm_f32 (pp ladder, FP32 MFMA)", "avg_ms": round(ms32, 4),
              "gbs": round(bytes32 / (ms32 * 1e-3) / 1e9, 1),
              "tflops": round(flops32 / (ms32 * 1e-3) / 1e12, 2)}

    study = learned_study(ccnm, solvers, net)
    twists = twist_batch(ccnm)
    t3 = triples_rate(ccnm)
    t3["ccdt1"] = ccdt1_rate(ccnm)
    print(f"# (T): {t3['elements']:.3e} elements in {t3['seconds']:.3f} s", file=sys.stderr, flush=True)
    c1 = t3["ccdt1"]
    print(f"# CCDT-1: {c1['elements']:.3e} elements per T3 sum in {c1['t3_sum_s']:.3f} s; solve {c1['solve_s']:.2f} s, "
          f"{c1['iterations']} iterations", file=sys.stderr, flush=True)
    print(f"# twists: sequential {twists['sequential_fp64_s']:.3f} s, batched FP64 {twists['batched_fp64_s']:.3f} s, "
          f"batched mixed {twists['batched_mixed_s']:.3f} s", file=sys.stderr, flush=True)
    cpu = None if args.no_cpu_baseline else cpu_baseline([r["iterations"] for r in plain])

    out = {
        "metric": "time to E/N converged at 1e-5 MeV vs FP64 oracle: mixed+learned vs plain GPU",
        "value": round(ms_per_step / 1e3, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic lattice inputs built on the GPU: plane-wave basis, channels and the "
                "Delta-full chiral NNLO(450) NN of the reference (chiral_NN_with_delta, default LECs)",
        "config": {"workload": "PNM CCD A=66 |n_i|<=4 (1458 sp states), rho 0.08/0.12/0.16/0.20, "
                               "chiral NNLO(450) NN, DIIS 8, tol 1e-8 MeV, MBPT2 start",
                   "arms": "fast = learned warm start + batched 4 densities + hipGraph, precision plan "
                           "chosen among fp64 / v32 (FP32-stored V_pppp until converged, FP64 after) / "
                           "fp32 (FP32 MFMA for the first DIIS cycle, FP64 after); plain = MBPT2 start, "
                           "sequential FP64; both to |dE| <= 1e-8 MeV",
                   "fast_plan": best,
                   "t2_unique": solvers[0].dims["t2_unique"], "t2_ordered": solvers[0].dims["t2_ordered"],
                   "vpppp_unique": solvers[0].dims["vpppp_unique"]},
        "plain_fp64_s": round(plain_s, 4),
        "speedup_vs_plain": round(plain_s / (ms_per_step / 1e3), 3),
        "iterations_plain": [r["iterations"] for r in plain],
        "iterations_fast": [r["iterations"] for r in fast],
        "fp32_iterations_fast": [r["fp32_iterations"] for r in fast],
        "fast_plans": plans,
        "at_tol_1e-5": {"plain_s": round(plain_t5_s, 4), "fast_s": round(fast_t5_s, 4),
                        "speedup": round(plain_t5_s / fast_t5_s, 3),
                        "iterations_plain": [r["iterations"] for r in plain_t5],
                        "iterations_fast": [r["iterations"] for r in fast_t5],
                        "fp32_iterations_fast": [r["fp32_iterations"] for r in fast_t5],
                        "max_err_per_n_vs_plain_1e-8": max(errs5)},
        "e_per_n_fp64": [round(r["eccd"].real / A, 8) for r in plain],
        "max_err_per_n": max(errs),
        "setup_s": round(setup_s, 3),
        "twist_batch": twists,
        "triples": t3,
        "warm_start": {"train_s": round(train_s, 3), "train_amplitudes": nsamples,
                       "train_solves": "24 x PNM A=66 |n_i|<=3 chiral NNLO(450), rho ~ U[0.06,0.22], "
                                       "twist ~ U[0,pi]^3, c1 and c3 ~ default x U[0.8,1.2]; and the 4 bench "
                                       "densities at |n_i|<=3, twist 0",
                       "final_loss": losses[-1],
                       "per_density": study},
        "roofline": roof,
        "roofline_fp32_ladder": roof32,
        "kernel_ms_fp64": {k: round(v[0], 4) for k, v in st64.items()},
        "tile_shapes": shapes,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
