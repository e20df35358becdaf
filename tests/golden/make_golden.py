"""Write the golden fixtures of tests/golden/ from the compiled reference.

Runs only where the reference sources were compiled (oracle/Makefile builds
oracle/_ref/ref_driver from the reference in place).  Each case runs the
reference's own setup, CCD loop and (T) through oracle/ref_driver.f90 and stores
the results as a compressed .npz: per-iteration correlation energies, E0, the
final CCD / CCD(T) energies, and for the small cases the orbitals, Fock matrix,
canonical V elements and converged T2 keyed by orbital index of the stored
orbital table.

Usage: python tests/golden/make_golden.py [case ...]
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
PROGRAM = os.path.join(ROOT, "oracle", "_ref", "ccm.exe")

# name: (A, matter, rho, nmax, order, interaction, subspace, step, tol, maxit,
#        twist, dump, triples, ranks[, cutoff MeV, default 450[, 3NF
#        approximation 0/1/2, default off]])
# twist is a vector, or ("tabc", ntwist, k): the k-th combination of the
# reference's own Gauss-Legendre twists, also run through the reference program
# ccm.exe with my_twists = k for its printed twist-weighted E0.
CASES = {
    "pnm14_n1_nnlo": (14, "pnm", 0.16, 1, 3, "chiral", 6, 6, 1e-6, 100, (0, 0, 0), 1, 1, 1),
    "pnm14_n1_minn": (14, "pnm", 0.16, 1, 3, "minnesota", 8, 8, 1e-8, 100, (0, 0, 0), 1, 1, 1),
    "pnm14_n1_minn_twist": (14, "pnm", 0.16, 1, 3, "minnesota", 6, 6, 1e-6, 100,
                            (0.3, 1.1, 2.0), 1, 1, 1),
    "snm4_n1_nnlo": (4, "snm", 0.16, 1, 3, "chiral", 6, 6, 1e-6, 100, (0, 0, 0), 1, 1, 1),
    "snm4_n1_minn": (4, "snm", 0.16, 1, 3, "minnesota", 6, 6, 1e-6, 100, (0, 0, 0), 1, 1, 1),
    "pnm66_n2_minn": (66, "pnm", 0.16, 2, 3, "minnesota", 8, 8, 1e-8, 100, (0, 0, 0), 0, 0, 1),
    "pnm66_n2_minn_ref": (66, "pnm", 0.16, 2, 3, "minnesota", 6, 6, 1e-6, 100, (0, 0, 0), 0, 0, 1),
    "pnm66_n2_minn_twist": (66, "pnm", 0.16, 2, 3, "minnesota", 8, 8, 1e-8, 100,
                            (0.7, 1.9, 2.6), 0, 0, 1),
    "pnm66_n2_nnlo": (66, "pnm", 0.16, 2, 3, "chiral", 8, 8, 1e-8, 100, (0, 0, 0), 0, 0, 8),
    "snm28_n2_minn": (28, "snm", 0.16, 2, 3, "minnesota", 8, 8, 1e-11, 300, (0, 0, 0), 0, 0, 1),
    # chiral orders and cutoffs of the Delta-full family (V records only)
    "snm4_n1_lo450": (4, "snm", 0.16, 1, 0, "chiral", 6, 6, 1e-6, 20, (0, 0, 0), 1, 0, 1),
    "snm4_n1_nlo450": (4, "snm", 0.16, 1, 2, "chiral", 6, 6, 1e-6, 20, (0, 0, 0), 1, 0, 1),
    "snm4_n1_nnlo394": (4, "snm", 0.16, 1, 3, "chiral", 6, 6, 1e-6, 20, (0, 0, 0), 1, 0, 1, 394),
    "snm4_n1_nnlo500": (4, "snm", 0.16, 1, 3, "chiral", 6, 6, 1e-6, 20, (0, 0, 0), 1, 0, 1, 500),
    "pnm14_n1_nnlo_twist": (14, "pnm", 0.12, 1, 3, "chiral", 6, 6, 1e-6, 100, (0.4, 1.3, 2.2), 1, 1, 1),
    # the NNLO three-nucleon force normal-ordered to two-body (and only to
    # zero- and one-body), NNLO and NLO (Delta terms only)
    "pnm14_n1_nnlo_3nf": (14, "pnm", 0.16, 1, 3, "chiral", 6, 6, 1e-6, 100, (0, 0, 0), 1, 1, 1, 450, 2),
    "snm4_n1_nnlo_3nf": (4, "snm", 0.16, 1, 3, "chiral", 6, 6, 1e-6, 100, (0, 0, 0), 1, 1, 1, 450, 2),
    "snm4_n1_nlo450_3nf": (4, "snm", 0.16, 1, 2, "chiral", 6, 6, 1e-6, 20, (0, 0, 0), 1, 0, 1, 450, 2),
    "pnm14_n1_nnlo_3nf1": (14, "pnm", 0.16, 1, 3, "chiral", 6, 6, 1e-6, 100, (0.3, 2.1, 1.2), 1, 0, 1, 450, 1),
    # tnf_approx 3: two-body normal ordering, and (T) with the residual 3NF
    # element in the triples amplitude (pert_triples2's et_5)
    "snm4_n1_nnlo_3nf3": (4, "snm", 0.16, 1, 3, "chiral", 6, 6, 1e-6, 100, (0, 0, 0), 1, 1, 1, 450, 3),
    "pnm14_n1_nnlo_3nf3": (14, "pnm", 0.16, 1, 3, "chiral", 6, 6, 1e-6, 100, (0.3, 1.1, 2.0), 0, 1, 1, 450, 3),
}



# every twist combination of the ntwist=2 TABC deck: PNM A=14, nmax=1, NNLO
# (the second also with its T2 and blocks dumped)
for _k in range(1, 5):
    CASES[f"pnm14_n1_nnlo_tabc2_{_k}"] = (14, "pnm", 0.16, 1, 3, "chiral", 6, 6, 1e-6, 100, ("tabc", 2, _k),
                                          int(_k == 2), 1, 1)
# one combination of the ntwist=3 deck (E0 and twist only)
CASES["pnm14_n1_nnlo_tabc3_5"] = (14, "pnm", 0.16, 1, 3, "chiral", 6, 6, 1e-6, 100, ("tabc", 3, 5), 0, 0, 1)
# PNM A=66 |n_i|<=4 (1458 states), chiral NNLO(450) NN, DIIS 8/8, tol 1e-8 MeV,
# at the four densities of the large-basis study: E0, the MBPT2 start and every
# iteration's correlation energy (no dumps: V_pppp alone is 7.7 GB ordered)
for _rho in (0.08, 0.12, 0.16, 0.20):
    CASES[f"pnm66_n4_nnlo_r{int(round(_rho * 100)):03d}"] = (66, "pnm", _rho, 4, 3, "chiral", 8, 8, 1e-8, 100,
                                                            (0, 0, 0), 0, 0, 8)

# Runs of the reference program itself (oracle/_ref/ccm.exe) on a deck, parsed
# from its standard output: E0, every 'CCD energy' line (E0 + the T2
# correlation energy: the MBPT2 start, then one per iteration), every
# 'CCDT-1 energy' line (the T3 energy: the start, then one per iteration) and
# the final 8-number result line.
# name: (A, matter, rho, nmax, order, cutoff, tnf approximation, boundary
#        conditions, ntwist, my_twists, cluster approximation)
PROGRAM_CASES = {
    "snm4_n1_nnlo_3nf_ccdt1": (4, "snm", 0.16, 1, 3, 450, 2, "PBC", 1, 1, "CCDT1"),
    "pnm14_n1_nnlo_3nf_ccdt1": (14, "pnm", 0.16, 1, 3, 450, 2, "PBC", 1, 1, "CCDT1"),
}
# the reference's own sample deck (deltafull-CCMkspace/ccm_in: SNM A=28,
# rho=0.165, |n_i|<=2, chiral NNLO(450) plus the 3NF normal-ordered to
# two-body, CCD), run unmodified from where it lies
PROGRAM_CASES["ccm_in_sample"] = (28, "snm", 0.165, 2, 3, 450, 2, "PBC", 1, 1, "CCD")
SAMPLE_DECK = "/root/reference/deltafull-CCMkspace/ccm_in"
# every combination of the ntwist=2 TABC deck with CCDT-1
for _k in range(1, 5):
    PROGRAM_CASES[f"pnm14_n1_nnlo_3nf_ccdt1_tabc2_{_k}"] = (14, "pnm", 0.16, 1, 3, 450, 2, "TABC", 2, _k, "CCDT1")


def run_program_case(name):
    A, matter, rho, nmax, order, cutoff, tnf_approx, bc, ntw, myt, cc = PROGRAM_CASES[name]
    with tempfile.TemporaryDirectory() as tmp:
        if name == "ccm_in_sample":
            deck = SAMPLE_DECK
        else:
            deck = os.path.join(tmp, "ccm_in")
            with open(deck, "w") as fh:
                fh.write(CCM_IN.format(order=order, cutoff=cutoff, A=A, matter=matter, rho=rho, nmax=nmax,
                                      tnf="T" if tnf_approx >= 0 else "F", tnf_approx=max(tnf_approx, 0),
                                      bc=bc, ntwist=ntw, my_twists=myt, cc=cc))
        env = dict(os.environ, OMP_NUM_THREADS="8", MKL_NUM_THREADS="1")
        res = subprocess.run([PROGRAM, deck], env=env, cwd=tmp, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"{name}: reference program failed\n{res.stdout[-2000:]}")

    def cplx(line):
        re_, im = line.split("(")[1].split(")")[0].split(",")
        return complex(float(re_), float(im))
    lines = res.stdout.splitlines()
    out = {"params": np.array([A, nmax, 6, 6, 100, order], np.int64), "cutoff": np.float64(cutoff),
           "tnf_approx": np.int64(tnf_approx), "matter": np.array(matter), "interaction": np.array("chiral"),
           "rho": np.float64(rho), "tol": np.float64(np.float32(1e-6)), "cc": np.array(cc),
           "boundary": np.array(bc), "ntwist": np.int64(ntw), "my_twists": np.int64(myt)}
    out["e0"] = np.complex128(cplx([x for x in lines if "Vacuum expectation value" in x][0]))
    out["ccd_energies"] = np.array([cplx(x) for x in lines if x.strip().startswith("CCD energy")], np.complex128)
    out["e3"] = np.array([cplx(x) for x in lines if "CCDT-1 energy" in x], np.complex128)
    tail = [x for x in lines if x.startswith("     0.01700")]
    out["result_line"] = np.array([float(v) for v in tail[-1].split()], np.float64)
    for key, label in (("e0_per_n", "Twist averaged E0/N"), ("mbpt2_per_n", "Twist averaged mbpt2"),
                       ("eccd_per_n", "Twist averaged CCSD")):
        hit = [x for x in lines if label in x]
        if hit:
            out[key] = np.complex128(cplx(hit[-1]))
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    e3 = out["e3"][-1] if len(out["e3"]) else 0.0
    print(name, "iterations", len(out["ccd_energies"]) - 1, "final", out["ccd_energies"][-1] + e3)


CCM_IN = """!Chiral order for Deltas(LO = 0,NLO=2,NNLO=3,N3LO=4) and cutoff
{order}, {cutoff}
! cE and cD 3nf parameters:
 0.017, 0.79
! number of particles
{A}
! specify: pnm/snm, input type: density/kfermi
{matter}, density
! specify boundary conditions (PBC/TABC/TABCsp)
{bc}
! dens/kf, ntwist,  nmax, my_twists
{rho},{ntwist},{nmax},{my_twists}
! specify cluster approximation: CCD, CCD(T)
{cc}
! tnf switch (T/F) and specify 3nf approximation: 0=tnf0b, 1=tnf1b, 2=tnf2b
{tnf}, {tnf_approx}
! 3nf cutoff(MeV),non-local reg. exp
450, 3
"""


def run_case(name):
    (A, matter, rho, nmax, order, inter, sub, step, tol, maxit, tw, dump, trip, ranks) = CASES[name][:14]
    cutoff = CASES[name][14] if len(CASES[name]) > 14 else 450
    tnf_approx = CASES[name][15] if len(CASES[name]) > 15 else -1
    with tempfile.TemporaryDirectory() as tmp:
        inp = os.path.join(tmp, "ccm_in")
        with open(inp, "w") as fh:
            fh.write(CCM_IN.format(order=order, cutoff=cutoff, A=A, matter=matter, rho=rho, nmax=nmax,
                                  tnf="T" if tnf_approx >= 0 else "F", tnf_approx=max(tnf_approx, 0),
                                  bc="PBC", ntwist=1, my_twists=1, cc="CCD"))
        tabc = tw[0] == "tabc"
        twargs = ["0.0", "0.0", "0.0", str(tw[1]), str(tw[2])] if tabc else [repr(x) for x in tw]
        cmd = [DRIVER, inp, tmp, inter, str(sub), str(step), repr(tol), str(maxit)] + twargs[:3] + \
              [str(dump), str(trip)] + twargs[3:]
        if ranks > 1:
            cmd = ["/opt/conda/bin/mpiexec", "-n", str(ranks)] + cmd
        env = dict(os.environ, OMP_NUM_THREADS="1", MKL_NUM_THREADS="1")
        res = subprocess.run(cmd, env=env, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"{name}: reference driver failed\n{res.stdout[-2000:]}\n{res.stderr[-2000:]}")
        out = {"params": np.array([A, nmax, sub, step, maxit, order], np.int64),
               "cutoff": np.float64(cutoff),
               "tnf_approx": np.int64(tnf_approx),
               "matter": np.array(matter), "interaction": np.array(inter),
               "rho": np.float64(rho), "tol": np.float64(np.float32(tol)),
               "twist": np.zeros(3)}
        its = []
        for line in open(os.path.join(tmp, "energies.txt")):
            f = line.split()
            if f[0] == "it":
                its.append(complex(float(f[2]), float(f[3])))
            elif f[0] in ("e0", "eccd", "eccdt"):
                out[f[0]] = np.complex128(complex(float(f[1]), float(f[2])))
            elif f[0] == "kf":
                out["kf"] = np.float64(float(f[1]))
            elif f[0] == "box":
                out["box"] = np.float64(float(f[1]))
                out["volume"] = np.float64(float(f[2]))
            elif f[0] == "twist":
                out["twist"] = np.array([float(x) for x in f[1:4]], np.float64)
                if tabc:
                    out["tabc_weight"] = np.float64(float(f[4]))
            elif f[0] == "looptime":
                out["looptime"] = np.float64(float(f[1]))
        out["energies"] = np.array(its, np.complex128)
        if tabc:
            # the reference program's own TABC branch for my_twists = k
            deck = os.path.join(tmp, "ccm_in_tabc")
            with open(deck, "w") as fh:
                fh.write(CCM_IN.format(order=order, cutoff=cutoff, A=A, matter=matter, rho=rho, nmax=nmax,
                                      tnf="T" if tnf_approx >= 0 else "F", tnf_approx=max(tnf_approx, 0),
                                      bc="TABC", ntwist=tw[1], my_twists=tw[2], cc="mbpt2"))
            res = subprocess.run([PROGRAM, deck], env=env, cwd=tmp, capture_output=True, text=True)
            line = [x for x in res.stdout.splitlines() if "Twist averaged E0/N" in x]
            if res.returncode != 0 or not line:
                raise RuntimeError(f"{name}: reference program failed\n{res.stdout[-2000:]}")
            re_, im = line[-1].split("(")[1].rstrip(")").split(",")
            out["tabc_e0_per_n"] = np.complex128(complex(float(re_), float(im)))
        if dump:
            with open(os.path.join(tmp, "orbitals.bin"), "rb") as fh:
                nsp, na = np.frombuffer(fh.read(8), np.int32)
                ints = np.frombuffer(fh.read(4 * 5 * nsp), np.int32).reshape(5, nsp)
                dbl = np.frombuffer(fh.read(8 * 4 * nsp), np.float64).reshape(4, nsp)
            out["orb_int"] = ints.copy()
            out["orb_dbl"] = dbl.copy()
            raw = np.fromfile(os.path.join(tmp, "fock.bin"), np.complex128)
            out["fock"] = raw[: nsp * nsp].reshape(nsp, nsp).T.copy()
            rec = np.dtype([("kind", "<i4"), ("p", "<i4"), ("q", "<i4"), ("r", "<i4"),
                            ("s", "<i4"), ("v", "<c16")])
            vb = np.fromfile(os.path.join(tmp, "vblocks.bin"), rec)
            out["v_idx"] = np.stack([vb["kind"], vb["p"], vb["q"], vb["r"], vb["s"]], 1).astype(np.int16)
            out["v_val"] = vb["v"].copy()
            trec = np.dtype([("a", "<i4"), ("b", "<i4"), ("i", "<i4"), ("j", "<i4"), ("v", "<c16")])
            t2 = np.fromfile(os.path.join(tmp, "t2.bin"), trec)
            out["t2_idx"] = np.stack([t2["a"], t2["b"], t2["i"], t2["j"]], 1).astype(np.int16)
            out["t2_val"] = t2["v"].copy()
            if trip:
                hrec = np.dtype([("k", "<i4"), ("d", "<i4"), ("c", "<i4"), ("b", "<i4"), ("v", "<c16")])
                hp = np.fromfile(os.path.join(tmp, "vhppp.bin"), hrec)
                out["hppp_idx"] = np.stack([hp["k"], hp["d"], hp["c"], hp["b"]], 1).astype(np.int16)
                out["hppp_val"] = hp["v"].copy()
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print(name, "iterations", len(out["energies"]) - 1, "eccd", out.get("eccd"), "eccdt", out.get("eccdt"))


if __name__ == "__main__":
    names = sys.argv[1:] or list(CASES) + list(PROGRAM_CASES)
    for n in names:
        if n in PROGRAM_CASES:
            run_program_case(n)
        else:
            run_case(n)
