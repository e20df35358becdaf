"""Host-side checks that need no GPU: the C ABI library loads and exports every
symbol include/ccgpu.h declares, and the native lattice plan (basis, channels,
Pandya channels) agrees with the oracle and with closed forms."""
import math
import os
import re

import numpy as np
import pytest

import ccd_oracle as co
from conftest import ROOT


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "ccgpu.h")).read()
    return sorted(set(re.findall(r"^int (ccgpu_\w+)\(", txt, re.M)))


def test_library_exports_every_header_symbol(ccnm):
    lib = ccnm.lib()
    syms = header_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(lib, s), s
        assert s in ccnm.SIGNATURES, s


@pytest.mark.parametrize("A,matter,rho,nmax,twist", [
    (14, "pnm", 0.16, 1, (0, 0, 0)), (66, "pnm", 0.16, 2, (0, 0, 0)), (4, "snm", 0.16, 1, (0, 0, 0)),
    (28, "snm", 0.165, 2, (0, 0, 0)), (66, "pnm", 0.16, 2, (0.7, 1.9, 2.6)), (14, "pnm", 0.08, 2, (0.3, 1.1, 2.0))])
def test_plan_matches_oracle_channels(ccnm, A, matter, rho, nmax, twist):
    d = ccnm.plan(A, ccnm.PNM if matter == "pnm" else ccnm.SNM, rho, nmax, twist)
    lat = co.Lattice.build(A, matter, rho, nmax, twist)
    ch = co.Channels(lat)
    assert d["nsp"] == lat.nsp == (2 if matter == "pnm" else 4) * (2 * nmax + 1) ** 3
    assert d["t2_ordered"] == ch.size
    nuniq, nchan, maxpp, maxhh = 0, 0, 0, 0
    for h, p in zip(ch.hh, ch.pp):
        nh = int(np.sum(h[:, 0] < h[:, 1]))
        npp = int(np.sum(p[:, 0] < p[:, 1]))
        if nh and npp:
            nchan += 1
            nuniq += nh * npp
            maxpp, maxhh = max(maxpp, npp), max(maxhh, nh)
    assert d["nchannels"] == nchan
    assert d["t2_unique"] == nuniq
    assert (d["max_npp"], d["max_nhh"]) == (maxpp, maxhh)
    assert d["nph_channels"] == len(ch.ph_keys)
    assert d["ring_elems"] == sum(len(a) * len(b) for a, b in zip(ch.ph, ch.hp))
    # closed forms: box L = (A/rho)^(1/3), k_F of the matter type
    assert abs(d["box"] - (A / rho) ** (1 / 3)) < 1e-12
    g = 1 if matter == "pnm" else 2
    assert abs(d["kf"] - (3 * co.PI ** 2 * rho / g) ** (1 / 3)) < 1e-12


def test_plan_cfg_sizes_match_brute_force_counts(ccnm):
    """PNM A=66 nmax=4: 1458 states, ordered |T2| = 8 067 968 (reference diis space),
    185 (K,Tz) channels; counts from brute-force enumeration of ordered pairs."""
    d = ccnm.plan(66, ccnm.PNM, 0.16, 4)
    assert d["nsp"] == 1458
    assert d["t2_ordered"] == 8067968
    assert d["nchannels"] == 185


def test_closed_shell_hole_sets(ccnm):
    """At zero twist the A lowest states fill whole momentum shells (1+6+12+8+6
    momenta for A=66 PNM), so the hole set is unique."""
    lat = co.Lattice.build(66, "pnm", 0.16, 2)
    n2 = np.sum(lat.n ** 2, 1)
    assert np.all(n2[:66] <= 4) and np.all(n2[66:] >= 4)
    assert np.sum(n2[:66] == 4) == 12 and np.sum(n2 <= 4) == 66 + 2 * 0 + (np.sum(n2[66:] == 4))


def test_plan_rejects_bad_input(ccnm):
    with pytest.raises(RuntimeError):
        ccnm.plan(0, ccnm.PNM, 0.16, 2)
    with pytest.raises(RuntimeError):
        ccnm.plan(66, 7, 0.16, 2)
    with pytest.raises(RuntimeError):
        ccnm.plan(66, ccnm.PNM, -1.0, 2)
    with pytest.raises(RuntimeError):
        ccnm.plan(5000, ccnm.PNM, 0.16, 1)
