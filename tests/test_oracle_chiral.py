"""The oracle's chiral NN (oracle/chiral_oracle.py) pinned against the
compiled reference's chiral_NN_with_delta elements, and against what the
mathematics fixes: hermiticity, twist independence, antisymmetry, the
quadrature and the 3j symbols."""
import itertools

import numpy as np
import pytest

import ccd_oracle as co
import chiral_oracle as ch
from conftest import fixture_lattice, golden

# every fixture holding chiral V records: (name, order, cutoff)
CHIRAL_CASES = [("pnm14_n1_nnlo", 3, 450.0), ("snm4_n1_nnlo", 3, 450.0), ("snm4_n1_lo450", 0, 450.0),
                ("snm4_n1_nlo450", 2, 450.0), ("snm4_n1_nnlo394", 3, 394.0), ("snm4_n1_nnlo500", 3, 500.0),
                ("pnm14_n1_nnlo_twist", 3, 450.0)]


@pytest.mark.parametrize("name,order,cutoff", CHIRAL_CASES)
def test_oracle_chiral_matches_reference_elements(name, order, cutoff):
    """Every stored <pq||rs> record of the reference's vnn_* blocks (LO, NLO
    and NNLO; cutoffs 394/450/500; PNM and SNM with np pairs; a twisted
    lattice) to rounding."""
    g = golden(name)
    assert int(g["params"][5]) == order and float(g.get("cutoff", 450.0)) == cutoff
    lat = fixture_lattice(g)
    it = ch.ChiralInteraction(lat, order, cutoff)
    idx = g["v_idx"].astype(np.int64)
    v = it(idx[:, 1] - 1, idx[:, 2] - 1, idx[:, 3] - 1, idx[:, 4] - 1)
    ref = g["v_val"]
    assert np.max(np.abs(v - ref)) < 1e-14 * max(1.0, np.max(np.abs(ref)))


def _random_quadruples(lat, n, seed):
    """n momentum- and charge-conserving quadruples (p, q, r, s) of a lattice."""
    rng = np.random.default_rng(seed)
    key = {}
    for x in range(lat.nsp):
        key.setdefault((tuple(lat.n[x]), int(lat.tz[x])), []).append(x)
    out = []
    while len(out) < n:
        p, q, r = rng.integers(0, lat.nsp, 3)
        nn = lat.n[p] + lat.n[q] - lat.n[r]
        tz = int(lat.tz[p] + lat.tz[q] - lat.tz[r])
        cands = key.get((tuple(nn), tz))
        if cands:
            out.append((p, q, r, cands[rng.integers(0, len(cands))]))
    return np.array(out, np.int64).T


def test_chiral_hermitian_and_antisymmetric():
    """<pq||rs> = conj <rs||pq> = -<qp||rs> = -<pq||sr> (regulator and
    operators symmetric, chiral_module_andreas_with_delta.f90:3358-3378)."""
    lat = co.Lattice.build(28, "snm", 0.16, 2)
    it = ch.ChiralInteraction(lat)
    p, q, r, s = _random_quadruples(lat, 3000, 7)
    v = it(p, q, r, s)
    assert np.max(np.abs(v)) > 1.0
    scale = np.max(np.abs(v))
    assert np.max(np.abs(v - np.conj(it(r, s, p, q)))) < 1e-13 * scale
    assert np.max(np.abs(v + it(q, p, r, s))) < 1e-13 * scale
    assert np.max(np.abs(v + it(p, q, s, r))) < 1e-13 * scale


def test_chiral_twist_independent():
    """Only relative momenta enter, so a twist leaves every element (by lattice
    labels) unchanged."""
    a = co.Lattice.build(14, "pnm", 0.16, 1)
    b = co.Lattice.build(14, "pnm", 0.16, 1, twist=(0.3, 1.7, 2.9))
    lab_b = {(tuple(b.n[x]), int(b.sz[x]), int(b.tz[x])): x for x in range(b.nsp)}
    amap = np.array([lab_b[(tuple(a.n[x]), int(a.sz[x]), int(a.tz[x]))] for x in range(a.nsp)])
    p, q, r, s = _random_quadruples(a, 2000, 3)
    va = ch.ChiralInteraction(a)(p, q, r, s)
    vb = ch.ChiralInteraction(b)(amap[p], amap[q], amap[r], amap[s])
    assert np.max(np.abs(va - vb)) < 1e-12 * np.max(np.abs(va))


def test_gauss_legendre_exact_for_polynomials():
    """The 25-point rule integrates every polynomial of degree <= 49 exactly.
    The reference's mu mesh (chiral_module_andreas_with_delta.f90:2606-2620)
    is that rule with its middle node and weight replaced by default-real
    (float32) literals; the weight literal is not the Gauss weight, so the
    restated mesh keeps the replacement rather than the exact rule."""
    a, b = 2.0 * ch.MPI[2], ch.SFR
    x, w = ch.gauss_legendre(25, a, b)
    for k in (0, 1, 7, 24, 49):
        exact = (b - a) / (k + 1)
        assert abs(np.sum(w * ((x - a) / (b - a)) ** k) - exact) < 1e-12 * exact
    assert np.all(np.diff(x) > 0)
    assert abs(x[12] - 0.5 * (a + b)) < 1e-12
    cc = ch.ChiralConstants()
    keep = np.arange(25) != 12
    assert np.array_equal(cc.mu[keep], x[keep]) and np.array_equal(cc.wmu[keep], w[keep])
    assert cc.mu[12] == float(np.float32(488.038986666667))
    assert cc.wmu[12] == float(np.float32(25.5194921127220))
    assert abs(cc.wmu[12] - w[12]) > 0.5


def test_three_j_values():
    """Closed forms and the orthogonality relation of the 3j symbols."""
    assert abs(ch.three_j(1, 1, 0, 1, -1, 0) - 1 / np.sqrt(2)) < 1e-15
    assert abs(ch.three_j(1, 1, 0, -1, 1, 0) + 1 / np.sqrt(2)) < 1e-15
    assert abs(ch.three_j(1, 1, 2, 1, -1, 0) - 1 / np.sqrt(6)) < 1e-15
    assert abs(ch.three_j(1, 1, 2, 1, 1, -2) + 1 / np.sqrt(3)) < 1e-15   # (-1)^(j1-j2-M3)/sqrt(2J+1)
    # sum_{m1 m2} (j1 j2 J; m1 m2 M)^2 = 1/(2J+1)
    for j1, j2 in ((1, 1), (2, 1), (3, 2)):
        for J in range(abs(j1 - j2), j1 + j2 + 1, 2):
            for M in range(-J, J + 1, 2):
                s = sum(ch.three_j(j1, j2, J, m1, m2, -M) ** 2
                        for m1, m2 in itertools.product(range(-j1, j1 + 1, 2), range(-j2, j2 + 1, 2)))
                assert abs(s - 1.0 / (J + 1)) < 1e-14


def test_lo_one_pion_exchange_closed_form():
    """LO nn element of two spin-up neutrons scattering by momentum transfer q
    along z: <++|(s1.q)(s2.q)|++> = q^2, tau.tau = 1, so the direct part is
    (q^2 OPE(q) + CS + CT) x regulator x relativity x hbarc^3/V with
    OPE(q) = -gA^2/(4 fpi^2)/(q^2 + m_pi^2) (average pion mass)."""
    lat = co.Lattice.build(14, "pnm", 0.16, 1)
    cc = ch.ChiralConstants(0, 450.0)
    lab = {(tuple(lat.n[x]), int(lat.sz[x])): x for x in range(lat.nsp)}
    p, q = lab[((0, 0, 1), 1)], lab[((0, 0, -1), 1)]
    r, s = lab[((0, 0, 0), 1)], lab[((0, 0, 0), 1)]
    # r == s makes the ket a Pauli-forbidden state only after antisymmetrizing;
    # the direct element itself is well defined
    v = ch.chiral_direct(lat, cc, np.array([p]), np.array([q]), np.array([r]), np.array([s]))[0]
    kz = lat.k[p][2] * ch.HBARC
    q2 = kz * kz                       # prel = (0, 0, kz), pprel = 0
    ope = -(ch.GA * ch.GA / (4.0 * ch.FPI * ch.FPI)) / (q2 + ch.MPI[2] ** 2)
    m = ch.MNUC[1]                     # nn pair: neutron mass
    rel = m / np.sqrt(np.sqrt((m ** 2 + q2) * m ** 2))
    reg = np.exp(-(q2 ** 3) / 450.0 ** 6)
    want = (q2 * ope + cc.CS[0] + cc.CT[0]) * rel * reg * ch.HBARC ** 3 / lat.volume
    assert abs(v - want) < 1e-13 * abs(want)


TNF_CASES = ["pnm14_n1_nnlo_3nf", "snm4_n1_nnlo_3nf", "snm4_n1_nlo450_3nf", "pnm14_n1_nnlo_3nf1"]


@pytest.mark.parametrize("name", TNF_CASES)
def test_oracle_3nf_normal_ordering_matches_reference(name):
    """The NNLO three-nucleon force normal-ordered with respect to the holes
    (normal_ordered_hamiltonian, setup_3nf_2bNO): E0 and the Fock matrix
    against the reference (zero- and one-body parts), and at the two-body
    level every stored block record (hhhh, hhpp, hphp, pppp with the NO2B
    term, hhhp without it), NNLO and NLO (Delta terms), PNM and SNM."""
    g = golden(name)
    lat = fixture_lattice(g)
    order, approx = int(g["params"][5]), int(g["tnf_approx"])
    nn = ch.ChiralInteraction(lat, order, 450.0)
    no = ch.NormalOrdered3NF(lat, nn.cc)
    A = lat.A
    i, j = (x.ravel() for x in np.meshgrid(np.arange(A), np.arange(A), indexing="ij"))
    e0 = np.sum(lat.e[:A]) + 0.5 * np.sum(nn(i, j, i, j)) + no.e0()
    assert abs(e0 - g["e0"]) < 1e-12 * abs(g["e0"])
    pd = np.arange(lat.nsp)
    f = lat.e.astype(np.complex128)
    for h in range(A):
        f = f + nn(pd, np.full(lat.nsp, h), pd, np.full(lat.nsp, h))
    if approx >= 1:
        f = f + no.fock(pd, pd)
    assert np.max(np.abs(f - np.diag(g["fock"]))) < 1e-12 * np.max(np.abs(f))
    if approx >= 2:
        idx = g["v_idx"].astype(np.int64)
        ref = g["v_val"]
        for kind in np.unique(idx[:, 0]):
            m = idx[:, 0] == kind
            p, q, r, s = (idx[m, c] - 1 for c in range(1, 5))
            v = nn(p, q, r, s) + no.block(int(kind), p, q, r, s)
            assert np.max(np.abs(v - ref[m])) < 1e-13 * max(1.0, np.max(np.abs(ref[m]))), kind
        # (T)'s <cb||kd> records (vnn_hppp) carry the NO2B term as well
        if "hppp_idx" in g:
            h = g["hppp_idx"].astype(np.int64) - 1
            k_, d_, c_, b_ = h.T
            v = nn(c_, b_, k_, d_) + no.block(5, c_, b_, k_, d_)
            assert np.max(np.abs(v - g["hppp_val"])) < 1e-13 * np.max(np.abs(g["hppp_val"]))


def test_3nf_pair_symmetry_broken_only_by_the_float32_two_thirds():
    """The reference's 3NF regulator uses the default-real (2./3.): the
    normal-ordered element is antisymmetric and Hermitian only to ~1e-7 (the
    reason the block conventions above matter), and exactly so with 2/3 in
    double precision."""
    lat = co.Lattice.build(14, "pnm", 0.16, 1)
    nn = ch.ChiralInteraction(lat)
    no = ch.NormalOrdered3NF(lat, nn.cc)
    p, q, r, s = _random_quadruples(lat, 300, 5)
    keep = (p != q) & (r != s)
    p, q, r, s = p[keep], q[keep], r[keep], s[keep]
    a = no.two_body(p, q, r, s)
    scale = np.max(np.abs(a))
    assert scale > 0.1
    asym = np.max(np.abs(a + no.two_body(q, p, r, s)))
    herm = np.max(np.abs(a - np.conj(no.two_body(r, s, p, q))))
    assert 1e-12 * scale < max(asym, herm) < 1e-5 * scale
    saved = ch.F32_TWO_THIRDS
    try:
        ch.F32_TWO_THIRDS = 2.0 / 3.0
        a = no.two_body(p, q, r, s)
        assert np.max(np.abs(a + no.two_body(q, p, r, s))) < 1e-13 * scale
        assert np.max(np.abs(a - np.conj(no.two_body(r, s, p, q)))) < 1e-13 * scale
    finally:
        ch.F32_TWO_THIRDS = saved
