"""CPU restatement of the reference's Delta-full chiral NN interaction.

TEST INFRASTRUCTURE ONLY.  This module is part of the oracle: tests/, the smoke
check and the CPU-baseline leg of bench.py may import it as the *checker*.  The
product (the HIP library and its hosts) never imports, links or calls it.

It evaluates chiral_NN_with_delta (chiral_module_andreas_with_delta.f90:6304-6668)
element by element, vectorized over arrays of orbital indices, with the
reference's arithmetic order where it matters for the last bits:

  constants and LECs   init_chp_constants, :122-808 (orders LO=0, NLO=2, NNLO=3;
                       cutoffs 394, 450, 500 MeV of the Delta-full family)
  Gauss-Legendre mesh  chp_setup_gauleg_mesh, :824-880; the 25-point mu mesh of
                       chp_setup_delta_loop_int_data, :2606-2620, including its
                       13th point overwritten with default-real (float32) literals
  loop functions       :1110-1300 (one-pion exchange, w, s, L, wtilde^2, A),
                       :2622-2680 (D, H), :1341-1515 (NNLO 1loop_d)
  Delta diagrams       :2683-2925
  spin/isospin         :2944-3300 (sigma.sigma, tau.tau, sigma1.q sigma2.q,
                       i/2 (sigma1+sigma2).(q x k))
  regulator            freg_delta, :3358-3378
  3j symbols           tjs, ccm_modules.f90:187-226 (standard Wigner 3j)

Pins (tests/test_oracle_chiral.py): every stored V record of the chiral
fixtures written by the compiled reference (tests/golden/*_nnlo*.npz and the
order/cutoff cases), hermiticity, twist independence, and the closed forms of
the one-pion-exchange limit.
"""
from __future__ import annotations

from math import factorial, sqrt

import numpy as np

HBARC = 197.326968
PI = 3.14159265358979

# chiral_constants (chiral_module_andreas_with_delta.f90:43-121)
GA = 1.28899999999999992
HA = 1.39999999999999991
FPI = 92.2000000000000028
PROTON_MASS = 938.272046000000046
NEUTRON_MASS = 939.565379000000007
DELTA_MASS = 1232.0
MNUC = {-1: PROTON_MASS, 0: 938.918267117927371, 1: NEUTRON_MASS}
MPI = {-1: 139.570179999999993, 0: 134.976599999999991, 1: 139.570179999999993}
MPI[2] = (MPI[-1] + MPI[0] + MPI[1]) / 3.0
SFR = 700.0

LO, NLO, NNLO = 0, 2, 3

# LECs of the Delta-full family (init_chp_constants, :172-530): c_i in GeV^-1,
# LO contacts c1s0/c3s1 per pair charge (-1 pp, 0 np, 1 nn) in 10^4 GeV^-2,
# NLO contacts (1S0, 3P0, 1P1, 3P1, 3S1, 3S1-3D1, 3P2) in 10^4 GeV^-4.
_CI_NNLO = (-7.39999999999999991e-01, -4.89999999999999991e-01, -6.50000000000000022e-01,
            9.59999999999999964e-01)
_LO_C = {
    500.0: ((-1.08522470229349993e-01,) * 3, (-6.84437111953200006e-02,) * 3),
    450.0: ((-1.12926682348719998e-01,) * 3, (-8.73398209265599940e-02,) * 3),
    394.0: ((-1.16797069688160005e-01,) * 3, (-1.03577257368989997e-01,) * 3),
}
_NLO_C = {
    500.0: ((-3.09617914433389985e-01, -3.10443063173440004e-01, -3.10256027522040012e-01),
            (-1.91012821791280013e-01,) * 3),
    450.0: ((-3.09893184381559994e-01, -3.10711872064980021e-01, -3.10510995610929985e-01),
            (-1.97950619113270027e-01,) * 3),
    394.0: ((-3.09925223545309991e-01, -3.10751704075070023e-01, -3.10539966723270000e-01),
            (-2.05462289226700012e-01,) * 3),
}
_NNLO_C = {
    500.0: ((-3.37302636943159984e-01, -3.38320181277209975e-01, -3.38223065406040024e-01),
            (-2.21721491662439985e-01,) * 3),
    450.0: ((-3.37136656357900022e-01, -3.38139465283629981e-01, -3.38023088490550017e-01),
            (-2.29310079440350006e-01,) * 3),
    394.0: ((-3.36845687915189984e-01, -3.37854531334379982e-01, -3.37722607113270012e-01),
            (-2.36959231730459990e-01,) * 3),
}
_NLO_PW = {
    (NLO, 500.0): (2.39537506986395998, 7.27048937924560046e-01, 1.52465155576100003e-02,
                   -9.51416990370669957e-01, 5.39378033769969956e-01, 3.58443125233390003e-01,
                   -7.93620896751420002e-01),
    (NLO, 450.0): (2.39163813513152990, 6.86901585796929948e-01, 4.81266565581999969e-03,
                   -1.00011242015441004, 5.58973150278149977e-01, 3.62094234568600015e-01,
                   -8.08073350643100041e-01),
    (NLO, 394.0): (2.39117370641156990, 6.65731825802269972e-01, -1.52505123405899980e-02,
                   -1.04638857865136004, 5.88755038823760013e-01, 3.81315710481519965e-01,
                   -8.18066377264660027e-01),
    (NNLO, 500.0): (2.48801946794199980, 6.98453619479979970e-01, -1.26513321381799999e-02,
                    -9.37263520317270116e-01, 6.75353391585620022e-01, 3.54479450076249980e-01,
                    -8.59526137679109903e-01),
    (NNLO, 450.0): (2.47658908242146980, 6.45550411071979968e-01, -2.85410030715300030e-02,
                    -1.02235931835912996, 6.95953209842609999e-01, 3.58329844893869986e-01,
                    -8.70203217397279905e-01),
    (NNLO, 394.0): (2.46900829544757006, 6.16935689120689967e-01, -5.26929817733700068e-02,
                    -1.09054771638107995, 7.22641309492930040e-01, 3.75968068144939938e-01,
                    -8.76265471802940010e-01),
}


def three_j(j1, j2, j3, m1, m2, m3):
    """Wigner 3j symbol with doubled arguments (Racah's formula); tjs of
    ccm_modules.f90:187-226 computes the same standard symbol."""
    if m1 + m2 + m3 != 0:
        return 0.0
    if any((j + m) % 2 for j, m in ((j1, m1), (j2, m2), (j3, m3))):
        return 0.0
    if j3 < abs(j1 - j2) or j3 > j1 + j2 or (j1 + j2 + j3) % 2:
        return 0.0
    if abs(m1) > j1 or abs(m2) > j2 or abs(m3) > j3:
        return 0.0
    a = lambda x: x // 2
    tri = (factorial(a(j1 + j2 - j3)) * factorial(a(j1 - j2 + j3)) * factorial(a(-j1 + j2 + j3))
           / factorial(a(j1 + j2 + j3) + 1))
    pre = sqrt(tri * factorial(a(j1 + m1)) * factorial(a(j1 - m1)) * factorial(a(j2 + m2))
               * factorial(a(j2 - m2)) * factorial(a(j3 + m3)) * factorial(a(j3 - m3)))
    s = 0.0
    for k in range(0, a(j1 + j2 - j3) + 1):
        d = (a(j1 + j2 - j3) - k, a(j1 - m1) - k, a(j2 + m2) - k, a(j3 - j2 + m1) + k,
             a(j3 - j1 - m2) + k)
        if min(d) < 0:
            continue
        den = factorial(k)
        for x in d:
            den *= factorial(x)
        s += (-1) ** k / den
    return (-1) ** a(j1 - j2 - m3) * pre * s


def gauss_legendre(n, x1, x2):
    """Nodes (ascending) and weights on [x1, x2] by Newton's method on the
    Legendre recurrence, the scheme of chp_setup_gauleg_mesh (:824-880)."""
    x = np.zeros(n)
    w = np.zeros(n)
    m = (n + 1) // 2
    xm, xl = 0.5 * (x2 + x1), 0.5 * (x2 - x1)
    for i in range(1, m + 1):
        z = np.cos(PI * (i - 0.25) / (n + 0.5))
        # at least one Newton step: for odd n the middle start z = cos(pi/2)
        # is already within the tolerance of the first z1 = 0, and the
        # reference's loop then keeps the derivative of the previous node (its
        # 25-point mesh replaces that node afterwards, ChiralConstants)
        while True:
            p1, p2 = 1.0, 0.0
            for j in range(1, n + 1):
                p3 = p2
                p2 = p1
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j
            pp = n * (z * p1 - p2) / (z * z - 1.0)
            z1 = z
            z = z - p1 / pp
            if abs(z - z1) <= 3e-14:
                break
        x[i - 1] = xm - xl * z
        x[n - i] = xm + xl * z
        w[i - 1] = 2.0 * xl / ((1.0 - z * z) * pp * pp)
        w[n - i] = w[i - 1]
    return x, w


class ChiralConstants:
    """init_chp_constants (:122-808) for one chiral order and cutoff, with
    optional overrides of the NN c1 and c3 (GeV^-1)."""

    def __init__(self, order=NNLO, cutoff=450.0, c1=None, c3=None):
        cutoff = float(cutoff)
        if order not in (LO, NLO, NNLO):
            raise ValueError(f"chiral order {order} (LO=0, NLO=2, NNLO=3)")
        if cutoff not in (394.0, 450.0, 500.0):
            raise ValueError(f"cutoff {cutoff} MeV (394, 450, 500)")
        self.order, self.cutoff = order, cutoff
        self.nexp = 4.0 if abs(cutoff - 394.0) < 1e-5 else 3.0
        ci = _CI_NNLO if order == NNLO else (0.0, 0.0, 0.0, 0.0)
        ci = list(ci)
        if c1 is not None:
            ci[0] = c1
        if c3 is not None:
            ci[2] = c3
        self.c1, self.c2, self.c3, self.c4 = (x * 1.0e-3 for x in ci)
        c1s0, c3s1 = {LO: _LO_C, NLO: _NLO_C, NNLO: _NNLO_C}[order][cutoff]
        pw = _NLO_PW.get((order, cutoff), (0.0,) * 7)
        pw = [x * 1.0e-08 for x in pw]
        self.CS = {}
        self.CT = {}
        for i, ch in enumerate((-1, 0, 1)):
            a, b = c1s0[i] * 0.01, c3s1[i] * 0.01
            self.CS[ch] = (a + 3.0 * b) / 16.0 / PI
            self.CT[ch] = (b - a) / 16.0 / PI
        s2 = sqrt(2.0)
        self.cnlo = [
            (-5.0 * pw[6] + 6.0 * pw[4] - 3.0 * pw[3] - 3.0 * pw[2] - pw[1] + 2.0 * pw[0]) / (64.0 * PI),
            (5.0 * pw[6] + 6.0 * pw[4] + 3.0 * pw[3] + 3.0 * pw[2] + pw[1] + 2.0 * pw[0]) / (16.0 * PI),
            -(2.0 * pw[6] - 2.0 * s2 * pw[5] - 2.0 * pw[4] - 3.0 * pw[2] + pw[1] + 2.0 * pw[0]) / (64.0 * PI),
            -(-2.0 * pw[6] - 2.0 * s2 * pw[5] - 2.0 * pw[4] + 3.0 * pw[2] - pw[1] + 2.0 * pw[0]) / (16.0 * PI),
            -(-5.0 * pw[6] + 3.0 * pw[3] + 2.0 * pw[1]) / (16.0 * PI),
            (pw[6] - 6.0 * s2 * pw[5] - 3.0 * pw[3] + 2.0 * pw[1]) / (64.0 * PI),
            -(pw[6] + 6.0 * s2 * pw[5] - 3.0 * pw[3] + 2.0 * pw[1]) / (16.0 * PI),
        ]
        ga2 = GA * GA
        ga4 = ga2 * ga2
        fpi2 = FPI * FPI
        fpi4 = fpi2 * fpi2
        pi2 = PI * PI
        self.ga2, self.ha2, self.ha4 = ga2, HA * HA, (HA * HA) * (HA * HA)
        self.k2 = ga2 / (4.0 * fpi2)
        self.k3 = 384.0 * pi2 * fpi4
        self.k4 = 5.0 * ga4 - 4.0 * ga2 - 1.0
        self.k5 = 23.0 * ga4 - 10.0 * ga2 - 1.0
        self.k6 = 48.0 * ga4
        self.k7 = 3.0 * ga4
        self.k8 = 64.0 * pi2 * fpi4
        self.k9 = 3.0 * ga2 / (16.0 * PI * fpi4)
        self.k11 = 2.0 * (2.0 * self.c1 - self.c3)
        self.k16 = ga2 / (32.0 * PI * fpi4)
        self.k22 = 1.0 / (pi2 * fpi4)
        self.k24 = 1.0 / (PI * fpi4)
        self.delta = DELTA_MASS - MNUC[0]
        self.mpi2 = {i: MPI[i] * MPI[i] for i in MPI}
        mu, wmu = gauss_legendre(25, 2.0 * MPI[2], SFR)
        mu[12] = float(np.float32(488.038986666667))
        wmu[12] = float(np.float32(25.5194921127220))
        self.mu, self.wmu = mu, wmu


def _pauli(m_out, m_in, v):
    """<m_out| sigma.v |m_in> for m = +-1 (spin up first, as the reference's
    chi vectors), v complex or real (..., 3)."""
    vx, vy, vz = v[..., 0], v[..., 1], v[..., 2]
    up_up = (m_out == 1) & (m_in == 1)
    dn_dn = (m_out == -1) & (m_in == -1)
    up_dn = (m_out == 1) & (m_in == -1)
    dn_up = (m_out == -1) & (m_in == 1)
    return (np.where(up_up, vz, 0.0) + np.where(dn_dn, -vz, 0.0)
            + np.where(up_dn, vx - 1j * vy, 0.0) + np.where(dn_up, vx + 1j * vy, 0.0))


def _pair_dot(m1, m2, m3, m4):
    """<m1 m2| sigma1.sigma2 |m3 m4> (chp_sigma_dot_sigma_mtx, :2944-2974)."""
    diag = (m1 == m3) & (m2 == m4)
    exch = (m1 == m4) & (m2 == m3) & (m1 != m2)
    return np.where(diag, (m1 * m2).astype(np.float64), 0.0) + np.where(exch, 2.0, 0.0)


def chiral_direct(lat, cc: ChiralConstants, p, q, r, s):
    """chiral_NN_with_delta(p,q,r,s) (:6304-6668), complex, in MeV."""
    p, q, r, s = (np.asarray(x, np.int64) for x in (p, q, r, s))
    k1, k2, k3, k4 = (lat.k[x] * HBARC for x in (p, q, r, s))
    m1, m2, m3, m4 = (lat.sz[x] for x in (p, q, r, s))
    t1, t2, t3, t4 = (lat.tz[x] for x in (p, q, r, s))
    ok = np.all(lat.n[p] + lat.n[q] == lat.n[r] + lat.n[s], axis=-1) & (t1 + t2 == t3 + t4)
    prel = 0.5 * (k1 - k2)
    pprel = 0.5 * (k3 - k4)
    p2 = np.sum(prel * prel, axis=-1)
    pp2 = np.sum(pprel * pprel, axis=-1)
    kmean = 0.5 * (prel + pprel)
    kmean2 = np.sum(kmean * kmean, axis=-1)
    charge = (t1 + t2) // 2
    mass = np.select([charge == -1, charge == 0], [MNUC[-1], MNUC[0]], MNUC[1])
    rel = mass / np.sqrt(np.sqrt((mass ** 2 + p2) * (mass ** 2 + pp2)))
    qt = prel - pprel
    q2 = np.sum(qt * qt, axis=-1)
    qabs = np.sqrt(q2)
    qxk = np.cross(qt, kmean)

    dm = ((m1 == m3) & (m2 == m4)).astype(np.float64)
    dt = ((t1 == t3) & (t2 == t4)).astype(np.float64)
    ss = _pair_dot(m1, m2, m3, m4)
    tt = _pair_dot(t1, t2, t3, t4)
    sqsq = _pauli(m1, m3, qt) * _pauli(m2, m4, qt)
    mpi2 = cc.mpi2

    def ope(qq, i):
        return -1.0 * cc.k2 / (qq + mpi2[i])

    vlo = np.zeros(p.shape, np.complex128)
    cont_lo = np.zeros(p.shape, np.complex128)
    cont_nlo = np.zeros(p.shape, np.complex128)
    vnlo = np.zeros(p.shape, np.complex128)
    vnnlo = np.zeros(p.shape, np.complex128)
    if cc.order == LO:
        vlo = sqsq * ope(q2, 2) * tt
        cont_lo = cc.CS[0] * dm * dt + cc.CT[0] * ss * dt
    else:
        fm = 4.0 * mpi2[2]
        w = np.sqrt(fm + q2)
        sl = np.sqrt(SFR * SFR - fm)
        # q = 0: the loop functions move q to 1e-8 (and q^2 to 1e-16) for every
        # later use in the element (INTENT(INOUT) arguments, :1244-1312)
        zero = qabs == 0.0
        qabs = np.where(zero, 1.0e-8, qabs)
        q2 = np.where(zero, qabs * qabs, q2)
        L = w * np.log((SFR * SFR * w * w + q2 * sl * sl + 2.0 * SFR * qabs * w * sl)
                       / (fm * (SFR * SFR + q2))) / (2.0 * qabs)
        wt2 = 2.0 * mpi2[2] + q2
        twompi = 2.0 * MPI[2]
        A = np.arctan(qabs * (SFR - twompi) / (q2 + SFR * twompi)) / (2.0 * qabs)
        de = cc.delta
        D = np.zeros(p.shape)
        for mu, wmu in zip(cc.mu, cc.wmu):
            mut = np.sqrt(mu * mu - fm) / (2.0 * de)
            D = D + wmu * (np.arctan(mut) / (mu * mu + q2))
        D = D / de
        S = 2.0 * mpi2[2] + q2 - 2.0 * de ** 2.0
        qdel = 2.0 * np.sqrt(de ** 2 - mpi2[2])
        qdel2 = qdel * qdel
        wdel = np.sqrt(qdel2 + fm)
        sdel = np.sqrt(SFR * SFR - fm)
        delL = wdel * np.log((SFR * SFR * wdel * wdel + qdel2 * sdel * sdel + 2.0 * SFR * qdel * wdel * sdel)
                             / (fm * (SFR * SFR + qdel2))) / (2.0 * qdel)
        H = 2.0 * S * (L - delL) / (w ** 2 - 4.0 * de ** 2)

        # one-pion exchange with charge dependence (:6534-6551)
        WT = np.zeros(p.shape, np.complex128)
        np_pair = (t1 + t2) == 0
        if np.any(np_pair):
            acc = np.zeros(p.shape)
            for T2 in (0, 2):
                cg1 = np.array([three_j(1, 1, T2, int(a), int(b), 0) for a, b in zip(t1.ravel(), t2.ravel())]
                               ).reshape(p.shape) * sqrt(T2 + 1.0)
                cg2 = np.array([three_j(1, 1, T2, int(a), int(b), 0) for a, b in zip(t3.ravel(), t4.ravel())]
                               ).reshape(p.shape) * sqrt(T2 + 1.0)
                ph = (-1) ** (T2 // 2 + 1)
                acc = acc + cg1 * cg2 * (-ope(q2, 0) + ph * 2.0 * ope(q2, 1))
            WT = np.where(np_pair, acc, WT)
        WT = np.where(np_pair, WT, ope(q2, 0))
        vlo = WT * sqsq
        cs = np.select([charge == -1, charge == 0], [cc.CS[-1], cc.CS[0]], cc.CS[1])
        ct = np.select([charge == -1, charge == 0], [cc.CT[-1], cc.CT[0]], cc.CT[1])
        cont_lo = cs * dm * dt + ct * ss * dt

        ga2, ha2, ha4, k22, k24 = cc.ga2, cc.ha2, cc.ha4, cc.k22, cc.k24
        # NLO two-pion exchange and the leading Delta diagrams (:6553-6598, 2683-2870)
        Wc = -L * (fm * cc.k4 + q2 * cc.k5 + cc.k6 * mpi2[2] * mpi2[2] / (w * w)) / cc.k3
        Vs = cc.k7 * L * q2 / cc.k8
        VT = -cc.k7 * L / cc.k8
        Wc = Wc + (-ha2 * k22 * ((6.0 * S - w * w) * L + 12.0 * de * de * S * D) / 216.0)
        Wc = Wc + (-ga2 * ha2 * k22 * ((12.0 * de * de - 20.0 * mpi2[2] - 11.0 * q2) * L
                                       + 6.0 * S * S * D) / 216.0)
        sb_vt = -ga2 * ha2 * k22 * (-2.0 * L + (w * w - 4.0 * de ** 2) * D) / 48.0
        Vs = Vs + (-q2 * sb_vt)
        VT = VT + sb_vt
        Vc = -ga2 * ha2 * k24 * (2.0 * mpi2[2] + q2) * (2.0 * mpi2[2] + q2) * A / (12.0 * de)
        sb_wt = -ga2 * ha2 * k24 * (w * w * A) / (144.0 * de)
        WT2 = sb_wt
        Ws = -q2 * sb_wt
        Wc = Wc + (-ha4 * k22 * ((12.0 * S - w * w) * L + 3.0 * S * (H + (8.0 * de * de - S) * D)) / 486.0)
        db_vt = -ha4 * k22 * (6.0 * L + (12.0 * de * de - w * w) * D) / 216.0
        Vs = Vs + (-q2 * db_vt)
        VT = VT + db_vt
        Vc = Vc + (-ha4 * k22 * (-4.0 * de * de * L + S * (H + (S + 8.0 * de * de) * D)) / 27.0)
        db_wt = -ha4 * k22 * (2.0 * L + (4.0 * de * de + w * w) * D) / 1296.0
        WT2 = WT2 + db_wt
        Ws = Ws + (-q2 * db_wt)

        Wc = Wc * tt * dm
        Vs = Vs * ss * dt
        VT = VT * sqsq * dt
        Vc = Vc * dm * dt
        WT2 = WT2 * sqsq * tt
        Ws = Ws * ss * tt
        vnlo = Wc + Vs + VT + Vc + WT2 + Ws

        # NLO contacts (:6600-6617); q2 here carries the q = 0 shift
        qxk_op = (-0.5j * _pauli(m1, m3, qxk) * (m2 == m4) + (-0.5j) * _pauli(m2, m4, qxk) * (m1 == m3))
        km_op = _pauli(m1, m3, kmean) * _pauli(m2, m4, kmean)
        c = cc.cnlo
        terms = [c[0] * q2 * dm * dt, c[1] * kmean2 * dm * dt, c[2] * q2 * ss * dt, c[3] * kmean2 * ss * dt,
                 c[4] * qxk_op * dt, c[5] * sqsq * dt, c[6] * km_op * dt]
        cont_nlo = np.zeros(p.shape, np.complex128)
        for t in terms:
            cont_nlo = cont_nlo + t

        if cc.order == NNLO:
            # subleading two-pion exchange and Delta triangles (:6620-6658)
            c1, c2, c3, c4 = cc.c1, cc.c2, cc.c3, cc.c4
            Vc = cc.k9 * (-(mpi2[2] * cc.k11 - q2 * c3) * wt2 * A)
            d_wt = -1.0 * cc.k16 * A * (c4 * w * w)
            WT3 = d_wt
            Ws = -q2 * d_wt
            Vc = Vc + (-ha2 * k22 * de * (6.0 * S * (4.0 * c1 * mpi2[2] - 2.0 * c2 * de * de
                                                      - c3 * (2.0 * de * de + S)) * D
                                           + (-24.0 * c1 * mpi2[2] + c2 * (w * w - 6.0 * S)
                                              + 6.0 * c3 * (2.0 * de * de + S)) * L) / 18.0)
            tr_wt = -c4 * ha2 * k22 * de * ((w * w - 4.0 * de * de) * D - 2.0 * L) / 72.0
            WT3 = WT3 + tr_wt
            Ws = Ws + (-q2 * tr_wt)
            vnnlo = Vc * dm * dt + WT3 * sqsq * tt + Ws * ss * tt

    lam = cc.cutoff
    reg = np.exp(-(p2 ** cc.nexp / lam ** (2.0 * cc.nexp) + pp2 ** cc.nexp / lam ** (2.0 * cc.nexp)))
    vdir = (cont_lo + vlo + cont_nlo + vnlo + vnnlo) * rel * reg
    v = HBARC ** 3 * vdir / lat.volume
    return np.where(ok, v, 0.0)


class ChiralInteraction:
    """Antisymmetrized <pq||rs> = v(p,q,r,s) - v(p,q,s,r) of the chiral NN, the
    combination every vnn_* block of setup_channel_structures stores
    (ccm_iter.f90:854, 917, 996, 1126, 1254)."""

    def __init__(self, lat, order=NNLO, cutoff=450.0, c1=None, c3=None):
        self.lat = lat
        self.cc = ChiralConstants(order, cutoff, c1, c3)

    def __call__(self, p, q, r, s):
        p, q, r, s = (np.asarray(x, np.int64) for x in (p, q, r, s))
        return chiral_direct(self.lat, self.cc, p, q, r, s) - chiral_direct(self.lat, self.cc, p, q, s, r)
