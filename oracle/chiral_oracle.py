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
    p, q, r, s = np.broadcast_arrays(*(np.asarray(x, np.int64) for x in (p, q, r, s)))
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


# --------------------------------------------------------------------------- #
# The NNLO three-nucleon force and its normal ordering
# --------------------------------------------------------------------------- #

F32_TWO_THIRDS = float(np.float32(2.0) / np.float32(3.0))   # the default-real (2./3.) of freg_3nfnonlocal


class ThreeBodyConstants:
    """The 3NF constants of init_chp_constants / init_cEcD
    (chiral_module_andreas_with_delta.f90:122-808, 1057-1066): cE, cD (zero
    below NNLO, ccm_main.f90:33-36), the 3NF c1, c3, c4 (those of the NN at
    NNLO, zero below) and the regulator cutoff and power of the deck's last
    line (freg_3nfnonlocal, :3400-3423)."""

    def __init__(self, cc: ChiralConstants, cE=0.017, cD=0.79, cutoff3=450.0, nexp3=3):
        if cc.order < NNLO:
            cE = cD = 0.0
            c1 = c3 = c4 = 0.0
        else:
            c1, c3, c4 = cc.c1, cc.c3, cc.c4
        fpi2 = FPI * FPI
        fpi4 = fpi2 * fpi2
        lam_chi = 700.0
        self.econst = cE / (fpi4 * lam_chi)
        self.dconst = cD / (fpi2 * lam_chi)
        self.k2 = GA * GA / (4.0 * fpi2)
        self.k21 = GA / (8.0 * fpi2)
        self.fpi2 = fpi2
        ha2 = HA * HA
        self.c1_bar = c1
        self.c3_bar = c3 - 4.0 * ha2 / (9.0 * cc.delta)
        self.c4_bar = c4 + 2.0 * ha2 / (9.0 * cc.delta)
        self.mpi2 = cc.mpi2[2]
        self.llambda = 4.0 * float(cutoff3) ** 2
        self.nexp = int(nexp3)


def _tau1_dot_tau_x_tau(t1, t2, t3, t4, t5, t6):
    """<t1|tau|t4> . (<t2|tau|t5> x <t3|tau|t6>) (chp_tau1_dot_tauXtau, :3544-3602)."""
    def comp(a, b, axis):
        up_up = (a == 1) & (b == 1)
        dn_dn = (a == -1) & (b == -1)
        up_dn = (a == 1) & (b == -1)
        dn_up = (a == -1) & (b == 1)
        if axis == 0:
            return np.where(up_dn | dn_up, 1.0 + 0j, 0j)
        if axis == 1:
            return np.where(up_dn, -1j, 0j) + np.where(dn_up, 1j, 0j)
        return np.where(up_up, 1.0 + 0j, 0j) + np.where(dn_dn, -1.0 + 0j, 0j)
    s2 = [comp(t2, t5, a) for a in range(3)]
    s3 = [comp(t3, t6, a) for a in range(3)]
    fx = s2[1] * s3[2] - s2[2] * s3[1]
    fy = s2[2] * s3[0] - s2[0] * s3[2]
    fz = s2[0] * s3[1] - s2[1] * s3[0]
    return fx * comp(t1, t4, 0) + fy * comp(t1, t4, 1) + fz * comp(t1, t4, 2)


def _freg3(c3: ThreeBodyConstants, k1, k2, k3):
    """freg_3nfnonlocal (:3400-3423), k in MeV."""
    p = 0.5 * (k1 - k2)
    q = F32_TWO_THIRDS * (k3 - 0.5 * (k1 + k2))
    q2 = 4.0 * np.sum(p * p, axis=-1) + 3.0 * np.sum(q * q, axis=-1)
    return np.exp(-((q2 / c3.llambda) ** c3.nexp))


def chiral3_direct(lat, c3: ThreeBodyConstants, p, q, r, s, t, u):
    """chiral_3nf_with_delta(p,q,r,s,t,u) (:6670-6902), complex, in MeV."""
    idx = [np.asarray(x, np.int64) for x in (p, q, r, s, t, u)]
    k = [lat.k[x] * HBARC for x in idx]
    m = [lat.sz[x] for x in idx]
    t_ = [lat.tz[x] for x in idx]
    ok = (t_[0] + t_[1] + t_[2] == t_[3] + t_[4] + t_[5]) & np.all(
        lat.n[idx[0]] + lat.n[idx[1]] + lat.n[idx[2]] == lat.n[idx[3]] + lat.n[idx[4]] + lat.n[idx[5]], axis=-1)
    q1, q2, q3 = k[3] - k[0], k[4] - k[1], k[5] - k[2]
    q1xq2, q1xq3, q2xq3 = np.cross(q1, q2), np.cross(q1, q3), np.cross(q2, q3)
    Q1, Q2, Q3 = (np.sum(x * x, axis=-1) for x in (q1, q2, q3))
    m1, m2, m3, m4, m5, m6 = m
    t1, t2, t3, t4, t5, t6 = t_
    sdq = lambda a, b, v: _pauli(a, b, v)
    q252, q142, q362 = sdq(m2, m5, q2), sdq(m1, m4, q2), sdq(m3, m6, q2)
    q251, q141, q361 = sdq(m2, m5, q1), sdq(m1, m4, q1), sdq(m3, m6, q1)
    q363, q143, q253 = sdq(m3, m6, q3), sdq(m1, m4, q3), sdq(m2, m5, q3)
    mp2 = c3.mpi2
    ope252, ope141, ope363 = q252 / (Q2 + mp2), q141 / (Q1 + mp2), q363 / (Q3 + mp2)
    s3_q1xq2, s2_q1xq3, s1_q2xq3 = sdq(m3, m6, q1xq2), sdq(m2, m5, q1xq3), sdq(m1, m4, q2xq3)
    d = lambda a, b: (a == b).astype(np.float64)
    k2, k21, fpi2 = c3.k2, c3.k21, c3.fpi2
    E, D = c3.econst, c3.dconst
    c1t = -c3.c1_bar * k2 * (4.0 * mp2 / fpi2)
    c3f = c3.c3_bar * k2 * (2.0 / fpi2)
    c4f = c3.c4_bar * k2 * (1.0 / fpi2) * 0.5
    sum1 = _pair_dot(t1, t2, t4, t5) * d(m3, m6) * d(t3, t6) * (
        E * d(m1, m4) * d(m2, m5)
        - D * k21 * (q252 * q142 / (Q2 + mp2) + q141 * q251 / (Q1 + mp2))
        + ope141 * ope252 * (c1t + c3f * np.sum(q1 * q2, axis=-1)))
    sum1 = sum1 + c4f * ope141 * ope252 * (s3_q1xq2 * (_tau1_dot_tau_x_tau(t3, t1, t2, t6, t4, t5)
                                                       - _tau1_dot_tau_x_tau(t3, t2, t1, t6, t5, t4)))
    sum2 = _pair_dot(t1, t3, t4, t6) * d(m2, m5) * d(t2, t5) * (
        E * d(m1, m4) * d(m3, m6)
        - D * k21 * (q363 * q143 / (Q3 + mp2) + q141 * q361 / (Q1 + mp2))
        + ope141 * ope363 * (c1t + c3f * np.sum(q1 * q3, axis=-1)))
    sum2 = sum2 + c4f * ope141 * ope363 * (s2_q1xq3 * (_tau1_dot_tau_x_tau(t2, t1, t3, t5, t4, t6)
                                                       - _tau1_dot_tau_x_tau(t2, t3, t1, t5, t6, t4)))
    sum3 = _pair_dot(t2, t3, t5, t6) * d(m1, m4) * d(t1, t4) * (
        E * d(m2, m5) * d(m3, m6)
        - D * k21 * (q363 * q253 / (Q3 + mp2) + q252 * q362 / (Q2 + mp2))
        + ope252 * ope363 * (c1t + c3f * np.sum(q2 * q3, axis=-1)))
    sum3 = sum3 + c4f * ope252 * ope363 * (s1_q2xq3 * (_tau1_dot_tau_x_tau(t1, t2, t3, t4, t5, t6)
                                                       - _tau1_dot_tau_x_tau(t1, t3, t2, t4, t6, t5)))
    v = sum1 + sum2 + sum3
    res = _freg3(c3, k[0], k[1], k[2]) * _freg3(c3, k[3], k[4], k[5]) * HBARC ** 6 * v / lat.volume ** 2
    return np.where(ok, res, 0.0)


def chiral3_asym(lat, c3, i, j, p, k, l, q):
    """chiral_3nf_with_delta_asym (:6904-6917): the ket antisymmetrized."""
    f = lambda a, b, c: chiral3_direct(lat, c3, i, j, p, a, b, c)
    return f(k, l, q) - f(l, k, q) - f(q, l, k) - f(k, q, l) + f(l, q, k) + f(q, k, l)


class NormalOrdered3NF:
    """The 3NF normal-ordered with respect to the hole states
    (normal_ordered_hamiltonian, ccm_main.f90:1687-2024; setup_3nf_2bNO,
    2025-2568): zero-body sum_{i<j<k} <ijk||ijk>, one-body
    f_pq += sum_{i<=j} <ijp||ijq>, two-body V_pqrs += sum_h <pqh||rsh>."""

    def __init__(self, lat, cc: ChiralConstants, cE=0.017, cD=0.79, cutoff3=450.0, nexp3=3):
        self.lat = lat
        self.c3 = ThreeBodyConstants(cc, cE, cD, cutoff3, nexp3)

    def e0(self):
        A = self.lat.A
        ijk = np.array([(i, j, k) for i in range(A) for j in range(i + 1, A) for k in range(j + 1, A)], np.int64)
        if len(ijk) == 0:
            return 0j
        i, j, k = ijk.T
        return np.sum(chiral3_asym(self.lat, self.c3, i, j, k, i, j, k))

    def fock(self, p, q):
        p, q = np.asarray(p, np.int64), np.asarray(q, np.int64)
        out = np.zeros(p.shape, np.complex128)
        A = self.lat.A
        for i in range(A):
            for j in range(i, A):
                ii, jj = np.full(p.shape, i), np.full(p.shape, j)
                out = out + chiral3_asym(self.lat, self.c3, ii, jj, p, ii, jj, q)
        return out

    def two_body(self, p, q, r, s):
        p, q, r, s = (np.asarray(x, np.int64) for x in (p, q, r, s))
        out = np.zeros(p.shape, np.complex128)
        for h in range(self.lat.A):
            hh = np.full(p.shape, h)
            out = out + chiral3_asym(self.lat, self.c3, p, q, hh, r, s, hh)
        return out

    def herm(self, p, q, r, s):
        """N(p,q,r,s) where the pair (p,q) comes first in lexicographic orbital
        order, else conj N(r,s,p,q): the Hermitian fill of the reference's
        hhhh, hphp and pppp blocks (ccm_main.f90:2046-2140, 2290-2540)."""
        p, q, r, s = (np.asarray(x, np.int64) for x in (p, q, r, s))
        first = (p < r) | ((p == r) & (q <= s))
        return np.where(first, self.two_body(p, q, r, s), np.conj(self.two_body(r, s, p, q)))

    def block(self, kind, p, q, r, s):
        """The normal-ordered 3NF part of a canonical record (p<q, r<s) of the
        reference's blocks, with its conventions: hhhh (1) and pppp (6) from
        the descending pair orientation with the Hermitian fill, hphp (4) with
        the Hermitian fill, hhpp (3) with the particle pair in the bra
        (ccm_main.f90:2160-2260), hppp <cb||kd> directly (ccm_iter.f90:
        1641-1690); hhhp (2) carries none."""
        p, q, r, s = (np.asarray(x, np.int64) for x in (p, q, r, s))
        if kind in (1, 6):
            return self.herm(q, p, s, r)
        if kind == 4:
            return self.herm(p, q, r, s)
        if kind == 3:
            return self.two_body(r, s, p, q)
        if kind == 5:
            return self.two_body(p, q, r, s)
        return np.zeros(p.shape, np.complex128)
