// Constants and tables of the built-in Delta-full chiral NN.  The functions
// restate init_chp_constants and the loop functions of
// chiral_module_andreas_with_delta.f90 (line numbers below) in the form the
// device evaluator (interaction.h) reads: radial functions of the momentum
// transfer on the integer lattice, one row per m = |dn|^2.
#include "chiral.h"

#include <cmath>
#include <stdexcept>
#include <string>

#include "interaction.h"

namespace ccgpu {
namespace {

const double kPi = 3.14159265358979;
// chiral_constants (:43-121)
const double kGA = 1.28899999999999992, kHA = 1.39999999999999991, kFpi = 92.2000000000000028;
const double kMnuc[3] = {938.272046000000046, 938.918267117927371, 939.565379000000007};   // pp, np, nn
const double kDeltaMass = 1232.0;
const double kMpiCharged = 139.570179999999993, kMpiNeutral = 134.976599999999991;
const double kSfr = 700.0;

struct LecSet {
  double cutoff;
  double lo_1s0[3], lo_3s1[3];        // LO contacts per pair charge, 10^4 GeV^-2
  double pw[7];                       // NLO contacts (1S0 3P0 1P1 3P1 3S1 3S1-3D1 3P2), 10^4 GeV^-4
};

// The Delta-full LECs of init_chp_constants (:172-530), per order.
const LecSet kLecLo[3] = {
    {500.0, {-1.08522470229349993e-01, -1.08522470229349993e-01, -1.08522470229349993e-01},
     {-6.84437111953200006e-02, -6.84437111953200006e-02, -6.84437111953200006e-02}, {0, 0, 0, 0, 0, 0, 0}},
    {450.0, {-1.12926682348719998e-01, -1.12926682348719998e-01, -1.12926682348719998e-01},
     {-8.73398209265599940e-02, -8.73398209265599940e-02, -8.73398209265599940e-02}, {0, 0, 0, 0, 0, 0, 0}},
    {394.0, {-1.16797069688160005e-01, -1.16797069688160005e-01, -1.16797069688160005e-01},
     {-1.03577257368989997e-01, -1.03577257368989997e-01, -1.03577257368989997e-01}, {0, 0, 0, 0, 0, 0, 0}},
};
const LecSet kLecNlo[3] = {
    {500.0, {-3.09617914433389985e-01, -3.10443063173440004e-01, -3.10256027522040012e-01},
     {-1.91012821791280013e-01, -1.91012821791280013e-01, -1.91012821791280013e-01},
     {2.39537506986395998, 7.27048937924560046e-01, 1.52465155576100003e-02, -9.51416990370669957e-01,
      5.39378033769969956e-01, 3.58443125233390003e-01, -7.93620896751420002e-01}},
    {450.0, {-3.09893184381559994e-01, -3.10711872064980021e-01, -3.10510995610929985e-01},
     {-1.97950619113270027e-01, -1.97950619113270027e-01, -1.97950619113270027e-01},
     {2.39163813513152990, 6.86901585796929948e-01, 4.81266565581999969e-03, -1.00011242015441004,
      5.58973150278149977e-01, 3.62094234568600015e-01, -8.08073350643100041e-01}},
    {394.0, {-3.09925223545309991e-01, -3.10751704075070023e-01, -3.10539966723270000e-01},
     {-2.05462289226700012e-01, -2.05462289226700012e-01, -2.05462289226700012e-01},
     {2.39117370641156990, 6.65731825802269972e-01, -1.52505123405899980e-02, -1.04638857865136004,
      5.88755038823760013e-01, 3.81315710481519965e-01, -8.18066377264660027e-01}},
};
const LecSet kLecNnlo[3] = {
    {500.0, {-3.37302636943159984e-01, -3.38320181277209975e-01, -3.38223065406040024e-01},
     {-2.21721491662439985e-01, -2.21721491662439985e-01, -2.21721491662439985e-01},
     {2.48801946794199980, 6.98453619479979970e-01, -1.26513321381799999e-02, -9.37263520317270116e-01,
      6.75353391585620022e-01, 3.54479450076249980e-01, -8.59526137679109903e-01}},
    {450.0, {-3.37136656357900022e-01, -3.38139465283629981e-01, -3.38023088490550017e-01},
     {-2.29310079440350006e-01, -2.29310079440350006e-01, -2.29310079440350006e-01},
     {2.47658908242146980, 6.45550411071979968e-01, -2.85410030715300030e-02, -1.02235931835912996,
      6.95953209842609999e-01, 3.58329844893869986e-01, -8.70203217397279905e-01}},
    {394.0, {-3.36845687915189984e-01, -3.37854531334379982e-01, -3.37722607113270012e-01},
     {-2.36959231730459990e-01, -2.36959231730459990e-01, -2.36959231730459990e-01},
     {2.46900829544757006, 6.16935689120689967e-01, -5.26929817733700068e-02, -1.09054771638107995,
      7.22641309492930040e-01, 3.75968068144939938e-01, -8.76265471802940010e-01}},
};
const double kCiNnlo[4] = {-7.39999999999999991e-01, -4.89999999999999991e-01, -6.50000000000000022e-01,
                           9.59999999999999964e-01};

double factorial(int n) {
  double f = 1.0;
  for (int i = 2; i <= n; ++i) f *= i;
  return f;
}

// Wigner 3j symbol with doubled arguments (Racah's formula); the reference's
// tjs (ccm_modules.f90:187-226) computes the same standard symbol.
double three_j(int j1, int j2, int j3, int m1, int m2, int m3) {
  if (m1 + m2 + m3 != 0) return 0.0;
  if (((j1 + m1) | (j2 + m2) | (j3 + m3)) & 1) return 0.0;
  if (j3 < std::abs(j1 - j2) || j3 > j1 + j2 || ((j1 + j2 + j3) & 1)) return 0.0;
  if (std::abs(m1) > j1 || std::abs(m2) > j2 || std::abs(m3) > j3) return 0.0;
  auto h = [](int x) { return x / 2; };
  const double tri = factorial(h(j1 + j2 - j3)) * factorial(h(j1 - j2 + j3)) * factorial(h(-j1 + j2 + j3)) /
                     factorial(h(j1 + j2 + j3) + 1);
  const double pre = std::sqrt(tri * factorial(h(j1 + m1)) * factorial(h(j1 - m1)) * factorial(h(j2 + m2)) *
                               factorial(h(j2 - m2)) * factorial(h(j3 + m3)) * factorial(h(j3 - m3)));
  double s = 0.0;
  for (int k = 0; k <= h(j1 + j2 - j3); ++k) {
    const int d[5] = {h(j1 + j2 - j3) - k, h(j1 - m1) - k, h(j2 + m2) - k, h(j3 - j2 + m1) + k, h(j3 - j1 - m2) + k};
    bool ok = true;
    double den = factorial(k);
    for (int x : d) {
      if (x < 0) ok = false;
      else den *= factorial(x);
    }
    if (ok) s += ((k & 1) ? -1.0 : 1.0) / den;
  }
  const int ph = h(j1 - j2 - m3);
  return ((ph & 1) ? -1.0 : 1.0) * pre * s;
}

// n-point Gauss-Legendre nodes (ascending) and weights on [x1, x2]: Newton's
// method on the Legendre recurrence, as chp_setup_gauleg_mesh (:824-880).
void gauss_legendre(int n, double x1, double x2, double *x, double *w) {
  const int m = (n + 1) / 2;
  const double xm = 0.5 * (x2 + x1), xl = 0.5 * (x2 - x1);
  for (int i = 1; i <= m; ++i) {
    // at least one Newton step (for odd n the middle start cos(pi/2) already
    // passes the stopping test; the mesh of chiral_radial_table replaces
    // that node with the reference's literals anyway)
    double z = std::cos(kPi * (i - 0.25) / (n + 0.5)), z1, pp;
    do {
      double p1 = 1.0, p2 = 0.0;
      for (int j = 1; j <= n; ++j) {
        const double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      z1 = z;
      z = z1 - p1 / pp;
    } while (std::fabs(z - z1) > 3.0e-14);
    x[i - 1] = xm - xl * z;
    x[n - i] = xm + xl * z;
    w[i - 1] = 2.0 * xl / ((1.0 - z * z) * pp * pp);
    w[n - i] = w[i - 1];
  }
}

const LecSet &lec_set(int order, double cutoff) {
  const LecSet *tab = order == 0 ? kLecLo : order == 2 ? kLecNlo : kLecNnlo;
  for (int i = 0; i < 3; ++i)
    if (std::fabs(tab[i].cutoff - cutoff) < 1e-5) return tab[i];
  throw std::runtime_error("chiral cutoff " + std::to_string(cutoff) + " MeV: the Delta-full family has 394, 450, 500");
}

}  // namespace

ChiralModel chiral_model(int order, double cutoff, const double *ci) {
  if (order != 0 && order != 2 && order != 3)
    throw std::runtime_error("chiral order " + std::to_string(order) + ": LO=0, NLO=2, NNLO=3");
  const LecSet &L = lec_set(order, cutoff);
  ChiralModel c;
  c.order = order;
  c.cutoff = cutoff;
  c.nexp = std::fabs(cutoff - 394.0) < 1e-5 ? 4.0 : 3.0;
  double lec[4] = {0, 0, 0, 0};
  if (order == 3)
    for (int i = 0; i < 4; ++i) lec[i] = kCiNnlo[i];
  if (ci)
    for (int i = 0; i < 4; ++i) lec[i] = ci[i];
  c.c1 = lec[0] * 1.0e-3;
  c.c2 = lec[1] * 1.0e-3;
  c.c3 = lec[2] * 1.0e-3;
  c.c4 = lec[3] * 1.0e-3;
  // LO contacts in the spin-isospin basis (:599-600)
  for (int i = 0; i < 3; ++i) {
    const double a = L.lo_1s0[i] * 0.01, b = L.lo_3s1[i] * 0.01;
    c.cs[i] = (a + 3.0 * b) / 16.0 / kPi;
    c.ct[i] = (b - a) / 16.0 / kPi;
  }
  // NLO contacts from partial waves (:604-614)
  double pw[7];
  for (int i = 0; i < 7; ++i) pw[i] = L.pw[i] * 1.0e-08;
  const double s2 = std::sqrt(2.0);
  c.cnlo[0] = (-5.0 * pw[6] + 6.0 * pw[4] - 3.0 * pw[3] - 3.0 * pw[2] - pw[1] + 2.0 * pw[0]) / (64.0 * kPi);
  c.cnlo[1] = (5.0 * pw[6] + 6.0 * pw[4] + 3.0 * pw[3] + 3.0 * pw[2] + pw[1] + 2.0 * pw[0]) / (16.0 * kPi);
  c.cnlo[2] = -(2.0 * pw[6] - 2.0 * s2 * pw[5] - 2.0 * pw[4] - 3.0 * pw[2] + pw[1] + 2.0 * pw[0]) / (64.0 * kPi);
  c.cnlo[3] = -(-2.0 * pw[6] - 2.0 * s2 * pw[5] - 2.0 * pw[4] + 3.0 * pw[2] - pw[1] + 2.0 * pw[0]) / (16.0 * kPi);
  c.cnlo[4] = -(-5.0 * pw[6] + 3.0 * pw[3] + 2.0 * pw[1]) / (16.0 * kPi);
  c.cnlo[5] = (pw[6] - 6.0 * s2 * pw[5] - 3.0 * pw[3] + 2.0 * pw[1]) / (64.0 * kPi);
  c.cnlo[6] = -(pw[6] + 6.0 * s2 * pw[5] - 3.0 * pw[3] + 2.0 * pw[1]) / (16.0 * kPi);
  // <1/2 t, 1/2 -t | T 0> = 3j * sqrt(2T+1) for t = +1, -1 (the CIB one-pion exchange, :6534-6551)
  for (int i = 0; i < 2; ++i) {
    const int t = 1 - 2 * i;
    c.cg0[i] = three_j(1, 1, 0, t, -t, 0) * std::sqrt(1.0);
    c.cg1[i] = three_j(1, 1, 2, t, -t, 0) * std::sqrt(3.0);
  }
  return c;
}

std::vector<double> chiral_radial_table(const ChiralModel &c, double qunit, int mmax) {
  const double ga2 = kGA * kGA, ga4 = ga2 * ga2, fpi2 = kFpi * kFpi, fpi4 = fpi2 * fpi2, pi2 = kPi * kPi;
  const double ha2 = kHA * kHA, ha4 = ha2 * ha2;
  // const(2..24) of init_chp_constants (:740-784)
  const double k2 = ga2 / (4.0 * fpi2), k3 = 384.0 * pi2 * fpi4, k4 = 5.0 * ga4 - 4.0 * ga2 - 1.0;
  const double k5 = 23.0 * ga4 - 10.0 * ga2 - 1.0, k6 = 48.0 * ga4, k7 = 3.0 * ga4, k8 = 64.0 * pi2 * fpi4;
  const double k9 = 3.0 * ga2 / (16.0 * kPi * fpi4), k11 = 2.0 * (2.0 * c.c1 - c.c3);
  const double k16 = ga2 / (32.0 * kPi * fpi4), k22 = 1.0 / (pi2 * fpi4), k24 = 1.0 / (kPi * fpi4);
  const double mpi_avg = (kMpiCharged + kMpiNeutral + kMpiCharged) / 3.0;
  const double mpi2[3] = {kMpiNeutral * kMpiNeutral, kMpiCharged * kMpiCharged, mpi_avg * mpi_avg};
  const double de = kDeltaMass - kMnuc[1];
  // 25-point mu mesh on [2 m_pi, Lambda_SFR] with its 13th point replaced by
  // default-real (float32) literals, as chp_setup_delta_loop_int_data (:2606-2620)
  double mu[25], wmu[25];
  gauss_legendre(25, 2.0 * mpi_avg, kSfr, mu, wmu);
  mu[12] = (double)488.038986666667f;
  wmu[12] = (double)25.5194921127220f;

  const double fm = 4.0 * mpi2[2], sfr2 = kSfr * kSfr, twompi = 2.0 * mpi_avg;
  const double sl = std::sqrt(sfr2 - fm);
  // H's q-independent part (:2656-2680)
  const double qdel = 2.0 * std::sqrt(de * de - mpi2[2]), qdel2 = qdel * qdel;
  const double wdel = std::sqrt(qdel2 + fm), sdel = std::sqrt(sfr2 - fm);
  const double delL = wdel * std::log((sfr2 * wdel * wdel + qdel2 * sdel * sdel + 2.0 * kSfr * qdel * wdel * sdel) /
                                      (fm * (sfr2 + qdel2))) / (2.0 * qdel);

  std::vector<double> tab((size_t)(mmax + 1) * kChiralRad, 0.0);
  for (int m = 0; m <= mmax; ++m) {
    double *R = &tab[(size_t)m * kChiralRad];
    double q2 = qunit * qunit * m;
    double q = std::sqrt(q2);
    if (c.order == 0) {
      R[kOpe2] = -1.0 * k2 / (q2 + mpi2[2]);
      R[kQ2] = q2;
      continue;
    }
    const double w = std::sqrt(fm + q2);
    // q = 0: the reference's loop functions shift q to 1e-8 for the rest of the element (:1244-1312)
    if (q == 0.0) {
      q = 1.0e-8;
      q2 = q * q;
    }
    const double L = w * std::log((sfr2 * w * w + q2 * sl * sl + 2.0 * kSfr * q * w * sl) / (fm * (sfr2 + q2))) /
                     (2.0 * q);
    const double wt2 = 2.0 * mpi2[2] + q2;
    const double A = std::atan(q * (kSfr - twompi) / (q2 + kSfr * twompi)) / (2.0 * q);
    double D = 0.0;
    for (int j = 0; j < 25; ++j) {
      const double mut = std::sqrt(mu[j] * mu[j] - fm) / (2.0 * de);
      D += wmu[j] * (std::atan(mut) / (mu[j] * mu[j] + q2));
    }
    D /= de;
    const double S = 2.0 * mpi2[2] + q2 - 2.0 * de * de;
    const double H = 2.0 * S * (L - delL) / (w * w - 4.0 * de * de);

    R[kOpe0] = -1.0 * k2 / (q2 + mpi2[0]);
    R[kOpe1] = -1.0 * k2 / (q2 + mpi2[1]);
    // NLO 2PE (:1127-1168) and the leading Delta diagrams (:2683-2870)
    double Wc = -L * (fm * k4 + q2 * k5 + k6 * mpi2[2] * mpi2[2] / (w * w)) / k3;
    double Vs = k7 * L * q2 / k8;
    double Vt = -k7 * L / k8;
    Wc += -ha2 * k22 * ((6.0 * S - w * w) * L + 12.0 * de * de * S * D) / 216.0;
    Wc += -ga2 * ha2 * k22 * ((12.0 * de * de - 20.0 * mpi2[2] - 11.0 * q2) * L + 6.0 * S * S * D) / 216.0;
    const double sb_vt = -ga2 * ha2 * k22 * (-2.0 * L + (w * w - 4.0 * de * de) * D) / 48.0;
    Vs += -q2 * sb_vt;
    Vt += sb_vt;
    double Vc = -ga2 * ha2 * k24 * (2.0 * mpi2[2] + q2) * (2.0 * mpi2[2] + q2) * A / (12.0 * de);
    const double sb_wt = -ga2 * ha2 * k24 * (w * w * A) / (144.0 * de);
    double Wt = sb_wt;
    double Ws = -q2 * sb_wt;
    Wc += -ha4 * k22 * ((12.0 * S - w * w) * L + 3.0 * S * (H + (8.0 * de * de - S) * D)) / 486.0;
    const double db_vt = -ha4 * k22 * (6.0 * L + (12.0 * de * de - w * w) * D) / 216.0;
    Vs += -q2 * db_vt;
    Vt += db_vt;
    Vc += -ha4 * k22 * (-4.0 * de * de * L + S * (H + (S + 8.0 * de * de) * D)) / 27.0;
    const double db_wt = -ha4 * k22 * (2.0 * L + (4.0 * de * de + w * w) * D) / 1296.0;
    Wt += db_wt;
    Ws += -q2 * db_wt;
    R[kWc] = Wc;
    R[kVs] = Vs;
    R[kVt] = Vt;
    R[kVc] = Vc;
    R[kWt] = Wt;
    R[kWs] = Ws;
    R[kQ2] = q2;
    if (c.order == 3) {
      // NNLO 2PE (:1341-1515) and the subleading Delta triangles (:2872-2925)
      double Vc3 = k9 * (-(mpi2[2] * k11 - q2 * c.c3) * wt2 * A);
      const double d_wt = -1.0 * k16 * A * (c.c4 * w * w);
      double Wt3 = d_wt;
      double Ws3 = -q2 * d_wt;
      Vc3 += -ha2 * k22 * de *
             (6.0 * S * (4.0 * c.c1 * mpi2[2] - 2.0 * c.c2 * de * de - c.c3 * (2.0 * de * de + S)) * D +
              (-24.0 * c.c1 * mpi2[2] + c.c2 * (w * w - 6.0 * S) + 6.0 * c.c3 * (2.0 * de * de + S)) * L) /
             18.0;
      const double tr_wt = -c.c4 * ha2 * k22 * de * ((w * w - 4.0 * de * de) * D - 2.0 * L) / 72.0;
      Wt3 += tr_wt;
      Ws3 += -q2 * tr_wt;
      R[kVc3] = Vc3;
      R[kWt3] = Wt3;
      R[kWs3] = Ws3;
    }
  }
  return tab;
}

std::vector<double> chiral_side_table(const ChiralModel &c, double qunit, int jmax) {
  std::vector<double> tab((size_t)3 * (jmax + 1));
  const double lam2n = std::pow(c.cutoff, 2.0 * c.nexp);
  for (int ch = 0; ch < 3; ++ch) {
    const double M = kMnuc[ch];
    for (int j = 0; j <= jmax; ++j) {
      const double p2 = qunit * qunit * j;
      // relativity_factor = M / sqrt(sqrt((M^2 + p^2)(M^2 + p'^2))) and
      // freg_delta = exp(-(p^2n + p'^2n) / Lambda^2n) (:6441-6442, 3358-3378), one side each
      tab[(size_t)ch * (jmax + 1) + j] =
          std::sqrt(M / std::sqrt(M * M + p2)) * std::exp(-std::pow(p2, c.nexp) / lam2n);
    }
  }
  return tab;
}

Chiral3Dev chiral3_constants(const ChiralModel &c, double cE, double cD, double cutoff3, int nexp3) {
  const double fpi2 = kFpi * kFpi, fpi4 = fpi2 * fpi2, lam_chi = 700.0;
  const double ga2 = kGA * kGA, ha2 = kHA * kHA;
  const double mpi_avg = (kMpiCharged + kMpiNeutral + kMpiCharged) / 3.0;
  const double de = kDeltaMass - kMnuc[1];
  const bool nnlo = c.order >= 3;
  // init_cEcD (:1057-1066); the 3NF c_i follow the NN ones at NNLO (:270-290)
  const double ce = nnlo ? cE : 0.0, cd = nnlo ? cD : 0.0;
  const double c1 = nnlo ? c.c1 : 0.0, c3 = nnlo ? c.c3 : 0.0, c4 = nnlo ? c.c4 : 0.0;
  const double k2 = ga2 / (4.0 * fpi2);
  Chiral3Dev t{};
  t.econst = ce / (fpi4 * lam_chi);
  t.dconst = cd / (fpi2 * lam_chi);
  t.k21 = kGA / (8.0 * fpi2);
  const double c1_bar = c1, c3_bar = c3 - 4.0 * ha2 / (9.0 * de), c4_bar = c4 + 2.0 * ha2 / (9.0 * de);
  t.c1t = -c1_bar * k2 * (4.0 * (mpi_avg * mpi_avg) / fpi2);
  t.c3f = c3_bar * k2 * (2.0 / fpi2);
  t.c4f = c4_bar * k2 * (1.0 / fpi2) * 0.5;
  t.mpi2 = mpi_avg * mpi_avg;
  t.llambda = 4.0 * cutoff3 * cutoff3;
  t.two_thirds = (double)(2.0f / 3.0f);
  t.nexp = nexp3;
  return t;
}

}  // namespace ccgpu
