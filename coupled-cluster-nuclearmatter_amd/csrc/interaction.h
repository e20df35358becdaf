// The built-in interaction generator shared by the block-fill kernels and the
// (T) kernel: the Minnesota potential in momentum space.
#pragma once
#include <hip/hip_runtime.h>

namespace ccgpu {

// ---------------------------------------------------------------------------
// Interaction: Minnesota potential in momentum space, vmom_minnesota
// (chiral_module_andreas_with_delta.f90:8980-9111).  The reference's
// default-real literals are reproduced as their float32 values.
// ---------------------------------------------------------------------------
struct Orb {
  const double *kx, *ky, *kz;
  const int *sz, *tz;
  double volume;
};

static __device__ __forceinline__ double sig_dot_sig(int m1, int m2, int m3, int m4) {
  // <m1 m2| s1.s2 |m3 m4> for m = +-1: m1*m2 on the diagonal, 2 for a spin exchange
  double r = 0.0;
  if (m1 == m3 && m2 == m4) r += (double)(m1 * m2);
  if (m1 == m4 && m2 == m3 && m1 != m2) r += 2.0;
  return r;
}

static __device__ __forceinline__ double minnesota_direct(const Orb &o, int p, int q, int r, int s) {
  const double PI = 3.14159265358979;
  const double v0r = 200.0, v0t = 178.0, v0s = (double)91.85f;
  const double kr = (double)1.487f, kt = (double)0.639f, ks = (double)0.465f;
  double qx = 0.5 * (o.kx[p] - o.kx[q]) - 0.5 * (o.kx[r] - o.kx[s]);
  double qy = 0.5 * (o.ky[p] - o.ky[q]) - 0.5 * (o.ky[r] - o.ky[s]);
  double qz = 0.5 * (o.kz[p] - o.kz[q]) - 0.5 * (o.kz[r] - o.kz[s]);
  double q2 = qx * qx + qy * qy + qz * qz;
  double pi15 = pow(PI, 1.5);
  double vr = v0r * pi15 * exp(-q2 / (4.0 * kr)) / pow(kr, 1.5);
  double vt = -v0t * pi15 * exp(-q2 / (4.0 * kt)) / pow(kt, 1.5);
  double vs = -v0s * pi15 * exp(-q2 / (4.0 * ks)) / pow(ks, 1.5);
  int m1 = o.sz[p], m2 = o.sz[q], m3 = o.sz[r], m4 = o.sz[s];
  int t1 = o.tz[p], t2 = o.tz[q], t3 = o.tz[r], t4 = o.tz[s];
  if (t1 + t2 != t3 + t4) return 0.0;
  double ds = (m1 == m3 && m2 == m4) ? 1.0 : 0.0;
  double dt = (t1 == t3 && t2 == t4) ? 1.0 : 0.0;
  double ss = sig_dot_sig(m1, m2, m3, m4);
  double tt = sig_dot_sig(t1, t2, t3, t4);
  vr = vr * (3.0 * ds * dt - tt * ds - ss * dt - tt * ss) / 8.0;
  vs = vs * (3.0 * ds * dt + tt * ds - 3.0 * ss * dt - tt * ss) / 16.0;
  vt = vt * (3.0 * ds * dt - 3.0 * tt * ds + ss * dt - tt * ss) / 16.0;
  return (vs + vr + vt) / o.volume;
}

// The same element with the radial factors read from a table over the integer
// momentum transfer m = |n_p - n_q - n_r + n_s|^2 (q = (pi/L) dn for every
// twist): rad[0..2] = v0 pi^1.5 exp(-q^2/(4 kappa)) / kappa^1.5 for the R, T
// and S Gaussians (the last two carry their minus sign).
static __device__ __forceinline__ double minnesota_direct_tab(const double *rad, int m1, int m2, int m3, int m4,
                                                              int t1, int t2, int t3, int t4, double volume) {
  if (t1 + t2 != t3 + t4) return 0.0;
  const double ds = (m1 == m3 && m2 == m4) ? 1.0 : 0.0;
  const double dt = (t1 == t3 && t2 == t4) ? 1.0 : 0.0;
  const double ss = sig_dot_sig(m1, m2, m3, m4);
  const double tt = sig_dot_sig(t1, t2, t3, t4);
  const double vr = rad[0] * (3.0 * ds * dt - tt * ds - ss * dt - tt * ss) / 8.0;
  const double vs = rad[2] * (3.0 * ds * dt + tt * ds - 3.0 * ss * dt - tt * ss) / 16.0;
  const double vt = rad[1] * (3.0 * ds * dt - 3.0 * tt * ds + ss * dt - tt * ss) / 16.0;
  return (vs + vr + vt) / volume;
}

// ---------------------------------------------------------------------------
// Interaction: the Delta-full chiral NN, chiral_NN_with_delta
// (chiral_module_andreas_with_delta.f90:6304-6668).  Every radial function
// depends on the momentum transfer only, q = (pi/L) dn with the integer
// dn = (n_p - n_q) - (n_r - n_s) (twists cancel), so the host tabulates them
// over m = |dn|^2; the regulator and the relativity factor split into one
// factor per side, tabulated over j = |n_p - n_q|^2.  The element is then a
// few table reads and the spin-isospin algebra.
// ---------------------------------------------------------------------------
constexpr int kChiralRad = 13;
enum ChiralSlot {
  kOpe0 = 0,  // one-pion exchange, neutral pion        (:1110-1124)
  kOpe1,      // charged pion
  kOpe2,      // average pion mass (LO)
  kWc,        // NLO 2PE + leading Delta: W_C, V_S, V_T, V_C, W_T, W_S (:6553-6598)
  kVs,
  kVt,
  kVc,
  kWt,
  kWs,
  kVc3,       // NNLO 2PE + Delta triangles: V_C, W_T, W_S (:6620-6658)
  kWt3,
  kWs3,
  kQ2         // q^2 as the contacts see it (1e-16 at q = 0 from NLO on)
};

struct ChiralDev {
  const double *rad;      // kChiralRad per m
  const double *side;     // 3 pair charges (pp, np, nn) x (jmax + 1)
  int jmax, order;
  double qunit;           // pi hbarc / L, MeV per unit of dn
  double scale;           // hbarc^3 / volume
  double cs[3], ct[3];    // LO contacts per pair charge
  double cnlo[7];         // NLO contacts
  double cg0[2], cg1[2];  // np isospin CG for T = 0, 1, index (1 - tz) / 2
};

// <mo| sigma.v |mi> for spin projections +-1
static __device__ __forceinline__ double2 pauli_dot(int mo, int mi, double vx, double vy, double vz) {
  if (mo == mi) return make_double2(mo > 0 ? vz : -vz, 0.0);
  return mo > 0 ? make_double2(vx, -vy) : make_double2(vx, vy);
}

static __device__ __forceinline__ double2 cmul_i(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// v(p,q,r,s) of the chiral NN in MeV, complex (the spin-orbit and tensor
// operators carry sigma_y); zero unless momentum and isospin are conserved.
static __device__ __forceinline__ double2 chiral_direct(const ChiralDev &c, const int *nx, const int *ny,
                                                       const int *nz, const int *sz, const int *tz, int p, int q,
                                                       int r, int s) {
  const int t1 = tz[p], t2 = tz[q], t3 = tz[r], t4 = tz[s];
  const int ax = nx[p] - nx[q], ay = ny[p] - ny[q], az = nz[p] - nz[q];
  const int bx = nx[r] - nx[s], by = ny[r] - ny[s], bz = nz[r] - nz[s];
  if (t1 + t2 != t3 + t4 || nx[p] + nx[q] != nx[r] + nx[s] || ny[p] + ny[q] != ny[r] + ny[s] ||
      nz[p] + nz[q] != nz[r] + nz[s])
    return make_double2(0.0, 0.0);
  const int dx = ax - bx, dy = ay - by, dz = az - bz;
  const double *R = c.rad + kChiralRad * (dx * dx + dy * dy + dz * dz);
  const int ch = (t1 + t2) / 2 + 1;
  const double f = c.side[ch * (c.jmax + 1) + ax * ax + ay * ay + az * az] *
                   c.side[ch * (c.jmax + 1) + bx * bx + by * by + bz * bz];
  const int m1 = sz[p], m2 = sz[q], m3 = sz[r], m4 = sz[s];
  const double dm = (m1 == m3 && m2 == m4) ? 1.0 : 0.0;
  const double dt = (t1 == t3 && t2 == t4) ? 1.0 : 0.0;
  const double ss = sig_dot_sig(m1, m2, m3, m4);
  const double tt = sig_dot_sig(t1, t2, t3, t4);
  const double u = c.qunit;
  const double qx = u * dx, qy = u * dy, qz = u * dz;
  const double2 sq = cmul_i(pauli_dot(m1, m3, qx, qy, qz), pauli_dot(m2, m4, qx, qy, qz));
  double2 v;
  if (c.order == 0) {
    // LO: (sigma1.q)(sigma2.q) OPE tau.tau + CS + CT sigma.sigma (np contacts)
    const double a = R[kOpe2] * tt;
    v = make_double2(sq.x * a + c.cs[1] * dm * dt + c.ct[1] * ss * dt, sq.y * a);
  } else {
    double wt1;
    if (t1 + t2 == 0) {
      const double a = c.cg0[(1 - t1) / 2] * c.cg0[(1 - t3) / 2], b = c.cg1[(1 - t1) / 2] * c.cg1[(1 - t3) / 2];
      wt1 = a * (-R[kOpe0] - 2.0 * R[kOpe1]) + b * (-R[kOpe0] + 2.0 * R[kOpe1]);
    } else {
      wt1 = R[kOpe0];
    }
    const double q2 = R[kQ2];
    const double kx = 0.5 * u * (ax + bx), ky = 0.5 * u * (ay + by), kz = 0.5 * u * (az + bz);
    const double k2 = kx * kx + ky * ky + kz * kz;
    // coefficients of the operator structures
    const double a_sq = wt1 + R[kVt] * dt + R[kWt] * tt + c.cnlo[5] * dt + R[kWt3] * tt;
    const double a_1 = c.cs[ch] * dm * dt + R[kVc] * dm * dt + R[kWc] * tt * dm +
                       (c.cnlo[0] * q2 + c.cnlo[1] * k2) * dm * dt + R[kVc3] * dm * dt;
    const double a_ss = (c.ct[ch] + R[kVs] + c.cnlo[2] * q2 + c.cnlo[3] * k2) * dt + (R[kWs] + R[kWs3]) * tt;
    v = make_double2(a_sq * sq.x + a_1 + a_ss * ss, a_sq * sq.y);
    if (dt != 0.0) {
      // C5 i/2 (sigma1 + sigma2).(k x q) and C7 (sigma1.k)(sigma2.k)
      const double wx = qy * kz - qz * ky, wy = qz * kx - qx * kz, wz = qx * ky - qy * kx;
      double2 so = make_double2(0.0, 0.0);
      if (m2 == m4) {
        const double2 e = pauli_dot(m1, m3, wx, wy, wz);
        so.x += e.x; so.y += e.y;
      }
      if (m1 == m3) {
        const double2 e = pauli_dot(m2, m4, wx, wy, wz);
        so.x += e.x; so.y += e.y;
      }
      // -i/2 so
      v.x += c.cnlo[4] * 0.5 * so.y;
      v.y -= c.cnlo[4] * 0.5 * so.x;
      const double2 kk = cmul_i(pauli_dot(m1, m3, kx, ky, kz), pauli_dot(m2, m4, kx, ky, kz));
      v.x += c.cnlo[6] * kk.x;
      v.y += c.cnlo[6] * kk.y;
    }
  }
  const double g = c.scale * f;
  return make_double2(g * v.x, g * v.y);
}

// ---------------------------------------------------------------------------
// The NNLO three-nucleon force, chiral_3nf_with_delta
// (chiral_module_andreas_with_delta.f90:6670-6902): cD and cE contacts,
// two-pion exchange with c1, c3, c4 shifted by the Delta (c3 - 4hA^2/9Delta,
// c4 + 2hA^2/9Delta), summed over the three choices of the middle nucleon, with
// the nonlocal regulator freg_3nfnonlocal (:3400-3423) on both sides.  The
// reference's default-real (2./3.) in the regulator is kept (float32 value).
// ---------------------------------------------------------------------------
struct Chiral3Dev {
  double econst, dconst;     // cE / (fpi^4 Lambda_chi), cD / (fpi^2 Lambda_chi)
  double k21;                // gA / (8 fpi^2)
  double c1t, c3f, c4f;      // -c1 gA^2/(4fpi^2) 4mpi^2/fpi^2, c3 gA^2/(4fpi^2) 2/fpi^2, c4 gA^2/(4fpi^2) 1/(2fpi^2)
  double mpi2;               // average pion mass squared
  double llambda;            // 4 cutoff^2
  double two_thirds;         // float32(2/3)
  double scale;              // hbarc^6 / volume^2
  int nexp;                  // regulator power
  const double *kx, *ky, *kz;   // single-particle momenta, fm^-1
  const int *nx, *ny, *nz, *sz, *tz;
};

static __device__ __forceinline__ double freg3(const Chiral3Dev &c, const double *k1, const double *k2,
                                               const double *k3) {
  double q2 = 0.0;
  double pp = 0.0, qq = 0.0;
  for (int x = 0; x < 3; ++x) {
    const double p = 0.5 * (k1[x] - k2[x]);
    const double q = c.two_thirds * (k3[x] - 0.5 * (k1[x] + k2[x]));
    pp += p * p;
    qq += q * q;
  }
  q2 = 4.0 * pp + 3.0 * qq;
  const double r = q2 / c.llambda;
  double e = 1.0;
  for (int i = 0; i < c.nexp; ++i) e *= r;
  return exp(-e);
}

// <t1|tau|t4> . (<t2|tau|t5> x <t3|tau|t6>)  (chp_tau1_dot_tauXtau, :3544-3602)
static __device__ __forceinline__ double2 tau_triple(int t1, int t2, int t3, int t4, int t5, int t6) {
  // Cartesian components of <a|tau|b>: x, y (complex), z
  auto comp = [](int a, int b, int axis) -> double2 {
    if (axis == 0) return make_double2(a != b ? 1.0 : 0.0, 0.0);
    if (axis == 1) return a != b ? make_double2(0.0, a > 0 ? -1.0 : 1.0) : make_double2(0.0, 0.0);
    return make_double2(a == b ? (double)a : 0.0, 0.0);
  };
  double2 s2[3], s3[3], s1[3];
  for (int a = 0; a < 3; ++a) {
    s2[a] = comp(t2, t5, a);
    s3[a] = comp(t3, t6, a);
    s1[a] = comp(t1, t4, a);
  }
  auto sub = [](double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); };
  const double2 fx = sub(cmul_i(s2[1], s3[2]), cmul_i(s2[2], s3[1]));
  const double2 fy = sub(cmul_i(s2[2], s3[0]), cmul_i(s2[0], s3[2]));
  const double2 fz = sub(cmul_i(s2[0], s3[1]), cmul_i(s2[1], s3[0]));
  const double2 a = cmul_i(fx, s1[0]), b = cmul_i(fy, s1[1]), e = cmul_i(fz, s1[2]);
  return make_double2(a.x + b.x + e.x, a.y + b.y + e.y);
}

// v3(p,q,r,s,t,u) in MeV; the bra regulator is passed in (it is the same for
// every ket permutation of the antisymmetrized element)
static __device__ __forceinline__ double2 chiral3_direct(const Chiral3Dev &c, double freg_bra, int p, int q, int r,
                                                         int s, int t, int u) {
  const int id[6] = {p, q, r, s, t, u};
  if (c.tz[p] + c.tz[q] + c.tz[r] != c.tz[s] + c.tz[t] + c.tz[u] ||
      c.nx[p] + c.nx[q] + c.nx[r] != c.nx[s] + c.nx[t] + c.nx[u] ||
      c.ny[p] + c.ny[q] + c.ny[r] != c.ny[s] + c.ny[t] + c.ny[u] ||
      c.nz[p] + c.nz[q] + c.nz[r] != c.nz[s] + c.nz[t] + c.nz[u])
    return make_double2(0.0, 0.0);
  const double hbarc = 197.326968;
  double k[6][3];
  for (int a = 0; a < 6; ++a) {
    k[a][0] = c.kx[id[a]] * hbarc;
    k[a][1] = c.ky[id[a]] * hbarc;
    k[a][2] = c.kz[id[a]] * hbarc;
  }
  double q1[3], q2[3], q3[3];
  for (int x = 0; x < 3; ++x) {
    q1[x] = k[3][x] - k[0][x];
    q2[x] = k[4][x] - k[1][x];
    q3[x] = k[5][x] - k[2][x];
  }
  auto cross = [](const double *a, const double *b, double *o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
  };
  double q1xq2[3], q1xq3[3], q2xq3[3];
  cross(q1, q2, q1xq2);
  cross(q1, q3, q1xq3);
  cross(q2, q3, q2xq3);
  const double Q1 = q1[0] * q1[0] + q1[1] * q1[1] + q1[2] * q1[2];
  const double Q2 = q2[0] * q2[0] + q2[1] * q2[1] + q2[2] * q2[2];
  const double Q3 = q3[0] * q3[0] + q3[1] * q3[1] + q3[2] * q3[2];
  const int m1 = c.sz[p], m2 = c.sz[q], m3 = c.sz[r], m4 = c.sz[s], m5 = c.sz[t], m6 = c.sz[u];
  const int t1 = c.tz[p], t2 = c.tz[q], t3 = c.tz[r], t4 = c.tz[s], t5 = c.tz[t], t6 = c.tz[u];
  auto sdq = [](int a, int b, const double *v) { return pauli_dot(a, b, v[0], v[1], v[2]); };
  const double2 q252 = sdq(m2, m5, q2), q142 = sdq(m1, m4, q2), q362 = sdq(m3, m6, q2);
  const double2 q251 = sdq(m2, m5, q1), q141 = sdq(m1, m4, q1), q361 = sdq(m3, m6, q1);
  const double2 q363 = sdq(m3, m6, q3), q143 = sdq(m1, m4, q3), q253 = sdq(m2, m5, q3);
  const double w1 = 1.0 / (Q1 + c.mpi2), w2 = 1.0 / (Q2 + c.mpi2), w3 = 1.0 / (Q3 + c.mpi2);
  auto scl = [](double a, double2 b) { return make_double2(a * b.x, a * b.y); };
  auto add = [](double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); };
  auto dl = [](int a, int b) { return a == b ? 1.0 : 0.0; };
  const double2 ope141 = scl(w1, q141), ope252 = scl(w2, q252), ope363 = scl(w3, q363);
  // one "middle nucleon" term; (x,y) the two pion-coupled nucleons
  auto term = [&](double ttf, double dspec, double e_d, double2 qa_b, double2 qc_d, double wab, double2 qe_f,
                  double2 qg_h, double wef, double2 opex, double2 opey, double qdot, double2 sig_qxq, double2 tau_a,
                  double2 tau_b) {
    double2 in = make_double2(c.econst * e_d, 0.0);
    const double2 dterm = add(scl(wab, cmul_i(qa_b, qc_d)), scl(wef, cmul_i(qe_f, qg_h)));
    in = make_double2(in.x - c.dconst * c.k21 * dterm.x, in.y - c.dconst * c.k21 * dterm.y);
    const double2 oo = cmul_i(opex, opey);
    in = add(in, scl(c.c1t + c.c3f * qdot, oo));
    double2 r = scl(ttf * dspec, in);
    const double2 tdiff = make_double2(tau_a.x - tau_b.x, tau_a.y - tau_b.y);
    r = add(r, scl(c.c4f, cmul_i(oo, cmul_i(sig_qxq, tdiff))));
    return r;
  };
  const double2 s1 = term(sig_dot_sig(t1, t2, t4, t5), dl(m3, m6) * dl(t3, t6), dl(m1, m4) * dl(m2, m5), q252, q142,
                          w2, q141, q251, w1, ope141, ope252, q1[0] * q2[0] + q1[1] * q2[1] + q1[2] * q2[2],
                          sdq(m3, m6, q1xq2), tau_triple(t3, t1, t2, t6, t4, t5), tau_triple(t3, t2, t1, t6, t5, t4));
  const double2 s2 = term(sig_dot_sig(t1, t3, t4, t6), dl(m2, m5) * dl(t2, t5), dl(m1, m4) * dl(m3, m6), q363, q143,
                          w3, q141, q361, w1, ope141, ope363, q1[0] * q3[0] + q1[1] * q3[1] + q1[2] * q3[2],
                          sdq(m2, m5, q1xq3), tau_triple(t2, t1, t3, t5, t4, t6), tau_triple(t2, t3, t1, t5, t6, t4));
  const double2 s3 = term(sig_dot_sig(t2, t3, t5, t6), dl(m1, m4) * dl(t1, t4), dl(m2, m5) * dl(m3, m6), q363, q253,
                          w3, q252, q362, w2, ope252, ope363, q2[0] * q3[0] + q2[1] * q3[1] + q2[2] * q3[2],
                          sdq(m1, m4, q2xq3), tau_triple(t1, t2, t3, t4, t5, t6), tau_triple(t1, t3, t2, t4, t6, t5));
  const double2 v = add(add(s1, s2), s3);
  const double f = freg_bra * freg3(c, k[3], k[4], k[5]) * c.scale;
  return make_double2(f * v.x, f * v.y);
}

static __device__ __forceinline__ double freg3_of(const Chiral3Dev &c, int p, int q, int r) {
  const double hbarc = 197.326968;
  const double k1[3] = {c.kx[p] * hbarc, c.ky[p] * hbarc, c.kz[p] * hbarc};
  const double k2[3] = {c.kx[q] * hbarc, c.ky[q] * hbarc, c.kz[q] * hbarc};
  const double k3[3] = {c.kx[r] * hbarc, c.ky[r] * hbarc, c.kz[r] * hbarc};
  return freg3(c, k1, k2, k3);
}

// chiral_3nf_with_delta_asym(i,j,p,k,l,q) (:6904-6917): the ket antisymmetrized
static __device__ double2 chiral3_asym(const Chiral3Dev &c, int i, int j, int p, int k, int l, int q) {
  const double fb = freg3_of(c, i, j, p);
  double2 r = chiral3_direct(c, fb, i, j, p, k, l, q);
  const double2 a = chiral3_direct(c, fb, i, j, p, l, k, q), b = chiral3_direct(c, fb, i, j, p, q, l, k);
  const double2 d = chiral3_direct(c, fb, i, j, p, k, q, l), e = chiral3_direct(c, fb, i, j, p, l, q, k);
  const double2 f = chiral3_direct(c, fb, i, j, p, q, k, l);
  r.x = r.x - a.x - b.x - d.x + e.x + f.x;
  r.y = r.y - a.y - b.y - d.y + e.y + f.y;
  return r;
}

// sum over the holes h of <pqh||rsh> (the normal-ordered two-body part)
static __device__ double2 no2b_sum(const Chiral3Dev &c, int A, int p, int q, int r, int s) {
  double2 acc = make_double2(0.0, 0.0);
  for (int h = 0; h < A; ++h) {
    const double2 v = chiral3_asym(c, p, q, h, r, s, h);
    acc.x += v.x;
    acc.y += v.y;
  }
  return acc;
}

}  // namespace ccgpu
