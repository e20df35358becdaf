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

}  // namespace ccgpu
