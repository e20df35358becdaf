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

}  // namespace ccgpu
