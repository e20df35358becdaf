// Perturbative triples (T) on the GPU: pert_triples2 (ccm_triples.f90:528-735)
// with three_body_oneparticle_1 / three_body_onehole_1 (ccm_triples.f90:174-294)
// and the (K3, Tz3) channels of setup_t3_channel_structures
// (ccm_t3full_channels.f90:4-284).
//
// E(T) = sum over channels, a<b<c and i<j<k of conj(W) W / D with
//   D = f_ii + f_jj + f_kk - f_aa - f_bb - f_cc            (diagonal Fock)
//   W = sum of 9 signed F terms - sum of 9 signed H terms
//   F(a,b,c,i,j,k) = sum_d conj(<cb||kd>) t(ad,ij),  n_d = n_i + n_j - n_a
//   H(a,b,c,i,j,k) = sum_l conj(<jk||lc>) t(ab,il),  n_l = n_k + n_j - n_c
// (d runs over the particle spin states of its momentum, l over the holes).
// The interaction elements are evaluated in place by the same generator that
// filled the V blocks, or, for handed-over interactions, read from a device
// hash table of the <cb||kd> (vnn_hppp) and <jk||lc> (vnn_hhhp) blocks.
//
// Work decomposition: one thread per particle triple (a<b<c): block x = the
// particle pair (a,b), threads stride over c.  A thread whose (K3,Tz3) has hole
// triples walks that channel's hole-triple list and accumulates its own sum in
// fixed order; blocks reduce with a fixed tree and a final pass adds the block
// partials in order, so the result is deterministic.
#include "device.h"
#include "interaction.h"

#include <algorithm>
#include <cstdlib>

namespace ccgpu {


__device__ __forceinline__ double minnesota_direct_t(const SolveDesc &d, double volume, int p, int q, int r, int s) {
  Orb o;
  o.kx = d.kx; o.ky = d.ky; o.kz = d.kz; o.sz = d.sz; o.tz = d.tz; o.volume = volume;
  return minnesota_direct(o, p, q, r, s);
}

__device__ __forceinline__ double2 tcmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ int orb_lookup(const TriplesDesc &td, int nx, int ny, int nz, int sz, int tz) {
  const int n = 2 * td.nmax + 1;
  if (nx < -td.nmax || nx > td.nmax || ny < -td.nmax || ny > td.nmax || nz < -td.nmax || nz > td.nmax) return -1;
  if (tz < -1 || tz > 1 || (tz & 1) == 0) return -1;
  int idx = (((nx + td.nmax) * n + (ny + td.nmax)) * n + (nz + td.nmax)) * 4 + ((sz + 1) / 2) * 2 + (tz + 1) / 2;
  return td.orbitof[idx];
}

// amplitude t(ab,ij) for any ordering, 0 when absent
__device__ __forceinline__ double2 t_of(const SolveDesc &d, int a, int b, int i, int j) {
  if (a == b || i == j) return make_double2(0.0, 0.0);
  DPair p = d.ppref[(size_t)(a - d.A) * d.np + (b - d.A)];
  DPair h = d.hhref[(size_t)i * d.A + j];
  if (p.ch < 0 || p.ch != h.ch) return make_double2(0.0, 0.0);
  double2 t = d.T[d.off_t[p.ch] + (int64_t)(p.v >> 1) * d.nhh[p.ch] + (h.v >> 1)];
  double sg = (((p.v ^ h.v) & 1) ? -1.0 : 1.0);
  return make_double2(sg * t.x, sg * t.y);
}

// antisymmetrized element <pq||rs>: from the table of handed-over elements,
// else from the built-in generator (chiral, or Minnesota which is real)
__device__ __forceinline__ double2 vas_t(const SolveDesc &d, const TriplesDesc &td, double volume, int p, int q,
                                         int r, int s) {
  if (td.vkeys) {
    const unsigned long long n = (unsigned long long)d.nsp;
    const unsigned long long key = (((unsigned long long)p * n + q) * n + r) * n + s;
    unsigned long long h = key;
    h ^= h >> 33;
    h *= 0xff51afd7ed558ccdull;
    h ^= h >> 33;
    h &= td.vmask;
    // the table is at most half full, so an empty slot ends every probe
    for (;;) {
      const unsigned long long k = td.vkeys[h];
      if (k == key) return td.vvals[h];
      if (k == ~0ull) return make_double2(0.0, 0.0);
      h = (h + 1) & td.vmask;
    }
  }
  if (td.chiral) {
    const double2 a = chiral_direct(td.ch, d.nx, d.ny, d.nz, d.sz, d.tz, p, q, r, s);
    const double2 b = chiral_direct(td.ch, d.nx, d.ny, d.nz, d.sz, d.tz, p, q, s, r);
    return make_double2(a.x - b.x, a.y - b.y);
  }
  if (td.mtab) {
    const int ax = d.nx[p] - d.nx[q], ay = d.ny[p] - d.ny[q], az = d.nz[p] - d.nz[q];
    const int dx1 = ax - d.nx[r] + d.nx[s], dy1 = ay - d.ny[r] + d.ny[s], dz1 = az - d.nz[r] + d.nz[s];
    const int dx2 = ax - d.nx[s] + d.nx[r], dy2 = ay - d.ny[s] + d.ny[r], dz2 = az - d.nz[s] + d.nz[r];
    const int m1 = dx1 * dx1 + dy1 * dy1 + dz1 * dz1, m2 = dx2 * dx2 + dy2 * dy2 + dz2 * dz2;
    const int sp = d.sz[p], sq = d.sz[q], sr = d.sz[r], ss = d.sz[s];
    const int tp = d.tz[p], tq = d.tz[q], tr = d.tz[r], ts = d.tz[s];
    const double v = minnesota_direct_tab(td.mtab + 3 * m1, sp, sq, sr, ss, tp, tq, tr, ts, volume) -
                     minnesota_direct_tab(td.mtab + 3 * m2, sp, sq, ss, sr, tp, tq, ts, tr, volume);
    return make_double2(v, 0.0);
  }
  return make_double2(minnesota_direct_t(d, volume, p, q, r, s) - minnesota_direct_t(d, volume, p, q, s, r), 0.0);
}

// F(a,b,c,i,j,k) = sum_d conj(<cb||kd>) t(ad,ij)  (three_body_oneparticle_1);
// TAB: elements from the dense tables, else evaluated in place
template <bool TAB>
__device__ double2 f_term(const SolveDesc &d, const TriplesDesc &td, double vol, int a, int b, int c, int i, int j,
                          int k) {
  double2 s = make_double2(0.0, 0.0);
  const int nx = d.nx[i] + d.nx[j] - d.nx[a], ny = d.ny[i] + d.ny[j] - d.ny[a], nz = d.nz[i] + d.nz[j] - d.nz[a];
  const int tz = d.tz[i] + d.tz[j] - d.tz[a];
  const int64_t row = TAB ? ((int64_t)(c - d.A) * d.np + (b - d.A)) * d.A + k : 0;
  for (int sz = -1; sz <= 1; sz += 2) {
    int dd = orb_lookup(td, nx, ny, nz, sz, tz);
    if (dd < d.A) continue;
    const double2 t = t_of(d, a, dd, i, j);
    double2 v;   // conj(v) t
    if constexpr (TAB) v = td.vhppp[2 * row + (sz + 1) / 2];
    else v = vas_t(d, td, vol, c, b, k, dd);
    s.x += v.x * t.x + v.y * t.y;
    s.y += v.x * t.y - v.y * t.x;
  }
  return s;
}

// H(a,b,c,i,j,k) = sum_l conj(<jk||lc>) t(ab,il)  (three_body_onehole_1)
template <bool TAB>
__device__ double2 h_term(const SolveDesc &d, const TriplesDesc &td, double vol, int a, int b, int c, int i, int j,
                          int k) {
  double2 s = make_double2(0.0, 0.0);
  const int nx = d.nx[k] + d.nx[j] - d.nx[c], ny = d.ny[k] + d.ny[j] - d.ny[c], nz = d.nz[k] + d.nz[j] - d.nz[c];
  const int tz = d.tz[k] + d.tz[j] - d.tz[c];
  const int64_t row = TAB ? ((int64_t)j * d.A + k) * d.np + (c - d.A) : 0;
  for (int sz = -1; sz <= 1; sz += 2) {
    int l = orb_lookup(td, nx, ny, nz, sz, tz);
    if (l < 0 || l >= d.A) continue;
    const double2 t = t_of(d, a, b, i, l);
    double2 v;   // conj(v) t
    if constexpr (TAB) v = td.vhhhp[2 * row + (sz + 1) / 2];
    else v = vas_t(d, td, vol, j, k, l, c);
    s.x += v.x * t.x + v.y * t.y;
    s.y += v.x * t.y - v.y * t.x;
  }
  return s;
}

// Dense element tables for (T): every <cb||kd> (particles c, b, hole k, the
// particle d of either spin that conserves momentum and isospin) and every
// <jk||lc>, evaluated once by the same element source as the in-place path.
// <cb||kd> = -<bc||kd> (the NN element and the normal-ordered 3NF sum are
// both antisymmetric in the bra): the fill evaluates c < b and writes both
// entries, c = b is zero.  Grid stride over the half (c < b) only.
__global__ __launch_bounds__(256) void k_t3_fill_hppp(SolveDesc d, TriplesDesc td, double vol, double2 *out) {
  const int np = d.np;
  const int64_t npair = (int64_t)np * (np - 1) / 2;   // c < b, in the order of pairs (b, c) with c < b
  const int64_t per = (int64_t)d.A * 2;
  const int64_t n = npair * per;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t q = e / per;
    const int rem = (int)(e - q * per);
    const int sd = (rem & 1) * 2 - 1, k = rem >> 1;
    // pair index q -> (c < b): b - A = largest x with x (x - 1) / 2 <= q
    int bx = (int)((1.0 + sqrt(1.0 + 8.0 * (double)q)) * 0.5);
    while ((int64_t)bx * (bx - 1) / 2 > q) --bx;
    while ((int64_t)(bx + 1) * bx / 2 <= q) ++bx;
    const int cx = (int)(q - (int64_t)bx * (bx - 1) / 2);
    const int b = bx + d.A, c = cx + d.A;
    // parity image (Pc, Pb, Pk): the element with the smaller (pair, k) key of
    // the two is evaluated and written to both places
    int pbx = -1, pcx = -1, pk = -1;
    if (td.parity) {
      const int ib = orb_lookup(td, -d.nx[b], -d.ny[b], -d.nz[b], d.sz[b], d.tz[b]) - d.A;
      const int ic = orb_lookup(td, -d.nx[c], -d.ny[c], -d.nz[c], d.sz[c], d.tz[c]) - d.A;
      pk = orb_lookup(td, -d.nx[k], -d.ny[k], -d.nz[k], d.sz[k], d.tz[k]);
      const int hi = max(ib, ic), lo = min(ib, ic);
      const int64_t qi = (int64_t)hi * (hi - 1) / 2 + lo;
      if (qi < q || (qi == q && pk < k)) continue;
      if (qi != q || pk != k) {
        pbx = ib;
        pcx = ic;
      }
    }
    const int dd = orb_lookup(td, d.nx[c] + d.nx[b] - d.nx[k], d.ny[c] + d.ny[b] - d.ny[k],
                              d.nz[c] + d.nz[b] - d.nz[k], sd, d.tz[c] + d.tz[b] - d.tz[k]);
    double2 v = make_double2(0.0, 0.0);
    if (dd >= d.A) {
      v = vas_t(d, td, vol, c, b, k, dd);
      if (td.tnf) {
        const double2 w = no2b_sum(td.c3, d.A, c, b, k, dd);
        v.x += w.x;
        v.y += w.y;
      }
    }
    const int64_t tail = (int64_t)k * 2 + (sd + 1) / 2;
    out[((int64_t)cx * np + bx) * per + tail] = v;
    out[((int64_t)bx * np + cx) * per + tail] = make_double2(-v.x, -v.y);
    if (pbx >= 0) {
      const int64_t ptail = (int64_t)pk * 2 + (sd + 1) / 2;
      out[((int64_t)pcx * np + pbx) * per + ptail] = v;
      out[((int64_t)pbx * np + pcx) * per + ptail] = make_double2(-v.x, -v.y);
    }
  }
  // the diagonal c = b
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < (int64_t)np * per;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t x = e / per;
    out[(x * np + x) * per + (e - x * per)] = make_double2(0.0, 0.0);
  }
}

__global__ __launch_bounds__(256) void k_t3_fill_hhhp(SolveDesc d, TriplesDesc td, double vol, double2 *out) {
  const int64_t n = (int64_t)d.A * d.A * d.np * 2;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int sl = (int)(e & 1) * 2 - 1;
    const int64_t r = e >> 1;
    const int c = (int)(r % d.np) + d.A;
    const int64_t jk = r / d.np;
    const int k = (int)(jk % d.A), j = (int)(jk / d.A);
    const int l = orb_lookup(td, d.nx[j] + d.nx[k] - d.nx[c], d.ny[j] + d.ny[k] - d.ny[c],
                             d.nz[j] + d.nz[k] - d.nz[c], sl, d.tz[j] + d.tz[k] - d.tz[c]);
    out[e] = (l >= 0 && l < d.A) ? vas_t(d, td, vol, j, k, l, c) : make_double2(0.0, 0.0);
  }
}

// The 18 signed terms of W (pert_triples2, ccm_triples.f90:611-700): kind
// (0 F, 1 H), sign, the particle order as indices into (a, b, c) and the hole
// order as indices into (i, j, k).
struct TermCode {
  signed char kind, sign, p0, p1, p2, h0, h1, h2;
};
__constant__ TermCode kTerms[18] = {
    {0, 1, 0, 1, 2, 0, 1, 2},  {0, -1, 1, 0, 2, 0, 1, 2}, {0, -1, 2, 1, 0, 0, 1, 2}, {0, -1, 0, 1, 2, 2, 1, 0},
    {0, -1, 0, 1, 2, 0, 2, 1}, {0, 1, 2, 1, 0, 0, 2, 1},  {0, 1, 2, 1, 0, 2, 1, 0},  {0, 1, 1, 0, 2, 0, 2, 1},
    {0, 1, 1, 0, 2, 2, 1, 0},  {1, -1, 0, 1, 2, 0, 1, 2}, {1, 1, 0, 2, 1, 0, 1, 2},  {1, 1, 2, 1, 0, 0, 1, 2},
    {1, 1, 0, 1, 2, 2, 1, 0},  {1, 1, 0, 1, 2, 1, 0, 2},  {1, -1, 2, 1, 0, 1, 0, 2}, {1, -1, 2, 1, 0, 2, 1, 0},
    {1, -1, 0, 2, 1, 1, 0, 2}, {1, -1, 0, 2, 1, 2, 1, 0}};

__device__ __forceinline__ int sel3(int s, int x0, int x1, int x2) { return s == 0 ? x0 : (s == 1 ? x1 : x2); }

template <bool TAB>
__global__ __launch_bounds__(256) void k_triples(SolveDesc d, TriplesDesc td, double volume, double2 *partial,
                                                 unsigned long long *count) {
  __shared__ double sre[256], sim[256];
  __shared__ unsigned long long scnt[256];
  double re = 0.0, im = 0.0;
  unsigned long long nel = 0;
  for (int64_t pr = blockIdx.x; pr < td.npairs; pr += gridDim.x) {
    const int a = td.pair_a[pr], b = td.pair_b[pr];
    for (int c = b + 1 + threadIdx.x; c < d.nsp; c += blockDim.x) {
      const int kx = d.nx[a] + d.nx[b] + d.nx[c], ky = d.ny[a] + d.ny[b] + d.ny[c], kz = d.nz[a] + d.nz[b] + d.nz[c];
      const int tz = d.tz[a] + d.tz[b] + d.tz[c];
      const int ix = kx - td.k0, iy = ky - td.k0, iz = kz - td.k0, it = tz - td.tz0;
      if (ix < 0 || iy < 0 || iz < 0 || ix >= td.nk || iy >= td.nk || iz >= td.nk || it < 0 || it >= td.ntz) continue;
      const int ch = td.chan[((it * td.nk + ix) * td.nk + iy) * td.nk + iz];
      if (ch < 0) continue;
      const double fa = d.fdiag[a].x, fb = d.fdiag[b].x, fc = d.fdiag[c].x;
      const double fai = d.fdiag[a].y, fbi = d.fdiag[b].y, fci = d.fdiag[c].y;
      nel += (unsigned long long)(td.hptr[ch + 1] - td.hptr[ch]);
      for (int h = td.hptr[ch]; h < td.hptr[ch + 1]; ++h) {
        const int i = td.hijk[3 * h], j = td.hijk[3 * h + 1], k = td.hijk[3 * h + 2];
        double2 w = make_double2(0.0, 0.0);
        auto add = [&](double sg, double2 v) { w.x += sg * v.x; w.y += sg * v.y; };
        if constexpr (TAB) {
          // the same 18 terms in the same order, one at a time (fewer live
          // registers than the unrolled form)
#pragma unroll 1
          for (int t = 0; t < 18; ++t) {
            const TermCode tc = kTerms[t];
            const int x = sel3(tc.p0, a, b, c), y = sel3(tc.p1, a, b, c), z = sel3(tc.p2, a, b, c);
            const int u = sel3(tc.h0, i, j, k), v = sel3(tc.h1, i, j, k), q = sel3(tc.h2, i, j, k);
            add((double)tc.sign, tc.kind == 0 ? f_term<true>(d, td, volume, x, y, z, u, v, q)
                                              : h_term<true>(d, td, volume, x, y, z, u, v, q));
          }
        } else {
        add(1, f_term<TAB>(d, td, volume, a, b, c, i, j, k));
        add(-1, f_term<TAB>(d, td, volume, b, a, c, i, j, k));
        add(-1, f_term<TAB>(d, td, volume, c, b, a, i, j, k));
        add(-1, f_term<TAB>(d, td, volume, a, b, c, k, j, i));
        add(-1, f_term<TAB>(d, td, volume, a, b, c, i, k, j));
        add(1, f_term<TAB>(d, td, volume, c, b, a, i, k, j));
        add(1, f_term<TAB>(d, td, volume, c, b, a, k, j, i));
        add(1, f_term<TAB>(d, td, volume, b, a, c, i, k, j));
        add(1, f_term<TAB>(d, td, volume, b, a, c, k, j, i));
        add(-1, h_term<TAB>(d, td, volume, a, b, c, i, j, k));
        add(1, h_term<TAB>(d, td, volume, a, c, b, i, j, k));
        add(1, h_term<TAB>(d, td, volume, c, b, a, i, j, k));
        add(1, h_term<TAB>(d, td, volume, a, b, c, k, j, i));
        add(1, h_term<TAB>(d, td, volume, a, b, c, j, i, k));
        add(-1, h_term<TAB>(d, td, volume, c, b, a, j, i, k));
        add(-1, h_term<TAB>(d, td, volume, c, b, a, k, j, i));
        add(-1, h_term<TAB>(d, td, volume, a, c, b, j, i, k));
        add(-1, h_term<TAB>(d, td, volume, a, c, b, k, j, i));
        }
        // conj(W) W / D with complex D
        const double dre = d.fdiag[i].x + d.fdiag[j].x + d.fdiag[k].x - fa - fb - fc;
        const double dim = d.fdiag[i].y + d.fdiag[j].y + d.fdiag[k].y - fai - fbi - fci;
        const double ww = w.x * w.x + w.y * w.y;
        const double den = dre * dre + dim * dim;
        re += ww * dre / den;
        im += -ww * dim / den;
      }
    }
  }
  sre[threadIdx.x] = re;
  sim[threadIdx.x] = im;
  scnt[threadIdx.x] = nel;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sre[threadIdx.x] += sre[threadIdx.x + s];
      sim[threadIdx.x] += sim[threadIdx.x + s];
      scnt[threadIdx.x] += scnt[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = make_double2(sre[0], sim[0]);
    count[blockIdx.x] = scnt[0];
  }
}

// chiral3_asym with the six ket permutations in a loop over one inlined copy
// of the direct element (fewer registers); the same terms summed in the same
// order, so the value is identical.
__device__ __forceinline__ double2 w3_element(const Chiral3Dev &c, int a, int b, int cc, int i, int j, int k) {
  const double fb = freg3_of(c, a, b, cc);
  const int ket[3] = {i, j, k};
  double2 r = make_double2(0.0, 0.0);
#pragma unroll 1
  for (int p = 0; p < 6; ++p) {
    // (k,l,q) +, (l,k,q) -, (q,l,k) -, (k,q,l) -, (l,q,k) +, (q,k,l) +
    const int x = p == 0 || p == 3 ? 0 : (p == 1 || p == 4 ? 1 : 2);
    const int y = p == 1 || p == 5 ? 0 : (p == 0 || p == 2 ? 1 : 2);
    const int z = 3 - x - y;
    const double2 v = chiral3_direct(c, fb, a, b, cc, ket[x], ket[y], ket[z]);
    if (p == 0 || p >= 4) {
      r.x += v.x;
      r.y += v.y;
    } else {
      r.x -= v.x;
      r.y -= v.y;
    }
  }
  return r;
}

// Table mode, one wave per particle triple: a block takes particle pairs
// (a,b), its four waves take c = b+1+wave, b+5+wave, ..., and the 64 lanes of
// a wave stride over the hole triples of that (K3,Tz3) channel.  Every lane of
// a wave reads the same <c.||..> and <..||.c> table rows (cache-friendly), and
// no lane waits for a longer hole-triple list than its wave's.
//
// With W the 18-term numerator (pert_triples2, ccm_triples.f90:611-700), W3 =
// <abc||ijk> the antisymmetrized three-nucleon element (chiral_3nf_asym) and D
// = f_i + f_j + f_k - f_a - f_b - f_c, the sum over a<b<c, i<j<k is
//   MODE 0: |W|^2 / D                 (T), et_4
//   MODE 1: (W3 + W) conj(W3) / D     CCDT-1 energy of T3 = (W3 + W)/D
//           (t3_intermediate, ccm_t3_eqn.f90:1-192; t3_energy,
//           ccm_t3full_channels.f90:570-667)
//   MODE 2: |W3|^2 / D                CCDT-1 energy of the start T3 = W3/D
//           (setup_t3, ccm_t3full_channels.f90:291-567)
//   MODE 3: |W + W3|^2 / D            (T) at tnf_approx 3, et_5 (:656-722)
template <int MODE>
__global__ __launch_bounds__(256) void k_triples_w(SolveDesc d, TriplesDesc td, double volume, double2 *partial,
                                                   unsigned long long *count) {
  __shared__ double sre[256], sim[256];
  __shared__ unsigned long long scnt[256];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double re = 0.0, im = 0.0;
  unsigned long long nel = 0;
  for (int64_t pr = blockIdx.x; pr < td.npairs; pr += gridDim.x) {
    const int a = td.pair_a[pr], b = td.pair_b[pr];
    for (int c = b + 1 + wave; c < d.nsp; c += 4) {
      const int kx = d.nx[a] + d.nx[b] + d.nx[c], ky = d.ny[a] + d.ny[b] + d.ny[c], kz = d.nz[a] + d.nz[b] + d.nz[c];
      const int tz = d.tz[a] + d.tz[b] + d.tz[c];
      const int ix = kx - td.k0, iy = ky - td.k0, iz = kz - td.k0, it = tz - td.tz0;
      if (ix < 0 || iy < 0 || iz < 0 || ix >= td.nk || iy >= td.nk || iz >= td.nk || it < 0 || it >= td.ntz) continue;
      const int ch = td.chan[((it * td.nk + ix) * td.nk + iy) * td.nk + iz];
      if (ch < 0) continue;
      const double fa = d.fdiag[a].x, fb = d.fdiag[b].x, fc = d.fdiag[c].x;
      const double fai = d.fdiag[a].y, fbi = d.fdiag[b].y, fci = d.fdiag[c].y;
      const int h0 = td.hptr[ch], h1 = td.hptr[ch + 1];
      if (lane == 0) nel += (unsigned long long)(h1 - h0);
      for (int h = h0 + lane; h < h1; h += 64) {
        const int i = td.hijk[3 * h], j = td.hijk[3 * h + 1], k = td.hijk[3 * h + 2];
        double2 w = make_double2(0.0, 0.0);
        if constexpr (MODE != 2) {
#pragma unroll 1
          for (int t = 0; t < 18; ++t) {
            const TermCode tc = kTerms[t];
            const int x = sel3(tc.p0, a, b, c), y = sel3(tc.p1, a, b, c), z = sel3(tc.p2, a, b, c);
            const int u = sel3(tc.h0, i, j, k), v = sel3(tc.h1, i, j, k), q = sel3(tc.h2, i, j, k);
            const double2 f = tc.kind == 0 ? f_term<true>(d, td, volume, x, y, z, u, v, q)
                                           : h_term<true>(d, td, volume, x, y, z, u, v, q);
            w.x += (double)tc.sign * f.x;
            w.y += (double)tc.sign * f.y;
          }
        }
        const double dre = d.fdiag[i].x + d.fdiag[j].x + d.fdiag[k].x - fa - fb - fc;
        const double dim = d.fdiag[i].y + d.fdiag[j].y + d.fdiag[k].y - fai - fbi - fci;
        const double den = dre * dre + dim * dim;
        // numerator n, then n / D = n conj(D) / |D|^2
        double nre, nim = 0.0;
        if constexpr (MODE == 0) {
          nre = w.x * w.x + w.y * w.y;
        } else {
          const double2 w3 = w3_element(td.c3w, a, b, c, i, j, k);
          if constexpr (MODE == 1) {
            const double xr = w3.x + w.x, xi = w3.y + w.y;
            nre = xr * w3.x + xi * w3.y;
            nim = xi * w3.x - xr * w3.y;
          } else if constexpr (MODE == 2) {
            nre = w3.x * w3.x + w3.y * w3.y;
          } else {
            const double xr = w3.x + w.x, xi = w3.y + w.y;
            nre = xr * xr + xi * xi;
          }
        }
        re += (nre * dre + nim * dim) / den;
        im += (nim * dre - nre * dim) / den;
      }
    }
  }
  sre[threadIdx.x] = re;
  sim[threadIdx.x] = im;
  scnt[threadIdx.x] = nel;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sre[threadIdx.x] += sre[threadIdx.x + s];
      sim[threadIdx.x] += sim[threadIdx.x + s];
      scnt[threadIdx.x] += scnt[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = make_double2(sre[0], sim[0]);
    count[blockIdx.x] = scnt[0];
  }
}

// Amplitude factor of every F term, t(x d, u v) for each particle x, ordered
// hole pair (u, v) and both spins of the particle d that conserves momentum
// and isospin (zero where d is a hole or absent): out[((x - A) A + u) A + v]
// holds the two spins side by side.  Rebuilt from T for every (T) sum.
__global__ __launch_bounds__(256) void k_t3_fill_td(SolveDesc d, TriplesDesc td, double2 *out) {
  const int64_t n = (int64_t)d.np * d.A * d.A * 2;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int sd = (int)(e & 1) * 2 - 1;
    const int64_t r = e >> 1;
    const int v = (int)(r % d.A), u = (int)((r / d.A) % d.A);
    const int x = (int)(r / ((int64_t)d.A * d.A)) + d.A;
    const int dd = orb_lookup(td, d.nx[u] + d.nx[v] - d.nx[x], d.ny[u] + d.ny[v] - d.ny[x],
                              d.nz[u] + d.nz[v] - d.nz[x], sd, d.tz[u] + d.tz[v] - d.tz[x]);
    out[e] = dd >= d.A ? t_of(d, x, dd, u, v) : make_double2(0.0, 0.0);
  }
}

// (T) from the element tables with the amplitudes staged: the same sum, terms
// and order as k_triples_w<0> (the same bits), with fewer gathers per
// element.  A wave's particle triple (a,b,c) fixes the three particle pairs
// (a,b), (a,c), (c,b) of the H terms: their rows of T (one per pair, the hole
// pairs of its channel) are copied into the wave's slice of LDS, and an H term
// reads t(xy, u l) there at the column of the hole pair (u, l).  The F terms
// read t(x d, u v) for both spins of d from the table above (one 32-byte
// load instead of the momentum lookup and three index loads per spin).
__device__ __forceinline__ void t3_wave_sync() {
  __builtin_amdgcn_s_waitcnt((15) | (3 << 14) | (7 << 4) | (0 << 8));   // lgkmcnt(0)
  __builtin_amdgcn_wave_barrier();
}

__global__ __launch_bounds__(256) void k_triples_s(SolveDesc d, TriplesDesc td, const double2 *__restrict__ tdt,
                                                   int maxnhh, double2 *partial, unsigned long long *count) {
  extern __shared__ double2 t3rows[];   // 4 waves x 3 pairs x maxnhh
  __shared__ double sre[256], sim[256];
  __shared__ unsigned long long scnt[256];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double2 *rows = t3rows + (int64_t)wave * 3 * maxnhh;
  const int A = d.A, np = d.np;
  double re = 0.0, im = 0.0;
  unsigned long long nel = 0;
  for (int64_t pr = blockIdx.x; pr < td.npairs; pr += gridDim.x) {
    const int a = td.pair_a[pr], b = td.pair_b[pr];
    for (int c = b + 1 + wave; c < d.nsp; c += 4) {
      const int kx = d.nx[a] + d.nx[b] + d.nx[c], ky = d.ny[a] + d.ny[b] + d.ny[c], kz = d.nz[a] + d.nz[b] + d.nz[c];
      const int tz = d.tz[a] + d.tz[b] + d.tz[c];
      const int ix = kx - td.k0, iy = ky - td.k0, iz = kz - td.k0, it = tz - td.tz0;
      if (ix < 0 || iy < 0 || iz < 0 || ix >= td.nk || iy >= td.nk || iz >= td.nk || it < 0 || it >= td.ntz) continue;
      const int ch = td.chan[((it * td.nk + ix) * td.nk + iy) * td.nk + iz];
      if (ch < 0) continue;
      const int h0 = td.hptr[ch], h1 = td.hptr[ch + 1];
      if (lane == 0) nel += (unsigned long long)(h1 - h0);
      // the pairs (a,b), (a,c), (c,b): channel, row, orientation; their rows
      int pch[3], psg[3];
#pragma unroll
      for (int p = 0; p < 3; ++p) {
        const int x = p == 2 ? c : a, y = p == 0 ? b : (p == 1 ? c : b);
        const DPair q = d.ppref[(size_t)(x - A) * np + (y - A)];
        pch[p] = q.ch;
        psg[p] = q.v & 1;
        if (q.ch >= 0) {
          const int nh = d.nhh[q.ch];
          const double2 *src = d.T + d.off_t[q.ch] + (int64_t)(q.v >> 1) * nh;
          for (int col = lane; col < nh; col += 64) rows[p * maxnhh + col] = src[col];
        }
      }
      t3_wave_sync();
      const double fa = d.fdiag[a].x, fb = d.fdiag[b].x, fc = d.fdiag[c].x;
      const double fai = d.fdiag[a].y, fbi = d.fdiag[b].y, fci = d.fdiag[c].y;
      for (int h = h0 + lane; h < h1; h += 64) {
        const int i = td.hijk[3 * h], j = td.hijk[3 * h + 1], k = td.hijk[3 * h + 2];
        double2 w = make_double2(0.0, 0.0);
#pragma unroll 1
        for (int t = 0; t < 18; ++t) {
          const TermCode tc = kTerms[t];
          const int x = sel3(tc.p0, a, b, c), y = sel3(tc.p1, a, b, c), z = sel3(tc.p2, a, b, c);
          const int u = sel3(tc.h0, i, j, k), v = sel3(tc.h1, i, j, k), q = sel3(tc.h2, i, j, k);
          double2 f = make_double2(0.0, 0.0);
          if (tc.kind == 0) {
            // F = sum_d conj(<zy||qd>) t(x d, u v)
            const double2 *vv = td.vhppp + 2 * (((int64_t)(z - A) * np + (y - A)) * A + q);
            const double2 *tt = tdt + 2 * (((int64_t)(x - A) * A + u) * A + v);
#pragma unroll
            for (int sd = 0; sd < 2; ++sd) {
              const double2 vx = vv[sd], tx = tt[sd];
              f.x += vx.x * tx.x + vx.y * tx.y;
              f.y += vx.x * tx.y - vx.y * tx.x;
            }
          } else {
            // H = sum_l conj(<vq||lz>) t(xy, u l)
            const int slot = tc.p0 == 2 ? 2 : (tc.p1 == 1 ? 0 : 1);
            const int nx = d.nx[q] + d.nx[v] - d.nx[z], ny = d.ny[q] + d.ny[v] - d.ny[z];
            const int nz = d.nz[q] + d.nz[v] - d.nz[z], tl = d.tz[q] + d.tz[v] - d.tz[z];
            const double2 *vv = td.vhhhp + 2 * (((int64_t)v * A + q) * np + (z - A));
#pragma unroll
            for (int sd = 0; sd < 2; ++sd) {
              const int l = orb_lookup(td, nx, ny, nz, 2 * sd - 1, tl);
              if (l < 0 || l >= A) continue;
              double2 tx = make_double2(0.0, 0.0);
              const DPair hr = d.hhref[(size_t)u * A + l];
              if (pch[slot] >= 0 && hr.ch == pch[slot]) {
                tx = rows[slot * maxnhh + (hr.v >> 1)];
                if ((psg[slot] ^ hr.v) & 1) tx = make_double2(-tx.x, -tx.y);
              }
              const double2 vx = vv[sd];
              f.x += vx.x * tx.x + vx.y * tx.y;
              f.y += vx.x * tx.y - vx.y * tx.x;
            }
          }
          w.x += (double)tc.sign * f.x;
          w.y += (double)tc.sign * f.y;
        }
        const double dre = d.fdiag[i].x + d.fdiag[j].x + d.fdiag[k].x - fa - fb - fc;
        const double dim = d.fdiag[i].y + d.fdiag[j].y + d.fdiag[k].y - fai - fbi - fci;
        const double den = dre * dre + dim * dim;
        const double nre = w.x * w.x + w.y * w.y, nim = 0.0;
        re += (nre * dre + nim * dim) / den;
        im += (nim * dre - nre * dim) / den;
      }
      // the rows are overwritten for the next triple only after every lane read them
      t3_wave_sync();
    }
  }
  sre[threadIdx.x] = re;
  sim[threadIdx.x] = im;
  scnt[threadIdx.x] = nel;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sre[threadIdx.x] += sre[threadIdx.x + s];
      sim[threadIdx.x] += sim[threadIdx.x + s];
      scnt[threadIdx.x] += scnt[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = make_double2(sre[0], sim[0]);
    count[blockIdx.x] = scnt[0];
  }
}

__global__ void k_triples_count(const unsigned long long *count, int n, unsigned long long *out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  unsigned long long s = 0;
  for (int k = 0; k < n; ++k) s += count[k];
  out[0] = s;
}

__global__ void k_triples_final(const double2 *partial, int n, double2 *out) {
  __shared__ double sre[256], sim[256];
  double re = 0.0, im = 0.0;
  for (int p = threadIdx.x; p < n; p += blockDim.x) {
    re += partial[p].x;
    im += partial[p].y;
  }
  sre[threadIdx.x] = re;
  sim[threadIdx.x] = im;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sre[threadIdx.x] += sre[threadIdx.x + s];
      sim[threadIdx.x] += sim[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = make_double2(sre[0], sim[0]);
}

void launch_t3_tables(const SolveDesc &d, const TriplesDesc &td, double volume, double2 *hppp, double2 *hhhp,
                      hipStream_t st) {
  hipLaunchKernelGGL(k_t3_fill_hppp, dim3(8192), dim3(256), 0, st, d, td, volume, hppp);
  hipLaunchKernelGGL(k_t3_fill_hhhp, dim3(1024), dim3(256), 0, st, d, td, volume, hhhp);
}

void launch_t3_td(const SolveDesc &d, const TriplesDesc &td, double2 *tdt, hipStream_t st) {
  hipLaunchKernelGGL(k_t3_fill_td, dim3(4096), dim3(256), 0, st, d, td, tdt);
}

// CCGPU_T3_KERNEL (studies and the tests that compare the forms): 1 (default)
// the staged kernel when its amplitude table tdt exists, else the table
// kernel k_triples_w; 3 k_triples_w; other values the thread-per-triple form.
void launch_triples(const SolveDesc &d, const TriplesDesc &td, double volume, double2 *partial,
                    unsigned long long *count, int nblocks, double2 *out, unsigned long long *count_out,
                    hipStream_t st, int mode, const double2 *tdt, int maxnhh) {
  const int variant = getenv("CCGPU_T3_KERNEL") ? atoi(getenv("CCGPU_T3_KERNEL")) : 1;
  if (mode == 0 && variant == 1 && tdt && td.vhppp && td.vhhhp) {
    const size_t lds = (size_t)4 * 3 * maxnhh * sizeof(double2);
    hipLaunchKernelGGL(k_triples_s, dim3(nblocks), dim3(256), lds, st, d, td, tdt, maxnhh, partial, count);
  } else if (mode == 1)
    hipLaunchKernelGGL(k_triples_w<1>, dim3(nblocks), dim3(256), 0, st, d, td, volume, partial, count);
  else if (mode == 2)
    hipLaunchKernelGGL(k_triples_w<2>, dim3(nblocks), dim3(256), 0, st, d, td, volume, partial, count);
  else if (mode == 3)
    hipLaunchKernelGGL(k_triples_w<3>, dim3(nblocks), dim3(256), 0, st, d, td, volume, partial, count);
  else if (td.vhppp && td.vhhhp && (variant == 1 || variant == 3))
    hipLaunchKernelGGL(k_triples_w<0>, dim3(nblocks), dim3(256), 0, st, d, td, volume, partial, count);
  else if (td.vhppp && td.vhhhp)
    hipLaunchKernelGGL(k_triples<true>, dim3(nblocks), dim3(256), 0, st, d, td, volume, partial, count);
  else
    hipLaunchKernelGGL(k_triples<false>, dim3(nblocks), dim3(256), 0, st, d, td, volume, partial, count);
  hipLaunchKernelGGL(k_triples_final, dim3(1), dim3(256), 0, st, partial, nblocks, out);
  hipLaunchKernelGGL(k_triples_count, dim3(1), dim3(64), 0, st, count, nblocks, count_out);
}

}  // namespace ccgpu
