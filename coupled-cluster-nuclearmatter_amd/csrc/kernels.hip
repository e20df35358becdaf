// HIP kernels of the CCD/(T) engine for gfx950.
//
// Storage (per solve, all complex interleaved double2, row-major):
//   T, R, R0 : per (K,Tz) channel a block Npp x Nhh of unique pairs (a<b, i<j)
//   Vpppp    : Npp x Npp per channel       Vhhpp : Nhh x Npp per channel
//   Vhhhh, X : Nhh x Nhh per channel       (X = Vhhpp T, I9 = X + Vhhhh)
//   Tt, Rt   : Pandya blocks Nph x Nhp     I4, W0 : Nhp x Nhp    Vt : Nhp x Nph
// Reference routines restated here are cited per kernel (paths relative to
// deltafull-CCMkspace/ of the reference).
#include "device.h"
#include "interaction.h"

#include <algorithm>
#include <type_traits>

namespace ccgpu {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cscale(double s, double2 a) { return make_double2(s * a.x, s * a.y); }
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cdiv(double2 a, double2 b) {
  double d = b.x * b.x + b.y * b.y;
  return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}

// Interaction generators: antisymmetrized <pq||rs> = v(p,q,r,s) - v(p,q,s,r)
// (the combination of every vnn_* block, ccm_iter.f90:854, 917, 996, 1126,
// 1254); momentum conservation holds for every element the callers request.
struct MinnesotaV {
  Orb o;
  __device__ __forceinline__ double2 operator()(int p, int q, int r, int s) const {
    return make_double2(minnesota_direct(o, p, q, r, s) - minnesota_direct(o, p, q, s, r), 0.0);
  }
};

struct ChiralV {
  ChiralDev c;
  const int *nx, *ny, *nz, *sz, *tz;
  __device__ __forceinline__ double2 operator()(int p, int q, int r, int s) const {
    const double2 a = chiral_direct(c, nx, ny, nz, sz, tz, p, q, r, s);
    const double2 b = chiral_direct(c, nx, ny, nz, sz, tz, p, q, s, r);
    return make_double2(a.x - b.x, a.y - b.y);
  }
};

// V blocks of setup_channel_structures (ccm_iter.f90:819-1311), unique pairs.
template <class V>
__global__ void k_fill_vpppp(SolveDesc d, V vas) {
  const int c = blockIdx.y;
  const int n = d.npp[c];
  const int64_t ne = (int64_t)n * n;
  const int *pa = d.pp_a + d.off_pp[c], *pb = d.pp_b + d.off_pp[c];
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
    int r = (int)(e / n), s = (int)(e % n);
    d.Vpppp[d.off_vpp[c] + e] = vas(pa[r], pb[r], pa[s], pb[s]);
  }
}

template <class V>
__global__ void k_fill_vhhpp(SolveDesc d, V vas) {
  const int c = blockIdx.y;
  const int np_ = d.npp[c], nh = d.nhh[c];
  const int64_t ne = (int64_t)np_ * nh;
  const int *pa = d.pp_a + d.off_pp[c], *pb = d.pp_b + d.off_pp[c];
  const int *hi = d.hh_i + d.off_hh[c], *hj = d.hh_j + d.off_hh[c];
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
    int s = (int)(e / np_), r = (int)(e % np_);
    d.Vhhpp[d.off_t[c] + e] = vas(hi[s], hj[s], pa[r], pb[r]);
  }
}

template <class V>
__global__ void k_fill_vhhhh(SolveDesc d, V vas) {
  const int c = blockIdx.y;
  const int nh = d.nhh[c];
  const int64_t ne = (int64_t)nh * nh;
  const int *hi = d.hh_i + d.off_hh[c], *hj = d.hh_j + d.off_hh[c];
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
    int s = (int)(e / nh), t = (int)(e % nh);
    d.Vhhhh[d.off_vhh[c] + e] = vas(hi[s], hj[s], hi[t], hj[t]);
  }
}

// W0((i,a),(j,b)) = -<ib||ja> (the recoupled hphp term of I4, ccm_t2_eqn.f90:740-780).
template <class V>
__global__ void k_fill_w0(SolveDesc d, V vas) {
  const int K = blockIdx.y;
  const int n = d.nhp[K];
  const int64_t ne = (int64_t)n * n;
  const int *hj = d.hp_j + d.off_hp[K], *hb = d.hp_b + d.off_hp[K];
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
    int r = (int)(e / n), s = (int)(e % n);
    int i = hj[r], a = hb[r], j = hj[s], b = hb[s];
    double2 v = vas(i, b, j, a);
    d.W0[d.off_i4[K] + e] = make_double2(-v.x, -v.y);
  }
}

// Fock matrix f_pq = t_p delta_pq + sum_h <ph||qh> for q in {p, partner(p)}
// (normal_ordered_hamiltonian, ccm_main.f90:1586-1657).
template <class V>
__global__ void k_fock(SolveDesc d, V vas) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= d.nsp) return;
  double2 fd = make_double2(d.e[p], 0.0), fp = make_double2(0.0, 0.0);
  int pp = d.partner[p];
  for (int h = 0; h < d.A; ++h) {
    fd = cadd(fd, vas(p, h, p, h));
    if (pp >= 0) fp = cadd(fp, vas(p, h, pp, h));
  }
  d.fdiag[p] = fd;
  d.fpart[p] = fp;
}

// E0 = sum_i t_ii + 1/2 sum_ij <ij||ij> (ccm_main.f90:1661-1684).  One block.
template <class V>
__global__ void k_e0(SolveDesc d, V vas) {
  __shared__ double sre[256], sim[256];
  double re = 0.0, im = 0.0;
  for (int64_t e = threadIdx.x; e < (int64_t)d.A * d.A; e += blockDim.x) {
    int i = (int)(e / d.A), j = (int)(e % d.A);
    double2 v = vas(i, j, i, j);
    re += 0.5 * v.x; im += 0.5 * v.y;
  }
  for (int i = threadIdx.x; i < d.A; i += blockDim.x) re += d.e[i];
  sre[threadIdx.x] = re; sim[threadIdx.x] = im;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) { sre[threadIdx.x] += sre[threadIdx.x + w]; sim[threadIdx.x] += sim[threadIdx.x + w]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) d.e0[0] = make_double2(sre[0], sim[0]);
}

// ---------------------------------------------------------------------------
// Normal-ordered NNLO three-nucleon force added to the NN blocks
// (normal_ordered_hamiltonian and setup_3nf_2bNO, ccm_main.f90:1687-2568).
// N(p,q,r,s) = sum_h <pqh||rsh>.  The reference's block conventions are kept:
//   hhhh, pppp: computed for the descending pair orientation and, between the
//     two pairs, for the bra that comes first (lexicographic orbital order),
//     the other triangle as the complex conjugate (ccm_main.f90:2046-2140,
//     2420-2540);
//   hphp: the same Hermitian fill on the (hole, particle) pairs (2290-2410);
//   hhpp: <ij||ab> receives N(a,b,i,j), the sum with the particle pair in the
//     bra (2160-2260).
// The float32 2/3 of the 3NF regulator makes the element symmetric under
// pair permutations only to ~1e-7, which is why the orientation matters.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool pair_le(int a, int b, int c, int d) { return a < c || (a == c && b <= d); }

__device__ __forceinline__ double2 herm_no2b(const Chiral3Dev &c, int A, int p, int q, int r, int s) {
  if (pair_le(p, q, r, s)) return no2b_sum(c, A, p, q, r, s);
  const double2 v = no2b_sum(c, A, r, s, p, q);
  return make_double2(v.x, -v.y);
}

// The Hermitian blocks (V_pppp, V_hhhh, W0) take herm_no2b, which evaluates
// every element from the orientation pair_le puts first, so element (s, r) is
// the conjugate of element (r, s) bit for bit.  One thread evaluates the pair
// r <= s and adds to both.
__device__ __forceinline__ void add_herm_pair(double2 *blk, int n, int r, int s, double2 v, double sg) {
  double2 &o = blk[(int64_t)r * n + s];
  o.x += sg * v.x;
  o.y += sg * v.y;
  if (r != s) {
    double2 &t = blk[(int64_t)s * n + r];
    t.x += sg * v.x;
    t.y -= sg * v.y;
  }
}

// upper-triangle index t = s (s + 1) / 2 + r, r <= s, to (r, s)
__device__ __forceinline__ void tri_rs(int64_t t, int &r, int &s) {
  int64_t x = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while (x * (x + 1) / 2 > t) --x;
  while ((x + 1) * (x + 2) / 2 <= t) ++x;
  s = (int)x;
  r = (int)(t - x * (x + 1) / 2);
}

// With d.pimg the element's parity image <Pb Pa||Pd Pc> equals it (the 3NF
// is even); the image of the stored pair r is sr times pair ri of channel
// chi, so V(chi)[ri][si] = sr ss v.  Of the two, the element with the
// smaller (channel, upper-triangle position) evaluates and writes both.
__global__ __launch_bounds__(256) void k_no2b_vpppp(SolveDesc d, Chiral3Dev c) {
  const int ch = blockIdx.y;
  const int n = d.npp[ch];
  const int64_t ne = (int64_t)n * (n + 1) / 2;
  const int *pa = d.pp_a + d.off_pp[ch], *pb = d.pp_b + d.off_pp[ch];
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
    int r, s;
    tri_rs(e, r, s);
    int chi = -1, ri = 0, si = 0;
    double sg = 1.0;
    if (d.pimg) {
      const DPair pr = d.ppref[(int64_t)(d.pimg[pb[r]] - d.A) * d.np + (d.pimg[pa[r]] - d.A)];
      const DPair ps = d.ppref[(int64_t)(d.pimg[pb[s]] - d.A) * d.np + (d.pimg[pa[s]] - d.A)];
      const int pri = pr.v >> 1, psi = ps.v >> 1;   // pair index; bit 0: the orientation flips
      const int lo = min(pri, psi), hi = max(pri, psi);
      if (pr.ch < ch || (pr.ch == ch && (lo < r || (lo == r && hi < s)))) continue;
      if (pr.ch != ch || lo != r || hi != s) {
        chi = pr.ch;
        ri = pri;
        si = psi;
        sg = ((pr.v ^ ps.v) & 1) ? -1.0 : 1.0;
      }
    }
    const double2 v = herm_no2b(c, d.A, pb[r], pa[r], pb[s], pa[s]);   // descending orientation
    add_herm_pair(d.Vpppp + d.off_vpp[ch], n, r, s, v, 1.0);
    if (chi >= 0) {
      const double2 w = make_double2(sg * v.x, sg * v.y);
      if (ri <= si) add_herm_pair(d.Vpppp + d.off_vpp[chi], d.npp[chi], ri, si, w, 1.0);
      else add_herm_pair(d.Vpppp + d.off_vpp[chi], d.npp[chi], si, ri, make_double2(w.x, -w.y), 1.0);
    }
  }
}

__global__ __launch_bounds__(256) void k_no2b_vhhhh(SolveDesc d, Chiral3Dev c) {
  const int ch = blockIdx.y;
  const int nh = d.nhh[ch];
  const int64_t ne = (int64_t)nh * (nh + 1) / 2;
  const int *hi = d.hh_i + d.off_hh[ch], *hj = d.hh_j + d.off_hh[ch];
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
    int s, t;
    tri_rs(e, s, t);
    const double2 v = herm_no2b(c, d.A, hj[s], hi[s], hj[t], hi[t]);
    add_herm_pair(d.Vhhhh + d.off_vhh[ch], nh, s, t, v, 1.0);
  }
}

__global__ __launch_bounds__(256) void k_no2b_vhhpp(SolveDesc d, Chiral3Dev c) {
  const int ch = blockIdx.y;
  const int np_ = d.npp[ch], nh = d.nhh[ch];
  const int64_t ne = (int64_t)np_ * nh;
  const int *pa = d.pp_a + d.off_pp[ch], *pb = d.pp_b + d.off_pp[ch];
  const int *hi = d.hh_i + d.off_hh[ch], *hj = d.hh_j + d.off_hh[ch];
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
    const int s = (int)(e / np_), r = (int)(e % np_);
    const double2 v = no2b_sum(c, d.A, pa[r], pb[r], hi[s], hj[s]);
    double2 &o = d.Vhhpp[d.off_t[ch] + e];
    o.x += v.x;
    o.y += v.y;
  }
}

// W0((i,a),(j,b)) = -<ib||ja>
__global__ __launch_bounds__(256) void k_no2b_w0(SolveDesc d, Chiral3Dev c) {
  const int K = blockIdx.y;
  const int n = d.nhp[K];
  const int64_t ne = (int64_t)n * (n + 1) / 2;
  const int *hj = d.hp_j + d.off_hp[K], *hb = d.hp_b + d.off_hp[K];
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
    int r, s;
    tri_rs(e, r, s);
    const int i = hj[r], a = hb[r], j = hj[s], b = hb[s];
    const double2 v = herm_no2b(c, d.A, i, b, j, a);
    add_herm_pair(d.W0 + d.off_i4[K], n, r, s, v, -1.0);
  }
}

// f_pq += sum_{i<=j} <ijp||ijq> for q in {p, partner(p)} (ccm_main.f90:1826-1995)
__global__ __launch_bounds__(256) void k_3nf_fock(SolveDesc d, Chiral3Dev c) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= d.nsp) return;
  const int pp = d.partner[p];
  double2 fd = make_double2(0.0, 0.0), fp = make_double2(0.0, 0.0);
  for (int i = 0; i < d.A; ++i)
    for (int j = i; j < d.A; ++j) {
      const double2 v = chiral3_asym(c, i, j, p, i, j, p);
      fd.x += v.x;
      fd.y += v.y;
      if (pp >= 0) {
        const double2 w = chiral3_asym(c, i, j, p, i, j, pp);
        fp.x += w.x;
        fp.y += w.y;
      }
    }
  d.fdiag[p] = make_double2(d.fdiag[p].x + fd.x, d.fdiag[p].y + fd.y);
  d.fpart[p] = make_double2(d.fpart[p].x + fp.x, d.fpart[p].y + fp.y);
}

// E0 += sum_{i<j<k} <ijk||ijk> (ccm_main.f90:1687-1800).  One block.
__global__ __launch_bounds__(256) void k_3nf_e0(SolveDesc d, Chiral3Dev c) {
  __shared__ double sre[256], sim[256];
  double re = 0.0, im = 0.0;
  const int64_t A = d.A, n = A * A * A;
  for (int64_t e = threadIdx.x; e < n; e += blockDim.x) {
    const int i = (int)(e / (A * A)), j = (int)((e / A) % A), k = (int)(e % A);
    if (!(i < j && j < k)) continue;
    const double2 v = chiral3_asym(c, i, j, k, i, j, k);
    re += v.x;
    im += v.y;
  }
  sre[threadIdx.x] = re;
  sim[threadIdx.x] = im;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sre[threadIdx.x] += sre[threadIdx.x + w];
      sim[threadIdx.x] += sim[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) d.e0[0] = make_double2(d.e0[0].x + sre[0], d.e0[0].y + sim[0]);
}

// <pq||rs> for a list of orbital quadruples; zero where momentum or isospin
// is not conserved.
template <class V>
__global__ void k_elements(SolveDesc d, V vas, const int4 *pqrs, int n, double2 *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int4 x = pqrs[i];
  const bool ok = d.nx[x.x] + d.nx[x.y] == d.nx[x.z] + d.nx[x.w] && d.ny[x.x] + d.ny[x.y] == d.ny[x.z] + d.ny[x.w] &&
                  d.nz[x.x] + d.nz[x.y] == d.nz[x.z] + d.nz[x.w] && d.tz[x.x] + d.tz[x.y] == d.tz[x.z] + d.tz[x.w];
  out[i] = ok ? vas(x.x, x.y, x.z, x.w) : make_double2(0.0, 0.0);
}

// R0(ab,ij) = conj(<ij||ab>): the V term of the residual (ccm_t2_eqn.f90:107-138),
// stored in amplitude layout.
__global__ void k_derive_r0(SolveDesc d) {
  const int c = blockIdx.y;
  const int np_ = d.npp[c], nh = d.nhh[c];
  const int64_t ne = (int64_t)np_ * nh;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
    int r = (int)(e / nh), s = (int)(e % nh);
    d.R0[d.off_t[c] + e] = cconj(d.Vhhpp[d.off_t[c] + (int64_t)s * np_ + r]);
  }
}

__device__ __forceinline__ int pref_idx(DPair p) { return p.v >> 1; }
__device__ __forceinline__ double pref_sign(DPair p) { return (p.v & 1) ? -1.0 : 1.0; }

// Vt((i,a),(b,j)) = <ij||ab> gathered from Vhhpp (the amat gather, ccm_t2_eqn.f90:694-716).
__global__ void k_derive_vt(SolveDesc d) {
  const int K = blockIdx.y;
  const int nhp = d.nhp[K], nph = d.nph[K];
  const int64_t ne = (int64_t)nhp * nph;
  const int *hj = d.hp_j + d.off_hp[K], *hb = d.hp_b + d.off_hp[K];
  const int *pa = d.ph_a + d.off_ph[K], *pi = d.ph_i + d.off_ph[K];
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
    int r = (int)(e / nph), s = (int)(e % nph);
    int i = hj[r], a = hb[r], b = pa[s], j = pi[s];
    double2 v = make_double2(0.0, 0.0);
    if (i != j && a != b) {
      DPair h = d.hhref[(size_t)i * d.A + j];
      DPair p = d.ppref[(size_t)(a - d.A) * d.np + (b - d.A)];
      if (h.ch >= 0 && p.ch == h.ch) {
        double sg = pref_sign(h) * pref_sign(p);
        v = cscale(sg, d.Vhhpp[d.off_t[h.ch] + (int64_t)pref_idx(h) * d.npp[h.ch] + pref_idx(p)]);
      }
    }
    d.Vt[d.off_vt[K] + e] = v;
  }
}

// Drop-in hand-over: dst[idx[e]] = val[e] for the canonical elements of one
// host block (ccgpu_set_block uploads only these pairs).
__global__ void k_scatter_values(double2 *dst, const int64_t *idx, const double2 *val, int64_t n) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    dst[idx[e]] = val[e];
}

// ---------------------------------------------------------------------------
// per-iteration kernels; blockIdx.y selects the solve of a batch
// ---------------------------------------------------------------------------
// orbitals of unique amplitude u from its packed label
__device__ __forceinline__ void amp_orbitals(const SolveDesc &d, int64_t u, int &a, int &b, int &i, int &j) {
  const uint64_t l = d.lab[u];
  a = (int)(l & 0xffff);
  b = (int)((l >> 16) & 0xffff);
  i = (int)((l >> 32) & 0xffff);
  j = (int)(l >> 48);
}
// dynamic energy denominator dh_i + dh_j - dp_a - dp_b of unique amplitude u
__device__ __forceinline__ double2 amp_denominator(const SolveDesc &d, int64_t u) {
  int a, b, i, j;
  amp_orbitals(d, u, a, b, i, j);
  const double2 hi = d.dh[i], hj = d.dh[j], pa = d.dp[a - d.A], pb = d.dp[b - d.A];
  return make_double2(hi.x + hj.x - pa.x - pb.x, hi.y + hj.y - pa.y - pb.y);
}
__device__ __forceinline__ void unique_labels(const SolveDesc &d, int64_t u, int &c, int &r, int &s) {
  c = d.t_ch[u];
  int64_t loc = u - d.off_t[c];
  int nh = d.nhh[c];
  r = (int)(loc / nh);
  s = (int)(loc % nh);
}

// MBPT2 start t(ab,ij) = conj(<ij||ab>)/(f_ii+f_jj-f_aa-f_bb), zero denominators
// skipped (setup_t_amplitudes, ccm_iter.f90:2351-2367).
// Shared pp ladder of a batch of twists of one lattice (one V_pppp over the
// union of their particle pairs, per (K, Tz) channel one GEMM whose columns
// are the hole pairs of every twist side by side).  maps[s][u] = 2 p + (sign
// < 0): the position p of solve s's unique amplitude u in the union block
// (row: the pair's union row, column: the solve's column range), with the
// orientation sign of the pair there.  Positions no solve owns stay zero.
__global__ void k_share_gather(const SolveDesc *descs, const uint32_t *const *maps, double2 *Tu) {
  const SolveDesc d = descs[blockIdx.y];   // a copy: stores cannot change it, so no reloads
  const uint32_t *m = maps[blockIdx.y];
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < d.nT; u += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t w = m[u];
    const double2 t = d.T[u];
    Tu[w >> 1] = (w & 1) ? make_double2(-t.x, -t.y) : t;
  }
}
// R += the union product at each amplitude's position (active solves only)
__global__ void k_share_scatter(const SolveDesc *descs, const uint32_t *const *maps, const double2 *Ru) {
  const SolveDesc d = descs[blockIdx.y];   // a copy: stores cannot change it, so no reloads
  if (!d.active) return;
  const uint32_t *m = maps[blockIdx.y];
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < d.nT; u += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t w = m[u];
    const double2 r = Ru[w >> 1];
    double2 x = d.R[u];
    if (w & 1) {
      x.x -= r.x;
      x.y -= r.y;
    } else {
      x.x += r.x;
      x.y += r.y;
    }
    d.R[u] = x;
  }
}

__global__ void k_mbpt2(const SolveDesc *descs) {
  const SolveDesc d = descs[blockIdx.y];   // a copy: stores cannot change it, so no reloads
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < d.nT; u += (int64_t)gridDim.x * blockDim.x) {
    int a, b, i, j;
    amp_orbitals(d, u, a, b, i, j);
    double den = -d.fdiag[a].x - d.fdiag[b].x + d.fdiag[i].x + d.fdiag[j].x;
    double2 t = make_double2(0.0, 0.0);
    if (fabs(den) != 0.0) { double2 v = d.R0[u]; t = make_double2(v.x / den, v.y / den); }
    d.T[u] = t;
  }
}

// Deterministic block reduction of one complex value per thread.
__device__ double2 block_sum(double2 v) {
  __shared__ double sre[256], sim[256];
  sre[threadIdx.x] = v.x; sim[threadIdx.x] = v.y;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) { sre[threadIdx.x] += sre[threadIdx.x + w]; sim[threadIdx.x] += sim[threadIdx.x + w]; }
    __syncthreads();
  }
  double2 r = make_double2(sre[0], sim[0]);
  __syncthreads();
  return r;
}

// Correlation energy 1/4 sum_ordered t(ab,ij) <ij||ab> = sum_unique t * conj(R0)
// (ccd_energy_save, ccm_energy.f90:31-48).  Pass 1: block partials.
__global__ void k_energy_partial(const SolveDesc *descs) {
  const SolveDesc d = descs[blockIdx.y];   // a copy: stores cannot change it, so no reloads
  double2 acc = make_double2(0.0, 0.0);
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < d.nT; u += (int64_t)gridDim.x * blockDim.x)
    acc = cadd(acc, cmul(d.T[u], cconj(d.R0[u])));
  double2 s = block_sum(acc);
  if (threadIdx.x == 0) d.red[blockIdx.x] = s;
}

// Pass 2: fixed-order sum of nparts partials into scal[slot].
__global__ void k_reduce_final(const SolveDesc *descs, int nparts, int stride, int nvals, int slot) {
  const SolveDesc d = descs[blockIdx.y];   // a copy: stores cannot change it, so no reloads
  for (int v = 0; v < nvals; ++v) {
    double2 acc = make_double2(0.0, 0.0);
    for (int p = threadIdx.x; p < nparts; p += blockDim.x) acc = cadd(acc, d.red[(int64_t)p * stride + v]);
    double2 s = block_sum(acc);
    if (threadIdx.x == 0) d.scal[slot + v] = s;
  }
}

// I7(k,i) = f_ki + 1/2 sum_l sum_cd <kl||cd> t(cd,il) for k in {i, partner(i)}, read
// from X = Vhhpp T (setup_I6I7_intermediate, ccm_t2_eqn.f90:1352-1392, 1507-1510).
// dh(i) = I7(i,i) is the dynamic hole denominator (ccm_t2_eqn.f90:81-84).
__device__ __forceinline__ double2 wave_sum(double2 v);

__global__ __launch_bounds__(256) void k_i7(const SolveDesc *descs) {
  const SolveDesc d = descs[blockIdx.y];   // a copy: stores cannot change it, so no reloads
  if (!d.active) return;
  const int lane = threadIdx.x & 63;
  const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= 2 * d.A) return;
  const int i = t >> 1, which = t & 1;
  const int k = which ? d.partner[i] : i;
  if (k < 0 || k >= d.A) {
    if (which && lane == 0) d.I7off[i] = make_double2(0.0, 0.0);
    return;
  }
  double2 acc = make_double2(0.0, 0.0);
  for (int l = lane; l < d.A; l += 64) {
    if (l == k || l == i) continue;
    DPair pk = d.hhref[(size_t)k * d.A + l], pi = d.hhref[(size_t)i * d.A + l];
    if (pk.ch < 0 || pk.ch != pi.ch) continue;
    int c = pk.ch;
    double2 x = d.X[d.off_vhh[c] + (int64_t)pref_idx(pk) * d.nhh[c] + pref_idx(pi)];
    acc = cadd(acc, cscale(pref_sign(pk) * pref_sign(pi), x));
  }
  acc = wave_sum(acc);
  if (lane == 0) {
    acc = cadd(acc, which ? d.fpart[k] : d.fdiag[i]);
    if (which) d.I7off[i] = acc; else d.dh[i] = acc;
  }
}

// I6(a,c) = f_ac - 1/2 sum_d sum_kl <kl||dc> t(da,kl) for c in {a, partner(a)}
// (ccm_t2_eqn.f90:1399-1441); dp(a) = I6(a,a) (ccm_t2_eqn.f90:86-89).
// Pass 1, one unique pp row (x<y) per 16 lanes: with V(kl,dc) = conj(R0(dc,kl))
// the row contributes -conj(R0_r).T_r to I6(x,x) and I6(y,y),
// -s_B conj(R0_(x,ybar)).T_r to I6(y,ybar) and +s_A conj(R0_(y,xbar)).T_r to
// I6(x,xbar).  Pass 2 sums each particle's list of partials in fixed order.
__device__ __forceinline__ double2 cdotc(const double2 *v, const double2 *t, int n, int l16) {
  double re = 0.0, im = 0.0;
  for (int s = l16; s < n; s += 16) {
    double2 a = v[s], b = t[s];
    re += a.x * b.x + a.y * b.y;
    im += a.x * b.y - a.y * b.x;
  }
  for (int off = 8; off > 0; off >>= 1) {
    re += __shfl_xor(re, off, 16);
    im += __shfl_xor(im, off, 16);
  }
  return make_double2(re, im);
}

__global__ __launch_bounds__(256) void k_i6_rows(const SolveDesc *descs) {
  const SolveDesc d = descs[blockIdx.y];   // a copy: stores cannot change it, so no reloads
  if (!d.active) return;
  const int l16 = threadIdx.x & 15;
  for (int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4; r < d.nrows;
       r += ((int64_t)gridDim.x * blockDim.x) >> 4) {
    const int c = d.row_ch[r], nh = d.nhh[c];
    const int64_t base = d.off_t[c], loc = r - d.off_pp[c];
    const double2 *tr = d.T + base + loc * nh;
    double2 d0 = cdotc(d.R0 + base + loc * nh, tr, nh, l16);
    double2 dB = make_double2(0.0, 0.0), dA = make_double2(0.0, 0.0);
    const int vB = d.ob_rowB[r], vA = d.ob_rowA[r];
    if (vB >= 0) dB = cscale((double)((vB & 3) - 1), cdotc(d.R0 + base + (int64_t)(vB >> 2) * nh, tr, nh, l16));
    if (vA >= 0) dA = cscale((double)((vA & 3) - 1), cdotc(d.R0 + base + (int64_t)(vA >> 2) * nh, tr, nh, l16));
    if (l16 == 0) {
      d.p6[3 * r + 0] = make_double2(-d0.x, -d0.y);
      d.p6[3 * r + 1] = make_double2(-dB.x, -dB.y);
      d.p6[3 * r + 2] = dA;
    }
  }
}

// one wavefront per target; lanes take every 64th list entry, then a fixed
// butterfly: deterministic for a given list
__device__ __forceinline__ double2 wave_sum(double2 v) {
  for (int off = 32; off > 0; off >>= 1) {
    v.x += __shfl_xor(v.x, off);
    v.y += __shfl_xor(v.y, off);
  }
  return v;
}

__global__ __launch_bounds__(256) void k_i6_reduce(const SolveDesc *descs) {
  const SolveDesc d = descs[blockIdx.y];   // a copy: stores cannot change it, so no reloads
  if (!d.active) return;
  const int lane = threadIdx.x & 63;
  const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= 2 * d.np) return;
  const int p = t >> 1, which = t & 1, a = d.A + p;
  if (which && d.partner[a] < d.A) {
    if (lane == 0) d.I6off[p] = make_double2(0.0, 0.0);
    return;
  }
  const int *ptr = which ? d.i6_optr : d.i6_dptr;
  const int *idx = which ? d.i6_oidx : d.i6_didx;
  double2 acc = make_double2(0.0, 0.0);
  for (int q = ptr[p] + lane; q < ptr[p + 1]; q += 64) acc = cadd(acc, d.p6[idx[q]]);
  acc = wave_sum(acc);
  if (lane == 0) {
    acc = cadd(acc, which ? d.fpart[a] : d.fdiag[a]);
    if (which) d.I6off[p] = acc; else d.dp[p] = acc;
  }
}

// I9 = Vhhhh + 1/2 Vhhpp T (ccm_t2_eqn.f90:431-464, 558-560): X[0,n) holds
// Vhhpp T (unique pairs absorb the 1/2), X[n,2n) receives I9.
__global__ void k_i9(const SolveDesc *descs) {
  const SolveDesc d = descs[blockIdx.y];   // a copy: stores cannot change it, so no reloads
  if (!d.active) return;
  int64_t n = d.off_vhh[d.nch];
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    d.X[n + e] = cadd(d.X[e], d.Vhhhh[e]);
}

// Residual start: R = conj(<ij||ab>) + P(ab) sum_c I6(b,c) t(ac,ij)
//                 - P(ij) sum_k I7(k,j) t(ab,ik), diagonals removed
// (ccm_t2_eqn.f90:81-211).  Owner-computes over unique elements.
__global__ void k_resid_init(const SolveDesc *descs, int onebody) {
  const SolveDesc d = descs[blockIdx.y];   // a copy: stores cannot change it, so no reloads
  if (!d.active) return;
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < d.nT; u += (int64_t)gridDim.x * blockDim.x) {
    int c, r, s;
    unique_labels(d, u, c, r, s);
    const int64_t o = d.off_t[c];
    const int nh = d.nhh[c];
    const int64_t pr = d.off_pp[c] + r, hs = d.off_hh[c] + s;
    int a = d.pp_a[pr], b = d.pp_b[pr], i = d.hh_i[hs], j = d.hh_j[hs];
    double2 acc = d.R0[u];
    if (!onebody) { d.R[u] = acc; continue; }
    int vB = d.ob_rowB[pr];   // (a, partner(b))
    if (vB >= 0) {
      double2 tv = d.T[o + (int64_t)(vB >> 2) * nh + s];
      acc = cadd(acc, cscale((double)((vB & 3) - 1), cmul(d.I6off[b - d.A], tv)));
    }
    int vA = d.ob_rowA[pr];   // (b, partner(a))
    if (vA >= 0) {
      double2 tv = d.T[o + (int64_t)(vA >> 2) * nh + s];
      acc = csub(acc, cscale((double)((vA & 3) - 1), cmul(d.I6off[a - d.A], tv)));
    }
    int vJ = d.ob_colJ[hs];   // (i, partner(j))
    if (vJ >= 0) {
      double2 tv = d.T[o + (int64_t)r * nh + (vJ >> 2)];
      acc = csub(acc, cscale((double)((vJ & 3) - 1), cmul(d.I7off[j], tv)));
    }
    int vI = d.ob_colI[hs];   // (j, partner(i))
    if (vI >= 0) {
      double2 tv = d.T[o + (int64_t)r * nh + (vI >> 2)];
      acc = cadd(acc, cscale((double)((vI & 3) - 1), cmul(d.I7off[i], tv)));
    }
    d.R[u] = acc;
  }
}

// Pandya gather T~((a,i),(j,b)) = t(ab,ij) (ccm_t2_eqn.f90:37-62); F32: into
// the FP32 blocks of the BF16x3 residual.  The four positions of an amplitude
// are one 16-byte load.
template <bool F32>
__global__ void k_ring_gather(const SolveDesc *descs) {
  const SolveDesc d = descs[blockIdx.y];   // a copy: stores cannot change it, so no reloads
  if (!d.active) return;
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < d.nT; u += (int64_t)gridDim.x * blockDim.x) {
    const double2 t = d.T[u];
    const int4 p = *(const int4 *)(d.ring4 + 4 * u);
    if constexpr (F32) {
      const float2 tf = make_float2((float)t.x, (float)t.y), mf = make_float2(-tf.x, -tf.y);
      d.Tt32[p.x] = tf;
      d.Tt32[p.y] = mf;
      d.Tt32[p.z] = mf;
      d.Tt32[p.w] = tf;
    } else {
      const double2 m = make_double2(-t.x, -t.y);
      d.Tt[p.x] = t;
      d.Tt[p.y] = m;
      d.Tt[p.z] = m;
      d.Tt[p.w] = t;
    }
  }
}

// Ring scatter with P(ab)P(ij) (ccm_t2_eqn.f90:1133-1185), owner-computes;
// F32: from the FP32 R~ of the BF16x3 residual (summed in FP64).  With hh it
// also adds the hh ladder's Rh first, R = (Rh + R) + ring: the order in which
// the single-stream residual summed them.
template <bool F32>
__global__ void k_ring_scatter(const SolveDesc *descs, int ring, int hh) {
  const SolveDesc d = descs[blockIdx.y];   // a copy: stores cannot change it, so no reloads
  if (!d.active) return;
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < d.nT; u += (int64_t)gridDim.x * blockDim.x) {
    double2 r = d.R[u];
    if (hh) r = cadd(d.Rh[u], r);
    if (ring) {
      const int4 p = *(const int4 *)(d.ring4 + 4 * u);
      double2 r0, r1, r2, r3;
      if constexpr (F32) {
        const float2 a = d.Rt32[p.x], b = d.Rt32[p.y], c = d.Rt32[p.z], e = d.Rt32[p.w];
        r0 = make_double2(a.x, a.y);
        r1 = make_double2(b.x, b.y);
        r2 = make_double2(c.x, c.y);
        r3 = make_double2(e.x, e.y);
      } else {
        r0 = d.Rt[p.x];
        r1 = d.Rt[p.y];
        r2 = d.Rt[p.z];
        r3 = d.Rt[p.w];
      }
      r = cadd(r, csub(csub(cadd(r0, r3), r1), r2));
    }
    d.R[u] = r;
  }
}

// Jacobi step with mixing, t <- 0.75 t + 0.25 R/(dh_i+dh_j-dp_a-dp_b), and the
// history write of diis_t1t2 (ccm_diis.f90:106-126).
__global__ void k_update(const SolveDesc *descs, const int *slot) {
  const SolveDesc d = descs[blockIdx.y];   // a copy: stores cannot change it, so no reloads
  if (!d.active) return;
  double2 *h = d.hist + (int64_t)slot[blockIdx.y] * d.nT;
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < d.nT; u += (int64_t)gridDim.x * blockDim.x) {
    const double2 den = amp_denominator(d, u);
    double2 t = cadd(cscale(0.75, d.T[u]), cscale(0.25, cdiv(d.R[u], den)));
    d.T[u] = t;
    h[u] = t;
  }
}

// DIIS Gram matrix B_pq = sum e_p e_q (no conjugation, ccm_diis.f90:134-147) of the
// error vectors e_p = t_{p+1} - t_p over the ring-buffer history, NE = nstep-1
// error vectors (at most 8).  Partials of the upper triangle go to
// red[block*36 + idx].
template <int NE>
__global__ __launch_bounds__(256) void k_gram_partial(const SolveDesc *descs, const int *slots) {
  const SolveDesc d = descs[blockIdx.y];   // a copy: stores cannot change it, so no reloads
  if (!d.active) return;
  constexpr int NP = NE * (NE + 1) / 2;
  double2 acc[NP];
#pragma unroll
  for (int q = 0; q < NP; ++q) acc[q] = make_double2(0.0, 0.0);
  const int *sl = slots + blockIdx.y * 16;
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < d.nT; u += (int64_t)gridDim.x * blockDim.x) {
    double2 e[NE];
    double2 prev = d.hist[(int64_t)sl[0] * d.nT + u];
#pragma unroll
    for (int p = 0; p < NE; ++p) {
      double2 cur = d.hist[(int64_t)sl[p + 1] * d.nT + u];
      e[p] = csub(cur, prev);
      prev = cur;
    }
    int q = 0;
#pragma unroll
    for (int p = 0; p < NE; ++p)
#pragma unroll
      for (int p2 = p; p2 < NE; ++p2) {
        acc[q] = cadd(acc[q], cmul(e[p], e[p2]));
        ++q;
      }
  }
  // Block sums of the NP values with block_sum's pairwise tree (t + 128, then
  // t + 64 across waves in LDS, then t + 32 .. t + 1 within wave 0, here as a
  // butterfly reduce-scatter, whose pairs are the same), eight values per pass:
  // 3 barriers per pass instead of 9 per value.
  __shared__ double2 sv[8][256];
  const int x = threadIdx.x;
#pragma unroll
  for (int q0 = 0; q0 < NP; q0 += 8) {
#pragma unroll
    for (int t = 0; t < 8; ++t)
      if (q0 + t < NP) sv[t][x] = acc[q0 + t];
    __syncthreads();
    if (x < 128) {
#pragma unroll
      for (int t = 0; t < 8; ++t)
        if (q0 + t < NP) sv[t][x] = cadd(sv[t][x], sv[t][x + 128]);
    }
    __syncthreads();
    if (x < 64) {
      double2 v[8];
#pragma unroll
      for (int t = 0; t < 8; ++t)
        v[t] = q0 + t < NP ? cadd(sv[t][x], sv[t][x + 64]) : make_double2(0.0, 0.0);
#pragma unroll
      for (int st = 0; st < 3; ++st) {   // lane L keeps value (L >> 3) & 7
        const int s = 32 >> st, h = 4 >> st;
        const bool up = (x & s) != 0;
#pragma unroll
        for (int t = 0; t < h; ++t) {
          const double2 lo = v[t], hi = v[t + h];
          const double2 keep = up ? hi : lo, send = up ? lo : hi;
          v[t].x = keep.x + __shfl_xor(send.x, s);
          v[t].y = keep.y + __shfl_xor(send.y, s);
        }
      }
#pragma unroll
      for (int s = 4; s > 0; s >>= 1) {
        v[0].x += __shfl_xor(v[0].x, s);
        v[0].y += __shfl_xor(v[0].y, s);
      }
      const int q = q0 + ((x >> 3) & 7);
      if ((x & 7) == 0 && q < NP) d.red[(int64_t)blockIdx.x * 36 + q] = v[0];
    }
    __syncthreads();   // sv is rewritten by the next pass
  }
}

// Anderson (Pulay) update of the fixed-point map g(T) = R(T)/D(T) with the
// dynamic denominators D = dh_i + dh_j - dp_a - dp_b of the residual:
// f = g(T) - T, y = T + beta f.  y goes to history slot sl[14] (FP64), f to
// histf (FP32), and each block writes its partial sums of Re<f_j, f> over
// the m = sl[15] held vectors j (slots sl[0..m-1], oldest first) and of
// <f, f> (values 0..m, real part in .x), and of the correlation energy of y,
// sum y conj(R0) (value 16).  The coefficients of the next iterate come from
// the host's solve of the bordered Gram system.  No counterpart in the
// reference, whose diis_t1t2 (ccm_diis.f90:74-249) mixes 0.75/0.25 every
// iteration and extrapolates every diis_step iterations.
constexpr int kAAMax = 9;   // held vectors + the new one
__global__ __launch_bounds__(256) void k_aa_update(const SolveDesc *descs, const int *slots, double beta) {
  const SolveDesc d = descs[blockIdx.y];   // a copy: stores cannot change it, so no reloads
  if (!d.active) return;
  const int *sl = slots + blockIdx.y * 16;
  const int m = sl[15];
  double2 *y = d.hist + (int64_t)sl[14] * d.nT;
  float2 *fn = d.histf + (int64_t)sl[14] * d.nT;
  double acc[kAAMax];
#pragma unroll
  for (int q = 0; q < kAAMax; ++q) acc[q] = 0.0;
  double2 ey = make_double2(0.0, 0.0);
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < d.nT; u += (int64_t)gridDim.x * blockDim.x) {
    const double2 den = amp_denominator(d, u);
    const double2 t = d.T[u];
    const double2 f = csub(cdiv(d.R[u], den), t);
    const double2 yv = cadd(t, cscale(beta, f));
    y[u] = yv;
    const float2 f32 = make_float2((float)f.x, (float)f.y);
    fn[u] = f32;
    const double fr = f32.x, fi = f32.y;
#pragma unroll
    for (int q = 0; q < kAAMax - 1; ++q)
      if (q < m) {
        const float2 g = d.histf[(int64_t)sl[q] * d.nT + u];
        acc[q] += (double)g.x * fr + (double)g.y * fi;
      }
    acc[kAAMax - 1] += fr * fr + fi * fi;
    ey = cadd(ey, cmul(yv, cconj(d.R0[u])));
  }
  // fixed-order block sums (pairwise tree over the 256 threads)
  __shared__ double sv[kAAMax + 2][256];
  const int x = threadIdx.x;
#pragma unroll
  for (int q = 0; q < kAAMax; ++q) sv[q][x] = acc[q];
  sv[kAAMax][x] = ey.x;
  sv[kAAMax + 1][x] = ey.y;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (x < w)
#pragma unroll
      for (int q = 0; q < kAAMax + 2; ++q) sv[q][x] += sv[q][x + w];
    __syncthreads();
  }
  if (x < 17) {
    double2 v = make_double2(0.0, 0.0);
    if (x < m) v.x = sv[x][0];
    else if (x == m) v.x = sv[kAAMax - 1][0];
    else if (x == 16) v = make_double2(sv[kAAMax][0], sv[kAAMax + 1][0]);
    d.red[(int64_t)blockIdx.x * 36 + x] = v;
  }
}

// Next iterate of the Anderson update: T = sum_j c_j y_j over the m = sl[15]
// held vectors (slots sl[0..m-1]); real coefficients coef[16 s + j].
__global__ void k_aa_combine(const SolveDesc *descs, const int *slots, const double *coef) {
  const SolveDesc d = descs[blockIdx.y];   // a copy: stores cannot change it, so no reloads
  if (!d.active) return;
  const int *sl = slots + blockIdx.y * 16;
  const double *cf = coef + blockIdx.y * 16;
  const int m = sl[15];
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < d.nT; u += (int64_t)gridDim.x * blockDim.x) {
    double2 acc = cscale(cf[0], d.hist[(int64_t)sl[0] * d.nT + u]);
    for (int p = 1; p < m; ++p) acc = cadd(acc, cscale(cf[p], d.hist[(int64_t)sl[p] * d.nT + u]));
    d.T[u] = acc;
  }
}

// t = sum_p c_p t_p over the first nstep-1 history vectors (ccm_diis.f90:186-212).
__global__ void k_diis_axpy(const SolveDesc *descs, const int *slots, const double2 *coef, int nvec) {
  const SolveDesc d = descs[blockIdx.y];   // a copy: stores cannot change it, so no reloads
  if (!d.active) return;
  const int *sl = slots + blockIdx.y * 16;
  const double2 *cf = coef + blockIdx.y * 16;
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < d.nT; u += (int64_t)gridDim.x * blockDim.x) {
    double2 acc = cmul(cf[0], d.hist[(int64_t)sl[0] * d.nT + u]);
    for (int p = 1; p < nvec; ++p) acc = cadd(acc, cmul(cf[p], d.hist[(int64_t)sl[p] * d.nT + u]));
    d.T[u] = acc;
  }
}

// ---------------------------------------------------------------------------
// Grouped complex GEMMs: C = alpha A B + beta C0, every group a different
// channel (and solve).  A complex multiply-accumulate is four real MFMAs in
// FP32 (re += ar br, re += (-ai) bi, im += ar bi, im += ai br) and three in
// FP64 (Gauss: ar br, ai bi, (ar+ai)(br+bi)).
// ---------------------------------------------------------------------------
// Loads through global-address-space pointers: the compiler then emits
// global_load (not flat_load) and can count outstanding loads precisely.
typedef double gdouble2 __attribute__((ext_vector_type(2)));
typedef float gfloat2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 gload(const double2 *p) {
  gdouble2 v = *(const __attribute__((address_space(1))) gdouble2 *)p;
  return make_double2(v.x, v.y);
}
__device__ __forceinline__ float2 gload(const float2 *p) {
  gfloat2 v = *(const __attribute__((address_space(1))) gfloat2 *)p;
  return make_float2(v.x, v.y);
}
template <typename AT>
__device__ __forceinline__ double2 load_a(const AT *p);
template <>
__device__ __forceinline__ double2 load_a<double2>(const double2 *p) { return gload(p); }
template <>
__device__ __forceinline__ double2 load_a<float2>(const float2 *p) { float2 v = gload(p); return make_double2(v.x, v.y); }

// The FP32 copy of V_pppp has rows of even length and 16-byte aligned starts:
// the element pair (k, k+1) at even k is one aligned 16-byte load.
typedef float gfloat4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void load_pair(const float2 *p, double2 &v0, double2 &v1) {
  const gfloat4 v = *(const __attribute__((address_space(1))) gfloat4 *)p;
  v0 = make_double2(v.x, v.y);
  v1 = make_double2(v.z, v.w);
}

// k index of MFMA step q (0..3) of a 16-wide chunk held by lane group lg: the
// four lane groups side by side (64 contiguous bytes of a row per FP64 load);
// with FP32 storage each lane loads the pair (k, k+1) for steps 2h and 2h+1
// (also 64 contiguous bytes per load).  Any assignment of the chunk's 16 k
// values works as long as A and B use the same one.
template <bool PAIR>
__device__ __forceinline__ int kstep(int q, int lg) {
  return PAIR ? 8 * (q >> 1) + 2 * lg + (q & 1) : 4 * q + lg;
}

// Plain FP64 VALU GEMM with the same contract, kept as an independent check of
// the MFMA lane layouts (selected with CCGPU_GEMM=valu).
__global__ __launch_bounds__(256) void k_zgemm_valu(const GemmGroup *__restrict__ groups,
                                                    const GemmTile *__restrict__ tiles,
                                                    const SolveDesc *__restrict__ descs,
                                                    const int *__restrict__ gsolve) {
  const GemmTile tl = tiles[blockIdx.x];
  if (descs && !descs[gsolve[tl.g]].active) return;
  const GemmGroup g = groups[tl.g];
  const double2 *A = (const double2 *)g.A;
  __shared__ double2 As[16][64 + 1];
  __shared__ double2 Bs[16][64 + 1];
  const int tid = threadIdx.x;
  const int tm = (tid >> 4) * 4, tn = (tid & 15) * 4;
  double2 acc[4][4];
  for (int x = 0; x < 4; ++x)
    for (int y = 0; y < 4; ++y) acc[x][y] = make_double2(0.0, 0.0);
  for (int k0 = 0; k0 < g.K; k0 += 16) {
    for (int e = tid; e < 64 * 16; e += 256) {
      int m = e >> 4, k = e & 15;
      int gm = tl.m0 + m, gk = k0 + k;
      As[k][m] = (gm < g.M && gk < g.K) ? A[(int64_t)gm * g.lda + gk] : make_double2(0.0, 0.0);
      int kb = e >> 6, n = e & 63;
      int gk2 = k0 + kb, gn = tl.n0 + n;
      Bs[kb][n] = (gk2 < g.K && gn < g.N) ? g.B[(int64_t)gk2 * g.ldb + gn] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    for (int k = 0; k < 16; ++k)
      for (int x = 0; x < 4; ++x)
        for (int y = 0; y < 4; ++y) acc[x][y] = cadd(acc[x][y], cmul(As[k][tm + x], Bs[k][tn + y]));
    __syncthreads();
  }
  for (int x = 0; x < 4; ++x)
    for (int y = 0; y < 4; ++y) {
      int m = tl.m0 + tm + x, n = tl.n0 + tn + y;
      if (m < g.M && n < g.N) {
        double2 v = cscale(g.alpha, acc[x][y]);
        if (g.beta != 0.0) v = cadd(v, cscale(g.beta, g.C0[(int64_t)m * g.ldc0 + n]));
        g.C[(int64_t)m * g.ldc + n] = v;
      }
    }
}

// Grouped complex GEMM with direct-to-register operand loads.  The streamed
// operand A (one read per element) never touches LDS; B (a channel's amplitude
// block, shared by every row tile of the channel) comes through L2.  The four
// wavefronts of a workgroup stack along M, each owning MB 16-row blocks and the
// workgroup's NB 16-column blocks.  The sum over k of one 16x16x4 MFMA is
// permutation invariant, so MFMA q of a 16-wide k chunk takes k0+4q..k0+4q+3,
// lane group g = lane>>4 holding k0+4q+g: each load instruction then reads 64
// contiguous bytes of every row (the four lane groups side by side), half the
// cache lines per instruction of a per-lane-contiguous layout, which matters
// because these loads are limited by the L1 tag path.
// FAM only names the instantiation (1 = pp ladder, 2 = pp ladder on the FP32
// copy of V with FP64 arithmetic) so profiles can tell the pp-ladder launches
// from other families with the same tile grid.
template <int MB, int NB, bool F32, typename AT, int FAM = 0>
__global__ __launch_bounds__(256) void k_zgemm_dl(const GemmGroup *__restrict__ groups,
                                                  const GemmTile *__restrict__ tiles,
                                                  const SolveDesc *__restrict__ descs,
                                                  const int *__restrict__ gsolve) {
  typedef typename std::conditional<F32, float, double>::type real;
  typedef typename std::conditional<F32, f4, d4>::type acc4;
  const GemmTile tl = tiles[blockIdx.x];
  if (descs && !descs[gsolve[tl.g]].active) return;
  const GemmGroup g = groups[tl.g];
  const AT *A = (const AT *)g.A;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int li = lane & 15, lg = lane >> 4;
  const int mw = tl.m0 + w * (MB * 16);
  // FP32: cre, cim with four real MFMAs per complex MAC.  FP64: three products
  // (Gauss): cre = sum ar br, cim = sum ai bi, cs = sum (ar+ai)(br+bi); then
  // re = cre - cim and im = cs - cre - cim, one MFMA in four saved.
  acc4 cre[MB][NB], cim[MB][NB], cs[MB][NB];
#pragma unroll
  for (int x = 0; x < MB; ++x)
#pragma unroll
    for (int y = 0; y < NB; ++y) {
      cre[x][y] = (acc4){0, 0, 0, 0};
      cim[x][y] = (acc4){0, 0, 0, 0};
      cs[x][y] = (acc4){0, 0, 0, 0};
    }
  double2 a0[MB][4], b0[NB][4], a1[MB][4], b1[NB][4];
  const int kb = tl.kb, ke = min(tl.ke, g.K);
  // branch-free loads: out-of-range indices are clamped to a valid element and
  // the value is zeroed, so all loads of a chunk issue back to back
  constexpr bool PAIR = std::is_same<AT, float2>::value;
  auto load = [&](double2 (&a)[MB][4], double2 (&b)[NB][4], int k0) {
#pragma unroll
    for (int x = 0; x < MB; ++x) {
      const int m = mw + x * 16 + li;
      const int mc = min(m, g.M - 1);
      if constexpr (PAIR) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int k = k0 + 8 * h + 2 * lg;
          double2 v0, v1;
          load_pair((const float2 *)A + (int64_t)mc * g.lda + min(k, (ke - 1) & ~1), v0, v1);
          const bool ok0 = m < g.M && k < ke, ok1 = m < g.M && k + 1 < ke;
          a[x][2 * h] = make_double2(ok0 ? v0.x : 0.0, ok0 ? v0.y : 0.0);
          a[x][2 * h + 1] = make_double2(ok1 ? v1.x : 0.0, ok1 ? v1.y : 0.0);
        }
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int k = k0 + kstep<false>(q, lg);
          const double2 v = load_a<AT>(A + (int64_t)mc * g.lda + min(k, ke - 1));
          const bool ok = m < g.M && k < ke;
          a[x][q] = make_double2(ok ? v.x : 0.0, ok ? v.y : 0.0);
        }
      }
    }
#pragma unroll
    for (int y = 0; y < NB; ++y) {
      const int n = tl.n0 + y * 16 + li;
      const int nc = min(n, g.N - 1);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int k = k0 + kstep<PAIR>(q, lg);
        const double2 v = gload(g.B + (int64_t)min(k, ke - 1) * g.ldb + nc);
        const bool ok = n < g.N && k < ke;
        b[y][q] = make_double2(ok ? v.x : 0.0, ok ? v.y : 0.0);
      }
    }
  };
  auto compute = [&](const double2 (&a)[MB][4], const double2 (&b)[NB][4]) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
#pragma unroll
      for (int x = 0; x < MB; ++x) {
        const real ar = (real)a[x][q].x, ai = (real)a[x][q].y;
#pragma unroll
        for (int y = 0; y < NB; ++y) {
          const real br = (real)b[y][q].x, bi = (real)b[y][q].y;
          if constexpr (F32) {
            cre[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, br, cre[x][y], 0, 0, 0);
            cre[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(-ai, bi, cre[x][y], 0, 0, 0);
            cim[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, bi, cim[x][y], 0, 0, 0);
            cim[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, br, cim[x][y], 0, 0, 0);
          } else {
            cre[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, cre[x][y], 0, 0, 0);
            cim[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, bi, cim[x][y], 0, 0, 0);
            cs[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar + ai, br + bi, cs[x][y], 0, 0, 0);
          }
        }
      }
    }
  };
  // two statically indexed register buffers: the next 16-wide k chunk is in
  // flight while the current one feeds the MFMAs
  load(a0, b0, kb);
  for (int k0 = kb; k0 < ke; k0 += 32) {
    if (k0 + 16 < ke) load(a1, b1, k0 + 16);
    compute(a0, b0);
    if (k0 + 16 >= ke) break;
    if (k0 + 32 < ke) load(a0, b0, k0 + 32);
    compute(a1, b1);
  }
  // result layouts of the 16x16x4 MFMAs: register r of lane l holds column l&15
  // and row (l>>4)+4r for f64, row 4(l>>4)+r for f32
#pragma unroll
  for (int x = 0; x < MB; ++x)
#pragma unroll
    for (int y = 0; y < NB; ++y)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = F32 ? (lg * 4 + r) : (lg + 4 * r);
        const int m = mw + x * 16 + row;
        const int n = tl.n0 + y * 16 + li;
        double vr, vi;
        if constexpr (F32) {
          vr = (double)cre[x][y][r];
          vi = (double)cim[x][y][r];
        } else {
          vr = cre[x][y][r] - cim[x][y][r];
          vi = cs[x][y][r] - cre[x][y][r] - cim[x][y][r];
        }
        if (m < g.M && n < g.N && g.nsplit > 1) {
          g.P[((int64_t)tl.ks * g.M + m) * g.N + n] = make_double2(vr, vi);
        } else if (m < g.M && n < g.N) {
          double2 v = make_double2(g.alpha * vr, g.alpha * vi);
          if (g.beta != 0.0) {
            double2 c0 = g.C0[(int64_t)m * g.ldc0 + n];
            v.x += g.beta * c0.x;
            v.y += g.beta * c0.y;
          }
          g.C[(int64_t)m * g.ldc + n] = v;
        }
      }
}

// The same grouped GEMM with the B operand staged once per workgroup in LDS
// (double-buffered, one barrier per 16-wide k chunk) instead of being loaded
// by each of the four waves: a quarter of the B traffic and no B registers.
// The k order of every sum is the one of k_zgemm_dl, so results are identical.
template <int MB, int NB, bool F32, typename AT, int FAM = 0>
__global__ __launch_bounds__(256) void k_zgemm_lb(const GemmGroup *__restrict__ groups,
                                                  const GemmTile *__restrict__ tiles,
                                                  const SolveDesc *__restrict__ descs,
                                                  const int *__restrict__ gsolve) {
  typedef typename std::conditional<F32, float, double>::type real;
  typedef typename std::conditional<F32, f4, d4>::type acc4;
  constexpr int BC = 16 * NB;                 // tile columns
  constexpr int BE = 16 * BC;                 // B elements per chunk
  constexpr int BPT = (BE + 255) / 256;       // B elements per thread
  __shared__ double2 Bs[2][16][BC + 1];
  const GemmTile tl = tiles[blockIdx.x];
  if (descs && !descs[gsolve[tl.g]].active) return;
  const GemmGroup g = groups[tl.g];
  const AT *A = (const AT *)g.A;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int li = lane & 15, lg = lane >> 4;
  const int mw = tl.m0 + w * (MB * 16);
  acc4 cre[MB][NB], cim[MB][NB], cs[MB][NB];
#pragma unroll
  for (int x = 0; x < MB; ++x)
#pragma unroll
    for (int y = 0; y < NB; ++y) {
      cre[x][y] = (acc4){0, 0, 0, 0};
      cim[x][y] = (acc4){0, 0, 0, 0};
      cs[x][y] = (acc4){0, 0, 0, 0};
    }
  double2 a0[MB][4], a1[MB][4], bn[BPT];
  const int kb = tl.kb, ke = min(tl.ke, g.K);
  constexpr bool PAIR = std::is_same<AT, float2>::value;
  auto load_a = [&](double2 (&a)[MB][4], int k0) {
#pragma unroll
    for (int x = 0; x < MB; ++x) {
      const int m = mw + x * 16 + li;
      const int mc = min(m, g.M - 1);
      if constexpr (PAIR) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int k = k0 + 8 * h + 2 * lg;
          double2 v0, v1;
          load_pair((const float2 *)A + (int64_t)mc * g.lda + min(k, (ke - 1) & ~1), v0, v1);
          const bool ok0 = m < g.M && k < ke, ok1 = m < g.M && k + 1 < ke;
          a[x][2 * h] = make_double2(ok0 ? v0.x : 0.0, ok0 ? v0.y : 0.0);
          a[x][2 * h + 1] = make_double2(ok1 ? v1.x : 0.0, ok1 ? v1.y : 0.0);
        }
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int k = k0 + kstep<false>(q, lg);
          const double2 v = ccgpu::load_a<AT>(A + (int64_t)mc * g.lda + min(k, ke - 1));
          const bool ok = m < g.M && k < ke;
          a[x][q] = make_double2(ok ? v.x : 0.0, ok ? v.y : 0.0);
        }
      }
    }
  };
  // cooperative B chunk load: element e = threadIdx.x + 256 j -> (k row, column)
  auto load_b = [&](int k0) {
#pragma unroll
    for (int j = 0; j < BPT; ++j) {
      const int e = threadIdx.x + 256 * j;
      const int kr = e / BC, c = e % BC;
      const int k = k0 + kr, n = tl.n0 + c;
      const double2 v = gload(g.B + (int64_t)min(k, ke - 1) * g.ldb + min(n, g.N - 1));
      const bool ok = e < BE && k < ke && n < g.N;
      bn[j] = make_double2(ok ? v.x : 0.0, ok ? v.y : 0.0);
    }
  };
  auto store_b = [&](int buf) {
#pragma unroll
    for (int j = 0; j < BPT; ++j) {
      const int e = threadIdx.x + 256 * j;
      if (e < BE) Bs[buf][e / BC][e % BC] = bn[j];
    }
  };
  auto compute = [&](const double2 (&a)[MB][4], int buf) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      double2 b[NB];
#pragma unroll
      for (int y = 0; y < NB; ++y) b[y] = Bs[buf][kstep<PAIR>(q, lg)][y * 16 + li];
#pragma unroll
      for (int x = 0; x < MB; ++x) {
        const real ar = (real)a[x][q].x, ai = (real)a[x][q].y;
#pragma unroll
        for (int y = 0; y < NB; ++y) {
          const real br = (real)b[y].x, bi = (real)b[y].y;
          if constexpr (F32) {
            cre[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, br, cre[x][y], 0, 0, 0);
            cre[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(-ai, bi, cre[x][y], 0, 0, 0);
            cim[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, bi, cim[x][y], 0, 0, 0);
            cim[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, br, cim[x][y], 0, 0, 0);
          } else {
            cre[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, cre[x][y], 0, 0, 0);
            cim[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, bi, cim[x][y], 0, 0, 0);
            cs[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar + ai, br + bi, cs[x][y], 0, 0, 0);
          }
        }
      }
    }
  };
  // prologue: chunk kb in LDS buffer 0 and registers a0
  load_b(kb);
  load_a(a0, kb);
  store_b(0);
  __syncthreads();
  int buf = 0;
  for (int k0 = kb; k0 < ke; k0 += 32) {
    const bool more1 = k0 + 16 < ke;
    if (more1) {
      load_b(k0 + 16);
      load_a(a1, k0 + 16);
    }
    compute(a0, buf);
    if (!more1) break;
    store_b(buf ^ 1);
    __syncthreads();
    buf ^= 1;
    const bool more2 = k0 + 32 < ke;
    if (more2) {
      load_b(k0 + 32);
      load_a(a0, k0 + 32);
    }
    compute(a1, buf);
    if (!more2) break;
    store_b(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
#pragma unroll
  for (int x = 0; x < MB; ++x)
#pragma unroll
    for (int y = 0; y < NB; ++y)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = F32 ? (lg * 4 + r) : (lg + 4 * r);
        const int m = mw + x * 16 + row;
        const int n = tl.n0 + y * 16 + li;
        double vr, vi;
        if constexpr (F32) {
          vr = (double)cre[x][y][r];
          vi = (double)cim[x][y][r];
        } else {
          vr = cre[x][y][r] - cim[x][y][r];
          vi = cs[x][y][r] - cre[x][y][r] - cim[x][y][r];
        }
        if (m < g.M && n < g.N && g.nsplit > 1) {
          g.P[((int64_t)tl.ks * g.M + m) * g.N + n] = make_double2(vr, vi);
        } else if (m < g.M && n < g.N) {
          double2 v = make_double2(g.alpha * vr, g.alpha * vi);
          if (g.beta != 0.0) {
            double2 c0 = g.C0[(int64_t)m * g.ldc0 + n];
            v.x += g.beta * c0.x;
            v.y += g.beta * c0.y;
          }
          g.C[(int64_t)m * g.ldc + n] = v;
        }
      }
}

// The same grouped GEMM with both operands staged in LDS: the 64 x 16 A chunk
// of a workgroup (MB = 1) is read row-coalesced from global memory (16
// consecutive threads read one row's 16 consecutive k, 1-2 cache lines per
// quarter-wave) and written to LDS, from where each lane reads its MFMA
// operand; B as in k_zgemm_lb.  The k order is that of k_zgemm_dl and
// k_zgemm_lb, so results are identical across the three forms.
template <int MB, int NB, bool F32, typename AT, int FAM = 0>
__global__ __launch_bounds__(256) void k_zgemm_la(const GemmGroup *__restrict__ groups,
                                                  const GemmTile *__restrict__ tiles,
                                                  const SolveDesc *__restrict__ descs,
                                                  const int *__restrict__ gsolve) {
  static_assert(MB == 1, "A chunk of 64 rows");
  typedef typename std::conditional<F32, float, double>::type real;
  typedef typename std::conditional<F32, f4, d4>::type acc4;
  constexpr bool PAIR = std::is_same<AT, float2>::value;
  constexpr int BC = 16 * NB;                 // tile columns
  constexpr int BE = 16 * BC;                 // B elements per chunk
  constexpr int BPT = (BE + 255) / 256;       // B elements per thread
  __shared__ double2 Bs[2][16][BC + 1];
  __shared__ double2 As[2][64][17];
  const GemmTile tl = tiles[blockIdx.x];
  if (descs && !descs[gsolve[tl.g]].active) return;
  const GemmGroup g = groups[tl.g];
  const AT *A = (const AT *)g.A;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int li = lane & 15, lg = lane >> 4;
  const int mw = tl.m0 + w * 16;
  acc4 cre[NB], cim[NB], cs[NB];
#pragma unroll
  for (int y = 0; y < NB; ++y) {
    cre[y] = (acc4){0, 0, 0, 0};
    cim[y] = (acc4){0, 0, 0, 0};
    cs[y] = (acc4){0, 0, 0, 0};
  }
  double2 an[4], bn[BPT];
  const int kb = tl.kb, ke = min(tl.ke, g.K);
  // A chunk element e = threadIdx.x + 256 j: row e >> 4 of the tile, k0 + (e & 15)
  auto load_a = [&](int k0) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int e = threadIdx.x + 256 * j;
      const int m = tl.m0 + (e >> 4), k = k0 + (e & 15);
      const double2 v = ccgpu::load_a<AT>(A + (int64_t)min(m, g.M - 1) * g.lda + min(k, ke - 1));
      const bool ok = m < g.M && k < ke;
      an[j] = make_double2(ok ? v.x : 0.0, ok ? v.y : 0.0);
    }
  };
  auto store_a = [&](int buf) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int e = threadIdx.x + 256 * j;
      As[buf][e >> 4][e & 15] = an[j];
    }
  };
  auto load_b = [&](int k0) {
#pragma unroll
    for (int j = 0; j < BPT; ++j) {
      const int e = threadIdx.x + 256 * j;
      const int kr = e / BC, c = e % BC;
      const int k = k0 + kr, n = tl.n0 + c;
      const double2 v = gload(g.B + (int64_t)min(k, ke - 1) * g.ldb + min(n, g.N - 1));
      const bool ok = e < BE && k < ke && n < g.N;
      bn[j] = make_double2(ok ? v.x : 0.0, ok ? v.y : 0.0);
    }
  };
  auto store_b = [&](int buf) {
#pragma unroll
    for (int j = 0; j < BPT; ++j) {
      const int e = threadIdx.x + 256 * j;
      if (e < BE) Bs[buf][e / BC][e % BC] = bn[j];
    }
  };
  auto compute = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int kq = kstep<PAIR>(q, lg);
      const double2 a = As[buf][w * 16 + li][kq];
      const real ar = (real)a.x, ai = (real)a.y;
#pragma unroll
      for (int y = 0; y < NB; ++y) {
        const double2 b = Bs[buf][kq][y * 16 + li];
        const real br = (real)b.x, bi = (real)b.y;
        if constexpr (F32) {
          cre[y] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, br, cre[y], 0, 0, 0);
          cre[y] = __builtin_amdgcn_mfma_f32_16x16x4f32(-ai, bi, cre[y], 0, 0, 0);
          cim[y] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, bi, cim[y], 0, 0, 0);
          cim[y] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, br, cim[y], 0, 0, 0);
        } else {
          cre[y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, cre[y], 0, 0, 0);
          cim[y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, bi, cim[y], 0, 0, 0);
          cs[y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar + ai, br + bi, cs[y], 0, 0, 0);
        }
      }
    }
  };
  load_b(kb);
  load_a(kb);
  store_b(0);
  store_a(0);
  __syncthreads();
  int buf = 0;
  for (int k0 = kb; k0 < ke; k0 += 16) {
    const bool more = k0 + 16 < ke;
    if (more) {
      load_b(k0 + 16);
      load_a(k0 + 16);
    }
    compute(buf);
    if (!more) break;
    store_b(buf ^ 1);
    store_a(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
#pragma unroll
  for (int y = 0; y < NB; ++y)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = F32 ? (lg * 4 + r) : (lg + 4 * r);
      const int m = mw + row;
      const int n = tl.n0 + y * 16 + li;
      double vr, vi;
      if constexpr (F32) {
        vr = (double)cre[y][r];
        vi = (double)cim[y][r];
      } else {
        vr = cre[y][r] - cim[y][r];
        vi = cs[y][r] - cre[y][r] - cim[y][r];
      }
      if (m < g.M && n < g.N && g.nsplit > 1) {
        g.P[((int64_t)tl.ks * g.M + m) * g.N + n] = make_double2(vr, vi);
      } else if (m < g.M && n < g.N) {
        double2 v = make_double2(g.alpha * vr, g.alpha * vi);
        if (g.beta != 0.0) {
          double2 c0 = g.C0[(int64_t)m * g.ldc0 + n];
          v.x += g.beta * c0.x;
          v.y += g.beta * c0.y;
        }
        g.C[(int64_t)m * g.ldc + n] = v;
      }
    }
}

// BF16x3 grouped complex GEMM: every real operand x (FP64 or FP32 in memory)
// is split once into two bfloat16 numbers, xh = bf16(x) and xl = bf16(x - xh)
// (both rounded to nearest, x - xh exact in FP32), and a product is summed as
// xh yh + xh yl + xl yh on v_mfma_f32_16x16x32_bf16 with FP32 accumulation; the
// dropped xl yl and the two roundings leave |x y - sum| <= 3 * 2^-16 |x y|.
// A complex MAC packs (re, im) as neighbouring k of one bf16 MFMA: lane slot
// 2c holds ar and 2c+1 ai of complex index c, the B operand (br, -bi) gives the
// real part and (bi, br) the imaginary part, so 16 complex k are one MFMA and
// three MFMAs per part cover the split.
// Both operands are staged in LDS already split and in MFMA lane order: the
// 64*MB x 16 A chunk row-coalesced (16 threads per row), B column-wise with
// the pair (br, -bi); each wave rebuilds (bi, br) from it with a rotate and a
// sign flip.  Loads run two 16-wide k chunks ahead of the MFMAs in registers,
// issued unconditionally (indices clamped, values zeroed past the k range) so
// the count of loads in flight is the same on every path, and the barriers
// wait for LDS only (a workgroup fence would drain the loads in flight).
// The k order is fixed by the tile, so repeated runs give the same bits.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void bf3_split(float x, float y, unsigned &hi, unsigned &lo) {
  const __bf16 xh = (__bf16)x, yh = (__bf16)y;
  const __bf16 xl = (__bf16)(x - (float)xh), yl = (__bf16)(y - (float)yh);
  hi = __builtin_bit_cast(unsigned, (bf16x2){xh, yh});
  lo = __builtin_bit_cast(unsigned, (bf16x2){xl, yl});
}
// (br, -bi) -> (bi, br): swap the halves, then flip the sign of the new low half
__device__ __forceinline__ u32x4 bf3_im_pairs(u32x4 v) {
  u32x4 r;
#pragma unroll
  for (int q = 0; q < 4; ++q) r[q] = ((v[q] >> 16) | (v[q] << 16)) ^ 0x8000u;
  return r;
}
// T split once per residual for the BF16x3 pp ladder (B operand (br, -bi)):
// the bits the ladder's in-kernel split gives, computed once per amplitude
// instead of once per wave and row tile
__global__ void k_t_split(const SolveDesc *descs) {
  const SolveDesc d = descs[blockIdx.y];   // a copy: stores cannot change it, so no reloads
  if (!d.active) return;
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < d.nT; u += (int64_t)gridDim.x * blockDim.x) {
    const double2 b = d.T[u];
    unsigned h, l;
    bf3_split((float)b.x, -(float)b.y, h, l);
    const u32x4 im = bf3_im_pairs(u32x4{h, l, 0u, 0u});
    d.Tbf[u] = make_uint4(h, l, im[0], im[1]);
  }
}

__device__ __forceinline__ void bf3_lds_sync() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt((15) | (3 << 14) | (7 << 4) | (0 << 8));   // lgkmcnt(0) only
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// F32IO: B, C and C0 are FP32 blocks (float2 behind the group's pointers).
template <int MB, int NB, typename AT, bool F32IO = false>
__global__ __launch_bounds__(256) void k_zgemm_bf3(const GemmGroup *__restrict__ groups,
                                                   const GemmTile *__restrict__ tiles,
                                                   const SolveDesc *__restrict__ descs,
                                                   const int *__restrict__ gsolve) {
  constexpr int ROWS = 64 * MB;               // tile rows
  constexpr int BC = 16 * NB;                 // tile columns
  constexpr int APT = ROWS * 16 / 256;        // A elements per thread per chunk
  constexpr int BPT = (16 * BC + 255) / 256;  // B elements per thread per chunk
  constexpr int LD = 20;                      // dwords per LDS row (16 + 4 pad: conflict-free b128 reads)
  __shared__ __align__(16) unsigned Ah[2][ROWS][LD], Al[2][ROWS][LD];
  __shared__ __align__(16) unsigned Bh[2][BC][LD], Bl[2][BC][LD];
  const GemmTile tl = tiles[blockIdx.x];
  if (descs && !descs[gsolve[tl.g]].active) return;
  const GemmGroup g = groups[tl.g];
  const AT *A = (const AT *)g.A;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int li = lane & 15, lg = lane >> 4;
  f4 cre[MB][NB], cim[MB][NB];
#pragma unroll
  for (int x = 0; x < MB; ++x)
#pragma unroll
    for (int y = 0; y < NB; ++y) {
      cre[x][y] = (f4){0, 0, 0, 0};
      cim[x][y] = (f4){0, 0, 0, 0};
    }
  const int kb = tl.kb, ke = min(tl.ke, g.K);
  // A element e = threadIdx.x + 256 j: tile row e >> 4, k0 + (e & 15);
  // B element e: k0 + e / BC, tile column e % BC.  The loads only fetch (from
  // clamped, valid addresses); elements outside the block are zeroed when the
  // chunk is split into LDS, so no instruction touches a load's result before
  // the chunk is needed.
  typedef typename std::conditional<F32IO, float2, double2>::type BT;
  const BT *Bp = (const BT *)g.B;
  auto load = [&](AT (&a)[APT], BT (&b)[BPT], int k0) {
#pragma unroll
    for (int j = 0; j < APT; ++j) {
      const int e = threadIdx.x + 256 * j;
      const int m = tl.m0 + (e >> 4), k = k0 + (e & 15);
      a[j] = gload(A + (int64_t)min(m, g.M - 1) * g.lda + min(k, ke - 1));
    }
#pragma unroll
    for (int j = 0; j < BPT; ++j) {
      const int e = threadIdx.x + 256 * j;
      const int k = k0 + e / BC, n = tl.n0 + e % BC;
      b[j] = gload(Bp + (int64_t)min(k, ke - 1) * g.ldb + min(n, g.N - 1));
    }
  };
  auto store = [&](const AT (&a)[APT], const BT (&b)[BPT], int k0, int buf) {
#pragma unroll
    for (int j = 0; j < APT; ++j) {
      const int e = threadIdx.x + 256 * j;
      const bool ok = tl.m0 + (e >> 4) < g.M && k0 + (e & 15) < ke;
      unsigned h, l;
      bf3_split(ok ? (float)a[j].x : 0.f, ok ? (float)a[j].y : 0.f, h, l);
      Ah[buf][e >> 4][e & 15] = h;
      Al[buf][e >> 4][e & 15] = l;
    }
#pragma unroll
    for (int j = 0; j < BPT; ++j) {
      const int e = threadIdx.x + 256 * j;
      if (16 * BC % 256 == 0 || e < 16 * BC) {
        const bool ok = k0 + e / BC < ke && tl.n0 + e % BC < g.N;
        unsigned h, l;
        bf3_split(ok ? (float)b[j].x : 0.f, ok ? -(float)b[j].y : 0.f, h, l);
        Bh[buf][e % BC][e / BC] = h;
        Bl[buf][e % BC][e / BC] = l;
      }
    }
  };
  auto compute = [&](int buf) {
    u32x4 ah[MB], al[MB];
#pragma unroll
    for (int x = 0; x < MB; ++x) {
      const int r = (w * MB + x) * 16 + li;
      ah[x] = *(const u32x4 *)&Ah[buf][r][4 * lg];
      al[x] = *(const u32x4 *)&Al[buf][r][4 * lg];
    }
#pragma unroll
    for (int y = 0; y < NB; ++y) {
      const u32x4 brh = *(const u32x4 *)&Bh[buf][y * 16 + li][4 * lg];
      const u32x4 brl = *(const u32x4 *)&Bl[buf][y * 16 + li][4 * lg];
      const u32x4 bih = bf3_im_pairs(brh), bil = bf3_im_pairs(brl);
#pragma unroll
      for (int x = 0; x < MB; ++x) {
        const bf16x8 xh = __builtin_bit_cast(bf16x8, ah[x]), xl = __builtin_bit_cast(bf16x8, al[x]);
        cre[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xl, __builtin_bit_cast(bf16x8, brh), cre[x][y], 0, 0, 0);
        cre[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh, __builtin_bit_cast(bf16x8, brl), cre[x][y], 0, 0, 0);
        cre[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh, __builtin_bit_cast(bf16x8, brh), cre[x][y], 0, 0, 0);
        cim[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xl, __builtin_bit_cast(bf16x8, bih), cim[x][y], 0, 0, 0);
        cim[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh, __builtin_bit_cast(bf16x8, bil), cim[x][y], 0, 0, 0);
        cim[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh, __builtin_bit_cast(bf16x8, bih), cim[x][y], 0, 0, 0);
      }
    }
  };
  AT a0[APT], a1[APT];
  BT b0[BPT], b1[BPT];
  // pairs of chunks; the loads and LDS stores of a pair run unconditionally
  // (past the k range they fetch valid addresses and store zeros), only the
  // MFMAs of a second chunk past the end are skipped
  const int nchunk = (ke - kb + 15) >> 4;
  load(a0, b0, kb);
  load(a1, b1, kb + 16);
  store(a0, b0, kb, 0);
  bf3_lds_sync();
  for (int c = 0, k0 = kb; c < nchunk; c += 2, k0 += 32) {
    // chunk k0 in LDS buffer 0, chunk k0 + 16 in flight in a1/b1
    load(a0, b0, k0 + 32);
    compute(0);
    store(a1, b1, k0 + 16, 1);
    bf3_lds_sync();
    // chunk k0 + 16 in LDS buffer 1, chunk k0 + 32 in flight in a0/b0
    load(a1, b1, k0 + 48);
    if (c + 1 < nchunk) compute(1);
    store(a0, b0, k0 + 32, 0);
    bf3_lds_sync();
  }
  // result layout of the 16x16 FP32 MFMAs: register r of lane l holds column
  // l & 15 and row 4 (l >> 4) + r
#pragma unroll
  for (int x = 0; x < MB; ++x)
#pragma unroll
    for (int y = 0; y < NB; ++y)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = tl.m0 + (w * MB + x) * 16 + lg * 4 + r;
        const int n = tl.n0 + y * 16 + li;
        const double vr = (double)cre[x][y][r], vi = (double)cim[x][y][r];
        if (m < g.M && n < g.N && g.nsplit > 1) {
          g.P[((int64_t)tl.ks * g.M + m) * g.N + n] = make_double2(vr, vi);
        } else if (m < g.M && n < g.N) {
          double2 v = make_double2(g.alpha * vr, g.alpha * vi);
          if constexpr (F32IO) {
            if (g.beta != 0.0) {
              const float2 c0 = ((const float2 *)g.C0)[(int64_t)m * g.ldc0 + n];
              v.x += g.beta * c0.x;
              v.y += g.beta * c0.y;
            }
            ((float2 *)g.C)[(int64_t)m * g.ldc + n] = make_float2((float)v.x, (float)v.y);
          } else {
            if (g.beta != 0.0) {
              double2 c0 = g.C0[(int64_t)m * g.ldc0 + n];
              v.x += g.beta * c0.x;
              v.y += g.beta * c0.y;
            }
            g.C[(int64_t)m * g.ldc + n] = v;
          }
        }
      }
}

// Thin GEMM for groups with N <= 4 columns (most pp-ladder channels with few
// hole pairs): FP64 VALU instead of 16-column MFMA tiles, which would be 75% or
// more padding.  A tile is 16 rows: each wave owns 4 rows, its 64 lanes walk
// the k range 64 consecutive elements at a time (one coalesced 1 KB read of a
// row per instruction), each lane reuses its B row (4 columns) for the 4 rows,
// and the 64 partial sums of every output are combined with a fixed xor tree.
// F32IO: B, C and C0 are FP32 blocks (float2 behind the group's pointers).
template <typename AT, int FAM, bool F32IO = false>
__global__ __launch_bounds__(256) void k_zgemm_thin(const GemmGroup *__restrict__ groups,
                                                    const GemmTile *__restrict__ tiles,
                                                    const SolveDesc *__restrict__ descs,
                                                    const int *__restrict__ gsolve) {
  const GemmTile tl = tiles[blockIdx.x];
  if (descs && !descs[gsolve[tl.g]].active) return;
  const GemmGroup g = groups[tl.g];
  const AT *A = (const AT *)g.A;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int m0 = tl.m0 + 4 * w;
  const int kb = tl.kb, ke = min(tl.ke, g.K);
  double2 acc[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int n = 0; n < 4; ++n) acc[r][n] = make_double2(0.0, 0.0);
  for (int k0 = kb; k0 < ke; k0 += 64) {
    const int k = k0 + lane;
    const bool okk = k < ke;
    const int kc = min(k, ke - 1);
    double2 b[4], a[4];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      double2 v;
      if constexpr (F32IO) {
        const float2 f = gload((const float2 *)g.B + (int64_t)kc * g.ldb + min(n, g.N - 1));
        v = make_double2(f.x, f.y);
      } else {
        v = gload(g.B + (int64_t)kc * g.ldb + min(n, g.N - 1));
      }
      const bool ok = okk && n < g.N;
      b[n] = make_double2(ok ? v.x : 0.0, ok ? v.y : 0.0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int m = m0 + r;
      const double2 v = load_a<AT>(A + (int64_t)min(m, g.M - 1) * g.lda + kc);
      const bool ok = okk && m < g.M;
      a[r] = make_double2(ok ? v.x : 0.0, ok ? v.y : 0.0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int n = 0; n < 4; ++n) {
        acc[r][n].x = fma(a[r].x, b[n].x, fma(-a[r].y, b[n].y, acc[r][n].x));
        acc[r][n].y = fma(a[r].x, b[n].y, fma(a[r].y, b[n].x, acc[r][n].y));
      }
  }
  // Butterfly reduce-scatter of the 16 sums over the 64 lanes: at the steps
  // s = 32, 16, 8, 4 a lane keeps half of its values and adds its partner's
  // copy of them, so lane L ends with value j = (L >> 2) & 15; the steps 2, 1
  // finish it.  Every value goes through the same pairwise tree (lane, lane ^ s)
  // as a full xor reduction of each value, so the sums are bit-identical to
  // that, with 17 exchanges per component instead of 96.
  double2 v[16];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int n = 0; n < 4; ++n) v[4 * r + n] = acc[r][n];
#pragma unroll
  for (int st = 0; st < 4; ++st) {
    const int s = 32 >> st, h = 8 >> st;
    const bool up = (lane & s) != 0;
#pragma unroll
    for (int t = 0; t < h; ++t) {
      const double2 lo = v[t], hi = v[t + h];
      const double2 keep = up ? hi : lo, send = up ? lo : hi;
      v[t].x = keep.x + __shfl_xor(send.x, s);
      v[t].y = keep.y + __shfl_xor(send.y, s);
    }
  }
#pragma unroll
  for (int s = 2; s > 0; s >>= 1) {
    v[0].x += __shfl_xor(v[0].x, s);
    v[0].y += __shfl_xor(v[0].y, s);
  }
  if ((lane & 3) == 0) {
    const int r = lane >> 4, n = (lane >> 2) & 3;
    const int m = m0 + r;
    const double2 v0 = v[0];
    if (m < g.M && n < g.N) {
      if (g.nsplit > 1) {
        g.P[((int64_t)tl.ks * g.M + m) * g.N + n] = v0;
      } else {
        double2 o = make_double2(g.alpha * v0.x, g.alpha * v0.y);
        if constexpr (F32IO) {
          if (g.beta != 0.0) {
            const float2 c0 = ((const float2 *)g.C0)[(int64_t)m * g.ldc0 + n];
            o.x += g.beta * c0.x;
            o.y += g.beta * c0.y;
          }
          ((float2 *)g.C)[(int64_t)m * g.ldc + n] = make_float2((float)o.x, (float)o.y);
        } else {
          if (g.beta != 0.0) {
            const double2 c0 = g.C0[(int64_t)m * g.ldc0 + n];
            o.x += g.beta * c0.x;
            o.y += g.beta * c0.y;
          }
          g.C[(int64_t)m * g.ldc + n] = o;
        }
      }
    }
  }
}

// Fixed-order sum of the split-K partials: C = alpha sum_s P_s + beta C0.
__global__ void k_splitk_reduce(const GemmGroup *__restrict__ groups, const int *__restrict__ gidx,
                                const SolveDesc *__restrict__ descs, const int *__restrict__ gsolve) {
  const int gi = gidx[blockIdx.y];
  if (descs && !descs[gsolve[gi]].active) return;
  const GemmGroup g = groups[gi];
  const int64_t mn = (int64_t)g.M * g.N;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < mn; e += (int64_t)gridDim.x * blockDim.x) {
    double2 acc = g.P[e];
    for (int s = 1; s < g.nsplit; ++s) acc = cadd(acc, g.P[s * mn + e]);
    const int m = (int)(e / g.N), n = (int)(e % g.N);
    double2 v = cscale(g.alpha, acc);
    if (g.beta != 0.0) v = cadd(v, cscale(g.beta, g.C0[(int64_t)m * g.ldc0 + n]));
    g.C[(int64_t)m * g.ldc + n] = v;
  }
}

void launch_splitk_reduce(const GemmGroup *groups, const int *gidx, int n, int64_t maxmn, const SolveDesc *dd,
                          const int *gsolve, hipStream_t st) {
  if (n <= 0) return;
  unsigned gx = (unsigned)std::max<int64_t>(1, std::min<int64_t>((maxmn + 255) / 256, 64));
  hipLaunchKernelGGL(k_splitk_reduce, dim3(gx, n), dim3(256), 0, st, groups, gidx, dd, gsolve);
}

// Down-conversion of the stored pppp blocks for the FP32 ladder.
__global__ void k_to_f32(const double2 *src, float2 *dst, int64_t n) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    double2 v = src[e];
    dst[e] = make_float2((float)v.x, (float)v.y);
  }
}

// ---------------------------------------------------------------------------
// launch wrappers (host side)
// ---------------------------------------------------------------------------
template <class V>
static void fill_blocks(const SolveDesc &d, const V &vas, int nch, int nphc, int max_npp, int max_nhh, int max_nhp,
                        hipStream_t st) {
  auto gx = [](int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192)); };
  hipLaunchKernelGGL(k_fill_vpppp<V>, dim3(gx((int64_t)max_npp * max_npp), nch), dim3(256), 0, st, d, vas);
  hipLaunchKernelGGL(k_fill_vhhpp<V>, dim3(gx((int64_t)max_npp * max_nhh), nch), dim3(256), 0, st, d, vas);
  hipLaunchKernelGGL(k_fill_vhhhh<V>, dim3(gx((int64_t)max_nhh * max_nhh), nch), dim3(256), 0, st, d, vas);
  hipLaunchKernelGGL(k_fill_w0<V>, dim3(gx((int64_t)max_nhp * max_nhp), nphc), dim3(256), 0, st, d, vas);
  hipLaunchKernelGGL(k_fock<V>, dim3((d.nsp + 255) / 256), dim3(256), 0, st, d, vas);
  hipLaunchKernelGGL(k_e0<V>, dim3(1), dim3(256), 0, st, d, vas);
}

void launch_fill_blocks(const SolveDesc &d, double volume, int nch, int nphc, int max_npp,
                        int max_nhh, int max_nhp, hipStream_t st) {
  MinnesotaV v;
  v.o.kx = d.kx; v.o.ky = d.ky; v.o.kz = d.kz; v.o.sz = d.sz; v.o.tz = d.tz; v.o.volume = volume;
  fill_blocks(d, v, nch, nphc, max_npp, max_nhh, max_nhp, st);
}

void launch_fill_blocks_chiral(const SolveDesc &d, const ChiralDev &c, int nch, int nphc, int max_npp,
                               int max_nhh, int max_nhp, hipStream_t st) {
  ChiralV v;
  v.c = c;
  v.nx = d.nx; v.ny = d.ny; v.nz = d.nz; v.sz = d.sz; v.tz = d.tz;
  fill_blocks(d, v, nch, nphc, max_npp, max_nhh, max_nhp, st);
}

// V_pppp blocks only (the union blocks of a shared pp ladder)
void launch_fill_vpppp(const SolveDesc &d, double volume, const ChiralDev *c, int nch, int max_npp, hipStream_t st) {
  const unsigned gx = (unsigned)std::max<int64_t>(1, std::min<int64_t>(((int64_t)max_npp * max_npp + 255) / 256, 8192));
  if (c) {
    ChiralV v;
    v.c = *c;
    v.nx = d.nx; v.ny = d.ny; v.nz = d.nz; v.sz = d.sz; v.tz = d.tz;
    hipLaunchKernelGGL(k_fill_vpppp<ChiralV>, dim3(gx, nch), dim3(256), 0, st, d, v);
  } else {
    MinnesotaV v;
    v.o.kx = d.kx; v.o.ky = d.ky; v.o.kz = d.kz; v.o.sz = d.sz; v.o.tz = d.tz; v.o.volume = volume;
    hipLaunchKernelGGL(k_fill_vpppp<MinnesotaV>, dim3(gx, nch), dim3(256), 0, st, d, v);
  }
}

void launch_elements(const SolveDesc &d, double volume, const ChiralDev *c, const int4 *pqrs, int n, double2 *out,
                     hipStream_t st) {
  const dim3 g((unsigned)std::max(1, (n + 255) / 256));
  if (c) {
    ChiralV v;
    v.c = *c;
    v.nx = d.nx; v.ny = d.ny; v.nz = d.nz; v.sz = d.sz; v.tz = d.tz;
    hipLaunchKernelGGL(k_elements<ChiralV>, g, dim3(256), 0, st, d, v, pqrs, n, out);
  } else {
    MinnesotaV v;
    v.o.kx = d.kx; v.o.ky = d.ky; v.o.kz = d.kz; v.o.sz = d.sz; v.o.tz = d.tz; v.o.volume = volume;
    hipLaunchKernelGGL(k_elements<MinnesotaV>, g, dim3(256), 0, st, d, v, pqrs, n, out);
  }
}

// approx 0: E0 only; 1: and the Fock matrix; 2: and the two-body blocks
void launch_3nf(const SolveDesc &d, const Chiral3Dev &c, int approx, int nch, int nphc, int max_npp, int max_nhh,
                int max_nhp, hipStream_t st) {
  auto gx = [](int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192)); };
  hipLaunchKernelGGL(k_3nf_e0, dim3(1), dim3(256), 0, st, d, c);
  if (approx < 1) return;
  hipLaunchKernelGGL(k_3nf_fock, dim3((d.nsp + 255) / 256), dim3(256), 0, st, d, c);
  if (approx < 2) return;
  hipLaunchKernelGGL(k_no2b_vpppp, dim3(gx((int64_t)max_npp * (max_npp + 1) / 2), nch), dim3(256), 0, st, d, c);
  hipLaunchKernelGGL(k_no2b_vhhpp, dim3(gx((int64_t)max_npp * max_nhh), nch), dim3(256), 0, st, d, c);
  hipLaunchKernelGGL(k_no2b_vhhhh, dim3(gx((int64_t)max_nhh * (max_nhh + 1) / 2), nch), dim3(256), 0, st, d, c);
  hipLaunchKernelGGL(k_no2b_w0, dim3(gx((int64_t)max_nhp * (max_nhp + 1) / 2), nphc), dim3(256), 0, st, d, c);
}

void launch_derive(const SolveDesc &d, int nch, int nphc, int max_npp, int max_nhh, int max_nph,
                   int max_nhp, hipStream_t st) {
  auto gx = [](int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192)); };
  hipLaunchKernelGGL(k_derive_r0, dim3(gx((int64_t)max_npp * max_nhh), nch), dim3(256), 0, st, d);
  hipLaunchKernelGGL(k_derive_vt, dim3(gx((int64_t)max_nhp * max_nph), nphc), dim3(256), 0, st, d);
}

// FP32 copy of the pppp blocks in the layout of the FP32-storage ladder: block
// c (npp x npp) at dst + off32[c] with rows of even length ld = npp rounded up,
// so every pair of elements (k, k+1) at even k is one aligned 16-byte load; the
// padding element of an odd row is zero.
__global__ void k_to_f32_rows(const double2 *src, float2 *dst, const int64_t *off, const int64_t *off32,
                              const int *npp) {
  const int c = blockIdx.y;
  const int n = npp[c];
  const int ld = (n + 1) & ~1;
  const int64_t m = (int64_t)n * ld;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(e / ld), col = (int)(e % ld);
    float2 v = make_float2(0.f, 0.f);
    if (col < n) {
      const double2 x = src[off[c] + (int64_t)r * n + col];
      v = make_float2((float)x.x, (float)x.y);
    }
    dst[off32[c] + e] = v;
  }
}

void launch_to_f32_rows(const double2 *src, float2 *dst, const int64_t *off, const int64_t *off32, const int *npp,
                        int nch, int max_npp, hipStream_t st) {
  const int64_t m = (int64_t)max_npp * ((max_npp + 1) & ~1);
  const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>((m + 255) / 256, 1024));
  hipLaunchKernelGGL(k_to_f32_rows, dim3(g, nch), dim3(256), 0, st, src, dst, off, off32, npp);
}

void launch_to_f32(const double2 *src, float2 *dst, int64_t n, hipStream_t st) {
  unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 16384));
  hipLaunchKernelGGL(k_to_f32, dim3(g), dim3(256), 0, st, src, dst, n);
}

// elementwise launch geometry: grid.x blocks of 256 per solve, grid.y = solves
static inline dim3 egrid(int64_t n, int nsolve) {
  return dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 2048)), nsolve);
}

void launch_mbpt2(const SolveDesc *dd, int nsolve, int64_t maxT, hipStream_t st) {
  hipLaunchKernelGGL(k_mbpt2, egrid(maxT, nsolve), dim3(256), 0, st, dd);
}

// energy into scal[slot]; uses red[0..ENERGY_BLOCKS)
constexpr int kRedBlocks = 512;
void launch_energy(const SolveDesc *dd, int nsolve, int slot, hipStream_t st) {
  hipLaunchKernelGGL(k_energy_partial, dim3(kRedBlocks, nsolve), dim3(256), 0, st, dd);
  hipLaunchKernelGGL(k_reduce_final, dim3(1, nsolve), dim3(256), 0, st, dd, kRedBlocks, 1, 1, slot);
}

void launch_i7(const SolveDesc *dd, int nsolve, int maxA, hipStream_t st) {
  hipLaunchKernelGGL(k_i7, dim3((2 * maxA + 3) / 4, nsolve), dim3(256), 0, st, dd);
}
void launch_i6(const SolveDesc *dd, int nsolve, int maxnp, int64_t maxrows, hipStream_t st) {
  unsigned gx = (unsigned)std::max<int64_t>(1, std::min<int64_t>((maxrows * 16 + 255) / 256, 4096));
  hipLaunchKernelGGL(k_i6_rows, dim3(gx, nsolve), dim3(256), 0, st, dd);
  hipLaunchKernelGGL(k_i6_reduce, dim3((2 * maxnp + 3) / 4, nsolve), dim3(256), 0, st, dd);
}
void launch_i9(const SolveDesc *dd, int nsolve, int64_t maxvhh, hipStream_t st) {
  hipLaunchKernelGGL(k_i9, egrid(maxvhh, nsolve), dim3(256), 0, st, dd);
}
void launch_resid_init(const SolveDesc *dd, int nsolve, int64_t maxT, bool onebody, hipStream_t st) {
  hipLaunchKernelGGL(k_resid_init, egrid(maxT, nsolve), dim3(256), 0, st, dd, onebody ? 1 : 0);
}
void launch_t_split(const SolveDesc *dd, int nsolve, int64_t maxT, hipStream_t st) {
  hipLaunchKernelGGL(k_t_split, egrid(maxT, nsolve), dim3(256), 0, st, dd);
}
void launch_ring_gather(const SolveDesc *dd, int nsolve, int64_t maxT, bool f32, hipStream_t st) {
  if (f32) hipLaunchKernelGGL(k_ring_gather<true>, egrid(maxT, nsolve), dim3(256), 0, st, dd);
  else hipLaunchKernelGGL(k_ring_gather<false>, egrid(maxT, nsolve), dim3(256), 0, st, dd);
}
void launch_ring_scatter(const SolveDesc *dd, int nsolve, int64_t maxT, bool f32, bool ring, bool hh,
                         hipStream_t st) {
  if (f32) hipLaunchKernelGGL(k_ring_scatter<true>, egrid(maxT, nsolve), dim3(256), 0, st, dd, (int)ring, (int)hh);
  else hipLaunchKernelGGL(k_ring_scatter<false>, egrid(maxT, nsolve), dim3(256), 0, st, dd, (int)ring, (int)hh);
}
void launch_update(const SolveDesc *dd, int nsolve, int64_t maxT, const int *slot, hipStream_t st) {
  hipLaunchKernelGGL(k_update, egrid(maxT, nsolve), dim3(256), 0, st, dd, slot);
}
void launch_gram(const SolveDesc *dd, int nsolve, const int *slots, int nstep, hipStream_t st) {
  int ne = nstep - 1;
  dim3 gr(kRedBlocks, nsolve);
  switch (ne) {
    case 1: hipLaunchKernelGGL(k_gram_partial<1>, gr, dim3(256), 0, st, dd, slots); break;
    case 2: hipLaunchKernelGGL(k_gram_partial<2>, gr, dim3(256), 0, st, dd, slots); break;
    case 3: hipLaunchKernelGGL(k_gram_partial<3>, gr, dim3(256), 0, st, dd, slots); break;
    case 4: hipLaunchKernelGGL(k_gram_partial<4>, gr, dim3(256), 0, st, dd, slots); break;
    case 5: hipLaunchKernelGGL(k_gram_partial<5>, gr, dim3(256), 0, st, dd, slots); break;
    case 6: hipLaunchKernelGGL(k_gram_partial<6>, gr, dim3(256), 0, st, dd, slots); break;
    case 7: hipLaunchKernelGGL(k_gram_partial<7>, gr, dim3(256), 0, st, dd, slots); break;
    case 8: hipLaunchKernelGGL(k_gram_partial<8>, gr, dim3(256), 0, st, dd, slots); break;
    default: return;
  }
  int nv = ne * (ne + 1) / 2;
  hipLaunchKernelGGL(k_reduce_final, dim3(1, nsolve), dim3(256), 0, st, dd, kRedBlocks, 36, nv, 4);
}
void launch_share_gather(const SolveDesc *dd, int nsolve, int64_t maxT, const uint32_t *const *maps, double2 *Tu,
                         hipStream_t st) {
  hipLaunchKernelGGL(k_share_gather, egrid(maxT, nsolve), dim3(256), 0, st, dd, maps, Tu);
}
void launch_share_scatter(const SolveDesc *dd, int nsolve, int64_t maxT, const uint32_t *const *maps,
                          const double2 *Ru, hipStream_t st) {
  hipLaunchKernelGGL(k_share_scatter, egrid(maxT, nsolve), dim3(256), 0, st, dd, maps, Ru);
}
void launch_scatter_values(double2 *dst, const int64_t *idx, const double2 *val, int64_t n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_scatter_values, egrid(n, 1), dim3(256), 0, st, dst, idx, val, n);
}
// Streaming read of n 16-byte elements: block b reads its contiguous span of
// per elements, lanes 16 bytes apart; the sum is stored only when it takes an
// impossible value, so the loads stay and nothing is written.
__global__ __launch_bounds__(256) void k_read_probe(const double2 *__restrict__ a, int64_t n, int64_t per,
                                                    double *out) {
  const int64_t b0 = (int64_t)blockIdx.x * per;
  const int64_t b1 = b0 + per < n ? b0 + per : n;
  double acc = 0.0;
  for (int64_t e = b0 + threadIdx.x; e < b1; e += blockDim.x) {
    const double2 v = a[e];
    acc += v.x + v.y;
  }
  if (acc == 123.456) out[blockIdx.x] = acc;
}
void launch_read_probe(const double2 *a, int64_t n, double *out, int nblocks, hipStream_t st) {
  const int64_t per = (n + nblocks - 1) / nblocks;
  hipLaunchKernelGGL(k_read_probe, dim3(nblocks), dim3(256), 0, st, a, n, per, out);
}

void launch_aa_update(const SolveDesc *dd, int nsolve, const int *slots, double beta, hipStream_t st) {
  constexpr int kAABlocks = 1024;   // partials: SolveDesc::red holds 1024 x 36
  hipLaunchKernelGGL(k_aa_update, dim3(kAABlocks, nsolve), dim3(256), 0, st, dd, slots, beta);
  // 17 values per solve into scal[24 .. 40]
  hipLaunchKernelGGL(k_reduce_final, dim3(1, nsolve), dim3(256), 0, st, dd, kAABlocks, 36, 17, 24);
}
void launch_aa_combine(const SolveDesc *dd, int nsolve, int64_t maxT, const int *slots, const double *coef,
                       hipStream_t st) {
  hipLaunchKernelGGL(k_aa_combine, egrid(maxT, nsolve), dim3(256), 0, st, dd, slots, coef);
}
void launch_diis_axpy(const SolveDesc *dd, int nsolve, int64_t maxT, const int *slots, const double2 *coef,
                      int nvec, hipStream_t st) {
  hipLaunchKernelGGL(k_diis_axpy, egrid(maxT, nsolve), dim3(256), 0, st, dd, slots, coef, nvec);
}

// Kernel kinds: 0 FP64, 1 FP32 (A double), 2 FP32 (A float), 3 FP64 pp ladder,
// 4 FP32 pp ladder (A float), 5 pp ladder with A float and FP64 arithmetic,
// 9 pp ladder on BF16x3 split products (A float), 10 the same with B the
// split copy of T (k_t_split);
// 6 BF16x3 split products (A double; the thin N <= 4 class stays FP64), tile
// shapes (1,1..6), (1,8), (2,1..4), (4,1); 7 the same with A float; 8 with A,
// B, C and C0 float (the FP32 ring blocks).
// Tile shapes (MB, NB): MB 16-row blocks per wave (4 waves stacked along M),
// NB 16-column blocks: (1,1) (1,2) (1,3) (1,4) (1,5) (2,1) (2,2) (4,1); and
// (0,4): the thin VALU kernel (16-row tiles, N <= 4).  lb = 1 selects the
// variant with B staged in LDS (k_zgemm_lb).
template <int MB, int NB>
static void launch_shape_lb(const GemmGroup *groups, const GemmTile *tiles, int ntiles, const SolveDesc *dd,
                            const int *gsolve, int kind, hipStream_t st) {
  switch (kind) {
    case 0: hipLaunchKernelGGL((k_zgemm_lb<MB, NB, false, double2, 0>), dim3(ntiles), dim3(256), 0, st, groups, tiles, dd, gsolve); break;
    case 1: hipLaunchKernelGGL((k_zgemm_lb<MB, NB, true, double2, 0>), dim3(ntiles), dim3(256), 0, st, groups, tiles, dd, gsolve); break;
    case 2: hipLaunchKernelGGL((k_zgemm_lb<MB, NB, true, float2, 0>), dim3(ntiles), dim3(256), 0, st, groups, tiles, dd, gsolve); break;
    case 3: hipLaunchKernelGGL((k_zgemm_lb<MB, NB, false, double2, 1>), dim3(ntiles), dim3(256), 0, st, groups, tiles, dd, gsolve); break;
    case 4: hipLaunchKernelGGL((k_zgemm_lb<MB, NB, true, float2, 1>), dim3(ntiles), dim3(256), 0, st, groups, tiles, dd, gsolve); break;
    case 5: hipLaunchKernelGGL((k_zgemm_lb<MB, NB, false, float2, 2>), dim3(ntiles), dim3(256), 0, st, groups, tiles, dd, gsolve); break;
    default: break;
  }
}

// ---------------------------------------------------------------------------
// pp ladder R += V_pppp T (ccm_t2_eqn.f90:574-607) with the V stream staged
// through LDS by direct-to-LDS loads (global_load_lds_dwordx4): a tile is 64
// rows x 16 NB columns over the whole k range of its channel (no split-K), and
// every workgroup keeps up to three chunks of 16 KB of V (16 k of FP64, or 32 k
// of the FP32 copy) plus the matching rows of T in flight, with no VGPRs spent on
// staging.  Each 16 lanes of a load instruction read one V row's 256
// contiguous bytes; the row's slots land XOR-swizzled by the row index, so the
// MFMA reads (16 rows at one k) hit 16 different LDS slots.  The k order of
// every sum is that of the other MFMA forms (step q takes k0 + 4q + lane
// group), so all unsplit forms give the same bits.
//   VK 0: FP64 V, FP64 MFMA (3 real products per complex MAC)
//   VK 1: FP32-stored V, FP64 MFMA
//   VK 2: FP32-stored V, FP32 MFMA (4 real products)
//   VK 3: FP32-stored V, BF16x3 split products on bf16 MFMA (FP32 sums)
//   VK 4: as 3, with T read already split (k_t_split; same bits)
// Waits are explicit: s_waitcnt vmcnt(n) for the chunk about to be read and
// lgkmcnt(0) for this wave's LDS reads of the stage about to be refilled, then
// a bare s_barrier (a workgroup fence would force vmcnt(0) and drain the
// pipeline); the empty asm statements keep the compiler from moving memory
// accesses across it.
// ---------------------------------------------------------------------------
typedef __attribute__((address_space(3))) void lds_void;
template <int VM>
__device__ __forceinline__ void ppl_sync() {
  static_assert(VM >= 0 && VM < 64, "vmcnt range");
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt((VM & 15) | ((VM >> 4) << 14) | (7 << 4) | (0 << 8));
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}
__device__ __forceinline__ void dma16(const void *src, char *lds) {
  __builtin_amdgcn_global_load_lds(src, (lds_void *)lds, 16, 0, 0);
}

template <int NB, int VK>
struct PplCfg {
  static constexpr bool VF32 = VK != 0;
  static constexpr int KC = VF32 ? 32 : 16;          // k per chunk: 64 rows x KC = 16 KB of V
  static constexpr int NT = 16 * NB;
  static constexpr int BBYTES = KC * NT * 16;        // T rows of a chunk (FP64)
  static constexpr int BIW = BBYTES / 4096;          // T load instructions per wave per chunk
  static constexpr int IPW = 4 + BIW;                // load instructions per wave per chunk
  // stages: as many as two workgroups per CU (160 KB of LDS) allow, at most 4
  static constexpr int S = 4 * (16384 + BBYTES) <= 81920 ? 4 : 3 * (16384 + BBYTES) <= 81920 ? 3 : 2;
};

template <int NB, int VK>
__global__ __launch_bounds__(256, 2) void k_ppl(const GemmGroup *__restrict__ groups,
                                              const GemmTile *__restrict__ tiles,
                                              const SolveDesc *__restrict__ descs,
                                              const int *__restrict__ gsolve) {
  typedef PplCfg<NB, VK> C;
  constexpr bool VF32 = C::VF32, MF32 = VK >= 2, BF3 = VK >= 3, TSPL = VK == 4;
  constexpr int KC = C::KC, NT = C::NT, S = C::S, IPW = C::IPW, BIW = C::BIW;
  static_assert(C::BBYTES % 4096 == 0, "T rows of a chunk in whole instructions per wave");
  typedef typename std::conditional<MF32, float, double>::type real;
  typedef typename std::conditional<MF32, f4, d4>::type acc4;
  // one object per stage: the compiler's wait insertion then tells the
  // stages apart and leaves the waits of the pipeline to the code below
  __shared__ __attribute__((aligned(16))) char A0[16384], A1[16384], A2[S >= 3 ? 16384 : 16],
      A3[S >= 4 ? 16384 : 16];
  __shared__ __attribute__((aligned(16))) char B0[C::BBYTES], B1[C::BBYTES], B2[S >= 3 ? C::BBYTES : 16],
      B3[S >= 4 ? C::BBYTES : 16];

  const GemmTile tl = tiles[blockIdx.x];
  if (descs && !descs[gsolve[tl.g]].active) return;
  const GemmGroup g = groups[tl.g];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int li = lane & 15, lg = lane >> 4;
  const int K = g.K, nchunk = (K + KC - 1) / KC;
  // per-lane source rows of the V loads: instruction i of wave w covers tile
  // rows 16 w + 4 i + lg, slot li of each row holding element (or FP32 pair)
  // li ^ (row & 15)
  const char *arow[4];
  int aslot[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int r = 16 * w + 4 * i + lg;
    const int m = min(tl.m0 + r, g.M - 1);
    aslot[i] = li ^ (r & 15);
    arow[i] = (const char *)g.A + (int64_t)m * g.lda * (VF32 ? 8 : 16);
  }
  // T loads: element e = 64 (BIW w + j) + lane of the chunk's KC x NT block
  int bkr[BIW], bcol[BIW];
#pragma unroll
  for (int j = 0; j < BIW; ++j) {
    const int e = 64 * (BIW * w + j) + lane;
    bkr[j] = e / NT;
    bcol[j] = min(tl.n0 + e % NT, g.N - 1);
  }
  // chunks past the last one (the pipeline runs whole rounds of S) re-load the
  // last chunk (L2 hits) and are masked to zero in compute
  auto issue = [&](int c, char *as, char *bs) {
    const int k0 = min(c, nchunk - 1) * KC;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const char *src;
      if constexpr (VF32) src = arow[i] + 8 * (int64_t)min(k0 + 2 * aslot[i], ((K + 1) & ~1) - 2);
      else src = arow[i] + 16 * (int64_t)min(k0 + aslot[i], K - 1);
      dma16(src, as + (16 * w + 4 * i) * 256);
    }
#pragma unroll
    for (int j = 0; j < BIW; ++j) {
      const int k = min(k0 + bkr[j], K - 1);
      dma16(g.B + (int64_t)k * g.ldb + bcol[j], bs + (BIW * w + j) * 1024);
    }
  };
  acc4 cre[NB], cim[NB], cs[NB];
#pragma unroll
  for (int y = 0; y < NB; ++y) {
    cre[y] = (acc4){0, 0, 0, 0};
    cim[y] = (acc4){0, 0, 0, 0};
    cs[y] = (acc4){0, 0, 0, 0};
  }
  const int row = 16 * w + li;
  auto mac = [&](double2 a, int kk, const char *bs, int kvalid) {
    const real ar = (real)a.x, ai = (real)a.y;
#pragma unroll
    for (int y = 0; y < NB; ++y) {
      double2 b = *(const double2 *)(bs + (kk * NT + 16 * y + li) * 16);
      if (kk >= kvalid) b = make_double2(0.0, 0.0);
      const real br = (real)b.x, bi = (real)b.y;
      if constexpr (MF32) {
        cre[y] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, br, cre[y], 0, 0, 0);
        cre[y] = __builtin_amdgcn_mfma_f32_16x16x4f32(-ai, bi, cre[y], 0, 0, 0);
        cim[y] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, bi, cim[y], 0, 0, 0);
        cim[y] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, br, cim[y], 0, 0, 0);
      } else {
        cre[y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, cre[y], 0, 0, 0);
        cim[y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, bi, cim[y], 0, 0, 0);
        cs[y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar + ai, br + bi, cs[y], 0, 0, 0);
      }
    }
  };
  auto compute = [&](const char *as, const char *bs, int kvalid) {
    if constexpr (BF3) {
      // two bf16 MFMA steps of 16 complex k: lane group lg takes complex
      // k = 16 s + 4 lg + 0..3, i.e. the FP32 pairs 8 s + 2 lg and 8 s + 2 lg + 1
      // of its row; A and B split into bf16 hi/lo, (re, im) side by side
#pragma unroll
      for (int st = 0; st < 2; ++st) {
        const int p0 = 8 * st + 2 * lg;
        const float4 x0 = *(const float4 *)(as + row * 256 + ((p0 ^ (row & 15)) * 16));
        const float4 x1 = *(const float4 *)(as + row * 256 + (((p0 + 1) ^ (row & 15)) * 16));
        unsigned hh[4], ll[4];
        bf3_split(x0.x, x0.y, hh[0], ll[0]);
        bf3_split(x0.z, x0.w, hh[1], ll[1]);
        bf3_split(x1.x, x1.y, hh[2], ll[2]);
        bf3_split(x1.z, x1.w, hh[3], ll[3]);
        const u32x4 ah = {hh[0], hh[1], hh[2], hh[3]}, al = {ll[0], ll[1], ll[2], ll[3]};
        const bf16x8 xh = __builtin_bit_cast(bf16x8, ah), xl = __builtin_bit_cast(bf16x8, al);
#pragma unroll
        for (int y = 0; y < NB; ++y) {
          u32x4 bh, bl, ih, il;
          if constexpr (TSPL) {
            // masked k: the split of b = 0 as the in-kernel form makes it
            unsigned zh, zl;
            bf3_split(0.f, -0.f, zh, zl);
            const u32x4 zi = bf3_im_pairs(u32x4{zh, zl, 0u, 0u});
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int kk = 16 * st + 4 * lg + q;
              uint4 e = *(const uint4 *)(bs + (kk * NT + 16 * y + li) * 16);
              if (kk >= kvalid) e = make_uint4(zh, zl, zi[0], zi[1]);
              bh[q] = e.x;
              bl[q] = e.y;
              ih[q] = e.z;
              il[q] = e.w;
            }
          } else {
            unsigned bhh[4], bll[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int kk = 16 * st + 4 * lg + q;
              double2 b = *(const double2 *)(bs + (kk * NT + 16 * y + li) * 16);
              if (kk >= kvalid) b = make_double2(0.0, 0.0);
              bf3_split((float)b.x, -(float)b.y, bhh[q], bll[q]);
            }
            bh = u32x4{bhh[0], bhh[1], bhh[2], bhh[3]};
            bl = u32x4{bll[0], bll[1], bll[2], bll[3]};
            ih = bf3_im_pairs(bh);
            il = bf3_im_pairs(bl);
          }
          cre[y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xl, __builtin_bit_cast(bf16x8, bh), cre[y], 0, 0, 0);
          cre[y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh, __builtin_bit_cast(bf16x8, bl), cre[y], 0, 0, 0);
          cre[y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh, __builtin_bit_cast(bf16x8, bh), cre[y], 0, 0, 0);
          cim[y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xl, __builtin_bit_cast(bf16x8, ih), cim[y], 0, 0, 0);
          cim[y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh, __builtin_bit_cast(bf16x8, il), cim[y], 0, 0, 0);
          cim[y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh, __builtin_bit_cast(bf16x8, ih), cim[y], 0, 0, 0);
        }
      }
    } else if constexpr (VF32) {
      // MFMA step (h, t) takes k = 2 (4 h + lane group) + t: each lane reads
      // its FP32 pair once for two steps (same k order as kstep<true>)
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        const int pr = 4 * h + lg;
        const float4 a4 = *(const float4 *)(as + row * 256 + ((pr ^ (row & 15)) * 16));
        mac(make_double2(a4.x, a4.y), 2 * pr, bs, kvalid);
        mac(make_double2(a4.z, a4.w), 2 * pr + 1, bs, kvalid);
      }
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int kk = 4 * q + lg;
        const double2 a = *(const double2 *)(as + row * 256 + ((kk ^ (row & 15)) * 16));
        mac(a, kk, bs, kvalid);
      }
    }
  };
  // One step of the pipeline: chunk c from stage (ac, bc); the stage (an, bn)
  // read in the previous step is refilled with chunk c + S - 1.  The loop body
  // is branch-free (whole rounds of S chunks, every step issues and waits for
  // the same count), so the compiler's wait insertion tracks each stage's
  // loads and adds no wait of its own before the LDS reads.
  auto step = [&](int c, const char *ac, const char *bc, char *an, char *bn) {
    ppl_sync<(S - 2) * IPW>();
    issue(c + S - 1, an, bn);
    compute(ac, bc, max(0, min(KC, K - c * KC)));
  };
  issue(0, A0, B0);
  if constexpr (S == 4) {
    issue(1, A1, B1);
    issue(2, A2, B2);
    for (int c = 0; c < nchunk; c += 4) {
      step(c, A0, B0, A3, B3);
      step(c + 1, A1, B1, A0, B0);
      step(c + 2, A2, B2, A1, B1);
      step(c + 3, A3, B3, A2, B2);
    }
  } else if constexpr (S == 3) {
    issue(1, A1, B1);
    for (int c = 0; c < nchunk; c += 3) {
      step(c, A0, B0, A2, B2);
      step(c + 1, A1, B1, A0, B0);
      step(c + 2, A2, B2, A1, B1);
    }
  } else {
    for (int c = 0; c < nchunk; c += 2) {
      step(c, A0, B0, A1, B1);
      step(c + 1, A1, B1, A0, B0);
    }
  }
  // no load into LDS may outlive the workgroup
  ppl_sync<0>();
  // result layouts of the 16x16x4 MFMAs: register r of lane l holds column l&15
  // and row (l>>4)+4r for f64, row 4(l>>4)+r for f32
#pragma unroll
  for (int y = 0; y < NB; ++y)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int rr = MF32 ? (lg * 4 + r) : (lg + 4 * r);
      const int m = tl.m0 + 16 * w + rr;
      const int n = tl.n0 + y * 16 + li;
      double vr, vi;
      if constexpr (MF32) {
        vr = (double)cre[y][r];
        vi = (double)cim[y][r];
      } else {
        vr = cre[y][r] - cim[y][r];
        vi = cs[y][r] - cre[y][r] - cim[y][r];
      }
      if (m < g.M && n < g.N) {
        double2 v = make_double2(g.alpha * vr, g.alpha * vi);
        if (g.beta != 0.0) {
          const double2 c0 = g.C0[(int64_t)m * g.ldc0 + n];
          v.x += g.beta * c0.x;
          v.y += g.beta * c0.y;
        }
        g.C[(int64_t)m * g.ldc + n] = v;
      }
    }
}

// The same pipeline with eight waves and one workgroup per CU: a tile is 64
// rows x 32 columns, waves 0-3 take columns 0-15 and waves 4-7 columns 16-31 of
// the same V chunk, so a channel with up to 32 hole pairs streams its V_pppp
// block once (16-column tiles stream it once per column tile).  With the
// whole LDS to itself the workgroup keeps five chunks in flight (FP64 V; three
// for the FP32 copy).  Same k order, so the same bits as the other forms.
template <int VK>
struct Ppl8Cfg {
  static constexpr bool VF32 = VK != 0;
  static constexpr int KC = VF32 ? 32 : 16;
  static constexpr int BBYTES = KC * 32 * 16;         // T rows of a chunk: KC x 32 columns
  static constexpr int BIW = BBYTES / 8192;           // T load instructions per wave per chunk
  static constexpr int IPW = 2 + BIW;                 // V: 16 instructions over 8 waves
  static constexpr int S = VF32 ? 4 : 6;              // stages (S x (16 KB + BBYTES) <= 160 KB)
};

template <int VK>
__global__ __launch_bounds__(512, 2) void k_ppl8(const GemmGroup *__restrict__ groups,
                                               const GemmTile *__restrict__ tiles,
                                               const SolveDesc *__restrict__ descs,
                                               const int *__restrict__ gsolve) {
  typedef Ppl8Cfg<VK> C;
  constexpr bool VF32 = C::VF32, MF32 = VK == 2;
  constexpr int KC = C::KC, S = C::S, IPW = C::IPW, BIW = C::BIW;
  static_assert(C::BBYTES % 8192 == 0 && S * (16384 + C::BBYTES) <= 163840, "LDS budget");
  typedef typename std::conditional<MF32, float, double>::type real;
  typedef typename std::conditional<MF32, f4, d4>::type acc4;
  __shared__ __attribute__((aligned(16))) char A0[16384], A1[16384], A2[16384], A3[16384],
      A4[S >= 5 ? 16384 : 16], A5[S >= 6 ? 16384 : 16];
  __shared__ __attribute__((aligned(16))) char B0[C::BBYTES], B1[C::BBYTES], B2[C::BBYTES], B3[C::BBYTES],
      B4[S >= 5 ? C::BBYTES : 16], B5[S >= 6 ? C::BBYTES : 16];
  const GemmTile tl = tiles[blockIdx.x];
  if (descs && !descs[gsolve[tl.g]].active) return;
  const GemmGroup g = groups[tl.g];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int li = lane & 15, lg = lane >> 4;
  const int wr = w & 3, wc = w >> 2;                 // row block and column block of this wave
  const int K = g.K, nchunk = (K + KC - 1) / KC;
  const char *arow[2];
  int aslot[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int r = 4 * (2 * w + i) + lg;
    aslot[i] = li ^ (r & 15);
    arow[i] = (const char *)g.A + (int64_t)min(tl.m0 + r, g.M - 1) * g.lda * (VF32 ? 8 : 16);
  }
  int bkr[BIW], bcol[BIW];
#pragma unroll
  for (int j = 0; j < BIW; ++j) {
    const int e = 64 * (BIW * w + j) + lane;
    bkr[j] = e >> 5;
    bcol[j] = min(tl.n0 + (e & 31), g.N - 1);
  }
  auto issue = [&](int c, char *as, char *bs) {
    const int k0 = min(c, nchunk - 1) * KC;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const char *src;
      if constexpr (VF32) src = arow[i] + 8 * (int64_t)min(k0 + 2 * aslot[i], ((K + 1) & ~1) - 2);
      else src = arow[i] + 16 * (int64_t)min(k0 + aslot[i], K - 1);
      dma16(src, as + (2 * w + i) * 1024);
    }
#pragma unroll
    for (int j = 0; j < BIW; ++j) {
      const int k = min(k0 + bkr[j], K - 1);
      dma16(g.B + (int64_t)k * g.ldb + bcol[j], bs + (BIW * w + j) * 1024);
    }
  };
  acc4 cre = (acc4){0, 0, 0, 0}, cim = cre, cs = cre;
  const int row = 16 * wr + li;
  auto mac = [&](double2 a, int kk, const char *bs, int kvalid) {
    const real ar = (real)a.x, ai = (real)a.y;
    double2 b = *(const double2 *)(bs + (kk * 32 + 16 * wc + li) * 16);
    if (kk >= kvalid) b = make_double2(0.0, 0.0);
    const real br = (real)b.x, bi = (real)b.y;
    if constexpr (MF32) {
      cre = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, br, cre, 0, 0, 0);
      cre = __builtin_amdgcn_mfma_f32_16x16x4f32(-ai, bi, cre, 0, 0, 0);
      cim = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, bi, cim, 0, 0, 0);
      cim = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, br, cim, 0, 0, 0);
    } else {
      cre = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, cre, 0, 0, 0);
      cim = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, bi, cim, 0, 0, 0);
      cs = __builtin_amdgcn_mfma_f64_16x16x4f64(ar + ai, br + bi, cs, 0, 0, 0);
    }
  };
  auto compute = [&](const char *as, const char *bs, int kvalid) {
    if constexpr (VF32) {
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        const int pr = 4 * h + lg;
        const float4 a4 = *(const float4 *)(as + row * 256 + ((pr ^ (row & 15)) * 16));
        mac(make_double2(a4.x, a4.y), 2 * pr, bs, kvalid);
        mac(make_double2(a4.z, a4.w), 2 * pr + 1, bs, kvalid);
      }
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int kk = 4 * q + lg;
        mac(*(const double2 *)(as + row * 256 + ((kk ^ (row & 15)) * 16)), kk, bs, kvalid);
      }
    }
  };
  auto step = [&](int c, const char *ac, const char *bc, char *an, char *bn) {
    ppl_sync<(S - 2) * IPW>();
    issue(c + S - 1, an, bn);
    compute(ac, bc, max(0, min(KC, K - c * KC)));
  };
  issue(0, A0, B0);
  issue(1, A1, B1);
  issue(2, A2, B2);
  if constexpr (S == 6) {
    issue(3, A3, B3);
    issue(4, A4, B4);
    for (int c = 0; c < nchunk; c += 6) {
      step(c, A0, B0, A5, B5);
      step(c + 1, A1, B1, A0, B0);
      step(c + 2, A2, B2, A1, B1);
      step(c + 3, A3, B3, A2, B2);
      step(c + 4, A4, B4, A3, B3);
      step(c + 5, A5, B5, A4, B4);
    }
  } else {
    for (int c = 0; c < nchunk; c += 4) {
      step(c, A0, B0, A3, B3);
      step(c + 1, A1, B1, A0, B0);
      step(c + 2, A2, B2, A1, B1);
      step(c + 3, A3, B3, A2, B2);
    }
  }
  ppl_sync<0>();
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int rr = MF32 ? (lg * 4 + r) : (lg + 4 * r);
    const int m = tl.m0 + 16 * wr + rr;
    const int n = tl.n0 + 16 * wc + li;
    double vr, vi;
    if constexpr (MF32) {
      vr = (double)cre[r];
      vi = (double)cim[r];
    } else {
      vr = cre[r] - cim[r];
      vi = cs[r] - cre[r] - cim[r];
    }
    if (m < g.M && n < g.N) {
      double2 v = make_double2(g.alpha * vr, g.alpha * vi);
      if (g.beta != 0.0) {
        const double2 c0 = g.C0[(int64_t)m * g.ldc0 + n];
        v.x += g.beta * c0.x;
        v.y += g.beta * c0.y;
      }
      g.C[(int64_t)m * g.ldc + n] = v;
    }
  }
}

static void launch_shape_ppl8(const GemmGroup *groups, const GemmTile *tiles, int ntiles, const SolveDesc *dd,
                              const int *gsolve, int kind, hipStream_t st) {
  switch (kind) {
    case 3: hipLaunchKernelGGL((k_ppl8<0>), dim3(ntiles), dim3(512), 0, st, groups, tiles, dd, gsolve); break;
    case 5: hipLaunchKernelGGL((k_ppl8<1>), dim3(ntiles), dim3(512), 0, st, groups, tiles, dd, gsolve); break;
    case 4: hipLaunchKernelGGL((k_ppl8<2>), dim3(ntiles), dim3(512), 0, st, groups, tiles, dd, gsolve); break;
    // BF16x3: the four-wave kernel on the same 64 x 32 tiles
    case 9: hipLaunchKernelGGL((k_ppl<2, 3>), dim3(ntiles), dim3(256), 0, st, groups, tiles, dd, gsolve); break;
    case 10: hipLaunchKernelGGL((k_ppl<2, 4>), dim3(ntiles), dim3(256), 0, st, groups, tiles, dd, gsolve); break;
    default: break;
  }
}

template <int NB>
static void launch_shape_ppl(const GemmGroup *groups, const GemmTile *tiles, int ntiles, const SolveDesc *dd,
                             const int *gsolve, int kind, hipStream_t st) {
  // kinds of the pp ladder: 3 FP64, 4 FP32 MFMA on the FP32 copy, 5 FP32 copy with FP64 MFMA,
  // 9 BF16x3, 10 BF16x3 on the split copy of T
  switch (kind) {
    case 3: hipLaunchKernelGGL((k_ppl<NB, 0>), dim3(ntiles), dim3(256), 0, st, groups, tiles, dd, gsolve); break;
    case 5: hipLaunchKernelGGL((k_ppl<NB, 1>), dim3(ntiles), dim3(256), 0, st, groups, tiles, dd, gsolve); break;
    case 4: hipLaunchKernelGGL((k_ppl<NB, 2>), dim3(ntiles), dim3(256), 0, st, groups, tiles, dd, gsolve); break;
    case 9: hipLaunchKernelGGL((k_ppl<NB, 3>), dim3(ntiles), dim3(256), 0, st, groups, tiles, dd, gsolve); break;
    case 10: hipLaunchKernelGGL((k_ppl<NB, 4>), dim3(ntiles), dim3(256), 0, st, groups, tiles, dd, gsolve); break;
    default: break;
  }
}

// variant with A and B staged in LDS (k_zgemm_la), MB = 1.
template <int NB>
static void launch_shape_la(const GemmGroup *groups, const GemmTile *tiles, int ntiles, const SolveDesc *dd,
                            const int *gsolve, int kind, hipStream_t st) {
  switch (kind) {
    case 0: hipLaunchKernelGGL((k_zgemm_la<1, NB, false, double2, 0>), dim3(ntiles), dim3(256), 0, st, groups, tiles, dd, gsolve); break;
    case 1: hipLaunchKernelGGL((k_zgemm_la<1, NB, true, double2, 0>), dim3(ntiles), dim3(256), 0, st, groups, tiles, dd, gsolve); break;
    case 2: hipLaunchKernelGGL((k_zgemm_la<1, NB, true, float2, 0>), dim3(ntiles), dim3(256), 0, st, groups, tiles, dd, gsolve); break;
    case 3: hipLaunchKernelGGL((k_zgemm_la<1, NB, false, double2, 1>), dim3(ntiles), dim3(256), 0, st, groups, tiles, dd, gsolve); break;
    case 4: hipLaunchKernelGGL((k_zgemm_la<1, NB, true, float2, 1>), dim3(ntiles), dim3(256), 0, st, groups, tiles, dd, gsolve); break;
    case 5: hipLaunchKernelGGL((k_zgemm_la<1, NB, false, float2, 2>), dim3(ntiles), dim3(256), 0, st, groups, tiles, dd, gsolve); break;
    default: break;
  }
}

template <int MB, int NB>
static void launch_shape(const GemmGroup *groups, const GemmTile *tiles, int ntiles, const SolveDesc *dd,
                         const int *gsolve, int kind, hipStream_t st) {
  switch (kind) {
    case 0: hipLaunchKernelGGL((k_zgemm_dl<MB, NB, false, double2, 0>), dim3(ntiles), dim3(256), 0, st, groups, tiles, dd, gsolve); break;
    case 1: hipLaunchKernelGGL((k_zgemm_dl<MB, NB, true, double2, 0>), dim3(ntiles), dim3(256), 0, st, groups, tiles, dd, gsolve); break;
    case 2: hipLaunchKernelGGL((k_zgemm_dl<MB, NB, true, float2, 0>), dim3(ntiles), dim3(256), 0, st, groups, tiles, dd, gsolve); break;
    case 3: hipLaunchKernelGGL((k_zgemm_dl<MB, NB, false, double2, 1>), dim3(ntiles), dim3(256), 0, st, groups, tiles, dd, gsolve); break;
    case 4: hipLaunchKernelGGL((k_zgemm_dl<MB, NB, true, float2, 1>), dim3(ntiles), dim3(256), 0, st, groups, tiles, dd, gsolve); break;
    case 5: hipLaunchKernelGGL((k_zgemm_dl<MB, NB, false, float2, 2>), dim3(ntiles), dim3(256), 0, st, groups, tiles, dd, gsolve); break;
    default: break;
  }
}

// kind 6: BF16x3 split products (k_zgemm_bf3), one form for every lb
template <int MB, int NB>
static void launch_shape_bf3(const GemmGroup *groups, const GemmTile *tiles, int ntiles, const SolveDesc *dd,
                             const int *gsolve, int kind, hipStream_t st) {
  if (kind == 8)
    hipLaunchKernelGGL((k_zgemm_bf3<MB, NB, float2, true>), dim3(ntiles), dim3(256), 0, st, groups, tiles, dd, gsolve);
  else if (kind == 7)
    hipLaunchKernelGGL((k_zgemm_bf3<MB, NB, float2>), dim3(ntiles), dim3(256), 0, st, groups, tiles, dd, gsolve);
  else
    hipLaunchKernelGGL((k_zgemm_bf3<MB, NB, double2>), dim3(ntiles), dim3(256), 0, st, groups, tiles, dd, gsolve);
}

void launch_gemm_shape(const GemmGroup *groups, const GemmTile *tiles, int ntiles, const SolveDesc *dd,
                       const int *gsolve, int kind, int mb, int nb, int lb, hipStream_t st) {
  if (ntiles <= 0) return;
  if ((kind == 6 || kind == 7 || kind == 8) && mb > 0) {
    switch (mb * 10 + nb) {
      case 11: launch_shape_bf3<1, 1>(groups, tiles, ntiles, dd, gsolve, kind, st); break;
      case 12: launch_shape_bf3<1, 2>(groups, tiles, ntiles, dd, gsolve, kind, st); break;
      case 13: launch_shape_bf3<1, 3>(groups, tiles, ntiles, dd, gsolve, kind, st); break;
      case 14: launch_shape_bf3<1, 4>(groups, tiles, ntiles, dd, gsolve, kind, st); break;
      case 15: launch_shape_bf3<1, 5>(groups, tiles, ntiles, dd, gsolve, kind, st); break;
      case 16: launch_shape_bf3<1, 6>(groups, tiles, ntiles, dd, gsolve, kind, st); break;
      case 18: launch_shape_bf3<1, 8>(groups, tiles, ntiles, dd, gsolve, kind, st); break;
      case 21: launch_shape_bf3<2, 1>(groups, tiles, ntiles, dd, gsolve, kind, st); break;
      case 22: launch_shape_bf3<2, 2>(groups, tiles, ntiles, dd, gsolve, kind, st); break;
      case 23: launch_shape_bf3<2, 3>(groups, tiles, ntiles, dd, gsolve, kind, st); break;
      case 24: launch_shape_bf3<2, 4>(groups, tiles, ntiles, dd, gsolve, kind, st); break;
      case 41: launch_shape_bf3<4, 1>(groups, tiles, ntiles, dd, gsolve, kind, st); break;
      default: break;
    }
    return;
  }
  if (lb == 4 && mb == 1 && nb == 2) {   // pp ladder, eight-wave LDS pipeline, 64 x 32 tiles
    launch_shape_ppl8(groups, tiles, ntiles, dd, gsolve, kind, st);
    return;
  }
  if (lb == 3 && mb == 1) {   // pp ladder through the LDS load pipeline
    switch (nb) {
      case 1: launch_shape_ppl<1>(groups, tiles, ntiles, dd, gsolve, kind, st); break;
      case 2: launch_shape_ppl<2>(groups, tiles, ntiles, dd, gsolve, kind, st); break;
      default: break;
    }
    return;
  }
  if (lb == 2 && mb == 1) {
    switch (nb) {
      case 1: launch_shape_la<1>(groups, tiles, ntiles, dd, gsolve, kind, st); break;
      case 2: launch_shape_la<2>(groups, tiles, ntiles, dd, gsolve, kind, st); break;
      case 3: launch_shape_la<3>(groups, tiles, ntiles, dd, gsolve, kind, st); break;
      default: break;
    }
    return;
  }
  if (lb && mb > 0) {
    switch (mb * 10 + nb) {
      case 11: launch_shape_lb<1, 1>(groups, tiles, ntiles, dd, gsolve, kind, st); break;
      case 12: launch_shape_lb<1, 2>(groups, tiles, ntiles, dd, gsolve, kind, st); break;
      case 13: launch_shape_lb<1, 3>(groups, tiles, ntiles, dd, gsolve, kind, st); break;
      case 14: launch_shape_lb<1, 4>(groups, tiles, ntiles, dd, gsolve, kind, st); break;
      case 15: launch_shape_lb<1, 5>(groups, tiles, ntiles, dd, gsolve, kind, st); break;
      case 21: launch_shape_lb<2, 1>(groups, tiles, ntiles, dd, gsolve, kind, st); break;
      case 22: launch_shape_lb<2, 2>(groups, tiles, ntiles, dd, gsolve, kind, st); break;
      case 41: launch_shape_lb<4, 1>(groups, tiles, ntiles, dd, gsolve, kind, st); break;
      default: break;
    }
    return;
  }
  if (mb == 0 && kind == 8) {   // thin VALU kernel on FP32 blocks
    hipLaunchKernelGGL((k_zgemm_thin<float2, 0, true>), dim3(ntiles), dim3(256), 0, st, groups, tiles, dd, gsolve);
    return;
  }
  if (mb == 0) {   // thin VALU kernel, N <= 4, FP64 arithmetic
    const bool af = kind == 2 || kind == 4 || kind == 5 || kind == 7;   // A stored as float2
    const bool pp = kind >= 3 && kind <= 5;
    if (af && pp) hipLaunchKernelGGL((k_zgemm_thin<float2, 1>), dim3(ntiles), dim3(256), 0, st, groups, tiles, dd, gsolve);
    else if (af) hipLaunchKernelGGL((k_zgemm_thin<float2, 0>), dim3(ntiles), dim3(256), 0, st, groups, tiles, dd, gsolve);
    else if (pp) hipLaunchKernelGGL((k_zgemm_thin<double2, 1>), dim3(ntiles), dim3(256), 0, st, groups, tiles, dd, gsolve);
    else hipLaunchKernelGGL((k_zgemm_thin<double2, 0>), dim3(ntiles), dim3(256), 0, st, groups, tiles, dd, gsolve);
    return;
  }
  const int code = mb * 10 + nb;
  switch (code) {
    case 11: launch_shape<1, 1>(groups, tiles, ntiles, dd, gsolve, kind, st); break;
    case 12: launch_shape<1, 2>(groups, tiles, ntiles, dd, gsolve, kind, st); break;
    case 13: launch_shape<1, 3>(groups, tiles, ntiles, dd, gsolve, kind, st); break;
    case 14: launch_shape<1, 4>(groups, tiles, ntiles, dd, gsolve, kind, st); break;
    case 15: launch_shape<1, 5>(groups, tiles, ntiles, dd, gsolve, kind, st); break;
    case 21: launch_shape<2, 1>(groups, tiles, ntiles, dd, gsolve, kind, st); break;
    case 22: launch_shape<2, 2>(groups, tiles, ntiles, dd, gsolve, kind, st); break;
    case 41: launch_shape<4, 1>(groups, tiles, ntiles, dd, gsolve, kind, st); break;
    default: break;
  }
}

// Plain FP64 VALU cross-check on uniform 64x64 tiles.
void launch_gemm_valu(const GemmGroup *groups, const GemmTile *tiles, int ntiles, const SolveDesc *dd,
                      const int *gsolve, hipStream_t st) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_zgemm_valu, dim3(ntiles), dim3(256), 0, st, groups, tiles, dd, gsolve);
}

}  // namespace ccgpu
