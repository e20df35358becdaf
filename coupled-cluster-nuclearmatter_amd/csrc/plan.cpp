// Host planning of the lattice: basis, channels, pair maps, Pandya maps.
//
// Basis: allocate_sp_data / setup_sp_data (ccm_main.f90:878-1197).
// Channels: setup_channel_structures (ccm_iter.f90:175-603) groups ordered pairs
// by K = n_a + n_b and Tz; here only unique pairs (a<b, i<j) are stored, the other
// orderings follow from antisymmetry.  Pandya channels follow
// setup_ph_channel_structures (ccm_iter.f90:1816-2171).
#include "plan.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <numeric>
#include <stdexcept>
#include <tuple>

namespace ccgpu {

namespace {
// Constants as declared by the reference (ccm_modules.f90:91-95): the nucleon
// mass is a default-real literal (float32 939.565).
const double kHbarc = 197.326968;
const double kPMass = (double)939.565f;
const double kPi = 3.14159265358979;
}  // namespace

// Ascending order of arr as the reference's `index` leaves it
// (ccm_main.f90:2654-2732): quicksort with the median of three as pivot,
// partitions below 7 elements finished by insertion, the larger part deferred
// on a stack.  The order of equal keys is that routine's, which decides which
// of several degenerate momenta become holes.  Returns 0-based positions.
std::vector<int> reference_index_order(const std::vector<double> &arr) {
  const int n = (int)arr.size();
  std::vector<int> ix(n + 1);  // 1-based positions, ix[r] = element at rank r
  for (int r = 0; r <= n; ++r) ix[r] = r;
  auto key = [&](int r) { return arr[ix[r] - 1]; };
  std::vector<int> stack;
  int lo = 1, hi = n;
  for (;;) {
    if (hi - lo < 7) {
      for (int j = lo + 1; j <= hi; ++j) {
        const int moving = ix[j];
        const double v = arr[moving - 1];
        int i = j - 1;
        for (; i >= 1 && key(i) > v; --i) ix[i + 1] = ix[i];
        ix[i + 1] = moving;
      }
      if (stack.empty()) break;
      hi = stack.back(); stack.pop_back();
      lo = stack.back(); stack.pop_back();
      continue;
    }
    std::swap(ix[(lo + hi) / 2], ix[lo + 1]);
    if (key(lo + 1) > key(hi)) std::swap(ix[lo + 1], ix[hi]);
    if (key(lo) > key(hi)) std::swap(ix[lo], ix[hi]);
    if (key(lo + 1) > key(lo)) std::swap(ix[lo + 1], ix[lo]);
    const int pivot = ix[lo];
    const double v = arr[pivot - 1];
    int i = lo + 1, j = hi;
    for (;;) {
      do ++i; while (key(i) < v);
      do --j; while (key(j) > v);
      if (j < i) break;
      std::swap(ix[i], ix[j]);
    }
    ix[lo] = ix[j];
    ix[j] = pivot;
    if (hi - i + 1 >= j - lo) {
      stack.push_back(i); stack.push_back(hi);
      hi = j - 1;
    } else {
      stack.push_back(lo); stack.push_back(j - 1);
      lo = i;
    }
  }
  std::vector<int> out(n);
  for (int r = 0; r < n; ++r) out[r] = ix[r + 1] - 1;
  return out;
}

Orbitals build_orbitals(int A, int matter, double rho, int nmax, const double twist[3]) {
  if (A <= 0) throw std::runtime_error("A must be positive");
  if (nmax < 0 || nmax > 12) throw std::runtime_error("nmax out of range [0,12]");
  if (!(rho > 0)) throw std::runtime_error("density must be positive");
  Orbitals o;
  std::vector<int> tzs;
  if (matter == 0) {
    tzs = {1};
    o.kf = std::pow(3.0 * kPi * kPi * rho, 1.0 / 3.0);
  } else if (matter == 1) {
    tzs = {-1, 1};
    o.kf = std::pow(3.0 * kPi * kPi * rho / 2.0, 1.0 / 3.0);
  } else {
    throw std::runtime_error("matter must be 0 (pnm) or 1 (snm)");
  }
  const double L = std::pow((double)A / rho, 1.0 / 3.0);
  std::vector<int> ns = {0};
  for (int i = 1; i <= nmax; ++i) { ns.push_back(-i); ns.push_back(i); }
  struct M { double k2; int n[3]; double k[3]; int order; };
  std::vector<M> mesh;
  int ord = 0;
  for (int x : ns)
    for (int y : ns)
      for (int z : ns) {
        M m;
        m.n[0] = x; m.n[1] = y; m.n[2] = z;
        m.k[0] = (2.0 * x * kPi + twist[0]) / L;
        m.k[1] = (2.0 * y * kPi + twist[1]) / L;
        m.k[2] = (2.0 * z * kPi + twist[2]) / L;
        m.k2 = m.k[0] * m.k[0] + m.k[1] * m.k[1] + m.k[2] * m.k[2];
        m.order = ord++;
        mesh.push_back(m);
      }
  // mesh points by k^2 in the reference's sort order (setup_sp_data,
  // ccm_main.f90:1100-1110)
  std::vector<double> k2(mesh.size());
  for (size_t q = 0; q < mesh.size(); ++q) k2[q] = mesh[q].k2;
  std::vector<M> sorted;
  for (int q : reference_index_order(k2)) sorted.push_back(mesh[q]);
  for (const M &m : sorted)
    for (int t : tzs)
      for (int s : {-1, 1}) {
        o.nx.push_back(m.n[0]); o.ny.push_back(m.n[1]); o.nz.push_back(m.n[2]);
        o.sz.push_back(s); o.tz.push_back(t);
        o.kx.push_back(m.k[0]); o.ky.push_back(m.k[1]); o.kz.push_back(m.k[2]);
        double k2 = m.k[0] * m.k[0] + m.k[1] * m.k[1] + m.k[2] * m.k[2];
        o.e.push_back(k2 * (kHbarc * kHbarc) / kPMass / 2.0);
      }
  o.nsp = (int)o.e.size();
  if (A >= o.nsp) throw std::runtime_error("A exceeds the number of lattice states");
  o.A = A;
  o.box = L;
  o.volume = L * L * L;
  o.rho = rho;
  o.nmax = nmax;
  return o;
}

namespace {
struct Key {
  int tz, x, y, z;
  bool operator<(const Key &o) const { return std::tie(tz, x, y, z) < std::tie(o.tz, o.x, o.y, o.z); }
};
}  // namespace

void build_plan(const Orbitals &orb, Plan &P) {
  P.orb = orb;
  const int A = orb.A, nsp = orb.nsp, np = nsp - A;
  P.A = A; P.nsp = nsp; P.np = np;
  if (A < 2 || np < 2) throw std::runtime_error("need at least two holes and two particles");

  // spin partners: same momentum and isospin, other spin
  P.partner.assign(nsp, -1);
  {
    std::map<std::tuple<int, int, int, int, int>, int> lut;
    for (int x = 0; x < nsp; ++x) lut[{orb.nx[x], orb.ny[x], orb.nz[x], orb.sz[x], orb.tz[x]}] = x;
    for (int x = 0; x < nsp; ++x) {
      auto it = lut.find({orb.nx[x], orb.ny[x], orb.nz[x], -orb.sz[x], orb.tz[x]});
      if (it != lut.end()) P.partner[x] = it->second;
    }
  }

  auto key2 = [&](int a, int b) {
    return Key{orb.tz[a] + orb.tz[b], orb.nx[a] + orb.nx[b], orb.ny[a] + orb.ny[b], orb.nz[a] + orb.nz[b]};
  };
  std::map<Key, std::vector<std::pair<int, int>>> hhm, ppm;
  for (int i = 0; i < A; ++i)
    for (int j = i + 1; j < A; ++j) hhm[key2(i, j)].push_back({i, j});
  for (int a = A; a < nsp; ++a)
    for (int b = a + 1; b < nsp; ++b) {
      Key k = key2(a, b);
      if (hhm.count(k)) ppm[k].push_back({a, b});
    }
  P.ppref.assign((size_t)np * np, PairRef());
  P.hhref.assign((size_t)A * A, PairRef());
  P.off_pp = {0}; P.off_hh = {0}; P.off_t = {0}; P.off_vpp = {0}; P.off_vhh = {0};
  int c = 0;
  for (auto &kv : hhm) {
    auto pit = ppm.find(kv.first);
    if (pit == ppm.end()) continue;
    const auto &hl = kv.second;
    const auto &pl = pit->second;
    P.key_tz.push_back(kv.first.tz); P.key_kx.push_back(kv.first.x);
    P.key_ky.push_back(kv.first.y); P.key_kz.push_back(kv.first.z);
    for (size_t r = 0; r < pl.size(); ++r) {
      int a = pl[r].first, b = pl[r].second;
      P.pp_a.push_back(a); P.pp_b.push_back(b);
      P.ppref[(size_t)(a - A) * np + (b - A)] = PairRef{c, (int)r, 1};
      P.ppref[(size_t)(b - A) * np + (a - A)] = PairRef{c, (int)r, -1};
    }
    for (size_t s = 0; s < hl.size(); ++s) {
      int i = hl[s].first, j = hl[s].second;
      P.hh_i.push_back(i); P.hh_j.push_back(j);
      P.hhref[(size_t)i * A + j] = PairRef{c, (int)s, 1};
      P.hhref[(size_t)j * A + i] = PairRef{c, (int)s, -1};
    }
    int npc = (int)pl.size(), nhc = (int)hl.size();
    P.npp.push_back(npc); P.nhh.push_back(nhc);
    P.off_pp.push_back(P.off_pp.back() + npc);
    P.off_hh.push_back(P.off_hh.back() + nhc);
    P.off_t.push_back(P.off_t.back() + (int64_t)npc * nhc);
    P.off_vpp.push_back(P.off_vpp.back() + (int64_t)npc * npc);
    P.off_vhh.push_back(P.off_vhh.back() + (int64_t)nhc * nhc);
    P.max_npp = std::max(P.max_npp, npc);
    P.max_nhh = std::max(P.max_nhh, nhc);
    // reference-shaped ordered count includes a=b and i=j pairs
    ++c;
  }
  P.nch = c;
  P.nT = P.off_t.back();
  P.nVpp = P.off_vpp.back();
  P.nVhh = P.off_vhh.back();
  if (P.nch == 0) throw std::runtime_error("no (K,Tz) channel has both hole and particle pairs");

  // ordered-pair count of the reference (all a,b and i,j, including a=b, i=j)
  {
    std::map<Key, int64_t> ho, po;
    for (int i = 0; i < A; ++i)
      for (int j = 0; j < A; ++j) ho[key2(i, j)]++;
    for (int a = A; a < nsp; ++a)
      for (int b = A; b < nsp; ++b) {
        Key k = key2(a, b);
        if (ho.count(k)) po[k]++;
      }
    P.t2_ordered = 0;
    for (auto &kv : po) P.t2_ordered += kv.second * ho[kv.first];
  }

  P.t_ch.resize(P.nT);
  for (int ch = 0; ch < P.nch; ++ch)
    for (int64_t u = P.off_t[ch]; u < P.off_t[ch + 1]; ++u) P.t_ch[u] = ch;

  // one-body maps
  auto pack = [](int idx, int sign) { return idx < 0 || sign == 0 ? -1 : idx * 4 + (sign + 1); };
  P.ob_rowB.resize(P.pp_a.size());
  P.ob_rowA.resize(P.pp_a.size());
  for (size_t r = 0; r < P.pp_a.size(); ++r) {
    int a = P.pp_a[r], b = P.pp_b[r];
    int bb = P.partner[b], aa = P.partner[a];
    int vB = -1, vA = -1;
    if (bb >= A && bb != a) { PairRef pr = P.pp(a, bb); if (pr.sign) vB = pack(pr.idx, pr.sign); }
    if (aa >= A && aa != b) { PairRef pr = P.pp(b, aa); if (pr.sign) vA = pack(pr.idx, pr.sign); }
    P.ob_rowB[r] = vB;
    P.ob_rowA[r] = vA;
  }
  P.ob_colJ.resize(P.hh_i.size());
  P.ob_colI.resize(P.hh_i.size());
  for (size_t s = 0; s < P.hh_i.size(); ++s) {
    int i = P.hh_i[s], j = P.hh_j[s];
    int jj = P.partner[j], ii = P.partner[i];
    int vJ = -1, vI = -1;
    if (jj >= 0 && jj < A && jj != i) { PairRef pr = P.hh(i, jj); if (pr.sign) vJ = pack(pr.idx, pr.sign); }
    if (ii >= 0 && ii < A && ii != j) { PairRef pr = P.hh(j, ii); if (pr.sign) vI = pack(pr.idx, pr.sign); }
    P.ob_colJ[s] = vJ;
    P.ob_colI[s] = vI;
  }

  // I6 partial lists
  {
    size_t nr = P.pp_a.size();
    P.row_ch.resize(nr);
    for (int ch = 0; ch < P.nch; ++ch)
      for (int64_t r = P.off_pp[ch]; r < P.off_pp[ch + 1]; ++r) P.row_ch[r] = ch;
    std::vector<std::vector<int>> dl(np), ol(np);
    for (size_t r = 0; r < nr; ++r) {
      int x = P.pp_a[r], y = P.pp_b[r];
      dl[y - A].push_back((int)(3 * r + 0));
      dl[x - A].push_back((int)(3 * r + 0));
      if (P.ob_rowB[r] >= 0) ol[y - A].push_back((int)(3 * r + 1));
      if (P.ob_rowA[r] >= 0) ol[x - A].push_back((int)(3 * r + 2));
    }
    P.i6_dptr = {0};
    P.i6_optr = {0};
    for (int a = 0; a < np; ++a) {
      for (int v : dl[a]) P.i6_didx.push_back(v);
      for (int v : ol[a]) P.i6_oidx.push_back(v);
      P.i6_dptr.push_back((int)P.i6_didx.size());
      P.i6_optr.push_back((int)P.i6_oidx.size());
    }
  }

  // Pandya channels: ph pairs (a,i) with K' = n_a - n_i, hp pairs (j,b) with n_j - n_b
  auto keyd = [&](int x, int y) {
    return Key{orb.tz[x] - orb.tz[y], orb.nx[x] - orb.nx[y], orb.ny[x] - orb.ny[y], orb.nz[x] - orb.nz[y]};
  };
  std::map<Key, std::vector<std::pair<int, int>>> phm, hpm;
  for (int a = A; a < nsp; ++a)
    for (int i = 0; i < A; ++i) phm[keyd(a, i)].push_back({a, i});
  for (int j = 0; j < A; ++j)
    for (int b = A; b < nsp; ++b) hpm[keyd(j, b)].push_back({j, b});
  std::vector<int> phrow((size_t)np * A, -1), phch((size_t)np * A, -1);
  std::vector<int> hpcol((size_t)A * np, -1);
  P.off_ph = {0}; P.off_hp = {0}; P.off_rt = {0}; P.off_i4 = {0}; P.off_vt = {0};
  int K = 0;
  for (auto &kv : phm) {
    auto hit = hpm.find(kv.first);
    if (hit == hpm.end()) continue;
    const auto &pl = kv.second;
    const auto &hl = hit->second;
    for (size_t r = 0; r < pl.size(); ++r) {
      P.ph_a.push_back(pl[r].first); P.ph_i.push_back(pl[r].second);
      phrow[(size_t)(pl[r].first - A) * A + pl[r].second] = (int)r;
      phch[(size_t)(pl[r].first - A) * A + pl[r].second] = K;
    }
    for (size_t s = 0; s < hl.size(); ++s) {
      P.hp_j.push_back(hl[s].first); P.hp_b.push_back(hl[s].second);
      hpcol[(size_t)hl[s].first * np + (hl[s].second - A)] = (int)s;
    }
    int nphk = (int)pl.size(), nhpk = (int)hl.size();
    P.nph.push_back(nphk); P.nhp.push_back(nhpk);
    P.off_ph.push_back(P.off_ph.back() + nphk);
    P.off_hp.push_back(P.off_hp.back() + nhpk);
    P.off_rt.push_back(P.off_rt.back() + (int64_t)nphk * nhpk);
    P.off_i4.push_back(P.off_i4.back() + (int64_t)nhpk * nhpk);
    P.off_vt.push_back(P.off_vt.back() + (int64_t)nhpk * nphk);
    ++K;
  }
  P.nphc = K;
  P.nRt = P.off_rt.back();
  P.nI4 = P.off_i4.back();

  auto pos = [&](int a, int i, int j, int b) -> int64_t {
    size_t pa = (size_t)(a - A) * A + i;
    int k = phch[pa];
    if (k < 0) return -1;
    int r = phrow[pa];
    int s = hpcol[(size_t)j * np + (b - A)];
    if (s < 0) return -1;
    return P.off_rt[k] + (int64_t)r * P.nhp[k] + s;
  };
  P.ring4.resize((size_t)P.nT * 4);
  for (int ch = 0; ch < P.nch; ++ch) {
    for (int r = 0; r < P.npp[ch]; ++r) {
      int a = P.pp_a[P.off_pp[ch] + r], b = P.pp_b[P.off_pp[ch] + r];
      for (int s = 0; s < P.nhh[ch]; ++s) {
        int i = P.hh_i[P.off_hh[ch] + s], j = P.hh_j[P.off_hh[ch] + s];
        int64_t u = P.off_t[ch] + (int64_t)r * P.nhh[ch] + s;
        P.ring4[u * 4 + 0] = pos(a, i, j, b);
        P.ring4[u * 4 + 1] = pos(b, i, j, a);
        P.ring4[u * 4 + 2] = pos(a, j, i, b);
        P.ring4[u * 4 + 3] = pos(b, j, i, a);
        for (int q = 0; q < 4; ++q)
          if (P.ring4[u * 4 + q] < 0) throw std::runtime_error("internal: Pandya position missing");
      }
    }
  }
}

}  // namespace ccgpu
