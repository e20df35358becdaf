// Host-side plan of one lattice solve: single-particle basis, (K,Tz) channels of
// unique pairs, Pandya channels and every index map the kernels read.
#pragma once
#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

namespace ccgpu {

struct Orbitals {
  int nsp = 0, A = 0;
  std::vector<int> nx, ny, nz, sz, tz;
  std::vector<double> kx, ky, kz, e;
  double box = 0, volume = 0, kf = 0, rho = 0;
  int nmax = 0;
};

// Ordered pair -> (channel, index within the channel, orientation sign).
// sign = +1 for the stored orientation (x<y), -1 for the swapped one, 0 if absent.
struct PairRef {
  int ch = -1;
  int idx = -1;
  int sign = 0;
};

struct Plan {
  Orbitals orb;
  int A = 0, nsp = 0, np = 0;
  std::vector<int> partner;  // other spin state of the same momentum and isospin

  // ---- (K, Tz) channels with both hole pairs and particle pairs ----
  int nch = 0;
  std::vector<int> npp, nhh;                  // unique pair counts per channel
  std::vector<int64_t> off_pp, off_hh;        // offsets into pair lists
  std::vector<int64_t> off_t;                 // T/R blocks: npp x nhh
  std::vector<int64_t> off_vpp;               // V_pppp blocks: npp x npp
  std::vector<int64_t> off_vhh;               // V_hhhh / X / I9 blocks: nhh x nhh
  std::vector<int> pp_a, pp_b, hh_i, hh_j;    // concatenated unique pairs
  std::vector<int> key_tz, key_kx, key_ky, key_kz;
  int64_t nT = 0, nVpp = 0, nVhh = 0;
  int max_npp = 0, max_nhh = 0;

  // ordered pair maps (particle index p-A, hole index)
  std::vector<PairRef> ppref;  // np * np
  std::vector<PairRef> hhref;  // A * A
  PairRef pp(int a, int b) const { return ppref[(size_t)(a - A) * np + (b - A)]; }
  PairRef hh(int i, int j) const { return hhref[(size_t)i * A + j]; }

  // per unique T element: channel id (for kernels that walk the flat vector)
  std::vector<int> t_ch;

  // one-body (I6/I7) maps: for each unique pp pair (a<b) within its channel the
  // rows of (a, partner(b)) and (b, partner(a)) with signs (0 = absent), and for
  // each unique hh pair (i<j) the columns of (i, partner(j)) and (j, partner(i)).
  std::vector<int> ob_rowB, ob_rowA, ob_colJ, ob_colI;   // packed: idx*4 + (sign+1)
  // I6 row partials (3 per unique pp row (x<y): diag of x and y, I6(y,ybar),
  // I6(x,xbar)) and, per particle a, the CSR lists of partials that sum into
  // I6(a,a) and I6(a,abar), in increasing row order
  std::vector<int> row_ch;
  std::vector<int> i6_dptr, i6_didx, i6_optr, i6_oidx;

  // ---- Pandya channels K' = n_a - n_i ----
  int nphc = 0;
  std::vector<int> nph, nhp;
  std::vector<int64_t> off_ph, off_hp;    // pair lists
  std::vector<int64_t> off_rt;            // T~/R~ / V~^T blocks: nph x nhp
  std::vector<int64_t> off_i4;            // I4/W0 blocks: nhp x nhp
  std::vector<int64_t> off_vt;            // V~ blocks: nhp x nph
  std::vector<int> ph_a, ph_i, hp_j, hp_b;
  int64_t nRt = 0, nI4 = 0;
  // four Pandya positions of every unique T element (a<b,i<j):
  //   ((a,i),(j,b)) +,  ((b,i),(j,a)) -,  ((a,j),(i,b)) -,  ((b,j),(i,a)) +
  std::vector<int64_t> ring4;   // nT * 4, -1 if absent

  int64_t t2_ordered = 0;
};

// Builds the orbital table exactly as setup_sp_data does, degenerate momenta in
// the order of the reference's quicksort.
Orbitals build_orbitals(int A, int matter, double rho, int nmax, const double twist[3]);
// Builds all channel structures and maps from an orbital table.
void build_plan(const Orbitals &orb, Plan &plan);

}  // namespace ccgpu
