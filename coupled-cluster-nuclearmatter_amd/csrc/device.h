// Device-side descriptors shared by the kernels and the host runtime.
#pragma once
#include "interaction.h"
#include <hip/hip_runtime.h>
#include <cstdint>

namespace ccgpu {

// Ordered pair reference packed for the device: ch = channel (-1 absent),
// v = idx * 2 + (sign < 0).
struct DPair {
  int ch;
  int v;
};

// Everything one solve's kernels read or write.  All pointers are device
// pointers; arrays indexed by unique amplitude u have length nT.
struct SolveDesc {
  int A, nsp, np, nch, nphc;
  int64_t nT, nRt, nI4;
  int active;  // 0: converged, kernels skip this solve
  // orbitals
  const double *kx, *ky, *kz, *e;
  const int *nx, *ny, *nz, *sz, *tz, *partner;
  // channels
  const int *t_ch, *npp, *nhh;
  const int64_t *off_t, *off_pp, *off_hh, *off_vpp, *off_vhh;
  const int *pp_a, *pp_b, *hh_i, *hh_j;
  // per unique amplitude its orbitals a | b << 16 | i << 32 | j << 48
  const uint64_t *lab;
  const int *ob_rowB, *ob_rowA, *ob_colJ, *ob_colI;
  // I6 by rows: channel of each unique pp row, CSR lists of partials per particle
  int64_t nrows;
  const int *row_ch;
  const int *i6_dptr, *i6_didx, *i6_optr, *i6_oidx;
  double2 *p6;
  const DPair *ppref, *hhref;
  // parity image of each orbital, n -> -n (null unless the lattice maps onto
  // itself with k -> -k bit for bit and holes onto holes): the 3NF fill of
  // V_pppp evaluates one element of each parity pair
  const int *pimg;
  // Pandya
  const int *nph, *nhp;
  const int64_t *off_ph, *off_hp, *off_rt, *off_i4, *off_vt;
  const int *ph_a, *ph_i, *hp_j, *hp_b;
  const int *ring4;   // four Pandya positions per unique amplitude (every ring block below 2^31 elements)
  // amplitudes and blocks
  double2 *T, *R, *R0, *Vhhpp, *Vhhhh, *X, *Vpppp;
  double2 *Rh;   // hh ladder T I9 of the residual, added to R by the ring scatter
  float2 *Vpppp32;
  double2 *Tt, *Rt, *I4, *W0, *Vt;
  // FP32 ring blocks of the BF16x3 residual (null until the FP32 copies are
  // made): T~, R~, I4 per iteration, V~ and W0 copied from the FP64 blocks
  float2 *Tt32, *Rt32, *I4_32, *W0_32, *Vt32;
  // T as the BF16x3 pp ladder's B operand, split once per residual: per
  // amplitude the bf16 hi and lo pairs of (re, -im) and of (im, re)
  uint4 *Tbf;
  // one-body
  double2 *fdiag, *fpart;   // f(p,p) and f(p,partner(p))
  double2 *dh, *I7off, *dp, *I6off;
  double2 *e0;              // device E0
  // DIIS history (depth x nT) and reductions.  With the Anderson update hist
  // holds the mixed iterates y_j = T_j + beta f_j and histf the residual
  // vectors f_j = R_j/D_j - T_j in FP32 (for the Gram matrix only).
  double2 *hist;
  float2 *histf;
  double2 *red;             // reduction partials
  double2 *scal;            // small scalar outputs (energy, Gram)
};

// One member of a grouped complex GEMM: C = alpha * A * B + beta * C0.
// With nsplit > 1 the K range is split over tiles that write raw partial
// products to P (nsplit blocks of M x N); k_splitk_reduce then forms C.
struct GemmGroup {
  const void *A;
  const double2 *B;
  double2 *C;
  const double2 *C0;
  double2 *P;
  int M, N, K;
  int lda, ldb, ldc, ldc0;
  int nsplit;
  double alpha, beta;
};

struct GemmTile {
  int g, m0, n0, ks;   // group, tile origin, split index
  int kb, ke, pad0, pad1;  // k range of this tile
};

// (T) channel tables and, for handed-over interactions, the open-addressing
// table of the <cb||kd> and <jk||lc> elements (key ((p*nsp+q)*nsp+r)*nsp+s,
// empty slots ~0, capacity a power of two at least twice the entries).
struct TriplesDesc {
  // channel lookup over (Tz3, K3): index = ((tz - tz0) * nk + (kx - k0)) * nk^2 + ...
  int k0, nk, tz0, ntz;
  const int *chan;        // dense table -> channel or -1
  const int *hptr;        // CSR over channels of hole triples
  const int *hijk;        // 3 ints per hole triple
  const int *orbitof;     // dense (n, sz, tz) -> orbital, -1 absent
  int nmax;
  const int *pair_a, *pair_b;  // every particle pair a<b (all momenta)
  int64_t npairs;
  const unsigned long long *vkeys;  // null: elements from the built-in generator
  const double2 *vvals;
  unsigned long long vmask;
  // built-in generator: radial factors of the three Minnesota Gaussians for
  // q^2 = (pi/L)^2 m, m = |n_p - n_q - n_r + n_s|^2 (twists cancel), 3 per m
  const double *mtab;
  int mmax;
  // built-in chiral NN (used when chiral != 0)
  int chiral;
  ChiralDev ch;
  // dense element tables (null: elements evaluated in place): vhppp holds
  // <cb||kd> at ((c-A) np + (b-A)) A + k, times 2 for the spin of d (d's
  // momentum and isospin follow from conservation); vhhhp holds <jk||lc> at
  // (j A + k) np + (c-A), times 2 for the spin of l
  const double2 *vhppp, *vhhhp;
  // 1: the lattice maps onto itself under n -> -n with k -> -k exactly (no
  // twist) and holes onto holes, so <cb||kd> = <Pc Pb||Pk Pd> (the NN and 3NF
  // elements are even under parity) and the table fill evaluates one of each
  // parity pair
  int parity;
  // normal-ordered 3NF in <cb||kd> (the reference adds it to vnn_hppp, not to
  // vnn_hhhp: ccm_iter.f90:1641-1690)
  int tnf;
  Chiral3Dev c3;
  // the antisymmetrized three-body element W3 = <abc||ijk> (chiral_3nf_asym)
  // of the CCDT-1 triples and of (T) at tnf_approx 3
  Chiral3Dev c3w;
};

// Warm-start network: number of parameters (flat FP32 vector, layout in
// mlp.hip) and the buffers of one training step: feature-major activations and
// pre-activation gradients of the minibatch (B columns, B a multiple of 64).
constexpr int kMlpParamCount = 9473;
struct MlpTrainBufs {
  uint16_t *x1t, *h1t, *h2t, *h3t;   // [16][B], [64][B] x 3
  uint16_t *dz1t, *dz2t, *dz3t;      // [64][B] x 3
  float *g;                          // [B] dL/dy
  float *loss_part;                  // [gridDim.x of the pass]
};

}  // namespace ccgpu
