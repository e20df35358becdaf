// C-ABI of the CCD/(T) engine: host runtime (device arenas, work lists, the
// iteration driver, DIIS solve, batching, hipGraph replay, timing).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <random>
#include <tuple>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/ccgpu.h"
#include "chiral.h"
#include "device.h"
#include "plan.h"

namespace ccgpu {
// kernels.hip
void launch_fill_blocks(const SolveDesc &d, double volume, int nch, int nphc, int max_npp, int max_nhh,
                        int max_nhp, hipStream_t st);
void launch_fill_blocks_chiral(const SolveDesc &d, const ChiralDev &c, int nch, int nphc, int max_npp,
                               int max_nhh, int max_nhp, hipStream_t st);
void launch_elements(const SolveDesc &d, double volume, const ChiralDev *c, const int4 *pqrs, int n, double2 *out,
                     hipStream_t st);
void launch_3nf(const SolveDesc &d, const Chiral3Dev &c, int approx, int nch, int nphc, int max_npp, int max_nhh,
                int max_nhp, hipStream_t st);
void launch_derive(const SolveDesc &d, int nch, int nphc, int max_npp, int max_nhh, int max_nph, int max_nhp,
                   hipStream_t st);
void launch_to_f32(const double2 *src, float2 *dst, int64_t n, hipStream_t st);
void launch_to_f32_rows(const double2 *src, float2 *dst, const int64_t *off, const int64_t *off32, const int *npp,
                        int nch, int max_npp, hipStream_t st);
void launch_mbpt2(const SolveDesc *dd, int nsolve, int64_t maxT, hipStream_t st);
void launch_energy(const SolveDesc *dd, int nsolve, int slot, hipStream_t st);
void launch_i7(const SolveDesc *dd, int nsolve, int maxA, hipStream_t st);
void launch_i6(const SolveDesc *dd, int nsolve, int maxnp, int64_t maxrows, hipStream_t st);
void launch_i9(const SolveDesc *dd, int nsolve, int64_t maxvhh, hipStream_t st);
void launch_resid_init(const SolveDesc *dd, int nsolve, int64_t maxT, bool onebody, hipStream_t st);
void launch_ring_gather(const SolveDesc *dd, int nsolve, int64_t maxT, bool f32, hipStream_t st);
void launch_t_split(const SolveDesc *dd, int nsolve, int64_t maxT, hipStream_t st);
void launch_ring_scatter(const SolveDesc *dd, int nsolve, int64_t maxT, bool f32, bool ring, bool hh,
                         hipStream_t st);
void launch_update(const SolveDesc *dd, int nsolve, int64_t maxT, const int *slot, hipStream_t st);
void launch_gram(const SolveDesc *dd, int nsolve, const int *slots, int nstep, hipStream_t st);
void launch_aa_update(const SolveDesc *dd, int nsolve, const int *slots, double beta, hipStream_t st);
void launch_scatter_values(double2 *dst, const int64_t *idx, const double2 *val, int64_t n, hipStream_t st);
void launch_fill_vpppp(const SolveDesc &d, double volume, const ChiralDev *c, int nch, int max_npp, hipStream_t st);
void launch_share_gather(const SolveDesc *dd, int nsolve, int64_t maxT, const uint32_t *const *maps, double2 *Tu,
                         hipStream_t st);
void launch_share_scatter(const SolveDesc *dd, int nsolve, int64_t maxT, const uint32_t *const *maps,
                          const double2 *Ru, hipStream_t st);
void launch_aa_combine(const SolveDesc *dd, int nsolve, int64_t maxT, const int *slots, const double *coef,
                       hipStream_t st);
void launch_diis_axpy(const SolveDesc *dd, int nsolve, int64_t maxT, const int *slots, const double2 *coef,
                      int nvec, hipStream_t st);
void launch_gemm_shape(const GemmGroup *groups, const GemmTile *tiles, int ntiles, const SolveDesc *dd,
                       const int *gsolve, int kind, int mb, int nb, int lb, hipStream_t st);
void launch_gemm_valu(const GemmGroup *groups, const GemmTile *tiles, int ntiles, const SolveDesc *dd,
                      const int *gsolve, hipStream_t st);
void launch_splitk_reduce(const GemmGroup *groups, const int *gidx, int n, int64_t maxmn, const SolveDesc *dd,
                          const int *gsolve, hipStream_t st);
void launch_t3_tables(const SolveDesc &d, const TriplesDesc &td, double volume, double2 *hppp, double2 *hhhp,
                      hipStream_t st);
void launch_triples(const SolveDesc &d, const TriplesDesc &td, double volume, double2 *partial,
                    unsigned long long *count, int nblocks, double2 *out, unsigned long long *count_out,
                    hipStream_t st, int mode, const double2 *tdt, int maxnhh);
void launch_t3_td(const SolveDesc &d, const TriplesDesc &td, double2 *tdt, hipStream_t st);
// mlp.hip
void launch_mlp_forward(const float *params, const uint16_t *feat, int64_t n, float *y, hipStream_t st);
void launch_mlp_loss_grad(const float *params, const uint16_t *feat, const float *target, const int64_t *idx,
                          int64_t n, int64_t npad, const MlpTrainBufs &tb, float *part, int max_split, float *grad,
                          double *loss_sum, hipStream_t st);
void launch_mlp_adam(float *th, const float *g, float *m, float *v, float lr, int step, hipStream_t st);
void launch_mlp_sample(int64_t *idx, int64_t B, int64_t ntot, uint64_t seed, hipStream_t st);
void launch_mlp_features(const SolveDesc &d, float kf, float ef, const float tw[3], float scale, uint16_t *feat,
                         float *target, hipStream_t st);
void launch_mlp_apply(const SolveDesc &d, const float *y, float scale, float ef, hipStream_t st);
void launch_mlp_infer(const float *params, const SolveDesc &d, float kf, float ef, const float tw[3], float scale,
                      hipStream_t st);
void launch_read_probe(const double2 *a, int64_t n, double *out, int nblocks, hipStream_t st);
}  // namespace ccgpu

using namespace ccgpu;

namespace {
thread_local std::string g_err;
}  // namespace

namespace ccgpu {
void set_last_error(const std::string &msg) { g_err = msg; }
}  // namespace ccgpu

namespace {

struct Error : std::runtime_error {
  explicit Error(const std::string &m) : std::runtime_error(m) {}
};

std::string fmt(const char *f, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, f);
  vsnprintf(buf, sizeof buf, f, ap);
  va_end(ap);
  return buf;
}

#define HIPCHK(x)                                                                         \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess) throw Error(fmt("%s failed: %s", #x, hipGetErrorString(e_))); \
  } while (0)

template <typename T>
T *dalloc(size_t n, std::vector<void *> &owned) {
  void *p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc(&p, n * sizeof(T));
  if (e != hipSuccess)
    throw Error(fmt("hipMalloc of %zu bytes failed: %s", n * sizeof(T), hipGetErrorString(e)));
  owned.push_back(p);
  return (T *)p;
}

template <typename T>
T *dupload(const std::vector<T> &v, std::vector<void *> &owned) {
  T *p = dalloc<T>(v.size(), owned);
  if (!v.empty()) HIPCHK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return p;
}

// Tiles reordered so that each channel's tiles go to one XCD (workgroup i runs
// on XCD i mod 8, which keeps the channel's B operand in that XCD's L2) and the
// XCDs get equal tile counts.
void xcd_order_list(std::vector<GemmTile> &tiles, size_t ngroups) {
  const int nx = 8;
  std::vector<std::vector<GemmTile>> bins(nx);
  std::vector<double> load(nx, 0.0);
  std::vector<std::vector<GemmTile>> per(ngroups);
  std::vector<double> cost(ngroups, 0.0);
  for (auto &t : tiles) {
    per[t.g].push_back(t);
    cost[t.g] += (double)(t.ke - t.kb);   // a tile's time goes with its k range
  }
  // groups largest first onto the least-loaded XCD: the longest tiles start
  // first and the tail of the launch is made of short ones
  std::vector<int> order(ngroups);
  for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
  for (int g : order) {
    if (per[g].empty()) continue;
    int best = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    for (auto &t : per[g]) bins[best].push_back(t);
    load[best] += cost[g];
  }
  std::vector<GemmTile> out;
  size_t maxlen = 0;
  for (auto &b : bins) maxlen = std::max(maxlen, b.size());
  for (size_t r = 0; r < maxlen; ++r)
    for (int x = 0; x < nx; ++x)
      if (r < bins[x].size()) out.push_back(bins[x][r]);
  tiles.swap(out);
}

// Work list of one GEMM family (one group per channel and solve).  Groups are
// classed by their column count N (<= 16, 32, 48, 64, 80; wider -> class 3),
// and every class is one launch with its own tile shape: MB 16-row blocks per
// wave (4 waves stacked along M, so 64*MB rows) by NB 16-column blocks; a group
// wider than 16*NB gets several column tiles.  Narrow amplitude blocks thus do
// not pay for padded MFMA columns.  The shapes come from a default table, the
// autotuner, or CCGPU_SHAPES for studies.  tiles64 is a uniform 64x64 tiling
// for the VALU cross-check kernel.
struct GemmShape {
  int mb, nb;
  int lb = 0;   // 1: B operand staged in LDS; 2: A and B in LDS (mb 1); 3: pp ladder with V and T
                // streamed through LDS by direct-to-LDS loads (mb 1, nb 1-2); 4: the same with eight
                // waves and one workgroup per CU (mb 1, nb 2); same tiles, same results
};
bool gemm_shape_ok(int mb, int nb) {
  return (mb == 0 && nb == 4) || (mb == 1 && nb >= 1 && nb <= 5) || (mb == 2 && (nb == 1 || nb == 2)) ||
         (mb == 4 && nb == 1);
}
// Shapes of the BF16x3 kernel (its accumulators are FP32 and its operands
// are staged in LDS already split, so it affords wider tiles): 64 rows by
// 16..128 columns, or 128 rows by 16..64 columns; (0,4) is the thin kernel.
bool bf3_shape_ok(int mb, int nb) {
  return (mb == 0 && nb == 4) || (mb == 1 && nb >= 1 && nb <= 8 && nb != 7) || (mb == 2 && nb >= 1 && nb <= 4);
}
// Forms a family's tile class may use.  Class 5 (N <= 4) has one form per
// family, so its summation order never depends on the tuner: the thin VALU
// kernel, or for the pp ladder the LDS-pipelined MFMA kernel.  The other
// classes take any MFMA form; the LDS pipeline (lb 3) serves the pp ladder only.
// List ids 10 + fam are the BF16x3 lists of a family (lb 0 only).
bool form_ok(int fam, int c, GemmShape s) {
  if (fam >= 10) return c == 5 ? (s.mb == 0 && s.nb == 4 && s.lb == 0) : (s.mb > 0 && s.lb == 0 && bf3_shape_ok(s.mb, s.nb));
  const bool ppl = s.mb == 1 && ((s.nb <= 2 && s.lb == 3) || (s.nb == 2 && s.lb == 4));
  if (c == 5) return fam == 1 ? ppl : (s.mb == 0 && s.nb == 4 && s.lb == 0);
  if (s.mb <= 0 || !gemm_shape_ok(s.mb, s.nb) || s.lb < 0 || s.lb > 4) return false;
  if (s.lb == 2) return s.mb == 1 && s.nb <= 3;
  if (s.lb >= 3) return fam == 1 && ppl;
  return true;
}
// rows and columns of one tile of a shape; (0,4) is the thin VALU kernel
int tile_rows(GemmShape s) { return s.mb == 0 ? 16 : 64 * s.mb; }
int tile_cols(GemmShape s) { return s.mb == 0 ? 4 : 16 * s.nb; }

struct GemmList {
  static constexpr int NCLS = 6;
  int fam = 0;   // 0 X = V_hhpp T, 1 pp ladder, 2 hh ladder, 3 ring I4, 4 ring R~, 5 shared pp ladder
  bool bf3 = false;   // the list of the BF16x3 kernel: its own classes, shapes and tuning
  int id() const { return bf3 ? 10 + fam : fam; }
  std::vector<GemmGroup> groups;
  std::vector<int> gsolve;
  GemmShape shape[NCLS] = {{2, 1}, {2, 2}, {1, 3}, {1, 4}, {1, 5}, {0, 4}};
  std::vector<GemmTile> tiles[NCLS];
  std::vector<GemmTile> tiles64;
  std::vector<int> split;        // groups reduced by k_splitk_reduce
  std::vector<int64_t> poff;     // partial-buffer offset per group
  int64_t psize = 0, maxmn = 0;
  bool xcd = false;              // order tiles channel-per-XCD
  bool ready = false;            // shapes chosen, tiles built and uploaded
  GemmGroup *d_groups = nullptr;
  GemmTile *d_tiles[NCLS] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  GemmTile *d_tiles64 = nullptr;
  int *d_gsolve = nullptr, *d_split = nullptr;
  double2 *d_part = nullptr;
  double flops = 0;       // algorithmic real flops per launch
  double bytes = 0;       // algorithmic bytes of the A operand per launch
  // class 5: N <= 4; 0: N <= 16; 1: <= 32; 2: <= 48; 3: <= 64 or > 80; 4: <= 80.
  // BF16x3 lists: 5: N <= 4; 0: <= 16; 1: <= 32; 2: <= 64; 4: <= 96; 3: wider
  static int cls_of(int N) {
    return N <= 4 ? 5 : N <= 16 ? 0 : N <= 32 ? 1 : N <= 48 ? 2 : N <= 64 ? 3 : N <= 80 ? 4 : 3;
  }
  int cls(int N) const {
    if (!bf3) return cls_of(N);
    return N <= 4 ? 5 : N <= 16 ? 0 : N <= 32 ? 1 : N <= 64 ? 2 : N <= 96 ? 4 : 3;
  }
  void add(GemmGroup g, int solve) {
    if (g.M <= 0 || g.N <= 0 || g.K <= 0) return;
    g.nsplit = 1;
    g.P = nullptr;
    groups.push_back(g);
    gsolve.push_back(solve);
    flops += 8.0 * g.M * (double)g.N * g.K;
  }
  int ntiles() const {
    size_t n = 0;
    for (int c = 0; c < NCLS; ++c) n += tiles[c].size();
    return (int)n;
  }
  bool has_class(int c) const {
    for (const GemmGroup &g : groups)
      if (cls(g.N) == c) return true;
    return false;
  }
  void apply_env_shapes() {
    // CCGPU_SHAPES="fam.class=mb.nb[.lb],..." (studies and the tests of the
    // kernel forms; lb 0 = B from L2, 1 = B in LDS, 2 = A and B in LDS)
    const char *e = getenv("CCGPU_SHAPES");
    if (!e) return;
    std::string spec(e);
    size_t pos = 0;
    while (pos < spec.size()) {
      size_t end = spec.find(',', pos);
      if (end == std::string::npos) end = spec.size();
      int f, c, mb, nb, lb = 0;
      const int got = sscanf(spec.substr(pos, end - pos).c_str(), "%d.%d=%d.%d.%d", &f, &c, &mb, &nb, &lb);
      if (got >= 4 && f == id() && c >= 0 && c < NCLS && form_ok(id(), c, GemmShape{mb, nb, got >= 5 ? lb : 0}))
        shape[c] = GemmShape{mb, nb, got >= 5 ? lb : 0};
      pos = end + 1;
    }
  }
  // Tiles of every class for the current shapes.  A group with few output
  // tiles and a long k range splits k over several workgroups; the partial
  // products are summed in fixed order by k_splitk_reduce.  The split depends
  // on the group alone (not on the tile shape or on the other groups of a
  // batch), and the MFMA kernels sum k in the same order for every shape, so
  // results do not depend on the autotuner's choice or on batching.
  bool nosplit = false;           // every group over its whole k range (the pp ladder)
  bool merged = false;            // one launch per distinct tile shape instead of per class
  static int split_of(const GemmGroup &g) {
    // CCGPU_SPLIT_TILES and CCGPU_SPLIT_K are read per list build (tests force
    // splits: a tile limit and a k quantum other than 128 and 256)
    const char *e = getenv("CCGPU_SPLIT_TILES");
    const int64_t lim = e && atoi(e) > 0 ? atoi(e) : 128;
    const char *q = getenv("CCGPU_SPLIT_K");
    const int64_t quantum = q && atoi(q) >= 16 ? atoi(q) : 256;
    const int64_t ref_tiles = (int64_t)((g.M + 63) / 64) * ((g.N + 15) / 16);
    if (g.K < 512 || ref_tiles >= lim) return 1;
    return (int)std::min<int64_t>({(2 * lim + ref_tiles - 1) / ref_tiles, (int64_t)(g.K / quantum), 8});
  }
  void build_tiles() {
    for (int c = 0; c < NCLS; ++c) tiles[c].clear();
    tiles64.clear();
    split.clear();
    poff.clear();
    psize = 0;
    maxmn = 0;
    for (size_t gi = 0; gi < groups.size(); ++gi) {
      GemmGroup &g = groups[gi];
      const int c = cls(g.N);
      const int nsp = nosplit ? 1 : split_of(g);
      int ksz = (g.K + nsp - 1) / nsp;
      ksz = (ksz + 15) / 16 * 16;
      g.nsplit = (g.K + ksz - 1) / ksz;
      g.P = nullptr;
      poff.push_back(g.nsplit > 1 ? psize : -1);
      if (g.nsplit > 1) {
        psize += (int64_t)g.nsplit * g.M * g.N;
        split.push_back((int)gi);
        maxmn = std::max<int64_t>(maxmn, (int64_t)g.M * g.N);
      }
      const int bm = tile_rows(shape[c]), bn = tile_cols(shape[c]);
      for (int ks = 0; ks < g.nsplit; ++ks)
        for (int m = 0; m < g.M; m += bm)
          for (int n = 0; n < g.N; n += bn)
            tiles[c].push_back(GemmTile{(int)gi, m, n, ks, ks * ksz, std::min(g.K, (ks + 1) * ksz), 0, 0});
      for (int m = 0; m < g.M; m += 64)
        for (int n = 0; n < g.N; n += 64) tiles64.push_back(GemmTile{(int)gi, m, n, 0, 0, g.K, 0, 0});
    }
    if (merged) {   // classes of one shape share one launch (that of the first class of the shape)
      for (int c = 1; c < NCLS; ++c)
        for (int c2 = 0; c2 < c; ++c2)
          if (shape[c2].mb == shape[c].mb && shape[c2].nb == shape[c].nb && shape[c2].lb == shape[c].lb) {
            tiles[c2].insert(tiles[c2].end(), tiles[c].begin(), tiles[c].end());
            tiles[c].clear();
            break;
          }
    }
    if (xcd)
      for (int c = 0; c < NCLS; ++c) xcd_order_list(tiles[c], groups.size());
  }
  // partial buffer for the split groups, then upload everything
  void upload(std::vector<void *> &owned) {
    d_part = psize > 0 ? dalloc<double2>(psize, owned) : nullptr;
    for (size_t g = 0; g < groups.size(); ++g) groups[g].P = poff[g] >= 0 ? d_part + poff[g] : nullptr;
    d_groups = dupload(groups, owned);
    for (int c = 0; c < NCLS; ++c) d_tiles[c] = dupload(tiles[c], owned);
    d_tiles64 = dupload(tiles64, owned);
    d_gsolve = dupload(gsolve, owned);
    d_split = dupload(split, owned);
  }
  void prepare(std::vector<void *> &owned) {
    apply_env_shapes();
    build_tiles();
    upload(owned);
  }
};

GemmGroup mkgroup(const void *A, const double2 *B, double2 *C, const double2 *C0, int M, int N, int K, int lda,
                  int ldb, int ldc, int ldc0, double alpha, double beta) {
  GemmGroup g{};
  g.A = A; g.B = B; g.C = C; g.C0 = C0; g.P = nullptr;
  g.M = M; g.N = N; g.K = K; g.lda = lda; g.ldb = ldb; g.ldc = ldc; g.ldc0 = ldc0;
  g.nsplit = 1; g.alpha = alpha; g.beta = beta;
  return g;
}

// XCD-aware reorder: tiles of one group stay on one of the 8 XCDs (blocks are
// dealt round-robin to XCDs), so the shared operand of a channel is read from
// one L2.  Groups are packed onto XCD bins by largest-first.

struct Solve {
  Plan plan;
  std::vector<void *> owned;
  SolveDesc desc{};
  SolveDesc *d_desc = nullptr;
  hipStream_t stream = nullptr;
  bool have_v = false, have_fock = false, derived = false, have_t = false;
  bool native_v = false;   // V from the built-in generator (needed by (T))
  uint64_t vgen = 0;       // counts changes of V (a shared pp ladder refills its union blocks)
  int vkind = -1;          // CCGPU_INT_* of the generated V
  ChiralModel chiral = chiral_model(3, 450.0, nullptr);   // constants of CCGPU_INT_CHIRAL
  ChiralDev chiral_dev{};  // its device tables (set by the generator)
  // three-nucleon force of CCGPU_INT_CHIRAL: -1 off, else the normal-ordering
  // level (0 E0, 1 and Fock, 2 and 3 and the two-body blocks; 3 also puts the
  // three-body element into (T)) and its constants; tnf_set once the
  // constants were given (they also define the residual three-body element
  // W3 of CCDT-1 when the force is not normal-ordered)
  int tnf = -1;
  bool tnf_set = false;
  double tnf_ce = 0.017, tnf_cd = 0.79, tnf_cutoff = 450.0;
  int tnf_nexp = 3;
  Chiral3Dev chiral3{};
  // FP32 copy of V_pppp (Vpppp32): block c at off_vpp32[c], rows padded to
  // even length; device copies of the offset tables
  std::vector<int64_t> off_vpp32;
  int64_t nVpp32 = 0;
  int64_t *d_off_vpp = nullptr, *d_off_vpp32 = nullptr;
  int *d_npp = nullptr;
  bool have_t3 = false;
  bool have_w3 = false;          // t3.c3w holds the W3 constants
  bool have_ccdt1 = false;       // setup_t3 done
  double2 ccdt1_e3{0.0, 0.0};    // energy of the last t3_intermediate
  double2 t3_et4{0.0, 0.0}, t3_et5{0.0, 0.0};   // parts of the last (T)
  TriplesDesc t3{};
  double2 *t3_partial = nullptr, *t3_out = nullptr;
  double2 *t3_td = nullptr;   // t(x d, u v) of the F terms, refilled from T for every (T) sum
  unsigned long long *t3_count = nullptr;   // per block, then the total at [8192]
  long long t3_elements = 0;
  double t3_seconds = 0.0;
  double2 e0{0, 0};
  // hhphp records of the drop-in path: key (i,b,j,a) -> <ib||ja>
  std::unordered_map<uint64_t, std::complex<double>> hphp;
  // <cb||kd> and <jk||lc> records of the drop-in path for (T): key (p,q,r,s)
  std::unordered_map<uint64_t, std::complex<double>> t3v;
  // correlation energies of the last solve: the start, then one per iteration
  std::vector<std::complex<double>> ehist;
  // DIIS state
  int depth = 0;
  int depth_f = 0;         // columns of desc.histf (Anderson residual vectors)
  std::vector<int> ring;  // logical column -> physical slot
  int nstep = 0, count = 0;
  int *d_slot = nullptr;   // [1] physical slot for k_update
  int *d_slots = nullptr;  // [16] physical slots in logical order
  double2 *d_coef = nullptr;
  // work lists for the five GEMM families
  GemmList g_x, g_pp, g_hh, g_i4, g_rt;
  bool vpppp32 = false;
  double2 *scal = nullptr;
  ~Solve() {
    for (void *p : owned) (void)hipFree(p);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

std::mutex g_mu;
std::map<int, std::unique_ptr<Solve>> g_solves;
int g_next = 1;

Solve &get(int h) {
  auto it = g_solves.find(h);
  if (it == g_solves.end()) throw Error(fmt("invalid handle %d", h));
  return *it->second;
}

int gemm_variant_override() {
  const char *e = getenv("CCGPU_GEMM");
  if (e && std::string(e) == "valu") return 3;
  return -1;
}

// ----- timing of kernel families -----
struct Stat {
  double ms = 0;
  long long n = 0;
};
std::map<std::string, Stat> g_stats;
std::vector<std::string> g_stat_order;

struct Timer {
  bool on;
  hipStream_t st;
  hipEvent_t a{}, b{};
  std::string name;
  Timer(bool on_, hipStream_t s, const char *nm) : on(on_), st(s), name(nm) {
    if (on) {
      (void)hipEventCreate(&a);
      (void)hipEventCreate(&b);
      (void)hipEventRecord(a, st);
    }
  }
  ~Timer() {
    if (!on) return;
    (void)hipEventRecord(b, st);
    (void)hipEventSynchronize(b);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, a, b);
    if (!g_stats.count(name)) g_stat_order.push_back(name);
    g_stats[name].ms += ms;
    g_stats[name].n += 1;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
  }
};

void build_device(Solve &S) {
  Plan &P = S.plan;
  auto &o = S.owned;
  SolveDesc &d = S.desc;
  HIPCHK(hipStreamCreateWithFlags(&S.stream, hipStreamNonBlocking));
  d.A = P.A; d.nsp = P.nsp; d.np = P.np; d.nch = P.nch; d.nphc = P.nphc;
  d.nT = P.nT; d.nRt = P.nRt; d.nI4 = P.nI4; d.active = 1;
  d.kx = dupload(P.orb.kx, o); d.ky = dupload(P.orb.ky, o); d.kz = dupload(P.orb.kz, o);
  d.e = dupload(P.orb.e, o);
  d.nx = dupload(P.orb.nx, o); d.ny = dupload(P.orb.ny, o); d.nz = dupload(P.orb.nz, o);
  d.sz = dupload(P.orb.sz, o); d.tz = dupload(P.orb.tz, o);
  d.partner = dupload(P.partner, o);
  d.t_ch = dupload(P.t_ch, o);
  d.npp = dupload(P.npp, o); d.nhh = dupload(P.nhh, o);
  d.off_t = dupload(P.off_t, o); d.off_pp = dupload(P.off_pp, o); d.off_hh = dupload(P.off_hh, o);
  d.off_vpp = dupload(P.off_vpp, o); d.off_vhh = dupload(P.off_vhh, o);
  d.pp_a = dupload(P.pp_a, o); d.pp_b = dupload(P.pp_b, o);
  d.hh_i = dupload(P.hh_i, o); d.hh_j = dupload(P.hh_j, o);
  {
    std::vector<uint64_t> lab(P.nT);
    for (int c = 0; c < P.nch; ++c)
      for (int r = 0; r < P.npp[c]; ++r)
        for (int q = 0; q < P.nhh[c]; ++q) {
          const int64_t pr = P.off_pp[c] + r, hs = P.off_hh[c] + q;
          lab[P.off_t[c] + (int64_t)r * P.nhh[c] + q] = (uint64_t)P.pp_a[pr] | ((uint64_t)P.pp_b[pr] << 16) |
                                                        ((uint64_t)P.hh_i[hs] << 32) | ((uint64_t)P.hh_j[hs] << 48);
        }
    d.lab = dupload(lab, o);
  }
  d.ob_rowB = dupload(P.ob_rowB, o); d.ob_rowA = dupload(P.ob_rowA, o);
  d.ob_colJ = dupload(P.ob_colJ, o); d.ob_colI = dupload(P.ob_colI, o);
  d.nrows = (int64_t)P.pp_a.size();
  d.row_ch = dupload(P.row_ch, o);
  d.i6_dptr = dupload(P.i6_dptr, o); d.i6_didx = dupload(P.i6_didx, o);
  d.i6_optr = dupload(P.i6_optr, o); d.i6_oidx = dupload(P.i6_oidx, o);
  d.p6 = dalloc<double2>(3 * P.pp_a.size(), o);
  std::vector<DPair> pr(P.ppref.size()), hr(P.hhref.size());
  for (size_t x = 0; x < pr.size(); ++x)
    pr[x] = DPair{P.ppref[x].sign ? P.ppref[x].ch : -1, P.ppref[x].idx * 2 + (P.ppref[x].sign < 0)};
  for (size_t x = 0; x < hr.size(); ++x)
    hr[x] = DPair{P.hhref[x].sign ? P.hhref[x].ch : -1, P.hhref[x].idx * 2 + (P.hhref[x].sign < 0)};
  d.ppref = dupload(pr, o);
  d.hhref = dupload(hr, o);
  d.nph = dupload(P.nph, o); d.nhp = dupload(P.nhp, o);
  d.off_ph = dupload(P.off_ph, o); d.off_hp = dupload(P.off_hp, o);
  d.off_rt = dupload(P.off_rt, o); d.off_i4 = dupload(P.off_i4, o); d.off_vt = dupload(P.off_vt, o);
  d.ph_a = dupload(P.ph_a, o); d.ph_i = dupload(P.ph_i, o);
  d.hp_j = dupload(P.hp_j, o); d.hp_b = dupload(P.hp_b, o);
  {
    if (P.nRt >= ((int64_t)1 << 31)) throw Error("ring blocks above 2^31 elements");
    std::vector<int> r4(P.ring4.begin(), P.ring4.end());
    d.ring4 = dupload(r4, o);
  }
  d.T = dalloc<double2>(P.nT, o);
  d.R = dalloc<double2>(P.nT, o);
  d.R0 = dalloc<double2>(P.nT, o);
  d.Rh = dalloc<double2>(P.nT, o);
  d.Vhhpp = dalloc<double2>(P.nT, o);
  d.Vhhhh = dalloc<double2>(P.nVhh, o);
  d.X = dalloc<double2>(2 * P.nVhh, o);
  d.Vpppp = dalloc<double2>(P.nVpp, o);
  d.Vpppp32 = nullptr;
  d.Tt = dalloc<double2>(P.nRt, o);
  d.Rt = dalloc<double2>(P.nRt, o);
  d.Vt = dalloc<double2>(P.nRt, o);
  d.I4 = dalloc<double2>(P.nI4, o);
  d.W0 = dalloc<double2>(P.nI4, o);
  d.fdiag = dalloc<double2>(P.nsp, o);
  d.fpart = dalloc<double2>(P.nsp, o);
  d.dh = dalloc<double2>(P.A, o);
  d.I7off = dalloc<double2>(P.A, o);
  d.dp = dalloc<double2>(P.np, o);
  d.I6off = dalloc<double2>(P.np, o);
  d.e0 = dalloc<double2>(1, o);
  d.red = dalloc<double2>(1024 * 36, o);
  d.scal = dalloc<double2>(64, o);
  d.hist = nullptr;
  S.scal = d.scal;
  HIPCHK(hipMemset(d.Tt, 0, P.nRt * sizeof(double2)));  // entries with a=b or i=j stay zero
  HIPCHK(hipMemset(d.Rh, 0, P.nT * sizeof(double2)));
  HIPCHK(hipMemset(d.T, 0, P.nT * sizeof(double2)));
  HIPCHK(hipMemset(d.Vpppp, 0, P.nVpp * sizeof(double2)));
  HIPCHK(hipMemset(d.Vhhpp, 0, P.nT * sizeof(double2)));
  HIPCHK(hipMemset(d.Vhhhh, 0, P.nVhh * sizeof(double2)));
  HIPCHK(hipMemset(d.W0, 0, P.nI4 * sizeof(double2)));
  S.d_desc = dalloc<SolveDesc>(1, o);
  S.d_slot = dalloc<int>(1, o);
  S.d_slots = dalloc<int>(16, o);
  S.d_coef = dalloc<double2>(16, o);
  HIPCHK(hipMemcpy(S.d_desc, &d, sizeof(SolveDesc), hipMemcpyHostToDevice));

  // GEMM work lists
  for (int c = 0; c < P.nch; ++c) {
    int np_ = P.npp[c], nh = P.nhh[c];
    // X = Vhhpp T
    S.g_x.add(mkgroup(d.Vhhpp + P.off_t[c], d.T + P.off_t[c], d.X + P.off_vhh[c], nullptr, nh, nh, np_, np_, nh, nh, 0, 1.0, 0.0), 0);
    // R += Vpppp T   (pp ladder, ccm_t2_eqn.f90:574-607)
    S.g_pp.add(mkgroup(d.Vpppp + P.off_vpp[c], d.T + P.off_t[c], d.R + P.off_t[c], d.R + P.off_t[c], np_, nh, np_, np_, nh, nh, nh, 1.0, 1.0), 0);
    // R += T I9      (hh ladder, ccm_t2_eqn.f90:614-647)
    // the hh ladder into Rh (beta 0), beside the pp ladder; the ring scatter adds it to R
    S.g_hh.add(mkgroup(d.T + P.off_t[c], d.X + P.nVhh + P.off_vhh[c], d.Rh + P.off_t[c], nullptr, np_, nh, nh, nh, nh, nh, nh, 1.0, 0.0), 0);
  }
  for (int K = 0; K < P.nphc; ++K) {
    int nph = P.nph[K], nhp = P.nhp[K];
    // I4 = W0 + 1/2 Vt Tt (ccm_t2_eqn.f90:657-780)
    S.g_i4.add(mkgroup(d.Vt + P.off_vt[K], d.Tt + P.off_rt[K], d.I4 + P.off_i4[K], d.W0 + P.off_i4[K], nhp, nhp, nph, nph, nhp, nhp, nhp, 0.5, 1.0), 0);
    // Rt = Tt I4 (ccm_t2_eqn.f90:995-1026)
    S.g_rt.add(mkgroup(d.Tt + P.off_rt[K], d.I4 + P.off_i4[K], d.Rt + P.off_rt[K], nullptr, nph, nhp, nhp, nhp, nhp, nhp, 0, 1.0, 0.0), 0);
  }
  // tiles are made per batch (make_batch), for the batch's own class sizes
  S.g_pp.bytes = 16.0 * (double)P.nVpp;
}

void alloc_history(Solve &S, int depth, bool anderson) {
  if (depth < 2 || depth > 9) throw Error(fmt("DIIS subspace %d outside [2, 9]", depth));
  if (anderson && depth > 8) throw Error(fmt("Anderson subspace %d outside [2, 8]", depth));
  bool changed = false;
  if (S.depth < depth) {
    // history has subspace columns (diis_setup, ccm_diis.f90:284-288)
    S.desc.hist = dalloc<double2>((size_t)depth * S.plan.nT, S.owned);
    S.depth = depth;
    S.depth_f = 0;   // the FP32 residual vectors pair with the FP64 columns
    changed = true;
  }
  if (anderson && S.depth_f < S.depth) {
    S.desc.histf = dalloc<float2>((size_t)S.depth * S.plan.nT, S.owned);
    S.depth_f = S.depth;
    changed = true;
  }
  if (changed) HIPCHK(hipMemcpy(S.d_desc, &S.desc, sizeof(SolveDesc), hipMemcpyHostToDevice));
}

void ensure_ready(Solve &S) {
  if (!S.have_v) throw Error("no interaction: call ccgpu_generate_interaction or ccgpu_set_block first");
  if (!S.derived) throw Error("setup not finalized: call ccgpu_finalize_setup");
}

// ---------- host complex linear solve (Gaussian elimination, partial pivoting
// by |re|+|im| as the reference's ZGESV does) ----------
std::vector<std::complex<double>> zsolve(std::vector<std::complex<double>> a, std::vector<std::complex<double>> b,
                                         int n, int *info) {
  *info = 0;
  for (int k = 0; k < n; ++k) {
    int piv = k;
    double best = -1;
    for (int i = k; i < n; ++i) {
      double v = std::fabs(a[i * n + k].real()) + std::fabs(a[i * n + k].imag());
      if (v > best) { best = v; piv = i; }
    }
    if (best == 0.0) { *info = k + 1; return b; }
    if (piv != k) {
      for (int j = 0; j < n; ++j) std::swap(a[k * n + j], a[piv * n + j]);
      std::swap(b[k], b[piv]);
    }
    for (int i = k + 1; i < n; ++i) {
      std::complex<double> f = a[i * n + k] / a[k * n + k];
      a[i * n + k] = f;
      for (int j = k + 1; j < n; ++j) a[i * n + j] -= f * a[k * n + j];
      b[i] -= f * b[k];
    }
  }
  for (int i = n - 1; i >= 0; --i) {
    std::complex<double> s = b[i];
    for (int j = i + 1; j < n; ++j) s -= a[i * n + j] * b[j];
    b[i] = s / a[i * n + i];
  }
  return b;
}

// ---------- one residual for a set of solves sharing work lists ----------
struct Batch {
  std::vector<Solve *> solves;
  SolveDesc *d_descs = nullptr;  // device array
  std::vector<SolveDesc> h_descs;
  GemmList g_x, g_pp, g_hh, g_i4, g_rt, g_pp32;
  GemmList g_xb, g_hhb, g_i4b, g_rtb;   // the same groups for the BF16x3 kernel
  GemmList g_i4f, g_rtf;   // BF16x3 ring GEMMs on the FP32 ring blocks (when every solve has them)
  bool ring32 = false;
  GemmList g_ppbf;         // BF16x3 pp ladder reading the split copy of T (when every solve has one)
  bool ppbf = false;
  // Shared pp ladder of a batch of twists of one lattice and interaction: the
  // NN element <ab||cd> depends on the lattice labels of a, b, c, d only (the
  // twist cancels in every momentum difference), so one V_pppp over the union
  // of the twists' particle pairs serves all of them.  Per (K, Tz) channel one
  // GEMM R_u = V_u T_u whose columns are the hole pairs of every twist side
  // by side; T_u holds each twist's amplitudes at its own pairs (zero where a
  // twist owns no pair), R_u goes back into each twist's R.  V_pppp is read
  // once per iteration instead of once per twist, and narrow per-twist
  // GEMMs become one wide GEMM per channel.
  struct Shared {
    bool on = false;
    std::vector<uint64_t> vgen;   // V generations the union blocks were filled from
    int nu = 0, max_npp = 0;
    std::vector<int> npp, ncol, pa, pb;   // union pairs in solve 0's orbital numbering
    std::vector<int64_t> off_pp, off_v, off_t;
    int64_t nV = 0, nTu = 0;
    double2 *V = nullptr, *Tu = nullptr, *Ru = nullptr;
    float2 *V32 = nullptr;
    int *d_npp = nullptr, *d_pa = nullptr, *d_pb = nullptr;
    int64_t *d_off_pp = nullptr, *d_off_v = nullptr;
    uint32_t **d_maps = nullptr;   // per solve: unique amplitude -> 2 * position + (sign < 0)
    GemmList g64, gbf;             // FP64 GEMMs on V_u, BF16x3 GEMMs on its FP32 copy
  } share;
  bool has_f32 = true;   // every solve has the FP32 copy of its pppp blocks
  int64_t maxT = 0, maxVhh = 0, maxRows = 0;
  int maxA = 0, maxNp = 0;
  int *d_slot = nullptr, *d_slots = nullptr;
  double2 *d_coef = nullptr;
  double2 *d_scal = nullptr;  // nsolve * 64
  std::vector<int> h_slot, h_slots;       // host staging of control arrays
  std::vector<double2> h_coef, h_scal;
  std::vector<int> status;                // CCGPU_STATUS_* per solve of the current run
  bool forced_done = false;               // CCGPU_FORCE_SINGULAR fired in this run
  // Anderson update: per solve the held vectors (physical history slots,
  // oldest first), their Gram matrix Re<f_j, f_k> and the energies of the
  // mixed iterates y_j; device copies of the slots and coefficients
  struct AA {
    std::vector<int> slot;
    std::vector<double> gram;   // kAAHeld x kAAHeld, logical order
    std::vector<std::complex<double>> ey;
  };
  std::vector<AA> aa;
  int *d_aaslots = nullptr;
  double *d_aacoef = nullptr;
  std::vector<int> h_aaslots;
  std::vector<double> h_aacoef;
  std::vector<void *> owned;
  hipStream_t stream = nullptr;
  // second stream (default priority) for the ring branch of the residual, and
  // the fork/join events
  hipStream_t stream2 = nullptr, stream3 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_join3 = nullptr, ev_x = nullptr, ev_hh = nullptr;
  hipGraph_t graph[4] = {nullptr, nullptr, nullptr, nullptr};        // per precision mode
  hipGraphExec_t gexec[4] = {nullptr, nullptr, nullptr, nullptr};
  // pinned staging of the small copies of every iteration, one buffer per
  // call site; each site's copy is followed by a stream synchronisation
  // before the host writes its buffer again
  void *pin[8] = {};
  size_t pin_bytes[8] = {};
  void *pinned(int site, size_t bytes) {
    if (pin_bytes[site] < bytes) {
      if (pin[site]) (void)hipHostFree(pin[site]);
      pin[site] = nullptr;
      pin_bytes[site] = 0;
      HIPCHK(hipHostMalloc(&pin[site], bytes, hipHostMallocDefault));
      pin_bytes[site] = bytes;
    }
    return pin[site];
  }
  ~Batch() {
    for (void *p : pin)
      if (p) (void)hipHostFree(p);
    for (int q = 0; q < 4; ++q) {
      if (gexec[q]) (void)hipGraphExecDestroy(gexec[q]);
      if (graph[q]) (void)hipGraphDestroy(graph[q]);
    }
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (ev_join3) (void)hipEventDestroy(ev_join3);
    if (ev_x) (void)hipEventDestroy(ev_x);
    if (ev_hh) (void)hipEventDestroy(ev_hh);
    if (stream2) (void)hipStreamDestroy(stream2);
    if (stream3) (void)hipStreamDestroy(stream3);
    for (void *p : owned) (void)hipFree(p);
  }
};

void merge(GemmList &dst, const GemmList &src, int solve) {
  for (size_t i = 0; i < src.groups.size(); ++i) {
    dst.groups.push_back(src.groups[i]);
    dst.gsolve.push_back(solve);
  }
  dst.flops += src.flops;
  dst.bytes += src.bytes;
}

// ---------------------------------------------------------------------------
// Tile-shape autotuner.  For every work list and tile class the candidate
// shapes (MB, NB) are timed with HIP events on the list's own groups (the
// launches overwrite only outputs that every residual recomputes), and the
// fastest is kept.  Results are cached by a signature of the list (family,
// kernel kind, group dimensions), so repeated batches of the same lattices
// reuse them.  CCGPU_AUTOTUNE=0 keeps the default table.
// ---------------------------------------------------------------------------
std::map<std::string, std::array<GemmShape, GemmList::NCLS>> g_tuned;
std::string g_tune_report;

std::string list_signature(const GemmList &L, int kind) {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](uint64_t v) { h = (h ^ v) * 1099511628211ull; };
  for (const GemmGroup &g : L.groups) {
    mix((uint64_t)g.M);
    mix((uint64_t)g.N);
    mix((uint64_t)g.K);
  }
  return fmt("%d.%d.%zu.%016llx", L.id(), kind, L.groups.size(), (unsigned long long)h);
}

std::vector<GemmShape> shape_candidates(int fam, int c) {
  std::vector<GemmShape> v;
  if (fam >= 10) {   // BF16x3 lists: up to the widest column count of the class
    if (c == 5) return {{0, 4}};
    const int w = c == 0 ? 1 : c == 1 ? 2 : c == 2 ? 4 : c == 4 ? 6 : 8;
    for (int nb = 1; nb <= w; ++nb)
      if (bf3_shape_ok(1, nb)) v.push_back({1, nb});
    for (int nb = 1; nb <= std::min(4, w); ++nb) v.push_back({2, nb});
    return v;
  }
  if (c == 5) return fam == 1 ? std::vector<GemmShape>{{1, 1, 3}} : std::vector<GemmShape>{{0, 4}};
  if (fam == 1) {   // the LDS-pipelined forms first: the tuner keeps the first of equal times
    v.push_back({1, 1, 3});
    v.push_back({1, 2, 4});
    if (c >= 1) v.push_back({1, 2, 3});
  }
  for (int nb = 1; nb <= std::min(5, c + 1); ++nb) v.push_back({1, nb});
  v.push_back({2, 1});
  if (c >= 1) v.push_back({2, 2});
  v.push_back({4, 1});
  const size_t n = v.size();
  for (size_t i = 0; i < n; ++i) v.push_back({v[i].mb, v[i].nb, 1});   // LDS-staged B variants
  for (int nb = 1; nb <= std::min(3, c + 1); ++nb) v.push_back({1, nb, 2});   // A and B in LDS
  return v;
}

// CCGPU_TUNE_CACHE=path: the tuned shapes persist across processes in a text
// file, one line per list signature ("sig mb nb lb" for each of the NCLS
// classes), read once and appended after each new tuning; runs that share the
// file (profiling passes, production restarts) use identical tiles.
void load_tune_cache() {
  static bool loaded = false;
  if (loaded) return;
  loaded = true;
  const char *path = getenv("CCGPU_TUNE_CACHE");
  if (!path) return;
  FILE *f = fopen(path, "r");
  if (!f) return;
  char sig[128];
  while (fscanf(f, "%127s", sig) == 1) {
    std::array<GemmShape, GemmList::NCLS> arr;
    bool ok = true;
    for (int c = 0; c < GemmList::NCLS && ok; ++c) {
      int mb, nb, lb;
      ok = fscanf(f, "%d %d %d", &mb, &nb, &lb) == 3;
      arr[c] = GemmShape{mb, nb, lb};
      int fam = -1;
      ok = ok && sscanf(sig, "%d.", &fam) == 1 && form_ok(fam, c, arr[c]);
    }
    if (!ok) break;
    g_tuned[sig] = arr;
  }
  fclose(f);
}

void save_tune_cache(const std::string &sig, const std::array<GemmShape, GemmList::NCLS> &arr) {
  const char *path = getenv("CCGPU_TUNE_CACHE");
  if (!path) return;
  FILE *f = fopen(path, "a");
  if (!f) return;
  fprintf(f, "%s", sig.c_str());
  for (int c = 0; c < GemmList::NCLS; ++c) fprintf(f, " %d %d %d", arr[c].mb, arr[c].nb, arr[c].lb);
  fprintf(f, "\n");
  fclose(f);
}

bool autotune_enabled() {
  const char *e = getenv("CCGPU_AUTOTUNE");
  return !(e && atoi(e) == 0) && !getenv("CCGPU_SHAPES");
}

void autotune(GemmList &L, int kind, const SolveDesc *d_descs, hipStream_t st) {
  if (!autotune_enabled() || L.groups.empty() || L.merged) return;
  const std::string sig = list_signature(L, kind);
  load_tune_cache();
  auto it = g_tuned.find(sig);
  if (it != g_tuned.end()) {
    for (int c = 0; c < GemmList::NCLS; ++c) L.shape[c] = it->second[c];
    return;
  }
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  std::string rep = fmt("fam%d/kind%d:", L.id(), kind);
  for (int c = 0; c < GemmList::NCLS; ++c) {
    if (!L.has_class(c) || shape_candidates(L.id(), c).size() < 2) continue;
    float best = 1e30f;
    GemmShape best_shape = L.shape[c];
    for (const GemmShape cand : shape_candidates(L.id(), c)) {
      GemmList T = L;   // copy of the groups; tiles for this candidate
      T.shape[c] = cand;
      std::vector<void *> tmp;
      T.build_tiles();
      T.upload(tmp);
      float ms = 0.f;
      for (int rep_ = 0; rep_ < 4; ++rep_) {
        HIPCHK(hipEventRecord(e0, st));
        launch_gemm_shape(T.d_groups, T.d_tiles[c], (int)T.tiles[c].size(), d_descs, T.d_gsolve, kind, cand.mb,
                          cand.nb, cand.lb, st);
        HIPCHK(hipEventRecord(e1, st));
        HIPCHK(hipEventSynchronize(e1));
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, e0, e1));
        if (rep_ > 0) ms += t;   // the first launch warms caches and code
      }
      HIPCHK(hipGetLastError());
      for (void *p : tmp) (void)hipFree(p);
      if (ms < best) {
        best = ms;
        best_shape = cand;
      }
    }
    L.shape[c] = best_shape;
    rep += fmt(" c%d=%dx%d%s(%.1fus)", c, best_shape.mb, best_shape.nb,
               best_shape.lb == 4 ? "W" : best_shape.lb == 3 ? "P" : best_shape.lb == 2 ? "A" : best_shape.lb ? "L" : "",
               1e3f * best / 3.f);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  std::array<GemmShape, GemmList::NCLS> arr;
  for (int c = 0; c < GemmList::NCLS; ++c) arr[c] = L.shape[c];
  g_tuned[sig] = arr;
  save_tune_cache(sig, arr);
  if (!g_tune_report.empty()) g_tune_report += "; ";
  g_tune_report += rep;
}

bool share_enabled();
bool share_eligible(const std::vector<Solve *> &ss);
void build_shared(Batch &B);
void check_launch();

// Copies of the solve loop through the batch's pinned buffers (site: one per
// call site): host to device without a pageable staging copy; device to host
// completes with the stream
void h2d_pinned(Batch &B, int site, void *dst, const void *src, size_t bytes, hipStream_t st) {
  void *p = B.pinned(site, bytes);
  memcpy(p, src, bytes);
  HIPCHK(hipMemcpyAsync(dst, p, bytes, hipMemcpyHostToDevice, st));
}
void d2h_pinned_sync(Batch &B, int site, void *dst, const void *src, size_t bytes, hipStream_t st) {
  void *p = B.pinned(site, bytes);
  HIPCHK(hipMemcpyAsync(p, src, bytes, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  memcpy(dst, p, bytes);
}

std::unique_ptr<Batch> make_batch(const std::vector<Solve *> &ss, hipStream_t st) {
  auto B = std::make_unique<Batch>();
  B->solves = ss;
  B->stream = st;
  int n = (int)ss.size();
  B->d_scal = dalloc<double2>((size_t)n * 64, B->owned);
  for (int s = 0; s < n; ++s) {
    Solve &S = *ss[s];
    SolveDesc d = S.desc;
    d.active = 1;
    d.scal = B->d_scal + (size_t)s * 64;
    B->h_descs.push_back(d);
    merge(B->g_x, S.g_x, s);
    merge(B->g_pp, S.g_pp, s);
    merge(B->g_hh, S.g_hh, s);
    merge(B->g_i4, S.g_i4, s);
    merge(B->g_rt, S.g_rt, s);
    B->maxT = std::max(B->maxT, S.plan.nT);
    B->maxVhh = std::max(B->maxVhh, S.plan.nVhh);
    B->maxRows = std::max<int64_t>(B->maxRows, (int64_t)S.plan.pp_a.size());
    B->maxA = std::max(B->maxA, S.plan.A);
    B->maxNp = std::max(B->maxNp, S.plan.np);
  }
  // FP32 pp ladder list reads the float copy of Vpppp (its own padded rows)
  B->g_pp32 = B->g_pp;
  for (size_t g = 0; g < B->g_pp32.groups.size(); ++g) {
    int s = B->g_pp32.gsolve[g];
    Solve &S = *ss[s];
    if (!S.desc.Vpppp32) {
      B->has_f32 = false;
      continue;
    }
    const int64_t off = (const double2 *)B->g_pp32.groups[g].A - S.desc.Vpppp;
    const auto &ov = S.plan.off_vpp;
    const int c = (int)(std::upper_bound(ov.begin(), ov.end(), off) - ov.begin()) - 1;
    if (c < 0 || c >= S.plan.nch || ov[c] != off) throw Error("pp ladder group does not start a V_pppp block");
    B->g_pp32.groups[g].A = S.desc.Vpppp32 + S.off_vpp32[c];
    B->g_pp32.groups[g].lda = (S.plan.npp[c] + 1) & ~1;
  }
  // the BF16x3 pp ladder on the split copy of T: the same groups with B moved
  // to Tbf (16 bytes per amplitude as T, so the same offsets)
  B->ppbf = B->has_f32;
  for (Solve *S : ss) B->ppbf = B->ppbf && S->desc.Tbf;
  if (B->ppbf) {
    B->g_ppbf = B->g_pp32;
    for (size_t g = 0; g < B->g_ppbf.groups.size(); ++g) {
      const SolveDesc &d = ss[B->g_ppbf.gsolve[g]]->desc;
      const double2 *q = B->g_ppbf.groups[g].B;
      if (q < d.T || q >= d.T + ss[B->g_ppbf.gsolve[g]]->plan.nT) throw Error("pp ladder B outside T");
      B->g_ppbf.groups[g].B = (const double2 *)(d.Tbf + (q - d.T));
    }
  }
  // BF16x3 lists: the groups of the X, hh and ring families, tiled for the
  // BF16x3 kernel (up to 128 columns, so a ring channel's Vt rows are read
  // once rather than once per 32 columns) and ordered channel-per-XCD
  B->g_xb = B->g_x;
  B->g_hhb = B->g_hh;
  B->g_i4b = B->g_i4;
  B->g_rtb = B->g_rt;
  // the ring GEMMs of the BF16x3 residual on the FP32 ring blocks: the same
  // groups with every pointer moved to the FP32 copy (same offsets); half the
  // ring's bytes, no change to the products (BF16x3 splits FP32 values)
  B->ring32 = true;
  for (Solve *S : ss) B->ring32 = B->ring32 && S->desc.Tt32;
  if (B->ring32) {
    B->g_i4f = B->g_i4;
    B->g_rtf = B->g_rt;
    auto to32 = [&](GemmList &L) {
      for (size_t g = 0; g < L.groups.size(); ++g) {
        const SolveDesc &d = ss[L.gsolve[g]]->desc;
        GemmGroup &G = L.groups[g];
        auto mv = [&](const void *p) -> const void * {
          const double2 *q = (const double2 *)p;
          if (!q) return nullptr;
          if (q >= d.Tt && q < d.Tt + ss[L.gsolve[g]]->plan.nRt) return d.Tt32 + (q - d.Tt);
          if (q >= d.Vt && q < d.Vt + ss[L.gsolve[g]]->plan.nRt) return d.Vt32 + (q - d.Vt);
          if (q >= d.Rt && q < d.Rt + ss[L.gsolve[g]]->plan.nRt) return d.Rt32 + (q - d.Rt);
          if (q >= d.I4 && q < d.I4 + ss[L.gsolve[g]]->plan.nI4) return d.I4_32 + (q - d.I4);
          if (q >= d.W0 && q < d.W0 + ss[L.gsolve[g]]->plan.nI4) return d.W0_32 + (q - d.W0);
          throw Error("ring GEMM operand outside the ring blocks");
        };
        G.A = mv(G.A);
        G.B = (const double2 *)mv(G.B);
        G.C = (double2 *)mv(G.C);
        G.C0 = (const double2 *)mv(G.C0);
      }
    };
    to32(B->g_i4f);
    to32(B->g_rtf);
  }
  const int fams[13] = {0, 1, 2, 3, 4, 1, 0, 2, 3, 4, 3, 4, 1};
  GemmList *lists[13] = {&B->g_x, &B->g_pp, &B->g_hh, &B->g_i4, &B->g_rt, &B->g_pp32,
                         &B->g_xb, &B->g_hhb, &B->g_i4b, &B->g_rtb, &B->g_i4f, &B->g_rtf, &B->g_ppbf};
  B->d_descs = dupload(B->h_descs, B->owned);
  for (int k = 0; k < 13; ++k) {
    lists[k]->fam = fams[k];
    lists[k]->bf3 = k >= 6 && k < 12;
    lists[k]->xcd = fams[k] == 1 || k >= 6;
    lists[k]->nosplit = lists[k]->nosplit || (k >= 10 && k < 12);   // FP32 outputs: no FP64 partials
    if (k >= 6 && k < 12) {
      const GemmShape def[GemmList::NCLS] = {{1, 1}, {1, 2}, {1, 4}, {1, 8}, {1, 6}, {0, 4}};
      for (int c = 0; c < GemmList::NCLS; ++c) lists[k]->shape[c] = def[c];
      // CCGPU_ONE_LAUNCH=1 (studies): every MFMA class of a BF16x3 family on
      // 64 x 64 tiles in one launch (the thin class keeps its own)
      if (getenv("CCGPU_ONE_LAUNCH") && atoi(getenv("CCGPU_ONE_LAUNCH")) == 1) {
        for (int c = 0; c < GemmList::NCLS - 1; ++c) lists[k]->shape[c] = GemmShape{1, 4};
        lists[k]->merged = true;
      }
    }
    if (fams[k] == 1) {
      // the pp ladder on 64 x 16 tiles of the LDS-pipelined kernel (four
      // waves, two workgroups per CU) for every class, all in one launch:
      // thousands of tiles, so no split-K and no partial products to HBM and
      // back, and the launch's tail is short tiles instead of one class's
      // last round.  Measured at PNM A=66 |n_i|<=4: 0.48 ms against 0.61 ms
      // per class with split-K; one launch per form (64 x 32 eight-wave
      // tiles for N > 16, which stream V once) took 0.54 ms, the wide launch
      // being too short to fill the GPU.
      lists[k]->nosplit = true;
      for (int c = 0; c < GemmList::NCLS; ++c) lists[k]->shape[c] = GemmShape{1, 1, 3};
      lists[k]->merged = true;
    }
    lists[k]->apply_env_shapes();
  }
  HIPCHK(hipStreamCreateWithFlags(&B->stream2, hipStreamNonBlocking));
  HIPCHK(hipStreamCreateWithFlags(&B->stream3, hipStreamNonBlocking));
  HIPCHK(hipEventCreateWithFlags(&B->ev_fork, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&B->ev_join, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&B->ev_join3, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&B->ev_x, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&B->ev_hh, hipEventDisableTiming));
  B->d_slot = dalloc<int>(n, B->owned);
  B->d_slots = dalloc<int>((size_t)n * 16, B->owned);
  B->d_coef = dalloc<double2>((size_t)n * 16, B->owned);
  B->d_aaslots = dalloc<int>((size_t)n * 16, B->owned);
  B->d_aacoef = dalloc<double>((size_t)n * 16, B->owned);
  B->status.assign(n, 0);
  if (share_enabled() && share_eligible(ss)) build_shared(*B);
  return B;
}

// ---------------------------------------------------------------------------
// Shared pp ladder (see Batch::Shared).  A batch qualifies when it holds two
// or more solves of one lattice (A, matter, density, nmax; twists free) whose
// V comes from the same built-in interaction with no three-nucleon part in
// V_pppp.  CCGPU_SHARE_V=0 keeps one ladder per solve.
// ---------------------------------------------------------------------------
bool share_enabled() {
  const char *e = getenv("CCGPU_SHARE_V");
  return !(e && atoi(e) == 0);
}

bool same_chiral(const ChiralModel &a, const ChiralModel &b) {
  auto eq = [](const double *x, const double *y, int n) {
    for (int i = 0; i < n; ++i)
      if (x[i] != y[i]) return false;
    return true;
  };
  return a.order == b.order && a.cutoff == b.cutoff && a.nexp == b.nexp && a.c1 == b.c1 && a.c2 == b.c2 &&
         a.c3 == b.c3 && a.c4 == b.c4 && eq(a.cs, b.cs, 3) && eq(a.ct, b.ct, 3) && eq(a.cnlo, b.cnlo, 7) &&
         eq(a.cg0, b.cg0, 2) && eq(a.cg1, b.cg1, 2);
}

bool share_eligible(const std::vector<Solve *> &ss) {
  if (ss.size() < 2) return false;
  const Solve &a = *ss[0];
  const Orbitals &oa = a.plan.orb;
  for (const Solve *p : ss) {
    const Solve &b = *p;
    const Orbitals &ob = b.plan.orb;
    if (!b.native_v || b.vkind != a.vkind || b.tnf >= 2 || a.tnf >= 2) return false;
    if (ob.A != oa.A || ob.nsp != oa.nsp || ob.nmax != oa.nmax || ob.rho != oa.rho || ob.kf != oa.kf ||
        ob.volume != oa.volume)
      return false;
    if (b.vkind == CCGPU_INT_CHIRAL && !same_chiral(a.chiral, b.chiral)) return false;
  }
  return true;
}

// Union channels, pairs and position maps; the lists of the wide GEMMs.
void build_shared(Batch &B) {
  Batch::Shared &X = B.share;
  const std::vector<Solve *> &ss = B.solves;
  const int n = (int)ss.size();
  const Orbitals &o0 = ss[0]->plan.orb;
  std::map<std::array<int, 5>, int> lab0;
  for (int x = 0; x < o0.nsp; ++x) lab0[{o0.nx[x], o0.ny[x], o0.nz[x], o0.sz[x], o0.tz[x]}] = x;
  std::vector<std::vector<int>> to0(n);
  for (int s = 0; s < n; ++s) {
    const Orbitals &o = ss[s]->plan.orb;
    to0[s].resize(o.nsp);
    for (int x = 0; x < o.nsp; ++x) {
      auto it = lab0.find({o.nx[x], o.ny[x], o.nz[x], o.sz[x], o.tz[x]});
      if (it == lab0.end()) throw Error("shared ladder: orbital labels differ between the solves");
      to0[s][x] = it->second;
    }
  }
  // union channels in key order, their pairs (x < y in solve 0's numbering)
  std::map<std::array<int, 4>, int> uch;
  std::vector<std::vector<int>> chof(n);   // solve channel -> union channel
  for (int s = 0; s < n; ++s) {
    const Plan &P = ss[s]->plan;
    chof[s].resize(P.nch);
    for (int c = 0; c < P.nch; ++c) uch.emplace(std::array<int, 4>{P.key_tz[c], P.key_kx[c], P.key_ky[c], P.key_kz[c]}, 0);
  }
  int u = 0;
  for (auto &kv : uch) kv.second = u++;
  X.nu = u;
  std::vector<std::vector<uint64_t>> pairs(X.nu);
  const uint64_t W = (uint64_t)o0.nsp;
  for (int s = 0; s < n; ++s) {
    const Plan &P = ss[s]->plan;
    for (int c = 0; c < P.nch; ++c) {
      const int uc = uch.at({P.key_tz[c], P.key_kx[c], P.key_ky[c], P.key_kz[c]});
      chof[s][c] = uc;
      for (int r = 0; r < P.npp[c]; ++r) {
        int x = to0[s][P.pp_a[P.off_pp[c] + r]], y = to0[s][P.pp_b[P.off_pp[c] + r]];
        if (x > y) std::swap(x, y);
        pairs[uc].push_back((uint64_t)x * W + y);
      }
    }
  }
  X.npp.assign(X.nu, 0);
  X.ncol.assign(X.nu, 0);
  X.off_pp = {0};
  X.off_v = {0};
  X.off_t = {0};
  X.pa.clear();
  X.pb.clear();
  std::vector<std::unordered_map<uint64_t, int>> row(X.nu);
  std::vector<std::vector<int>> col(X.nu, std::vector<int>(n, -1));
  for (int s = 0; s < n; ++s) {
    const Plan &P = ss[s]->plan;
    for (int c = 0; c < P.nch; ++c) {
      const int uc = chof[s][c];
      col[uc][s] = X.ncol[uc];
      X.ncol[uc] += P.nhh[c];
    }
  }
  X.max_npp = 0;
  for (int uc = 0; uc < X.nu; ++uc) {
    auto &pl = pairs[uc];
    std::sort(pl.begin(), pl.end());
    pl.erase(std::unique(pl.begin(), pl.end()), pl.end());
    X.npp[uc] = (int)pl.size();
    X.max_npp = std::max(X.max_npp, X.npp[uc]);
    for (size_t r = 0; r < pl.size(); ++r) {
      row[uc][pl[r]] = (int)r;
      X.pa.push_back((int)(pl[r] / W));
      X.pb.push_back((int)(pl[r] % W));
    }
    X.off_pp.push_back(X.off_pp.back() + X.npp[uc]);
    X.off_v.push_back(X.off_v.back() + (int64_t)X.npp[uc] * X.npp[uc]);
    X.off_t.push_back(X.off_t.back() + (int64_t)X.npp[uc] * X.ncol[uc]);
  }
  X.nV = X.off_v.back();
  X.nTu = X.off_t.back();
  if (X.nTu >= (int64_t)1 << 31) throw Error("shared ladder: union blocks above 2^31 elements");
  // position maps
  std::vector<uint32_t *> maps(n);
  for (int s = 0; s < n; ++s) {
    const Plan &P = ss[s]->plan;
    std::vector<uint32_t> m(P.nT);
    for (int c = 0; c < P.nch; ++c) {
      const int uc = chof[s][c], nh = P.nhh[c];
      for (int r = 0; r < P.npp[c]; ++r) {
        const int x = to0[s][P.pp_a[P.off_pp[c] + r]], y = to0[s][P.pp_b[P.off_pp[c] + r]];
        const int ur = row[uc].at((uint64_t)std::min(x, y) * W + std::max(x, y));
        const int64_t base = X.off_t[uc] + (int64_t)ur * X.ncol[uc] + col[uc][s];
        for (int j = 0; j < nh; ++j) m[P.off_t[c] + (int64_t)r * nh + j] = (uint32_t)(2 * (base + j) + (x > y));
      }
    }
    maps[s] = dupload(m, B.owned);
  }
  X.d_maps = dupload(maps, B.owned);
  X.d_npp = dupload(X.npp, B.owned);
  X.d_pa = dupload(X.pa, B.owned);
  X.d_pb = dupload(X.pb, B.owned);
  X.d_off_pp = dupload(X.off_pp, B.owned);
  X.d_off_v = dupload(X.off_v, B.owned);
  X.V = dalloc<double2>(std::max<int64_t>(X.nV, 1), B.owned);
  X.V32 = dalloc<float2>(std::max<int64_t>(X.nV, 1), B.owned);
  X.Tu = dalloc<double2>(std::max<int64_t>(X.nTu, 1), B.owned);
  X.Ru = dalloc<double2>(std::max<int64_t>(X.nTu, 1), B.owned);
  HIPCHK(hipMemset(X.Tu, 0, std::max<int64_t>(X.nTu, 1) * sizeof(double2)));   // positions no solve owns
  // the wide GEMMs: R_u = V_u T_u per union channel
  X.g64 = GemmList();
  for (int uc = 0; uc < X.nu; ++uc) {
    const int m_ = X.npp[uc], nc = X.ncol[uc];
    X.g64.add(mkgroup(X.V + X.off_v[uc], X.Tu + X.off_t[uc], X.Ru + X.off_t[uc], nullptr, m_, nc, m_, m_, nc, nc, 0,
                      1.0, 0.0),
              0);
  }
  X.gbf = X.g64;
  for (size_t g = 0; g < X.gbf.groups.size(); ++g)
    X.gbf.groups[g].A = X.V32 + (((const double2 *)X.g64.groups[g].A) - X.V);
  X.g64.fam = X.gbf.fam = 5;
  X.gbf.bf3 = true;
  X.g64.xcd = X.gbf.xcd = true;
  const GemmShape def[GemmList::NCLS] = {{1, 1}, {1, 2}, {1, 4}, {1, 8}, {1, 6}, {0, 4}};
  for (int c = 0; c < GemmList::NCLS; ++c) X.gbf.shape[c] = def[c];
  X.g64.apply_env_shapes();
  X.gbf.apply_env_shapes();
  X.vgen.assign(n, ~0ull);
  X.on = true;
}

// (Re)fill the union V_pppp and its FP32 copy when a solve's V changed.
void refresh_shared(Batch &B) {
  Batch::Shared &X = B.share;
  if (!X.on) return;
  bool stale = false;
  for (size_t s = 0; s < B.solves.size(); ++s) stale = stale || X.vgen[s] != B.solves[s]->vgen;
  if (!stale) return;
  if (!share_eligible(B.solves)) throw Error("shared pp ladder: the solves no longer share one interaction");
  Solve &S0 = *B.solves[0];
  SolveDesc dv = S0.desc;
  dv.npp = X.d_npp;
  dv.off_pp = X.d_off_pp;
  dv.pp_a = X.d_pa;
  dv.pp_b = X.d_pb;
  dv.off_vpp = X.d_off_v;
  dv.Vpppp = X.V;
  launch_fill_vpppp(dv, S0.plan.orb.volume, S0.vkind == CCGPU_INT_CHIRAL ? &S0.chiral_dev : nullptr, X.nu,
                    X.max_npp, B.stream);
  check_launch();
  launch_to_f32(X.V, X.V32, X.nV, B.stream);
  check_launch();
  HIPCHK(hipStreamSynchronize(B.stream));
  for (size_t s = 0; s < B.solves.size(); ++s) X.vgen[s] = B.solves[s]->vgen;
}

// Shapes, tiles and device copies of the work lists a residual in precision
// prec needs, made on first use (outside any graph capture): lists a batch
// never contracts (a single-solve batch used only for its MBPT2 start or
// energy, the FP32 ladder of an FP64 solve) are never tuned.
void prepare_lists(Batch &B, int prec) {
  const int kinds[13] = {0, 3, 0, 0, 0, 4, 6, 6, 6, 6, 8, 8, 10};   // the kernel kind each list is tuned for
  GemmList *lists[13] = {&B.g_x, &B.g_pp, &B.g_hh, &B.g_i4, &B.g_rt, &B.g_pp32, &B.g_xb, &B.g_hhb, &B.g_i4b, &B.g_rtb,
                         &B.g_i4f, &B.g_rtf, &B.g_ppbf};
  for (int k = 0; k < 13; ++k) {
    GemmList &L = *lists[k];
    if (L.ready) continue;
    if (k == 5 && (prec == CCGPU_PREC_FP64 || !B.has_f32)) continue;
    if (k >= 6 && prec != CCGPU_PREC_BF16X3) continue;
    if (k >= 10 && k < 12 && !B.ring32) continue;
    if (k == 12 && !B.ppbf) continue;
    autotune(L, kinds[k], B.d_descs, B.stream);
    L.build_tiles();
    L.upload(B.owned);
    L.ready = true;
  }
  if (B.share.on) {
    refresh_shared(B);
    GemmList *sl[2] = {&B.share.g64, &B.share.gbf};
    const int sk[2] = {0, 7};
    for (int k = 0; k < 2; ++k) {
      GemmList &L = *sl[k];
      if (L.ready || (k == 1 && prec == CCGPU_PREC_FP64)) continue;
      autotune(L, sk[k], nullptr, B.stream);
      L.build_tiles();
      L.upload(B.owned);
      L.ready = true;
    }
  }
}

void set_active(Batch &B, int s, int active) {
  B.h_descs[s].active = active;
  HIPCHK(hipMemcpyAsync(&B.d_descs[s].active, &B.h_descs[s].active, sizeof(int), hipMemcpyHostToDevice, B.stream));
  HIPCHK(hipStreamSynchronize(B.stream));
}

// CCGPU_TERMS (bit mask, default all): 1 one-body, 2 pp ladder, 4 hh ladder,
// 8 ring.  Used by tests to compare single residual terms with the oracle.
int term_mask() {
  const char *e = getenv("CCGPU_TERMS");
  return e ? atoi(e) : 15;
}

// One GEMM family: three launches (one per tile class), or the VALU check.
// prec here is the arithmetic of this family: FP64, FP32, V32 (pp ladder
// only: FP32-stored A, FP64 arithmetic) or BF16x3 (the pp ladder on the FP32
// copy of V, the other families on FP64-stored A)
int gemm_kind(int prec, bool a_f32, bool pp) {
  if (pp) return prec == CCGPU_PREC_FP64 ? 3 : prec == CCGPU_PREC_V32 ? 5 : prec == CCGPU_PREC_BF16X3 ? 9 : 4;
  if (prec == CCGPU_PREC_BF16X3) return 6;
  return prec == CCGPU_PREC_FP64 ? 0 : (a_f32 ? 2 : 1);
}
bool known_precision(int p) {
  return p == CCGPU_PREC_FP64 || p == CCGPU_PREC_FP32 || p == CCGPU_PREC_V32 || p == CCGPU_PREC_BF16X3;
}

void run_gemm(const GemmList &L, const Batch &B, int prec, bool a_f32, hipStream_t st, bool pp = false,
              int kind_override = -1) {
  // the shared pp ladder's groups span every solve of the batch: no per-solve skip
  const SolveDesc *dd = L.fam == 5 ? nullptr : B.d_descs;
  if (gemm_variant_override() == 3) {
    // the VALU check reads FP64-stored operands only
    if (a_f32) throw Error("CCGPU_GEMM=valu checks contractions on FP64-stored operands only");
    launch_gemm_valu(L.d_groups, L.d_tiles64, (int)L.tiles64.size(), dd, L.d_gsolve, st);
  } else {
    const int kind = kind_override >= 0 ? kind_override : gemm_kind(prec, a_f32, pp);
    // CCGPU_TIMING=2 also times every tile class of every family
    const bool detail = getenv("CCGPU_TIMING") && atoi(getenv("CCGPU_TIMING")) == 2;
    for (int c = GemmList::NCLS - 1; c >= 0; --c) {
      if (L.tiles[c].empty()) continue;
      char nm[48];
      snprintf(nm, sizeof nm, "fam%d.k%d.c%d.%dx%d%s", L.fam, kind, c, L.shape[c].mb, L.shape[c].nb,
               L.shape[c].lb == 4 ? "W" : L.shape[c].lb == 3 ? "P" : L.shape[c].lb == 2 ? "A" : L.shape[c].lb ? "L" : "");
      Timer t(detail, st, nm);
      launch_gemm_shape(L.d_groups, L.d_tiles[c], (int)L.tiles[c].size(), dd, L.d_gsolve, kind, L.shape[c].mb,
                        L.shape[c].nb, L.shape[c].lb, st);
    }
    // split groups: the MFMA and thin kernels wrote partials; the VALU check
    // above runs every group over its whole k range and writes C itself
    launch_splitk_reduce(L.d_groups, L.d_split, (int)L.split.size(), L.maxmn, dd, L.d_gsolve, st);
  }
}

// Families in low precision under CCGPU_PREC_FP32 and CCGPU_PREC_BF16X3 (bit 0
// X/I9 GEMM, 1 pp ladder, 2 hh ladder, 3 ring GEMMs); all by default,
// CCGPU_FP32_FAMILIES narrows it down for precision studies.  The low-precision
// pp ladder reads the FP32 copy of V_pppp.
int fp32_families() {
  const char *e = getenv("CCGPU_FP32_FAMILIES");
  return e ? atoi(e) : 15;
}

// CCGPU_TAIL_LOW=mask (precision studies): in FP64 and V32 residuals, the
// families of the mask (bits as CCGPU_FP32_FAMILIES) run in BF16x3
int tail_low_mask() {
  const char *e = getenv("CCGPU_TAIL_LOW");
  return e ? atoi(e) : 0;
}

bool tsplit_enabled() {
  const char *e = getenv("CCGPU_PPL_TSPLIT");
  return !(e && atoi(e) == 0);
}

bool fork_enabled() {
  const char *e = getenv("CCGPU_FORK");
  return !(e && atoi(e) == 0);
}

void residual_launches(Batch &B, int prec, bool timing) {
  const int n = (int)B.solves.size();
  hipStream_t st = B.stream;
  const int terms = term_mask();
  if (prec != CCGPU_PREC_FP64 && !B.has_f32) throw Error("FP32/V32 residual without the FP32 copy of Vpppp");
  const bool low = prec == CCGPU_PREC_FP32 || prec == CCGPU_PREC_BF16X3;
  const int fam = low ? fp32_families() : 0;
  const int tlow = low ? 0 : tail_low_mask();
  auto pf = [&](int bit) {
    if ((tlow >> bit) & 1) return (int)CCGPU_PREC_BF16X3;
    return (fam >> bit) & 1 ? prec : (int)CCGPU_PREC_FP64;
  };
  const int pp_prec = (tlow >> 1) & 1 ? CCGPU_PREC_BF16X3 : prec == CCGPU_PREC_V32 ? CCGPU_PREC_V32 : pf(1);
  // the BF16x3 kernel runs its own lists (wider tiles)
  auto list = [&](GemmList &L64, GemmList &Lbf, int bit) -> const GemmList & {
    return pf(bit) == CCGPU_PREC_BF16X3 ? Lbf : L64;
  };
  // Three streams outside the timing mode (fork/join through events, also
  // inside a captured graph).  The ring chain (T -> T~, I4, R~) reads only T
  // and the ring blocks and touches R only in the final scatter: a second
  // stream from the start of the residual, beside the one-body intermediates,
  // whose small launches leave most of the chip idle, and beside the ladders,
  // which stream V_pppp from HBM while the ring GEMMs keep the matrix cores
  // busy.  A third stream takes I6 (rows of V* against rows of T) and the split
  // copy of T for the BF16x3 pp ladder, both joined before the residual start,
  // then the hh ladder into Rh once X is done, joined before the scatter that
  // adds it: R = (Rh + R) + ring, the order of the single-stream sums.
  const bool fork = !timing && (terms & 8) && fork_enabled();
  const bool fork6 = !timing && fork_enabled();
  hipStream_t sr = fork ? B.stream2 : st;
  hipStream_t s6 = fork6 ? B.stream3 : st;
  // BF16x3 ring on the FP32 ring blocks when every solve has them
  const bool ring32 = pf(3) == CCGPU_PREC_BF16X3 && B.ring32;
  // BF16x3 reads T split once (k_t_split) rather than per wave and row tile;
  // CCGPU_PPL_TSPLIT=0 keeps the in-kernel split (same bits)
  const bool tsplit = (terms & 2) && !B.share.on && pp_prec == CCGPU_PREC_BF16X3 && B.ppbf && tsplit_enabled();
  if (fork || fork6) HIPCHK(hipEventRecord(B.ev_fork, st));
  if (fork) HIPCHK(hipStreamWaitEvent(sr, B.ev_fork, 0));
  if (fork6) HIPCHK(hipStreamWaitEvent(s6, B.ev_fork, 0));
  // launch (and graph node) order: the one-body chain first, the ring after it
  {
    Timer t(timing, s6, "i6");
    launch_i6(B.d_descs, n, B.maxNp, B.maxRows, s6);
  }
  if (tsplit) {
    Timer t(timing, s6, "pp_t_split");
    launch_t_split(B.d_descs, n, B.maxT, s6);
  }
  if (fork6) HIPCHK(hipEventRecord(B.ev_join3, s6));
  {
    Timer t(timing, st, "gemm_x_vhhpp_t");
    run_gemm(list(B.g_x, B.g_xb, 0), B, pf(0), false, st);
  }
  {
    Timer t(timing, st, "i7_i9");
    launch_i7(B.d_descs, n, B.maxA, st);
    launch_i9(B.d_descs, n, B.maxVhh, st);
  }
  if (terms & 4) {
    if (fork6) {
      HIPCHK(hipEventRecord(B.ev_x, st));
      HIPCHK(hipStreamWaitEvent(s6, B.ev_x, 0));
    }
    Timer t(timing, s6, "gemm_hh_ladder");
    run_gemm(list(B.g_hh, B.g_hhb, 2), B, pf(2), false, s6);
  }
  if (fork6) HIPCHK(hipEventRecord(B.ev_hh, s6));
  if (terms & 8) {
    {
      Timer t(timing, sr, "ring_gather");
      launch_ring_gather(B.d_descs, n, B.maxT, ring32, sr);
    }
    {
      Timer t(timing, sr, "gemm_ring_i4");
      if (ring32) run_gemm(B.g_i4f, B, pf(3), true, sr, false, 8);
      else run_gemm(list(B.g_i4, B.g_i4b, 3), B, pf(3), false, sr);
    }
    {
      Timer t(timing, sr, "gemm_ring_rt");
      if (ring32) run_gemm(B.g_rtf, B, pf(3), true, sr, false, 8);
      else run_gemm(list(B.g_rt, B.g_rtb, 3), B, pf(3), false, sr);
    }
  }
  if (fork6) HIPCHK(hipStreamWaitEvent(st, B.ev_join3, 0));
  {
    Timer t(timing, st, "resid_init_onebody");
    launch_resid_init(B.d_descs, n, B.maxT, (terms & 1) != 0, st);
  }
  if ((terms & 2) && B.share.on) {
    // shared pp ladder: FP64 GEMMs on V_u, or BF16x3 on its FP32 copy in every
    // low-precision mode
    {
      Timer t(timing, st, "pp_share_gather");
      launch_share_gather(B.d_descs, n, B.maxT, B.share.d_maps, B.share.Tu, st);
    }
    {
      Timer t(timing, st, "gemm_pp_ladder");
      if (pp_prec == CCGPU_PREC_FP64) run_gemm(B.share.g64, B, CCGPU_PREC_FP64, false, st, false, 0);
      else run_gemm(B.share.gbf, B, CCGPU_PREC_BF16X3, true, st, false, 7);
    }
    {
      Timer t(timing, st, "pp_share_scatter");
      launch_share_scatter(B.d_descs, n, B.maxT, B.share.d_maps, B.share.Ru, st);
    }
  } else if (terms & 2) {
    Timer t(timing, st, "gemm_pp_ladder");
    if (tsplit) run_gemm(B.g_ppbf, B, pp_prec, true, st, true, 10);
    else if (pp_prec != CCGPU_PREC_FP64) run_gemm(B.g_pp32, B, pp_prec, true, st, true);
    else run_gemm(B.g_pp, B, pp_prec, false, st, true);
  }
  if (fork) {
    HIPCHK(hipEventRecord(B.ev_join, sr));
    HIPCHK(hipStreamWaitEvent(st, B.ev_join, 0));
  }
  if (fork6) HIPCHK(hipStreamWaitEvent(st, B.ev_hh, 0));
  if (terms & 12) {
    Timer t(timing, st, "ring_scatter");
    launch_ring_scatter(B.d_descs, n, B.maxT, ring32, (terms & 8) != 0, (terms & 4) != 0, st);
  }
}

void check_launch() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw Error(fmt("kernel launch failed: %s", hipGetErrorString(e)));
}

// CCGPU_FORCE_SINGULAR=s (tests of the failure handling): the first DIIS
// extrapolation (or Anderson system of two or more pairs) of solve s of a
// batch in a run is treated as singular, as a duplicate history would make it.
int forced_singular() {
  const char *e = getenv("CCGPU_FORCE_SINGULAR");
  return e ? atoi(e) : -1;
}

// Update of every active solve (diis_t1t2, ccm_diis.f90:74-249); all active solves
// of a batch share count/nstep because they start together.
void update_launches(Batch &B, int &count, int &nstep, int subspace, int diis_step, bool timing) {
  const int n = (int)B.solves.size();
  hipStream_t st = B.stream;
  // history column handling: shift when full (ring buffer)
  std::vector<int> &slot = B.h_slot, &slots = B.h_slots;
  slot.assign(n, 0);
  slots.assign((size_t)n * 16, 0);
  for (int s = 0; s < n; ++s) {
    Solve &S = *B.solves[s];
    if (nstep > subspace) {
      // drop the oldest column
      std::rotate(S.ring.begin(), S.ring.begin() + 1, S.ring.begin() + subspace);
    }
    slot[s] = S.ring[std::min(nstep, subspace) - 1];
  }
  if (nstep > subspace) nstep = subspace;
  HIPCHK(hipMemcpyAsync(B.d_slot, slot.data(), n * sizeof(int), hipMemcpyHostToDevice, st));
  {
    Timer t(timing, st, "update");
    launch_update(B.d_descs, n, B.maxT, B.d_slot, st);
  }
  if (count != diis_step) return;
  count = 0;
  if (nstep < 2) return;
  for (int s = 0; s < n; ++s)
    for (int k = 0; k < nstep; ++k) slots[(size_t)s * 16 + k] = B.solves[s]->ring[k];
  h2d_pinned(B, 4, B.d_slots, slots.data(), slots.size() * sizeof(int), st);
  std::vector<double2> &sc = B.h_scal;
  sc.assign((size_t)n * 64, double2{0, 0});
  {
    Timer t(timing, st, "diis_gram");
    launch_gram(B.d_descs, n, B.d_slots, nstep, st);
    check_launch();
    d2h_pinned_sync(B, 3, sc.data(), B.d_scal, sc.size() * sizeof(double2), st);
  }
  std::vector<double2> &coef = B.h_coef;
  coef.assign((size_t)n * 16, double2{0, 0});
  const int ne = nstep - 1;
  for (int s = 0; s < n; ++s) {
    if (!B.h_descs[s].active) continue;
    std::vector<std::complex<double>> M((size_t)nstep * nstep, 0.0), rhs(nstep, 0.0);
    int q = 0;
    for (int p = 0; p < ne; ++p)
      for (int p2 = p; p2 < ne; ++p2) {
        double2 v = sc[(size_t)s * 64 + 4 + q++];
        M[p * nstep + p2] = M[p2 * nstep + p] = std::complex<double>(v.x, v.y);
      }
    for (int p = 0; p < ne; ++p) M[p * nstep + ne] = M[ne * nstep + p] = -1.0;
    M[ne * nstep + ne] = 0.0;
    rhs[ne] = -1.0;
    int info = 0;
    auto sol = zsolve(M, rhs, nstep, &info);
    const bool forced = s == forced_singular() && !B.forced_done;
    if (forced) B.forced_done = true;
    bool finite = info == 0 && !forced;
    for (int p = 0; p < ne && finite; ++p) finite = std::isfinite(sol[p].real()) && std::isfinite(sol[p].imag());
    if (!finite) {
      // this solve keeps its Jacobi-updated amplitudes (the last history
      // column) for this step; the other solves of the batch extrapolate
      slots[(size_t)s * 16] = B.solves[s]->ring[nstep - 1];
      coef[(size_t)s * 16] = double2{1.0, 0.0};
      B.status[s] |= CCGPU_STATUS_DIIS_FALLBACK;
      continue;
    }
    for (int p = 0; p < ne; ++p) coef[(size_t)s * 16 + p] = double2{sol[p].real(), sol[p].imag()};
  }
  h2d_pinned(B, 5, B.d_slots, slots.data(), slots.size() * sizeof(int), st);
  h2d_pinned(B, 6, B.d_coef, coef.data(), coef.size() * sizeof(double2), st);
  {
    Timer t(timing, st, "diis_axpy");
    launch_diis_axpy(B.d_descs, n, B.maxT, B.d_slots, B.d_coef, ne, st);
  }
}

// one sum over the triples, defined with the (T) entry points below
double2 triples_sum(Solve &S, int mode);

std::vector<std::complex<double>> energies(Batch &B, bool timing) {
  const int n = (int)B.solves.size();
  {
    Timer t(timing, B.stream, "energy");
    launch_energy(B.d_descs, n, 0, B.stream);
  }
  check_launch();
  std::vector<double2> &sc = B.h_scal;
  sc.assign((size_t)n * 64, double2{0, 0});
  d2h_pinned_sync(B, 3, sc.data(), B.d_scal, sc.size() * sizeof(double2), B.stream);
  std::vector<std::complex<double>> e(n);
  for (int s = 0; s < n; ++s) e[s] = std::complex<double>(sc[(size_t)s * 64].x, sc[(size_t)s * 64].y);
  return e;
}

// Solution of min |sum_j c_j f_j|^2 subject to sum_j c_j = 1 for the Gram
// matrix G (m x m, row-major): the system G c = 1 scaled by the diagonal, then
// c normalized to unit sum.  Returns false when a pivot of the scaled matrix
// falls below 1e-13 (near-parallel residual vectors) or the result is not
// finite.
bool anderson_coefficients(const std::vector<double> &G, int m, std::vector<double> &c) {
  std::vector<double> sc(m), a((size_t)m * m), b(m, 0.0);
  for (int j = 0; j < m; ++j) {
    const double gjj = G[(size_t)j * m + j];
    if (!(gjj > 0.0) || !std::isfinite(gjj)) return false;
    sc[j] = 1.0 / std::sqrt(gjj);
  }
  for (int j = 0; j < m; ++j) {
    b[j] = sc[j];
    for (int k = 0; k < m; ++k) a[(size_t)j * m + k] = sc[j] * G[(size_t)j * m + k] * sc[k];
  }
  for (int k = 0; k < m; ++k) {
    int piv = k;
    for (int i = k + 1; i < m; ++i)
      if (std::fabs(a[(size_t)i * m + k]) > std::fabs(a[(size_t)piv * m + k])) piv = i;
    if (!(std::fabs(a[(size_t)piv * m + k]) > 1e-13)) return false;
    if (piv != k) {
      for (int j = 0; j < m; ++j) std::swap(a[(size_t)k * m + j], a[(size_t)piv * m + j]);
      std::swap(b[k], b[piv]);
    }
    for (int i = k + 1; i < m; ++i) {
      const double f = a[(size_t)i * m + k] / a[(size_t)k * m + k];
      for (int j = k; j < m; ++j) a[(size_t)i * m + j] -= f * a[(size_t)k * m + j];
      b[i] -= f * b[k];
    }
  }
  c.assign(m, 0.0);
  for (int i = m - 1; i >= 0; --i) {
    double v = b[i];
    for (int j = i + 1; j < m; ++j) v -= a[(size_t)i * m + j] * c[j];
    c[i] = v / a[(size_t)i * m + i];
  }
  double sum = 0.0;
  for (int j = 0; j < m; ++j) {
    c[j] *= sc[j];
    sum += c[j];
  }
  if (!std::isfinite(sum) || sum == 0.0) return false;
  for (int j = 0; j < m; ++j) c[j] /= sum;
  return true;
}

// One Anderson step of every active solve after its residual: k_aa_update
// forms f, y and the new Gram row; the host appends the row, solves for the
// coefficients (dropping the oldest vectors while the scaled system is near
// singular) and uploads them for k_aa_combine, which forms the next iterate at
// the start of the next iteration.  Returns the correlation energy of the next
// iterate, sum_j c_j E(y_j) (the energy is linear in T).
std::vector<std::complex<double>> anderson_step(Batch &B, int depth, double beta, bool timing) {
  const int n = (int)B.solves.size();
  hipStream_t st = B.stream;
  B.h_aaslots.assign((size_t)n * 16, 0);
  for (int s = 0; s < n; ++s) {
    Batch::AA &q = B.aa[s];
    int m = (int)q.slot.size();
    int fresh = -1;
    if (m == depth) {   // drop the oldest held vector, reuse its slot
      fresh = q.slot[0];
      std::vector<double> g((size_t)(m - 1) * (m - 1));
      for (int j = 1; j < m; ++j)
        for (int k = 1; k < m; ++k) g[(size_t)(j - 1) * (m - 1) + (k - 1)] = q.gram[(size_t)j * m + k];
      q.gram.swap(g);
      q.slot.erase(q.slot.begin());
      q.ey.erase(q.ey.begin());
      --m;
    } else {
      for (int c = 0; c < depth && fresh < 0; ++c)
        if (std::find(q.slot.begin(), q.slot.end(), c) == q.slot.end()) fresh = c;
    }
    for (int j = 0; j < m; ++j) B.h_aaslots[(size_t)s * 16 + j] = q.slot[j];
    B.h_aaslots[(size_t)s * 16 + 14] = fresh;
    B.h_aaslots[(size_t)s * 16 + 15] = m;
  }
  h2d_pinned(B, 0, B.d_aaslots, B.h_aaslots.data(), B.h_aaslots.size() * sizeof(int), st);
  {
    Timer t(timing, st, "anderson_update");
    launch_aa_update(B.d_descs, n, B.d_aaslots, beta, st);
  }
  check_launch();
  std::vector<double2> &sc = B.h_scal;
  sc.assign((size_t)n * 64, double2{0, 0});
  d2h_pinned_sync(B, 3, sc.data(), B.d_scal, sc.size() * sizeof(double2), st);
  std::vector<std::complex<double>> e(n, 0.0);
  B.h_aacoef.assign((size_t)n * 16, 0.0);
  for (int s = 0; s < n; ++s) {
    if (!B.h_descs[s].active) continue;
    Batch::AA &q = B.aa[s];
    const int m = (int)q.slot.size();
    const double2 *v = &sc[(size_t)s * 64 + 24];
    std::vector<double> g((size_t)(m + 1) * (m + 1));
    for (int j = 0; j < m; ++j)
      for (int k = 0; k < m; ++k) g[(size_t)j * (m + 1) + k] = q.gram[(size_t)j * m + k];
    for (int j = 0; j < m; ++j) g[(size_t)j * (m + 1) + m] = g[(size_t)m * (m + 1) + j] = v[j].x;
    g[(size_t)m * (m + 1) + m] = v[m].x;
    q.gram.swap(g);
    q.slot.push_back(B.h_aaslots[(size_t)s * 16 + 14]);
    q.ey.emplace_back(v[16].x, v[16].y);
    std::vector<double> c;
    const bool forced = s == forced_singular() && !B.forced_done && q.slot.size() >= 2;
    if (forced) B.forced_done = true;
    while (forced || !anderson_coefficients(q.gram, (int)q.slot.size(), c)) {
      B.status[s] |= CCGPU_STATUS_DIIS_FALLBACK;
      const int mm = (int)q.slot.size();
      if (mm == 1) {   // a single vector: the plain mixed step
        c.assign(1, 1.0);
        break;
      }
      if (forced) {    // keep only the newest pair
        c.assign(1, 1.0);
        q.gram.assign(1, q.gram[(size_t)mm * mm - 1]);
        q.slot.erase(q.slot.begin(), q.slot.end() - 1);
        q.ey.erase(q.ey.begin(), q.ey.end() - 1);
        break;
      }
      std::vector<double> g2((size_t)(mm - 1) * (mm - 1));
      for (int j = 1; j < mm; ++j)
        for (int k = 1; k < mm; ++k) g2[(size_t)(j - 1) * (mm - 1) + (k - 1)] = q.gram[(size_t)j * mm + k];
      q.gram.swap(g2);
      q.slot.erase(q.slot.begin());
      q.ey.erase(q.ey.begin());
    }
    const int mm = (int)q.slot.size();
    for (int j = 0; j < mm; ++j) {
      B.h_aaslots[(size_t)s * 16 + j] = q.slot[j];
      B.h_aacoef[(size_t)s * 16 + j] = c[j];
      e[s] += c[j] * q.ey[j];
    }
    B.h_aaslots[(size_t)s * 16 + 15] = mm;
  }
  h2d_pinned(B, 1, B.d_aaslots, B.h_aaslots.data(), B.h_aaslots.size() * sizeof(int), st);
  h2d_pinned(B, 2, B.d_aacoef, B.h_aacoef.data(), B.h_aacoef.size() * sizeof(double), st);
  return e;
}

// The ccd_iter loop (ccm_iter.f90:60-135) over a batch of solves.
void run_batch(Batch &B, const ccgpu_solve_opts &op, std::vector<ccgpu_solve_result> &res) {
  const int n = (int)B.solves.size();
  const bool timing = getenv("CCGPU_TIMING") && atoi(getenv("CCGPU_TIMING")) != 0;
  if (timing) { g_stats.clear(); g_stat_order.clear(); }
  if (op.accel != CCGPU_ACCEL_REFERENCE && op.accel != CCGPU_ACCEL_ANDERSON)
    throw Error(fmt("unknown accelerator %d", op.accel));
  if (op.confirm < 0 || op.confirm > 4) throw Error(fmt("confirm %d: 0 (default) or 1-4", op.confirm));
  const bool aa = op.accel == CCGPU_ACCEL_ANDERSON;
  if (aa && op.ccdt1) throw Error("CCDT-1 runs with the reference update (accel 0)");
  if (aa && !(op.mix >= 0.0 && op.mix <= 1.0)) throw Error(fmt("mix %g outside [0, 1]", op.mix));
  const double beta = op.mix > 0.0 ? op.mix : 1.0;
  for (Solve *S : B.solves) {
    alloc_history(*S, op.subspace, aa);
    S->ring.resize(op.subspace);
    for (int k = 0; k < op.subspace; ++k) S->ring[k] = k;
  }
  B.status.assign(n, 0);
  B.forced_done = false;
  B.aa.assign(n, Batch::AA{});
  // descs may have changed (history pointer)
  for (int s = 0; s < n; ++s) {
    B.h_descs[s].hist = B.solves[s]->desc.hist;
    B.h_descs[s].histf = B.solves[s]->desc.histf;
    B.h_descs[s].active = 1;
  }
  HIPCHK(hipMemcpy(B.d_descs, B.h_descs.data(), n * sizeof(SolveDesc), hipMemcpyHostToDevice));
  if (!known_precision(op.precision)) throw Error(fmt("unknown precision %d", op.precision));
  if (op.tail != CCGPU_PREC_FP64 && op.tail != CCGPU_PREC_V32)
    throw Error(fmt("tail precision %d: FP64 or V32 only", op.tail));
  if (op.precision != CCGPU_PREC_FP64 || op.tail != CCGPU_PREC_FP64) {
    for (Solve *S : B.solves)
      if (!S->desc.Vpppp32) throw Error("FP32/V32 ladder requested before the FP32 copy of Vpppp was made");
  }
  prepare_lists(B, op.precision);
  if (op.tail != CCGPU_PREC_FP64) prepare_lists(B, op.tail);
  auto e1 = energies(B, false);
  res.assign(n, ccgpu_solve_result{});
  std::vector<std::complex<double>> dener(n, 4.0), elast = e1;
  std::vector<int> iters(n, 0), conv(n, 0), iters32(n, 0);
  for (int s = 0; s < n; ++s) {
    Solve &S = *B.solves[s];
    res[s].e0[0] = S.e0.x; res[s].e0[1] = S.e0.y;
    res[s].embpt2[0] = S.e0.x + e1[s].real();
    res[s].embpt2[1] = S.e0.y + e1[s].imag();
  }
  // CCDT-1 (the CCDT1 branch of ccd_iter, ccm_iter.f90:26-92): the start adds
  // the energy of T3 = W3/D; every iteration builds T3 from the T2 before the
  // update and adds its energy to that of the updated T2
  const bool ccdt1 = op.ccdt1 != 0;
  std::vector<std::complex<double>> e3(n, 0.0);
  if (ccdt1)
    for (int s = 0; s < n; ++s) {
      Solve &S = *B.solves[s];
      S.ccdt1_e3 = triples_sum(S, 2);
      S.have_ccdt1 = true;
      e1[s] += std::complex<double>(S.ccdt1_e3.x, S.ccdt1_e3.y);
      elast[s] = e1[s];
    }
  for (int s = 0; s < n; ++s) B.solves[s]->ehist.assign(1, e1[s]);
  int count = 0, nstep = 0;
  auto t0 = std::chrono::steady_clock::now();
  bool graph_ok = op.use_graph != 0 && !timing && gemm_variant_override() < 0;
  int refine = op.refine_every > 0 ? op.refine_every : 0;
  // Mixed precision: low-precision contractions (FP32 or V32, with the usual
  // DIIS schedule) until
  //  (a) a DIIS extrapolation that follows an iteration with max |dE| <=
  //      fp32_switch (fp32_switch > 0), or
  //  (b) every solve has met the stopping test |dE| <= tol in low precision
  //      at least once (solves that met it keep iterating meanwhile).
  // From the next iteration on the contractions are FP64 and the DIIS history
  // carries on; at least one FP64 iteration follows, so the stopping test that
  // ends a solve is always an FP64 one.  No solve ends inside the first phase.
  bool fp32_phase = op.precision != CCGPU_PREC_FP64;
  double prev_maxd = 1e300;
  std::vector<char> reached(n, 0), failed(n, 0);
  // stopping test: |dE| <= tol on `need` consecutive iterations.  The
  // reference's update stops at the first (ccm_iter.f90:71-74); by default the
  // Anderson update waits for a second one, since one small energy change can
  // fall between two larger ones before the iterate has settled
  const int need = op.confirm > 0 ? op.confirm : (aa ? 2 : 1);
  std::vector<int> small(n, 0);
  bool pending = false;   // Anderson: coefficients uploaded, next iterate not yet formed
  for (int it = 1; it <= op.maxit; ++it) {
    if (pending) {
      Timer t(timing, B.stream, "anderson_combine");
      launch_aa_combine(B.d_descs, n, B.maxT, B.d_aaslots, B.d_aacoef, B.stream);
      check_launch();
      pending = false;
    }
    bool any = false;
    for (int s = 0; s < n; ++s) {
      bool act = !failed[s] && (fp32_phase ? B.h_descs[s].active != 0 : small[s] < need);
      if (!act && B.h_descs[s].active) {
        set_active(B, s, 0);
        conv[s] = !failed[s];
      }
      any = any || act;
    }
    if (!any) break;
    // the low-precision phase, then the tail precision (FP64, or V32)
    int prec = op.tail;
    bool low = false;
    if (fp32_phase) {
      prec = op.precision;
      low = true;
      if (refine && it % refine == 0) {
        prec = CCGPU_PREC_FP64;
        low = false;
      }
    }
    if (ccdt1)
      for (int s = 0; s < n; ++s) {
        if (!B.h_descs[s].active) continue;
        Solve &S = *B.solves[s];
        S.ccdt1_e3 = triples_sum(S, 1);
        e3[s] = std::complex<double>(S.ccdt1_e3.x, S.ccdt1_e3.y);
      }
    if (graph_ok) {
      const int q = prec;
      if (!B.gexec[q]) {
        HIPCHK(hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal));
        residual_launches(B, prec, false);
        HIPCHK(hipStreamEndCapture(B.stream, &B.graph[q]));
        HIPCHK(hipGraphInstantiate(&B.gexec[q], B.graph[q], nullptr, nullptr, 0));
      }
      HIPCHK(hipGraphLaunch(B.gexec[q], B.stream));
    } else {
      residual_launches(B, prec, timing);
    }
    check_launch();
    ++count;
    ++nstep;
    std::vector<std::complex<double>> e2;
    bool extrapolated;
    if (aa) {
      e2 = anderson_step(B, op.subspace, beta, timing);
      pending = true;
      extrapolated = true;
    } else {
      update_launches(B, count, nstep, op.subspace, op.diis_step, timing);
      check_launch();
      extrapolated = count == 0;
      e2 = energies(B, timing);
    }
    double maxd = 0;
    for (int s = 0; s < n; ++s) {
      if (!B.h_descs[s].active) continue;
      if (ccdt1) e2[s] += e3[s];
      if (!std::isfinite(e2[s].real()) || !std::isfinite(e2[s].imag())) {
        // a diverged solve stops here; the rest of the batch goes on
        failed[s] = 1;
        B.status[s] |= CCGPU_STATUS_NONFINITE;
        iters[s] = it;
        if (op.verbose) printf("solve %d: non-finite energy at iteration %d, stopped\n", s, it);
        continue;
      }
      dener[s] = e2[s] - elast[s];
      elast[s] = e2[s];
      small[s] = std::abs(dener[s]) <= op.tol ? small[s] + 1 : 0;
      B.solves[s]->ehist.push_back(e2[s]);
      iters[s] = it;
      if (low) ++iters32[s];
      maxd = std::max(maxd, std::abs(dener[s]));
      if (std::abs(dener[s]) <= op.tol) reached[s] = 1;
      if (op.verbose) printf("dener %5d %5d %16.8E  E %.12f%s\n", s, it, dener[s].real(), e2[s].real(),
                             prec == CCGPU_PREC_FP32     ? "  fp32"
                             : prec == CCGPU_PREC_V32    ? "  v32"
                             : prec == CCGPU_PREC_BF16X3 ? "  bf16x3"
                                                         : "");
    }
    if (fp32_phase) {
      const bool at_boundary =
          op.fp32_switch > 0.0 && extrapolated && (aa ? maxd : prev_maxd) <= op.fp32_switch;
      bool at_tol = true;
      for (int s = 0; s < n; ++s) at_tol = at_tol && (reached[s] || !B.h_descs[s].active);
      if (at_boundary || at_tol) {
        fp32_phase = false;
        // an FP32 |dE| is no FP64 convergence test: force FP64 steps
        for (int s = 0; s < n; ++s)
          if (B.h_descs[s].active) {
            dener[s] = std::max(std::abs(dener[s]), 2.0 * op.tol);
            small[s] = 0;
          }
      }
    }
    if (!extrapolated || aa) prev_maxd = maxd;
  }
  if (pending) {   // the iterate of the last Anderson step (iteration cap)
    launch_aa_combine(B.d_descs, n, B.maxT, B.d_aaslots, B.d_aacoef, B.stream);
    check_launch();
  }
  HIPCHK(hipStreamSynchronize(B.stream));
  double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  for (int s = 0; s < n; ++s) {
    if (!failed[s] && small[s] >= need) conv[s] = 1;
    if (failed[s]) conv[s] = 0;
    if (!conv[s] && !failed[s]) B.status[s] |= CCGPU_STATUS_MAXIT;
    res[s].status = B.status[s];
    Solve &S = *B.solves[s];
    res[s].iterations = iters[s];
    res[s].converged = conv[s];
    res[s].ecorr_last[0] = elast[s].real();
    res[s].ecorr_last[1] = elast[s].imag();
    res[s].eccd[0] = S.e0.x + elast[s].real();
    res[s].eccd[1] = S.e0.y + elast[s].imag();
    res[s].seconds = secs;
    res[s].fp32_iterations = iters32[s];
    S.have_t = true;
  }
  // leave every solve active for later single calls
  for (int s = 0; s < n; ++s) set_active(B, s, 1);
  HIPCHK(hipStreamSynchronize(B.stream));
}

std::unique_ptr<Batch> single_batch(Solve &S) {
  std::vector<Solve *> v{&S};
  return make_batch(v, S.stream);
}

// per-handle cached batch of one (for the step-by-step API)
std::map<int, std::unique_ptr<Batch>> g_single;
Batch &single(int h) {
  auto it = g_single.find(h);
  if (it != g_single.end()) return *it->second;
  Solve &S = get(h);
  g_single[h] = single_batch(S);
  return *g_single[h];
}
// cached multi-solve batches, keyed by their handle lists
std::map<std::vector<int>, std::unique_ptr<Batch>> g_batches;
void drop_single(int h) {
  g_single.erase(h);
  for (auto it = g_batches.begin(); it != g_batches.end();) {
    if (std::find(it->first.begin(), it->first.end(), h) != it->first.end()) it = g_batches.erase(it);
    else ++it;
  }
}

// (Re)fill the FP32 copy of the pppp blocks from the FP64 blocks.
void fill_f32(Solve &S) {
  const Plan &P = S.plan;
  launch_to_f32_rows(S.desc.Vpppp, S.desc.Vpppp32, S.d_off_vpp, S.d_off_vpp32, S.d_npp, P.nch, P.max_npp, S.stream);
  check_launch();
  if (S.desc.Vt32) {
    launch_to_f32(S.desc.Vt, S.desc.Vt32, P.nRt, S.stream);
    launch_to_f32(S.desc.W0, S.desc.W0_32, P.nI4, S.stream);
    check_launch();
  }
  HIPCHK(hipStreamSynchronize(S.stream));
}

// FP32 copy of the stored pppp blocks, made on the first FP32 request; the
// cached single-solve batch is rebuilt so that its FP32 ladder list reads it.
void ensure_f32(Solve &S, int handle) {
  if (S.desc.Vpppp32) return;
  const Plan &P = S.plan;
  S.off_vpp32.assign(1, 0);
  for (int c = 0; c < P.nch; ++c) {
    const int64_t n = P.npp[c];
    S.off_vpp32.push_back(S.off_vpp32.back() + n * ((n + 1) & ~1));
  }
  S.nVpp32 = S.off_vpp32.back();
  S.d_off_vpp = dupload(P.off_vpp, S.owned);
  S.d_off_vpp32 = dupload(S.off_vpp32, S.owned);
  S.d_npp = dupload(P.npp, S.owned);
  S.desc.Vpppp32 = dalloc<float2>(S.nVpp32, S.owned);
  // FP32 ring blocks of the BF16x3 residual; T~ entries with a=b or i=j stay zero
  S.desc.Tt32 = dalloc<float2>(std::max<int64_t>(P.nRt, 1), S.owned);
  S.desc.Rt32 = dalloc<float2>(std::max<int64_t>(P.nRt, 1), S.owned);
  S.desc.Vt32 = dalloc<float2>(std::max<int64_t>(P.nRt, 1), S.owned);
  S.desc.I4_32 = dalloc<float2>(std::max<int64_t>(P.nI4, 1), S.owned);
  S.desc.W0_32 = dalloc<float2>(std::max<int64_t>(P.nI4, 1), S.owned);
  S.desc.Tbf = dalloc<uint4>(std::max<int64_t>(P.nT, 1), S.owned);
  HIPCHK(hipMemset(S.desc.Tt32, 0, std::max<int64_t>(P.nRt, 1) * sizeof(float2)));
  fill_f32(S);
  HIPCHK(hipMemcpy(S.d_desc, &S.desc, sizeof(SolveDesc), hipMemcpyHostToDevice));
  drop_single(handle);
}

template <typename F>
int guarded(F &&f) {
  std::lock_guard<std::mutex> lk(g_mu);
  try {
    f();
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return 1;
  } catch (...) {
    g_err = "unknown error";
    return 1;
  }
}

void fill_dims(const Plan &P, ccgpu_dims *d) {
  d->nsp = P.nsp; d->nholes = P.A; d->nparticles = P.np;
  d->nchannels = P.nch; d->nph_channels = P.nphc;
  d->max_npp = P.max_npp; d->max_nhh = P.max_nhh;
  d->t2_unique = P.nT; d->t2_ordered = P.t2_ordered; d->vpppp_unique = P.nVpp;
  d->ring_elems = P.nRt;
  d->box = P.orb.box; d->volume = P.orb.volume; d->kf = P.orb.kf;
  double fpp = 0, fhh = 0, fr = 0;
  for (int c = 0; c < P.nch; ++c) {
    double np_ = P.npp[c], nh = P.nhh[c];
    fpp += 8.0 * np_ * np_ * nh;
    fhh += 8.0 * 2.0 * np_ * nh * nh;
  }
  for (int K = 0; K < P.nphc; ++K) {
    double a = P.nph[K], b = P.nhp[K];
    fr += 8.0 * 2.0 * a * b * b;
  }
  d->pp_ladder_flops = fpp; d->hh_ladder_flops = fhh; d->ring_flops = fr;
}

int64_t t_index(const Plan &P, int a, int b, int i, int j, double &sign) {
  sign = 0;
  if (a < P.A || b < P.A || i >= P.A || j >= P.A || a >= P.nsp || b >= P.nsp || i < 0 || j < 0) return -1;
  if (a == b || i == j) return -1;
  PairRef p = P.pp(a, b), h = P.hh(i, j);
  if (!p.sign || !h.sign || p.ch != h.ch) return -1;
  sign = p.sign * h.sign;
  return P.off_t[p.ch] + (int64_t)p.idx * P.nhh[p.ch] + h.idx;
}

// ---------------------------------------------------------------------------
// Warm-start network state: FP32 master parameters with Adam moments, the
// device-resident training set (BF16 features and FP32 targets of every
// amplitude of the solves added so far) and the buffers of one minibatch.
// ---------------------------------------------------------------------------
constexpr float kMlpScale = 0.01f;   // t = t_MBPT2 + kMlpScale * y * e_F / |D|

uint16_t host_bf16(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

struct Mlp {
  hipStream_t stream = nullptr;
  std::vector<void *> owned;
  float *params = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
  double *loss = nullptr;
  int step = 0;
  // training set
  uint16_t *feat = nullptr;
  float *target = nullptr;
  int64_t n = 0, cap = 0;
  // minibatch buffers
  int64_t bcap = 0;
  MlpTrainBufs tb{};
  int64_t *idx = nullptr;
  float *part = nullptr;
  static constexpr int kMaxSplit = 256;
  std::vector<void *> batch_owned;
  // scratch for inference / host-provided data
  int64_t scap = 0;
  uint16_t *sfeat = nullptr;
  float *sy = nullptr, *st = nullptr;
  std::vector<void *> scratch_owned;
  ~Mlp() {
    for (void *p : owned) (void)hipFree(p);
    for (void *p : batch_owned) (void)hipFree(p);
    for (void *p : scratch_owned) (void)hipFree(p);
    if (feat) (void)hipFree(feat);
    if (target) (void)hipFree(target);
    if (stream) (void)hipStreamDestroy(stream);
  }
  void ensure_batch(int64_t B) {
    if (B <= bcap) return;
    HIPCHK(hipStreamSynchronize(stream));
    for (void *p : batch_owned) (void)hipFree(p);
    batch_owned.clear();
    tb.x1t = dalloc<uint16_t>((size_t)16 * B, batch_owned);
    tb.h1t = dalloc<uint16_t>((size_t)64 * B, batch_owned);
    tb.h2t = dalloc<uint16_t>((size_t)64 * B, batch_owned);
    tb.h3t = dalloc<uint16_t>((size_t)64 * B, batch_owned);
    tb.dz1t = dalloc<uint16_t>((size_t)64 * B, batch_owned);
    tb.dz2t = dalloc<uint16_t>((size_t)64 * B, batch_owned);
    tb.dz3t = dalloc<uint16_t>((size_t)64 * B, batch_owned);
    tb.g = dalloc<float>((size_t)B, batch_owned);
    tb.loss_part = dalloc<float>(2048, batch_owned);
    idx = dalloc<int64_t>((size_t)B, batch_owned);
    part = dalloc<float>((size_t)kMaxSplit * kMlpParamCount, batch_owned);
    bcap = B;
  }
  void ensure_scratch(int64_t n_) {
    if (n_ <= scap) return;
    HIPCHK(hipStreamSynchronize(stream));
    for (void *p : scratch_owned) (void)hipFree(p);
    scratch_owned.clear();
    sfeat = dalloc<uint16_t>((size_t)16 * n_, scratch_owned);
    sy = dalloc<float>((size_t)n_, scratch_owned);
    st = dalloc<float>((size_t)n_, scratch_owned);
    scap = n_;
  }
  void grow_set(int64_t need) {
    if (need <= cap) return;
    int64_t nc = std::max<int64_t>(need, cap + cap / 2);
    uint16_t *f2 = nullptr;
    float *t2 = nullptr;
    HIPCHK(hipMalloc((void **)&f2, (size_t)nc * 16 * sizeof(uint16_t)));
    HIPCHK(hipMalloc((void **)&t2, (size_t)nc * sizeof(float)));
    if (n > 0) {
      HIPCHK(hipMemcpyAsync(f2, feat, (size_t)n * 16 * sizeof(uint16_t), hipMemcpyDeviceToDevice, stream));
      HIPCHK(hipMemcpyAsync(t2, target, (size_t)n * sizeof(float), hipMemcpyDeviceToDevice, stream));
    }
    HIPCHK(hipStreamSynchronize(stream));
    if (feat) (void)hipFree(feat);
    if (target) (void)hipFree(target);
    feat = f2;
    target = t2;
    cap = nc;
  }
};

std::map<int, std::unique_ptr<Mlp>> g_mlps;
int g_next_mlp = 1;

Mlp &get_mlp(int h) {
  auto it = g_mlps.find(h);
  if (it == g_mlps.end()) throw Error(fmt("invalid network handle %d", h));
  return *it->second;
}

// kF, e_F and twist/pi of a solve's lattice for the features
void lattice_scalars(const Solve &S, float &kf, float &ef, float tw[3]) {
  const Orbitals &o = S.plan.orb;
  const double hbarc = 197.326968, pmass = (double)939.565f, pi = 3.14159265358979;
  kf = (float)o.kf;
  ef = (float)(hbarc * hbarc * o.kf * o.kf / (2.0 * pmass));
  const double th[3] = {o.kx[0] * o.box - 2.0 * pi * o.nx[0], o.ky[0] * o.box - 2.0 * pi * o.ny[0],
                        o.kz[0] * o.box - 2.0 * pi * o.nz[0]};
  for (int k = 0; k < 3; ++k) tw[k] = (float)(th[k] / pi);
}

}  // namespace

extern "C" {

int ccgpu_last_error(char *buf, int len) {
  if (buf && len > 0) {
    strncpy(buf, g_err.c_str(), len - 1);
    buf[len - 1] = 0;
  }
  return (int)g_err.size();
}

int ccgpu_plan(const ccgpu_lattice_params *p, ccgpu_dims *dims) {
  return guarded([&] {
    if (!p || !dims) throw Error("null argument");
    Orbitals o = build_orbitals(p->A, p->matter, p->rho, p->nmax, p->twist);
    Plan P;
    build_plan(o, P);
    fill_dims(P, dims);
  });
}

int ccgpu_create(const ccgpu_lattice_params *p, int *handle) {
  return guarded([&] {
    if (!p || !handle) throw Error("null argument");
    auto S = std::make_unique<Solve>();
    Orbitals o = build_orbitals(p->A, p->matter, p->rho, p->nmax, p->twist);
    build_plan(o, S->plan);
    build_device(*S);
    int h = g_next++;
    g_solves[h] = std::move(S);
    *handle = h;
  });
}

int ccgpu_create_from_orbitals(int nsp, int nholes, const int *nx, const int *ny, const int *nz, const int *sz,
                               const int *tz, const double *kx, const double *ky, const double *kz,
                               const double *e, double box, double volume, int *handle) {
  return guarded([&] {
    if (nsp <= 0 || nholes <= 0 || nholes >= nsp) throw Error("invalid orbital counts");
    Orbitals o;
    o.nsp = nsp; o.A = nholes;
    o.nx.assign(nx, nx + nsp); o.ny.assign(ny, ny + nsp); o.nz.assign(nz, nz + nsp);
    o.sz.assign(sz, sz + nsp); o.tz.assign(tz, tz + nsp);
    o.kx.assign(kx, kx + nsp); o.ky.assign(ky, ky + nsp); o.kz.assign(kz, kz + nsp);
    o.e.assign(e, e + nsp);
    o.box = box; o.volume = volume;
    int nm = 0;
    for (int x = 0; x < nsp; ++x) nm = std::max({nm, std::abs(o.nx[x]), std::abs(o.ny[x]), std::abs(o.nz[x])});
    o.nmax = nm;
    for (int x = 0; x < nsp; ++x)
      if (std::abs(o.sz[x]) != 1 || std::abs(o.tz[x]) != 1) throw Error("sz and tz must be +-1");
    auto S = std::make_unique<Solve>();
    build_plan(o, S->plan);
    build_device(*S);
    int h = g_next++;
    g_solves[h] = std::move(S);
    *handle = h;
  });
}

int ccgpu_destroy(int handle) {
  return guarded([&] {
    get(handle);
    drop_single(handle);
    g_solves.erase(handle);
  });
}

int ccgpu_get_dims(int handle, ccgpu_dims *dims) {
  return guarded([&] { fill_dims(get(handle).plan, dims); });
}

int ccgpu_get_orbitals(int handle, int *nx, int *ny, int *nz, int *sz, int *tz, double *e) {
  return guarded([&] {
    const Orbitals &o = get(handle).plan.orb;
    for (int x = 0; x < o.nsp; ++x) {
      if (nx) nx[x] = o.nx[x];
      if (ny) ny[x] = o.ny[x];
      if (nz) nz[x] = o.nz[x];
      if (sz) sz[x] = o.sz[x];
      if (tz) tz[x] = o.tz[x];
      if (e) e[x] = o.e[x];
    }
  });
}

}  // extern "C"
namespace {
// Device tables of the chiral NN for this lattice: radial functions over
// m = |dn|^2 <= 3 (4 nmax)^2 and the per-side factors over j <= 3 (2 nmax)^2.
void make_chiral_tables(Solve &S) {
  const Orbitals &o = S.plan.orb;
  int nm = 0;
  for (int x = 0; x < o.nsp; ++x) nm = std::max({nm, std::abs(o.nx[x]), std::abs(o.ny[x]), std::abs(o.nz[x])});
  const double hbarc = 197.326968, pi = 3.14159265358979;
  const int mmax = 3 * (4 * nm) * (4 * nm), jmax = 3 * (2 * nm) * (2 * nm);
  const double qunit = pi * hbarc / o.box;
  const ChiralModel &c = S.chiral;
  std::vector<double> rad = chiral_radial_table(c, qunit, mmax);
  std::vector<double> side = chiral_side_table(c, qunit, jmax);
  ChiralDev &d = S.chiral_dev;
  d.rad = dupload(rad, S.owned);
  d.side = dupload(side, S.owned);
  d.jmax = jmax;
  d.order = c.order;
  d.qunit = qunit;
  d.scale = std::pow(hbarc, 3) / o.volume;
  for (int i = 0; i < 3; ++i) { d.cs[i] = c.cs[i]; d.ct[i] = c.ct[i]; }
  for (int i = 0; i < 7; ++i) d.cnlo[i] = c.cnlo[i];
  for (int i = 0; i < 2; ++i) { d.cg0[i] = c.cg0[i]; d.cg1[i] = c.cg1[i]; }
}
}  // namespace
extern "C" {

int ccgpu_set_chiral(int handle, int order, double cutoff, const double *ci) {
  return guarded([&] {
    Solve &S = get(handle);
    S.chiral = chiral_model(order, cutoff, ci);
  });
}

int ccgpu_set_3nf(int handle, int approx, double cE, double cD, double cutoff, int nexp) {
  return guarded([&] {
    Solve &S = get(handle);
    if (approx < -1 || approx > 3) throw Error(fmt("3NF approximation %d: -1 off, 0, 1, 2 or 3", approx));
    if (cutoff <= 0.0 || nexp < 1) throw Error("3NF regulator needs cutoff > 0 and nexp >= 1");
    S.tnf = approx;
    S.tnf_set = true;
    S.have_w3 = false;
    S.have_ccdt1 = false;
    S.tnf_ce = cE;
    S.tnf_cd = cD;
    S.tnf_cutoff = cutoff;
    S.tnf_nexp = nexp;
  });
}

int ccgpu_generate_interaction(int handle, int kind) {
  return guarded([&] {
    Solve &S = get(handle);
    if (kind != CCGPU_INT_MINNESOTA && kind != CCGPU_INT_CHIRAL) throw Error(fmt("unknown interaction kind %d", kind));
    const Plan &P = S.plan;
    int max_nph = 0, max_nhp = 0;
    for (int K = 0; K < P.nphc; ++K) { max_nph = std::max(max_nph, P.nph[K]); max_nhp = std::max(max_nhp, P.nhp[K]); }
    if (kind != CCGPU_INT_CHIRAL && S.tnf >= 0)
      throw Error("the three-nucleon force belongs to the chiral interaction (CCGPU_INT_CHIRAL)");
    if (kind == CCGPU_INT_CHIRAL) {
      make_chiral_tables(S);
      launch_fill_blocks_chiral(S.desc, S.chiral_dev, P.nch, P.nphc, P.max_npp, P.max_nhh, max_nhp, S.stream);
      if (S.tnf >= 0) {
        Chiral3Dev &c = S.chiral3;
        c = chiral3_constants(S.chiral, S.tnf_ce, S.tnf_cd, S.tnf_cutoff, S.tnf_nexp);
        c.kx = S.desc.kx; c.ky = S.desc.ky; c.kz = S.desc.kz;
        c.nx = S.desc.nx; c.ny = S.desc.ny; c.nz = S.desc.nz; c.sz = S.desc.sz; c.tz = S.desc.tz;
        c.scale = std::pow(197.326968, 6) / (P.orb.volume * P.orb.volume);
        // parity images for the V_pppp fill: with no twist n -> -n maps the
        // lattice onto itself with k -> -k bit for bit and the closed Fermi
        // sea onto itself (CCGPU_3NF_PARITY=0 evaluates every element)
        S.desc.pimg = nullptr;
        if (!(getenv("CCGPU_3NF_PARITY") && atoi(getenv("CCGPU_3NF_PARITY")) == 0)) {
          const Orbitals &o = P.orb;
          std::map<std::tuple<int, int, int, int, int>, int> at;
          for (int x = 0; x < o.nsp; ++x) at[std::make_tuple(o.nx[x], o.ny[x], o.nz[x], o.sz[x], o.tz[x])] = x;
          std::vector<int> img(o.nsp, -1);
          bool ok = true;
          for (int x = 0; x < o.nsp && ok; ++x) {
            auto it = at.find(std::make_tuple(-o.nx[x], -o.ny[x], -o.nz[x], o.sz[x], o.tz[x]));
            if (it == at.end()) {
              ok = false;
              break;
            }
            const int y = it->second;
            ok = (x < P.A) == (y < P.A) && o.kx[y] == -o.kx[x] && o.ky[y] == -o.ky[x] && o.kz[y] == -o.kz[x];
            img[x] = y;
          }
          if (ok) S.desc.pimg = dupload(img, S.owned);
        }
        check_launch();
        launch_3nf(S.desc, c, S.tnf, P.nch, P.nphc, P.max_npp, P.max_nhh, max_nhp, S.stream);
      }
    } else {
      launch_fill_blocks(S.desc, P.orb.volume, P.nch, P.nphc, P.max_npp, P.max_nhh, max_nhp, S.stream);
    }
    check_launch();
    launch_derive(S.desc, P.nch, P.nphc, P.max_npp, P.max_nhh, max_nph, max_nhp, S.stream);
    check_launch();
    HIPCHK(hipMemcpyAsync(&S.e0, S.desc.e0, sizeof(double2), hipMemcpyDeviceToHost, S.stream));
    HIPCHK(hipStreamSynchronize(S.stream));
    S.have_v = S.have_fock = S.derived = true;
    S.native_v = true;
    ++S.vgen;
    S.vkind = kind;
    S.have_t3 = S.have_w3 = S.have_ccdt1 = false;
    // V changed: an FP32 copy and cached work lists of this handle are stale
    if (S.desc.Vpppp32) fill_f32(S);
  });
}

int ccgpu_vnn_elements(int handle, int n, const int *p, const int *q, const int *r, const int *s, double *v) {
  return guarded([&] {
    Solve &S = get(handle);
    if (S.vkind < 0) throw Error("no generated interaction: call ccgpu_generate_interaction first");
    if (n < 0) throw Error("negative element count");
    if (n == 0) return;
    const int nsp = S.plan.nsp;
    std::vector<int4> h(n);
    for (int x = 0; x < n; ++x) {
      for (int y : {p[x], q[x], r[x], s[x]})
        if (y < 0 || y >= nsp) throw Error(fmt("orbital index %d out of range", y));
      h[x] = int4{p[x], q[x], r[x], s[x]};
    }
    int4 *d_in = nullptr;
    double2 *d_out = nullptr;
    HIPCHK(hipMalloc(&d_in, n * sizeof(int4)));
    HIPCHK(hipMalloc(&d_out, n * sizeof(double2)));
    HIPCHK(hipMemcpy(d_in, h.data(), n * sizeof(int4), hipMemcpyHostToDevice));
    launch_elements(S.desc, S.plan.orb.volume, S.vkind == CCGPU_INT_CHIRAL ? &S.chiral_dev : nullptr, d_in, n,
                    d_out, S.stream);
    const hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
      (void)hipStreamSynchronize(S.stream);
      (void)hipMemcpy(v, d_out, n * sizeof(double2), hipMemcpyDeviceToHost);
    }
    (void)hipFree(d_in);
    (void)hipFree(d_out);
    if (e != hipSuccess) throw Error(std::string("kernel launch: ") + hipGetErrorString(e));
  });
}

// Host staging of reference-layout blocks (drop-in path).
int ccgpu_set_block(int handle, int kind, int nbra, const int *bra_p, const int *bra_q, int nket,
                    const int *ket_r, const int *ket_s, const double *val, int ld) {
  return guarded([&] {
    Solve &S = get(handle);
    const Plan &P = S.plan;
    if (ld < nbra) throw Error("ld < nbra");
    auto V = [&](int r, int c) {
      const double *p = val + 2 * ((size_t)c * ld + r);
      return double2{p[0], p[1]};
    };
    auto chk = [&](int x) {
      if (x < 1 || x > P.nsp) throw Error(fmt("orbital index %d out of range", x));
      return x - 1;
    };
    std::vector<std::pair<int64_t, double2>> up;
    double2 *dst = nullptr;
    switch (kind) {
      case CCGPU_BLOCK_PPPP:
      case CCGPU_BLOCK_HHPP:
      case CCGPU_BLOCK_HHHH: {
        for (int r = 0; r < nbra; ++r) {
          int p = chk(bra_p[r]), q = chk(bra_q[r]);
          if (p >= q) continue;
          for (int c = 0; c < nket; ++c) {
            int x = chk(ket_r[c]), y = chk(ket_s[c]);
            if (x >= y) continue;
            PairRef br, kr;
            if (kind == CCGPU_BLOCK_PPPP) {
              if (p < P.A || x < P.A) throw Error("pppp block with hole indices");
              br = P.pp(p, q); kr = P.pp(x, y);
              if (br.sign != 1 || kr.sign != 1 || br.ch != kr.ch) continue;
              up.push_back({P.off_vpp[br.ch] + (int64_t)br.idx * P.npp[br.ch] + kr.idx, V(r, c)});
              dst = S.desc.Vpppp;
            } else if (kind == CCGPU_BLOCK_HHPP) {
              if (p >= P.A || x < P.A) throw Error("hhpp block with wrong index classes");
              br = P.hh(p, q); kr = P.pp(x, y);
              if (br.sign != 1 || kr.sign != 1 || br.ch != kr.ch) continue;
              up.push_back({P.off_t[br.ch] + (int64_t)br.idx * P.npp[br.ch] + kr.idx, V(r, c)});
              dst = S.desc.Vhhpp;
            } else {
              if (p >= P.A || x >= P.A) throw Error("hhhh block with particle indices");
              br = P.hh(p, q); kr = P.hh(x, y);
              if (br.sign != 1 || kr.sign != 1 || br.ch != kr.ch) continue;
              up.push_back({P.off_vhh[br.ch] + (int64_t)br.idx * P.nhh[br.ch] + kr.idx, V(r, c)});
              dst = S.desc.Vhhhh;
            }
          }
        }
        break;
      }
      case CCGPU_BLOCK_HPHP: {
        for (int r = 0; r < nbra; ++r) {
          int i = chk(bra_p[r]), b = chk(bra_q[r]);
          for (int c = 0; c < nket; ++c) {
            int j = chk(ket_r[c]), a = chk(ket_s[c]);
            uint64_t key = (((uint64_t)i * P.nsp + b) * P.nsp + j) * P.nsp + a;
            double2 v = V(r, c);
            S.hphp[key] = std::complex<double>(v.x, v.y);
          }
        }
        break;
      }
      case CCGPU_BLOCK_HHHP:
      case CCGPU_BLOCK_HPPP: {
        // HHHP: rows hole pairs (j,k), columns (hole l, particle c), element
        // <jk||lc> (vnn_hhhp, ccm_iter.f90:903-920).  HPPP: rows (hole k,
        // particle d), columns particle pairs (c,b), element <cb||kd>
        // (vnn_hppp, ccm_iter.f90:1621-1711).  Stored for (T) keyed by the
        // element's own index order, with both orders of the like pair.
        const uint64_t n = (uint64_t)P.nsp;
        auto key = [&](int p, int q, int r, int t) { return (((uint64_t)p * n + q) * n + r) * n + t; };
        for (int r = 0; r < nbra; ++r) {
          int p = chk(bra_p[r]), q = chk(bra_q[r]);
          for (int c = 0; c < nket; ++c) {
            int x = chk(ket_r[c]), y = chk(ket_s[c]);
            const double2 v = V(r, c);
            const std::complex<double> z(v.x, v.y);
            if (kind == CCGPU_BLOCK_HHHP) {
              if (p >= P.A || q >= P.A || x >= P.A || y < P.A) throw Error("hhhp block with wrong index classes");
              S.t3v[key(p, q, x, y)] = z;
              S.t3v.emplace(key(q, p, x, y), -z);
            } else {
              if (p >= P.A || q < P.A || x < P.A || y < P.A) throw Error("hppp block with wrong index classes");
              S.t3v[key(x, y, p, q)] = z;
              S.t3v.emplace(key(y, x, p, q), -z);
            }
          }
        }
        S.have_t3 = S.have_w3 = S.have_ccdt1 = false;
        break;
      }
      default:
        throw Error(fmt("block kind %d not accepted by ccgpu_set_block", kind));
    }
    if (dst && !up.empty()) {
      // only the canonical (index, value) pairs cross PCIe; a kernel scatters
      // them into the device array
      std::vector<int64_t> idx(up.size());
      std::vector<double2> val(up.size());
      for (size_t e = 0; e < up.size(); ++e) {
        idx[e] = up[e].first;
        val[e] = up[e].second;
      }
      std::vector<void *> tmp;
      int64_t *d_idx = dupload(idx, tmp);
      double2 *d_val = dupload(val, tmp);
      launch_scatter_values(dst, d_idx, d_val, (int64_t)up.size(), S.stream);
      hipError_t e = hipGetLastError();
      if (e == hipSuccess) e = hipStreamSynchronize(S.stream);
      for (void *p : tmp) (void)hipFree(p);
      if (e != hipSuccess) throw Error(fmt("block scatter failed: %s", hipGetErrorString(e)));
    }
    S.have_v = true;
    S.native_v = false;
    ++S.vgen;
    S.vkind = -1;
    S.derived = false;
  });
}

int ccgpu_set_fock(int handle, const double *fock, double e0_re, double e0_im) {
  return guarded([&] {
    Solve &S = get(handle);
    const Plan &P = S.plan;
    std::vector<double2> fd(P.nsp), fp(P.nsp);
    for (int p = 0; p < P.nsp; ++p) {
      const double *x = fock + 2 * ((size_t)p * P.nsp + p);
      fd[p] = double2{x[0], x[1]};
      int q = P.partner[p];
      if (q >= 0) {
        const double *y = fock + 2 * ((size_t)q * P.nsp + p);  // f(p,q), column-major
        fp[p] = double2{y[0], y[1]};
      } else {
        fp[p] = double2{0, 0};
      }
    }
    HIPCHK(hipMemcpy(S.desc.fdiag, fd.data(), P.nsp * sizeof(double2), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(S.desc.fpart, fp.data(), P.nsp * sizeof(double2), hipMemcpyHostToDevice));
    S.e0 = double2{e0_re, e0_im};
    HIPCHK(hipMemcpy(S.desc.e0, &S.e0, sizeof(double2), hipMemcpyHostToDevice));
    S.have_fock = true;
  });
}

int ccgpu_finalize_setup(int handle) {
  return guarded([&] {
    Solve &S = get(handle);
    const Plan &P = S.plan;
    if (!S.have_v || !S.have_fock) throw Error("set V blocks and the Fock matrix first");
    // W0 from the staged hphp records: W0((i,a),(j,b)) = -<ib||ja>
    if (!S.hphp.empty()) {
      std::vector<double2> w(P.nI4);
      for (int K = 0; K < P.nphc; ++K)
        for (int r = 0; r < P.nhp[K]; ++r)
          for (int s = 0; s < P.nhp[K]; ++s) {
            int i = P.hp_j[P.off_hp[K] + r], a = P.hp_b[P.off_hp[K] + r];
            int j = P.hp_j[P.off_hp[K] + s], b = P.hp_b[P.off_hp[K] + s];
            uint64_t key = (((uint64_t)i * P.nsp + b) * P.nsp + j) * P.nsp + a;
            auto it = S.hphp.find(key);
            std::complex<double> v = it == S.hphp.end() ? 0.0 : it->second;
            w[P.off_i4[K] + (int64_t)r * P.nhp[K] + s] = double2{-v.real(), -v.imag()};
          }
      HIPCHK(hipMemcpy(S.desc.W0, w.data(), P.nI4 * sizeof(double2), hipMemcpyHostToDevice));
    }
    int max_nph = 0, max_nhp = 0;
    for (int K = 0; K < P.nphc; ++K) { max_nph = std::max(max_nph, P.nph[K]); max_nhp = std::max(max_nhp, P.nhp[K]); }
    launch_derive(S.desc, P.nch, P.nphc, P.max_npp, P.max_nhh, max_nph, max_nhp, S.stream);
    check_launch();
    HIPCHK(hipStreamSynchronize(S.stream));
    // the FP32 copy of V_pppp follows the blocks handed over since it was made
    if (S.desc.Vpppp32) fill_f32(S);
    S.derived = true;
  });
}

int ccgpu_get_e0(int handle, double *e0) {
  return guarded([&] {
    Solve &S = get(handle);
    e0[0] = S.e0.x;
    e0[1] = S.e0.y;
  });
}

int ccgpu_get_fock_diag(int handle, double *f) {
  return guarded([&] {
    Solve &S = get(handle);
    HIPCHK(hipMemcpy(f, S.desc.fdiag, S.plan.nsp * sizeof(double2), hipMemcpyDeviceToHost));
  });
}

int ccgpu_mbpt2_start(int handle, double *ecorr) {
  return guarded([&] {
    Solve &S = get(handle);
    ensure_ready(S);
    Batch &B = single(handle);
    launch_mbpt2(B.d_descs, 1, S.plan.nT, B.stream);
    check_launch();
    auto e = energies(B, false);
    S.have_t = true;
    S.nstep = 0;
    S.count = 0;
    if (ecorr) { ecorr[0] = e[0].real(); ecorr[1] = e[0].imag(); }
  });
}

int ccgpu_residual(int handle, int precision) {
  return guarded([&] {
    Solve &S = get(handle);
    ensure_ready(S);
    if (!S.have_t) throw Error("no amplitudes: call ccgpu_mbpt2_start or ccgpu_set_t2 first");
    if (!known_precision(precision)) throw Error(fmt("unknown precision %d", precision));
    if (precision != CCGPU_PREC_FP64) ensure_f32(S, handle);
    Batch &B = single(handle);
    prepare_lists(B, precision);
    residual_launches(B, precision, false);
    check_launch();
    HIPCHK(hipStreamSynchronize(B.stream));
  });
}

int ccgpu_update(int handle, int *count, int *nstep, int subspace, int diis_step) {
  return guarded([&] {
    Solve &S = get(handle);
    ensure_ready(S);
    Batch &B = single(handle);
    if (S.depth < subspace) {
      alloc_history(S, subspace, false);
      B.h_descs[0].hist = S.desc.hist;
      HIPCHK(hipMemcpy(B.d_descs, B.h_descs.data(), sizeof(SolveDesc), hipMemcpyHostToDevice));
    }
    if ((int)S.ring.size() != subspace) {
      S.ring.resize(subspace);
      for (int k = 0; k < subspace; ++k) S.ring[k] = k;
    }
    int c = *count, n = *nstep;
    update_launches(B, c, n, subspace, diis_step, false);
    check_launch();
    HIPCHK(hipStreamSynchronize(B.stream));
    *count = c;
    *nstep = n;
  });
}

int ccgpu_energy(int handle, double *ecorr) {
  return guarded([&] {
    Solve &S = get(handle);
    ensure_ready(S);
    Batch &B = single(handle);
    auto e = energies(B, false);
    ecorr[0] = e[0].real();
    ecorr[1] = e[0].imag();
  });
}

int ccgpu_solve(int handle, const ccgpu_solve_opts *opts, ccgpu_solve_result *res) {
  return guarded([&] {
    Solve &S = get(handle);
    ensure_ready(S);
    if (!S.have_t) throw Error("no amplitudes: call ccgpu_mbpt2_start or ccgpu_set_t2 first");
    if (opts->precision != CCGPU_PREC_FP64 || opts->tail != CCGPU_PREC_FP64) ensure_f32(S, handle);
    Batch &B = single(handle);
    std::vector<ccgpu_solve_result> r;
    run_batch(B, *opts, r);
    *res = r[0];
  });
}

int ccgpu_solve_energies(int handle, double *e, int max, int *count) {
  return guarded([&] {
    Solve &S = get(handle);
    const int n = (int)S.ehist.size();
    *count = n;
    for (int k = 0; k < std::min(n, max); ++k) {
      e[2 * k] = S.ehist[k].real();
      e[2 * k + 1] = S.ehist[k].imag();
    }
  });
}

int ccgpu_batch_solve(int n, const int *handles, const ccgpu_solve_opts *opts, ccgpu_solve_result *results) {
  return guarded([&] {
    if (n <= 0) throw Error("empty batch");
    std::vector<Solve *> ss;
    for (int s = 0; s < n; ++s) {
      Solve &S = get(handles[s]);
      ensure_ready(S);
      if (!S.have_t) throw Error(fmt("handle %d has no amplitudes", handles[s]));
      if (opts->precision != CCGPU_PREC_FP64 || opts->tail != CCGPU_PREC_FP64) ensure_f32(S, handles[s]);
      ss.push_back(&S);
    }
    // the work lists, tiles and captured graphs of a handle set are kept for
    // the next call on the same handles (dropped when one of them changes)
    std::vector<int> key(handles, handles + n);
    auto it = g_batches.find(key);
    if (it == g_batches.end()) {
      if (g_batches.size() >= 8) g_batches.clear();
      it = g_batches.emplace(key, make_batch(ss, ss[0]->stream)).first;
    }
    std::vector<ccgpu_solve_result> r;
    run_batch(*it->second, *opts, r);
    for (int s = 0; s < n; ++s) results[s] = r[s];
  });
}

int ccgpu_get_t2(int handle, int n, const int *a, const int *b, const int *i, const int *j, double *t) {
  return guarded([&] {
    Solve &S = get(handle);
    std::vector<double2> h(S.plan.nT);
    HIPCHK(hipMemcpy(h.data(), S.desc.T, S.plan.nT * sizeof(double2), hipMemcpyDeviceToHost));
    for (int x = 0; x < n; ++x) {
      double sg;
      int64_t u = t_index(S.plan, a[x], b[x], i[x], j[x], sg);
      t[2 * x] = u < 0 ? 0.0 : sg * h[u].x;
      t[2 * x + 1] = u < 0 ? 0.0 : sg * h[u].y;
    }
  });
}

int ccgpu_set_t2(int handle, int n, const int *a, const int *b, const int *i, const int *j, const double *t) {
  return guarded([&] {
    Solve &S = get(handle);
    std::vector<double2> h(S.plan.nT);
    HIPCHK(hipMemcpy(h.data(), S.desc.T, S.plan.nT * sizeof(double2), hipMemcpyDeviceToHost));
    for (int x = 0; x < n; ++x) {
      double sg;
      int64_t u = t_index(S.plan, a[x], b[x], i[x], j[x], sg);
      if (u < 0) continue;
      h[u] = double2{sg * t[2 * x], sg * t[2 * x + 1]};
    }
    HIPCHK(hipMemcpy(S.desc.T, h.data(), S.plan.nT * sizeof(double2), hipMemcpyHostToDevice));
    S.have_t = true;
  });
}

int ccgpu_get_t2_unique(int handle, double *t) {
  return guarded([&] {
    Solve &S = get(handle);
    HIPCHK(hipMemcpy(t, S.desc.T, S.plan.nT * sizeof(double2), hipMemcpyDeviceToHost));
  });
}

int ccgpu_set_t2_unique(int handle, const double *t) {
  return guarded([&] {
    Solve &S = get(handle);
    HIPCHK(hipMemcpy(S.desc.T, t, S.plan.nT * sizeof(double2), hipMemcpyHostToDevice));
    S.have_t = true;
  });
}

int ccgpu_get_t2_labels(int handle, int *abij) {
  return guarded([&] {
    const Plan &P = get(handle).plan;
    for (int c = 0; c < P.nch; ++c)
      for (int r = 0; r < P.npp[c]; ++r)
        for (int s = 0; s < P.nhh[c]; ++s) {
          int64_t u = P.off_t[c] + (int64_t)r * P.nhh[c] + s;
          abij[4 * u + 0] = P.pp_a[P.off_pp[c] + r];
          abij[4 * u + 1] = P.pp_b[P.off_pp[c] + r];
          abij[4 * u + 2] = P.hh_i[P.off_hh[c] + s];
          abij[4 * u + 3] = P.hh_j[P.off_hh[c] + s];
        }
  });
}

}  // extern "C"
namespace {
// (K3,Tz3) hole-triple channels, orbital lookup and particle-pair list of
// setup_t3_channel_structures (ccm_t3full_channels.f90:4-284), built once.
void build_triples(Solve &S) {
  const Plan &P = S.plan;
  const Orbitals &o = P.orb;
  int nh = 0;
  for (int i = 0; i < P.A; ++i) nh = std::max({nh, std::abs(o.nx[i]), std::abs(o.ny[i]), std::abs(o.nz[i])});
  TriplesDesc &td = S.t3;
  td.k0 = -3 * nh;
  td.nk = 6 * nh + 1;
  td.tz0 = -3;
  td.ntz = 7;
  int nm = 0;   // lattice extent from the orbitals themselves (drop-in tables too)
  for (int x = 0; x < P.nsp; ++x) nm = std::max({nm, std::abs(o.nx[x]), std::abs(o.ny[x]), std::abs(o.nz[x])});
  td.nmax = nm;
  std::vector<int> chan((size_t)td.ntz * td.nk * td.nk * td.nk, -1);
  std::map<std::tuple<int, int, int, int>, std::vector<int>> trip;
  for (int i = 0; i < P.A; ++i)
    for (int j = i + 1; j < P.A; ++j)
      for (int k = j + 1; k < P.A; ++k) {
        auto key = std::make_tuple(o.tz[i] + o.tz[j] + o.tz[k], o.nx[i] + o.nx[j] + o.nx[k],
                                   o.ny[i] + o.ny[j] + o.ny[k], o.nz[i] + o.nz[j] + o.nz[k]);
        auto &v = trip[key];
        v.push_back(i); v.push_back(j); v.push_back(k);
      }
  std::vector<int> hptr = {0}, hijk;
  int ch = 0;
  for (auto &kv : trip) {
    int tz, kx, ky, kz;
    std::tie(tz, kx, ky, kz) = kv.first;
    size_t idx = (((size_t)(tz - td.tz0) * td.nk + (kx - td.k0)) * td.nk + (ky - td.k0)) * td.nk + (kz - td.k0);
    chan[idx] = ch++;
    hijk.insert(hijk.end(), kv.second.begin(), kv.second.end());
    hptr.push_back((int)(hijk.size() / 3));
  }
  const int n = 2 * nm + 1;
  std::vector<int> orbitof((size_t)n * n * n * 4, -1);
  for (int x = 0; x < P.nsp; ++x) {
    size_t idx = (((size_t)(o.nx[x] + nm) * n + (o.ny[x] + nm)) * n + (o.nz[x] + nm)) * 4 +
                 ((o.sz[x] + 1) / 2) * 2 + (o.tz[x] + 1) / 2;
    orbitof[idx] = x;
  }
  std::vector<int> pa, pb;
  for (int a = P.A; a < P.nsp; ++a)
    for (int b = a + 1; b < P.nsp; ++b) { pa.push_back(a); pb.push_back(b); }
  td.chan = dupload(chan, S.owned);
  td.hptr = dupload(hptr, S.owned);
  td.hijk = dupload(hijk, S.owned);
  td.orbitof = dupload(orbitof, S.owned);
  td.pair_a = dupload(pa, S.owned);
  td.pair_b = dupload(pb, S.owned);
  td.npairs = (int64_t)pa.size();
  td.vkeys = nullptr;
  td.vvals = nullptr;
  td.vmask = 0;
  td.mtab = nullptr;
  td.mmax = 0;
  td.chiral = 0;
  td.tnf = 0;
  if (S.native_v && S.vkind == CCGPU_INT_CHIRAL) {
    td.chiral = 1;
    td.ch = S.chiral_dev;
    td.tnf = S.tnf >= 2 ? 1 : 0;
    td.c3 = S.chiral3;
  } else if (S.native_v) {
    // radial Minnesota factors over the integer momentum transfer (the
    // constants of the generator, float32 literals as in the reference)
    const double pi = 3.14159265358979;
    const double v0[3] = {200.0, -178.0, -(double)91.85f};
    const double kap[3] = {(double)1.487f, (double)0.639f, (double)0.465f};
    const int dmax = 4 * nm;
    td.mmax = 3 * dmax * dmax;
    const double f = pi / o.box;
    std::vector<double> tab((size_t)3 * (td.mmax + 1));
    for (int m = 0; m <= td.mmax; ++m) {
      const double q2 = f * f * (double)m;
      for (int g = 0; g < 3; ++g)
        tab[3 * m + g] = v0[g] * std::pow(pi, 1.5) * std::exp(-q2 / (4.0 * kap[g])) / std::pow(kap[g], 1.5);
    }
    td.mtab = dupload(tab, S.owned);
  }
  if (!S.native_v) {
    if (S.t3v.empty())
      throw Error("(T) of a handed-over interaction needs its HHHP and HPPP blocks (ccgpu_set_block)");
    uint64_t cap = 1;
    while (cap < 2 * S.t3v.size()) cap <<= 1;
    std::vector<unsigned long long> keys(cap, ~0ull);
    std::vector<double2> vals(cap, double2{0.0, 0.0});
    for (auto &kv : S.t3v) {
      // the probe start of the device lookup (murmur3 finalizer, low bits)
      uint64_t h = kv.first;
      h ^= h >> 33;
      h *= 0xff51afd7ed558ccdull;
      h ^= h >> 33;
      h &= cap - 1;
      while (keys[h] != ~0ull) h = (h + 1) & (cap - 1);
      keys[h] = kv.first;
      vals[h] = double2{kv.second.real(), kv.second.imag()};
    }
    td.vkeys = dupload(keys, S.owned);
    td.vvals = dupload(vals, S.owned);
    td.vmask = cap - 1;
  }
  // parity: with no twist, n -> -n maps the lattice onto itself with k -> -k
  // bit for bit and the closed Fermi sea onto itself; the built-in elements
  // are even under it (CCGPU_T3_PARITY=0 evaluates every element)
  td.parity = 0;
  if (S.native_v && !(getenv("CCGPU_T3_PARITY") && atoi(getenv("CCGPU_T3_PARITY")) == 0)) {
    bool ok = true;
    for (int x = 0; x < P.nsp && ok; ++x) {
      const size_t idx = (((size_t)(nm - o.nx[x]) * n + (nm - o.ny[x])) * n + (nm - o.nz[x])) * 4 +
                         ((o.sz[x] + 1) / 2) * 2 + (o.tz[x] + 1) / 2;
      const int y = orbitof[idx];
      ok = y >= 0 && (x < P.A) == (y < P.A) && o.kx[y] == -o.kx[x] && o.ky[y] == -o.ky[x] && o.kz[y] == -o.kz[x];
    }
    td.parity = ok ? 1 : 0;
  }
  // dense <cb||kd> and <jk||lc> tables, evaluated once per interaction, when
  // they fit in CCGPU_T3_TABLE_GB (default 16 GB); otherwise the kernel
  // evaluates the elements in place
  td.vhppp = nullptr;
  td.vhhhp = nullptr;
  {
    const double lim = (getenv("CCGPU_T3_TABLE_GB") ? atof(getenv("CCGPU_T3_TABLE_GB")) : 16.0) * 1e9;
    const int64_t nh = (int64_t)P.np * P.np * P.A * 2, nl = (int64_t)P.A * P.A * P.np * 2;
    if ((double)(nh + nl) * sizeof(double2) <= lim && nh > 0 && nl > 0) {
      double2 *hppp = dalloc<double2>(nh, S.owned), *hhhp = dalloc<double2>(nl, S.owned);
      launch_t3_tables(S.desc, td, P.orb.volume, hppp, hhhp, S.stream);
      check_launch();
      HIPCHK(hipStreamSynchronize(S.stream));
      td.vhppp = hppp;
      td.vhhhp = hhhp;
    }
    if (td.tnf && !td.vhppp)
      throw Error("(T) with the three-nucleon force needs its element tables: raise CCGPU_T3_TABLE_GB");
  }
  S.t3_partial = dalloc<double2>(8192, S.owned);
  S.t3_out = dalloc<double2>(1, S.owned);
  S.t3_count = dalloc<unsigned long long>(8193, S.owned);
  S.have_t3 = true;
}

// One sum over the triples (modes of k_triples_w): returns it and counts the
// elements.
double2 triples_sum(Solve &S, int mode) {
  if (!S.have_t3) build_triples(S);
  if (mode != 0 && !S.have_w3) {
    if (!(S.native_v && S.vkind == CCGPU_INT_CHIRAL))
      throw Error("the three-nucleon element W3 needs the built-in chiral interaction");
    if (!S.tnf_set) throw Error("the three-nucleon element W3 needs its constants: call ccgpu_set_3nf");
    const Plan &P = S.plan;
    Chiral3Dev c = chiral3_constants(S.chiral, S.tnf_ce, S.tnf_cd, S.tnf_cutoff, S.tnf_nexp);
    c.kx = S.desc.kx; c.ky = S.desc.ky; c.kz = S.desc.kz;
    c.nx = S.desc.nx; c.ny = S.desc.ny; c.nz = S.desc.nz; c.sz = S.desc.sz; c.tz = S.desc.tz;
    c.scale = std::pow(197.326968, 6) / (P.orb.volume * P.orb.volume);
    S.t3.c3w = c;
    S.have_w3 = true;
  }
  if (mode != 2 && !(S.t3.vhppp && S.t3.vhhhp) && mode != 0)
    throw Error("CCDT-1 and the three-body (T) need the element tables: raise CCGPU_T3_TABLE_GB");
  int nb = (int)std::min<int64_t>(8192, std::max<int64_t>(1, S.t3.npairs));
  // the staged (T) kernel: the F terms' amplitude table (np A^2 x 2 spins),
  // made once per handle when the element tables exist and it fits with them
  // in CCGPU_T3_TABLE_GB, refilled from the current T for every sum
  const Plan &P = S.plan;
  if (mode == 0 && S.t3.vhppp && S.t3.vhhhp && !S.t3_td) {
    const double lim = (getenv("CCGPU_T3_TABLE_GB") ? atof(getenv("CCGPU_T3_TABLE_GB")) : 16.0) * 1e9;
    const int64_t ntab = (int64_t)P.np * P.np * P.A * 2 + (int64_t)P.A * P.A * P.np * 2;
    const int64_t ntd = (int64_t)P.np * P.A * P.A * 2;
    if ((double)(ntab + ntd) * sizeof(double2) <= lim && (size_t)4 * 3 * P.max_nhh * sizeof(double2) <= 65536)
      S.t3_td = dalloc<double2>(ntd, S.owned);
  }
  if (mode == 0 && S.t3_td) {
    launch_t3_td(S.desc, S.t3, S.t3_td, S.stream);
    check_launch();
  }
  launch_triples(S.desc, S.t3, S.plan.orb.volume, S.t3_partial, S.t3_count, nb, S.t3_out, S.t3_count + 8192,
                 S.stream, mode, mode == 0 ? S.t3_td : nullptr, P.max_nhh);
  check_launch();
  double2 r;
  unsigned long long cnt = 0;
  HIPCHK(hipMemcpyAsync(&r, S.t3_out, sizeof(double2), hipMemcpyDeviceToHost, S.stream));
  HIPCHK(hipMemcpyAsync(&cnt, S.t3_count + 8192, sizeof(cnt), hipMemcpyDeviceToHost, S.stream));
  HIPCHK(hipStreamSynchronize(S.stream));
  S.t3_elements = (long long)cnt;
  return r;
}
}  // namespace
extern "C" {

int ccgpu_triples(int handle, double *et) {
  return guarded([&] {
    Solve &S = get(handle);
    ensure_ready(S);
    if (!S.have_t) throw Error("no amplitudes: solve CCD first");
    auto t0 = std::chrono::steady_clock::now();
    S.t3_et4 = triples_sum(S, 0);
    S.t3_et5 = double2{0.0, 0.0};
    if (S.tnf == 3) S.t3_et5 = triples_sum(S, 3);
    S.t3_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const double2 r = S.tnf == 3 ? S.t3_et5 : S.t3_et4;
    et[0] = r.x;
    et[1] = r.y;
  });
}

int ccgpu_triples_parts(int handle, double *et4, double *et5) {
  return guarded([&] {
    Solve &S = get(handle);
    et4[0] = S.t3_et4.x; et4[1] = S.t3_et4.y;
    et5[0] = S.t3_et5.x; et5[1] = S.t3_et5.y;
  });
}

int ccgpu_setup_t3(int handle, double *e3) {
  return guarded([&] {
    Solve &S = get(handle);
    ensure_ready(S);
    const double2 r = triples_sum(S, 2);
    S.have_ccdt1 = true;
    S.ccdt1_e3 = r;
    e3[0] = r.x;
    e3[1] = r.y;
  });
}

int ccgpu_t3_intermediate(int handle) {
  return guarded([&] {
    Solve &S = get(handle);
    ensure_ready(S);
    if (!S.have_ccdt1) throw Error("CCDT-1 not set up: call ccgpu_setup_t3 first");
    if (!S.have_t) throw Error("no amplitudes: call ccgpu_mbpt2_start first");
    S.ccdt1_e3 = triples_sum(S, 1);
  });
}

int ccgpu_t3_energy(int handle, double *e3) {
  return guarded([&] {
    Solve &S = get(handle);
    if (!S.have_ccdt1) throw Error("CCDT-1 not set up: call ccgpu_setup_t3 first");
    e3[0] = S.ccdt1_e3.x;
    e3[1] = S.ccdt1_e3.y;
  });
}

int ccgpu_triples_stats(int handle, long long *elements, double *seconds) {
  return guarded([&] {
    Solve &S = get(handle);
    if (elements) *elements = S.t3_elements;
    if (seconds) *seconds = S.t3_seconds;
  });
}

int ccgpu_debug_fetch(int handle, int what, double *out, long long n) {
  return guarded([&] {
    Solve &S = get(handle);
    const Plan &P = S.plan;
    const SolveDesc &d = S.desc;
    struct Arr { const void *p; long long n; };
    Arr a[] = {{d.R, P.nT}, {d.dh, P.A}, {d.I7off, P.A}, {d.dp, P.np}, {d.I6off, P.np}, {d.X, 2 * P.nVhh},
               {d.Rt, P.nRt}, {d.Tt, P.nRt}, {d.I4, P.nI4}, {d.T, P.nT}, {d.R0, P.nT}, {d.fdiag, P.nsp},
               {d.fpart, P.nsp}, {d.W0, P.nI4}, {d.Vt, P.nRt}};
    if (what < 0 || what >= (int)(sizeof a / sizeof a[0])) throw Error(fmt("unknown debug array %d", what));
    if (n != a[what].n) throw Error(fmt("debug array %d has %lld elements, caller asked %lld", what, a[what].n, n));
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(out, a[what].p, n * sizeof(double2), hipMemcpyDeviceToHost));
  });
}

// ---------------------------------------------------------------------------
// warm-start network
// ---------------------------------------------------------------------------
int ccgpu_mlp_create(unsigned long long seed, int *mlp) {
  return guarded([&] {
    auto M = std::make_unique<Mlp>();
    HIPCHK(hipStreamCreateWithFlags(&M->stream, hipStreamNonBlocking));
    M->params = dalloc<float>(kMlpParamCount, M->owned);
    M->m = dalloc<float>(kMlpParamCount, M->owned);
    M->v = dalloc<float>(kMlpParamCount, M->owned);
    M->grad = dalloc<float>(kMlpParamCount, M->owned);
    M->loss = dalloc<double>(1, M->owned);
    // He-normal hidden weights, small output layer, zero biases
    std::vector<float> th(kMlpParamCount, 0.f);
    std::mt19937_64 rng(seed);
    auto fill = [&](int off, int fan_in, int count, double gain) {
      std::normal_distribution<double> nd(0.0, gain / std::sqrt((double)fan_in));
      for (int k = 0; k < count; ++k) th[off + k] = (float)nd(rng);
    };
    fill(0, 16, 16 * 64, std::sqrt(2.0));
    fill(1088, 64, 64 * 64, std::sqrt(2.0));
    fill(5248, 64, 64 * 64, std::sqrt(2.0));
    fill(9408, 64, 64, 0.1);
    HIPCHK(hipMemcpy(M->params, th.data(), th.size() * sizeof(float), hipMemcpyHostToDevice));
    std::vector<float> z(kMlpParamCount, 0.f);
    HIPCHK(hipMemcpy(M->m, z.data(), z.size() * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(M->v, z.data(), z.size() * sizeof(float), hipMemcpyHostToDevice));
    int h = g_next_mlp++;
    g_mlps[h] = std::move(M);
    *mlp = h;
  });
}

int ccgpu_mlp_destroy(int mlp) {
  return guarded([&] {
    get_mlp(mlp);
    g_mlps.erase(mlp);
  });
}

int ccgpu_mlp_get_params(int mlp, float *theta, int n) {
  return guarded([&] {
    Mlp &M = get_mlp(mlp);
    if (n != kMlpParamCount) throw Error(fmt("the network has %d parameters, caller asked %d", kMlpParamCount, n));
    HIPCHK(hipStreamSynchronize(M.stream));
    HIPCHK(hipMemcpy(theta, M.params, n * sizeof(float), hipMemcpyDeviceToHost));
  });
}

int ccgpu_mlp_set_params(int mlp, const float *theta, int n) {
  return guarded([&] {
    Mlp &M = get_mlp(mlp);
    if (n != kMlpParamCount) throw Error(fmt("the network has %d parameters, caller gave %d", kMlpParamCount, n));
    HIPCHK(hipStreamSynchronize(M.stream));
    HIPCHK(hipMemcpy(M.params, theta, n * sizeof(float), hipMemcpyHostToDevice));
    std::vector<float> z(kMlpParamCount, 0.f);
    HIPCHK(hipMemcpy(M.m, z.data(), z.size() * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(M.v, z.data(), z.size() * sizeof(float), hipMemcpyHostToDevice));
    M.step = 0;
  });
}

// host features (n x 16, rounded to BF16 here) into the scratch buffer
static void upload_features(Mlp &M, int n, const float *x) {
  if (n <= 0) throw Error("no samples");
  M.ensure_scratch(n);
  std::vector<uint16_t> xb((size_t)n * 16);
  for (size_t k = 0; k < xb.size(); ++k) xb[k] = host_bf16(x[k]);
  HIPCHK(hipMemcpyAsync(M.sfeat, xb.data(), xb.size() * sizeof(uint16_t), hipMemcpyHostToDevice, M.stream));
  HIPCHK(hipStreamSynchronize(M.stream));
}

int ccgpu_mlp_forward(int mlp, int n, const float *x, float *y) {
  return guarded([&] {
    Mlp &M = get_mlp(mlp);
    upload_features(M, n, x);
    launch_mlp_forward(M.params, M.sfeat, n, M.sy, M.stream);
    check_launch();
    HIPCHK(hipMemcpyAsync(y, M.sy, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, M.stream));
    HIPCHK(hipStreamSynchronize(M.stream));
  });
}

int ccgpu_mlp_loss_grad(int mlp, int n, const float *x, const float *t, double *loss, float *grad) {
  return guarded([&] {
    Mlp &M = get_mlp(mlp);
    upload_features(M, n, x);
    HIPCHK(hipMemcpyAsync(M.st, t, (size_t)n * sizeof(float), hipMemcpyHostToDevice, M.stream));
    const int64_t npad = ((int64_t)n + 63) / 64 * 64;
    M.ensure_batch(npad);
    launch_mlp_loss_grad(M.params, M.sfeat, M.st, nullptr, n, npad, M.tb, M.part, Mlp::kMaxSplit, M.grad, M.loss,
                         M.stream);
    check_launch();
    double l = 0;
    HIPCHK(hipMemcpyAsync(&l, M.loss, sizeof(double), hipMemcpyDeviceToHost, M.stream));
    HIPCHK(hipMemcpyAsync(grad, M.grad, kMlpParamCount * sizeof(float), hipMemcpyDeviceToHost, M.stream));
    HIPCHK(hipStreamSynchronize(M.stream));
    *loss = l / n;
  });
}

int ccgpu_mlp_features(int mlp, int handle, float *feat, float *target) {
  return guarded([&] {
    Mlp &M = get_mlp(mlp);
    Solve &S = get(handle);
    ensure_ready(S);
    if (target && !S.have_t) throw Error("no amplitudes for the targets");
    const int64_t nT = S.plan.nT;
    M.ensure_scratch(nT);
    float kf, ef, tw[3];
    lattice_scalars(S, kf, ef, tw);
    HIPCHK(hipStreamSynchronize(S.stream));
    launch_mlp_features(S.desc, kf, ef, tw, kMlpScale, M.sfeat, target ? M.st : nullptr, M.stream);
    check_launch();
    std::vector<uint16_t> fb((size_t)nT * 16);
    HIPCHK(hipMemcpyAsync(fb.data(), M.sfeat, fb.size() * sizeof(uint16_t), hipMemcpyDeviceToHost, M.stream));
    if (target) HIPCHK(hipMemcpyAsync(target, M.st, (size_t)nT * sizeof(float), hipMemcpyDeviceToHost, M.stream));
    HIPCHK(hipStreamSynchronize(M.stream));
    for (size_t k = 0; k < fb.size(); ++k) {
      uint32_t u = (uint32_t)fb[k] << 16;
      std::memcpy(&feat[k], &u, 4);
    }
  });
}

int ccgpu_mlp_add_solve(int mlp, int handle, long long *nsamples) {
  return guarded([&] {
    Mlp &M = get_mlp(mlp);
    Solve &S = get(handle);
    ensure_ready(S);
    if (!S.have_t) throw Error("ccgpu_mlp_add_solve needs the solve's converged amplitudes");
    const int64_t nT = S.plan.nT;
    M.grow_set(M.n + nT);
    float kf, ef, tw[3];
    lattice_scalars(S, kf, ef, tw);
    HIPCHK(hipStreamSynchronize(S.stream));
    launch_mlp_features(S.desc, kf, ef, tw, kMlpScale, M.feat + M.n * 16, M.target + M.n, M.stream);
    check_launch();
    HIPCHK(hipStreamSynchronize(M.stream));
    M.n += nT;
    if (nsamples) *nsamples = M.n;
  });
}

int ccgpu_mlp_train(int mlp, int steps, int batch, double lr, unsigned long long seed, double *loss_last) {
  return guarded([&] {
    Mlp &M = get_mlp(mlp);
    if (M.n <= 0) throw Error("empty training set: call ccgpu_mlp_add_solve first");
    if (batch <= 0 || batch % 64 != 0) throw Error(fmt("batch %d is not a positive multiple of 64", batch));
    if (steps <= 0) throw Error("steps must be positive");
    M.ensure_batch(batch);
    for (int k = 0; k < steps; ++k) {
      launch_mlp_sample(M.idx, batch, M.n, seed + 0x51ED27ull * (uint64_t)(M.step + 1), M.stream);
      launch_mlp_loss_grad(M.params, M.feat, M.target, M.idx, batch, batch, M.tb, M.part, Mlp::kMaxSplit, M.grad,
                           M.loss, M.stream);
      ++M.step;
      launch_mlp_adam(M.params, M.grad, M.m, M.v, (float)lr, M.step, M.stream);
    }
    check_launch();
    double l = 0;
    HIPCHK(hipMemcpyAsync(&l, M.loss, sizeof(double), hipMemcpyDeviceToHost, M.stream));
    HIPCHK(hipStreamSynchronize(M.stream));
    if (loss_last) *loss_last = l / batch;
  });
}

// The learned start of one solve on stream st: one fused pass (features,
// network, t = t_MBPT2 + scale y), or with CCGPU_MLP_FUSED=0 the three kernels
// of the training path (same bits)
void mlp_start(Mlp &M, Solve &S, hipStream_t st) {
  ensure_ready(S);
  float kf, ef, tw[3];
  lattice_scalars(S, kf, ef, tw);
  const char *e = getenv("CCGPU_MLP_FUSED");
  if (e && atoi(e) == 0) {
    const int64_t nT = S.plan.nT;
    M.ensure_scratch(nT);
    launch_mlp_features(S.desc, kf, ef, tw, kMlpScale, M.sfeat, nullptr, st);
    launch_mlp_forward(M.params, M.sfeat, nT, M.sy, st);
    launch_mlp_apply(S.desc, M.sy, kMlpScale, ef, st);
  } else {
    launch_mlp_infer(M.params, S.desc, kf, ef, tw, kMlpScale, st);
  }
  check_launch();
  S.have_t = true;
  S.nstep = 0;
  S.count = 0;
}

int ccgpu_mlp_warm_start(int mlp, int handle, double *ecorr) {
  return guarded([&] {
    Mlp &M = get_mlp(mlp);
    Solve &S = get(handle);
    HIPCHK(hipStreamSynchronize(M.stream));
    // on the solve's stream
    mlp_start(M, S, S.stream);
    Batch &B = single(handle);
    auto e = energies(B, false);
    if (ecorr) { ecorr[0] = e[0].real(); ecorr[1] = e[0].imag(); }
  });
}

int ccgpu_mlp_warm_start_batch(int mlp, const int *handles, int n) {
  return guarded([&] {
    if (n < 1 || !handles) throw Error("warm_start_batch: no solves");
    Mlp &M = get_mlp(mlp);
    std::vector<Solve *> ss;
    for (int k = 0; k < n; ++k) ss.push_back(&get(handles[k]));
    // a solve's own stream may still hold work on its T
    for (Solve *S : ss) HIPCHK(hipStreamSynchronize(S->stream));
    HIPCHK(hipStreamSynchronize(M.stream));
    for (Solve *S : ss) mlp_start(M, *S, M.stream);
    HIPCHK(hipStreamSynchronize(M.stream));
  });
}

int ccgpu_autotune_report(char *buf, int len) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (buf && len > 0) {
    strncpy(buf, g_tune_report.c_str(), len - 1);
    buf[len - 1] = 0;
  }
  return (int)g_tune_report.size();
}

int ccgpu_hbm_read_probe(long long bytes, int reps, double *gbs) {
  return guarded([&] {
    if (bytes < (1ll << 20) || reps < 1 || !gbs) throw Error("hbm_read_probe: bytes >= 1 MiB, reps >= 1");
    const int64_t n = bytes / (int64_t)sizeof(double2);
    const int nblocks = 4096;
    std::vector<void *> own;
    double2 *a = dalloc<double2>((size_t)n, own);
    double *out = dalloc<double>(nblocks, own);
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double best = 1e300;
    try {
      HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
      HIPCHK(hipEventCreate(&e0));
      HIPCHK(hipEventCreate(&e1));
      HIPCHK(hipMemsetAsync(a, 0, (size_t)n * sizeof(double2), st));
      launch_read_probe(a, n, out, nblocks, st);   // warm-up
      check_launch();
      for (int r = 0; r < reps; ++r) {
        HIPCHK(hipEventRecord(e0, st));
        launch_read_probe(a, n, out, nblocks, st);
        HIPCHK(hipEventRecord(e1, st));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        best = std::min(best, (double)ms);
      }
      check_launch();
    } catch (...) {
      if (st) (void)hipStreamSynchronize(st);
      for (void *p : own) (void)hipFree(p);
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
      if (st) (void)hipStreamDestroy(st);
      throw;
    }
    HIPCHK(hipStreamSynchronize(st));
    for (void *p : own) (void)hipFree(p);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    *gbs = (double)n * sizeof(double2) / (best * 1e-3) / 1e9;
  });
}

int ccgpu_kernel_stats(char *names, int len, double *avg_ms, long long *launches, int max) {
  std::lock_guard<std::mutex> lk(g_mu);
  std::string all;
  int k = 0;
  for (auto &nm : g_stat_order) {
    if (k >= max) break;
    const Stat &s = g_stats[nm];
    if (avg_ms) avg_ms[k] = s.n ? s.ms / s.n : 0.0;
    if (launches) launches[k] = s.n;
    if (!all.empty()) all += ";";
    all += nm;
    ++k;
  }
  if (names && len > 0) {
    strncpy(names, all.c_str(), len - 1);
    names[len - 1] = 0;
  }
  return k;
}

}  // extern "C"
