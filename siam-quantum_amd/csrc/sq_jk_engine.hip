// sq_jk_engine.hip - host side of the two-electron engine and its C-ABI.
//
// Per molecule (sqjk_create):
//   * reference shells: consecutive basis functions on one centre with one
//     exponent list (the grouping of src/matrix.c:796-822, 2457-2541); they
//     are the screening unit, exactly as in the reference;
//   * pure-l sub-shells: a Pople L shell is split into its s and p parts so
//     that every kernel class has one angular momentum per shell;
//   * primitive pairs (p, P, coefficient*K_ab/p, K_ab/p) in HBM;
//   * Schwarz bounds sqrt|(pq|pq)| computed on the device unless given;
//   * unique sub-shell pairs binned by pair class (ss,ps,pp,ds,dp,dd) and
//     sorted by the parent-level Schwarz bound, descending.
//   * within a class, pairs with the same number of primitive pairs form a
//     group (sorted by Schwarz bound inside the group) whose primitive data is
//     stored primitive-major.
// Per build (sqjk_build):
//   * for every class pair (X>=Y), bra group and ket group, the surviving
//     quartets of bra pair i are the prefix j < cnt_i of the ket group; the
//     rows and their offsets (a segment) depend only on the cut-offs and are
//     kept on the device until the cut-offs change;
//   * integrals are digested into half-accumulators A (J = A + A^T) with
//     atomics; a finalize kernel forms J, -K and the host applies the
//     reference's accumulate/restricted/symmetrise semantics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sq_jk.h"
#include "../../include/sq_grad.h"
#include "../../include/sq_mp2.h"
#include "sq_jk_classes.h"
#include "sq_grad_classes.h"
#include "sq_mp2_classes.h"
#include "sq_jk_wd.h"
#include "sq_one_e.h"
#include "sq_internal.h"

using namespace sqjk;

namespace {

thread_local std::string g_err;
int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(call)                                                               \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      return fail(SQJK_EHIP, std::string(#call " failed: ") + hipGetErrorString(e_)); \
  } while (0)

// ------------------------------------------------------------- Boys table
// F_m(k*step) for m < kBoysNM and exp(-k*step) in column kBoysNM: series for
// the top order in long double
// (F_m(T) = e^-T sum_i (2T)^i / ((2m+1)(2m+3)...(2m+2i+1))), then the
// stable downward recursion F_m = (2T F_{m+1} + e^-T)/(2m+1).
std::vector<double> make_boys_table() {
  std::vector<double> tab((size_t)kBoysNT * kBoysCols);
  for (int k = 0; k < kBoysNT; ++k) {
    const long double T = (long double)k * (long double)kBoysStep;
    const int top = kBoysNM - 1 + 20;  // start a few orders above, recur down
    long double term = 1.0L / (2 * top + 1), sum = 0.0L;
    for (int i = 1; i < 2000; ++i) {
      sum += term;
      term *= 2.0L * T / (2 * top + 2 * i + 1);
      if (term < 1e-30L * sum) break;
    }
    const long double ex = expl(-T);
    long double f = ex * sum;
    std::vector<long double> F(top + 1);
    F[top] = f;
    for (int m = top - 1; m >= 0; --m) F[m] = (2.0L * T * F[m + 1] + ex) / (2 * m + 1);
    for (int m = 0; m < kBoysNM; ++m) tab[(size_t)k * kBoysCols + m] = (double)F[m];
    tab[(size_t)k * kBoysCols + kBoysNM] = (double)ex;
  }
  return tab;
}

// ------------------------------------------------------------- classes
constexpr int kNPairClass = 6;  // (0,0) (1,0) (1,1) (2,0) (2,1) (2,2)
int pair_class(int la, int lb) {  // la >= lb
  static const int tab[3][3] = {{0, -1, -1}, {1, 2, -1}, {3, 4, 5}};
  return (la <= 2 && lb <= la) ? tab[la][lb] : -1;
}
const int kPairL[kNPairClass][2] = {{0, 0}, {1, 0}, {1, 1}, {2, 0}, {2, 1}, {2, 2}};

ClassLauncher launcher_for(int la, int lb, int lc, int ld) {
#define SQJK_CASE(a, b, c, d) \
  if (la == a && lb == b && lc == c && ld == d) return &sqjk_launch_##a##b##c##d;
  SQJK_CLASSES(SQJK_CASE)
#undef SQJK_CASE
  return nullptr;
}

GradLauncher grad_launcher_for(int la, int lb, int lc, int ld) {
#define SQGRAD_CASE(a, b, c, d) \
  if (la == a && lb == b && lc == c && ld == d) return &sqgrad_launch_##a##b##c##d;
  SQJK_CLASSES(SQGRAD_CASE)
#undef SQGRAD_CASE
  return nullptr;
}

StoreLauncher store_launcher_for(int la, int lb, int lc, int ld) {
#define SQMP2_CASE(a, b, c, d) \
  if (la == a && lb == b && lc == c && ld == d) return &sqmp2_store_##a##b##c##d;
  SQJK_CLASSES(SQMP2_CASE)
#undef SQMP2_CASE
  return nullptr;
}

struct RefShell {
  int first, count;  // basis functions
};

__global__ void k_fill(double* p, size_t n, double v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// J = A + A^T ; K = -(A + A^T)
__global__ void k_finalize(const double* __restrict__ AJ, const double* __restrict__ AKA,
                           const double* __restrict__ AKB, double* J, double* KA, double* KB, int n,
                           int uhf) {
  const size_t nn = (size_t)n * n;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nn; e += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / n), j = (int)(e - (size_t)i * n);
    const size_t t = (size_t)j * n + i;
    J[e] = AJ[e] + AJ[t];
    KA[e] = -(AKA[e] + AKA[t]);
    if (uhf) KB[e] = -(AKB[e] + AKB[t]);
  }
}

// D_J = P_A + P_B
__global__ void k_density_sum(const double* __restrict__ DA, const double* __restrict__ DB, double* DJ, size_t nn) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nn; e += (size_t)gridDim.x * blockDim.x)
    DJ[e] = DA[e] + DB[e];
}

// PM[P][Q] = max |D| over the (P,Q) block of reference shells for D = D_J,
// P_A, P_B (the density factor of the screening test, src/matrix.c:3921-3924)
__global__ void k_pair_max(const double* __restrict__ DJ, const double* __restrict__ DA,
                           const double* __restrict__ DB, const int* __restrict__ ref, int nr, int n,
                           double* PM) {
  const long long tot = (long long)nr * nr;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < tot;
       t += (long long)gridDim.x * blockDim.x) {
    const int P = (int)(t / nr), Q = (int)(t - (long long)P * nr);
    const int p0 = ref[2 * P], np = ref[2 * P + 1], q0 = ref[2 * Q], nq = ref[2 * Q + 1];
    double r = 0.0;
    for (int p = p0; p < p0 + np; ++p)
      for (int q = q0; q < q0 + nq; ++q) {
        const size_t e = (size_t)p * n + q;
        r = fmax(r, fmax(fabs(DJ[e]), fmax(fabs(DA[e]), fabs(DB[e]))));
      }
    PM[t] = r;
  }
}

// A_J[fa+ia][fb+ib] += Jp[ab][i] for every pair i of one class.  Every
// (pair, component) owns one element of A_J, so plain read-modify-write is
// safe (the class kernels have finished).
__global__ void k_reduce_j(const double* __restrict__ Jp, const PairRec* __restrict__ pr,
                           const SubShell* __restrict__ sh, int np, int nb, int nab, double* AJ, int n) {
  const long long tot = (long long)np * nab;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot;
       e += (long long)gridDim.x * blockDim.x) {
    const double v = Jp[e];
    if (v == 0.0) continue;
    const int ab = (int)(e / np), i = (int)(e - (long long)ab * np);
    const PairRec p = pr[i];
    AJ[(size_t)(sh[p.a].bf_off + ab / nb) * n + sh[p.b].bf_off + ab % nb] += v;
  }
}

// Exchange staging of the per-thread kernels, one ket class per slot:
// S[x][comp][j] belongs to row x of A_K at column bf(c of pair j) + comp
// (comp < nc) or bf(d of pair j) + comp - nc.
// Exchange staging of a fold: per ket class Y, S[Y] holds [n][ncomp][np]
// for the ket pairs [p0, p0 + np) of the class, of which the first cnt exist.
struct KStage {
  const double* S[6];
  const PairRec* pr[6];
  int np[6], nc[6], ncomp[6], p0[6], cnt[6];
};

// One workgroup per row x of A_K: the row is summed in LDS, then added to A_K.
// The staging is streamed with four independent loads per lane in flight
// (1024-lane workgroups), since the reduction is bound by HBM reads.
constexpr int kReduceKBlock = 1024;
__global__ void __launch_bounds__(kReduceKBlock) k_reduce_k(KStage R, const SubShell* __restrict__ sh, double* AK,
                                                            int n) {
  extern __shared__ double row[];
  for (int x = blockIdx.x; x < n; x += gridDim.x) {
    for (int y = threadIdx.x; y < n; y += blockDim.x) row[y] = 0.0;
    __syncthreads();
    for (int Y = 0; Y < 6; ++Y) {
      if (!R.S[Y]) continue;
      const int np = R.np[Y], nc = R.nc[Y], ncomp = R.ncomp[Y], cnt = R.cnt[Y];
      const double* Sx = R.S[Y] + (size_t)x * ncomp * np;
      const PairRec* prY = R.pr[Y] + R.p0[Y];
      const long long tot = (long long)ncomp * np;
      const long long bs = blockDim.x;
      for (long long e0 = threadIdx.x; e0 < tot; e0 += 4 * bs) {
        double v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = (e0 + k * bs < tot) ? Sx[e0 + k * bs] : 0.0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          if (v[k] == 0.0) continue;
          const long long e = e0 + k * bs;
          const int comp = (int)(e / np), j = (int)(e - (long long)comp * np);
          if (j >= cnt) continue;
          const PairRec p = prY[j];
          const int col = comp < nc ? sh[p.a].bf_off + comp : sh[p.b].bf_off + comp - nc;
          __hip_atomic_fetch_add(&row[col], v[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
      }
    }
    __syncthreads();
    for (int y = threadIdx.x; y < n; y += blockDim.x) AK[(size_t)x * n + y] += row[y];
    __syncthreads();
  }
}

// Work that survives the screening in one launch of a class: contracted
// quartets that pass the density-weighted test, and their primitive
// quartets whose bare prefactor passes the 1e-15 cut.  Same tile walk as
// the class kernels (one wavefront per tile, lane = ket position); a
// quartet is counted once, by the tile that holds the first ket primitive
// pair of its segment (tiles of a split segment repeat the same quartets).
// cnt[0] += quartets, cnt[1] += primitive quartets.
__global__ void __launch_bounds__(256) k_count_work(ClassArgs A, unsigned long long* cnt) {
  const int lane = threadIdx.x & 63;
  const long long wstride = (long long)gridDim.x * (blockDim.x >> 6);
  double q = 0.0, p = 0.0;
  for (long long w = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); w < A.total; w += wstride) {
    const Tile tl = A.tile[w];
    const Segment sg = A.seg[tl.seg];
    const View V = make_view(A, sg.inst);
    const int j = tl.k0 + lane;
    for (int r = tl.r0; r < tl.r1; ++r) {
      if (j < A.rlo[r] || j >= A.rhi[r]) continue;
      const int ib = A.rbra[r], ik = sg.ket0 + j;
      const PairRec bp = A.bra[ib], kp = A.ket[ik];
      if (!density_screen_pass(V, bp, kp)) continue;
      if (tl.v0 == 0) q += 1.0;
      for (int u = 0; u < sg.nppb; ++u) {
        const PrimPair P = A.pp[bp.pp_off + u * bp.pp_stride];
        for (int v = tl.v0; v < tl.v1; ++v) {
          const PrimPair Q = A.pp[kp.pp_off + v * kp.pp_stride];
          if (!(kTwoPi25 * P.eb * Q.eb / sqrt(P.p + Q.p) < kPrimCutoff)) p += 1.0;
        }
      }
    }
  }
  q = wave_sum(q);
  p = wave_sum(p);
  if (lane == 0) {
    if (q > 0.0) atomicAdd(&cnt[0], (unsigned long long)q);
    if (p > 0.0) atomicAdd(&cnt[1], (unsigned long long)p);
  }
}

// F_0..F_L at nT points by the class kernels' Boys code (boys<L, R> from
// the LDS table), for checking it against independent values.
template <int L, typename R>
__global__ void __launch_bounds__(256) k_boys_eval(const double* __restrict__ tab, const double* __restrict__ T,
                                                   double* F, int nT) {
  extern __shared__ double lds[];
  boys_load_lds<L>(lds, tab);
  __syncthreads();
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nT; i += gridDim.x * blockDim.x) {
    R f[L + 1];
    boys<L, R>(R(T[i]), f, lds);
    for (int m = 0; m <= L; ++m) F[(size_t)i * (L + 1) + m] = (double)f[m];
  }
}

template <typename R>
int launch_boys_eval(int L, const double* tab, const double* T, double* F, int nT, hipStream_t st) {
  const unsigned grid = (unsigned)std::max(1, std::min(1024, (nT + 255) / 256));
  switch (L) {
#define SQJK_BOYS_CASE(l)                                                                                    \
  case l:                                                                                                    \
    hipLaunchKernelGGL((k_boys_eval<l, R>), dim3(grid), dim3(256), sizeof(double) * boys_rows<l>() * kBoysLdsW, \
                       st, tab, T, F, nT);                                                                   \
    return 0;
    SQJK_BOYS_CASE(0) SQJK_BOYS_CASE(1) SQJK_BOYS_CASE(2) SQJK_BOYS_CASE(3) SQJK_BOYS_CASE(4)
    SQJK_BOYS_CASE(5) SQJK_BOYS_CASE(6) SQJK_BOYS_CASE(7) SQJK_BOYS_CASE(8) SQJK_BOYS_CASE(9)
#undef SQJK_BOYS_CASE
    default: return -1;
  }
}

}  // namespace

struct sqjk_ctx {
  int device = 0, n = 0, nref = 0, nsub = 0;
  std::vector<RefShell> ref;
  std::vector<SubShell> sub;
  std::vector<double> bscale;
  std::vector<PrimPair> pp;
  std::vector<PairRec> pairs[kNPairClass];
  std::vector<double> sch_basis;  // n x n
  std::vector<double> sch_ref;    // nref x nref
  std::vector<double> pexp, pcoef;  // primitive exponents and coefficients per sub-shell
  // device
  hipStream_t st = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  SubShell* d_sub = nullptr;
  PrimPair* d_pp = nullptr;
  PairRec* d_pairs[kNPairClass] = {};
  double *d_boys = nullptr, *d_bscale = nullptr;
  double *d_DJ = nullptr, *d_DA = nullptr, *d_DB = nullptr, *d_PM = nullptr;
  double *d_AJ = nullptr, *d_AKA = nullptr, *d_AKB = nullptr;
  double *d_J = nullptr, *d_KA = nullptr, *d_KB = nullptr, *d_sch = nullptr;
  struct Group { int begin, count, npp; };
  std::vector<Group> groups[kNPairClass];
  // work plan for the cut-offs (thr, thr64) of the last build
  // lr0, nlr: its (chunk, tile begin, tile end) ranges in h_lrange when the
  // exchange staging is held in ket-pair chunks
  struct Launch { int X, Y, fp32, seg0, tile0; long long ntile, quartets; int lr0, nlr; };
  // plan_ok: the device plan (d_seg ... d_rhi) holds the host plan below,
  // made for the cut-offs (plan_thr, plan_thr64); hplan_ok: the host plan
  // was made for (hplan_thr, hplan_thr64)
  bool plan_ok = false, hplan_ok = false;
  double plan_thr = 0.0, plan_thr64 = 0.0, hplan_thr = 0.0, hplan_thr64 = 0.0;
  std::vector<Launch> launches;
  std::vector<Segment> h_seg;
  std::vector<Tile> h_tile;
  std::vector<int> h_rbra, h_rlo, h_rhi;
  Segment* d_seg = nullptr;
  Tile* d_tile = nullptr;
  int *d_rbra = nullptr, *d_rlo = nullptr, *d_rhi = nullptr;
  size_t seg_cap = 0, tile_cap = 0, row_cap = 0;
  long long plan_q = 0, plan_q32 = 0;
  // pair-indexed digestion buffers per class: Jp [NAB][np], S [n][NC+ND][np]
  bool profile = false;
  std::vector<hipEvent_t> prof_ev;                      // 2 per recorded launch
  std::vector<int> prof_launch;                         // launch of each event pair
  double cls_ms[kNPairClass][kNPairClass] = {};
  // Exchange staging in ket-pair chunks, when the staging of all classes
  // would take more than stage_budget bytes per spin: the launches of ket
  // class Y run chunk by chunk (chunk_len[Y] pairs each, held with a stride
  // of chunk_stride[Y] columns so that a tile starting in the chunk fits),
  // sharing one buffer per spin (d_SC, d_SCB, sc_len doubles), and each
  // chunk is folded into A_K before the next.
  size_t stage_budget = size_t(256) << 20;
  bool stage_chunked = false;
  int chunk_len[kNPairClass] = {}, chunk_stride[kNPairClass] = {}, chunk_cnt[kNPairClass] = {};
  double *d_SC = nullptr, *d_SCB = nullptr;
  size_t sc_len = 0;
  std::vector<std::array<int, 3>> h_lrange;
  double* d_Jp[kNPairClass] = {};
  double* d_SA[kNPairClass] = {};
  double* d_SB[kNPairClass] = {};
  size_t jp_len[kNPairClass] = {}, s_len[kNPairClass] = {};
  double mixed_thr = 0.0;  // Schwarz products below this run in FP32
  // pinned host staging of the densities (in) and of J, K (out)
  double *hp_A = nullptr, *hp_B = nullptr, *hp_J = nullptr, *hp_KA = nullptr, *hp_KB = nullptr;
  int* d_ref = nullptr;  // reference shells: first basis function, count
  // Hermite coefficient tables of the dd-class primitive pairs (their
  // primitive-pair indices are [herm_base, herm_base + herm_count))
  double* d_herm = nullptr;
  int herm_base = 0, herm_count = 0;
  bool no_wd = false;  // SQJK_DD_KERNEL=wave: the cooperative (dd|dd) kernel instead
  sqjk_stats last{};
  // per class pair (X,Y) of the last build: contracted and primitive quartets
  long long cls_q[kNPairClass][kNPairClass] = {};
  long long cls_p[kNPairClass][kNPairClass] = {};
  // work-plan tuning per class pair: bra rows per tile, integral-primitive
  // products per lane and tile before a segment is split, and the grid's
  // multiple of the resident workgroups (0: the kernel's default)
  int rows[kNPairClass][kNPairClass] = {};
  int split_budget[kNPairClass][kNPairClass] = {};
  int grid_over[kNPairClass][kNPairClass] = {};
  double last_fixed = 0.0;  // density-weighted cut-off of the last build
  bool timing_pending = false;  // the last build's events are not yet read
  bool built = false;       // a build has run (PM holds its density maxima)
};

namespace {

// Device and pinned host blocks released by a context are kept for the next
// one (an SCF creates and releases its context once per solve; hipFree and
// hipHostFree synchronise the device and cost more than the work they end).
// A request takes the smallest cached block of the same kind and device that
// holds it, at most twice its size; blocks beyond 4 GiB in total are freed.
struct Block { void* p; size_t bytes; int device; bool host; };
std::mutex g_block_mu;
std::vector<Block> g_blocks;                   // cached, free
std::vector<Block> g_live;                     // handed out
size_t g_block_bytes = 0;
constexpr size_t kBlockCacheMax = size_t(4) << 30;

void* block_get(size_t bytes, bool host) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(g_block_mu);
  size_t best = g_blocks.size();
  for (size_t i = 0; i < g_blocks.size(); ++i) {
    const Block& b = g_blocks[i];
    if (b.host == host && (host || b.device == dev) && b.bytes >= bytes && b.bytes <= 2 * bytes &&
        (best == g_blocks.size() || b.bytes < g_blocks[best].bytes))
      best = i;
  }
  if (best < g_blocks.size()) {
    Block b = g_blocks[best];
    g_blocks.erase(g_blocks.begin() + (long)best);
    g_block_bytes -= b.bytes;
    g_live.push_back(b);
    return b.p;
  }
  void* p = nullptr;
  const hipError_t e = host ? hipHostMalloc(&p, bytes, hipHostMallocDefault) : hipMalloc(&p, bytes);
  if (e != hipSuccess) return nullptr;
  g_live.push_back({p, bytes, dev, host});
  return p;
}

void block_put(void* p) {
  if (!p) return;
  std::lock_guard<std::mutex> lk(g_block_mu);
  for (size_t i = 0; i < g_live.size(); ++i)
    if (g_live[i].p == p) {
      Block b = g_live[i];
      g_live.erase(g_live.begin() + (long)i);
      if (g_block_bytes + b.bytes <= kBlockCacheMax) {
        g_blocks.push_back(b);
        g_block_bytes += b.bytes;
      } else if (b.host) {
        (void)hipHostFree(p);
      } else {
        (void)hipFree(p);
      }
      return;
    }
}

template <typename T>
int dalloc(T** p, size_t count) {
  if (count == 0) count = 1;
  void* v = block_get(count * sizeof(T), false);
  if (!v) return fail(SQJK_ENOMEM, "device allocation failed");
  *p = static_cast<T*>(v);
  return SQJK_OK;
}

int grid_for(long long work, int per_block) {
  long long b = (work + per_block - 1) / per_block;
  if (b < 1) b = 1;
  if (b > 4096) b = 4096;
  return (int)b;
}

bool same_shell(const GTOBasis_t* g, int i, int j) {
  if (g[i].x0 != g[j].x0 || g[i].y0 != g[j].y0 || g[i].z0 != g[j].z0) return false;
  if (g[i].nContract != g[j].nContract) return false;
  for (int k = 0; k < g[i].nContract; ++k)
    if (g[i].exp[k] != g[j].exp[k]) return false;
  return true;
}

// Reference shells, pure-l sub-shells, component scales, primitive data.
int plan_shells(sqjk_ctx* c, int n, const GTOBasis_t* g, std::vector<double>& pexp,
                std::vector<double>& pcoef) {
  for (int p = 0; p < n;) {
    int q = p;
    while (q < n && same_shell(g, p, q)) ++q;
    c->ref.push_back({p, q - p});
    p = q;
  }
  c->nref = (int)c->ref.size();
  c->bscale.assign(n, 0.0);
  for (int r = 0; r < c->nref; ++r) {
    const RefShell& rs = c->ref[r];
    for (int f = rs.first; f < rs.first + rs.count;) {
      const int l = g[f].l + g[f].m + g[f].n;
      if (l > 2)
        return fail(SQJK_EBASIS, "angular momentum above d is not supported by this build (function " +
                                     std::to_string(f) + ")");
      const int nc = sqmd::ncart(l);
      if (f + nc > rs.first + rs.count)
        return fail(SQJK_EBASIS, "incomplete Cartesian shell at function " + std::to_string(f));
      for (int i = 0; i < nc; ++i) {
        const sqmd::Cart ct = sqmd::cart(l, i);
        if (g[f + i].l != ct.x || g[f + i].m != ct.y || g[f + i].n != ct.z)
          return fail(SQJK_EBASIS, "Cartesian components out of the reference order at function " +
                                       std::to_string(f + i));
      }
      SubShell s{};
      s.x = g[f].x0; s.y = g[f].y0; s.z = g[f].z0;
      s.l = l; s.nprim = g[f].nContract; s.prim_off = (int)pexp.size(); s.bf_off = f; s.parent = r; s.gbf = f;
      // contraction factorises as coef*norm = ctilde[k] * scale[component]
      for (int k = 0; k < s.nprim; ++k) {
        const double c0 = g[f].coef[k] * g[f].norm[k];
        if (c0 == 0.0) return fail(SQJK_EBASIS, "zero contraction coefficient at function " + std::to_string(f));
        pexp.push_back(g[f].exp[k]);
        pcoef.push_back(c0);
      }
      for (int i = 0; i < nc; ++i) {
        const double r0 = (g[f + i].coef[0] * g[f + i].norm[0]) / (g[f].coef[0] * g[f].norm[0]);
        for (int k = 1; k < s.nprim; ++k) {
          const double rk = (g[f + i].coef[k] * g[f + i].norm[k]) / (g[f].coef[k] * g[f].norm[k]);
          if (std::fabs(rk - r0) > 1e-10 * std::fabs(r0))
            return fail(SQJK_EBASIS, "contraction of function " + std::to_string(f + i) +
                                         " does not factor into a radial part and a component scale");
        }
        c->bscale[f + i] = r0;
      }
      c->sub.push_back(s);
      f += nc;
    }
  }
  c->nsub = (int)c->sub.size();
  return SQJK_OK;
}

// all unique sub-shell pairs with primitive-pair data; sch filled later
void plan_pairs(sqjk_ctx* c, const std::vector<double>& pexp, const std::vector<double>& pcoef) {
  for (int X = 0; X < kNPairClass; ++X) c->pairs[X].clear();
  c->pp.clear();
  for (int i = 0; i < c->nsub; ++i)
    for (int j = 0; j <= i; ++j) {
      int a = i, b = j;
      if (c->sub[b].l > c->sub[a].l) std::swap(a, b);
      const SubShell &A = c->sub[a], &B = c->sub[b];
      PairRec pr{};
      pr.a = a; pr.b = b; pr.pa = A.parent; pr.pb = B.parent;
      pr.pp_off = (int)c->pp.size(); pr.pp_stride = 1; pr.npp = A.nprim * B.nprim;
      const double ab2 = (A.x - B.x) * (A.x - B.x) + (A.y - B.y) * (A.y - B.y) + (A.z - B.z) * (A.z - B.z);
      for (int ka = 0; ka < A.nprim; ++ka)
        for (int kb = 0; kb < B.nprim; ++kb) {
          const double ea = pexp[A.prim_off + ka], eb = pexp[B.prim_off + kb];
          const double p = ea + eb;
          const double K = std::exp(-ea * eb * ab2 / p) / p;
          PrimPair q{};
          q.p = p;
          q.Px = (ea * A.x + eb * B.x) / p; q.Py = (ea * A.y + eb * B.y) / p; q.Pz = (ea * A.z + eb * B.z) / p;
          q.eb = K;
          q.cc = pcoef[A.prim_off + ka] * pcoef[B.prim_off + kb] * K;
          q.hp = 0.5 / p;
          q.ea = ea;
          c->pp.push_back(q);
        }
      c->pairs[pair_class(A.l, B.l)].push_back(pr);
    }
}

const std::vector<double>& boys_table() {
  static std::vector<double> boys;
  static std::once_flag once;
  std::call_once(once, [] { boys = make_boys_table(); });
  return boys;
}

int upload_static(sqjk_ctx* c) {
  const int n = c->n;
  const size_t nn = (size_t)n * n;
  int rc;
  if ((rc = dalloc(&c->d_sub, c->sub.size()))) return rc;
  if ((rc = dalloc(&c->d_pp, c->pp.size()))) return rc;
  if ((rc = dalloc(&c->d_bscale, (size_t)n))) return rc;
  if ((rc = dalloc(&c->d_boys, (size_t)kBoysNT * kBoysCols))) return rc;
  double** mats[] = {&c->d_DJ, &c->d_DA, &c->d_DB, &c->d_AJ, &c->d_AKA, &c->d_AKB,
                     &c->d_J,  &c->d_KA, &c->d_KB, &c->d_sch};
  for (double** m : mats)
    if ((rc = dalloc(m, nn))) return rc;
  if ((rc = dalloc(&c->d_PM, (size_t)c->nref * c->nref))) return rc;
  if ((rc = dalloc(&c->d_ref, (size_t)2 * c->nref))) return rc;
  double** pins[] = {&c->hp_A, &c->hp_B, &c->hp_J, &c->hp_KA, &c->hp_KB};
  for (double** h : pins)
    if (!(*h = static_cast<double*>(block_get(nn * sizeof(double), true))))
      return fail(SQJK_ENOMEM, "pinned host allocation failed");
  std::vector<int> ref2((size_t)2 * c->nref);
  for (int P = 0; P < c->nref; ++P) {
    ref2[2 * P] = c->ref[P].first;
    ref2[2 * P + 1] = c->ref[P].count;
  }
  HIPCHK(hipMemcpy(c->d_ref, ref2.data(), ref2.size() * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpyAsync(c->d_sub, c->sub.data(), c->sub.size() * sizeof(SubShell), hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->d_pp, c->pp.data(), c->pp.size() * sizeof(PrimPair), hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->d_bscale, c->bscale.data(), n * sizeof(double), hipMemcpyHostToDevice, c->st));
  const std::vector<double>& boys = boys_table();
  HIPCHK(hipMemcpyAsync(c->d_boys, boys.data(), boys.size() * sizeof(double), hipMemcpyHostToDevice, c->st));
  return SQJK_OK;
}

int upload_pairs(sqjk_ctx* c) {
  for (int X = 0; X < kNPairClass; ++X) {
    if (c->d_pairs[X]) { block_put(c->d_pairs[X]); c->d_pairs[X] = nullptr; }
    int rc = dalloc(&c->d_pairs[X], c->pairs[X].size());
    if (rc) return rc;
    if (!c->pairs[X].empty())
      HIPCHK(hipMemcpyAsync(c->d_pairs[X], c->pairs[X].data(), c->pairs[X].size() * sizeof(PairRec),
                            hipMemcpyHostToDevice, c->st));
  }
  return SQJK_OK;
}

// Hermite coefficient tables of the dd-class primitive pairs (geometry
// only), for the wavefront-per-quartet (dd|dd) kernel.  The class's primitive
// pairs are contiguous in the grouped primitive-major layout.
int build_herm_tables(sqjk_ctx* c) {
  constexpr int X = kNPairClass - 1;  // dd
  using HL = HermTabLayout<2, 2>;
  const auto& prs = c->pairs[X];
  if (prs.empty()) return SQJK_OK;
  int lo = INT32_MAX, hi = 0;
  for (const PairRec& pr : prs) {
    lo = std::min(lo, pr.pp_off);
    hi = std::max(hi, pr.pp_off + (pr.npp - 1) * pr.pp_stride + 1);
    if (pr.npp > 64) return fail(SQJK_EBASIS, "a dd shell pair has more than 64 primitive pairs");
  }
  const int count = hi - lo;
  if (!c->d_herm || count > c->herm_count) {
    if (c->d_herm) block_put(c->d_herm);
    c->d_herm = nullptr;
    if (int rc = dalloc(&c->d_herm, (size_t)count * HL::SIZE)) return rc;
  }
  c->herm_base = lo;
  c->herm_count = count;
  const long long work = (long long)prs.size() * 64;
  hipLaunchKernelGGL((k_herm_table<2, 2>), dim3(grid_for(work, 256)), dim3(256), 0, c->st, c->d_pairs[X],
                     (int)prs.size(), c->d_pp, c->d_sub, lo, c->d_herm);
  HIPCHK(hipGetLastError());
  return SQJK_OK;
}

// Lengths of the pair-indexed digestion buffers: Jp [NAB][np] and the
// exchange staging S [n][NC+ND][np] of every pair class.
void stage_lengths(sqjk_ctx* c) {
  for (int X = 0; X < kNPairClass; ++X) {
    const size_t np = c->pairs[X].size();
    const int nab = sqmd::ncart(kPairL[X][0]) * sqmd::ncart(kPairL[X][1]);
    const int ncomp = sqmd::ncart(kPairL[X][0]) + sqmd::ncart(kPairL[X][1]);
    c->jp_len[X] = np * nab;
    c->s_len[X] = (size_t)c->n * ncomp * np;
  }
}

// Exchange staging budget per spin: 256 MiB, or SQJK_STAGE_MB.
size_t stage_budget_bytes() {
  if (const char* e = std::getenv("SQJK_STAGE_MB")) {
    const long long mb = std::atoll(e);
    if (mb > 0) return (size_t)mb << 20;
  }
  return size_t(256) << 20;
}

// Whether the exchange staging is held in ket-pair chunks, and their sizes:
// whole classes when all of it fits in stage_budget bytes per spin, else
// per ket class the longest chunk (a multiple of 64 pairs) whose buffer of
// n x ncomp x (chunk + 64) doubles fits.
void plan_stage(sqjk_ctx* c) {
  stage_lengths(c);
  size_t total = 0;
  for (int X = 0; X < kNPairClass; ++X) total += c->s_len[X] * sizeof(double);
  c->stage_chunked = total > c->stage_budget;
  c->sc_len = 0;
  for (int Y = 0; Y < kNPairClass; ++Y) {
    const int np = (int)c->pairs[Y].size();
    const size_t ncomp = sqmd::ncart(kPairL[Y][0]) + sqmd::ncart(kPairL[Y][1]);
    if (!c->stage_chunked || np == 0) {
      c->chunk_len[Y] = c->chunk_stride[Y] = np;
      c->chunk_cnt[Y] = np > 0 ? 1 : 0;
      continue;
    }
    const long long cols = (long long)(c->stage_budget / (sizeof(double) * (size_t)c->n * ncomp));
    int len = (int)std::max<long long>(64, (cols - 64) / 64 * 64);
    if (len >= np) len = np;
    c->chunk_len[Y] = len;
    c->chunk_stride[Y] = len >= np ? np : len + 64;
    c->chunk_cnt[Y] = (np + len - 1) / len;
    c->sc_len = std::max(c->sc_len, (size_t)c->n * ncomp * c->chunk_stride[Y]);
  }
}

// Pair-indexed digestion buffers (alpha always, beta for UHF builds).
int ensure_stage(sqjk_ctx* c, bool uhf) {
  stage_lengths(c);
  if (c->stage_chunked) {
    int rc;
    for (int X = 0; X < kNPairClass; ++X)
      if (!c->pairs[X].empty() && !c->d_Jp[X] && (rc = dalloc(&c->d_Jp[X], c->jp_len[X]))) return rc;
    if (!c->d_SC && (rc = dalloc(&c->d_SC, c->sc_len))) return rc;
    if (uhf && !c->d_SCB && (rc = dalloc(&c->d_SCB, c->sc_len))) return rc;
    return SQJK_OK;
  }
  for (int X = 0; X < kNPairClass; ++X) {
    if (c->pairs[X].empty()) continue;
    int rc;
    if (!c->d_Jp[X] && (rc = dalloc(&c->d_Jp[X], c->jp_len[X]))) return rc;
    if (!c->d_SA[X] && (rc = dalloc(&c->d_SA[X], c->s_len[X]))) return rc;
    if (uhf && !c->d_SB[X] && (rc = dalloc(&c->d_SB[X], c->s_len[X]))) return rc;
  }
  return SQJK_OK;
}

ClassArgs base_args(sqjk_ctx* c) {
  ClassArgs a{};
  a.sh = c->d_sub; a.pp = c->d_pp; a.boys = c->d_boys; a.bscale = c->d_bscale;
  a.DJ = c->d_DJ; a.DA = c->d_DA; a.DB = c->d_DB; a.PM = c->d_PM;
  a.AJ = c->d_AJ; a.AKA = c->d_AKA; a.AKB = c->d_AKB; a.sch_out = c->d_sch;
  a.n = c->n; a.nref = c->nref;
  return a;
}

// Schwarz matrix on the device: the diagonal blocks (ab|ab) of every pair.
int compute_schwarz(sqjk_ctx* c) {
  const size_t nn = (size_t)c->n * c->n;
  hipLaunchKernelGGL(k_fill, dim3(grid_for((long long)nn, 256)), dim3(256), 0, c->st, c->d_sch, nn, 0.0);
  HIPCHK(hipGetLastError());
  for (int X = 0; X < kNPairClass; ++X) {
    if (c->pairs[X].empty()) continue;
    ClassArgs a = base_args(c);
    a.bra = a.ket = c->d_pairs[X];
    a.total = (long long)c->pairs[X].size();
    a.same = 1;
    ClassLauncher f = launcher_for(kPairL[X][0], kPairL[X][1], kPairL[X][0], kPairL[X][1]);
    if (!f || f(a, kSchwarz, 0, 0, c->st) < 0)
      return fail(SQJK_EHIP, "Schwarz kernel launch failed");
  }
  c->sch_basis.resize(nn);
  HIPCHK(hipMemcpyAsync(c->sch_basis.data(), c->d_sch, nn * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return SQJK_OK;
}

void finish_pairs(sqjk_ctx* c) {
  const int n = c->n, nr = c->nref;
  c->sch_ref.assign((size_t)nr * nr, 0.0);
  for (int P = 0; P < nr; ++P)
    for (int Q = 0; Q < nr; ++Q) {
      double r = 0.0;
      for (int p = c->ref[P].first; p < c->ref[P].first + c->ref[P].count; ++p)
        for (int q = c->ref[Q].first; q < c->ref[Q].first + c->ref[Q].count; ++q)
          r = std::max(r, c->sch_basis[(size_t)p * n + q]);
      c->sch_ref[(size_t)P * nr + Q] = r;
    }
  // groups: most primitive pairs first, Schwarz bound descending inside
  std::vector<PrimPair> ppT;
  ppT.reserve(c->pp.size());
  for (int X = 0; X < kNPairClass; ++X) {
    auto& prs = c->pairs[X];
    for (PairRec& pr : prs) pr.sch = c->sch_ref[(size_t)pr.pa * nr + pr.pb];
    std::stable_sort(prs.begin(), prs.end(), [](const PairRec& x, const PairRec& y) {
      return x.npp != y.npp ? x.npp > y.npp : x.sch > y.sch;
    });
    c->groups[X].clear();
    for (int i = 0; i < (int)prs.size();) {
      int j = i;
      while (j < (int)prs.size() && prs[j].npp == prs[i].npp) ++j;
      c->groups[X].push_back({i, j - i, prs[i].npp});
      // primitive-major copy: record (v, pair) at base + v * count + pair
      const int base = (int)ppT.size(), cnt = j - i;
      ppT.resize(ppT.size() + (size_t)cnt * prs[i].npp);
      for (int k = 0; k < cnt; ++k) {
        PairRec& pr = prs[i + k];
        for (int v = 0; v < pr.npp; ++v) ppT[(size_t)base + (size_t)v * cnt + k] = c->pp[pr.pp_off + v];
        pr.pp_off = base + k;
        pr.pp_stride = cnt;
      }
      i = j;
    }
  }
  c->pp.swap(ppT);
  c->plan_ok = c->hplan_ok = false;
}

// Rows and tiles of every launch for the cut-offs (thr, thr64): quartets with
// Schwarz product >= thr64 run in FP64, those in [thr, thr64) in FP32.
// Bra rows are taken in (first shell, Schwarz) order; a tile is up to
// kTileRows rows of one first shell that meet one chunk of 64 ket positions.
// Host only; make_plan uploads the result.  Rows per tile come from
// tile_rows(X, Y) below.
// Runs fn(t, nt) on nt host threads (t = 0..nt-1), nt = min(maxt, hardware
// threads); the calling thread takes t = 0.
template <class F>
void host_parallel(int maxt, F&& fn) {
  static const int hw = std::max(1u, std::thread::hardware_concurrency());
  const int nt = std::max(1, std::min(maxt, hw));
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(fn, t, nt);
  fn(0, nt);
  for (auto& th : pool) th.join();
}

// Primitive quartets per lane and tile above which a tile's ket primitive
// pairs are split over several tiles, for a class of nint integrals per
// quartet: about 576 integral-primitive products, between 8 (32 for classes
// of more than 81 integrals, whose digestion is dearer than their short
// primitive loops) and 64 primitive quartets.  The (ss|ss) blocks of two
// six-primitive core shells (1296 primitive quartets per lane) or the (pp|pp)
// blocks of contracted p shells otherwise set the length of a whole launch
// for small molecules.
int split_prims(int nint, int budget) {
  return std::max(nint > 81 ? 32 : 8, std::min(64, budget / std::max(1, nint)));
}
constexpr int kSplitBudget = 576;

// Bra rows per tile for the class pair (X, Y) (pair classes ss, ps, pp, ds,
// dp, dd).  Two rows keep J_cd and the a-side exchange of the lane's ket pair
// in registers across both rows and halve their atomics; that pays where the
// bra's second shell is s and for most dp bras (three for (ps|ps) and the ds
// bras), while (pp|pp), (pp|ps), (dp|pp) and the dd bras lose more to
// register pressure than they save.
// Measured per class at 1, 2 and 3 rows on MI355X, (H2O)32 6-31G*: 16.7 ms
// per build with one row everywhere, 14.8 ms with two rows everywhere.
// SQJK_TILE_ROWS overrides the table.
int default_tile_rows(int X, int Y) {
  static const int env = [] {
    const char* e = std::getenv("SQJK_TILE_ROWS");
    const int r = e ? std::atoi(e) : 0;
    return r < 0 ? 0 : (r > 256 ? 256 : r);
  }();
  if (env > 0) return env;
  static const int kRows[kNPairClass][kNPairClass] = {
      // Y: ss ps pp ds dp dd
      {2, 0, 0, 0, 0, 0},   // X = ss
      {2, 3, 0, 0, 0, 0},   // X = ps
      {2, 1, 1, 0, 0, 0},   // X = pp
      {3, 3, 3, 3, 0, 0},   // X = ds
      {2, 2, 1, 2, 2, 0},   // X = dp
      {1, 1, 1, 1, 1, 1}};  // X = dd
  return kRows[X][Y] > 0 ? kRows[X][Y] : 1;
}

void init_tuning(sqjk_ctx* c) {
  for (int X = 0; X < kNPairClass; ++X)
    for (int Y = 0; Y < kNPairClass; ++Y) {
      c->rows[X][Y] = default_tile_rows(X, Y);
      c->split_budget[X][Y] = kSplitBudget;
      c->grid_over[X][Y] = 0;
    }
}

void build_plan(sqjk_ctx* c, double thr, double thr64) {
  c->plan_ok = false;  // the device plan no longer holds the host plan
  c->hplan_ok = true;
  c->hplan_thr = thr;
  c->hplan_thr64 = thr64;
  c->launches.clear();
  c->h_lrange.clear();
  c->h_seg.clear();
  c->h_tile.clear();
  c->h_rbra.clear();
  c->h_rlo.clear();
  c->h_rhi.clear();
  std::memset(c->cls_q, 0, sizeof(c->cls_q));
  std::memset(c->cls_p, 0, sizeof(c->cls_p));
  c->plan_q = c->plan_q32 = 0;
  struct Row { int i, lo, hi; };
  std::vector<Row> rows, sel;
  std::vector<int> order, pass, pass64;
  for (int X = 0; X < kNPairClass; ++X)
    for (int Y = 0; Y <= X; ++Y)
      for (int fp = 0; fp < 2; ++fp) {
        // FP32 evaluation exists for classes of at most 81 integrals per
        // quartet; larger classes keep their whole range in FP64
        const int nint = sqmd::ncart(kPairL[X][0]) * sqmd::ncart(kPairL[X][1]) * sqmd::ncart(kPairL[Y][0]) *
                         sqmd::ncart(kPairL[Y][1]);
        const double thr64c = nint <= 81 ? thr64 : thr;
        if (fp == 1 && !(thr64c > thr)) continue;
        const auto &bra = c->pairs[X], &ket = c->pairs[Y];
        sqjk_ctx::Launch L{X, Y, fp, (int)c->h_seg.size(), (int)c->h_tile.size(), 0, 0};
        for (size_t a = 0; a < c->groups[X].size(); ++a) {
          // bra rows of group a in (first shell, Schwarz) order, shared by all ket groups
          {
            const auto& g1 = c->groups[X][a];
            order.resize(g1.count);
            for (int k = 0; k < g1.count; ++k) order[k] = g1.begin + k;
            std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return bra[x].a < bra[y].a; });
          }
          for (size_t b = 0; b < c->groups[Y].size(); ++b) {
            if (X == Y && b > a) continue;  // canonical: ket pair index <= bra pair index
            const auto &g1 = c->groups[X][a], &g2 = c->groups[Y][b];
            // kets of the group passing with bra Schwarz si: a prefix (sorted by sch)
            // kets passing with bra i: a prefix of the ket group (sorted by
            // Schwarz, descending), whose length cannot grow along the bra
            // group (also sorted descending), so one sweep per cut-off
            auto sweep = [&](double t, std::vector<int>& out) {
              out.resize(g1.count);
              int p = g2.count;
              for (int k = 0; k < g1.count; ++k) {
                const double si = bra[g1.begin + k].sch;
                while (p > 0 && !(si * ket[g2.begin + p - 1].sch >= t)) --p;
                out[k] = p;
              }
            };
            sweep(thr, pass);
            sweep(thr64c, pass64);
            rows.clear();
            long long acc = 0;
            for (int i : order) {
              int cnt = pass[i - g1.begin], c64 = std::min(pass64[i - g1.begin], cnt);
              if (X == Y && a == b) {
                const int lim = i - g1.begin + 1;
                cnt = std::min(cnt, lim);
                c64 = std::min(c64, lim);
              }
              const int lo = fp ? c64 : 0, hi = fp ? cnt : c64;
              if (hi <= lo) continue;
              rows.push_back({i, lo, hi});
              acc += hi - lo;
            }
            if (acc == 0) continue;
            const int seg = (int)c->h_seg.size();
            c->h_seg.push_back({g2.begin, g1.npp, g2.npp, 0});
            for (size_t r = 0; r < rows.size();) {
              size_t e = r;  // run of rows with the same first shell
              int mlo = rows[r].lo, mhi = rows[r].hi;
              while (e < rows.size() && bra[rows[e].i].a == bra[rows[r].i].a) {
                mlo = std::min(mlo, rows[e].lo);
                mhi = std::max(mhi, rows[e].hi);
                ++e;
              }
              for (int k0 = mlo; k0 < mhi; k0 += 64) {  // chunks start at the run's first ket
                sel.clear();
                for (size_t q = r; q < e; ++q)
                  if (rows[q].lo < k0 + 64 && rows[q].hi > k0) sel.push_back(rows[q]);
                const size_t tr = (size_t)c->rows[X][Y];
                for (size_t q = 0; q < sel.size(); q += tr) {
                  const int r0 = (int)c->h_rbra.size();
                  for (size_t u = q; u < std::min(sel.size(), q + tr); ++u) {
                    c->h_rbra.push_back(sel[u].i);
                    c->h_rlo.push_back(sel[u].lo);
                    c->h_rhi.push_back(sel[u].hi);
                  }
                  // a segment of more than kSplitPrims primitive quartets per
                  // quartet is cut along its ket primitive pairs
                  const int sp = split_prims(nint, c->split_budget[X][Y]);
                  const int nch = std::min(g2.npp, (g1.npp * g2.npp + sp - 1) / sp);
                  for (int ch = 0; ch < nch; ++ch)
                    c->h_tile.push_back({seg, r0, (int)c->h_rbra.size(), k0, g2.npp * ch / nch,
                                         g2.npp * (ch + 1) / nch});
                }
              }
              r = e;
            }
            L.quartets += acc;
            c->cls_p[X][Y] += acc * g1.npp * g2.npp;
          }
        }
        L.ntile = (long long)c->h_tile.size() - L.tile0;
        if (L.ntile == 0) continue;
        if (c->stage_chunked) {
          // tiles by the chunk of their first ket pair, one range per chunk
          const int cl = c->chunk_len[Y];
          auto chunk_of = [&](const Tile& t) { return (c->h_seg[t.seg].ket0 + t.k0) / cl; };
          std::stable_sort(c->h_tile.begin() + L.tile0, c->h_tile.end(),
                           [&](const Tile& p, const Tile& q) { return chunk_of(p) < chunk_of(q); });
          L.lr0 = (int)c->h_lrange.size();
          for (int t = L.tile0; t < (int)c->h_tile.size();) {
            int e = t;
            const int ch = chunk_of(c->h_tile[t]);
            while (e < (int)c->h_tile.size() && chunk_of(c->h_tile[e]) == ch) ++e;
            c->h_lrange.push_back({ch, t - L.tile0, e - L.tile0});
            t = e;
          }
          L.nlr = (int)c->h_lrange.size() - L.lr0;
        }
        c->launches.push_back(L);
        c->cls_q[X][Y] += L.quartets;
        c->plan_q += L.quartets;
        if (fp) c->plan_q32 += L.quartets;
      }
}

// Host plan for the cut-offs (thr, thr64), made only when they changed.
void host_plan(sqjk_ctx* c, double thr, double thr64) {
  if (!(c->hplan_ok && c->hplan_thr == thr && c->hplan_thr64 == thr64)) build_plan(c, thr, thr64);
}

int make_plan(sqjk_ctx* c, double thr, double thr64) {
  if (c->plan_ok && c->plan_thr == thr && c->plan_thr64 == thr64) return SQJK_OK;
  host_plan(c, thr, thr64);
  if (c->h_rbra.size() > (size_t)INT32_MAX) return fail(SQJK_ENOMEM, "work plan too large");
  auto grow = [&](auto** d, size_t& cap, size_t need) -> int {
    if (cap >= need && *d) return SQJK_OK;
    if (*d) { block_put(*d); *d = nullptr; }
    cap = std::max<size_t>(need, 1024) * 3 / 2;
    return dalloc(d, cap);
  };
  int rc;
  const size_t nrow = c->h_rbra.size();
  if ((rc = grow(&c->d_seg, c->seg_cap, c->h_seg.size()))) return rc;
  if ((rc = grow(&c->d_tile, c->tile_cap, c->h_tile.size()))) return rc;
  if (c->row_cap < nrow || !c->d_rbra) {
    int** rowbufs[] = {&c->d_rbra, &c->d_rlo, &c->d_rhi};
    for (int** d : rowbufs) {
      if (*d) block_put(*d);
      *d = nullptr;
    }
    c->row_cap = std::max<size_t>(nrow, 1024) * 3 / 2;
    for (int** d : rowbufs)
      if ((rc = dalloc(d, c->row_cap))) return rc;
  }
  if (!c->h_tile.empty()) {
    HIPCHK(hipMemcpyAsync(c->d_seg, c->h_seg.data(), c->h_seg.size() * sizeof(Segment), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->d_tile, c->h_tile.data(), c->h_tile.size() * sizeof(Tile), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->d_rbra, c->h_rbra.data(), nrow * sizeof(int), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->d_rlo, c->h_rlo.data(), nrow * sizeof(int), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->d_rhi, c->h_rhi.data(), nrow * sizeof(int), hipMemcpyHostToDevice, c->st));
  }
  c->plan_ok = true;
  c->plan_thr = thr;
  c->plan_thr64 = thr64;
  return SQJK_OK;
}

}  // namespace

extern "C" {

const char* sqjk_errmsg(void) { return g_err.c_str(); }

int sqjk_create(int nBasis, const GTOBasis_t* gto, const double* schwarz_basis, int device, sqjk_ctx** out) {
  if (!out || nBasis <= 0 || !gto) return fail(SQJK_EARG, "sqjk_create: invalid arguments");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    return fail(SQJK_ENODEV, "no HIP device is visible; this library has no CPU path");
  if (device < 0 || device >= ndev) return fail(SQJK_EARG, "sqjk_create: device ordinal out of range");
  HIPCHK(hipSetDevice(device));
  sqjk_ctx* c = new sqjk_ctx();
  c->device = device;
  c->n = nBasis;
  init_tuning(c);
  if (const char* e = std::getenv("SQJK_DD_KERNEL")) c->no_wd = std::string(e) == "wave";
  c->stage_budget = stage_budget_bytes();
  std::vector<double> pexp, pcoef;
  int rc = plan_shells(c, nBasis, gto, pexp, pcoef);
  if (rc) { delete c; return rc; }
  plan_pairs(c, pexp, pcoef);
  c->pexp = pexp;
  c->pcoef = pcoef;
  if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    sqjk_destroy(c);
    return fail(SQJK_EHIP, "stream/event creation failed");
  }
  if ((rc = upload_static(c)) || (rc = upload_pairs(c))) { sqjk_destroy(c); return rc; }
  if (schwarz_basis) {
    c->sch_basis.assign(schwarz_basis, schwarz_basis + (size_t)nBasis * nBasis);
  } else if ((rc = compute_schwarz(c))) {
    sqjk_destroy(c);
    return rc;
  }
  finish_pairs(c);
  if ((rc = upload_pairs(c))) { sqjk_destroy(c); return rc; }
  // primitive data in the grouped, primitive-major order (same size)
  if (hipMemcpyAsync(c->d_pp, c->pp.data(), c->pp.size() * sizeof(PrimPair), hipMemcpyHostToDevice, c->st) !=
      hipSuccess) {
    sqjk_destroy(c);
    return fail(SQJK_EHIP, "primitive pair upload failed");
  }
  if ((rc = build_herm_tables(c))) { sqjk_destroy(c); return rc; }
  plan_stage(c);
  if (hipStreamSynchronize(c->st) != hipSuccess) { sqjk_destroy(c); return fail(SQJK_EHIP, "sync failed"); }
  *out = c;
  return SQJK_OK;
}

long long sqjk_footprint_bytes(int nBasis, const GTOBasis_t* gto, int uhf, int with_scf) {
  if (nBasis <= 0 || !gto) return fail(SQJK_EARG, "sqjk_footprint_bytes: invalid arguments");
  sqjk_ctx t;
  std::vector<double> pexp, pcoef;
  if (int rc = plan_shells(&t, nBasis, gto, pexp, pcoef)) return rc;
  plan_pairs(&t, pexp, pcoef);
  const long long nn = (long long)nBasis * nBasis;
  long long b = (long long)t.sub.size() * sizeof(SubShell) + (long long)t.pp.size() * sizeof(PrimPair) +
                (long long)kBoysNT * kBoysCols * 8 + (long long)nBasis * 8 + (long long)t.nref * t.nref * 8;
  b += 10 * nn * 8;  // densities, half-accumulators, results, Schwarz
  t.n = nBasis;
  t.stage_budget = stage_budget_bytes();
  plan_stage(&t);
  for (int X = 0; X < kNPairClass; ++X) {
    const long long np = (long long)t.pairs[X].size();
    const int nab = sqmd::ncart(kPairL[X][0]) * sqmd::ncart(kPairL[X][1]);
    b += np * sizeof(PairRec) + np * nab * 8;
    if (!t.stage_chunked) b += (uhf ? 2 : 1) * (long long)t.s_len[X] * 8;
  }
  if (t.stage_chunked) b += (uhf ? 2 : 1) * (long long)t.sc_len * 8;
  for (const PairRec& pr : t.pairs[kNPairClass - 1]) b += (long long)pr.npp * HermTabLayout<2, 2>::SIZE * 8;
  if (with_scf) b += 38 * nn * 8;  // SCF matrices, DIIS history, eigen-solver work
  return b + b / 8;                 // work plan and allocator slack
}

int sqjk_update_geometry(sqjk_ctx* c, int nBasis, const GTOBasis_t* gto) {
  if (!c || !gto) return fail(SQJK_EARG, "sqjk_update_geometry: null argument");
  if (nBasis != c->n) return fail(SQJK_EARG, "sqjk_update_geometry: the basis has a different number of functions");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->st));
  // the new basis must have the same shells, primitives and coefficients;
  // only the centres may move
  sqjk_ctx t;
  std::vector<double> pexp, pcoef;
  int rc = plan_shells(&t, c->n, gto, pexp, pcoef);
  if (rc) return rc;
  bool same = t.nref == c->nref && t.nsub == c->nsub;
  for (int r = 0; same && r < c->nref; ++r) same = t.ref[r].first == c->ref[r].first && t.ref[r].count == c->ref[r].count;
  for (int i = 0; same && i < c->nsub; ++i) {
    const SubShell &a = t.sub[i], &b = c->sub[i];
    same = a.l == b.l && a.nprim == b.nprim && a.bf_off == b.bf_off && a.parent == b.parent;
  }
  for (int f = 0; same && f < c->n; ++f) same = t.bscale[f] == c->bscale[f];
  same = same && pexp == c->pexp && pcoef == c->pcoef;
  if (!same) return fail(SQJK_EARG, "sqjk_update_geometry: the basis differs in more than its centres");
  // new centres: primitive pairs (pair order), Schwarz on the device, then
  // groups and the primitive-major layout as at creation; every buffer and
  // the stream are kept, the work plan is rebuilt at the next build
  c->sub = t.sub;
  plan_pairs(c, pexp, pcoef);
  HIPCHK(hipMemcpyAsync(c->d_sub, c->sub.data(), c->sub.size() * sizeof(SubShell), hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->d_pp, c->pp.data(), c->pp.size() * sizeof(PrimPair), hipMemcpyHostToDevice, c->st));
  for (int X = 0; X < kNPairClass; ++X)
    if (!c->pairs[X].empty())
      HIPCHK(hipMemcpyAsync(c->d_pairs[X], c->pairs[X].data(), c->pairs[X].size() * sizeof(PairRec),
                            hipMemcpyHostToDevice, c->st));
  if ((rc = compute_schwarz(c))) return rc;
  finish_pairs(c);
  for (int X = 0; X < kNPairClass; ++X)
    if (!c->pairs[X].empty())
      HIPCHK(hipMemcpyAsync(c->d_pairs[X], c->pairs[X].data(), c->pairs[X].size() * sizeof(PairRec),
                            hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->d_pp, c->pp.data(), c->pp.size() * sizeof(PrimPair), hipMemcpyHostToDevice, c->st));
  if ((rc = build_herm_tables(c))) return rc;
  HIPCHK(hipStreamSynchronize(c->st));
  c->plan_ok = c->hplan_ok = false;
  c->built = false;
  return SQJK_OK;
}

int sqjk_one_electron(int nBasis, const GTOBasis_t* gto, const Molecule_t* mol, int device, double* T, double* V,
                      double* S) {
  if (nBasis <= 0 || !gto || !mol || mol->nAtom < 0) return fail(SQJK_EARG, "sqjk_one_electron: invalid arguments");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    return fail(SQJK_ENODEV, "no HIP device is visible; this library has no CPU path");
  if (device < 0 || device >= ndev) return fail(SQJK_EARG, "sqjk_one_electron: device ordinal out of range");
  HIPCHK(hipSetDevice(device));
  sqjk_ctx c;
  c.n = nBasis;
  std::vector<double> pexp, pcoef;
  int rc = plan_shells(&c, nBasis, gto, pexp, pcoef);
  if (rc) return rc;
  plan_pairs(&c, pexp, pcoef);
  std::vector<double> atoms((size_t)4 * std::max(mol->nAtom, 1));
  for (int k = 0; k < mol->nAtom; ++k) {
    atoms[4 * k + 0] = mol->x[k]; atoms[4 * k + 1] = mol->y[k]; atoms[4 * k + 2] = mol->z[k];
    atoms[4 * k + 3] = (double)mol->Z[k];
  }
  const size_t nn = (size_t)nBasis * nBasis;
  // a stream per host thread, so that one-electron matrices of molecules
  // served side by side do not wait for each other's work
  thread_local hipStream_t st = nullptr;
  thread_local int st_dev = -1;
  if (!st || st_dev != device) {
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    st_dev = device;
  }
  std::vector<void*> bufs;
  bool up_ok = true;
  auto up = [&](const void* h, size_t bytes) -> void* {
    void* d = nullptr;
    if (!(d = block_get(std::max<size_t>(bytes, 8), false))) { up_ok = false; return nullptr; }
    bufs.push_back(d);
    if (h && bytes && hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, st) != hipSuccess) up_ok = false;
    return d;
  };
  auto release = [&] {
    (void)hipStreamSynchronize(st);
    for (void* d : bufs) block_put(d);
  };
  const std::vector<double>& boys = boys_table();
  OneEArgs a{};
  a.sh = (const SubShell*)up(c.sub.data(), c.sub.size() * sizeof(SubShell));
  a.pexp = (const double*)up(pexp.data(), pexp.size() * 8);
  a.pcoef = (const double*)up(pcoef.data(), pcoef.size() * 8);
  a.bscale = (const double*)up(c.bscale.data(), c.bscale.size() * 8);
  a.atom = (const double*)up(atoms.data(), atoms.size() * 8);
  a.boys = (const double*)up(boys.data(), boys.size() * 8);
  a.T = T ? (double*)up(nullptr, nn * 8) : nullptr;
  a.V = V ? (double*)up(nullptr, nn * 8) : nullptr;
  a.S = S ? (double*)up(nullptr, nn * 8) : nullptr;
  a.natom = mol->nAtom;
  a.n = nBasis;
  if (!up_ok) {
    release();
    return fail(SQJK_ENOMEM, "sqjk_one_electron: device allocation or upload failed");
  }
  for (int X = 0; X < kNPairClass; ++X) {
    a.pr = (const PairRec*)up(c.pairs[X].data(), c.pairs[X].size() * sizeof(PairRec));
    a.np = (int)c.pairs[X].size();
    if (!up_ok || launch_one_electron(X, a, st) != 0) {
      release();
      return fail(SQJK_EHIP, "sqjk_one_electron: kernel launch failed");
    }
  }
  hipError_t e = hipSuccess;
  if (T) e = hipMemcpyAsync(T, a.T, nn * 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess && V) e = hipMemcpyAsync(V, a.V, nn * 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess && S) e = hipMemcpyAsync(S, a.S, nn * 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  release();
  if (e != hipSuccess) return fail(SQJK_EHIP, std::string("sqjk_one_electron: ") + hipGetErrorString(e));
  return SQJK_OK;
}

}  // extern "C"

namespace {

// Device buffers of one build: the work plan and the pair-indexed digestion
// buffers of every class (exchange staging per spin).
struct BuildBufs {
  Segment* seg = nullptr;
  Tile* tile = nullptr;
  int *rbra = nullptr, *rlo = nullptr, *rhi = nullptr;
  double* Jp[kNPairClass] = {};
  double* SA[kNPairClass] = {};
  double* SB[kNPairClass] = {};
};

// The context's own buffers (make_plan and ensure_stage have run).
BuildBufs context_bufs(const sqjk_ctx* c) {
  BuildBufs B;
  B.seg = c->d_seg; B.tile = c->d_tile; B.rbra = c->d_rbra; B.rlo = c->d_rlo; B.rhi = c->d_rhi;
  for (int X = 0; X < kNPairClass; ++X) { B.Jp[X] = c->d_Jp[X]; B.SA[X] = c->d_SA[X]; B.SB[X] = c->d_SB[X]; }
  return B;
}

// The same buffers carved out of a caller-provided region of `cap` bytes
// (256-byte aligned pieces), with the host plan for the cut-offs uploaded
// into it.  Nothing is allocated on the device.
int carve_bufs(sqjk_ctx* c, double thr, double thr64, bool uhf, char* region, size_t cap, BuildBufs* B) {
  host_plan(c, thr, thr64);
  stage_lengths(c);
  const size_t nrow = c->h_rbra.size();
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off += (bytes + 255) / 256 * 256;
    return at;
  };
  const size_t o_seg = take(c->h_seg.size() * sizeof(Segment)), o_tile = take(c->h_tile.size() * sizeof(Tile));
  const size_t o_rbra = take(nrow * sizeof(int)), o_rlo = take(nrow * sizeof(int)), o_rhi = take(nrow * sizeof(int));
  size_t o_jp[kNPairClass], o_sa[kNPairClass], o_sb[kNPairClass];
  for (int X = 0; X < kNPairClass; ++X) {
    o_jp[X] = take(c->jp_len[X] * sizeof(double));
    o_sa[X] = take(c->s_len[X] * sizeof(double));
    o_sb[X] = uhf ? take(c->s_len[X] * sizeof(double)) : 0;
  }
  if (off > cap)
    return fail(SQJK_ENOMEM, "sqjk_build_device_noalloc: this build needs " + std::to_string(off >> 20) +
                             " MiB of scratch, above the " + std::to_string(cap >> 20) + " MiB the library holds");
  B->seg = reinterpret_cast<Segment*>(region + o_seg);
  B->tile = reinterpret_cast<Tile*>(region + o_tile);
  B->rbra = reinterpret_cast<int*>(region + o_rbra);
  B->rlo = reinterpret_cast<int*>(region + o_rlo);
  B->rhi = reinterpret_cast<int*>(region + o_rhi);
  for (int X = 0; X < kNPairClass; ++X) {
    B->Jp[X] = reinterpret_cast<double*>(region + o_jp[X]);
    B->SA[X] = reinterpret_cast<double*>(region + o_sa[X]);
    B->SB[X] = uhf ? reinterpret_cast<double*>(region + o_sb[X]) : nullptr;
  }
  if (!c->h_tile.empty()) {
    HIPCHK(hipMemcpyAsync(B->seg, c->h_seg.data(), c->h_seg.size() * sizeof(Segment), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(B->tile, c->h_tile.data(), c->h_tile.size() * sizeof(Tile), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(B->rbra, c->h_rbra.data(), nrow * sizeof(int), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(B->rlo, c->h_rlo.data(), nrow * sizeof(int), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(B->rhi, c->h_rhi.data(), nrow * sizeof(int), hipMemcpyHostToDevice, c->st));
  }
  return SQJK_OK;
}

// The device part of a build.  The densities are in c->d_DA, c->d_DB; the
// results J and -K (full, symmetric) are written to oJ, oKA and, for UHF,
// oKB.  Everything is queued on the context's stream; nothing waits.  The
// work plan and digestion buffers are the context's own (allocated on first
// use) unless `region` (cap bytes of device memory) is given: then they are
// carved out of it, nothing is allocated, and per-class profiling is off.
int build_on_device(sqjk_ctx* c, double fixedCutoff, double scfCutoff, int uhf, double* oJ, double* oKA,
                    double* oKB, char* region = nullptr, size_t cap = 0) {
  const int n = c->n, nr = c->nref;
  const size_t nn = (size_t)n * n;
  if ((size_t)n * sizeof(double) > 64 * 1024) return fail(SQJK_EARG, "nBasis above 8192 is not supported");
  hipLaunchKernelGGL(k_density_sum, dim3(grid_for((long long)nn, 256)), dim3(256), 0, c->st, c->d_DA, c->d_DB,
                     c->d_DJ, nn);
  hipLaunchKernelGGL(k_pair_max, dim3(grid_for((long long)nr * nr, 256)), dim3(256), 0, c->st, c->d_DJ, c->d_DA,
                     c->d_DB, c->d_ref, nr, n, c->d_PM);
  HIPCHK(hipGetLastError());

  // work plan (rows/segments), rebuilt only when the cut-offs change
  const double thr = std::max(fixedCutoff, scfCutoff);
  const double thr64 = std::max(thr, c->mixed_thr);
  BuildBufs B;
  int rc;
  if (region) {
    if ((rc = carve_bufs(c, thr, thr64, uhf != 0, region, cap, &B))) return rc;
  } else {
    if ((rc = make_plan(c, thr, thr64))) return rc;
    if ((rc = ensure_stage(c, uhf))) return rc;
    B = context_bufs(c);
  }
  const bool prof = c->profile && !region;
  HIPCHK(hipEventRecord(c->ev0, c->st));
  const int fill_grid = grid_for((long long)nn, 256);
  hipLaunchKernelGGL(k_fill, dim3(fill_grid), dim3(256), 0, c->st, c->d_AJ, nn, 0.0);
  hipLaunchKernelGGL(k_fill, dim3(fill_grid), dim3(256), 0, c->st, c->d_AKA, nn, 0.0);
  if (uhf) hipLaunchKernelGGL(k_fill, dim3(fill_grid), dim3(256), 0, c->st, c->d_AKB, nn, 0.0);
  for (int X = 0; X < kNPairClass; ++X) {
    if (c->pairs[X].empty()) continue;
    hipLaunchKernelGGL(k_fill, dim3(grid_for((long long)c->jp_len[X], 256)), dim3(256), 0, c->st, B.Jp[X],
                       c->jp_len[X], 0.0);
    if (c->stage_chunked && !region) continue;  // the chunk buffer is filled per chunk
    const int sg = std::max(1, (int)std::min<size_t>((c->s_len[X] + 255) / 256, 16384));
    hipLaunchKernelGGL(k_fill, dim3(sg), dim3(256), 0, c->st, B.SA[X], c->s_len[X], 0.0);
    if (uhf) hipLaunchKernelGGL(k_fill, dim3(sg), dim3(256), 0, c->st, B.SB[X], c->s_len[X], 0.0);
  }
  HIPCHK(hipGetLastError());
  int launches = 0;
  const bool chunked = c->stage_chunked && !region;
  c->prof_launch.clear();
  // one class launch over the tiles [tb, te) of launch li, with the
  // exchange staging SA/SB holding the ket pairs from s_k0 with stride s_np
  auto launch = [&](size_t li, long long tb, long long te, double* SA, double* SB, int s_k0, int s_np) -> int {
    const sqjk_ctx::Launch& L = c->launches[li];
    size_t ev = 0;
    if (prof) {
      ev = 2 * c->prof_launch.size();
      c->prof_launch.push_back((int)li);
      while (c->prof_ev.size() < ev + 2) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        c->prof_ev.push_back(e);
      }
      HIPCHK(hipEventRecord(c->prof_ev[ev], c->st));
    }
    ClassArgs a = base_args(c);
    a.bra = c->d_pairs[L.X];
    a.ket = c->d_pairs[L.Y];
    a.seg = B.seg;
    a.tile = B.tile + L.tile0 + tb;
    a.rbra = B.rbra;
    a.rlo = B.rlo;
    a.rhi = B.rhi;
    a.same = (L.X == L.Y);
    a.total = te - tb;
    a.fixed_cut = fixedCutoff;
    a.JpB = B.Jp[L.X];
    a.JpK = B.Jp[L.Y];
    a.SA = SA;
    a.SB = SB;
    a.s_k0 = s_k0;
    a.s_np = s_np;
    a.npb = (int)c->pairs[L.X].size();
    a.npk = (int)c->pairs[L.Y].size();
    if (L.X == kNPairClass - 1 && L.Y == kNPairClass - 1 && c->d_herm && !c->no_wd) {
      a.herm = c->d_herm;
      a.herm_base = c->herm_base;
    }
    ClassLauncher f = launcher_for(kPairL[L.X][0], kPairL[L.X][1], kPairL[L.Y][0], kPairL[L.Y][1]);
    const int k = f ? f(a, uhf ? kUHF : kRHF, L.fp32, -c->grid_over[L.X][L.Y], c->st) : -1;
    if (k < 0) return fail(SQJK_EHIP, "class kernel launch failed");
    launches += k;
    if (prof) HIPCHK(hipEventRecord(c->prof_ev[ev + 1], c->st));
    return SQJK_OK;
  };
  auto stage_of = [&](int Y, double* S, double* SBp, int p0, int np, int cnt) {
    KStage R{};
    R.S[Y] = S;
    R.pr[Y] = c->d_pairs[Y];
    R.np[Y] = np;
    R.p0[Y] = p0;
    R.cnt[Y] = cnt;
    R.nc[Y] = sqmd::ncart(kPairL[Y][0]);
    R.ncomp[Y] = sqmd::ncart(kPairL[Y][0]) + sqmd::ncart(kPairL[Y][1]);
    (void)SBp;
    return R;
  };
  if (!chunked) {
    for (size_t li = 0; li < c->launches.size(); ++li) {
      const sqjk_ctx::Launch& L = c->launches[li];
      if ((rc = launch(li, 0, L.ntile, B.SA[L.Y], B.SB[L.Y], 0, (int)c->pairs[L.Y].size()))) return rc;
    }
  } else {
    // ket class by ket class, chunk by chunk: fill, launch the tiles whose
    // first ket pair is in the chunk, fold into A_K
    for (int Y = 0; Y < kNPairClass; ++Y) {
      const int np = (int)c->pairs[Y].size();
      if (np == 0) continue;
      const size_t ncomp = sqmd::ncart(kPairL[Y][0]) + sqmd::ncart(kPairL[Y][1]);
      const int cl = c->chunk_len[Y], stride = c->chunk_stride[Y];
      const size_t len = (size_t)n * ncomp * stride;
      const int sg = std::max(1, (int)std::min<size_t>((len + 255) / 256, 16384));
      for (int ch = 0; ch < c->chunk_cnt[Y]; ++ch) {
        bool any = false;
        for (const sqjk_ctx::Launch& L : c->launches)
          if (L.Y == Y)
            for (int r = L.lr0; r < L.lr0 + L.nlr; ++r) any = any || c->h_lrange[r][0] == ch;
        if (!any) continue;
        hipLaunchKernelGGL(k_fill, dim3(sg), dim3(256), 0, c->st, c->d_SC, len, 0.0);
        if (uhf) hipLaunchKernelGGL(k_fill, dim3(sg), dim3(256), 0, c->st, c->d_SCB, len, 0.0);
        HIPCHK(hipGetLastError());
        for (size_t li = 0; li < c->launches.size(); ++li) {
          const sqjk_ctx::Launch& L = c->launches[li];
          if (L.Y != Y) continue;
          for (int r = L.lr0; r < L.lr0 + L.nlr; ++r) {
            const auto& lr = c->h_lrange[r];
            if (lr[0] != ch) continue;
            if ((rc = launch(li, lr[1], lr[2], c->d_SC, c->d_SCB, ch * cl, stride))) return rc;
          }
        }
        const int p0 = ch * cl, cnt = std::min(stride, np - p0);
        const KStage RA = stage_of(Y, c->d_SC, nullptr, p0, stride, cnt);
        hipLaunchKernelGGL(k_reduce_k, dim3(n), dim3(kReduceKBlock), (size_t)n * sizeof(double), c->st, RA,
                           c->d_sub, c->d_AKA, n);
        if (uhf) {
          const KStage RB = stage_of(Y, c->d_SCB, nullptr, p0, stride, cnt);
          hipLaunchKernelGGL(k_reduce_k, dim3(n), dim3(kReduceKBlock), (size_t)n * sizeof(double), c->st, RB,
                             c->d_sub, c->d_AKB, n);
        }
        HIPCHK(hipGetLastError());
      }
    }
  }
  // fold the pair-indexed buffers into the half-accumulators
  KStage RA{}, RB{};
  for (int X = 0; X < kNPairClass; ++X) {
    const int np = (int)c->pairs[X].size();
    if (np == 0) continue;
    const int nb = sqmd::ncart(kPairL[X][1]), nab = sqmd::ncart(kPairL[X][0]) * nb;
    hipLaunchKernelGGL(k_reduce_j, dim3(grid_for((long long)np * nab, 256)), dim3(256), 0, c->st, B.Jp[X],
                       c->d_pairs[X], c->d_sub, np, nb, nab, c->d_AJ, n);
    RA.S[X] = B.SA[X];
    RB.S[X] = B.SB[X];
    RA.pr[X] = RB.pr[X] = c->d_pairs[X];
    RA.np[X] = RB.np[X] = np;
    RA.cnt[X] = RB.cnt[X] = np;
    RA.nc[X] = RB.nc[X] = sqmd::ncart(kPairL[X][0]);
    RA.ncomp[X] = RB.ncomp[X] = sqmd::ncart(kPairL[X][0]) + nb;
  }
  if (!chunked) {
    hipLaunchKernelGGL(k_reduce_k, dim3(n), dim3(kReduceKBlock), (size_t)n * sizeof(double), c->st, RA, c->d_sub,
                       c->d_AKA, n);
    if (uhf)
      hipLaunchKernelGGL(k_reduce_k, dim3(n), dim3(kReduceKBlock), (size_t)n * sizeof(double), c->st, RB, c->d_sub,
                         c->d_AKB, n);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(c->ev1, c->st));
  hipLaunchKernelGGL(k_finalize, dim3(fill_grid), dim3(256), 0, c->st, c->d_AJ, c->d_AKA, c->d_AKB, oJ, oKA, oKB,
                     n, uhf);
  HIPCHK(hipGetLastError());
  c->last.quartets = c->plan_q;
  c->last.quartets_fp32 = c->plan_q32;
  c->last.launches = launches;
  c->last_fixed = fixedCutoff;
  c->built = true;
  c->timing_pending = true;
  return SQJK_OK;
}

// Kernel time of the last build from its events (the build must have
// completed), and the per-class event times in profile mode.
int collect_times(sqjk_ctx* c) {
  if (!c->timing_pending) return SQJK_OK;
  float ms = 0.f;
  HIPCHK(hipEventSynchronize(c->ev1));
  HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  std::memset(c->cls_ms, 0, sizeof(c->cls_ms));
  if (c->profile)
    for (size_t k = 0; k < c->prof_launch.size(); ++k) {
      float lm = 0.f;
      HIPCHK(hipEventElapsedTime(&lm, c->prof_ev[2 * k], c->prof_ev[2 * k + 1]));
      const sqjk_ctx::Launch& L = c->launches[c->prof_launch[k]];
      c->cls_ms[L.X][L.Y] += lm;
    }
  c->last.kernel_ms = ms;
  c->timing_pending = false;
  return SQJK_OK;
}

}  // namespace

extern "C" {

int sqjk_build(sqjk_ctx* c, const double* PA, const double* PB, double fixedCutoff, double scfCutoff,
               int restricted, double* JT, double* KA, double* KB) {
  if (!c || !PA || !PB || !JT || !KA || !KB) return fail(SQJK_EARG, "sqjk_build: null argument");
  const auto t0 = std::chrono::steady_clock::now();
  HIPCHK(hipSetDevice(c->device));
  const int n = c->n;
  const size_t nn = (size_t)n * n;
  const int uhf = restricted ? 0 : 1;
  // densities: J uses PA+PB, K uses PA (and PB).  Both go up through the
  // pinned staging buffers (the previous build has synchronised, so they are
  // free); D_J and the screening maxima are formed on the device.
  // The copy of P_B into its staging buffer overlaps the upload of P_A; a
  // caller passing the same matrix twice gets one upload and a device copy.
  auto stage = [&](double* dst, const double* src) {
    host_parallel(8, [&](int t, int nt) {
      const size_t e0 = nn * t / nt, e1 = nn * (t + 1) / nt;
      std::memcpy(dst + e0, src + e0, (e1 - e0) * sizeof(double));
    });
  };
  stage(c->hp_A, PA);
  HIPCHK(hipMemcpyAsync(c->d_DA, c->hp_A, nn * 8, hipMemcpyHostToDevice, c->st));
  if (PB == PA) {
    HIPCHK(hipMemcpyAsync(c->d_DB, c->d_DA, nn * 8, hipMemcpyDeviceToDevice, c->st));
  } else {
    stage(c->hp_B, PB);
    HIPCHK(hipMemcpyAsync(c->d_DB, c->hp_B, nn * 8, hipMemcpyHostToDevice, c->st));
  }
  int rc = build_on_device(c, fixedCutoff, scfCutoff, uhf, c->d_J, c->d_KA, c->d_KB);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(c->hp_J, c->d_J, nn * 8, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(c->hp_KA, c->d_KA, nn * 8, hipMemcpyDeviceToHost, c->st));
  if (uhf) HIPCHK(hipMemcpyAsync(c->hp_KB, c->d_KB, nn * 8, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  // reference semantics: accumulate into the lower triangle, KB := KA if
  // restricted, upper := lower; done in 64 x 64 blocks so that the mirrored
  // writes stay in cache
  const double* hK2 = uhf ? c->hp_KB : c->hp_KA;
  constexpr int kBlk = 64;
  const int nblk = (n + kBlk - 1) / kBlk;
  // block rows dealt out round-robin; the (p0, q0) blocks and their mirrors
  // are disjoint, so threads never write the same element
  host_parallel(8, [&](int t, int nt) {
  for (int pb = t; pb < nblk; pb += nt)
    for (int q0 = 0, p0 = pb * kBlk; q0 <= p0; q0 += kBlk) {
      const int p1 = std::min(n, p0 + kBlk), q1 = std::min(n, q0 + kBlk);
      for (int p = p0; p < p1; ++p)
        for (int q = q0; q < std::min(q1, p + 1); ++q) {
          const size_t e = (size_t)p * n + q, tr = (size_t)q * n + p;
          const double j = JT[e] + c->hp_J[e], ka = KA[e] + c->hp_KA[e];
          const double kb = uhf ? KB[e] + hK2[e] : ka;
          JT[e] = JT[tr] = j;
          KA[e] = KA[tr] = ka;
          KB[e] = KB[tr] = kb;
        }
    }
  });
  if ((rc = collect_times(c))) return rc;
  c->last.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return SQJK_OK;
}

int sqjk_build_device(sqjk_ctx* c, const double* dPA, const double* dPB, double fixedCutoff, double scfCutoff,
                      int restricted, double* dJ, double* dKA, double* dKB) {
  if (!c || !dPA || !dPB || !dJ || !dKA || !dKB) return fail(SQJK_EARG, "sqjk_build_device: null argument");
  HIPCHK(hipSetDevice(c->device));
  const size_t nn = (size_t)c->n * c->n;
  const int uhf = restricted ? 0 : 1;
  HIPCHK(hipMemcpyAsync(c->d_DA, dPA, nn * 8, hipMemcpyDeviceToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->d_DB, dPB, nn * 8, hipMemcpyDeviceToDevice, c->st));
  int rc = build_on_device(c, fixedCutoff, scfCutoff, uhf, dJ, dKA, dKB);
  if (rc) return rc;
  if (!uhf && dKB != dKA) HIPCHK(hipMemcpyAsync(dKB, dKA, nn * 8, hipMemcpyDeviceToDevice, c->st));
  c->last.total_ms = 0.0;
  return SQJK_OK;
}

}  // extern "C"

// Device memory of the allocation-free build (sqjk_build_device_noalloc):
// its work plan and digestion buffers are carved out of this array.  Every
// GPU the library runs on holds its own copy, at its own address, for the
// life of the process.  In host code `scratch` names a host-side stand-in,
// so kernels are given the device address from hipGetSymbolAddress, kept per
// device below.
constexpr size_t kScratchBytes = SQJK_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's scratch array (the device must be current).
int scratch_region(int device, char** out) {
  if (device < 0 || device >= kMaxDevices) return fail(SQJK_EARG, "device ordinal out of range for the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(SQJK_EHIP, std::string("scratch address lookup failed: ") + hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<char*>(p);
  }
  *out = g_scratch_ptr[device];
  return SQJK_OK;
}
}  // namespace

extern "C" {

int sqjk_build_device_noalloc(sqjk_ctx* c, const double* dPA, const double* dPB, double fixedCutoff,
                              double scfCutoff, int restricted, double* dJ, double* dKA, double* dKB) {
  if (!c || !dPA || !dPB || !dJ || !dKA || !dKB) return fail(SQJK_EARG, "sqjk_build_device_noalloc: null argument");
  HIPCHK(hipSetDevice(c->device));
  char* region = nullptr;
  int rc = scratch_region(c->device, &region);
  if (rc) return rc;
  const size_t nn = (size_t)c->n * c->n;
  const int uhf = restricted ? 0 : 1;
  HIPCHK(hipMemcpyAsync(c->d_DA, dPA, nn * 8, hipMemcpyDeviceToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->d_DB, dPB, nn * 8, hipMemcpyDeviceToDevice, c->st));
  if ((rc = build_on_device(c, fixedCutoff, scfCutoff, uhf, dJ, dKA, dKB, region, kScratchBytes))) return rc;
  if (!uhf && dKB != dKA) HIPCHK(hipMemcpyAsync(dKB, dKA, nn * 8, hipMemcpyDeviceToDevice, c->st));
  c->last.total_ms = 0.0;
  return SQJK_OK;
}

// Two-electron part of the energy gradient (src/grad.c:1734-2090).  The J/K
// work plan of the context, made for a Schwarz cut-off of cutoff/16 (the
// largest density-block sum of a normal density is well below 16), and the
// exact density-weighted test per quartet in the kernels.
int sqjk_grad_eri(sqjk_ctx* c, const double* PA, const double* PB, double Kfactor, double cutoff,
                  const int* basis2Atom, int nAtom, double* G) {
  if (!c || !PA || !PB || !basis2Atom || !G || nAtom <= 0 || !(cutoff >= 0.0))
    return fail(SQJK_EARG, "sqjk_grad_eri: invalid arguments");
  const int n = c->n;
  for (int p = 0; p < n; ++p)
    if (basis2Atom[p] < 0 || basis2Atom[p] >= nAtom) return fail(SQJK_EARG, "sqjk_grad_eri: basis2Atom out of range");
  HIPCHK(hipSetDevice(c->device));
  const size_t nn = (size_t)n * n;
  HIPCHK(hipMemcpyAsync(c->d_DA, PA, nn * 8, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->d_DB, PB, nn * 8, hipMemcpyHostToDevice, c->st));
  hipLaunchKernelGGL(k_density_sum, dim3(grid_for((long long)nn, 256)), dim3(256), 0, c->st, c->d_DA, c->d_DB,
                     c->d_DJ, nn);
  hipLaunchKernelGGL(k_pair_max, dim3(grid_for((long long)c->nref * c->nref, 256)), dim3(256), 0, c->st, c->d_DJ,
                     c->d_DA, c->d_DB, c->d_ref, c->nref, n, c->d_PM);
  HIPCHK(hipGetLastError());
  int rc;
  const double thr = cutoff / 16.0;
  if ((rc = make_plan(c, thr, thr))) return rc;
  double* d_g = nullptr;
  if ((rc = dalloc(&d_g, (size_t)3 * c->nsub))) return rc;
  hipLaunchKernelGGL(k_fill, dim3(grid_for(3LL * c->nsub, 256)), dim3(256), 0, c->st, d_g, (size_t)3 * c->nsub, 0.0);
  HIPCHK(hipEventRecord(c->ev0, c->st));
  for (const sqjk_ctx::Launch& L : c->launches) {
    GradArgs a{};
    a.c = base_args(c);
    a.c.bra = c->d_pairs[L.X];
    a.c.ket = c->d_pairs[L.Y];
    a.c.seg = c->d_seg;
    a.c.tile = c->d_tile + L.tile0;
    a.c.rbra = c->d_rbra;
    a.c.rlo = c->d_rlo;
    a.c.rhi = c->d_rhi;
    a.c.same = (L.X == L.Y);
    a.c.total = L.ntile;
    a.c.fixed_cut = cutoff;
    a.c.npb = (int)c->pairs[L.X].size();
    a.c.npk = (int)c->pairs[L.Y].size();
    a.gsub = d_g;
    a.kfac = Kfactor;
    GradLauncher f = grad_launcher_for(kPairL[L.X][0], kPairL[L.X][1], kPairL[L.Y][0], kPairL[L.Y][1]);
    if (!f || f(a, 0, c->st) < 0) {
      block_put(d_g);
      return fail(SQJK_EHIP, "gradient kernel launch failed");
    }
  }
  HIPCHK(hipEventRecord(c->ev1, c->st));
  std::vector<double> gs((size_t)3 * c->nsub);
  hipError_t e = hipMemcpyAsync(gs.data(), d_g, gs.size() * 8, hipMemcpyDeviceToHost, c->st);
  if (e == hipSuccess) e = hipStreamSynchronize(c->st);
  block_put(d_g);
  if (e != hipSuccess) return fail(SQJK_EHIP, std::string("sqjk_grad_eri: ") + hipGetErrorString(e));
  float ms = 0.0f;
  HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last.kernel_ms = ms;
  for (int k = 0; k < c->nsub; ++k) {
    const int at = basis2Atom[c->sub[k].bf_off];
    for (int x = 0; x < 3; ++x) G[3 * at + x] += gs[3 * k + x];
  }
  return SQJK_OK;
}

// Integrals of every shell quartet of the plan at `cutoff` (no density
// screening) into the dense pair matrix dV (m x m, m = n(n+1)/2), zeroed
// here.  Returns the quartets evaluated and the kernel time.
static int store_pairs(sqjk_ctx* c, double cutoff, double* dV, long long* quartets, float* ms) {
  const size_t m = (size_t)c->n * (c->n + 1) / 2;
  hipLaunchKernelGGL(k_fill, dim3(grid_for((long long)std::min<size_t>(m * m, (size_t)1 << 40), 256)), dim3(256), 0,
                     c->st, dV, m * m, 0.0);
  HIPCHK(hipGetLastError());
  int rc;
  if ((rc = make_plan(c, cutoff, cutoff))) return rc;
  HIPCHK(hipEventRecord(c->ev0, c->st));
  for (const sqjk_ctx::Launch& L : c->launches) {
    ClassArgs a = base_args(c);
    a.bra = c->d_pairs[L.X];
    a.ket = c->d_pairs[L.Y];
    a.seg = c->d_seg;
    a.tile = c->d_tile + L.tile0;
    a.rbra = c->d_rbra;
    a.rlo = c->d_rlo;
    a.rhi = c->d_rhi;
    a.same = (L.X == L.Y);
    a.total = L.ntile;
    a.fixed_cut = 0.0;
    a.npb = (int)c->pairs[L.X].size();
    a.npk = (int)c->pairs[L.Y].size();
    a.sch_out = dV;
    StoreLauncher f = store_launcher_for(kPairL[L.X][0], kPairL[L.X][1], kPairL[L.Y][0], kPairL[L.Y][1]);
    if (!f || f(a, c->st) < 0) return fail(SQJK_EHIP, "integral store kernel launch failed");
  }
  HIPCHK(hipEventRecord(c->ev1, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  HIPCHK(hipEventElapsedTime(ms, c->ev0, c->ev1));
  *quartets = c->plan_q;
  return SQJK_OK;
}

int sqjk_eri_pairs(sqjk_ctx* c, double cutoff, double* V) {
  if (!c || !V || !(cutoff >= 0.0)) return fail(SQJK_EARG, "sqjk_eri_pairs: invalid arguments");
  HIPCHK(hipSetDevice(c->device));
  const size_t m = (size_t)c->n * (c->n + 1) / 2;
  double* dV = nullptr;
  if (hipMalloc(&dV, m * m * 8) != hipSuccess)
    return fail(SQJK_ENOMEM, "sqjk_eri_pairs: the pair matrix needs " + std::to_string(m * m * 8 >> 20) + " MiB");
  long long q = 0;
  float ms = 0.0f;
  int rc = store_pairs(c, cutoff, dV, &q, &ms);
  if (rc == SQJK_OK) {
    hipError_t e = hipMemcpy(V, dV, m * m * 8, hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = fail(SQJK_EHIP, std::string("sqjk_eri_pairs: ") + hipGetErrorString(e));
  }
  (void)hipFree(dV);
  c->last.kernel_ms = ms;
  return rc;
}

// MP2 of one (closed shell) or two (alpha, beta) spin blocks: nOcc[k]
// occupied orbitals of C[k] (orbital-major), the first nCore frozen.
static int mp2_run(sqjk_ctx* c, int nspin, const int* nOcc, int nCore, const double* const* e, const double* const* C,
                   double cutoff, sqjk_mp2_info* info, const char* who) {
  const int n = c->n;
  for (int k = 0; k < nspin; ++k) {
    if (!e[k] || !C[k]) return fail(SQJK_EARG, std::string(who) + ": invalid arguments");
    if (nCore < 0 || nOcc[k] < nCore || nOcc[k] > n)
      return fail(SQJK_EARG, std::string(who) + ": need 0 <= nCore <= nOcc <= nBasis");
  }
  HIPCHK(hipSetDevice(c->device));
  sqjk_mp2_info r{};
  int nc[2] = {0, 0}, nv[2] = {0, 0};
  size_t ncv = 0;
  bool any = false;
  for (int k = 0; k < nspin; ++k) {
    nc[k] = nOcc[k] - nCore;
    nv[k] = n - nOcc[k];
    ncv += (size_t)nc[k] * nv[k];
    any = any || (nc[k] > 0 && nv[k] > 0);
  }
  r.device_bytes = (long long)sq_mp2_bytes(n, ncv);
  if (any) {
    size_t freeb = 0, totb = 0;
    HIPCHK(hipMemGetInfo(&freeb, &totb));
    const size_t need = (size_t)r.device_bytes + ((size_t)1 << 30);
    if (need > freeb)
      return fail(SQJK_ENOMEM, std::string(who) + ": needs " + std::to_string(need >> 20) +
                                   " MiB of device memory (the pair matrix and the half-transformed integrals), " +
                                   std::to_string(freeb >> 20) + " MiB free");
    const size_t m = (size_t)n * (n + 1) / 2;
    std::vector<double*> bufs;
    double* dV = nullptr;
    auto release = [&] {
      if (dV) (void)hipFree(dV);
      for (double* p : bufs) (void)hipFree(p);
    };
    auto up = [&](const double* h, size_t count) -> double* {
      double* d = nullptr;
      if (hipMalloc(&d, std::max<size_t>(count, 1) * 8) != hipSuccess) return nullptr;
      bufs.push_back(d);
      if (count && hipMemcpyAsync(d, h, count * 8, hipMemcpyHostToDevice, c->st) != hipSuccess) return nullptr;
      return d;
    };
    SqMp2Spin sp[2] = {};
    bool ok = hipMalloc(&dV, m * m * 8) == hipSuccess;
    for (int k = 0; ok && k < nspin; ++k) {
      // orbital-major C: the correlated and virtual orbitals are contiguous
      // column blocks of the column-major n x n matrix
      sp[k].nc = nc[k];
      sp[k].nv = nv[k];
      ok = (sp[k].dCo = up(C[k] + (size_t)nCore * n, (size_t)n * nc[k])) &&
           (sp[k].dCv = up(C[k] + (size_t)nOcc[k] * n, (size_t)n * nv[k])) &&
           (sp[k].dec = up(e[k] + nCore, (size_t)nc[k])) && (sp[k].dev = up(e[k] + nOcc[k], (size_t)nv[k]));
    }
    if (!ok) {
      release();
      return fail(SQJK_ENOMEM, std::string(who) + ": device allocation or upload failed");
    }
    float ms = 0.0f;
    int rc = store_pairs(c, cutoff, dV, &r.quartets, &ms);
    if (rc) {
      release();
      return rc;
    }
    r.store_ms = ms;
    SqMp2Result t{};
    std::string err;
    rc = sq_mp2_from_pairs(c->st, n, nspin, sp, &dV, &t, &err);
    release();
    if (rc) return fail(SQJK_EHIP, std::string(who) + ": " + err);
    r.e2 = t.e2;
    r.half1_ms = t.half1_ms;
    r.half2_ms = t.half2_ms;
    r.half1_flop = t.half1_flop;
    r.half2_flop = t.half2_flop;
  }
  if (info) *info = r;
  return SQJK_OK;
}

int sqjk_mp2_rhf(sqjk_ctx* c, int nOcc, int nCore, const double* e, const double* C, double cutoff,
                 sqjk_mp2_info* info) {
  if (!c || !(cutoff >= 0.0)) return fail(SQJK_EARG, "sqjk_mp2_rhf: invalid arguments");
  const double* ee[1] = {e};
  const double* cc[1] = {C};
  return mp2_run(c, 1, &nOcc, nCore, ee, cc, cutoff, info, "sqjk_mp2_rhf");
}

int sqjk_mp2_uhf(sqjk_ctx* c, int nEA, int nEB, int nCore, const double* eA, const double* eB, const double* CA,
                 const double* CB, double cutoff, sqjk_mp2_info* info) {
  if (!c || !(cutoff >= 0.0)) return fail(SQJK_EARG, "sqjk_mp2_uhf: invalid arguments");
  const int no[2] = {nEA, nEB};
  const double* ee[2] = {eA, eB};
  const double* cc[2] = {CA, CB};
  return mp2_run(c, 2, no, nCore, ee, cc, cutoff, info, "sqjk_mp2_uhf");
}

int sqjk_grad_one_electron(int nBasis, const GTOBasis_t* gto, const Molecule_t* mol, const double* PT,
                           const double* QT, const int* basis2Atom, int device, double* G) {
  if (nBasis <= 0 || !gto || !mol || mol->nAtom <= 0 || !PT || !QT || !basis2Atom || !G)
    return fail(SQJK_EARG, "sqjk_grad_one_electron: invalid arguments");
  for (int p = 0; p < nBasis; ++p)
    if (basis2Atom[p] < 0 || basis2Atom[p] >= mol->nAtom)
      return fail(SQJK_EARG, "sqjk_grad_one_electron: basis2Atom out of range");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    return fail(SQJK_ENODEV, "no HIP device is visible; this library has no CPU path");
  if (device < 0 || device >= ndev) return fail(SQJK_EARG, "sqjk_grad_one_electron: device ordinal out of range");
  HIPCHK(hipSetDevice(device));
  sqjk_ctx c;
  c.n = nBasis;
  std::vector<double> pexp, pcoef;
  int rc = plan_shells(&c, nBasis, gto, pexp, pcoef);
  if (rc) return rc;
  plan_pairs(&c, pexp, pcoef);
  const int na = mol->nAtom;
  std::vector<double> atoms((size_t)4 * na);
  for (int k = 0; k < na; ++k) {
    atoms[4 * k + 0] = mol->x[k]; atoms[4 * k + 1] = mol->y[k]; atoms[4 * k + 2] = mol->z[k];
    atoms[4 * k + 3] = (double)mol->Z[k];
  }
  const size_t nn = (size_t)nBasis * nBasis;
  thread_local hipStream_t st = nullptr;
  thread_local int st_dev = -1;
  if (!st || st_dev != device) {
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    st_dev = device;
  }
  std::vector<void*> bufs;
  bool up_ok = true;
  auto up = [&](const void* h, size_t bytes) -> void* {
    void* d = nullptr;
    if (!(d = block_get(std::max<size_t>(bytes, 8), false))) { up_ok = false; return nullptr; }
    bufs.push_back(d);
    if (h && bytes && hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, st) != hipSuccess) up_ok = false;
    return d;
  };
  auto release = [&] {
    (void)hipStreamSynchronize(st);
    for (void* d : bufs) block_put(d);
  };
  const std::vector<double>& boys = boys_table();
  std::vector<double> zero((size_t)3 * std::max(c.nsub, na), 0.0);
  OneEGradArgs a{};
  a.sh = (const SubShell*)up(c.sub.data(), c.sub.size() * sizeof(SubShell));
  a.pexp = (const double*)up(pexp.data(), pexp.size() * 8);
  a.pcoef = (const double*)up(pcoef.data(), pcoef.size() * 8);
  a.bscale = (const double*)up(c.bscale.data(), c.bscale.size() * 8);
  a.atom = (const double*)up(atoms.data(), atoms.size() * 8);
  a.boys = (const double*)up(boys.data(), boys.size() * 8);
  a.PT = (const double*)up(PT, nn * 8);
  a.QT = (const double*)up(QT, nn * 8);
  a.gsub = (double*)up(zero.data(), (size_t)3 * c.nsub * 8);
  a.gatom = (double*)up(zero.data(), (size_t)3 * na * 8);
  a.natom = na;
  a.n = nBasis;
  if (!up_ok) {
    release();
    return fail(SQJK_ENOMEM, "sqjk_grad_one_electron: device allocation or upload failed");
  }
  for (int X = 0; X < kNPairClass; ++X) {
    a.pr = (const PairRec*)up(c.pairs[X].data(), c.pairs[X].size() * sizeof(PairRec));
    a.np = (int)c.pairs[X].size();
    if (!up_ok || launch_one_electron_grad(X, a, st) != 0) {
      release();
      return fail(SQJK_EHIP, "sqjk_grad_one_electron: kernel launch failed");
    }
  }
  std::vector<double> gs((size_t)3 * c.nsub), ga((size_t)3 * na);
  hipError_t e = hipMemcpyAsync(gs.data(), a.gsub, gs.size() * 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(ga.data(), a.gatom, ga.size() * 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  release();
  if (e != hipSuccess) return fail(SQJK_EHIP, std::string("sqjk_grad_one_electron: ") + hipGetErrorString(e));
  for (int k = 0; k < c.nsub; ++k) {
    const int at = basis2Atom[c.sub[k].bf_off];
    for (int x = 0; x < 3; ++x) G[3 * at + x] += gs[3 * k + x];
  }
  for (int k = 0; k < 3 * na; ++k) G[k] += ga[k];
  return SQJK_OK;
}

int sqjk_sync(sqjk_ctx* c) {
  if (!c) return fail(SQJK_EARG, "sqjk_sync: null context");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->st));
  return collect_times(c);
}

int sqjk_boys(int L, int nT, const double* T, int fp32, int device, double* F) {
  if (L < 0 || L > 9 || nT < 0 || (nT > 0 && (!T || !F))) return fail(SQJK_EARG, "sqjk_boys: invalid arguments");
  if (nT == 0) return SQJK_OK;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    return fail(SQJK_ENODEV, "no HIP device is visible; this library has no CPU path");
  if (device < 0 || device >= ndev) return fail(SQJK_EARG, "sqjk_boys: device ordinal out of range");
  HIPCHK(hipSetDevice(device));
  for (int i = 0; i < nT; ++i)
    if (!(T[i] >= 0.0)) return fail(SQJK_EARG, "sqjk_boys: T must be >= 0");
  const std::vector<double>& boys = boys_table();
  double *d_tab = nullptr, *d_T = nullptr, *d_F = nullptr;
  int rc;
  if ((rc = dalloc(&d_tab, boys.size())) || (rc = dalloc(&d_T, (size_t)nT)) ||
      (rc = dalloc(&d_F, (size_t)nT * (L + 1)))) {
    block_put(d_tab); block_put(d_T); block_put(d_F);
    return rc;
  }
  hipError_t e = hipMemcpy(d_tab, boys.data(), boys.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_T, T, (size_t)nT * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    const int k = fp32 ? launch_boys_eval<float>(L, d_tab, d_T, d_F, nT, 0)
                       : launch_boys_eval<double>(L, d_tab, d_T, d_F, nT, 0);
    e = k ? hipErrorInvalidValue : hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(F, d_F, (size_t)nT * (L + 1) * sizeof(double), hipMemcpyDeviceToHost);
  block_put(d_tab); block_put(d_T); block_put(d_F);
  if (e != hipSuccess) return fail(SQJK_EHIP, std::string("sqjk_boys: ") + hipGetErrorString(e));
  return SQJK_OK;
}

int sqjk_set_mixed(sqjk_ctx* c, double threshold) {
  if (!c || threshold < 0.0) return fail(SQJK_EARG, "sqjk_set_mixed: invalid arguments");
  c->mixed_thr = threshold;
  return SQJK_OK;
}

int sqjk_schwarz(sqjk_ctx* c, double* out) {
  if (!c || !out) return fail(SQJK_EARG, "sqjk_schwarz: null argument");
  std::memcpy(out, c->sch_basis.data(), c->sch_basis.size() * sizeof(double));
  return SQJK_OK;
}

int sqjk_last_stats(const sqjk_ctx* c, sqjk_stats* st) {
  if (!c || !st) return fail(SQJK_EARG, "sqjk_last_stats: null argument");
  int rc = collect_times(const_cast<sqjk_ctx*>(c));
  if (rc) return rc;
  *st = c->last;
  return SQJK_OK;
}

int sqjk_class_stats(const sqjk_ctx* c, int* lq, long long* quartets, long long* prim_quartets) {
  if (!c) return fail(SQJK_EARG, "sqjk_class_stats: null context");
  int k = 0;
  for (int X = 0; X < kNPairClass; ++X)
    for (int Y = 0; Y <= X; ++Y, ++k) {
      if (lq) {
        lq[4 * k + 0] = kPairL[X][0]; lq[4 * k + 1] = kPairL[X][1];
        lq[4 * k + 2] = kPairL[Y][0]; lq[4 * k + 3] = kPairL[Y][1];
      }
      if (quartets) quartets[k] = c->cls_q[X][Y];
      if (prim_quartets) prim_quartets[k] = c->cls_p[X][Y];
    }
  return k;
}

int sqjk_class_work(sqjk_ctx* c, long long* quartets, long long* prim_quartets, long long* quartets_fp32,
                    long long* prim_quartets_fp32) {
  if (!c) return fail(SQJK_EARG, "sqjk_class_work: null context");
  if (!c->built) return fail(SQJK_EARG, "sqjk_class_work: no build has run on this context");
  if (!c->plan_ok)
    return fail(SQJK_EARG, "sqjk_class_work: the last build's work plan is not held by the context "
                           "(a batched or allocation-free build ran since)");
  HIPCHK(hipSetDevice(c->device));
  const size_t nl = c->launches.size();
  unsigned long long* d_cnt = nullptr;
  int rc = dalloc(&d_cnt, 2 * std::max<size_t>(nl, 1));
  if (rc) return rc;
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(256), 0, c->st, reinterpret_cast<double*>(d_cnt), 2 * nl, 0.0);
  for (size_t li = 0; li < nl; ++li) {
    const sqjk_ctx::Launch& L = c->launches[li];
    ClassArgs a = base_args(c);
    a.bra = c->d_pairs[L.X];
    a.ket = c->d_pairs[L.Y];
    a.seg = c->d_seg;
    a.tile = c->d_tile + L.tile0;
    a.rbra = c->d_rbra; a.rlo = c->d_rlo; a.rhi = c->d_rhi;
    a.total = L.ntile;
    a.fixed_cut = c->last_fixed;
    const int grid = (int)std::min<long long>((L.ntile + 3) / 4, 8192);
    hipLaunchKernelGGL(k_count_work, dim3(std::max(grid, 1)), dim3(256), 0, c->st, a, d_cnt + 2 * li);
  }
  std::vector<unsigned long long> h(2 * nl);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && nl) e = hipMemcpyAsync(h.data(), d_cnt, 2 * nl * sizeof(unsigned long long),
                                                hipMemcpyDeviceToHost, c->st);
  if (e == hipSuccess) e = hipStreamSynchronize(c->st);
  block_put(d_cnt);
  if (e != hipSuccess) return fail(SQJK_EHIP, std::string("sqjk_class_work: ") + hipGetErrorString(e));
  long long q[kNPairClass][kNPairClass][2] = {}, p[kNPairClass][kNPairClass][2] = {};
  for (size_t li = 0; li < nl; ++li) {
    const sqjk_ctx::Launch& L = c->launches[li];
    q[L.X][L.Y][L.fp32] += (long long)h[2 * li];
    p[L.X][L.Y][L.fp32] += (long long)h[2 * li + 1];
  }
  int k = 0;
  for (int X = 0; X < kNPairClass; ++X)
    for (int Y = 0; Y <= X; ++Y, ++k) {
      if (quartets) quartets[k] = q[X][Y][0];
      if (prim_quartets) prim_quartets[k] = p[X][Y][0];
      if (quartets_fp32) quartets_fp32[k] = q[X][Y][1];
      if (prim_quartets_fp32) prim_quartets_fp32[k] = p[X][Y][1];
    }
  return k;
}

int sqjk_set_tuning(sqjk_ctx* c, const int* rows, const int* split_budget, const int* grid_over) {
  if (!c) return fail(SQJK_EARG, "sqjk_set_tuning: null context");
  for (int k = 0; k < kNPairClass * kNPairClass; ++k) {
    if (rows && (rows[k] < 1 || rows[k] > 64)) return fail(SQJK_EARG, "sqjk_set_tuning: rows per tile outside [1, 64]");
    if (split_budget && (split_budget[k] < 1 || split_budget[k] > 1 << 20))
      return fail(SQJK_EARG, "sqjk_set_tuning: split budget outside [1, 2^20]");
    if (grid_over && (grid_over[k] < 0 || grid_over[k] > 64))
      return fail(SQJK_EARG, "sqjk_set_tuning: grid multiple outside [0, 64]");
  }
  if (!rows && !split_budget && !grid_over) {
    init_tuning(c);
  } else {
    for (int X = 0; X < kNPairClass; ++X)
      for (int Y = 0; Y < kNPairClass; ++Y) {
        if (rows) c->rows[X][Y] = rows[X * kNPairClass + Y];
        if (split_budget) c->split_budget[X][Y] = split_budget[X * kNPairClass + Y];
        if (grid_over) c->grid_over[X][Y] = grid_over[X * kNPairClass + Y];
      }
  }
  c->plan_ok = c->hplan_ok = false;
  return SQJK_OK;
}

int sqjk_get_tuning(const sqjk_ctx* c, int* rows, int* split_budget, int* grid_over) {
  if (!c) return fail(SQJK_EARG, "sqjk_get_tuning: null context");
  for (int X = 0; X < kNPairClass; ++X)
    for (int Y = 0; Y < kNPairClass; ++Y) {
      if (rows) rows[X * kNPairClass + Y] = c->rows[X][Y];
      if (split_budget) split_budget[X * kNPairClass + Y] = c->split_budget[X][Y];
      if (grid_over) grid_over[X * kNPairClass + Y] = c->grid_over[X][Y];
    }
  return SQJK_OK;
}

int sqjk_set_profile(sqjk_ctx* c, int on) {
  if (!c) return fail(SQJK_EARG, "sqjk_set_profile: null context");
  c->profile = on != 0;
  return SQJK_OK;
}

int sqjk_class_times(const sqjk_ctx* c, double* kernel_ms) {
  if (!c || !kernel_ms) return fail(SQJK_EARG, "sqjk_class_times: null argument");
  if (int rc = collect_times(const_cast<sqjk_ctx*>(c))) return rc;
  int k = 0;
  for (int X = 0; X < kNPairClass; ++X)
    for (int Y = 0; Y <= X; ++Y, ++k) kernel_ms[k] = c->cls_ms[X][Y];
  return k;
}

int sqjk_dims(const sqjk_ctx* c, int* nBasis, int* nShell, int* nSubShell) {
  if (!c) return fail(SQJK_EARG, "sqjk_dims: null context");
  if (nBasis) *nBasis = c->n;
  if (nShell) *nShell = c->nref;
  if (nSubShell) *nSubShell = c->nsub;
  return SQJK_OK;
}

void sqjk_destroy(sqjk_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->st) (void)hipStreamSynchronize(c->st);
  void* ptrs[] = {c->d_sub, c->d_pp, c->d_boys, c->d_bscale, c->d_DJ, c->d_DA, c->d_DB, c->d_PM, c->d_AJ,
                  c->d_AKA, c->d_AKB, c->d_J,   c->d_KA,    c->d_KB, c->d_sch, c->d_seg, c->d_tile, c->d_rbra, c->d_rlo, c->d_rhi};
  for (void* p : ptrs)
    block_put(p);
  block_put(c->d_ref);
  block_put(c->d_herm);
  block_put(c->d_SC);
  block_put(c->d_SCB);
  double* pins[] = {c->hp_A, c->hp_B, c->hp_J, c->hp_KA, c->hp_KB};
  for (double* h : pins)
    block_put(h);
  for (int X = 0; X < kNPairClass; ++X) {
    block_put(c->d_pairs[X]);
    block_put(c->d_Jp[X]);
    block_put(c->d_SA[X]);
    block_put(c->d_SB[X]);
  }
  for (hipEvent_t e : c->prof_ev) (void)hipEventDestroy(e);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->st) (void)hipStreamDestroy(c->st);
  delete c;
}

// ------------------------------------------------------------ batches
// Many molecules' J/K builds as one set of launches: one launch per (class
// pair, precision, spin mode) covers the tiles of every molecule of the
// batch.  The molecules' pair lists, primitive data and component scales
// are concatenated (pair records carry their molecule, sub-shells their
// offset in the concatenated scales), the tiles of each molecule keep their
// own work plan, and a segment names its molecule, whose matrices,
// screening maxima and exchange staging the kernels find through a
// per-molecule descriptor.  The whole build (density preparation, fills,
// class launches, reductions, finalize) is captured once into a HIP graph
// and replayed while the work plan stays the same.

namespace {

// Per-molecule offsets of the concatenated helper arrays
struct BatchGeom {
  long long mat_off;  // n x n blocks
  long long pm_off;   // nref x nref blocks
  int ref0;           // first reference shell in the concatenated list
  int row0;           // first basis function (rows of all molecules)
  int n, nref, uhf;
};

__host__ __device__ inline int find_inst(const long long* off, int count, long long e) {  // last i with off[i] <= e
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= e) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__global__ void k_pair_max_batch(const double* __restrict__ DJ, const double* __restrict__ DA,
                                 const double* __restrict__ DB, const int* __restrict__ ref,
                                 const BatchGeom* __restrict__ geo, const long long* __restrict__ pm_off, int count,
                                 long long total, double* PM) {
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const int i = find_inst(pm_off, count, t);
    const BatchGeom g = geo[i];
    const long long l = t - g.pm_off;
    const int P = (int)(l / g.nref), Q = (int)(l - (long long)P * g.nref);
    const int* rf = ref + 2 * g.ref0;
    const int p0 = rf[2 * P], np = rf[2 * P + 1], q0 = rf[2 * Q], nq = rf[2 * Q + 1];
    const double* dj = DJ + g.mat_off;
    const double* da = DA + g.mat_off;
    const double* db = DB + g.mat_off;
    double r = 0.0;
    for (int p = p0; p < p0 + np; ++p)
      for (int q = q0; q < q0 + nq; ++q) {
        const size_t e = (size_t)p * g.n + q;
        r = fmax(r, fmax(fabs(dj[e]), fmax(fabs(da[e]), fabs(db[e]))));
      }
    PM[t] = r;
  }
}

// A_J of each pair's molecule += Jp (see k_reduce_j)
__global__ void k_reduce_j_batch(const double* __restrict__ Jp, const PairRec* __restrict__ pr,
                                 const SubShell* __restrict__ sh, const InstDesc* __restrict__ inst, int np, int nb,
                                 int nab) {
  const long long tot = (long long)np * nab;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot;
       e += (long long)gridDim.x * blockDim.x) {
    const double v = Jp[e];
    if (v == 0.0) continue;
    const int ab = (int)(e / np), i = (int)(e - (long long)ab * np);
    const PairRec p = pr[i];
    const InstDesc& d = inst[p.inst];
    d.AJ[(size_t)(sh[p.a].bf_off + ab / nb) * d.n + sh[p.b].bf_off + ab % nb] += v;
  }
}

// One workgroup per row of every molecule's A_K (see k_reduce_k); spin 1
// rows of RHF molecules are skipped.
__global__ void __launch_bounds__(kReduceKBlock) k_reduce_k_batch(const PairRec* const* __restrict__ prs,
                                                                  const int* __restrict__ ncomp_c,
                                                                  const int* __restrict__ nc_c,
                                                                  const SubShell* __restrict__ sh,
                                                                  const InstDesc* __restrict__ inst,
                                                                  const BatchGeom* __restrict__ geo,
                                                                  const long long* __restrict__ row0, int count,
                                                                  int rows, int spin) {
  extern __shared__ double row[];
  for (int xg = blockIdx.x; xg < rows; xg += gridDim.x) {
    const int i = find_inst(row0, count, xg);
    const BatchGeom g = geo[i];
    if (spin == 1 && !g.uhf) continue;  // uniform in the workgroup
    const int x = xg - g.row0, n = g.n;
    const InstDesc& d = inst[i];
    for (int y = threadIdx.x; y < n; y += blockDim.x) row[y] = 0.0;
    __syncthreads();
    for (int Y = 0; Y < 6; ++Y) {
      const double* S = spin ? d.SB[Y] : d.SA[Y];
      const int np = d.np[Y];
      if (!S || np == 0) continue;
      const int nc = nc_c[Y], ncomp = ncomp_c[Y];
      const double* Sx = S + (size_t)x * ncomp * np;
      const PairRec* prY = prs[Y] + d.p0[Y];
      const long long tot = (long long)ncomp * np;
      for (long long e = threadIdx.x; e < tot; e += blockDim.x) {
        const double v = Sx[e];
        if (v == 0.0) continue;
        const int comp = (int)(e / np), j = (int)(e - (long long)comp * np);
        const PairRec p = prY[j];
        const int col = comp < nc ? sh[p.a].bf_off + comp : sh[p.b].bf_off + comp - nc;
        __hip_atomic_fetch_add(&row[col], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
    }
    __syncthreads();
    double* AK = spin ? d.AKB : d.AKA;
    for (int y = threadIdx.x; y < n; y += blockDim.x) AK[(size_t)x * n + y] += row[y];
    __syncthreads();
  }
}

// J = A + A^T, K = -(A + A^T) per molecule; KB := KA for RHF molecules
__global__ void k_finalize_batch(const double* __restrict__ AJ, const double* __restrict__ AKA,
                                 const double* __restrict__ AKB, double* J, double* KA, double* KB,
                                 const BatchGeom* __restrict__ geo, const long long* __restrict__ mat_off, int count,
                                 long long total) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int i = find_inst(mat_off, count, e);
    const BatchGeom g = geo[i];
    const long long l = e - g.mat_off;
    const int r = (int)(l / g.n), c = (int)(l - (long long)r * g.n);
    const long long t = g.mat_off + (long long)c * g.n + r;
    J[e] = AJ[e] + AJ[t];
    const double ka = -(AKA[e] + AKA[t]);
    KA[e] = ka;
    KB[e] = g.uhf ? -(AKB[e] + AKB[t]) : ka;
  }
}

}  // namespace

struct sqjk_batch {
  int device = 0, count = 0;
  hipStream_t st = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<sqjk_ctx*> ctx;
  std::vector<int> uhf;
  std::vector<double> fixed, scf;  // per-molecule cut-offs of the next build
  double mixed_thr = 0.0;
  // concatenated device data
  SubShell* d_sub = nullptr;
  PrimPair* d_pp = nullptr;
  PairRec* d_pairs[kNPairClass] = {};
  const PairRec** d_prs = nullptr;  // the six class lists (for the K reduction)
  int* d_ncomp = nullptr;           // per pair class: components, c-components
  int* d_nc = nullptr;
  double *d_bscale = nullptr, *d_boys = nullptr;
  int* d_ref = nullptr;
  long long mat_total = 0, pm_total = 0;
  int rows_total = 0;
  double *d_DA = nullptr, *d_DB = nullptr, *d_DJ = nullptr, *d_PM = nullptr;
  double *d_AJ = nullptr, *d_AKA = nullptr, *d_AKB = nullptr, *d_J = nullptr, *d_KA = nullptr, *d_KB = nullptr;
  double* d_Jp[kNPairClass] = {};
  size_t jp_len[kNPairClass] = {};
  int np_total[kNPairClass] = {};
  double *d_SA = nullptr, *d_SB = nullptr;
  size_t s_total = 0;
  std::vector<BatchGeom> geo;
  BatchGeom* d_geo = nullptr;
  long long *d_mat_off = nullptr, *d_pm_off = nullptr, *d_row0 = nullptr;
  std::vector<InstDesc> desc;
  InstDesc* d_inst = nullptr;
  // merged work plan
  struct Launch { int X, Y, fp32, mode; long long tile0, ntile; };
  std::vector<Launch> launches;
  Segment* d_seg = nullptr;
  Tile* d_tile = nullptr;
  int *d_rbra = nullptr, *d_rlo = nullptr, *d_rhi = nullptr;
  bool plan_ok = false;
  long long plan_q = 0, plan_q32 = 0;
  // graph of the build for the current plan
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool profile = false;
  std::vector<hipEvent_t> prof_ev;
  double cls_ms[kNPairClass][kNPairClass] = {};
  double last_ms = 0.0;
  int last_launches = 0;
};

namespace {

void batch_free_plan(sqjk_batch* b) {
  void* p[] = {b->d_seg, b->d_tile, b->d_rbra, b->d_rlo, b->d_rhi};
  for (void* q : p) block_put(q);
  b->d_seg = nullptr; b->d_tile = nullptr; b->d_rbra = b->d_rlo = b->d_rhi = nullptr;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->plan_ok = false;
}

// Per-molecule plans for the molecules' cut-offs, merged per (class pair,
// precision, spin mode): segments get their molecule and the class-list
// offset of its ket pairs, rows the offset of its bra pairs, tiles the
// offsets of the merged segments and rows.
int batch_plan(sqjk_batch* b) {
  batch_free_plan(b);
  std::vector<Segment> seg;
  std::vector<Tile> tile;
  std::vector<int> rbra, rlo, rhi;
  b->launches.clear();
  b->plan_q = b->plan_q32 = 0;
  std::vector<std::array<int, kNPairClass>> p0(b->count);
  {
    std::array<int, kNPairClass> acc{};
    for (int i = 0; i < b->count; ++i) {
      p0[i] = acc;
      for (int X = 0; X < kNPairClass; ++X) acc[X] += (int)b->ctx[i]->pairs[X].size();
    }
  }
  for (int i = 0; i < b->count; ++i) {
    sqjk_ctx* c = b->ctx[i];
    c->mixed_thr = b->mixed_thr;
    const double thr = std::max(b->fixed[i], b->scf[i]);
    build_plan(c, thr, std::max(thr, c->mixed_thr));
    c->plan_ok = false;  // the context's own device plan is not this one
  }
  for (int mode = 0; mode < 2; ++mode)
    for (int X = 0; X < kNPairClass; ++X)
      for (int Y = 0; Y <= X; ++Y)
        for (int fp = 0; fp < 2; ++fp) {
          sqjk_batch::Launch L{X, Y, fp, mode, (long long)tile.size(), 0};
          for (int i = 0; i < b->count; ++i) {
            if (b->uhf[i] != mode) continue;
            const sqjk_ctx* c = b->ctx[i];
            for (const sqjk_ctx::Launch& l : c->launches) {
              if (l.X != X || l.Y != Y || l.fp32 != fp) continue;
              const int seg_base = (int)seg.size() - l.seg0, row_base = (int)rbra.size();
              const int seg_end = (&l == &c->launches.back()) ? (int)c->h_seg.size() : (&l + 1)->seg0;
              for (int sgi = l.seg0; sgi < seg_end; ++sgi) {
                Segment sgm = c->h_seg[sgi];
                sgm.ket0 += p0[i][Y];
                sgm.inst = i;
                seg.push_back(sgm);
              }
              const long long t_end = l.tile0 + l.ntile;
              int r_first = INT32_MAX, r_last = 0;
              for (long long t = l.tile0; t < t_end; ++t) {
                r_first = std::min(r_first, c->h_tile[t].r0);
                r_last = std::max(r_last, c->h_tile[t].r1);
              }
              for (int r = r_first; r < r_last; ++r) {
                rbra.push_back(c->h_rbra[r] + p0[i][X]);
                rlo.push_back(c->h_rlo[r]);
                rhi.push_back(c->h_rhi[r]);
              }
              for (long long t = l.tile0; t < t_end; ++t) {
                Tile tl = c->h_tile[t];
                tl.seg += seg_base;
                tl.r0 += row_base - r_first;
                tl.r1 += row_base - r_first;
                tile.push_back(tl);
              }
              b->plan_q += l.quartets;
              if (fp) b->plan_q32 += l.quartets;
            }
          }
          L.ntile = (long long)tile.size() - L.tile0;
          if (L.ntile > 0) b->launches.push_back(L);
        }
  if (rbra.size() > (size_t)INT32_MAX || seg.size() > (size_t)INT32_MAX)
    return fail(SQJK_ENOMEM, "batched work plan too large");
  int rc;
  if ((rc = dalloc(&b->d_seg, seg.size())) || (rc = dalloc(&b->d_tile, tile.size())) ||
      (rc = dalloc(&b->d_rbra, rbra.size())) || (rc = dalloc(&b->d_rlo, rlo.size())) ||
      (rc = dalloc(&b->d_rhi, rhi.size())))
    return rc;
  if (!tile.empty()) {
    HIPCHK(hipMemcpyAsync(b->d_seg, seg.data(), seg.size() * sizeof(Segment), hipMemcpyHostToDevice, b->st));
    HIPCHK(hipMemcpyAsync(b->d_tile, tile.data(), tile.size() * sizeof(Tile), hipMemcpyHostToDevice, b->st));
    HIPCHK(hipMemcpyAsync(b->d_rbra, rbra.data(), rbra.size() * sizeof(int), hipMemcpyHostToDevice, b->st));
    HIPCHK(hipMemcpyAsync(b->d_rlo, rlo.data(), rlo.size() * sizeof(int), hipMemcpyHostToDevice, b->st));
    HIPCHK(hipMemcpyAsync(b->d_rhi, rhi.data(), rhi.size() * sizeof(int), hipMemcpyHostToDevice, b->st));
  }
  for (int i = 0; i < b->count; ++i) b->desc[i].fixed_cut = b->fixed[i];
  HIPCHK(hipMemcpyAsync(b->d_inst, b->desc.data(), b->desc.size() * sizeof(InstDesc), hipMemcpyHostToDevice, b->st));
  HIPCHK(hipStreamSynchronize(b->st));
  b->plan_ok = true;
  return SQJK_OK;
}

// The build on the batch's stream (captured into the graph, or eager; the
// per-class events of profile mode only when eager).
int batch_enqueue(sqjk_batch* b, bool prof) {
  const long long mt = b->mat_total;
  const int fill_grid = grid_for(mt, 256);
  hipLaunchKernelGGL(k_density_sum, dim3(fill_grid), dim3(256), 0, b->st, b->d_DA, b->d_DB, b->d_DJ, (size_t)mt);
  hipLaunchKernelGGL(k_pair_max_batch, dim3(grid_for(b->pm_total, 256)), dim3(256), 0, b->st, b->d_DJ, b->d_DA,
                     b->d_DB, b->d_ref, b->d_geo, b->d_pm_off, b->count, b->pm_total, b->d_PM);
  hipLaunchKernelGGL(k_fill, dim3(fill_grid), dim3(256), 0, b->st, b->d_AJ, (size_t)mt, 0.0);
  hipLaunchKernelGGL(k_fill, dim3(fill_grid), dim3(256), 0, b->st, b->d_AKA, (size_t)mt, 0.0);
  hipLaunchKernelGGL(k_fill, dim3(fill_grid), dim3(256), 0, b->st, b->d_AKB, (size_t)mt, 0.0);
  for (int X = 0; X < kNPairClass; ++X)
    if (b->jp_len[X])
      hipLaunchKernelGGL(k_fill, dim3(grid_for((long long)b->jp_len[X], 256)), dim3(256), 0, b->st, b->d_Jp[X],
                         b->jp_len[X], 0.0);
  const int sg = std::max(1, (int)std::min<size_t>((b->s_total + 255) / 256, 16384));
  hipLaunchKernelGGL(k_fill, dim3(sg), dim3(256), 0, b->st, b->d_SA, b->s_total, 0.0);
  hipLaunchKernelGGL(k_fill, dim3(sg), dim3(256), 0, b->st, b->d_SB, b->s_total, 0.0);
  HIPCHK(hipGetLastError());
  int launches = 0;
  for (size_t li = 0; li < b->launches.size(); ++li) {
    const sqjk_batch::Launch& L = b->launches[li];
    if (prof) HIPCHK(hipEventRecord(b->prof_ev[2 * li], b->st));
    ClassArgs a{};
    a.sh = b->d_sub; a.pp = b->d_pp; a.boys = b->d_boys; a.bscale = b->d_bscale;
    a.bra = b->d_pairs[L.X];
    a.ket = b->d_pairs[L.Y];
    a.seg = b->d_seg;
    a.tile = b->d_tile + L.tile0;
    a.rbra = b->d_rbra; a.rlo = b->d_rlo; a.rhi = b->d_rhi;
    a.same = (L.X == L.Y);
    a.total = L.ntile;
    a.JpB = b->d_Jp[L.X];
    a.JpK = b->d_Jp[L.Y];
    a.npb = b->np_total[L.X];
    a.npk = b->np_total[L.Y];
    a.inst = b->d_inst;
    a.ycls = L.Y;
    const int go = b->ctx[0]->grid_over[L.X][L.Y];
    ClassLauncher f = launcher_for(kPairL[L.X][0], kPairL[L.X][1], kPairL[L.Y][0], kPairL[L.Y][1]);
    const int k = f ? f(a, L.mode ? kUHF : kRHF, L.fp32, -go, b->st) : -1;
    if (k < 0) return fail(SQJK_EHIP, "batched class kernel launch failed");
    launches += k;
    if (prof) HIPCHK(hipEventRecord(b->prof_ev[2 * li + 1], b->st));
  }
  for (int X = 0; X < kNPairClass; ++X) {
    const int np = b->np_total[X];
    if (np == 0) continue;
    const int nb = sqmd::ncart(kPairL[X][1]), nab = sqmd::ncart(kPairL[X][0]) * nb;
    hipLaunchKernelGGL(k_reduce_j_batch, dim3(grid_for((long long)np * nab, 256)), dim3(256), 0, b->st, b->d_Jp[X],
                       b->d_pairs[X], b->d_sub, b->d_inst, np, nb, nab);
  }
  int nmax = 0;
  for (const BatchGeom& g : b->geo) nmax = std::max(nmax, g.n);
  const int kgrid = std::min(b->rows_total, 65535);
  for (int spin = 0; spin < 2; ++spin)
    hipLaunchKernelGGL(k_reduce_k_batch, dim3(kgrid), dim3(kReduceKBlock), (size_t)nmax * sizeof(double), b->st,
                       b->d_prs, b->d_ncomp, b->d_nc, b->d_sub, b->d_inst, b->d_geo, b->d_row0, b->count,
                       b->rows_total, spin);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_finalize_batch, dim3(fill_grid), dim3(256), 0, b->st, b->d_AJ, b->d_AKA, b->d_AKB, b->d_J,
                     b->d_KA, b->d_KB, b->d_geo, b->d_mat_off, b->count, mt);
  HIPCHK(hipGetLastError());
  b->last_launches = launches;
  return SQJK_OK;
}

}  // namespace

extern "C" {

int sqjk_batch_create(int count, sqjk_ctx* const* ctxs, const int* restricted, sqjk_batch** out) {
  if (!out || count <= 0 || !ctxs || !restricted) return fail(SQJK_EARG, "sqjk_batch_create: invalid arguments");
  *out = nullptr;
  const int dev = ctxs[0]->device;
  for (int i = 0; i < count; ++i)
    if (!ctxs[i] || ctxs[i]->device != dev) return fail(SQJK_EARG, "sqjk_batch_create: contexts must share one device");
  HIPCHK(hipSetDevice(dev));
  sqjk_batch* b = new sqjk_batch();
  b->device = dev;
  b->count = count;
  b->ctx.assign(ctxs, ctxs + count);
  b->uhf.resize(count);
  for (int i = 0; i < count; ++i) b->uhf[i] = restricted[i] ? 0 : 1;
  b->fixed.assign(count, 0.0);
  b->scf.assign(count, 0.0);
  auto bail = [&](int rc) { sqjk_batch_destroy(b); return rc; };
  if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess)
    return bail(fail(SQJK_EHIP, "batch stream/event creation failed"));
  // concatenated sub-shells, pairs, primitive data, scales, reference shells
  std::vector<SubShell> sub;
  std::vector<PrimPair> pp;
  std::vector<PairRec> pairs[kNPairClass];
  std::vector<double> bscale;
  std::vector<int> ref;
  std::vector<long long> mat_off, pm_off, row0;
  b->geo.resize(count);
  b->desc.resize(count);
  long long mat = 0, pm = 0;
  int rows = 0;
  std::vector<std::array<size_t, kNPairClass>> s_off(count);
  size_t s_acc = 0;
  int np_acc[kNPairClass] = {};
  for (int i = 0; i < count; ++i) {
    const sqjk_ctx* c = ctxs[i];
    const int sub0 = (int)sub.size(), pp0 = (int)pp.size(), bf0 = (int)bscale.size();
    for (SubShell s : c->sub) {
      s.gbf = s.bf_off + bf0;
      sub.push_back(s);
    }
    pp.insert(pp.end(), c->pp.begin(), c->pp.end());
    bscale.insert(bscale.end(), c->bscale.begin(), c->bscale.end());
    BatchGeom& g = b->geo[i];
    g.mat_off = mat; g.pm_off = pm; g.ref0 = (int)(ref.size() / 2); g.row0 = rows;
    g.n = c->n; g.nref = c->nref; g.uhf = b->uhf[i];
    for (const RefShell& r : c->ref) { ref.push_back(r.first); ref.push_back(r.count); }
    mat_off.push_back(mat); pm_off.push_back(pm); row0.push_back(rows);
    InstDesc& d = b->desc[i];
    d.n = c->n; d.nref = c->nref; d.fixed_cut = 0.0;
    for (int X = 0; X < kNPairClass; ++X) {
      d.p0[X] = np_acc[X];
      d.np[X] = (int)c->pairs[X].size();
      for (PairRec pr : c->pairs[X]) {
        pr.a += sub0; pr.b += sub0; pr.pp_off += pp0; pr.inst = i;
        pairs[X].push_back(pr);
      }
      np_acc[X] += d.np[X];
      const int ncomp = sqmd::ncart(kPairL[X][0]) + sqmd::ncart(kPairL[X][1]);
      s_off[i][X] = s_acc;
      s_acc += (size_t)c->n * ncomp * d.np[X];
    }
    mat += (long long)c->n * c->n;
    pm += (long long)c->nref * c->nref;
    rows += c->n;
  }
  b->mat_total = mat; b->pm_total = pm; b->rows_total = rows; b->s_total = std::max<size_t>(s_acc, 1);
  int rc;
  if ((rc = dalloc(&b->d_sub, sub.size())) || (rc = dalloc(&b->d_pp, pp.size())) ||
      (rc = dalloc(&b->d_bscale, bscale.size())) || (rc = dalloc(&b->d_boys, (size_t)kBoysNT * kBoysCols)) ||
      (rc = dalloc(&b->d_ref, ref.size())) || (rc = dalloc(&b->d_geo, (size_t)count)) ||
      (rc = dalloc(&b->d_mat_off, (size_t)count)) || (rc = dalloc(&b->d_pm_off, (size_t)count)) ||
      (rc = dalloc(&b->d_row0, (size_t)count)) || (rc = dalloc(&b->d_inst, (size_t)count)) ||
      (rc = dalloc(&b->d_prs, (size_t)kNPairClass)) || (rc = dalloc(&b->d_ncomp, (size_t)kNPairClass)) ||
      (rc = dalloc(&b->d_nc, (size_t)kNPairClass)))
    return bail(rc);
  double** mats[] = {&b->d_DA, &b->d_DB, &b->d_DJ, &b->d_AJ, &b->d_AKA, &b->d_AKB, &b->d_J, &b->d_KA, &b->d_KB};
  for (double** m : mats)
    if ((rc = dalloc(m, (size_t)mat))) return bail(rc);
  if ((rc = dalloc(&b->d_PM, (size_t)std::max(pm, 1LL))) || (rc = dalloc(&b->d_SA, b->s_total)) ||
      (rc = dalloc(&b->d_SB, b->s_total)))
    return bail(rc);
  const PairRec* prs[kNPairClass];
  int ncomp[kNPairClass], nc[kNPairClass];
  for (int X = 0; X < kNPairClass; ++X) {
    b->np_total[X] = (int)pairs[X].size();
    const int nab = sqmd::ncart(kPairL[X][0]) * sqmd::ncart(kPairL[X][1]);
    b->jp_len[X] = pairs[X].size() * nab;
    if ((rc = dalloc(&b->d_pairs[X], pairs[X].size()))) return bail(rc);
    if (b->jp_len[X] && (rc = dalloc(&b->d_Jp[X], b->jp_len[X]))) return bail(rc);
    if (!pairs[X].empty())
      HIPCHK(hipMemcpyAsync(b->d_pairs[X], pairs[X].data(), pairs[X].size() * sizeof(PairRec), hipMemcpyHostToDevice,
                            b->st));
    prs[X] = b->d_pairs[X];
    nc[X] = sqmd::ncart(kPairL[X][0]);
    ncomp[X] = nc[X] + sqmd::ncart(kPairL[X][1]);
  }
  for (int i = 0; i < count; ++i) {
    InstDesc& d = b->desc[i];
    const BatchGeom& g = b->geo[i];
    d.DJ = b->d_DJ + g.mat_off; d.DA = b->d_DA + g.mat_off; d.DB = b->d_DB + g.mat_off;
    d.PM = b->d_PM + g.pm_off;
    d.AJ = b->d_AJ + g.mat_off; d.AKA = b->d_AKA + g.mat_off; d.AKB = b->d_AKB + g.mat_off;
    for (int X = 0; X < kNPairClass; ++X) {
      d.SA[X] = b->d_SA + s_off[i][X];
      d.SB[X] = b->uhf[i] ? b->d_SB + s_off[i][X] : nullptr;
    }
  }
  const std::vector<double>& boys = boys_table();
  HIPCHK(hipMemcpyAsync(b->d_sub, sub.data(), sub.size() * sizeof(SubShell), hipMemcpyHostToDevice, b->st));
  HIPCHK(hipMemcpyAsync(b->d_pp, pp.data(), pp.size() * sizeof(PrimPair), hipMemcpyHostToDevice, b->st));
  HIPCHK(hipMemcpyAsync(b->d_bscale, bscale.data(), bscale.size() * sizeof(double), hipMemcpyHostToDevice, b->st));
  HIPCHK(hipMemcpyAsync(b->d_boys, boys.data(), boys.size() * sizeof(double), hipMemcpyHostToDevice, b->st));
  HIPCHK(hipMemcpyAsync(b->d_ref, ref.data(), ref.size() * sizeof(int), hipMemcpyHostToDevice, b->st));
  HIPCHK(hipMemcpyAsync(b->d_geo, b->geo.data(), count * sizeof(BatchGeom), hipMemcpyHostToDevice, b->st));
  HIPCHK(hipMemcpyAsync(b->d_mat_off, mat_off.data(), count * sizeof(long long), hipMemcpyHostToDevice, b->st));
  HIPCHK(hipMemcpyAsync(b->d_pm_off, pm_off.data(), count * sizeof(long long), hipMemcpyHostToDevice, b->st));
  HIPCHK(hipMemcpyAsync(b->d_row0, row0.data(), count * sizeof(long long), hipMemcpyHostToDevice, b->st));
  HIPCHK(hipMemcpyAsync(b->d_prs, prs, sizeof(prs), hipMemcpyHostToDevice, b->st));
  HIPCHK(hipMemcpyAsync(b->d_ncomp, ncomp, sizeof(ncomp), hipMemcpyHostToDevice, b->st));
  HIPCHK(hipMemcpyAsync(b->d_nc, nc, sizeof(nc), hipMemcpyHostToDevice, b->st));
  hipLaunchKernelGGL(k_fill, dim3(grid_for(mat, 256)), dim3(256), 0, b->st, b->d_DA, (size_t)mat, 0.0);
  hipLaunchKernelGGL(k_fill, dim3(grid_for(mat, 256)), dim3(256), 0, b->st, b->d_DB, (size_t)mat, 0.0);
  HIPCHK(hipStreamSynchronize(b->st));
  *out = b;
  return SQJK_OK;
}

int sqjk_batch_set_density(sqjk_batch* b, int i, const double* PA, const double* PB, int on_device) {
  if (!b || i < 0 || i >= b->count || !PA || !PB) return fail(SQJK_EARG, "sqjk_batch_set_density: invalid arguments");
  HIPCHK(hipSetDevice(b->device));
  const BatchGeom& g = b->geo[i];
  const size_t bytes = (size_t)g.n * g.n * sizeof(double);
  const hipMemcpyKind k = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  HIPCHK(hipMemcpyAsync(b->d_DA + g.mat_off, PA, bytes, k, b->st));
  HIPCHK(hipMemcpyAsync(b->d_DB + g.mat_off, b->uhf[i] ? PB : PA, bytes, k, b->st));
  if (!on_device) HIPCHK(hipStreamSynchronize(b->st));  // host arrays may be reused at once
  return SQJK_OK;
}

int sqjk_batch_set_cutoffs(sqjk_batch* b, int i, double fixedCutoff, double scfCutoff) {
  if (!b || i < -1 || i >= b->count) return fail(SQJK_EARG, "sqjk_batch_set_cutoffs: invalid arguments");
  for (int k = (i < 0 ? 0 : i); k < (i < 0 ? b->count : i + 1); ++k) {
    if (b->fixed[k] != fixedCutoff || b->scf[k] != scfCutoff) b->plan_ok = false;
    b->fixed[k] = fixedCutoff;
    b->scf[k] = scfCutoff;
  }
  return SQJK_OK;
}

int sqjk_batch_set_mixed(sqjk_batch* b, double threshold) {
  if (!b || threshold < 0.0) return fail(SQJK_EARG, "sqjk_batch_set_mixed: invalid arguments");
  if (threshold != b->mixed_thr) b->plan_ok = false;
  b->mixed_thr = threshold;
  return SQJK_OK;
}

int sqjk_batch_set_tuning(sqjk_batch* b, const int* rows, const int* split_budget, const int* grid_over) {
  if (!b) return fail(SQJK_EARG, "sqjk_batch_set_tuning: null batch");
  for (sqjk_ctx* c : b->ctx)
    if (int rc = sqjk_set_tuning(c, rows, split_budget, grid_over)) return rc;
  b->plan_ok = false;
  return SQJK_OK;
}

int sqjk_batch_set_profile(sqjk_batch* b, int on) {
  if (!b) return fail(SQJK_EARG, "sqjk_batch_set_profile: null batch");
  if ((on != 0) != b->profile) b->plan_ok = false;  // the graph holds the events or not
  b->profile = on != 0;
  return SQJK_OK;
}

int sqjk_batch_build(sqjk_batch* b, int use_graph) {
  if (!b) return fail(SQJK_EARG, "sqjk_batch_build: null batch");
  HIPCHK(hipSetDevice(b->device));
  int rc;
  if (!b->plan_ok && (rc = batch_plan(b))) return rc;
  const bool prof = b->profile && !use_graph;
  if (prof && b->prof_ev.size() < 2 * b->launches.size()) {
    const size_t have = b->prof_ev.size();
    b->prof_ev.resize(2 * b->launches.size());
    for (size_t e = have; e < b->prof_ev.size(); ++e) HIPCHK(hipEventCreate(&b->prof_ev[e]));
  }
  HIPCHK(hipEventRecord(b->ev0, b->st));
  if (!use_graph || !b->exec) {
    // eager; the first graph build of a plan also runs eagerly (the
    // launchers size their grids from device queries then) and captures the
    // graph that the next builds replay
    if ((rc = batch_enqueue(b, prof))) return rc;
  } else {
    HIPCHK(hipGraphLaunch(b->exec, b->st));
  }
  HIPCHK(hipEventRecord(b->ev1, b->st));
  HIPCHK(hipStreamSynchronize(b->st));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->last_ms = ms;
  std::memset(b->cls_ms, 0, sizeof(b->cls_ms));
  if (prof)
    for (size_t li = 0; li < b->launches.size(); ++li) {
      float lm = 0.f;
      HIPCHK(hipEventElapsedTime(&lm, b->prof_ev[2 * li], b->prof_ev[2 * li + 1]));
      b->cls_ms[b->launches[li].X][b->launches[li].Y] += lm;
    }
  if (use_graph && !b->exec) {
    HIPCHK(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
    rc = batch_enqueue(b, false);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->st, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail(SQJK_EHIP, std::string("graph capture failed: ") + hipGetErrorString(e));
    b->graph = g;
    HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  }
  return SQJK_OK;
}

int sqjk_batch_result(sqjk_batch* b, int i, double* J, double* KA, double* KB, int on_device) {
  if (!b || i < 0 || i >= b->count) return fail(SQJK_EARG, "sqjk_batch_result: invalid arguments");
  HIPCHK(hipSetDevice(b->device));
  const BatchGeom& g = b->geo[i];
  const size_t bytes = (size_t)g.n * g.n * sizeof(double);
  const hipMemcpyKind k = on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  if (J) HIPCHK(hipMemcpyAsync(J, b->d_J + g.mat_off, bytes, k, b->st));
  if (KA) HIPCHK(hipMemcpyAsync(KA, b->d_KA + g.mat_off, bytes, k, b->st));
  if (KB) HIPCHK(hipMemcpyAsync(KB, b->d_KB + g.mat_off, bytes, k, b->st));
  HIPCHK(hipStreamSynchronize(b->st));
  return SQJK_OK;
}

int sqjk_batch_stats(const sqjk_batch* b, double* kernel_ms, long long* quartets, int* launches,
                     double* class_ms) {
  if (!b) return fail(SQJK_EARG, "sqjk_batch_stats: null batch");
  if (kernel_ms) *kernel_ms = b->last_ms;
  if (quartets) *quartets = b->plan_q;
  if (launches) *launches = b->last_launches;
  if (class_ms) {
    int k = 0;
    for (int X = 0; X < kNPairClass; ++X)
      for (int Y = 0; Y <= X; ++Y, ++k) class_ms[k] = b->cls_ms[X][Y];
  }
  return SQJK_OK;
}

void sqjk_batch_destroy(sqjk_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  if (b->st) (void)hipStreamSynchronize(b->st);
  batch_free_plan(b);
  void* ptrs[] = {b->d_sub, b->d_pp, b->d_bscale, b->d_boys, b->d_ref, b->d_geo, b->d_mat_off, b->d_pm_off,
                  b->d_row0, b->d_inst, b->d_prs, b->d_ncomp, b->d_nc, b->d_DA, b->d_DB, b->d_DJ, b->d_PM,
                  b->d_AJ, b->d_AKA, b->d_AKB, b->d_J, b->d_KA, b->d_KB, b->d_SA, b->d_SB};
  for (void* p : ptrs) block_put(p);
  for (int X = 0; X < kNPairClass; ++X) {
    block_put(b->d_pairs[X]);
    block_put(b->d_Jp[X]);
  }
  for (hipEvent_t e : b->prof_ev) (void)hipEventDestroy(e);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->st) (void)hipStreamDestroy(b->st);
  delete b;
}

}  // extern "C"

// ------------------------------------------------------------ drop-ins
static sqjk_ctx* g_ctx = nullptr;  // the reference keeps this state in statics too
static double g_dropin_mixed = 0.0;

}  // extern "C"

hipStream_t sqjk_stream_internal(sqjk_ctx* c) { return c ? c->st : nullptr; }
void* sq_device_block(size_t bytes) { return block_get(bytes, false); }
void sq_device_release(void* p) { block_put(p); }
double sqjk_dropin_mixed_internal() { return g_dropin_mixed; }

extern "C" {

double sqjk_get_dropin_mixed(void) { return g_dropin_mixed; }

void sqjk_set_dropin_mixed(double threshold) {
  g_dropin_mixed = threshold > 0.0 ? threshold : 0.0;
  if (g_ctx) g_ctx->mixed_thr = g_dropin_mixed;
}

static void die(const char* where, int rc) {
  std::printf("%s - error %d: %s\n", where, rc, sqjk_errmsg());
  std::fflush(stdout);
  std::exit(-1);
}

void GTO_JK_Matrix_Quartet_Parallel(int childID, int nBasis, const double* PA, const double* PB,
                                    const GTOBasis_t* gto, const double* schwarz_basis, double fixedCutoff,
                                    double* JT, double* KA, double* KB, option_t* opt) {
  if (nBasis == 0) {  // reset call (src/matrix.c:3817-3828)
    sqjk_destroy(g_ctx);
    g_ctx = nullptr;
    return;
  }
  if (!g_ctx) {
    int rc = sqjk_create(nBasis, gto, schwarz_basis, 0, &g_ctx);
    if (rc) die("GTO_JK_Matrix_Quartet_Parallel", rc);
    g_ctx->mixed_thr = g_dropin_mixed;
  }
  if (opt->nCPU > 1 && childID != 0) {
    // all work is done by child 0; keep the caller's arrays in the state the
    // reference leaves them in after a child adds nothing
    const size_t nn = (size_t)nBasis * nBasis;
    if (opt->restricted) std::memcpy(KB, KA, nn * sizeof(double));
    for (int p = 0; p < nBasis; ++p)
      for (int q = 0; q < p; ++q) {
        JT[(size_t)q * nBasis + p] = JT[(size_t)p * nBasis + q];
        KA[(size_t)q * nBasis + p] = KA[(size_t)p * nBasis + q];
        KB[(size_t)q * nBasis + p] = KB[(size_t)p * nBasis + q];
      }
    return;
  }
  int rc = sqjk_build(g_ctx, PA, PB, fixedCutoff, opt->SCFCutoff, opt->restricted, JT, KA, KB);
  if (rc) die("GTO_JK_Matrix_Quartet_Parallel", rc);
}

void uhf_getGMatrix(int nBasis, const GTOBasis_t* gto, const double* Schwarz, double cutoff, const double* PA,
                    const double* PB, double* JT, double* KA, double* KB, option_t* opt) {
  const size_t nn = (size_t)nBasis * nBasis;
  std::memset(JT, 0, nn * sizeof(double));
  std::memset(KA, 0, nn * sizeof(double));
  std::memset(KB, 0, nn * sizeof(double));
  GTO_JK_Matrix_Quartet_Parallel(0, nBasis, PA, PB, gto, Schwarz, cutoff, JT, KA, KB, opt);
}

double* create_Schwarz(int nBasis, const GTOBasis_t* gto) {
  sqjk_ctx* c = nullptr;
  int rc = sqjk_create(nBasis, gto, nullptr, 0, &c);
  if (rc) die("create_Schwarz", rc);
  double* s = (double*)std::malloc((size_t)nBasis * nBasis * sizeof(double));
  if (!s) {
    std::printf("Cannot allocate Schwarz matrix\n");
    std::exit(-1);
  }
  sqjk_schwarz(c, s);
  sqjk_destroy(c);
  return s;
}

}  // extern "C"
