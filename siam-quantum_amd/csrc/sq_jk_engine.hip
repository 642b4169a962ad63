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
// Per build (sqjk_build):
//   * for every class pair (X>=Y) the surviving quartets of bra pair i are
//     the prefix j < cnt_i of the sorted ket list; offsets are uploaded and
//     each class kernel walks its quartets with a grid-stride loop;
//   * integrals are digested into half-accumulators A (J = A + A^T) with
//     atomics; a finalize kernel forms J, -K and the host applies the
//     reference's accumulate/restricted/symmetrise semantics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sq_jk.h"
#include "sq_jk_classes.h"

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
// F_m(k*step) for m < kBoysNM: series for the top order in long double
// (F_m(T) = e^-T sum_i (2T)^i / ((2m+1)(2m+3)...(2m+2i+1))), then the
// stable downward recursion F_m = (2T F_{m+1} + e^-T)/(2m+1).
std::vector<double> make_boys_table() {
  std::vector<double> tab((size_t)kBoysNT * kBoysNM);
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
    for (int m = 0; m < kBoysNM; ++m) tab[(size_t)k * kBoysNM + m] = (double)F[m];
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
  long long* d_off = nullptr;
  int* d_kstart = nullptr;
  size_t off_cap = 0, ks_cap = 0;
  std::vector<long long> h_off;
  std::vector<int> h_kstart;
  double mixed_thr = 0.0;  // Schwarz products below this run in FP32
  std::vector<double> h_DJ, h_PM, h_J, h_KA, h_KB;
  sqjk_stats last{};
  // per class pair (X,Y) of the last build: contracted and primitive quartets
  long long cls_q[kNPairClass][kNPairClass] = {};
  long long cls_p[kNPairClass][kNPairClass] = {};
};

namespace {

template <typename T>
int dalloc(T** p, size_t count) {
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void**)p, count * sizeof(T));
  if (e != hipSuccess) return fail(SQJK_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
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
      s.l = l; s.nprim = g[f].nContract; s.prim_off = (int)pexp.size(); s.bf_off = f; s.parent = r;
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
      pr.pp_off = (int)c->pp.size(); pr.npp = A.nprim * B.nprim;
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
          c->pp.push_back(q);
        }
      c->pairs[pair_class(A.l, B.l)].push_back(pr);
    }
}

int upload_static(sqjk_ctx* c) {
  const int n = c->n;
  const size_t nn = (size_t)n * n;
  int rc;
  if ((rc = dalloc(&c->d_sub, c->sub.size()))) return rc;
  if ((rc = dalloc(&c->d_pp, c->pp.size()))) return rc;
  if ((rc = dalloc(&c->d_bscale, (size_t)n))) return rc;
  if ((rc = dalloc(&c->d_boys, (size_t)kBoysNT * kBoysNM))) return rc;
  double** mats[] = {&c->d_DJ, &c->d_DA, &c->d_DB, &c->d_AJ, &c->d_AKA, &c->d_AKB,
                     &c->d_J,  &c->d_KA, &c->d_KB, &c->d_sch};
  for (double** m : mats)
    if ((rc = dalloc(m, nn))) return rc;
  if ((rc = dalloc(&c->d_PM, (size_t)c->nref * c->nref))) return rc;
  HIPCHK(hipMemcpyAsync(c->d_sub, c->sub.data(), c->sub.size() * sizeof(SubShell), hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->d_pp, c->pp.data(), c->pp.size() * sizeof(PrimPair), hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->d_bscale, c->bscale.data(), n * sizeof(double), hipMemcpyHostToDevice, c->st));
  static std::vector<double> boys;
  static std::once_flag once;
  std::call_once(once, [] { boys = make_boys_table(); });
  HIPCHK(hipMemcpyAsync(c->d_boys, boys.data(), boys.size() * sizeof(double), hipMemcpyHostToDevice, c->st));
  return SQJK_OK;
}

int upload_pairs(sqjk_ctx* c) {
  for (int X = 0; X < kNPairClass; ++X) {
    if (c->d_pairs[X]) { (void)hipFree(c->d_pairs[X]); c->d_pairs[X] = nullptr; }
    int rc = dalloc(&c->d_pairs[X], c->pairs[X].size());
    if (rc) return rc;
    if (!c->pairs[X].empty())
      HIPCHK(hipMemcpyAsync(c->d_pairs[X], c->pairs[X].data(), c->pairs[X].size() * sizeof(PairRec),
                            hipMemcpyHostToDevice, c->st));
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
    a.nbra = (int)c->pairs[X].size();
    a.same = 1;
    ClassLauncher f = launcher_for(kPairL[X][0], kPairL[X][1], kPairL[X][0], kPairL[X][1]);
    if (!f || f(a, kSchwarz, 0, grid_for(a.total, 256), c->st) < 0)
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
  for (int X = 0; X < kNPairClass; ++X) {
    for (PairRec& pr : c->pairs[X]) pr.sch = c->sch_ref[(size_t)pr.pa * nr + pr.pb];
    std::stable_sort(c->pairs[X].begin(), c->pairs[X].end(),
                     [](const PairRec& x, const PairRec& y) { return x.sch > y.sch; });
  }
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
  std::vector<double> pexp, pcoef;
  int rc = plan_shells(c, nBasis, gto, pexp, pcoef);
  if (rc) { delete c; return rc; }
  plan_pairs(c, pexp, pcoef);
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
  if (hipStreamSynchronize(c->st) != hipSuccess) { sqjk_destroy(c); return fail(SQJK_EHIP, "sync failed"); }
  *out = c;
  return SQJK_OK;
}

int sqjk_build(sqjk_ctx* c, const double* PA, const double* PB, double fixedCutoff, double scfCutoff,
               int restricted, double* JT, double* KA, double* KB) {
  if (!c || !PA || !PB || !JT || !KA || !KB) return fail(SQJK_EARG, "sqjk_build: null argument");
  const auto t0 = std::chrono::steady_clock::now();
  HIPCHK(hipSetDevice(c->device));
  const int n = c->n, nr = c->nref;
  const size_t nn = (size_t)n * n;
  const int uhf = restricted ? 0 : 1;
  // densities: J uses PA+PB, K uses PA (and PB)
  c->h_DJ.resize(nn);
  for (size_t e = 0; e < nn; ++e) c->h_DJ[e] = PA[e] + PB[e];
  c->h_PM.assign((size_t)nr * nr, 0.0);
  for (int P = 0; P < nr; ++P)
    for (int Q = 0; Q < nr; ++Q) {
      double r = 0.0;
      for (int p = c->ref[P].first; p < c->ref[P].first + c->ref[P].count; ++p)
        for (int q = c->ref[Q].first; q < c->ref[Q].first + c->ref[Q].count; ++q) {
          const size_t e = (size_t)p * n + q;
          r = std::max(r, std::fabs(c->h_DJ[e]));
          r = std::max(r, std::fabs(PA[e]));
          r = std::max(r, std::fabs(PB[e]));
        }
      c->h_PM[(size_t)P * nr + Q] = r;
    }
  HIPCHK(hipMemcpyAsync(c->d_DJ, c->h_DJ.data(), nn * 8, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->d_DA, PA, nn * 8, hipMemcpyHostToDevice, c->st));
  if (uhf) HIPCHK(hipMemcpyAsync(c->d_DB, PB, nn * 8, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->d_PM, c->h_PM.data(), (size_t)nr * nr * 8, hipMemcpyHostToDevice, c->st));

  // surviving quartet ranges per class pair: ket pairs [0, c64) in FP64 and
  // [c64, cnt) in FP32, where c64 counts Schwarz products >= mixed_thr
  const double thr = std::max(fixedCutoff, scfCutoff);
  const double thr64 = std::max(thr, c->mixed_thr);
  struct Launch { int X, Y, fp32; size_t off, ks; long long total; };
  std::vector<Launch> plan;
  c->h_off.clear();
  c->h_kstart.clear();
  long long nq = 0, nq32 = 0;
  std::memset(c->cls_q, 0, sizeof(c->cls_q));
  std::memset(c->cls_p, 0, sizeof(c->cls_p));
  std::vector<long long> kpre;
  std::vector<size_t> c64v;
  for (int X = 0; X < kNPairClass; ++X)
    for (int Y = 0; Y <= X; ++Y) {
      const auto &bra = c->pairs[X], &ket = c->pairs[Y];
      if (bra.empty() || ket.empty()) continue;
      const size_t base = c->h_off.size();
      long long acc = 0, prim = 0;
      kpre.assign(ket.size() + 1, 0);  // prefix sums of ket primitive pairs
      for (size_t j = 0; j < ket.size(); ++j) kpre[j + 1] = kpre[j] + ket[j].npp;
      c64v.assign(bra.size(), 0);
      // both lists are sorted by sch descending, so the number of ket pairs
      // with sch_i*sch_j >= thr can only shrink as i grows: one sweep
      size_t lo = ket.size(), lo64 = ket.size();
      for (size_t i = 0; i < bra.size(); ++i) {
        c->h_off.push_back(acc);
        while (lo > 0 && !(bra[i].sch * ket[lo - 1].sch >= thr)) --lo;
        while (lo64 > 0 && !(bra[i].sch * ket[lo64 - 1].sch >= thr64)) --lo64;
        size_t cnt = lo, c64 = lo64;
        if (X == Y) { cnt = std::min(cnt, i + 1); c64 = std::min(c64, i + 1); }
        c64v[i] = c64;
        acc += (long long)c64;
        prim += (long long)bra[i].npp * kpre[cnt];
      }
      c->h_off.push_back(acc);
      if (acc > 0) plan.push_back({X, Y, 0, base, 0, acc});
      // FP32 range
      const size_t base32 = c->h_off.size(), ks = c->h_kstart.size();
      long long acc32 = 0;
      lo = ket.size();
      for (size_t i = 0; i < bra.size(); ++i) {
        c->h_off.push_back(acc32);
        c->h_kstart.push_back((int)c64v[i]);
        while (lo > 0 && !(bra[i].sch * ket[lo - 1].sch >= thr)) --lo;
        size_t cnt = lo;
        if (X == Y) cnt = std::min(cnt, i + 1);
        acc32 += (long long)(cnt > c64v[i] ? cnt - c64v[i] : 0);
      }
      c->h_off.push_back(acc32);
      if (acc32 > 0) plan.push_back({X, Y, 1, base32, ks, acc32});
      c->cls_q[X][Y] = acc + acc32;
      c->cls_p[X][Y] = prim;
      nq += acc + acc32;
      nq32 += acc32;
    }
  if (c->off_cap < c->h_off.size()) {
    if (c->d_off) (void)hipFree(c->d_off);
    c->off_cap = std::max<size_t>(c->h_off.size(), 1024) * 2;
    int rc = dalloc(&c->d_off, c->off_cap);
    if (rc) return rc;
  }
  if (c->ks_cap < c->h_kstart.size() || !c->d_kstart) {
    if (c->d_kstart) (void)hipFree(c->d_kstart);
    c->ks_cap = std::max<size_t>(c->h_kstart.size(), 1024) * 2;
    int rc = dalloc(&c->d_kstart, c->ks_cap);
    if (rc) return rc;
  }
  if (!c->h_off.empty())
    HIPCHK(hipMemcpyAsync(c->d_off, c->h_off.data(), c->h_off.size() * sizeof(long long),
                          hipMemcpyHostToDevice, c->st));
  if (!c->h_kstart.empty())
    HIPCHK(hipMemcpyAsync(c->d_kstart, c->h_kstart.data(), c->h_kstart.size() * sizeof(int),
                          hipMemcpyHostToDevice, c->st));
  const int fill_grid = grid_for((long long)nn, 256);
  hipLaunchKernelGGL(k_fill, dim3(fill_grid), dim3(256), 0, c->st, c->d_AJ, nn, 0.0);
  hipLaunchKernelGGL(k_fill, dim3(fill_grid), dim3(256), 0, c->st, c->d_AKA, nn, 0.0);
  if (uhf) hipLaunchKernelGGL(k_fill, dim3(fill_grid), dim3(256), 0, c->st, c->d_AKB, nn, 0.0);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(c->ev0, c->st));
  int launches = 0;
  for (const Launch& L : plan) {
    ClassArgs a = base_args(c);
    a.bra = c->d_pairs[L.X];
    a.ket = c->d_pairs[L.Y];
    a.off = c->d_off + L.off;
    a.kstart = L.fp32 ? c->d_kstart + L.ks : nullptr;
    a.nbra = (int)c->pairs[L.X].size();
    a.same = (L.X == L.Y);
    a.total = L.total;
    a.fixed_cut = fixedCutoff;
    ClassLauncher f = launcher_for(kPairL[L.X][0], kPairL[L.X][1], kPairL[L.Y][0], kPairL[L.Y][1]);
    const int k = f ? f(a, uhf ? kUHF : kRHF, L.fp32, grid_for(L.total, 256), c->st) : -1;
    if (k < 0) return fail(SQJK_EHIP, "class kernel launch failed");
    launches += k;
  }
  HIPCHK(hipEventRecord(c->ev1, c->st));
  hipLaunchKernelGGL(k_finalize, dim3(fill_grid), dim3(256), 0, c->st, c->d_AJ, c->d_AKA, c->d_AKB, c->d_J,
                     c->d_KA, c->d_KB, n, uhf);
  HIPCHK(hipGetLastError());
  c->h_J.resize(nn); c->h_KA.resize(nn); c->h_KB.resize(nn);
  HIPCHK(hipMemcpyAsync(c->h_J.data(), c->d_J, nn * 8, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(c->h_KA.data(), c->d_KA, nn * 8, hipMemcpyDeviceToHost, c->st));
  if (uhf) HIPCHK(hipMemcpyAsync(c->h_KB.data(), c->d_KB, nn * 8, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  // reference semantics: accumulate, KB := KA if restricted, upper := lower
  for (size_t e = 0; e < nn; ++e) {
    JT[e] += c->h_J[e];
    KA[e] += c->h_KA[e];
    if (uhf) KB[e] += c->h_KB[e];
  }
  if (!uhf) std::memcpy(KB, KA, nn * sizeof(double));
  for (int p = 0; p < n; ++p)
    for (int q = 0; q < p; ++q) {
      JT[(size_t)q * n + p] = JT[(size_t)p * n + q];
      KA[(size_t)q * n + p] = KA[(size_t)p * n + q];
      KB[(size_t)q * n + p] = KB[(size_t)p * n + q];
    }
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last.kernel_ms = ms;
  c->last.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  c->last.quartets = nq;
  c->last.quartets_fp32 = nq32;
  c->last.launches = launches;
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
                  c->d_AKA, c->d_AKB, c->d_J,   c->d_KA,    c->d_KB, c->d_sch, c->d_off, c->d_kstart};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (int X = 0; X < kNPairClass; ++X)
    if (c->d_pairs[X]) (void)hipFree(c->d_pairs[X]);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->st) (void)hipStreamDestroy(c->st);
  delete c;
}

// ------------------------------------------------------------ drop-ins
static sqjk_ctx* g_ctx = nullptr;  // the reference keeps this state in statics too
static double g_dropin_mixed = 0.0;

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
