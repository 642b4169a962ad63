// sq_scf_dev.hip - the SCF iteration of uhf() with all of its state on the
// device (src/uhf.c:1464-1783, density DIIS of src/conv.c:40-844).
//
// Per SCF context: the one-electron matrices H and S, the densities, the
// accumulated J and K, the Fock matrices, the DIIS history and the orbitals
// live in HBM; one J/K context serves both the Schwarz bounds and every
// build.  An iteration is
//   J/K build on the density change dP (device in, device out),
//   JT += dJ, K += dK, F = H + JT + K                    (one kernel),
//   P = density of the lowest nOcc solutions of F C = S C e (spectral
//       projection; eigenvectors when it does not settle),
//   energy, rms dP and the new DIIS row                  (one reduction),
//   DIIS extrapolation or damping and the next dP        (one kernel),
// and only the reduction's few scalars cross to the host, where the
// convergence test, the three-step cut-offs and the DIIS coefficients
// (a 5 x 5 solve at most) are decided exactly as in the reference.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sq_jk.h"
#include "../../include/sq_scf.h"
#include "sq_internal.h"
#include "sq_linalg_dev.h"

namespace {

constexpr int kRedBlocks = 512, kRedThreads = 256, kMaxSums = 8;
constexpr int kMaxHist = 4;  // DIIS4

__global__ void k_zero(double* p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

// JT += dJ, KA += dKA, KB += dKB, Pold = P, FA = H + JT + KA, FB = H + JT + KB
// (src/uhf.c:1600-1619).  Restricted: KB, FB, PB alias KA, FA, PA.
__global__ void k_fock(size_t nn, const double* __restrict__ H, double* JT, double* KA, double* KB,
                       const double* __restrict__ dJ, const double* __restrict__ dKA, const double* __restrict__ dKB,
                       const double* __restrict__ PA, const double* __restrict__ PB, double* PoldA, double* PoldB,
                       double* FA, double* FB, int uhf) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nn; e += (size_t)gridDim.x * blockDim.x) {
    const double j = JT[e] + dJ[e], ka = KA[e] + dKA[e], h = H[e];
    JT[e] = j;
    KA[e] = ka;
    FA[e] = h + j + ka;
    PoldA[e] = PA[e];
    if (uhf) {
      const double kb = KB[e] + dKB[e];
      KB[e] = kb;
      FB[e] = h + j + kb;
      PoldB[e] = PB[e];
    }
  }
}

struct PostArgs {
  const double *H, *JT, *KA, *KB, *PA, *PB;  // new densities (before DIIS)
  const double *P3A, *P3B;                   // newest stored DIIS input
  double *ErrA, *ErrB;                       // out: current error cur - P3
  const double* hErrA[kMaxHist];             // older errors (dot products with the current one)
  const double* hErrB[kMaxHist];
  int nold;                                  // number of older errors
  int uhf;
  size_t nn;
};

// Block partial sums, deterministic order:
//   s0 = sum (H + JT/2)(PA + PB), s1 = sum KA PA, s2 = sum KB PB,
//   s3 = sum (errA^2 + errB^2), s4+a = sum (errA_a errA + errB_a errB).
// Restricted (uhf = 0): spin B terms are left out; the host doubles them.
__global__ void __launch_bounds__(kRedThreads) k_post(PostArgs a, double* partial) {
  __shared__ double sh[kMaxSums][kRedThreads];
  double s[kMaxSums];
  for (int k = 0; k < kMaxSums; ++k) s[k] = 0.0;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < a.nn; e += (size_t)gridDim.x * blockDim.x) {
    const double pa = a.PA[e], hj = a.H[e] + 0.5 * a.JT[e];
    const double ea = pa - a.P3A[e];
    a.ErrA[e] = ea;
    s[0] += hj * pa;
    s[1] += a.KA[e] * pa;
    s[3] += ea * ea;
    for (int k = 0; k < a.nold; ++k) s[4 + k] += a.hErrA[k][e] * ea;
    if (a.uhf) {
      const double pb = a.PB[e], eb = pb - a.P3B[e];
      a.ErrB[e] = eb;
      s[0] += hj * pb;
      s[2] += a.KB[e] * pb;
      s[3] += eb * eb;
      for (int k = 0; k < a.nold; ++k) s[4 + k] += a.hErrB[k][e] * eb;
    }
  }
  for (int k = 0; k < kMaxSums; ++k) sh[k][threadIdx.x] = s[k];
  __syncthreads();
  for (int o = kRedThreads / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int k = 0; k < kMaxSums; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < kMaxSums) partial[blockIdx.x * kMaxSums + threadIdx.x] = sh[threadIdx.x][0];
}

// partial sums of all elements of PA (slot 0) and PB (slot 1)
__global__ void __launch_bounds__(kRedThreads) k_elem_sums(const double* __restrict__ PA,
                                                           const double* __restrict__ PB, size_t nn,
                                                           double* partial) {
  __shared__ double sh[2][kRedThreads];
  double a = 0.0, b = 0.0;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nn; e += (size_t)gridDim.x * blockDim.x) {
    a += PA[e];
    b += PB[e];
  }
  sh[0][threadIdx.x] = a;
  sh[1][threadIdx.x] = b;
  __syncthreads();
  for (int o = kRedThreads / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + o];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x < kMaxSums) partial[blockIdx.x * kMaxSums + threadIdx.x] = threadIdx.x < 2 ? sh[threadIdx.x][0] : 0.0;
}

// out[k] = sum over blocks of partial[b][k], in block order
__global__ void k_post_sum(const double* __restrict__ partial, int nblocks, double* out) {
  if (threadIdx.x < kMaxSums) {
    double t = 0.0;
    for (int b = 0; b < nblocks; ++b) t += partial[b * kMaxSums + threadIdx.x];
    out[threadIdx.x] = t;
  }
}

struct UpdArgs {
  double *PA, *PB;                    // in: cur, out: extrapolated density
  const double *PoldA, *PoldB;        // density of the previous iteration
  double *dPA, *dPB;                  // out: density change for the next J/K build
  double *NewA, *NewB;                // out: newest DIIS input slot
  const double* HPA[kMaxHist];        // older inputs P_a (extrapolation) or P3 (damping)
  const double* HPB[kMaxHist];
  const double* HEA[kMaxHist];        // their errors
  const double* HEB[kMaxHist];
  double c[kMaxHist];                 // coefficients of the older slots
  double ccur;                        // coefficient of cur
  int nold, damp, uhf;
  size_t nn;
};

// damp:  cur = ccur cur + c[0] P3                         (src/conv.c:108-118, 749-760)
// else:  cur = ccur cur + sum_a c[a] (P_a + err_a)         (src/conv.c:250-262, 430-446, 797-832)
// then   dP = cur - Pold, newest slot = cur                (src/uhf.c:1733-1737)
__global__ void k_update(UpdArgs a) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < a.nn; e += (size_t)gridDim.x * blockDim.x) {
    for (int sp = 0; sp < (a.uhf ? 2 : 1); ++sp) {
      double* P = sp ? a.PB : a.PA;
      double v = a.ccur * P[e];
      if (a.damp) {
        v += a.c[0] * (sp ? a.HPB[0] : a.HPA[0])[e];
      } else {
        for (int k = 0; k < a.nold; ++k)
          v += a.c[k] * ((sp ? a.HPB[k] : a.HPA[k])[e] + (sp ? a.HEB[k] : a.HEA[k])[e]);
      }
      P[e] = v;
      (sp ? a.NewB : a.NewA)[e] = v;
      (sp ? a.dPB : a.dPA)[e] = v - (sp ? a.PoldB : a.PoldA)[e];
    }
  }
}

unsigned grid_nn(size_t nn) { return (unsigned)std::max<size_t>(1, std::min<size_t>((nn + 255) / 256, 8192)); }

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// host side of the 5 x 5 DIIS4 system (the reference hands it to dgesv,
// src/conv.c:775-786): Gaussian elimination with partial pivoting
void solve_small(int n, double* A, double* b) {
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(A[r * n + c]) > std::fabs(A[p * n + c])) p = r;
    if (p != c) {
      for (int k = 0; k < n; ++k) std::swap(A[c * n + k], A[p * n + k]);
      std::swap(b[c], b[p]);
    }
    for (int r = c + 1; r < n; ++r) {
      const double f = A[r * n + c] / A[c * n + c];
      for (int k = c; k < n; ++k) A[r * n + k] -= f * A[c * n + k];
      b[r] -= f * b[c];
    }
  }
  for (int r = n - 1; r >= 0; --r) {
    double s = b[r];
    for (int k = r + 1; k < n; ++k) s -= A[r * n + k] * b[k];
    b[r] = s / A[r * n + r];
  }
}

}  // namespace

struct sqscf_ctx {
  int device = 0, n = 0, cap = 0;
  hipStream_t st = nullptr;  // the J/K context's stream
  sqjk_ctx* jk = nullptr;
  sqla::Work lw;
  // n x n device matrices
  double *H = nullptr, *S = nullptr, *PA = nullptr, *PB = nullptr, *PoldA = nullptr, *PoldB = nullptr;
  double *dPA = nullptr, *dPB = nullptr, *JT = nullptr, *KA = nullptr, *KB = nullptr;
  double *Jd = nullptr, *KAd = nullptr, *KBd = nullptr, *FA = nullptr, *FB = nullptr;
  double *CA = nullptr, *CB = nullptr, *eA = nullptr, *eB = nullptr;
  double* hp[kMaxHist][2] = {};  // DIIS inputs
  double* he[kMaxHist][2] = {};  // DIIS errors
  double *partial = nullptr, *sums = nullptr;
  bool have_density = false, have_orbitals = false;
  int nEA = 0, nEB = 0, restricted = 1;
  std::string err;
};

namespace {

thread_local std::string g_scf_err;
int scf_fail(int code, const std::string& m) {
  g_scf_err = m;
  return code;
}
#define SCFCHK(call)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess) return scf_fail(SQJK_EHIP, std::string(#call ": ") + hipGetErrorString(e_)); \
  } while (0)

void free_mats(sqscf_ctx* c) {
  double** m[] = {&c->H, &c->S, &c->PA, &c->PB, &c->PoldA, &c->PoldB, &c->dPA, &c->dPB, &c->JT, &c->KA,
                  &c->KB, &c->Jd, &c->KAd, &c->KBd, &c->FA, &c->FB, &c->CA, &c->CB, &c->eA, &c->eB};
  for (double** p : m) {
    sq_device_release(*p);
    *p = nullptr;
  }
  for (int s = 0; s < kMaxHist; ++s)
    for (int k = 0; k < 2; ++k) {
      sq_device_release(c->hp[s][k]);
      sq_device_release(c->he[s][k]);
      c->hp[s][k] = c->he[s][k] = nullptr;
    }
  c->cap = 0;
}

int alloc_mats(sqscf_ctx* c, int n) {
  if (c->cap >= n) return SQJK_OK;
  free_mats(c);
  const size_t nn = (size_t)n * n;
  double** m[] = {&c->H, &c->S, &c->PA, &c->PB, &c->PoldA, &c->PoldB, &c->dPA, &c->dPB, &c->JT,
                  &c->KA, &c->KB, &c->Jd, &c->KAd, &c->KBd, &c->FA, &c->FB, &c->CA, &c->CB};
  auto get = [&](double** p, size_t bytes) { return (*p = static_cast<double*>(sq_device_block(bytes))) != nullptr; };
  for (double** p : m)
    if (!get(p, nn * 8)) return scf_fail(SQJK_ENOMEM, "SCF device allocation failed");
  if (!get(&c->eA, (size_t)n * 8) || !get(&c->eB, (size_t)n * 8))
    return scf_fail(SQJK_ENOMEM, "SCF device allocation failed");
  for (int s = 0; s < kMaxHist; ++s)
    for (int k = 0; k < 2; ++k)
      if (!get(&c->hp[s][k], nn * 8) || !get(&c->he[s][k], nn * 8))
        return scf_fail(SQJK_ENOMEM, "SCF device allocation failed");
  if (!c->partial && (hipMalloc(&c->partial, kRedBlocks * kMaxSums * 8) != hipSuccess ||
                      hipMalloc(&c->sums, kMaxSums * 8) != hipSuccess))
    return scf_fail(SQJK_ENOMEM, "SCF device allocation failed");
  c->cap = n;
  return SQJK_OK;
}

// density of F's lowest nOcc solutions into P (projection, else orbitals
// into C, e and P = C C^T)
int density_from_fock(sqscf_ctx* c, int nOcc, const double* F, double* P, double* C, double* e, bool eigen_always) {
  const int n = c->n;
  if (!eigen_always) {
    const int rc = sqla::density_sp2(c->lw, n, nOcc, F, P);
    if (rc == 0) return SQJK_OK;
    if (rc < 0) return scf_fail(SQJK_EHIP, "device density projection failed");
  }
  if (sqla::eigen(c->lw, n, F, e, C)) return scf_fail(SQJK_EHIP, "device eigensolver failed");
  if (sqla::density(c->lw, n, nOcc, C, P)) return scf_fail(SQJK_EHIP, "device density build failed");
  return SQJK_OK;
}

// S-normalise the first nOcc orbitals (column-major C) in the current S
// and form P = C_occ C_occ^T (the CACB guess, src/uhf.c:1358-1364)
__global__ void k_orb_norm(double* C, const double* __restrict__ S, int n, int nocc) {
  __shared__ double sh[256];
  for (int k = blockIdx.x; k < nocc; k += gridDim.x) {
    const double* ck = C + (size_t)k * n;
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      double t = 0.0;
      for (int j = 0; j < n; ++j) t += S[(size_t)i * n + j] * ck[j];
      s += ck[i] * t;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
      __syncthreads();
    }
    const double f = 1.0 / sqrt(sh[0]);
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) C[(size_t)k * n + i] *= f;
    __syncthreads();
  }
}

}  // namespace

extern "C" {

const char* sqscf_errmsg(void) { return g_scf_err.c_str(); }

int sqscf_create(int device, sqscf_ctx** out) {
  if (!out) return scf_fail(SQJK_EARG, "sqscf_create: null argument");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    return scf_fail(SQJK_ENODEV, "no HIP device is visible; this library has no CPU path");
  if (device < 0 || device >= ndev) return scf_fail(SQJK_EARG, "sqscf_create: device ordinal out of range");
  sqscf_ctx* c = new sqscf_ctx();
  c->device = device;
  *out = c;
  return SQJK_OK;
}

int sqscf_set_system(sqscf_ctx* c, int n, const struct GTOBasis_t* gto, const double* H, const double* S) {
  if (!c || n <= 0 || !gto || !H || !S) return scf_fail(SQJK_EARG, "sqscf_set_system: invalid arguments");
  SCFCHK(hipSetDevice(c->device));
  int rc;
  if (c->jk && c->n == n && sqjk_update_geometry(c->jk, n, gto) == SQJK_OK) {
    // same basis at a new geometry: the J/K context is kept
  } else {
    // the linear-algebra workspace (its rocBLAS handle) is bound to the old
    // context's stream: release it while that stream still exists
    if (c->lw.st) sqla::release(c->lw);
    if (c->jk) sqjk_destroy(c->jk);
    c->jk = nullptr;
    if ((rc = sqjk_create(n, gto, nullptr, c->device, &c->jk))) return scf_fail(rc, sqjk_errmsg());
    c->have_density = c->have_orbitals = false;
  }
  c->st = sqjk_stream_internal(c->jk);
  if (c->lw.st && c->lw.st != c->st) sqla::release(c->lw);
  if (sqla::ensure(c->lw, n, c->st)) return scf_fail(SQJK_ENOMEM, "SCF linear algebra workspace failed");
  if ((rc = alloc_mats(c, n))) return rc;
  if (c->n != n) c->have_density = c->have_orbitals = false;
  c->n = n;
  const size_t nn = (size_t)n * n;
  SCFCHK(hipMemcpyAsync(c->H, H, nn * 8, hipMemcpyHostToDevice, c->st));
  SCFCHK(hipMemcpyAsync(c->S, S, nn * 8, hipMemcpyHostToDevice, c->st));
  if (int e = sqla::set_metric(c->lw, n, c->S))
    return scf_fail(SQJK_EHIP, std::string("metric from the overlap matrix (n = ") + std::to_string(n) + "): " +
                                   sqla::what(c->lw, e));
  SCFCHK(hipStreamSynchronize(c->st));
  return SQJK_OK;
}

int sqscf_set_density(sqscf_ctx* c, const double* PA, const double* PB) {
  if (!c || !c->n || !PA || !PB) return scf_fail(SQJK_EARG, "sqscf_set_density: invalid arguments");
  const size_t nn = (size_t)c->n * c->n;
  SCFCHK(hipSetDevice(c->device));
  SCFCHK(hipMemcpyAsync(c->PA, PA, nn * 8, hipMemcpyHostToDevice, c->st));
  SCFCHK(hipMemcpyAsync(c->PB, PB, nn * 8, hipMemcpyHostToDevice, c->st));
  SCFCHK(hipStreamSynchronize(c->st));
  c->have_density = true;
  return SQJK_OK;
}

int sqscf_guess_from_orbitals(sqscf_ctx* c, int nEA, int nEB) {
  if (!c || !c->n) return scf_fail(SQJK_EARG, "sqscf_guess_from_orbitals: no system");
  if (!c->have_orbitals) return scf_fail(SQJK_EARG, "sqscf_guess_from_orbitals: no orbitals of an earlier solve");
  const int n = c->n;
  SCFCHK(hipSetDevice(c->device));
  if (nEA) hipLaunchKernelGGL(k_orb_norm, dim3(std::min(nEA, 1024)), dim3(256), 0, c->st, c->CA, c->S, n, nEA);
  if (nEB) hipLaunchKernelGGL(k_orb_norm, dim3(std::min(nEB, 1024)), dim3(256), 0, c->st, c->CB, c->S, n, nEB);
  SCFCHK(hipGetLastError());
  if (sqla::density(c->lw, n, nEA, c->CA, c->PA) || sqla::density(c->lw, n, nEB, c->CB, c->PB))
    return scf_fail(SQJK_EHIP, "device density build failed");
  SCFCHK(hipStreamSynchronize(c->st));
  c->have_density = true;
  return SQJK_OK;
}

int sqscf_run(sqscf_ctx* c, const sqscf_params* prm, sqscf_result* res) {
  if (!c || !prm || !res || !c->n) return scf_fail(SQJK_EARG, "sqscf_run: invalid arguments");
  if (!c->have_density) return scf_fail(SQJK_EARG, "sqscf_run: no initial density");
  const int n = c->n;
  const size_t nn = (size_t)n * n;
  const int uhf = prm->restricted ? 0 : 1;
  if (prm->restricted && prm->nEA != prm->nEB)
    return scf_fail(SQJK_EARG, "restricted SCF needs the same number of electrons in each spin");
  if (prm->SCFDrag < 0.0 || prm->SCFDrag > 1.0) return scf_fail(SQJK_EARG, "SCFDrag outside [0, 1]");
  int K;  // DIIS history: the damping and the 2-, 3- and 4-point density DIIS
  switch (prm->convMethod) {
    case CONVMETHOD_DAMPING: K = 1; break;
    case CONVMETHOD_DIIS2: K = 2; break;
    case CONVMETHOD_DIIS3: K = 3; break;
    case CONVMETHOD_DIIS4: K = 4; break;
    default: return scf_fail(SQJK_EARG, "unknown SCF convergence method");
  }
  SCFCHK(hipSetDevice(c->device));
  if (sqjk_set_mixed(c->jk, prm->mixedThreshold)) return scf_fail(SQJK_EARG, sqjk_errmsg());
  c->nEA = prm->nEA;
  c->nEB = prm->nEB;
  c->restricted = prm->restricted;
  const double t_start = now_s();
  // restricted: one spin (the beta matrices are not touched until the end)
  // cut-offs from the initial density (src/uhf.c:1465-1472, 1573-1589)
  double hs[kMaxSums];
  hipLaunchKernelGGL(k_elem_sums, dim3(kRedBlocks), dim3(kRedThreads), 0, c->st, c->PA, uhf ? c->PB : c->PA, nn,
                     c->partial);
  hipLaunchKernelGGL(k_post_sum, dim3(1), dim3(64), 0, c->st, c->partial, kRedBlocks, c->sums);
  SCFCHK(hipGetLastError());
  SCFCHK(hipMemcpyAsync(hs, c->sums, sizeof(hs), hipMemcpyDeviceToHost, c->st));
  SCFCHK(hipStreamSynchronize(c->st));
  const double sA = hs[0], sB = hs[1];
  auto getcutoff = [&](double acc) { return acc / (sA * sA + sA * sB + sB * sB) / 50.0; };
  const double realCutoff = getcutoff(prm->SCFConv);
  double cutA = 0.0, cutB = 0.0, cutC = 0.0, thisCutoff = realCutoff;
  if (prm->threeStep) {
    cutA = getcutoff(SCFACCURACY_3STEP_A);
    cutB = getcutoff(SCFACCURACY_3STEP_B);
    cutC = getcutoff(SCFACCURACY_3STEP_C);
    thisCutoff = cutA;
  }
  res->realCutoff = realCutoff;

  // DIIS history (src/conv.c): inputs I[0..K-1] (I[K-1] the newest stored
  // density) and errors E[0..K-1]; an iteration writes its error cur - I[K-1]
  // into E[K-1] (the buffer freed by the last rotation), extrapolates with
  // E[0..K-2] and I[0..K-2], then rotates both lists by one and stores cur
  // in the freed input buffer.  G holds the Gram matrix <E[a], E[b]> of the
  // slots 0..K-2 after the last rotation; each iteration adds one row.
  double* I[kMaxHist][2];
  double* E[kMaxHist][2];
  for (int s = 0; s < kMaxHist; ++s)
    for (int k = 0; k < 2; ++k) { I[s][k] = c->hp[s][k]; E[s][k] = c->he[s][k]; }
  for (int s = 0; s < K; ++s)
    for (int k = 0; k < 1 + uhf; ++k)
      hipLaunchKernelGGL(k_zero, dim3(grid_nn(nn)), dim3(256), 0, c->st, E[s][k], nn);
  // first call of the DIIS: store the input density (nIter 0)
  SCFCHK(hipMemcpyAsync(I[K - 1][0], c->PA, nn * 8, hipMemcpyDeviceToDevice, c->st));
  if (uhf) SCFCHK(hipMemcpyAsync(I[K - 1][1], c->PB, nn * 8, hipMemcpyDeviceToDevice, c->st));
  // zero J/K accumulators, dP = P
  hipLaunchKernelGGL(k_zero, dim3(grid_nn(nn)), dim3(256), 0, c->st, c->JT, nn);
  hipLaunchKernelGGL(k_zero, dim3(grid_nn(nn)), dim3(256), 0, c->st, c->KA, nn);
  hipLaunchKernelGGL(k_zero, dim3(grid_nn(nn)), dim3(256), 0, c->st, c->KB, nn);
  SCFCHK(hipGetLastError());
  SCFCHK(hipMemcpyAsync(c->dPA, c->PA, nn * 8, hipMemcpyDeviceToDevice, c->st));
  if (uhf) SCFCHK(hipMemcpyAsync(c->dPB, c->PB, nn * 8, hipMemcpyDeviceToDevice, c->st));

  const double drag = prm->SCFDrag;
  double Etot = 0.0, dE = 0.0, avgdP = 0.0;
  double G[kMaxHist][kMaxHist] = {};
  int iter = 0, notConverged = 1, ncall = 0;
  do {
    iter++;
    ncall++;
    int rc = sqjk_build_device(c->jk, c->dPA, uhf ? c->dPB : c->dPA, thisCutoff, realCutoff, prm->restricted,
                               c->Jd, c->KAd, uhf ? c->KBd : c->KAd);
    if (rc) return scf_fail(rc, sqjk_errmsg());
    hipLaunchKernelGGL(k_fock, dim3(grid_nn(nn)), dim3(256), 0, c->st, nn, c->H, c->JT, c->KA, c->KB, c->Jd,
                       c->KAd, c->KBd, c->PA, c->PB, c->PoldA, c->PoldB, c->FA, c->FB, uhf);
    SCFCHK(hipGetLastError());
    // new densities from the Fock matrices
    if ((rc = density_from_fock(c, prm->nEA, c->FA, c->PA, c->CA, c->eA, prm->eigenEveryIter))) return rc;
    if (uhf && (rc = density_from_fock(c, prm->nEB, c->FB, c->PB, c->CB, c->eB, prm->eigenEveryIter))) return rc;
    if (prm->eigenEveryIter) c->have_orbitals = true;
    // energy, rms dP and the new row <E[a], cur error>
    const int nold = K - 1;
    PostArgs pa{};
    pa.H = c->H; pa.JT = c->JT; pa.KA = c->KA; pa.KB = c->KB; pa.PA = c->PA; pa.PB = c->PB;
    pa.P3A = I[K - 1][0]; pa.P3B = I[K - 1][1];
    pa.ErrA = E[K - 1][0]; pa.ErrB = E[K - 1][1];
    for (int a = 0; a < nold; ++a) { pa.hErrA[a] = E[a][0]; pa.hErrB[a] = E[a][1]; }
    pa.nold = nold;
    pa.uhf = uhf;
    pa.nn = nn;
    hipLaunchKernelGGL(k_post, dim3(kRedBlocks), dim3(kRedThreads), 0, c->st, pa, c->partial);
    hipLaunchKernelGGL(k_post_sum, dim3(1), dim3(64), 0, c->st, c->partial, kRedBlocks, c->sums);
    SCFCHK(hipGetLastError());
    SCFCHK(hipMemcpyAsync(hs, c->sums, sizeof(hs), hipMemcpyDeviceToHost, c->st));
    SCFCHK(hipStreamSynchronize(c->st));
    if (!uhf) {  // the beta terms equal the alpha ones
      hs[0] *= 2.0; hs[2] = hs[1]; hs[3] *= 2.0;
      for (int a = 0; a < nold; ++a) hs[4 + a] *= 2.0;
    }
    dE = Etot;
    Etot = prm->Econst + hs[0] + 0.5 * hs[1] + 0.5 * hs[2];  // src/uhf.c:246-277
    dE = Etot - dE;
    avgdP = std::sqrt(hs[3] / n / n / 2.0);
    // Gram matrix over the slots 0..K-1 (K-1 = the current error)
    double Bn[kMaxHist][kMaxHist] = {};
    for (int a = 0; a < K - 1; ++a)
      for (int b = 0; b < K - 1; ++b) Bn[a][b] = G[a][b];
    for (int a = 0; a < K - 1; ++a) Bn[a][K - 1] = Bn[K - 1][a] = hs[4 + a];
    Bn[K - 1][K - 1] = hs[3];
    // DIIS coefficients: damping for the first K-1 calls and always for the
    // damping method, then the K-point extrapolation
    UpdArgs ua{};
    ua.PA = c->PA; ua.PB = c->PB; ua.PoldA = c->PoldA; ua.PoldB = c->PoldB; ua.dPA = c->dPA; ua.dPB = c->dPB;
    ua.uhf = uhf;
    ua.nn = nn;
    if (K == 1 || ncall <= K - 1) {  // src/conv.c:108-118, 234-242, 406-412, 749-760
      ua.damp = 1;
      ua.ccur = drag;
      ua.c[0] = 1.0 - drag;
      ua.HPA[0] = I[K - 1][0];
      ua.HPB[0] = I[K - 1][1];
    } else {
      double cc[kMaxHist] = {};
      if (K == 2) {  // src/conv.c:243-249
        double c1 = (Bn[0][0] - Bn[0][1]) / (Bn[1][1] - Bn[0][1]);
        double c0 = 1.0 / (1.0 + c1);
        c1 *= c0;
        if (Bn[1][1] > Bn[0][0]) { c1 = drag; c0 = 1.0 - drag; }
        cc[0] = c0; cc[1] = c1;
      } else if (K == 3) {  // src/conv.c:413-428
        const double B00 = Bn[0][0], B11 = Bn[1][1], B22 = Bn[2][2], B01 = Bn[0][1], B02 = Bn[0][2], B12 = Bn[1][2];
        const double den = B12 * B12 - B01 * B12 + B02 * B11 + B01 * B22 - B02 * B12 - B11 * B22;
        double c1 = (B02 * B02 + B00 * B12 - B01 * B02 - B00 * B22 + B01 * B22 - B02 * B12) / den;
        double c2 = (B01 * B01 - B00 * B11 - B01 * B02 + B02 * B11 + B12 * B00 - B12 * B01) / den;
        double c0 = 1.0 / (1.0 + c1 + c2);
        c1 *= c0;
        c2 *= c0;
        if (B22 > B11) { c2 = drag; c1 = 1.0 - drag; c0 = 0.0; }
        cc[0] = c0; cc[1] = c1; cc[2] = c2;
      } else {  // src/conv.c:775-795: [B -1; -1 0] (c, l) = (0 0 0 0 -1)
        double A[25], r[5] = {0, 0, 0, 0, -1};
        for (int a = 0; a < 5; ++a)
          for (int b = 0; b < 5; ++b) A[a * 5 + b] = (a < 4 && b < 4) ? Bn[a][b] : ((a == 4 && b == 4) ? 0.0 : -1.0);
        solve_small(5, A, r);
        for (int a = 0; a < 4; ++a) cc[a] = r[a];
        if (Bn[3][3] > Bn[2][2]) { cc[3] = drag; cc[2] = 1.0 - drag; cc[1] = 0.0; cc[0] = 0.0; }
      }
      ua.damp = 0;
      ua.ccur = cc[K - 1];
      ua.nold = K - 1;
      for (int a = 0; a < K - 1; ++a) {
        ua.c[a] = cc[a];
        ua.HPA[a] = I[a][0]; ua.HPB[a] = I[a][1];
        ua.HEA[a] = E[a][0]; ua.HEB[a] = E[a][1];
      }
    }
    // rotate by one; the freed oldest input buffer takes the new density
    for (int a = 0; a < K - 1; ++a)
      for (int b = 0; b < K - 1; ++b) G[a][b] = Bn[a + 1][b + 1];
    for (int k = 0; k < 2; ++k) {
      double* oe = E[0][k];
      double* oi = I[0][k];
      for (int s = 0; s + 1 < K; ++s) { E[s][k] = E[s + 1][k]; I[s][k] = I[s + 1][k]; }
      E[K - 1][k] = oe;
      I[K - 1][k] = oi;
    }
    ua.NewA = I[K - 1][0];
    ua.NewB = I[K - 1][1];
    hipLaunchKernelGGL(k_update, dim3(grid_nn(nn)), dim3(256), 0, c->st, ua);
    SCFCHK(hipGetLastError());
    notConverged = std::fabs(dE) > prm->SCFConv || avgdP > prm->SCFConv;
    if (prm->onIter) {
      SCFCHK(hipStreamSynchronize(c->st));
      prm->onIter(prm->user, c, iter, Etot, dE, avgdP);
    }
    if (iter >= prm->SCFMax) break;
    if (prm->threeStep) {
      if (!notConverged) break;
      if ((thisCutoff == cutA && std::fabs(dE) <= SCFACCURACY_3STEP_A && avgdP <= SCFACCURACY_3STEP_A) ||
          (thisCutoff == cutB && std::fabs(dE) <= SCFACCURACY_3STEP_B && avgdP <= SCFACCURACY_3STEP_B) ||
          (thisCutoff == cutC && std::fabs(dE) <= SCFACCURACY_3STEP_C && avgdP <= SCFACCURACY_3STEP_C)) {
        if (thisCutoff == cutA) thisCutoff = cutB;
        else if (thisCutoff == cutB) thisCutoff = cutC;
        else if (thisCutoff == cutC) thisCutoff = realCutoff;
      }
    }
  } while (notConverged);
  // orbitals and orbital energies of the last Fock matrices
  if (sqla::eigen(c->lw, n, c->FA, c->eA, c->CA)) return scf_fail(SQJK_EHIP, "device eigensolver failed");
  if (uhf) {
    if (sqla::eigen(c->lw, n, c->FB, c->eB, c->CB)) return scf_fail(SQJK_EHIP, "device eigensolver failed");
  } else {
    SCFCHK(hipMemcpyAsync(c->CB, c->CA, nn * 8, hipMemcpyDeviceToDevice, c->st));
    SCFCHK(hipMemcpyAsync(c->eB, c->eA, (size_t)n * 8, hipMemcpyDeviceToDevice, c->st));
    SCFCHK(hipMemcpyAsync(c->PB, c->PA, nn * 8, hipMemcpyDeviceToDevice, c->st));
  }
  SCFCHK(hipStreamSynchronize(c->st));
  c->have_orbitals = true;
  res->energy = Etot;
  res->dE = dE;
  res->rmsDP = avgdP;
  res->iterations = iter;
  res->converged = !notConverged;
  res->seconds = now_s() - t_start;
  return SQJK_OK;
}

int sqscf_orbitals(sqscf_ctx* c, double* CA, double* CB, double* eA, double* eB) {
  if (!c || !c->n) return scf_fail(SQJK_EARG, "sqscf_orbitals: no system");
  if (!c->have_orbitals) return scf_fail(SQJK_EARG, "sqscf_orbitals: no orbitals yet");
  const size_t nn = (size_t)c->n * c->n;
  SCFCHK(hipSetDevice(c->device));
  if (CA) SCFCHK(hipMemcpyAsync(CA, c->CA, nn * 8, hipMemcpyDeviceToHost, c->st));
  if (CB) SCFCHK(hipMemcpyAsync(CB, c->restricted ? c->CA : c->CB, nn * 8, hipMemcpyDeviceToHost, c->st));
  if (eA) SCFCHK(hipMemcpyAsync(eA, c->eA, (size_t)c->n * 8, hipMemcpyDeviceToHost, c->st));
  if (eB) SCFCHK(hipMemcpyAsync(eB, c->restricted ? c->eA : c->eB, (size_t)c->n * 8, hipMemcpyDeviceToHost, c->st));
  SCFCHK(hipStreamSynchronize(c->st));
  return SQJK_OK;
}

int sqscf_density(sqscf_ctx* c, double* PA, double* PB) {
  if (!c || !c->n || !c->have_density) return scf_fail(SQJK_EARG, "sqscf_density: no density");
  const size_t nn = (size_t)c->n * c->n;
  SCFCHK(hipSetDevice(c->device));
  if (PA) SCFCHK(hipMemcpyAsync(PA, c->PA, nn * 8, hipMemcpyDeviceToHost, c->st));
  if (PB) SCFCHK(hipMemcpyAsync(PB, c->restricted ? c->PA : c->PB, nn * 8, hipMemcpyDeviceToHost, c->st));
  SCFCHK(hipStreamSynchronize(c->st));
  return SQJK_OK;
}

int sqscf_jk_context(sqscf_ctx* c, sqjk_ctx** jk) {
  if (!c || !jk) return scf_fail(SQJK_EARG, "sqscf_jk_context: null argument");
  *jk = c->jk;
  return SQJK_OK;
}

void sqscf_destroy(sqscf_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->st) (void)hipStreamSynchronize(c->st);
  sqla::release(c->lw);
  free_mats(c);
  if (c->partial) (void)hipFree(c->partial);
  if (c->sums) (void)hipFree(c->sums);
  if (c->jk) sqjk_destroy(c->jk);
  delete c;
}

}  // extern "C"
