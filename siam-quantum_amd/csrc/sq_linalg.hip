// sq_linalg.hip - dense FP64 linear algebra of the SCF on the device.
//
// Generalised symmetric eigenproblem F C = S C e on the device, behind the
// reference's src/lin.h surface, and the same operations on device-resident
// matrices for the SCF that stays on the device (sqla:: in sq_linalg_dev.h).  S is fixed during an SCF, so its Cholesky
// factor S = U^T U and X = U^-1 are computed once per metric (rocSOLVER
// dpotrf/dtrtri) and every solve is X^T F X (two DGEMMs), rocSOLVER dsyevd
// and C = X C' (one DGEMM):
//   gen_sym_eigen      <- src/lin.h:20 (src/lin.c:66-177, LAPACK dspgvd)
//   gen_sym_SomeEigen  <- src/lin.h:24 (src/lin.c:195-307, LAPACK dspgvx)
//   linear_solver      <- src/lin.h:30 (src/lin.c:320-391, LAPACK dgesv)
// Orbitals are returned orbital-major, C[k*n + i] = coefficient of basis
// function i in orbital k, eigenvalues ascending, as in the reference.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/sq_scf.h"
#include "sq_linalg_dev.h"

namespace {

// strictly lower triangle (column-major) set to zero
__global__ void k_zero_lower(double* X, int n) {
  const size_t nn = (size_t)n * n;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nn; e += (size_t)gridDim.x * blockDim.x) {
    const int j = (int)(e / n), i = (int)(e - (size_t)j * n);
    if (i > j) X[e] = 0.0;
  }
}

// One workgroup: red[0] = min_i (F_ii - sum_{j!=i} |F_ij|), red[1] = max_i
// (F_ii + sum_{j!=i} |F_ij|), Gershgorin bounds of the spectrum of F.
constexpr int kRedBlock = 1024;
__global__ void __launch_bounds__(kRedBlock) k_gershgorin(const double* F, int n, double* red) {
  __shared__ double lo[kRedBlock], hi[kRedBlock];
  double l = 1e300, h = -1e300;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    double r = 0.0;
    for (int j = 0; j < n; ++j)
      if (j != i) r += fabs(F[(size_t)j * n + i]);
    const double d = F[(size_t)i * n + i];
    l = fmin(l, d - r);
    h = fmax(h, d + r);
  }
  lo[threadIdx.x] = l;
  hi[threadIdx.x] = h;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      lo[threadIdx.x] = fmin(lo[threadIdx.x], lo[threadIdx.x + o]);
      hi[threadIdx.x] = fmax(hi[threadIdx.x], hi[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { red[0] = lo[0]; red[1] = hi[0]; }
}

// X = (e_max I - F) / (e_max - e_min): the spectrum mapped into [0, 1],
// lowest eigenvalues of F nearest 1
__global__ void k_sp2_init(double* X, const double* F, int n, const double* red) {
  const double emin = red[0], emax = red[1], s = 1.0 / (emax - emin);
  const size_t nn = (size_t)n * n;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nn; e += (size_t)gridDim.x * blockDim.x) {
    const int j = (int)(e / n), i = (int)(e - (size_t)j * n);
    X[e] = ((i == j ? emax : 0.0) - F[e]) * s;
  }
}

// One workgroup: red[2] = tr X, red[3] = tr X2 (X2 = X X from the DGEMM)
__global__ void __launch_bounds__(kRedBlock) k_sp2_trace(const double* X, const double* X2, int n, double* red) {
  __shared__ double t1[kRedBlock], t2[kRedBlock];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    a += X[(size_t)i * n + i];
    b += X2[(size_t)i * n + i];
  }
  t1[threadIdx.x] = a;
  t2[threadIdx.x] = b;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      t1[threadIdx.x] += t1[threadIdx.x + o];
      t2[threadIdx.x] += t2[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { red[2] = t1[0]; red[3] = t2[0]; }
}

// SP2 step: X <- X^2 if that brings the trace nearer nOcc, else 2X - X^2
__global__ void k_sp2_update(double* X, const double* X2, size_t nn, const double* red, int nocc) {
  const double tr = red[2], tr2 = red[3];
  const bool sq = fabs(tr2 - nocc) < fabs(2.0 * tr - tr2 - nocc);
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nn; e += (size_t)gridDim.x * blockDim.x)
    X[e] = sq ? X2[e] : 2.0 * X[e] - X2[e];
}

}  // namespace

namespace sqla {

namespace {
int read_info(Work& w) {
  rocblas_int info = 0;
  hipError_t e = hipMemcpyAsync(&info, w.info, sizeof(info), hipMemcpyDeviceToHost, w.st);
  if (e == hipSuccess) e = hipStreamSynchronize(w.st);
  if (e != hipSuccess) {
    w.last_hip = e;
    return -3;
  }
  w.last_info = info;
  return info != 0 ? -5 : 0;
}

unsigned grid_nn(size_t nn) { return (unsigned)std::min<size_t>((nn + 255) / 256, 4096); }

int gemm(Work& w, rocblas_operation ta, rocblas_operation tb, int n, int k, const double* a, const double* b,
         double* c) {
  const double one = 1.0, zero = 0.0;
  return rocblas_dgemm(w.h, ta, tb, n, n, k, &one, a, n, b, n, &zero, c, n) == rocblas_status_success ? 0 : -4;
}
}  // namespace

int ensure(Work& w, int n, hipStream_t st) {
  if (!w.st) {
    if (st) {
      w.st = st;
    } else {
      if (hipStreamCreateWithFlags(&w.st, hipStreamNonBlocking) != hipSuccess) return -1;
      w.own_stream = true;
    }
  }
  if (!w.h) {
    if (rocblas_create_handle(&w.h) != rocblas_status_success) return -1;
    if (rocblas_set_stream(w.h, w.st) != rocblas_status_success) return -1;
  }
  if (w.n >= n) return 0;
  if (w.A) {
    (void)hipStreamSynchronize(w.st);
    (void)hipFree(w.A); (void)hipFree(w.B); (void)hipFree(w.W); (void)hipFree(w.E);
    (void)hipFree(w.P); (void)hipFree(w.X); (void)hipFree(w.info); (void)hipFree(w.red);
    w.A = w.B = w.W = w.E = w.P = w.X = w.red = nullptr;
    w.info = nullptr;
  }
  w.metric_n = 0;
  const size_t nn = (size_t)n * n;
  if (hipMalloc(&w.A, nn * 8) != hipSuccess || hipMalloc(&w.B, nn * 8) != hipSuccess ||
      hipMalloc(&w.P, nn * 8) != hipSuccess || hipMalloc(&w.X, nn * 8) != hipSuccess ||
      hipMalloc(&w.W, (size_t)n * 8) != hipSuccess || hipMalloc(&w.E, (size_t)n * 8) != hipSuccess ||
      hipMalloc(&w.info, sizeof(rocblas_int)) != hipSuccess || hipMalloc(&w.red, 4 * sizeof(double)) != hipSuccess) {
    w.n = 0;
    return -1;
  }
  w.n = n;
  return 0;
}

void release(Work& w) {
  if (w.st) (void)hipStreamSynchronize(w.st);
  void* p[] = {w.A, w.B, w.P, w.X, w.W, w.E, w.red, w.info};
  for (void* q : p)
    if (q) (void)hipFree(q);
  if (w.h) (void)rocblas_destroy_handle(w.h);
  if (w.own_stream && w.st) (void)hipStreamDestroy(w.st);
  w = Work{};
}

const char* what(const Work& w, int rc) {
  static thread_local char buf[160];
  switch (rc) {
    case -3: std::snprintf(buf, sizeof(buf), "device copy or synchronisation failed: %s", hipGetErrorString(w.last_hip)); break;
    case -4: std::snprintf(buf, sizeof(buf), "rocBLAS/rocSOLVER call failed"); break;
    case -5: std::snprintf(buf, sizeof(buf), "factorisation info %d (matrix not positive definite)", w.last_info); break;
    default: std::snprintf(buf, sizeof(buf), "linear algebra error %d", rc); break;
  }
  return buf;
}

int set_metric(Work& w, int n, const double* dS) {
  const size_t nn = (size_t)n * n;
  w.metric_n = 0;
  if ((w.last_hip = hipMemcpyAsync(w.X, dS, nn * 8, hipMemcpyDeviceToDevice, w.st)) != hipSuccess) return -3;
  if (rocsolver_dpotrf(w.h, rocblas_fill_upper, n, w.X, n, w.info) != rocblas_status_success) return -4;
  if (int rc = read_info(w)) return rc;
  if (rocsolver_dtrtri(w.h, rocblas_fill_upper, rocblas_diagonal_non_unit, n, w.X, n, w.info) !=
      rocblas_status_success)
    return -4;
  if (int rc = read_info(w)) return rc;
  hipLaunchKernelGGL(k_zero_lower, dim3(grid_nn(nn)), dim3(256), 0, w.st, w.X, n);
  if (hipGetLastError() != hipSuccess) return -3;
  w.metric_n = n;
  return 0;
}

int eigen(Work& w, int n, const double* dF, double* de, double* dC) {
  if (w.metric_n != n) return -1;
  // A' = X^T F X, dsyevd, C = X C'
  if (int rc = gemm(w, rocblas_operation_none, rocblas_operation_none, n, n, dF, w.X, w.P)) return rc;
  if (int rc = gemm(w, rocblas_operation_transpose, rocblas_operation_none, n, n, w.X, w.P, w.A)) return rc;
  if (rocsolver_dsyevd(w.h, rocblas_evect_original, rocblas_fill_upper, n, w.A, n, de, w.E, w.info) !=
      rocblas_status_success)
    return -4;
  if (int rc = read_info(w)) return rc;
  return gemm(w, rocblas_operation_none, rocblas_operation_none, n, n, w.X, w.A, dC);
}

int density(Work& w, int n, int nOcc, const double* dC, double* dP) {
  if (nOcc == 0) return hipMemsetAsync(dP, 0, (size_t)n * n * 8, w.st) == hipSuccess ? 0 : -3;
  // column-major C: orbital k in column k, so P = C_occ C_occ^T
  return gemm(w, rocblas_operation_none, rocblas_operation_transpose, n, nOcc, dC, dC, dP);
}

int density_sp2(Work& w, int n, int nOcc, const double* dF, double* dP) {
  const size_t nn = (size_t)n * n;
  if (nOcc == 0) return hipMemsetAsync(dP, 0, nn * 8, w.st) == hipSuccess ? 0 : -3;
  if (nOcc == n) return 1;  // full occupation: P = S^-1, left to the eigenvector path
  if (w.metric_n != n) return -1;
  if (int rc = gemm(w, rocblas_operation_none, rocblas_operation_none, n, n, dF, w.X, w.P)) return rc;
  if (int rc = gemm(w, rocblas_operation_transpose, rocblas_operation_none, n, n, w.X, w.P, w.A)) return rc;
  hipLaunchKernelGGL(k_gershgorin, dim3(1), dim3(kRedBlock), 0, w.st, w.A, n, w.red);
  hipLaunchKernelGGL(k_sp2_init, dim3(grid_nn(nn)), dim3(256), 0, w.st, w.P, w.A, n, w.red);
  if (hipGetLastError() != hipSuccess) return -3;
  // P holds the iterate, B its square.  The idempotency error tr X - tr X^2
  // of the iterate entering a step is read every 8 steps; once below 1e-11
  // four more steps take it to rounding level (each step squares it).
  constexpr int kMaxSteps = 96, kCheck = 8;
  int extra = -1;
  double red[4] = {0, 0, 0, 0};
  auto square = [&]() -> int {
    if (int rc = gemm(w, rocblas_operation_none, rocblas_operation_none, n, n, w.P, w.P, w.B)) return rc;
    hipLaunchKernelGGL(k_sp2_trace, dim3(1), dim3(kRedBlock), 0, w.st, w.P, w.B, n, w.red);
    return hipGetLastError() == hipSuccess ? 0 : -3;
  };
  auto read_red = [&]() -> int {
    if (hipMemcpyAsync(red, w.red, sizeof(red), hipMemcpyDeviceToHost, w.st) != hipSuccess ||
        hipStreamSynchronize(w.st) != hipSuccess)
      return -3;
    return 0;
  };
  for (int it = 0; it < kMaxSteps; ++it) {
    if (int rc = square()) return rc;
    if (extra < 0 && (it + 1) % kCheck == 0) {
      if (int rc = read_red()) return rc;
      if (!(red[2] == red[2]) || !(red[3] == red[3])) return 1;  // NaN
      // a spectrum of rounding-level width has no meaningful lowest nOcc
      if (!(red[1] - red[0] > 1e-10 * std::max(1.0, std::fabs(red[1])))) return 1;
      if (std::fabs(red[2] - red[3]) < 1e-11) extra = 4;
    }
    hipLaunchKernelGGL(k_sp2_update, dim3(grid_nn(nn)), dim3(256), 0, w.st, w.P, w.B, nn, w.red, nOcc);
    if (hipGetLastError() != hipSuccess) return -3;
    if (extra > 0 && --extra == 0) break;
  }
  if (extra != 0) return 1;
  if (int rc = square()) return rc;
  if (int rc = read_red()) return rc;
  if (!(std::fabs(red[2] - red[3]) < 1e-11) || !(std::fabs(red[2] - nOcc) < 1e-8)) return 1;
  // P_AO = X P' X^T
  if (int rc = gemm(w, rocblas_operation_none, rocblas_operation_none, n, n, w.X, w.P, w.B)) return rc;
  return gemm(w, rocblas_operation_none, rocblas_operation_transpose, n, n, w.B, w.X, dP);
}

}  // namespace sqla

// ------------------------------------------------- host-array entry points
namespace {
std::mutex g_mu;
sqla::Work g_w;                   // shared by the host-array entry points
std::vector<double> g_metric;     // host copy of the metric factored in g_w

// device copies of host inputs, metric from a host S (cached while S repeats)
int host_setup(int n, const double* S) {
  if (sqla::ensure(g_w, n)) return -2;
  const size_t nn = (size_t)n * n;
  if (g_w.metric_n == n && g_metric.size() == nn && std::memcmp(g_metric.data(), S, nn * sizeof(double)) == 0)
    return 0;
  if (hipMemcpyAsync(g_w.B, S, nn * 8, hipMemcpyHostToDevice, g_w.st) != hipSuccess) return -3;
  if (int rc = sqla::set_metric(g_w, n, g_w.B)) return rc;
  g_metric.assign(S, S + nn);
  return 0;
}

int download(void* dst, const void* src, size_t bytes) {
  if (hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g_w.st) != hipSuccess ||
      hipStreamSynchronize(g_w.st) != hipSuccess)
    return -3;
  return 0;
}
}  // namespace

extern "C" {

int sqla_gen_eigen(int n, const double* F, const double* S, double* e, double* C) {
  if (n < 1) return -1;
  std::lock_guard<std::mutex> lk(g_mu);
  if (int rc = host_setup(n, S)) return rc;
  const size_t nn = (size_t)n * n;
  // F is staged in B; eigen() reads it once, before it writes C into B
  if (hipMemcpyAsync(g_w.B, F, nn * 8, hipMemcpyHostToDevice, g_w.st) != hipSuccess) return -3;
  int rc = sqla::eigen(g_w, n, g_w.B, g_w.W, g_w.B);
  if (!rc) rc = download(C, g_w.B, nn * 8);
  if (!rc) rc = download(e, g_w.W, (size_t)n * 8);
  return rc;
}

int sqla_density(int n, int nOcc, const double* C, double* P) {
  if (n < 1 || nOcc < 0 || nOcc > n) return -1;
  const size_t nn = (size_t)n * n;
  if (nOcc == 0) {
    std::memset(P, 0, nn * sizeof(double));
    return 0;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (sqla::ensure(g_w, n)) return -2;
  if (hipMemcpyAsync(g_w.A, C, (size_t)nOcc * n * 8, hipMemcpyHostToDevice, g_w.st) != hipSuccess) return -3;
  if (int rc = sqla::density(g_w, n, nOcc, g_w.A, g_w.B)) return rc;
  return download(P, g_w.B, nn * 8);
}

// Density P = sum_{k < nOcc} C_k C_k^T of the lowest nOcc solutions of
// F C = S C e, without the eigenvectors: second-order spectral projection
// (SP2, Niklasson, Phys. Rev. B 66, 155115 (2002)) of F' = X^T F X in the
// orthonormal basis, then P = X P' X^T.  The iteration maps the spectrum
// into [0, 1] by the Gershgorin bounds and squares towards the projector
// (X^2 or 2X - X^2, whichever brings the trace nearer nOcc), one DGEMM per
// step.  Returns 1 (P untouched) when the projector does not settle within
// the step limit: no gap at nOcc (degenerate frontier orbitals) or a very
// small one; the caller then uses the eigenvectors.
int sqla_density_sp2(int n, int nOcc, const double* F, const double* S, double* P) {
  if (n < 1 || nOcc < 0 || nOcc > n) return -1;
  const size_t nn = (size_t)n * n;
  if (nOcc == 0 || nOcc == n) {
    if (nOcc == 0) { std::memset(P, 0, nn * sizeof(double)); return 0; }
    return 1;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (int rc = host_setup(n, S)) return rc;
  // F is staged in B (read once, before B holds the squares); the result
  // lands in A (free after the first steps)
  if (hipMemcpyAsync(g_w.B, F, nn * 8, hipMemcpyHostToDevice, g_w.st) != hipSuccess) return -3;
  int rc = sqla::density_sp2(g_w, n, nOcc, g_w.B, g_w.A);
  if (rc == 0) rc = download(P, g_w.A, nn * 8);
  return rc;
}

int gen_sym_eigen(int nDim, const double* A, const double* S, double* e, double* C) {
  int rc = sqla_gen_eigen(nDim, A, S, e, C);
  if (rc) {
    std::printf("gen_sym_eigen: Error - device eigensolver failed (%d)\n", rc);
    std::exit(EXIT_FAILURE);
  }
  return 1;
}

int gen_sym_SomeEigen(int nDim, int nEigen, const double* A, const double* S, double* e, double* C) {
  if (nEigen == 0) return 1;
  return gen_sym_eigen(nDim, A, S, e, C);
}

// A X = B with A nRow x nRow and B nRow x nCol, both row-major (the
// reference passes its 5x5 DIIS system here).  Small: solved on the host
// with partial pivoting; A is not modified.
int linear_solver(int nRow, int nCol, const double* Ain, double* B) {
  const int n = nRow;
  if (nRow < 1 || nCol < 1) {
    std::printf("linear_solver: Error - invalid dimension range\n");
    std::exit(-1);
  }
  std::vector<double> Av(Ain, Ain + (size_t)n * n);
  double* A = Av.data();
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(A[r * n + c]) > std::fabs(A[p * n + c])) p = r;
    if (A[p * n + c] == 0.0) {
      std::printf("linear_solver: Error - singular matrix\n");
      std::exit(EXIT_FAILURE);
    }
    if (p != c) {
      for (int k = 0; k < n; ++k) std::swap(A[c * n + k], A[p * n + k]);
      for (int k = 0; k < nCol; ++k) std::swap(B[c * nCol + k], B[p * nCol + k]);
    }
    for (int r = c + 1; r < n; ++r) {
      const double f = A[r * n + c] / A[c * n + c];
      for (int k = c; k < n; ++k) A[r * n + k] -= f * A[c * n + k];
      for (int k = 0; k < nCol; ++k) B[r * nCol + k] -= f * B[c * nCol + k];
    }
  }
  for (int k = 0; k < nCol; ++k)
    for (int r = n - 1; r >= 0; --r) {
      double s = B[r * nCol + k];
      for (int j = r + 1; j < n; ++j) s -= A[r * n + j] * B[j * nCol + k];
      B[r * nCol + k] = s / A[r * n + r];
    }
  return 1;
}

}  // extern "C"
