// sq_linalg.hip - dense FP64 linear algebra of the SCF on the device.
//
// Generalised symmetric eigenproblem F C = S C e on the device, behind the
// reference's src/lin.h surface.  S is fixed during an SCF, so its Cholesky
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

namespace {

struct EigWork {
  int n = 0;
  rocblas_handle h = nullptr;
  double *A = nullptr, *B = nullptr, *W = nullptr, *E = nullptr, *P = nullptr, *X = nullptr;
  double* red = nullptr;  // Gershgorin bounds and traces of the purification
  rocblas_int* info = nullptr;
  int metric_n = 0;             // X = U^-1 is valid for this n and the host copy below
  std::vector<double> metric;   // S of the cached factor
};
std::mutex g_mu;
EigWork g_w;

int ensure(int n) {
  if (!g_w.h && rocblas_create_handle(&g_w.h) != rocblas_status_success) return -1;
  if (g_w.n >= n) return 0;
  if (g_w.A) {
    (void)hipFree(g_w.A); (void)hipFree(g_w.B); (void)hipFree(g_w.W); (void)hipFree(g_w.E);
    (void)hipFree(g_w.P); (void)hipFree(g_w.X); (void)hipFree(g_w.info); (void)hipFree(g_w.red);
  }
  g_w.metric_n = 0;
  const size_t nn = (size_t)n * n;
  if (hipMalloc(&g_w.A, nn * 8) != hipSuccess || hipMalloc(&g_w.B, nn * 8) != hipSuccess ||
      hipMalloc(&g_w.P, nn * 8) != hipSuccess || hipMalloc(&g_w.X, nn * 8) != hipSuccess ||
      hipMalloc(&g_w.W, (size_t)n * 8) != hipSuccess || hipMalloc(&g_w.E, (size_t)n * 8) != hipSuccess ||
      hipMalloc(&g_w.info, sizeof(rocblas_int)) != hipSuccess || hipMalloc(&g_w.red, 4 * sizeof(double)) != hipSuccess) {
    g_w.n = 0;
    return -1;
  }
  g_w.n = n;
  return 0;
}

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

int read_info() {
  rocblas_int info = 0;
  if (hipMemcpy(&info, g_w.info, sizeof(info), hipMemcpyDeviceToHost) != hipSuccess) return -3;
  return info != 0 ? -5 : 0;
}

// X = U^-1 with S = U^T U (upper Cholesky factor), unless S is the cached metric
int set_metric(int n, const double* S) {
  const size_t nn = (size_t)n * n;
  if (g_w.metric_n == n && std::memcmp(g_w.metric.data(), S, nn * sizeof(double)) == 0) return 0;
  g_w.metric_n = 0;
  if (hipMemcpy(g_w.X, S, nn * 8, hipMemcpyHostToDevice) != hipSuccess) return -3;
  if (rocsolver_dpotrf(g_w.h, rocblas_fill_upper, n, g_w.X, n, g_w.info) != rocblas_status_success) return -4;
  if (int rc = read_info()) return rc;
  if (rocsolver_dtrtri(g_w.h, rocblas_fill_upper, rocblas_diagonal_non_unit, n, g_w.X, n, g_w.info) !=
      rocblas_status_success)
    return -4;
  if (int rc = read_info()) return rc;
  hipLaunchKernelGGL(k_zero_lower, dim3((unsigned)std::min<size_t>((nn + 255) / 256, 4096)), dim3(256), 0, 0, g_w.X, n);
  if (hipGetLastError() != hipSuccess) return -3;
  g_w.metric.assign(S, S + nn);
  g_w.metric_n = n;
  return 0;
}

}  // namespace

extern "C" {

int sqla_gen_eigen(int n, const double* F, const double* S, double* e, double* C) {
  if (n < 1) return -1;
  std::lock_guard<std::mutex> lk(g_mu);
  if (ensure(n)) return -2;
  if (int rc = set_metric(n, S)) return rc;
  const size_t nn = (size_t)n * n;
  const double one = 1.0, zero = 0.0;
  // symmetric input: row-major == column-major.  A' = X^T F X
  if (hipMemcpy(g_w.A, F, nn * 8, hipMemcpyHostToDevice) != hipSuccess) return -3;
  if (rocblas_dgemm(g_w.h, rocblas_operation_none, rocblas_operation_none, n, n, n, &one, g_w.A, n, g_w.X, n,
                    &zero, g_w.P, n) != rocblas_status_success ||
      rocblas_dgemm(g_w.h, rocblas_operation_transpose, rocblas_operation_none, n, n, n, &one, g_w.X, n, g_w.P, n,
                    &zero, g_w.A, n) != rocblas_status_success)
    return -4;
  if (rocsolver_dsyevd(g_w.h, rocblas_evect_original, rocblas_fill_upper, n, g_w.A, n, g_w.W, g_w.E, g_w.info) !=
      rocblas_status_success)
    return -4;
  if (int rc = read_info()) return rc;
  // C = X C'; column k (column-major) is orbital k: C[k*n + i]
  if (rocblas_dgemm(g_w.h, rocblas_operation_none, rocblas_operation_none, n, n, n, &one, g_w.X, n, g_w.A, n,
                    &zero, g_w.P, n) != rocblas_status_success)
    return -4;
  if (hipMemcpy(C, g_w.P, nn * 8, hipMemcpyDeviceToHost) != hipSuccess) return -3;
  if (hipMemcpy(e, g_w.W, (size_t)n * 8, hipMemcpyDeviceToHost) != hipSuccess) return -3;
  return 0;
}

// P = sum_{k < nOcc} C_k C_k^T with C orbital-major (C[k*n + i]); one DGEMM
// on the device (matrix cores).  Read as column-major, C is n x n with
// orbital k in column k, so P = C_occ C_occ^T.
int sqla_density(int n, int nOcc, const double* C, double* P) {
  if (n < 1 || nOcc < 0 || nOcc > n) return -1;
  const size_t nn = (size_t)n * n;
  if (nOcc == 0) {
    std::memset(P, 0, nn * sizeof(double));
    return 0;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (ensure(n)) return -2;
  if (hipMemcpy(g_w.A, C, (size_t)nOcc * n * 8, hipMemcpyHostToDevice) != hipSuccess) return -3;
  const double one = 1.0, zero = 0.0;
  if (rocblas_dgemm(g_w.h, rocblas_operation_none, rocblas_operation_transpose, n, n, nOcc, &one, g_w.A, n,
                    g_w.A, n, &zero, g_w.P, n) != rocblas_status_success)
    return -4;
  if (hipMemcpy(P, g_w.P, nn * 8, hipMemcpyDeviceToHost) != hipSuccess) return -3;
  return 0;
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
    return 1;  // full occupation: P = S^-1, left to the eigenvector path
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (ensure(n)) return -2;
  if (int rc = set_metric(n, S)) return rc;
  const double one = 1.0, zero = 0.0;
  if (hipMemcpy(g_w.A, F, nn * 8, hipMemcpyHostToDevice) != hipSuccess) return -3;
  if (rocblas_dgemm(g_w.h, rocblas_operation_none, rocblas_operation_none, n, n, n, &one, g_w.A, n, g_w.X, n,
                    &zero, g_w.P, n) != rocblas_status_success ||
      rocblas_dgemm(g_w.h, rocblas_operation_transpose, rocblas_operation_none, n, n, n, &one, g_w.X, n, g_w.P, n,
                    &zero, g_w.A, n) != rocblas_status_success)
    return -4;
  const unsigned grid = (unsigned)std::min<size_t>((nn + 255) / 256, 4096);
  hipLaunchKernelGGL(k_gershgorin, dim3(1), dim3(kRedBlock), 0, 0, g_w.A, n, g_w.red);
  hipLaunchKernelGGL(k_sp2_init, dim3(grid), dim3(256), 0, 0, g_w.P, g_w.A, n, g_w.red);
  if (hipGetLastError() != hipSuccess) return -3;
  // P holds the iterate, B its square.  The idempotency error tr X - tr X^2
  // of the iterate entering a step is read every 8 steps; once below 1e-11
  // four more steps take it to rounding level (each step squares it).
  constexpr int kMaxSteps = 96, kCheck = 8;
  int extra = -1;
  double red[4] = {0, 0, 0, 0};
  auto square = [&]() -> int {
    if (rocblas_dgemm(g_w.h, rocblas_operation_none, rocblas_operation_none, n, n, n, &one, g_w.P, n, g_w.P, n,
                      &zero, g_w.B, n) != rocblas_status_success)
      return -4;
    hipLaunchKernelGGL(k_sp2_trace, dim3(1), dim3(kRedBlock), 0, 0, g_w.P, g_w.B, n, g_w.red);
    return hipGetLastError() == hipSuccess ? 0 : -3;
  };
  for (int it = 0; it < kMaxSteps; ++it) {
    if (int rc = square()) return rc;
    if (extra < 0 && (it + 1) % kCheck == 0) {
      if (hipMemcpy(red, g_w.red, sizeof(red), hipMemcpyDeviceToHost) != hipSuccess) return -3;
      if (!(red[2] == red[2]) || !(red[3] == red[3])) return 1;  // NaN
      // a spectrum of rounding-level width has no meaningful lowest nOcc
      if (!(red[1] - red[0] > 1e-10 * std::max(1.0, std::fabs(red[1])))) return 1;
      if (std::fabs(red[2] - red[3]) < 1e-11) extra = 4;
    }
    hipLaunchKernelGGL(k_sp2_update, dim3(grid), dim3(256), 0, 0, g_w.P, g_w.B, nn, g_w.red, nOcc);
    if (hipGetLastError() != hipSuccess) return -3;
    if (extra > 0 && --extra == 0) break;
  }
  if (extra != 0) return 1;
  if (int rc = square()) return rc;
  if (hipMemcpy(red, g_w.red, sizeof(red), hipMemcpyDeviceToHost) != hipSuccess) return -3;
  if (!(std::fabs(red[2] - red[3]) < 1e-11) || !(std::fabs(red[2] - nOcc) < 1e-8)) return 1;
  // P_AO = X P' X^T
  if (rocblas_dgemm(g_w.h, rocblas_operation_none, rocblas_operation_none, n, n, n, &one, g_w.X, n, g_w.P, n,
                    &zero, g_w.B, n) != rocblas_status_success ||
      rocblas_dgemm(g_w.h, rocblas_operation_none, rocblas_operation_transpose, n, n, n, &one, g_w.B, n, g_w.X, n,
                    &zero, g_w.A, n) != rocblas_status_success)
    return -4;
  if (hipMemcpy(P, g_w.A, nn * 8, hipMemcpyDeviceToHost) != hipSuccess) return -3;
  return 0;
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
