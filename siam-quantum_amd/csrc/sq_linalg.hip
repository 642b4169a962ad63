// sq_linalg.hip - dense FP64 linear algebra of the SCF on the device.
//
// Generalised symmetric eigenproblem F C = S C e with rocSOLVER (dsygvd),
// behind the reference's src/lin.h surface:
//   gen_sym_eigen      <- src/lin.h:20 (src/lin.c:66-177, LAPACK dspgvd)
//   gen_sym_SomeEigen  <- src/lin.h:24 (src/lin.c:195-307, LAPACK dspgvx)
//   linear_solver      <- src/lin.h:30 (src/lin.c:320-391, LAPACK dgesv)
// Orbitals are returned orbital-major, C[k*n + i] = coefficient of basis
// function i in orbital k, eigenvalues ascending, as in the reference.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

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
  double *A = nullptr, *B = nullptr, *W = nullptr, *E = nullptr, *P = nullptr;
  rocblas_int* info = nullptr;
};
std::mutex g_mu;
EigWork g_w;

int ensure(int n) {
  if (!g_w.h && rocblas_create_handle(&g_w.h) != rocblas_status_success) return -1;
  if (g_w.n >= n) return 0;
  if (g_w.A) {
    (void)hipFree(g_w.A); (void)hipFree(g_w.B); (void)hipFree(g_w.W); (void)hipFree(g_w.E);
    (void)hipFree(g_w.P); (void)hipFree(g_w.info);
  }
  const size_t nn = (size_t)n * n;
  if (hipMalloc(&g_w.A, nn * 8) != hipSuccess || hipMalloc(&g_w.B, nn * 8) != hipSuccess ||
      hipMalloc(&g_w.P, nn * 8) != hipSuccess ||
      hipMalloc(&g_w.W, (size_t)n * 8) != hipSuccess || hipMalloc(&g_w.E, (size_t)n * 8) != hipSuccess ||
      hipMalloc(&g_w.info, sizeof(rocblas_int)) != hipSuccess) {
    g_w.n = 0;
    return -1;
  }
  g_w.n = n;
  return 0;
}

}  // namespace

extern "C" {

int sqla_gen_eigen(int n, const double* F, const double* S, double* e, double* C) {
  if (n < 1) return -1;
  std::lock_guard<std::mutex> lk(g_mu);
  if (ensure(n)) return -2;
  const size_t nn = (size_t)n * n;
  // symmetric input: row-major == column-major
  if (hipMemcpy(g_w.A, F, nn * 8, hipMemcpyHostToDevice) != hipSuccess) return -3;
  if (hipMemcpy(g_w.B, S, nn * 8, hipMemcpyHostToDevice) != hipSuccess) return -3;
  rocblas_status st = rocsolver_dsygvd(g_w.h, rocblas_eform_ax, rocblas_evect_original, rocblas_fill_upper, n,
                                       g_w.A, n, g_w.B, n, g_w.W, g_w.E, g_w.info);
  if (st != rocblas_status_success) return -4;
  rocblas_int info = 0;
  if (hipMemcpy(&info, g_w.info, sizeof(info), hipMemcpyDeviceToHost) != hipSuccess) return -3;
  if (info != 0) return -5;
  // column k of A (column-major) is orbital k: A[k*n + i]
  if (hipMemcpy(C, g_w.A, nn * 8, hipMemcpyDeviceToHost) != hipSuccess) return -3;
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
