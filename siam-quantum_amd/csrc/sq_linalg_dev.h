// sq_linalg_dev.h - dense FP64 linear algebra of the SCF on device-resident
// matrices (internal to the library).
//
// A Work holds the buffers, the rocBLAS handle and the stream of one user
// (the host drop-ins share one; every device-resident SCF owns its own, so
// that several SCFs can run side by side).  All matrices are n x n FP64 in
// device memory; symmetric inputs read the same in row- and column-major
// order.  Orbitals are orbital-major, C[k*n + i] (column k of a
// column-major matrix), eigenvalues ascending, as in the reference.
#pragma once
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

namespace sqla {

struct Work {
  int n = 0;                        // capacity
  hipStream_t st = nullptr;         // stream of every call (owned when own_stream)
  bool own_stream = false;
  rocblas_handle h = nullptr;
  double *A = nullptr, *B = nullptr, *P = nullptr, *X = nullptr;  // n x n
  double *W = nullptr, *E = nullptr;                              // n
  double* red = nullptr;            // Gershgorin bounds and traces (4)
  rocblas_int* info = nullptr;
  int metric_n = 0;                 // X = U^-1 holds the metric of this n
  int last_info = 0;                // LAPACK-style info of the last factorisation
  hipError_t last_hip = hipSuccess; // HIP error of the last failed call
};
// Text of the last failure of w (return code rc of the call).
const char* what(const Work& w, int rc);

// Buffers for size n on stream st (nullptr: a stream of its own).
int ensure(Work& w, int n, hipStream_t st = nullptr);
void release(Work& w);
// X = U^-1 with S = U^T U (upper Cholesky factor) from a device S.
int set_metric(Work& w, int n, const double* dS);
// F C = S C e for the metric set last: e (n), C (n x n, orbital-major).
int eigen(Work& w, int n, const double* dF, double* de, double* dC);
// P = sum_{k < nOcc} C_k C_k^T (one DGEMM).
int density(Work& w, int n, int nOcc, const double* dC, double* dP);
// Density of the lowest nOcc solutions of F C = S C e by second-order
// spectral projection; 0 on success, 1 when it does not settle (no gap at
// nOcc; dP untouched), < 0 on device errors.
int density_sp2(Work& w, int n, int nOcc, const double* dF, double* dP);

}  // namespace sqla
