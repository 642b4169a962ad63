// sq_mp2.hip - closed-shell MP2 energy from the dense AO pair matrix.
//
// Input: V[P][Q] = (pq|rs) for the packed pairs P = pair(p,q), Q = pair(r,s)
// (n(n+1)/2 on a side, written by the integral-store kernels), the
// correlated occupied orbitals Co (n x nc) and the virtual orbitals Cv
// (n x nv), column-major with leading dimension n (the reference's
// orbital-major C[k*n + i]), and their orbital energies.
//
// First half (per block of ket pairs j = pair(l,s)): unpack V[j][.] into the
// symmetric n x n matrix M_j, then X_j = Co^T M_j Cv (nc x nv) with two
// DGEMMs, so X[(a,r)][j] = (ar|ls).  Second half (per block of virtual
// orbitals s): K[j][(b,s)] = Co[l,b] Cv[s',s] + Co[s',b] Cv[l,s] for
// j = pair(l,s') (one term when l = s'), and Y = X K, Y[(a,r)][(b,s)] =
// (ar|bs), one DGEMM on the matrix cores.  The energy
//   E2 = sum_{a,b,r,s} (ar|bs) [2 (ar|bs) - (as|br)] / (e_a + e_b - e_r - e_s)
// has the same (r,s) and (s,r) terms (swap a<->b, r<->s), so a block of
// virtual orbitals s in [s0, s1) needs only the rows r >= s0 of Y: it sums
// the r > s terms twice and the r = s terms once, reading (as|br) =
// Y[(b,r)][(a,s)] from the same block.  With eight or more blocks this
// computes a little over half of Y.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "sq_internal.h"

namespace {

__device__ __forceinline__ size_t pair_at(int p, int q) {
  return p >= q ? (size_t)p * (p + 1) / 2 + q : (size_t)q * (q + 1) / 2 + p;
}

// M[b][p][q] = V[(j0 + b) * ld + pair(p, q)], b < nb
__global__ void k_unpack(const double* __restrict__ V, size_t ld, long long j0, int nb, int n, double* __restrict__ M) {
  const size_t nn = (size_t)n * n, total = nn * nb;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const size_t b = e / nn, r = e - b * nn;
    const int p = (int)(r / n), q = (int)(r - (size_t)p * n);
    M[e] = V[(j0 + b) * ld + pair_at(p, q)];
  }
}

// K[col][j], col = (s - s0) nc + b, for the pairs j = (l, m) in lm[j]
__global__ void k_kmat(const double* __restrict__ Co, const double* __restrict__ Cv, const int2* __restrict__ lm,
                       int n, int nc, int s0, int ns, size_t npr, double* __restrict__ K) {
  const size_t total = npr * (size_t)ns * nc;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const size_t col = e / npr, j = e - col * npr;
    const int s = s0 + (int)(col / nc), b = (int)(col % nc);
    const int2 t = lm[j];
    const double* cb = Co + (size_t)b * n;
    const double* cs = Cv + (size_t)s * n;
    K[e] = t.x == t.y ? cb[t.x] * cs[t.x] : cb[t.x] * cs[t.y] + cb[t.y] * cs[t.x];
  }
}

// part[block] += sum over the block of Y (rows r >= s0, columns s in
// [s0, s0 + ns), leading dimension ld = (nv - s0) nc) of
// w_rs (ar|bs)[2(ar|bs) - (as|br)] / D, w = 2 for r > s, 1 for r = s, 0 below
constexpr int kEnergyBlock = 256;
__global__ void __launch_bounds__(kEnergyBlock) k_energy(const double* __restrict__ Y, int nc, int nv, int s0, int ns,
                                                        const double* __restrict__ ec, const double* __restrict__ ev,
                                                        double* __restrict__ part) {
  const size_t ld = (size_t)(nv - s0) * nc, total = ld * (size_t)ns * nc;
  double acc = 0.0;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const size_t col = e / ld, row = e - col * ld;
    const int rl = (int)(row / nc), a = (int)(row % nc);
    const int sl = (int)(col / nc), b = (int)(col % nc);
    if (rl < sl) continue;
    const double v = Y[e];
    const double x = Y[((size_t)sl * nc + a) * ld + (size_t)rl * nc + b];
    const double w = rl == sl ? 1.0 : 2.0;
    acc += w * v * (2.0 * v - x) / (ec[a] + ec[b] - ev[s0 + rl] - ev[s0 + sl]);
  }
  __shared__ double red[kEnergyBlock];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kEnergyBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] += red[0];
}

__global__ void k_zero(double* p, size_t n) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) p[e] = 0.0;
}

int grid_of(size_t work, int bs) { return (int)std::max<size_t>(1, std::min<size_t>((work + bs - 1) / bs, 1 << 16)); }

struct Bufs {
  std::vector<void*> p;
  ~Bufs() {
    for (void* q : p)
      if (q) (void)hipFree(q);
  }
  template <class T>
  bool get(T** d, size_t count) {
    void* v = nullptr;
    if (hipMalloc(&v, std::max<size_t>(count * sizeof(T), 8)) != hipSuccess) return false;
    p.push_back(v);
    *d = (T*)v;
    return true;
  }
};

}  // namespace

size_t sq_mp2_bytes(int n, int nc, int nv) {
  const size_t npr = (size_t)n * (n + 1) / 2, ncv = (size_t)nc * nv;
  return npr * npr * 8 + npr * ncv * 8;
}

int sq_mp2_from_pairs(hipStream_t st, int n, int nc, int nv, const double* dCo, const double* dCv, const double* dec,
                      const double* dev, double** dV, SqMp2Result* out, std::string* err) {
  auto bad = [&](const std::string& m) {
    *err = m;
    return -1;
  };
  *out = SqMp2Result{};
  const size_t npr = (size_t)n * (n + 1) / 2, ncv = (size_t)nc * nv;
  if (nc <= 0 || nv <= 0) {  // nothing to correlate
    (void)hipFree(*dV);
    *dV = nullptr;
    return 0;
  }
  rocblas_handle h = nullptr;
  if (rocblas_create_handle(&h) != rocblas_status_success) return bad("rocblas_create_handle failed");
  struct HandleGuard {
    rocblas_handle h;
    ~HandleGuard() { (void)rocblas_destroy_handle(h); }
  } hg{h};
  if (rocblas_set_stream(h, st) != rocblas_status_success) return bad("rocblas_set_stream failed");

  hipEvent_t e0, e1, e2;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess || hipEventCreate(&e2) != hipSuccess)
    return bad("hipEventCreate failed");
  struct EventGuard {
    hipEvent_t a, b, c;
    ~EventGuard() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); (void)hipEventDestroy(c); }
  } eg{e0, e1, e2};

  Bufs keep;  // buffers of the whole call
  double* dX = nullptr;
  double* dPart = nullptr;
  int2* dLM = nullptr;
  const int G = 1024;  // energy partial sums (one per workgroup)
  if (!keep.get(&dX, npr * ncv) || !keep.get(&dPart, G) || !keep.get(&dLM, npr))
    return bad("device allocation of the half-transformed integrals failed (" +
               std::to_string(npr * ncv * 8 / 1048576) + " MiB)");
  {
    std::vector<int2> lm(npr);
    for (int l = 0, j = 0; l < n; ++l)
      for (int m = 0; m <= l; ++m, ++j) lm[j] = make_int2(l, m);
    if (hipMemcpyAsync(dLM, lm.data(), npr * sizeof(int2), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return bad("pair table upload failed");
  }
  hipLaunchKernelGGL(k_zero, dim3(4), dim3(256), 0, st, dPart, (size_t)G);
  const double one = 1.0, zero = 0.0;

  // first half, blocks of ket pairs: M (nb n^2) and T (nc nb n)
  size_t freeb = 0, totb = 0;
  if (hipMemGetInfo(&freeb, &totb) != hipSuccess) return bad("hipMemGetInfo failed");
  const size_t per_pair = (size_t)n * n * 8 + (size_t)nc * n * 8;
  const size_t blk_budget = std::min<size_t>(freeb / 4, (size_t)2 << 30);
  const int nb = (int)std::max<size_t>(1, std::min<size_t>({npr, blk_budget / per_pair, (size_t)INT32_MAX / n}));
  double *dM = nullptr, *dT = nullptr;
  {
    Bufs blk;
    if (!blk.get(&dM, (size_t)nb * n * n) || !blk.get(&dT, (size_t)nc * nb * n))
      return bad("device allocation of the unpack blocks failed");
    if (hipEventRecord(e0, st) != hipSuccess) return bad("hipEventRecord failed");
    for (size_t j0 = 0; j0 < npr; j0 += nb) {
      const int bc = (int)std::min<size_t>(nb, npr - j0);
      hipLaunchKernelGGL(k_unpack, dim3(grid_of((size_t)bc * n * n, 256)), dim3(256), 0, st, *dV, npr, (long long)j0,
                         bc, n, dM);
      // T = Co^T [M_0 ... M_bc-1]   (nc x bc n)
      if (rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, nc, bc * n, n, &one, dCo, n, dM, n,
                        &zero, dT, nc) != rocblas_status_success)
        return bad("rocblas_dgemm (first quarter) failed");
      // X_j = T_j Cv   (nc x nv each)
      if (rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, nc, nv, n, &one, dT, nc,
                                        (rocblas_stride)nc * n, dCv, n, 0, &zero, dX + j0 * ncv, nc,
                                        (rocblas_stride)ncv, bc) != rocblas_status_success)
        return bad("rocblas_dgemm_strided_batched (second quarter) failed");
    }
    if (hipEventRecord(e1, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
      return bad(std::string("first half-transform failed: ") + hipGetErrorString(hipGetLastError()));
  }
  // V is no longer needed
  (void)hipFree(*dV);
  *dV = nullptr;

  // second half, blocks of virtual orbitals s: K (npr x ns nc), Y (rows r >=
  // s0: at most ncv x ns nc); at least eight blocks, more if memory asks
  if (hipMemGetInfo(&freeb, &totb) != hipSuccess) return bad("hipMemGetInfo failed");
  const size_t per_s = (size_t)nc * (npr + ncv) * 8;
  const int ns = (int)std::max<size_t>(1, std::min<size_t>({(size_t)(nv + 7) / 8, (size_t)nv, (freeb / 2) / per_s}));
  double *dK = nullptr, *dY = nullptr;
  Bufs blk;
  if (!blk.get(&dK, npr * (size_t)ns * nc) || !blk.get(&dY, ncv * (size_t)ns * nc))
    return bad("device allocation of the second-half blocks failed");
  double flop2 = 0.0;
  for (int s0 = 0; s0 < nv; s0 += ns) {
    const int sc = std::min(ns, nv - s0);
    const size_t cols = (size_t)sc * nc, rows = (size_t)(nv - s0) * nc;
    hipLaunchKernelGGL(k_kmat, dim3(grid_of(npr * cols, 256)), dim3(256), 0, st, dCo, dCv, dLM, n, nc, s0, sc, npr, dK);
    // Y = X[rows (a, r >= s0)] K   (rows x cols)
    if (rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)rows, (rocblas_int)cols,
                      (rocblas_int)npr, &one, dX + (size_t)s0 * nc, (rocblas_int)ncv, dK, (rocblas_int)npr, &zero, dY,
                      (rocblas_int)rows) != rocblas_status_success)
      return bad("rocblas_dgemm (second half) failed");
    flop2 += 2.0 * (double)rows * cols * npr;
    hipLaunchKernelGGL(k_energy, dim3(G), dim3(kEnergyBlock), 0, st, dY, nc, nv, s0, sc, dec, dev, dPart);
  }
  if (hipEventRecord(e2, st) != hipSuccess) return bad("hipEventRecord failed");
  std::vector<double> part(G);
  if (hipMemcpyAsync(part.data(), dPart, G * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return bad(std::string("second half-transform failed: ") + hipGetErrorString(hipGetLastError()));
  double e2sum = 0.0;
  for (double v : part) e2sum += v;
  float ms1 = 0.0f, ms2 = 0.0f;
  (void)hipEventElapsedTime(&ms1, e0, e1);
  (void)hipEventElapsedTime(&ms2, e1, e2);
  out->e2 = e2sum;
  out->half1_ms = ms1;
  out->half2_ms = ms2;
  out->half1_flop = 2.0 * nc * (double)n * n * npr + 2.0 * nc * (double)nv * n * npr;
  out->half2_flop = flop2;
  return 0;
}
