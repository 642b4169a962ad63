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

// M[b][p][q] = V[(j0 + b) * ld + pair(p, q)], b < nb.  A workgroup (32 x 8
// lanes) takes one 32 x 32 tile p in [P0, P0+32), q in [Q0, Q0+32), Q0 <= P0,
// of one pair row: it reads the row's packed lower triangle (contiguous in q
// for each p), writes M[p][q] row by row and, through an LDS transpose, the
// mirrored tile M[q][p] row by row as well.
constexpr int kTile = 32;
__global__ void __launch_bounds__(256) k_unpack(const double* __restrict__ V, size_t ld, long long j0, int n,
                                                double* __restrict__ M) {
  __shared__ double t[kTile][kTile + 1];
  const int nt = (n + kTile - 1) / kTile;
  const int b = blockIdx.y;
  // tile index -> (ti >= tj) over the lower triangle of tiles
  int ti = (int)((sqrt(8.0 * blockIdx.x + 1.0) - 1.0) * 0.5);
  while ((ti + 1) * (ti + 2) / 2 <= (int)blockIdx.x) ++ti;
  while (ti * (ti + 1) / 2 > (int)blockIdx.x) --ti;
  const int tj = (int)blockIdx.x - ti * (ti + 1) / 2;
  if (ti >= nt) return;
  const int P0 = ti * kTile, Q0 = tj * kTile;
  const double* row = V + (size_t)(j0 + b) * ld;
  double* Mb = M + (size_t)b * n * n;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int i = ty; i < kTile; i += 8) {
    const int p = P0 + i, q = Q0 + tx;
    double v = 0.0;
    if (p < n && q < n && q <= p) {
      v = row[(size_t)p * (p + 1) / 2 + q];
      Mb[(size_t)p * n + q] = v;
    }
    t[i][tx] = v;
  }
  __syncthreads();
  // mirrored tile: M[q][p] for q in [Q0, Q0+32) (rows), p in [P0, P0+32)
  for (int i = ty; i < kTile; i += 8) {
    const int q = Q0 + i, p = P0 + tx;
    if (p < n && q < n && q < p) Mb[(size_t)q * n + p] = t[tx][i];
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

// part[block] += sum over one block of Y = (ar|bs) (columns: ket virtual
// orbitals s in [s0, s0 + ns) x ket occupied b; rows: bra virtual r x bra
// occupied a, from r = r0) of
//   MODE 0 (closed shell, rows r >= s0):  w (ar|bs)[2(ar|bs) - (as|br)] / D,
//          w = 2 for r > s, 1 for r = s (the (s,r) terms equal the (r,s) ones)
//   MODE 1 (one spin, rows r >= s0):      [(ar|bs) - (as|br)]^2 / (2 D), r > s
//   MODE 2 (two spins, all rows):         (ar|bs)^2 / D
// with D = e_a + e_b - e_r - e_s.
constexpr int kEnergyBlock = 256;
template <int MODE>
__global__ void __launch_bounds__(kEnergyBlock) k_energy(const double* __restrict__ Y, int nca, int nva, int ncb,
                                                        int s0, int ns, const double* __restrict__ eca,
                                                        const double* __restrict__ eva, const double* __restrict__ ecb,
                                                        const double* __restrict__ evb, double* __restrict__ part) {
  const int r0 = MODE == 2 ? 0 : s0;
  const size_t ld = (size_t)(nva - r0) * nca, total = ld * (size_t)ns * ncb;
  double acc = 0.0;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const size_t col = e / ld, row = e - col * ld;
    const int rl = (int)(row / nca), a = (int)(row % nca);
    const int sl = (int)(col / ncb), b = (int)(col % ncb);
    const double v = Y[e];
    const double d = eca[a] + ecb[b] - eva[r0 + rl] - evb[s0 + sl];
    if constexpr (MODE == 2) {
      acc += v * v / d;
    } else {
      if (rl < sl || (MODE == 1 && rl == sl)) continue;
      const double x = Y[((size_t)sl * nca + a) * ld + (size_t)rl * nca + b];
      if constexpr (MODE == 0)
        acc += (rl == sl ? 1.0 : 2.0) * v * (2.0 * v - x) / d;
      else
        acc += 0.5 * (v - x) * (v - x) / d;
    }
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

size_t sq_mp2_bytes(int n, size_t ncv_sum) {
  const size_t npr = (size_t)n * (n + 1) / 2;
  return npr * npr * 8 + npr * ncv_sum * 8;
}

namespace {

constexpr int kParts = 1024;  // energy partial sums (one per workgroup)

struct Ctx {
  hipStream_t st;
  rocblas_handle h;
  int n;
  size_t npr;
  const int2* lm;
  double* part;
  std::string* err;
  int bad(const std::string& m) {
    *err = m;
    return -1;
  }
};

// X[(a,r)][j] = (ar|ls) for every ket pair j = (l,s) of the pair matrix V:
// per block of pairs, unpack and two DGEMMs (Co^T M, then T Cv)
int first_half(Ctx& c, const double* dV, const SqMp2Spin& sp, double* dX) {
  const int n = c.n, nc = sp.nc, nv = sp.nv;
  const size_t npr = c.npr, ncv = (size_t)nc * nv;
  const double one = 1.0, zero = 0.0;
  size_t freeb = 0, totb = 0;
  if (hipMemGetInfo(&freeb, &totb) != hipSuccess) return c.bad("hipMemGetInfo failed");
  const size_t per_pair = (size_t)n * n * 8 + (size_t)nc * n * 8;
  const size_t blk_budget = std::min<size_t>(freeb / 4, (size_t)2 << 30);
  // (pairs per block: the unpack grid holds them in its y dimension)
  const int nb = (int)std::max<size_t>(1, std::min<size_t>({npr, blk_budget / per_pair, (size_t)INT32_MAX / n, (size_t)65535}));
  double *dM = nullptr, *dT = nullptr;
  Bufs blk;
  if (!blk.get(&dM, (size_t)nb * n * n) || !blk.get(&dT, (size_t)nc * nb * n))
    return c.bad("device allocation of the unpack blocks failed");
  for (size_t j0 = 0; j0 < npr; j0 += nb) {
    const int bc = (int)std::min<size_t>(nb, npr - j0);
    const int nt = (n + kTile - 1) / kTile;
    hipLaunchKernelGGL(k_unpack, dim3(nt * (nt + 1) / 2, bc), dim3(256), 0, c.st, dV, npr, (long long)j0, n, dM);
    // T = Co^T [M_0 ... M_bc-1]   (nc x bc n)
    if (rocblas_dgemm(c.h, rocblas_operation_transpose, rocblas_operation_none, nc, bc * n, n, &one, sp.dCo, n, dM, n,
                      &zero, dT, nc) != rocblas_status_success)
      return c.bad("rocblas_dgemm (first quarter) failed");
    // X_j = T_j Cv   (nc x nv each)
    if (rocblas_dgemm_strided_batched(c.h, rocblas_operation_none, rocblas_operation_none, nc, nv, n, &one, dT, nc,
                                      (rocblas_stride)nc * n, sp.dCv, n, 0, &zero, dX + j0 * ncv, nc,
                                      (rocblas_stride)ncv, bc) != rocblas_status_success)
      return c.bad("rocblas_dgemm_strided_batched (second quarter) failed");
  }
  if (hipStreamSynchronize(c.st) != hipSuccess)
    return c.bad(std::string("first half-transform failed: ") + hipGetErrorString(hipGetLastError()));
  return 0;
}

// Y = X_bra K_ket block by block over the ket virtual orbitals, summed into
// c.part by k_energy<MODE> (MODE 0/1: bra and ket the same spin, rows r >= s0
// only; MODE 2: all rows)
template <int MODE>
int second_half(Ctx& c, const double* dX, const SqMp2Spin& bra, const SqMp2Spin& ket, double* flop) {
  const size_t npr = c.npr, ncv = (size_t)bra.nc * bra.nv;
  const double one = 1.0, zero = 0.0;
  size_t freeb = 0, totb = 0;
  if (hipMemGetInfo(&freeb, &totb) != hipSuccess) return c.bad("hipMemGetInfo failed");
  const size_t per_s = (size_t)ket.nc * (npr + ncv) * 8;
  const size_t nblk_min = MODE == 2 ? 1 : 8;  // blocks restrict the rows only for one spin
  const int ns = (int)std::max<size_t>(
      1, std::min<size_t>({((size_t)ket.nv + nblk_min - 1) / nblk_min, (size_t)ket.nv, (freeb / 2) / per_s}));
  double *dK = nullptr, *dY = nullptr;
  Bufs blk;
  if (!blk.get(&dK, npr * (size_t)ns * ket.nc) || !blk.get(&dY, ncv * (size_t)ns * ket.nc))
    return c.bad("device allocation of the second-half blocks failed");
  for (int s0 = 0; s0 < ket.nv; s0 += ns) {
    const int sc = std::min(ns, ket.nv - s0);
    const int r0 = MODE == 2 ? 0 : s0;
    const size_t cols = (size_t)sc * ket.nc, rows = (size_t)(bra.nv - r0) * bra.nc;
    hipLaunchKernelGGL(k_kmat, dim3(grid_of(npr * cols, 256)), dim3(256), 0, c.st, ket.dCo, ket.dCv, c.lm, c.n, ket.nc,
                       s0, sc, npr, dK);
    // Y = X[rows (a, r >= r0)] K   (rows x cols)
    if (rocblas_dgemm(c.h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)rows, (rocblas_int)cols,
                      (rocblas_int)npr, &one, dX + (size_t)r0 * bra.nc, (rocblas_int)ncv, dK, (rocblas_int)npr, &zero,
                      dY, (rocblas_int)rows) != rocblas_status_success)
      return c.bad("rocblas_dgemm (second half) failed");
    *flop += 2.0 * (double)rows * cols * npr;
    hipLaunchKernelGGL(k_energy<MODE>, dim3(kParts), dim3(kEnergyBlock), 0, c.st, dY, bra.nc, bra.nv, ket.nc, s0, sc,
                       bra.dec, bra.dev, ket.dec, ket.dev, c.part);
  }
  if (hipStreamSynchronize(c.st) != hipSuccess)
    return c.bad(std::string("second half-transform failed: ") + hipGetErrorString(hipGetLastError()));
  return 0;
}

}  // namespace

int sq_mp2_from_pairs(hipStream_t st, int n, int nspin, const SqMp2Spin* sp, double** dV, SqMp2Result* out,
                      std::string* err) {
  *out = SqMp2Result{};
  const size_t npr = (size_t)n * (n + 1) / 2;
  auto fail = [&](const std::string& m) {
    *err = m;
    return -1;
  };
  bool any = false;
  for (int k = 0; k < nspin; ++k) any = any || (sp[k].nc > 0 && sp[k].nv > 0);
  if (!any) {  // nothing to correlate
    (void)hipFree(*dV);
    *dV = nullptr;
    return 0;
  }
  rocblas_handle h = nullptr;
  if (rocblas_create_handle(&h) != rocblas_status_success) return fail("rocblas_create_handle failed");
  struct HandleGuard {
    rocblas_handle h;
    ~HandleGuard() { (void)rocblas_destroy_handle(h); }
  } hg{h};
  if (rocblas_set_stream(h, st) != rocblas_status_success) return fail("rocblas_set_stream failed");
  hipEvent_t e0, e1, e2;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess || hipEventCreate(&e2) != hipSuccess)
    return fail("hipEventCreate failed");
  struct EventGuard {
    hipEvent_t a, b, c;
    ~EventGuard() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); (void)hipEventDestroy(c); }
  } eg{e0, e1, e2};

  Bufs keep;  // buffers of the whole call
  double* dX[2] = {nullptr, nullptr};
  double* dPart = nullptr;
  int2* dLM = nullptr;
  if (!keep.get(&dPart, kParts) || !keep.get(&dLM, npr)) return fail("device allocation failed");
  for (int k = 0; k < nspin; ++k)
    if (!keep.get(&dX[k], npr * (size_t)sp[k].nc * sp[k].nv))
      return fail("device allocation of the half-transformed integrals failed (" +
                  std::to_string(npr * (size_t)sp[k].nc * sp[k].nv * 8 / 1048576) + " MiB)");
  {
    std::vector<int2> lm(npr);
    for (int l = 0, j = 0; l < n; ++l)
      for (int m = 0; m <= l; ++m, ++j) lm[j] = make_int2(l, m);
    if (hipMemcpyAsync(dLM, lm.data(), npr * sizeof(int2), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return fail("pair table upload failed");
  }
  hipLaunchKernelGGL(k_zero, dim3(4), dim3(256), 0, st, dPart, (size_t)kParts);
  Ctx c{st, h, n, npr, dLM, dPart, err};

  if (hipEventRecord(e0, st) != hipSuccess) return fail("hipEventRecord failed");
  for (int k = 0; k < nspin; ++k)
    if (sp[k].nc > 0 && sp[k].nv > 0)
      if (int rc = first_half(c, *dV, sp[k], dX[k])) return rc;
  if (hipEventRecord(e1, st) != hipSuccess) return fail("hipEventRecord failed");
  (void)hipFree(*dV);  // V is no longer needed
  *dV = nullptr;

  double flop2 = 0.0;
  if (nspin == 1) {
    if (int rc = second_half<0>(c, dX[0], sp[0], sp[0], &flop2)) return rc;
  } else {
    for (int k = 0; k < 2; ++k)
      if (sp[k].nc > 1 && sp[k].nv > 1)
        if (int rc = second_half<1>(c, dX[k], sp[k], sp[k], &flop2)) return rc;
    if (sp[0].nc > 0 && sp[0].nv > 0 && sp[1].nc > 0 && sp[1].nv > 0)
      if (int rc = second_half<2>(c, dX[0], sp[0], sp[1], &flop2)) return rc;
  }
  if (hipEventRecord(e2, st) != hipSuccess) return fail("hipEventRecord failed");
  std::vector<double> part(kParts);
  if (hipMemcpyAsync(part.data(), dPart, kParts * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return fail(std::string("energy sum failed: ") + hipGetErrorString(hipGetLastError()));
  double e2sum = 0.0;
  for (double v : part) e2sum += v;
  float ms1 = 0.0f, ms2 = 0.0f;
  (void)hipEventElapsedTime(&ms1, e0, e1);
  (void)hipEventElapsedTime(&ms2, e1, e2);
  out->e2 = e2sum;
  out->half1_ms = ms1;
  out->half2_ms = ms2;
  double f1 = 0.0;
  for (int k = 0; k < nspin; ++k)
    f1 += 2.0 * sp[k].nc * (double)n * n * npr + 2.0 * sp[k].nc * (double)sp[k].nv * n * npr;
  out->half1_flop = f1;
  out->half2_flop = flop2;
  return 0;
}
