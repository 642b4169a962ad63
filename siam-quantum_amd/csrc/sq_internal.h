// sq_internal.h - library-internal accessors between the engine and the
// device-resident SCF (not part of the C-ABI).
#pragma once
#include <hip/hip_runtime.h>

struct sqjk_ctx;

// stream on which a J/K context queues its work
hipStream_t sqjk_stream_internal(sqjk_ctx* c);
// mixed-precision threshold set for the drop-in entry points
double sqjk_dropin_mixed_internal();
// device blocks from the library's cache (no hipFree per release, which
// would wait for the whole device)
void* sq_device_block(size_t bytes);
void sq_device_release(void* p);

// MP2 energy from the dense AO pair matrix *dV (sq_mp2.hip): one spin block
// (closed shell) or two (alpha, beta: same-spin and opposite-spin terms).
// Frees *dV after the first half-transform and sets it to nullptr.  Returns
// 0, or -1 with the reason in *err.
#include <string>
struct SqMp2Spin {
  int nc, nv;                     // correlated occupied, virtual orbitals
  const double *dCo, *dCv;        // n x nc, n x nv (column-major, ld n)
  const double *dec, *dev;        // their orbital energies
};
struct SqMp2Result {
  double e2;                     // MP2 correlation energy
  double half1_ms, half2_ms;     // first / second half-transform (with the energy sum)
  double half1_flop, half2_flop; // DGEMM flops of the two halves
};
int sq_mp2_from_pairs(hipStream_t st, int n, int nspin, const SqMp2Spin* sp, double** dV, SqMp2Result* out,
                      std::string* err);
// device bytes of the pair matrix plus the half-transformed integrals
// (ncv_sum = sum over the spin blocks of nc nv)
size_t sq_mp2_bytes(int n, size_t ncv_sum);
