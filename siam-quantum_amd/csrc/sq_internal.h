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

// Closed-shell MP2 energy from the dense AO pair matrix *dV (sq_mp2.hip);
// frees *dV after the first half-transform and sets it to nullptr.  Returns
// 0, or -1 with the reason in *err.
#include <string>
struct SqMp2Result {
  double e2;                     // MP2 correlation energy
  double half1_ms, half2_ms;     // first / second half-transform (with the energy sum)
  double half1_flop, half2_flop; // DGEMM flops of the two halves
};
int sq_mp2_from_pairs(hipStream_t st, int n, int nc, int nv, const double* dCo, const double* dCv, const double* dec,
                      const double* dev, double** dV, SqMp2Result* out, std::string* err);
// device bytes of the pair matrix plus the half-transformed integrals
size_t sq_mp2_bytes(int n, int nc, int nv);
