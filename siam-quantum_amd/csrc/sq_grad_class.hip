// sq_grad_class.hip - two-electron gradient kernel and launcher of one
// angular class, selected at compile time with -DSQ_LA=.. -DSQ_LB=..
// -DSQ_LC=.. -DSQ_LD=.. (the build compiles this file once per class).
#include <algorithm>
#include <cstdlib>
#include <string>

#include "sq_grad_classes.h"

#if !defined(SQ_LA) || !defined(SQ_LB) || !defined(SQ_LC) || !defined(SQ_LD)
#error "define SQ_LA SQ_LB SQ_LC SQ_LD"
#endif

#define SQ_CAT2(a, b, c, d, e) a##b##c##d##e
#define SQ_CAT(a, b, c, d, e) SQ_CAT2(a, b, c, d, e)
#define SQ_GRAD_LAUNCHER SQ_CAT(sqgrad_launch_, SQ_LA, SQ_LB, SQ_LC, SQ_LD)

namespace {
using namespace sqjk;
constexpr int LA = SQ_LA, LB = SQ_LB, LC = SQ_LC, LD = SQ_LD;
using SI = GradSlice<LA, LB, LC, LD>;

constexpr bool kWave = grad_wave<LA, LB, LC, LD>();
constexpr bool kBoth = grad_both<LA, LB, LC, LD>();

// workgroups the current device holds at once (occupancy x CUs)
template <class K>
long long resident_blocks(K kernel, int bs, size_t lds) {
  int dev = 0, ncu = 0, nb = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
    ncu = 1;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, bs, lds) != hipSuccess || nb < 1) nb = 1;
  return (long long)nb * ncu;
}
}  // namespace

// a.c.total = tiles.  Per-thread kernel: one wavefront per tile and slice, at
// most twice the workgroups the device holds at once over all slices.
// Wavefront kernel: one wavefront per (tile, ket position), at most four
// times the resident workgroups.  grid > 0: that many workgroups.
namespace {
// a template, so that only the selected kernel is instantiated
template <bool W>
int launch_grad(const GradArgs& a, int grid, hipStream_t st) {
  if constexpr (W) {
    static const long long res = resident_blocks(k_grad_wave<LA, LB, LC, LD>, 64 * kGWaves, 0);
    const long long need = (a.c.total * 64 + kGWaves - 1) / kGWaves;
    const long long g = grid > 0 ? grid : std::min(need, 4 * res);
    hipLaunchKernelGGL((k_grad_wave<LA, LB, LC, LD>), dim3((int)std::max(1LL, g)), dim3(64 * kGWaves), 0, st, a);
  } else {
    constexpr size_t lds = grad_lds_bytes<LA, LB, LC, LD>();
    constexpr int bs = grad_block<LA, LB, LC, LD>();
    static const long long res = resident_blocks(k_grad<LA, LB, LC, LD, bs>, bs, lds);
    const long long b = (a.c.total * 64 + bs - 1) / bs;
    const long long bmax = std::max(1LL, 2 * res / SI::NSLICE);
    const long long g = grid > 0 ? grid : std::min(b, bmax);
    hipLaunchKernelGGL((k_grad<LA, LB, LC, LD, bs>), dim3((int)std::max(1LL, g), SI::NSLICE), dim3(bs), lds, st, a);
  }
  return hipGetLastError() == hipSuccess ? 1 : -1;
}
}  // namespace

int SQ_GRAD_LAUNCHER(const sqjk::GradArgs& a, int grid, hipStream_t st) {
  if constexpr (kBoth) {
    static const bool thread = [] {
      const char* e = std::getenv("SQJK_GRAD_KERNEL");
      if (e && std::string(e) == "thread") return true;
      if (e && std::string(e) == "wave") return false;
      return grad_thread_default<LA, LB, LC, LD>();
    }();
    return thread ? launch_grad<false>(a, grid, st) : launch_grad<true>(a, grid, st);
  } else {
    return launch_grad<kWave>(a, grid, st);
  }
}
