// sq_mp2_class.hip - integral-store launcher of one angular class, selected
// at compile time with -DSQ_LA=.. -DSQ_LB=.. -DSQ_LC=.. -DSQ_LD=.. (the build
// compiles this file once per class).
#include <algorithm>

#include "sq_jk_wave.h"
#include "sq_mp2_classes.h"

#if !defined(SQ_LA) || !defined(SQ_LB) || !defined(SQ_LC) || !defined(SQ_LD)
#error "define SQ_LA SQ_LB SQ_LC SQ_LD"
#endif

#define SQ_CAT2(a, b, c, d, e) a##b##c##d##e
#define SQ_CAT(a, b, c, d, e) SQ_CAT2(a, b, c, d, e)
#define SQ_STORE_LAUNCHER SQ_CAT(sqmp2_store_, SQ_LA, SQ_LB, SQ_LC, SQ_LD)

namespace {
using namespace sqjk;
constexpr int LA = SQ_LA, LB = SQ_LB, LC = SQ_LC, LD = SQ_LD;
using SI = SliceInfo<LA, LB, LC, LD>;
// (dd|dd) evaluates on the one-wavefront-per-quartet kernel, as in the J/K
// build; every other class on the per-thread kernel
constexpr bool kWave = SI::NAB * SI::NCD > 648;
constexpr int kBS = class_block<LA, LB, LC, LD>();
constexpr size_t kLDS = class_lds_bytes<LA, LB, LC, LD>();

// Per-thread kernel: one wavefront per tile and slice, at most twice the
// workgroups the device holds at once over all slices.  Wave kernel: one
// 128-lane workgroup per (tile, ket position), at most 2^20 of them.
template <bool W>
int launch_store(const ClassArgs& a, hipStream_t st) {
  if constexpr (W) {
    const long long items = a.total * 64;
    const long long g = std::min(items, 1LL << 20);
    hipLaunchKernelGGL((k_class_wave<LA, LB, LC, LD, kStore, double>), dim3((int)std::max(1LL, g)), dim3(kWT), 0, st,
                       a);
  } else {
    static const long long res = [] {
      int dev = 0, ncu = 0, nb = 0;
      if (hipGetDevice(&dev) != hipSuccess ||
          hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
        ncu = 1;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_store<LA, LB, LC, LD, kBS>, kBS, kLDS) != hipSuccess ||
          nb < 1)
        nb = 1;
      return (long long)nb * ncu;
    }();
    const long long b = (a.total * 64 + kBS - 1) / kBS;
    const long long g = std::min(b, std::max(1LL, 2 * res / SI::NSLICE));
    hipLaunchKernelGGL((k_store<LA, LB, LC, LD, kBS>), dim3((int)std::max(1LL, g), SI::NSLICE), dim3(kBS), kLDS, st,
                       a);
  }
  return hipGetLastError() == hipSuccess ? 1 : -1;
}
}  // namespace

int SQ_STORE_LAUNCHER(const sqjk::ClassArgs& a, hipStream_t st) { return launch_store<kWave>(a, st); }
