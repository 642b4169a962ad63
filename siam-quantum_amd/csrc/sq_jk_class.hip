// sq_jk_class.hip - kernels and launcher of one angular class, selected at
// compile time with -DSQ_LA=.. -DSQ_LB=.. -DSQ_LC=.. -DSQ_LD=.. (the build
// compiles this file once per class, in parallel).
#include <algorithm>

#include "sq_jk_classes.h"
#include "sq_jk_wave.h"
#include "sq_jk_wd.h"

#if !defined(SQ_LA) || !defined(SQ_LB) || !defined(SQ_LC) || !defined(SQ_LD)
#error "define SQ_LA SQ_LB SQ_LC SQ_LD"
#endif

#define SQ_CAT2(a, b, c, d, e) a##b##c##d##e
#define SQ_CAT(a, b, c, d, e) SQ_CAT2(a, b, c, d, e)
#define SQ_LAUNCHER SQ_CAT(sqjk_launch_, SQ_LA, SQ_LB, SQ_LC, SQ_LD)

// Everything up to SQ_LAUNCHER stays in the unnamed namespace: each class
// translation unit defines these templates with a different body, so external
// (weak) linkage would let the linker keep one class's launcher for all.
namespace {
using namespace sqjk;
constexpr int LA = SQ_LA, LB = SQ_LB, LC = SQ_LC, LD = SQ_LD;
using SI = SliceInfo<LA, LB, LC, LD>;

// Classes up to 648 integrals per quartet use the per-thread kernel (one
// launch per ket-component slice); (dd|dd) the cooperative
// one-wavefront-per-quartet kernel.
constexpr bool kWave = SI::NAB * SI::NCD > 648;
// FP32 evaluation is compiled for classes of at most 81 integrals; larger
// classes evaluate their FP32 range in FP64 (FP32 and FP64 vector rates are
// equal on this target without packed FP32, so this costs no time).
constexpr bool kFP32 = SI::NAB * SI::NCD <= 81;

// all slices in one launch: grid (gx, NSLICE)
template <int MODE, typename R>
int launch_thread(const ClassArgs& a, int gx, hipStream_t st) {
  if constexpr (kWave) {
    return 0;
  } else {
    constexpr size_t lds = class_lds_bytes<LA, LB, LC, LD>();
    constexpr int bs = class_block<LA, LB, LC, LD>();
    hipLaunchKernelGGL((k_class<LA, LB, LC, LD, MODE, R, bs>), dim3(gx, SI::NSLICE), dim3(bs), lds, st, a);
    return hipGetLastError() == hipSuccess ? 1 : -1;
  }
}

template <int MODE, typename R>
int launch_wave(const ClassArgs& a, int grid, hipStream_t st) {
  hipLaunchKernelGGL((k_class_wave<LA, LB, LC, LD, MODE, R>), dim3(grid), dim3(kWT), 0, st, a);
  return hipGetLastError() == hipSuccess ? 1 : -1;
}

// Wavefront-per-quartet kernel with precomputed Hermite tables (RHF/UHF
// builds of one molecule): grid = at most `over` times the workgroups the
// device holds at once, each of its four wavefronts striding over the
// (tile, ket position) items.
template <int MODE>
int launch_wd(const ClassArgs& a, int grid, hipStream_t st) {
  if constexpr (MODE == kSchwarz || !kWave) {
    return -1;
  } else {
    static const long long resident = [] {
      int dev = 0, ncu = 0, nb = 0;
      if (hipGetDevice(&dev) != hipSuccess ||
          hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
        ncu = 1;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_class_wd<LA, LB, LC, LD, MODE>, 64 * kWDWaves, 0) !=
              hipSuccess || nb < 1)
        nb = 1;
      return (long long)nb * ncu;
    }();
    const long long items = a.total * 64;
    const long long need = (items + kWDWaves - 1) / kWDWaves;
    const long long over = grid < 0 ? -grid : 4;
    const long long cap = std::max(1LL, over * resident);
    const long long g = grid > 0 ? grid : (need < cap ? need : cap);
    hipLaunchKernelGGL((k_class_wd<LA, LB, LC, LD, MODE>), dim3((int)(g > 0 ? g : 1)), dim3(64 * kWDWaves), 0, st,
                       a);
    return hipGetLastError() == hipSuccess ? 1 : -1;
  }
}

// Workgroups of the per-thread kernel that the current device holds at once
// (occupancy x CUs), computed once per kernel.
template <int MODE, typename R>
long long resident_blocks() {
  static const long long v = [] {
    constexpr size_t lds = class_lds_bytes<LA, LB, LC, LD>();
    constexpr int bs = class_block<LA, LB, LC, LD>();
    int dev = 0, ncu = 0, nb = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
      ncu = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_class<LA, LB, LC, LD, MODE, R, bs>, bs, lds) !=
            hipSuccess || nb < 1)
      nb = 1;
    return (long long)nb * ncu;
  }();
  return v;
}

// grid > 0: that many workgroups; grid <= 0: size the grid from the work
// (with -grid as the multiple of resident workgroups when grid < 0), a.total = tiles (RHF/UHF) or pairs
// (Schwarz).  Per-thread kernel: one wavefront per tile (one lane per pair)
// and slice, at most a few times the workgroups the device holds at once
// over all slices, so each workgroup fills its LDS Boys table once and then
// strides over many tiles (4x for the 512-lane classes, whose tiles differ
// most in cost, so that later workgroups even out the end; 2x otherwise);
// wave kernel: one 128-lane workgroup per (tile, ket position) or pair, at
// most 2^20.
template <int MODE, typename R>
int launch_mode(const ClassArgs& a, int grid, hipStream_t st) {
  if constexpr (kWave) {
    if constexpr (MODE != kSchwarz && std::is_same<R, double>::value)
      if (a.herm) return launch_wd<MODE>(a, grid, st);
    const long long items = MODE == kSchwarz ? a.total : a.total * 64;
    const long long cap = grid < 0 ? (long long)(-grid) << 18 : (1 << 20);
    const long long g = grid > 0 ? grid : (items < cap ? items : cap);
    return launch_wave<MODE, R>(a, (int)(g > 0 ? g : 1), st);
  } else {
    constexpr int bs = class_block<LA, LB, LC, LD>();
    const long long lanes = MODE == kSchwarz ? a.total : a.total * 64;
    const long long b = (lanes + bs - 1) / bs;
    const long long over = grid < 0 ? -grid : (bs >= 512 ? 4 : 2);
    const long long bmax = std::max(1LL, over * resident_blocks<MODE, R>() / SI::NSLICE);
    const long long g = grid > 0 ? grid : (b < bmax ? b : bmax);
    return launch_thread<MODE, R>(a, (int)(g > 0 ? g : 1), st);
  }
}
}  // namespace

int SQ_LAUNCHER(const sqjk::ClassArgs& a, int mode, int fp32, int grid, hipStream_t st) {
  switch (mode) {
    case sqjk::kRHF:
      if constexpr (kFP32)
        if (fp32) return launch_mode<sqjk::kRHF, float>(a, grid, st);
      return launch_mode<sqjk::kRHF, double>(a, grid, st);
    case sqjk::kUHF:
      if constexpr (kFP32)
        if (fp32) return launch_mode<sqjk::kUHF, float>(a, grid, st);
      return launch_mode<sqjk::kUHF, double>(a, grid, st);
#if SQ_LA == SQ_LC && SQ_LB == SQ_LD
    case sqjk::kSchwarz: return fp32 ? -1 : launch_mode<sqjk::kSchwarz, double>(a, grid, st);
#endif
    default: return -1;
  }
}
