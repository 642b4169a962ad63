// sq_jk_classes.h - the angular classes the engine launches and their
// host-side launchers (one translation unit per class, see sq_jk_class.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "sq_jk_dev.h"

// X(LA,LB,LC,LD): canonical quartet classes for l <= 2.
#define SQJK_CLASSES(X)                                                                  \
  X(0, 0, 0, 0) X(1, 0, 0, 0) X(1, 0, 1, 0) X(1, 1, 0, 0) X(1, 1, 1, 0) X(1, 1, 1, 1)     \
  X(2, 0, 0, 0) X(2, 0, 1, 0) X(2, 0, 1, 1) X(2, 0, 2, 0) X(2, 1, 0, 0) X(2, 1, 1, 0)     \
  X(2, 1, 1, 1) X(2, 1, 2, 0) X(2, 1, 2, 1) X(2, 2, 0, 0) X(2, 2, 1, 0) X(2, 2, 1, 1)     \
  X(2, 2, 2, 0) X(2, 2, 2, 1) X(2, 2, 2, 2)

namespace sqjk {
// Launches every slice of the class in `mode` (kRHF, kUHF or kSchwarz; the
// latter only for LA==LC, LB==LD), evaluating integrals in FP32 when fp32 is
// set (kRHF/kUHF only).  grid <= 0 sizes the grid from args.total.  Returns
// the number of launches, or -1 if a launch failed.
typedef int (*ClassLauncher)(const ClassArgs& a, int mode, int fp32, int grid, hipStream_t st);
}

#define SQJK_DECL(a, b, c, d) \
  int sqjk_launch_##a##b##c##d(const sqjk::ClassArgs& args, int mode, int fp32, int grid, hipStream_t st);
SQJK_CLASSES(SQJK_DECL)
#undef SQJK_DECL
