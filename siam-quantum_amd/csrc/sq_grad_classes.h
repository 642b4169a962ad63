// sq_grad_classes.h - host-side launchers of the two-electron gradient
// kernel, one per angular class (one translation unit per class, see
// sq_grad_class.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "sq_grad_dev.h"
#include "sq_jk_classes.h"

namespace sqjk {
// Launches every ket-component slice of the class over a.c.total tiles.
// grid <= 0 sizes the grid from the work.  Returns the number of launches,
// or -1 if a launch failed.
typedef int (*GradLauncher)(const GradArgs& a, int grid, hipStream_t st);
}

#define SQGRAD_DECL(a, b, c, d) int sqgrad_launch_##a##b##c##d(const sqjk::GradArgs& args, int grid, hipStream_t st);
SQJK_CLASSES(SQGRAD_DECL)
#undef SQGRAD_DECL
