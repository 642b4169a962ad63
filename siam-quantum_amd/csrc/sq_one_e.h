// sq_one_e.h - launcher of the one-electron kernels (sq_one_e.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "sq_jk_dev.h"

namespace sqjk {

struct OneEArgs {
  const SubShell* __restrict__ sh;
  const PairRec* __restrict__ pr;     // all unique sub-shell pairs of one class
  const double* __restrict__ pexp;    // primitive exponents (SubShell::prim_off)
  const double* __restrict__ pcoef;   // primitive coefficient x norm of the first component
  const double* __restrict__ bscale;  // per basis function component scale
  const double* __restrict__ atom;    // x, y, z, Z per nucleus
  const double* __restrict__ boys;    // Boys table (global)
  double* T;                          // n x n outputs (NULL: skipped)
  double* V;
  double* S;
  int np, natom, n;
};

// pair_class: 0 (s,s) 1 (p,s) 2 (p,p) 3 (d,s) 4 (d,p) 5 (d,d).  0 on success.
int launch_one_electron(int pair_class, const OneEArgs& a, hipStream_t st);

// One-electron part of the energy gradient (sq_grad_one.hip).
struct OneEGradArgs {
  const SubShell* __restrict__ sh;
  const PairRec* __restrict__ pr;     // all unique sub-shell pairs of one class
  const double* __restrict__ pexp;
  const double* __restrict__ pcoef;
  const double* __restrict__ bscale;
  const double* __restrict__ atom;    // x, y, z, Z per nucleus
  const double* __restrict__ boys;
  const double* __restrict__ PT;      // total density (n x n)
  const double* __restrict__ QT;      // energy-weighted density (n x n)
  double* gsub;                       // [nsub][3] gradient on basis-function centres (atomics)
  double* gatom;                      // [natom][3] gradient on nuclei, operator part (atomics)
  int np, natom, n;
};

int launch_one_electron_grad(int pair_class, const OneEGradArgs& a, hipStream_t st);

}  // namespace sqjk
