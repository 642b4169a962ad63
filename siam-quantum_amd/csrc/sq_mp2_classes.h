// sq_mp2_classes.h - integral-store kernel of the MP2 path and its per-class
// launchers (one translation unit per class, see sq_mp2_class.hip).
//
// k_store runs the J/K work plan of a context (tiles of bra rows x 64 ket
// pairs, ket-component slices) with the class kernels' quartet evaluation in
// kStore mode: every surviving shell quartet writes its contracted integrals
// into the dense matrix V[pair(p,q)][pair(r,s)], n(n+1)/2 on a side, passed in
// ClassArgs::sch_out.  No density screening: MP2 needs every quartet above
// the Schwarz cut-off of the plan.
#pragma once
#include <hip/hip_runtime.h>

#include "sq_jk_classes.h"

namespace sqjk {

template <int LA, int LB, int LC, int LD, int BS>
__global__ void __launch_bounds__(BS) k_store(ClassArgs A) {
  using SI = SliceInfo<LA, LB, LC, LD>;
  constexpr int L = LA + LB + LC + LD;
  extern __shared__ double lds[];
  boys_load_lds<L>(lds, A.boys);
  __syncthreads();
  const int slice = blockIdx.y;
  sfor<SI::NSLICE>([&](auto sI) {
    constexpr int S = decltype(sI)::value;
    constexpr int CD0 = S * SI::SZ;
    constexpr int CD1 = (CD0 + SI::SZ < SI::NCD) ? CD0 + SI::SZ : SI::NCD;
    static_assert(CD0 < CD1, "empty ket-component slice");
    if (S != slice) return;
    const int lane = threadIdx.x & 63;
    const int wpb = blockDim.x >> 6;
    const long long wstride = (long long)gridDim.x * wpb;
    const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (long long w = (long long)blockIdx.x * wpb + wib; w < A.total; w += wstride) {
      const Tile tl = uload(A.tile + w);
      const Segment sg = uload(A.seg + tl.seg);
      const View V = make_view(A, sg.inst);
      const int j = tl.k0 + lane;  // ket position in the group
      for (int r = tl.r0; r < tl.r1; ++r) {
        const int ib = uload(A.rbra + r);
        const int lo = uload(A.rlo + r);
        const bool in = j >= lo && j < uload(A.rhi + r);
        if (!__any(in)) continue;              // uniform
        const int ik = sg.ket0 + (in ? j : lo);  // idle lanes evaluate a valid pair and store nothing
        quartet<LA, LB, LC, LD, CD0, CD1, kStore, double>(A, V, ib, ik, sg.nppb, tl.v0, tl.v1, in, lds);
      }
    }
  });
}

// Launches the store kernel of one class over args.total tiles (every slice).
// Returns the number of launches, or -1 if a launch failed.
typedef int (*StoreLauncher)(const ClassArgs& a, hipStream_t st);

}  // namespace sqjk

#define SQMP2_DECL(a, b, c, d) int sqmp2_store_##a##b##c##d(const sqjk::ClassArgs& args, hipStream_t st);
SQJK_CLASSES(SQMP2_DECL)
#undef SQMP2_DECL
