// sq_jk_wd.h - (dd|dd) kernel: one wavefront per contracted quartet, lanes
// over Hermite and Cartesian components, no workgroup barriers.
//
// Per primitive quartet (all quantities wavefront-uniform except the lane's
// component):
//   R_{tuv}  by level recursion in the wavefront's LDS slice,
//   W[cd][t] = sum_T (-1)^|T| E^{cd}_T R_{t+T}       lane = bra Hermite index t,
//   I[ab][cd] += pref sum_t E^{ab}_t W[cd][t]          lane = ket component cd.
// The Hermite coefficients E of every dd primitive pair are precomputed
// (k_herm_table, geometry-only) and read through the constant address space,
// so the two contractions are register FMAs with scalar operands; a lane
// reads from LDS only the 35 R (or W) values of its own component.  The
// contracted block stays in registers (lane cd holds column cd) until the
// digestion, which goes through LDS with one atomic per J/K element touched,
// straight into the n x n half-accumulators.  The four wavefronts of a
// workgroup work on different quartets and share nothing but the workgroup.
#pragma once
#include "sq_jk_wave.h"

namespace sqjk {

constexpr int kWDWaves = 4;  // wavefronts per workgroup

// Hermite coefficients of one primitive pair of classes (LI, LJ): per axis
// E[i][j][t], i <= LI, j <= LJ, t <= LI + LJ (zero above i + j).
template <int LI, int LJ>
struct HermTabLayout {
  static constexpr int PER_AXIS = (LI + 1) * (LJ + 1) * (LI + LJ + 1);
  static constexpr int SIZE = 3 * PER_AXIS;
  __host__ __device__ static constexpr int at(int axis, int i, int j, int t) {
    return axis * PER_AXIS + (i * (LJ + 1) + j) * (LI + LJ + 1) + t;
  }
};

// One lane per primitive pair of the class-(LI, LJ) pair list: the x, y, z
// coefficient tables of primitive pair v of pair k at
// out + (pp_index - base) * SIZE, pp_index = pp_off + v * pp_stride.
template <int LI, int LJ>
__global__ void k_herm_table(const PairRec* __restrict__ pairs, int npair, const PrimPair* __restrict__ pp,
                             const SubShell* __restrict__ sh, int base, double* __restrict__ out) {
  using HL = HermTabLayout<LI, LJ>;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;; e += (long long)gridDim.x * blockDim.x) {
    // e enumerates (pair, primitive) with at most 64 primitive pairs per pair
    const long long k = e >> 6;
    const int v = (int)(e & 63);
    if (k >= npair) break;
    const PairRec pr = pairs[k];
    if (v >= pr.npp) continue;
    const int idx = pr.pp_off + v * pr.pp_stride;
    const PrimPair P = pp[idx];
    const SubShell sA = sh[pr.a], sB = sh[pr.b];
    double* o = out + (size_t)(idx - base) * HL::SIZE;
    const double xa[3] = {P.Px - sA.x, P.Py - sA.y, P.Pz - sA.z};
    const double xb[3] = {P.Px - sB.x, P.Py - sB.y, P.Pz - sB.z};
    for (int ax = 0; ax < 3; ++ax) {
      Herm1<LI, LJ, double> E;
      herm_coef<LI, LJ, double>(E, xa[ax], xb[ax], P.hp);
      for (int i = 0; i <= LI; ++i)
        for (int j = 0; j <= LJ; ++j)
          for (int t = 0; t <= LI + LJ; ++t) o[HL::at(ax, i, j, t)] = t <= i + j ? E.e[i][j][t] : 0.0;
    }
  }
}

template <int LA, int LB, int LC, int LD>
struct WDLayout {
  static constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  static constexpr int LAB = LA + LB, LCD = LC + LD, L = LAB + LCD;
  static constexpr int NAB = NA * NB, NCD = NC * ND, NINT = NAB * NCD;
  static constexpr int NH = nherm(L), NHAB = nherm(LAB), NHCD = nherm(LCD);
  static constexpr int NW = NHAB * NCD;
  static constexpr int NBIG = NW > NINT ? NW : NINT;
  static constexpr int NDX = NB * ND + NA * ND + NB * NC + NA * NC;  // exchange density blocks
  // per-wavefront LDS (doubles)
  static constexpr int oR0 = 0, oR1 = NH, oG = 2 * NH, oW = oG + L + 1, oD = oW + NBIG;
  static constexpr int PER_WAVE = oD + NAB + NCD + NDX;
  static constexpr int NSL = (NH + 63) / 64;  // R entries per lane
  // workgroup-shared tables (ints): R recursion step per entry (4 ints),
  // W-stage gather offsets [T][lane]
  static constexpr int oStep = 0, oRoff = 4 * NH, SHARED_INTS = oRoff + NHCD * 64;
  static_assert(NHAB <= 64 && NCD <= 64 && NAB <= 64, "one lane per component");
};

// LDS written by some lanes of the wavefront and read by others: the
// wavefront runs in lockstep and its LDS operations complete in order, so a
// fence that keeps the compiler from moving memory operations across is
// enough.
__device__ __forceinline__ void wsync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// RHF/UHF only (the Schwarz pass uses k_class_wave).  A.herm holds the
// Hermite tables of this class's primitive pairs (bra and ket are the same
// pair class), A.herm_base the primitive-pair index of its first entry.
template <int LA, int LB, int LC, int LD, int MODE>
__global__ void __launch_bounds__(64 * kWDWaves, 2) k_class_wd(ClassArgs A) {
  static_assert(MODE == kRHF || MODE == kUHF, "Schwarz mode uses k_class_wave");
  using W_ = WDLayout<LA, LB, LC, LD>;
  using HA = HermTabLayout<LA, LB>;
  using HC = HermTabLayout<LC, LD>;
  constexpr int NA = W_::NA, NB = W_::NB, NC = W_::NC, ND = W_::ND;
  constexpr int NAB = W_::NAB, NCD = W_::NCD, L = W_::L, LAB = W_::LAB, LCD = W_::LCD;
  constexpr int NHAB = W_::NHAB, NHCD = W_::NHCD;
  __shared__ double lds_all[kWDWaves * W_::PER_WAVE];
  __shared__ int tabs[W_::SHARED_INTS];
  // the wavefront index through readfirstlane, so that the compiler sees it
  // (and all work-list reads derived from it) as uniform
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  double* lds = lds_all + wave * W_::PER_WAVE;
  double* g = lds + W_::oG;
  double* Wt = lds + W_::oW;

  // Recursion step of entry e of R:
  //   R^n_{tuv} = P_ax R^{n+1}_{..-1..} + c2 R^{n+1}_{..-2..}
  // along the first axis with a nonzero index; step[e] = (ax + 1 | order << 4,
  // i1, i2, c2), ax = -1 for e = 0.  W-stage gather offsets: lane t reads
  // R_{t+T} at roff[T * 64 + t].
  int* step = tabs + W_::oStep;
  int* roff = tabs + W_::oRoff;
  for (int e = threadIdx.x; e < W_::NH; e += blockDim.x) {
    const int t = kHerm.t[e], u = kHerm.u[e], v = kHerm.v[e];
    int ax = -1, i1 = 0, i2 = 0, c2 = 0;
    if (t > 0) {
      ax = 0; i1 = hidx(t - 1, u, v); i2 = t > 1 ? hidx(t - 2, u, v) : 0; c2 = t - 1;
    } else if (u > 0) {
      ax = 1; i1 = hidx(t, u - 1, v); i2 = u > 1 ? hidx(t, u - 2, v) : 0; c2 = u - 1;
    } else if (v > 0) {
      ax = 2; i1 = hidx(t, u, v - 1); i2 = v > 1 ? hidx(t, u, v - 2) : 0; c2 = v - 1;
    }
    step[4 * e] = (ax + 1) | ((t + u + v) << 4);
    step[4 * e + 1] = i1;
    step[4 * e + 2] = i2;
    step[4 * e + 3] = c2;
  }
  for (int e = threadIdx.x; e < NHCD * 64; e += blockDim.x) {
    const int T = e >> 6, l = e & 63;
    const int h = l < NHAB ? l : 0;
    roff[e] = hidx(kHerm.t[h] + kHerm.t[T], kHerm.u[h] + kHerm.u[T], kHerm.v[h] + kHerm.v[T]);
  }
  __syncthreads();  // the only workgroup barrier: the wavefronts run independently from here

  const long long items = A.total * 64;
  for (long long it = (long long)blockIdx.x * kWDWaves + wave; it < items; it += (long long)gridDim.x * kWDWaves) {
    const Tile tl = uload(A.tile + (it >> 6));
    const Segment sg = uload(A.seg + tl.seg);
    const int j = tl.k0 + (int)(it & 63);
    const int nppb = sg.nppb, v0 = tl.v0, v1 = tl.v1, ket0 = sg.ket0;
    const View V = make_view(A, sg.inst);
    const int n = V.n;
    for (int r = tl.r0; r < tl.r1; ++r) {
      if (j < uload(A.rlo + r) || j >= uload(A.rhi + r)) continue;  // uniform
      const int ib = uload(A.rbra + r), ik = ket0 + j;
      const PairRec bp = uload(A.bra + ib), kp = uload(A.ket + ik);
      if (!density_screen_pass(V, bp, kp)) continue;  // uniform
      const SubShell sA = uload(A.sh + bp.a), sB = uload(A.sh + bp.b), sC = uload(A.sh + kp.a),
                     sD = uload(A.sh + kp.b);

      // lane cd (< NCD) accumulates column cd of the contracted block
      double acc[NAB];
      sfor<NAB>([&](auto I) { acc[I.value] = 0.0; });

      for (int u = 0; u < nppb; ++u) {
        const int pi = bp.pp_off + u * bp.pp_stride;
        const PrimPair P = uload(A.pp + pi);
        const double* Eab0 = A.herm + (size_t)(pi - A.herm_base) * HA::SIZE;
        for (int v = v0; v < v1; ++v) {
          const int qi = kp.pp_off + v * kp.pp_stride;
          const PrimPair Q = uload(A.pp + qi);
          const double s = P.p + Q.p;
          const double rs = rsqrt_nr(s);
          if (kTwoPi25 * P.eb * Q.eb * rs < kPrimCutoff) continue;  // uniform
          const double* Ecd = A.herm + (size_t)(qi - A.herm_base) * HC::SIZE;
          const double alpha = P.p * Q.p * (rs * rs);
          const double X = P.Px - Q.Px, Y = P.Py - Q.Py, Z = P.Pz - Q.Pz;
          const double T = alpha * (X * X + Y * Y + Z * Z);
          const double pref = kTwoPi25 * P.cc * Q.cc * rs;

          // g_m = (-2 alpha)^m F_m(T) = R^m_{000}, one lane each
          if (lane <= L) {
            double pw = 1.0;
            for (int m = 0; m < lane; ++m) pw *= -2.0 * alpha;
            g[lane] = pw * boys_one(lane, T, A.boys);
          }
          wsync();
          // R^0_{tuv} over the levels n = L-1 .. 0 (level n holds the orders
          // up to L-n), alternating buffers so that level 0 lands in R0
#pragma unroll 1
          for (int lev = L - 1; lev >= 0; --lev) {
            const double* prev = (lev == L - 1) ? g + L : (((lev + 1) & 1) ? lds + W_::oR1 : lds + W_::oR0);
            double* next = (lev & 1) ? lds + W_::oR1 : lds + W_::oR0;
#pragma unroll
            for (int sl = 0; sl < W_::NSL; ++sl) {
              const int e = lane + 64 * sl;
              if (e < W_::NH) {
                const int s0 = step[4 * e];
                if ((s0 >> 4) <= L - lev) {
                  const int ax = (s0 & 15) - 1;
                  const double pa = ax == 0 ? X : (ax == 1 ? Y : Z);
                  next[e] = ax < 0 ? g[lev] : pa * prev[step[4 * e + 1]] + double(step[4 * e + 3]) * prev[step[4 * e + 2]];
                }
              }
            }
            wsync();
          }
          // W[cd][t] = sum_T (-1)^{|T|} E^{cd}_T R_{t+T}, lane = t
          // the E tables are read where they are used: an opaque copy of the
          // pointer keeps the compiler from hoisting the (invariant) loads out
          // of the loops, where they would all be live at once
          const double* Ecw = Ecd;
          asm volatile("" : "+s"(Ecw));
          if (lane < NHAB) {
            const double* R0 = lds + W_::oR0;
            double rg[NHCD];
            sfor<NHCD>([&](auto TI) { rg[TI.value] = R0[roff[TI.value * 64 + lane]]; });
            sfor<NCD>([&](auto cdI) {
              __builtin_amdgcn_sched_barrier(0);
              constexpr int cd = decltype(cdI)::value;
              constexpr Cart c = cart(LC, cd / ND), d = cart(LD, cd % ND);
              double w = 0.0;
              sfor<c.x + d.x + 1>([&](auto TI) {
                constexpr int T_ = decltype(TI)::value;
                double wy = 0.0;
                sfor<c.y + d.y + 1>([&](auto UI) {
                  constexpr int U_ = decltype(UI)::value;
                  double wz = 0.0;
                  sfor<c.z + d.z + 1>([&](auto VI) {
                    constexpr int V_ = decltype(VI)::value;
                    constexpr double sgn = ((T_ + U_ + V_) & 1) ? -1.0 : 1.0;
                    wz += (sgn * uload(Ecw + HC::at(2, c.z, d.z, V_))) * rg[hidx(T_, U_, V_)];
                  });
                  wy += uload(Ecw + HC::at(1, c.y, d.y, U_)) * wz;
                });
                w += uload(Ecw + HC::at(0, c.x, d.x, T_)) * wy;
              });
              Wt[cd * NHAB + lane] = w;
            });
          }
          wsync();
          // I[ab][cd] += pref sum_t E^{ab}_t W[cd][t], lane = cd
          const double* Eab = Eab0;
          asm volatile("" : "+s"(Eab));
          if (lane < NCD) {
            double wc[NHAB];
            sfor<NHAB>([&](auto tI) { wc[tI.value] = Wt[lane * NHAB + tI.value]; });
            sfor<NAB>([&](auto abI) {
              __builtin_amdgcn_sched_barrier(0);
              constexpr int ab = decltype(abI)::value;
              constexpr Cart a = cart(LA, ab / NB), b = cart(LB, ab % NB);
              double val = 0.0;
              sfor<a.x + b.x + 1>([&](auto tI) {
                constexpr int t_ = decltype(tI)::value;
                double vy = 0.0;
                sfor<a.y + b.y + 1>([&](auto uI) {
                  constexpr int u_ = decltype(uI)::value;
                  double vz = 0.0;
                  sfor<a.z + b.z + 1>([&](auto vI) {
                    constexpr int v_ = decltype(vI)::value;
                    vz += uload(Eab + HA::at(2, a.z, b.z, v_)) * wc[hidx(t_, u_, v_)];
                  });
                  vy += uload(Eab + HA::at(1, a.y, b.y, u_)) * vz;
                });
                val += uload(Eab + HA::at(0, a.x, b.x, t_)) * vy;
              });
              acc[ab] += pref * val;
            });
          }
          wsync();
        }
      }

      // ---- digestion (the formulas of k_class_wave) ----
      const int fa = sA.bf_off, fb = sB.bf_off, fc = sC.bf_off, fd = sD.bf_off;
      const int ga = sA.gbf, gb = sB.gbf, gc = sC.gbf, gd = sD.gbf;
      double deg = 1.0;
      if (bp.a == bp.b) deg *= 0.5;
      if (kp.a == kp.b) deg *= 0.5;
      if (A.same && ib == ik) deg *= 0.5;
      double* EE = Wt;  // the contracted, scaled block [ab][cd]
      double* Dst = lds + W_::oD;
      if (lane < NCD) {
        const double scd = deg * A.bscale[gc + lane / ND] * A.bscale[gd + lane % ND];
        sfor<NAB>([&](auto abI) {
          constexpr int ab = decltype(abI)::value;
          EE[ab * NCD + lane] = acc[ab] * scd * A.bscale[ga + ab / NB] * A.bscale[gb + ab % NB];
        });
      }
      if (lane < NAB) Dst[lane] = V.DJ[(fa + lane / NB) * n + fb + lane % NB];
      if (lane < NCD) Dst[NAB + lane] = V.DJ[(fc + lane / ND) * n + fd + lane % ND];
      wsync();
      // Coulomb: A_J[a,b] += 2 sum_cd v D[c,d];  A_J[c,d] += 2 sum_ab v D[a,b]
      if (lane < NAB) {
        double s = 0.0;
#pragma unroll 6
        for (int cd = 0; cd < NCD; ++cd) s += EE[lane * NCD + cd] * Dst[NAB + cd];
        fatomic(&V.AJ[(fa + lane / NB) * n + fb + lane % NB], 2.0 * s);
      }
      if (lane < NCD) {
        double s = 0.0;
#pragma unroll 6
        for (int ab = 0; ab < NAB; ++ab) s += EE[ab * NCD + lane] * Dst[ab];
        fatomic(&V.AJ[(fc + lane / ND) * n + fd + lane % ND], 2.0 * s);
      }
      // Exchange per spin: A_K[a,c] += v D[b,d]; A_K[b,c] += v D[a,d];
      //                    A_K[a,d] += v D[b,c]; A_K[b,d] += v D[a,c]
      constexpr int nspin = (MODE == kUHF) ? 2 : 1;
#pragma unroll 1
      for (int sp = 0; sp < nspin; ++sp) {
        const double* D = sp ? V.DB : V.DA;
        double* AK = sp ? V.AKB : V.AKA;
        double* dbd = Dst + NAB + NCD;  // D[b,d]  NB x ND
        double* dad = dbd + NB * ND;    // D[a,d]  NA x ND
        double* dbc = dad + NA * ND;    // D[b,c]  NB x NC
        double* dac = dbc + NB * NC;    // D[a,c]  NA x NC
        wsync();
        for (int e = lane; e < NB * ND; e += 64) dbd[e] = D[(fb + e / ND) * n + fd + e % ND];
        for (int e = lane; e < NA * ND; e += 64) dad[e] = D[(fa + e / ND) * n + fd + e % ND];
        for (int e = lane; e < NB * NC; e += 64) dbc[e] = D[(fb + e / NC) * n + fc + e % NC];
        for (int e = lane; e < NA * NC; e += 64) dac[e] = D[(fa + e / NC) * n + fc + e % NC];
        wsync();
        constexpr int NK = NA * NC + NB * NC + NA * ND + NB * ND;
        for (int e = lane; e < NK; e += 64) {
          double s = 0.0;
          if (e < NA * NC) {  // K[a,c] += sum_bd v D[b,d]
            const int a = e / NC, c = e % NC;
#pragma unroll 1
            for (int b = 0; b < NB; ++b)
              for (int d = 0; d < ND; ++d) s += EE[(a * NB + b) * NCD + c * ND + d] * dbd[b * ND + d];
            fatomic(&AK[(fa + a) * n + fc + c], s);
          } else if (e < NA * NC + NB * NC) {  // K[b,c] += sum_ad v D[a,d]
            const int q = e - NA * NC, b = q / NC, c = q % NC;
#pragma unroll 1
            for (int a = 0; a < NA; ++a)
              for (int d = 0; d < ND; ++d) s += EE[(a * NB + b) * NCD + c * ND + d] * dad[a * ND + d];
            fatomic(&AK[(fb + b) * n + fc + c], s);
          } else if (e < NA * NC + NB * NC + NA * ND) {  // K[a,d] += sum_bc v D[b,c]
            const int q = e - NA * NC - NB * NC, a = q / ND, d = q % ND;
#pragma unroll 1
            for (int b = 0; b < NB; ++b)
              for (int c = 0; c < NC; ++c) s += EE[(a * NB + b) * NCD + c * ND + d] * dbc[b * NC + c];
            fatomic(&AK[(fa + a) * n + fd + d], s);
          } else {  // K[b,d] += sum_ac v D[a,c]
            const int q = e - NA * NC - NB * NC - NA * ND, b = q / ND, d = q % ND;
#pragma unroll 1
            for (int a = 0; a < NA; ++a)
              for (int c = 0; c < NC; ++c) s += EE[(a * NB + b) * NCD + c * ND + d] * dac[a * NC + c];
            fatomic(&AK[(fb + b) * n + fd + d], s);
          }
        }
      }
      wsync();
    }  // rows
  }
}

}  // namespace sqjk
