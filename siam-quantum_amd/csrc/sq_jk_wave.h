// sq_jk_wave.h - cooperative kernel for the high-angular-momentum classes.
//
// One 128-lane workgroup (two wavefronts) owns one contracted shell quartet
// at a time.  Per primitive quartet the lanes build, in LDS,
//   R_{TUV}   (level recursion, a few entries per lane and level),
//   E^{ab}, E^{cd} Hermite coefficients (one axis per lane),
//   W[cd][tuv] = sum_{TUV} (-1)^{T+U+V} E^{cd}_{TUV} R_{t+T,u+U,v+V},
// and every lane accumulates its share of the (ab|cd) block in LDS.
// The contracted block is then staged in LDS and digested into J/K with one
// atomic per J/K element touched by the quartet.
#pragma once
#include "sq_jk_dev.h"

namespace sqjk {

// (t,u,v) of the compact Hermite index, up to total order 9
struct HermTab {
  unsigned char t[nherm(9)], u[nherm(9)], v[nherm(9)];
};
__host__ __device__ constexpr HermTab make_herm_tab() {
  HermTab h{};
  for (int n = 0; n <= 9; ++n)
    for (int m = 0; m <= n; ++m)
      for (int v = 0; v <= m; ++v) {
        const int u = m - v, t = n - m;
        const int i = hidx(t, u, v);
        h.t[i] = (unsigned char)t; h.u[i] = (unsigned char)u; h.v[i] = (unsigned char)v;
      }
  return h;
}
__constant__ constexpr HermTab kHerm = make_herm_tab();

// Threads of the cooperative workgroup (two wavefronts on one quartet).
constexpr int kWT = 128;

__device__ __forceinline__ void wave_sync() { __syncthreads(); }

// F_n(T) for one n straight from the global table (Taylor about the grid point)
__device__ __forceinline__ double boys_one(int n, double T, const double* __restrict__ tab) {
  if (T < kBoysTMax - 0.5 * kBoysStep) {
    const int k = (int)(T * (1.0 / kBoysStep) + 0.5);
    const double d = k * kBoysStep - T;
    const double* row = tab + k * kBoysCols + n;
    double s = row[kBoysTaylor - 1];
#pragma unroll
    for (int j = kBoysTaylor - 2; j >= 0; --j) s = s * (d * (1.0 / (j + 1))) + row[j];
    return s;
  }
  // asymptotic: (2n-1)!!/(2T)^n * sqrt(pi/T)/2 minus a term below 1e-17 here
  double f = 0.88622692545275801365 * sqrt(1.0 / T);
  const double e = exp(-T), h = 0.5 / T;
  for (int m = 0; m < n; ++m) f = ((2 * m + 1) * f - e) * h;
  return f;
}

template <int LI, int LJ, typename R>
__device__ __forceinline__ void herm_axis_to_lds(R* out, R xpa, R xpb, R h2p) {
  Herm1<LI, LJ, R> E;
  herm_coef<LI, LJ, R>(E, xpa, xpb, h2p);
  sfor<LI + 1>([&](auto i) {
    sfor<LJ + 1>([&](auto j) {
      sfor<LI + LJ + 1>([&](auto t) {
        out[(i.value * (LJ + 1) + j.value) * (LI + LJ + 1) + t.value] =
            (t.value <= i.value + j.value) ? E.e[i.value][j.value][t.value] : R(0);
      });
    });
  });
}

template <int LA, int LB, int LC, int LD>
struct WaveLayout {
  static constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  static constexpr int LAB = LA + LB, LCD = LC + LD, L = LAB + LCD;
  static constexpr int NAB = NA * NB, NCD = NC * ND, NINT = NAB * NCD;
  static constexpr int NH = nherm(L), NHAB = nherm(LAB);
  static constexpr int EA = (LA + 1) * (LB + 1) * (LAB + 1);  // per axis
  static constexpr int EC = (LC + 1) * (LD + 1) * (LCD + 1);
  static constexpr int NW = NCD * NHAB;
  static constexpr int NBIG = NW > NINT ? NW : NINT;
  // LDS offsets (doubles)
  static constexpr int oR = 0, oR2 = oR + NH, oEa = oR2 + NH, oEc = oEa + 3 * EA, oW = oEc + 3 * EC;
  static constexpr int oD = oW + NBIG;                 // staged density blocks
  static constexpr int ND6 = NAB + NCD + NB * ND + NA * ND + NB * NC + NA * NC;
  static constexpr int oH = oD + 2 * ND6;              // (t,u,v) of the Hermite index, one int each
  static constexpr int oAcc = oH + (NH + 1) / 2;       // contracted (ab|cd) block, lane-owned entries
  static constexpr int TOTAL = oAcc + NINT;
  static constexpr int NSL = (NH + kWT - 1) / kWT;     // R entries per lane
};

// R: precision of the integral evaluation; staging of the contracted block
// and the digestion are FP64.
template <int LA, int LB, int LC, int LD, int MODE, typename R = double>
__global__ void __launch_bounds__(kWT, 2) k_class_wave(ClassArgs A) {
  using W_ = WaveLayout<LA, LB, LC, LD>;
  constexpr int NA = W_::NA, NB = W_::NB, NC = W_::NC, ND = W_::ND;
  constexpr int NAB = W_::NAB, NCD = W_::NCD, NINT = W_::NINT, L = W_::L, LAB = W_::LAB;
  __shared__ double lds[W_::TOTAL];
  R* Rh = reinterpret_cast<R*>(lds + W_::oR);    // views sized in doubles,
  R* Rh2 = reinterpret_cast<R*>(lds + W_::oR2);  // so an FP32 view fits
  R* Ea = reinterpret_cast<R*>(lds + W_::oEa);
  R* Ec = reinterpret_cast<R*>(lds + W_::oEc);
  R* Wt = reinterpret_cast<R*>(lds + W_::oW);
  int* hdec = reinterpret_cast<int*>(lds + W_::oH);
  // Each lane accumulates the entries o = lane + kWT k of the block in LDS
  // (runtime loops over them keep the register count low).
  R* acc = reinterpret_cast<R*>(lds + W_::oAcc);
  const int lane = threadIdx.x;

  // Hermite index decode table in LDS (t | u << 8 | v << 16), and per lane
  // the recursion step of each R entry it owns, e = lane + kWT s:
  //   R^n_{tuv} = P_ax R^{n+1}_{..-1..} + c2 R^{n+1}_{..-2..}
  // along the first axis with a nonzero index (ax = -1 for e = 0).
  for (int e = lane; e < W_::NH; e += kWT) hdec[e] = kHerm.t[e] | (kHerm.u[e] << 8) | (kHerm.v[e] << 16);
  int r_ax[W_::NSL], r_i1[W_::NSL], r_i2[W_::NSL], r_ord[W_::NSL];
  R r_c2[W_::NSL];
#pragma unroll
  for (int sl = 0; sl < W_::NSL; ++sl) {
    const int e = lane + kWT * sl;
    r_ax[sl] = -1; r_i1[sl] = 0; r_i2[sl] = 0; r_c2[sl] = R(0); r_ord[sl] = 1 << 20;
    if (e < W_::NH) {
      const int t = kHerm.t[e], u = kHerm.u[e], v = kHerm.v[e];
      r_ord[sl] = t + u + v;
      if (t > 0) {
        r_ax[sl] = 0; r_i1[sl] = hidx(t - 1, u, v); r_i2[sl] = t > 1 ? hidx(t - 2, u, v) : 0; r_c2[sl] = R(t - 1);
      } else if (u > 0) {
        r_ax[sl] = 1; r_i1[sl] = hidx(t, u - 1, v); r_i2[sl] = u > 1 ? hidx(t, u - 2, v) : 0; r_c2[sl] = R(u - 1);
      } else if (v > 0) {
        r_ax[sl] = 2; r_i1[sl] = hidx(t, u, v - 1); r_i2[sl] = v > 1 ? hidx(t, u, v - 2) : 0; r_c2[sl] = R(v - 1);
      }
    }
  }
  wave_sync();

  // RHF/UHF: grid-stride over (tile, ket position) items, the rows of the
  // tile one after the other; Schwarz: grid-stride over the pairs.
  const long long items = MODE == kSchwarz ? A.total : A.total * 64;
  for (long long t = blockIdx.x; t < items; t += gridDim.x) {
    // ket primitive pairs [v0, v1): the whole range in Schwarz mode, the
    // tile's share of a segment split along its ket primitive pairs otherwise
    int r0 = 0, r1 = 1, j = 0, nppb = 0, v0 = 0, v1 = 0, ket0 = 0, inst = 0;
    if constexpr (MODE == kSchwarz) {
      nppb = v1 = uload(A.bra + t).npp;
    } else {
      const Tile tl = uload(A.tile + (t >> 6));
      const Segment sg = uload(A.seg + tl.seg);
      r0 = tl.r0; r1 = tl.r1; j = tl.k0 + (int)(t & 63);
      nppb = sg.nppb; v0 = tl.v0; v1 = tl.v1; ket0 = sg.ket0; inst = sg.inst;
    }
    const View V = make_view(A, inst);
    const int n = V.n;
    for (int r = r0; r < r1; ++r) {
    int ib, ik;
    if constexpr (MODE == kSchwarz) {
      ib = ik = (int)t;
    } else {
      if (j < uload(A.rlo + r) || j >= uload(A.rhi + r)) continue;  // uniform
      ib = uload(A.rbra + r);
      ik = ket0 + j;
      if constexpr (MODE != kStore)
        if (!density_screen_pass(V, uload(A.bra + ib), uload(A.ket + ik))) continue;  // uniform
    }
    const PairRec bp = uload(A.bra + ib), kp = uload(A.ket + ik);
    const SubShell sA = uload(A.sh + bp.a), sB = uload(A.sh + bp.b), sC = uload(A.sh + kp.a), sD = uload(A.sh + kp.b);

    for (int o = lane; o < NINT; o += kWT) acc[o] = R(0);

    for (int u = 0; u < nppb; ++u) {
      const PrimPair P = uload(A.pp + (bp.pp_off + u * bp.pp_stride));
      for (int v = v0; v < v1; ++v) {
        const PrimPair Q = uload(A.pp + (kp.pp_off + v * kp.pp_stride));
        const double s = P.p + Q.p;
        const double rs = rsqrt_nr(s);
        if constexpr (MODE != kSchwarz) {
          if (kTwoPi25 * P.eb * Q.eb * rs < kPrimCutoff) continue;  // uniform
        }
        const double alpha_d = P.p * Q.p * (rs * rs);
        const double Xd = P.Px - Q.Px, Yd = P.Py - Q.Py, Zd = P.Pz - Q.Pz;
        const double T = alpha_d * (Xd * Xd + Yd * Yd + Zd * Zd);
        const R X = R(Xd), Y = R(Yd), Z = R(Zd);
        const R pref = R(kTwoPi25 * P.cc * Q.cc * rs);

        // g_m = (-2 alpha)^m F_m(T) = R^m_{000} for m <= L, one lane each, via LDS scratch
        R* g = Wt;  // W is rebuilt below, so its space is free here
        if (lane <= L) {
          double pw = 1.0;
          for (int m = 0; m < lane; ++m) pw *= -2.0 * alpha_d;
          g[lane] = R(pw * boys_one(lane, T, A.boys));
        }
        const R h2p = R(P.hp), h2q = R(Q.hp);
        if (lane == 60) herm_axis_to_lds<LA, LB, R>(Ea + 0 * W_::EA, R(P.Px - sA.x), R(P.Px - sB.x), h2p);
        if (lane == 61) herm_axis_to_lds<LA, LB, R>(Ea + 1 * W_::EA, R(P.Py - sA.y), R(P.Py - sB.y), h2p);
        if (lane == 62) herm_axis_to_lds<LA, LB, R>(Ea + 2 * W_::EA, R(P.Pz - sA.z), R(P.Pz - sB.z), h2p);
        if (lane == 57) herm_axis_to_lds<LC, LD, R>(Ec + 0 * W_::EC, R(Q.Px - sC.x), R(Q.Px - sD.x), h2q);
        if (lane == 58) herm_axis_to_lds<LC, LD, R>(Ec + 1 * W_::EC, R(Q.Py - sC.y), R(Q.Py - sD.y), h2q);
        if (lane == 59) herm_axis_to_lds<LC, LD, R>(Ec + 2 * W_::EC, R(Q.Pz - sC.z), R(Q.Pz - sD.z), h2q);
        wave_sync();
        // R^0_{tuv} by the McMurchie-Davidson recursion over the levels
        // n = L-1 .. 0 (level n holds the orders up to L-n), alternating
        // between two LDS buffers so that level 0 lands in Rh
#pragma unroll 1
        for (int lev = L - 1; lev >= 0; --lev) {
          // level L holds R^L_{000} = g_L only, which the order-1 entries of
          // level L-1 read (their second index is 0 with weight 0)
          const R* prev = (lev == L - 1) ? g + L : (((lev + 1) & 1) ? Rh2 : Rh);
          R* next = (lev & 1) ? Rh2 : Rh;
#pragma unroll
          for (int sl = 0; sl < W_::NSL; ++sl) {
            const int e = lane + kWT * sl;
            if (e < W_::NH && r_ord[sl] <= L - lev) {
              const int ax = r_ax[sl];
              const R pa = ax == 0 ? X : (ax == 1 ? Y : Z);
              next[e] = ax < 0 ? g[lev] : pa * prev[r_i1[sl]] + r_c2[sl] * prev[r_i2[sl]];
            }
          }
          wave_sync();
        }
        // W[cd][tuv]
        for (int e = lane; e < W_::NW; e += kWT) {
          const int cd = e / W_::NHAB, h = e - cd * W_::NHAB;
          const Cart c = cart(LC, cd / ND), d = cart(LD, cd % ND);
          const int hd = hdec[h], t0 = hd & 255, u0 = (hd >> 8) & 255, w0 = hd >> 16;
          const R* ex = Ec + 0 * W_::EC + (c.x * (LD + 1) + d.x) * (LC + LD + 1);
          const R* ey = Ec + 1 * W_::EC + (c.y * (LD + 1) + d.y) * (LC + LD + 1);
          const R* ez = Ec + 2 * W_::EC + (c.z * (LD + 1) + d.z) * (LC + LD + 1);
          R w = R(0);
          for (int T_ = 0; T_ <= c.x + d.x; ++T_)
            for (int U_ = 0; U_ <= c.y + d.y; ++U_) {
              R wz = R(0);
              for (int V_ = 0; V_ <= c.z + d.z; ++V_) {
                const R sg = ((T_ + U_ + V_) & 1) ? R(-1) : R(1);
                wz += sg * ez[V_] * Rh[hidx(t0 + T_, u0 + U_, w0 + V_)];
              }
              w += ex[T_] * ey[U_] * wz;
            }
          Wt[e] = w;
        }
        wave_sync();
        // (ab|cd) += pref sum_{tuv} E^{ab}_{tuv} W[cd][tuv]
#pragma unroll 1
        for (int o = lane; o < NINT; o += kWT) {
          {
            const int ab = o / NCD, cd = o - ab * NCD;
            const Cart a = cart(LA, ab / NB), b = cart(LB, ab % NB);
            const R* ex = Ea + 0 * W_::EA + (a.x * (LB + 1) + b.x) * (LAB + 1);
            const R* ey = Ea + 1 * W_::EA + (a.y * (LB + 1) + b.y) * (LAB + 1);
            const R* ez = Ea + 2 * W_::EA + (a.z * (LB + 1) + b.z) * (LAB + 1);
            const R* w = Wt + cd * W_::NHAB;
            R val = R(0);
            for (int t_ = 0; t_ <= a.x + b.x; ++t_)
              for (int u_ = 0; u_ <= a.y + b.y; ++u_) {
                R vz = R(0);
                for (int v_ = 0; v_ <= a.z + b.z; ++v_) vz += ez[v_] * w[hidx(t_, u_, v_)];
                val += ex[t_] * ey[u_] * vz;
              }
            acc[o] += pref * val;
          }
        }
        wave_sync();
      }
    }

    const int fa = sA.bf_off, fb = sB.bf_off, fc = sC.bf_off, fd = sD.bf_off;
    const int ga = sA.gbf, gb = sB.gbf, gc = sC.gbf, gd = sD.gbf;  // component scale indices
    if constexpr (MODE == kSchwarz) {
      for (int o = lane; o < NINT; o += kWT) {
        {
          const int ab = o / NCD, cd = o - ab * NCD;
          if (ab == cd) {
            const int ia = ab / NB, jb = ab % NB;
            const double sc = A.bscale[ga + ia] * A.bscale[gb + jb];
            const double v = sqrt(fabs((double)acc[o] * sc * sc));
            A.sch_out[(fa + ia) * n + fb + jb] = v;
            A.sch_out[(fb + jb) * n + fa + ia] = v;
          }
        }
      }
    } else if constexpr (MODE == kStore) {
      // the per-thread kernel's kStore rule, one entry per lane and step
      const size_t ld = (size_t)n * (n + 1) / 2;
      const bool sb = bp.a == bp.b, sk = kp.a == kp.b, diag = A.same && ib == ik;
      const bool whole = v0 == 0 && v1 == kp.npp;
      for (int o = lane; o < NINT; o += kWT) {
        const int ab = o / NCD, cd = o - ab * NCD;
        const int p = fa + ab / NB, q = fb + ab % NB, rr = fc + cd / ND, ss = fd + cd % ND;
        if ((sb && q > p) || (sk && ss > rr)) continue;
        const double v = (double)acc[o] * A.bscale[ga + ab / NB] * A.bscale[gb + ab % NB] * A.bscale[gc + cd / ND] *
                         A.bscale[gd + cd % ND];
        const size_t P = pair_index(p, q), Q = pair_index(rr, ss);
        eri_put(A.sch_out, P * ld + Q, v, whole);
        if (!diag) eri_put(A.sch_out, Q * ld + P, v, whole);
      }
      wave_sync();
    } else {
      double deg = 1.0;
      if (bp.a == bp.b) deg *= 0.5;
      if (kp.a == kp.b) deg *= 0.5;
      if (A.same && ib == ik) deg *= 0.5;
      double* EE = lds + W_::oW;
      for (int o = lane; o < NINT; o += kWT) {
        {
          const int ab = o / NCD, cd = o - ab * NCD;
          EE[o] = (double)acc[o] * deg * A.bscale[ga + ab / NB] * A.bscale[gb + ab % NB] *
                  A.bscale[gc + cd / ND] * A.bscale[gd + cd % ND];
        }
      }
      wave_sync();
      // Coulomb
      double* Dst = lds + W_::oD;
      for (int e = lane; e < NAB; e += kWT) Dst[e] = V.DJ[(fa + e / NB) * n + fb + e % NB];
      for (int e = lane; e < NCD; e += kWT) Dst[NAB + e] = V.DJ[(fc + e / ND) * n + fd + e % ND];
      wave_sync();
      for (int e = lane; e < NAB + NCD; e += kWT) {
        double s = 0.0;
        if (e < NAB) {
#pragma unroll 6
          for (int cd = 0; cd < NCD; ++cd) s += EE[e * NCD + cd] * Dst[NAB + cd];
          fatomic(&V.AJ[(fa + e / NB) * n + fb + e % NB], 2.0 * s);
        } else {
          const int cd = e - NAB;
#pragma unroll 6
          for (int ab = 0; ab < NAB; ++ab) s += EE[ab * NCD + cd] * Dst[ab];
          fatomic(&V.AJ[(fc + cd / ND) * n + fd + cd % ND], 2.0 * s);
        }
      }
      // Exchange, per spin
      constexpr int nspin = (MODE == kUHF) ? 2 : 1;
      for (int sp = 0; sp < nspin; ++sp) {
        const double* D = sp ? V.DB : V.DA;
        double* AK = sp ? V.AKB : V.AKA;
        double* dbd = Dst + NAB + NCD;        // D[b,d]  NB x ND
        double* dad = dbd + NB * ND;          // D[a,d]  NA x ND
        double* dbc = dad + NA * ND;          // D[b,c]  NB x NC
        double* dac = dbc + NB * NC;          // D[a,c]  NA x NC
        wave_sync();
        for (int e = lane; e < NB * ND; e += kWT) dbd[e] = D[(fb + e / ND) * n + fd + e % ND];
        for (int e = lane; e < NA * ND; e += kWT) dad[e] = D[(fa + e / ND) * n + fd + e % ND];
        for (int e = lane; e < NB * NC; e += kWT) dbc[e] = D[(fb + e / NC) * n + fc + e % NC];
        for (int e = lane; e < NA * NC; e += kWT) dac[e] = D[(fa + e / NC) * n + fc + e % NC];
        wave_sync();
        constexpr int NK = NA * NC + NB * NC + NA * ND + NB * ND;
        for (int e = lane; e < NK; e += kWT) {
          double s = 0.0;
          if (e < NA * NC) {                       // K[a,c] += sum_bd v D[b,d]
            const int a = e / NC, c = e % NC;
#pragma unroll 1
            for (int b = 0; b < NB; ++b)
              for (int d = 0; d < ND; ++d) s += EE[(a * NB + b) * NCD + c * ND + d] * dbd[b * ND + d];
            fatomic(&AK[(fa + a) * n + fc + c], s);
          } else if (e < NA * NC + NB * NC) {      // K[b,c] += sum_ad v D[a,d]
            const int r = e - NA * NC, b = r / NC, c = r % NC;
#pragma unroll 1
            for (int a = 0; a < NA; ++a)
              for (int d = 0; d < ND; ++d) s += EE[(a * NB + b) * NCD + c * ND + d] * dad[a * ND + d];
            fatomic(&AK[(fb + b) * n + fc + c], s);
          } else if (e < NA * NC + NB * NC + NA * ND) {  // K[a,d] += sum_bc v D[b,c]
            const int r = e - NA * NC - NB * NC, a = r / ND, d = r % ND;
#pragma unroll 1
            for (int b = 0; b < NB; ++b)
              for (int c = 0; c < NC; ++c) s += EE[(a * NB + b) * NCD + c * ND + d] * dbc[b * NC + c];
            fatomic(&AK[(fa + a) * n + fd + d], s);
          } else {                                 // K[b,d] += sum_ac v D[a,c]
            const int r = e - NA * NC - NB * NC - NA * ND, b = r / ND, d = r % ND;
#pragma unroll 1
            for (int a = 0; a < NA; ++a)
              for (int c = 0; c < NC; ++c) s += EE[(a * NB + b) * NCD + c * ND + d] * dac[a * NC + c];
            fatomic(&AK[(fb + b) * n + fd + d], s);
          }
        }
      }
      wave_sync();
    }
    }  // rows
  }
}

}  // namespace sqjk
