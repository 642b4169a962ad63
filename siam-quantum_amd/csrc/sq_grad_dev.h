// sq_grad_dev.h - two-electron part of the energy gradient over shell
// quartets (the work of src/grad.c:1734-2090, GradEri_ShellSet_Parallel).
//
// Work unit and work plan are the J/K build's (sq_jk_dev.h): one launch per
// class pair, tiles of bra rows with one first shell and 64 consecutive ket
// pairs, a lane per (quartet, ket-component slice).  Per quartet the lane
// forms the density weights of its integrals,
//   w_abcd = deg [4 PT_ab PT_cd - 2 k (PA_ac PA_bd + PA_ad PA_bc + PB_ac PB_bd + PB_ad PB_bc)],
// (the eight permutations of the reference's 0.5 PT PT - k/4 (PT PT + PS PS)
// summed; deg halves repeated pairs as in the J/K digestion), and contracts
// them with the derivative integrals primitive quartet by primitive quartet.
//
// McMurchie-Davidson derivatives, with E the Hermite expansion coefficients
// of a pair and R the Hermite Coulomb integrals:
//   d/dA_x (ab|cd) = sum (2 alpha E^{a+1x,b} - a_x E^{a-1x,b})_tuv W^{cd}_tuv
//   (d/dA + d/dB)_x (ab|cd) = sum E^{ab}_tuv W^{cd}_{t+1,u,v}     (shift of P)
//   d/dC_x (ab|cd) = sum E^{ab}_tuv sum (-1)^{T+U+V} (2 gamma E^{c+1x,d} - c_x E^{c-1x,d})_TUV R_{t+T,u+U,v+V}
// with W^{cd}_tuv = sum (-1)^{T+U+V} E^{cd}_TUV R_{t+T,u+U,v+V}; D follows
// from translational invariance, d/dD = -(d/dA + d/dB) - d/dC.  The bra
// side is first contracted with the weights into the Hermite density
// H_tuv = sum_ab w_ab E^{ab}_tuv, so the P and C terms cost one pass over
// the Hermite indices per ket component.
#pragma once
#include "sq_jk_wd.h"

namespace sqjk {

struct GradArgs {
  ClassArgs c;   // work plan, pairs, DJ = PA + PB, DA = PA, DB = PB, PM, cut-off
  double* gsub;  // [nsub][3] gradient per sub-shell centre (atomics)
  double kfac;   // exchange weight of the energy (1 for Hartree-Fock)
};

// Ket-component slices of the gradient kernel: a lane holds the weights of
// all NAB bra components for SZ ket components (SZ = 1 for bras of six or
// more components; at most nine weights for the small bras).
template <int LA, int LB, int LC, int LD>
struct GradSlice {
  static constexpr int NAB = ncart(LA) * ncart(LB);
  static constexpr int NCD = ncart(LC) * ncart(LD);
  static constexpr int PER = (9 / NAB) > 0 ? (9 / NAB) : 1;
  static constexpr int NSLICE = (NCD + PER - 1) / PER;
  static constexpr int SZ = (NCD + NSLICE - 1) / NSLICE;
};

// Gradient contributions of the contracted quartet (ib, ik), ket components
// [CD0, CD1), bra primitive pairs [0, nppb), ket primitive pairs [v0, v1):
// g[0..2] d/dA, g[3..5] d/dA + d/dB, g[6..8] d/dC, added to g.  All lanes of
// a wavefront call this together; `valid` gates the lane's weights.
template <int LA, int LB, int LC, int LD, int CD0, int CD1>
__host__ __device__ __forceinline__ void grad_quartet(const GradArgs& G, const View& V, int ib, int ik, int nppb,
                                                      int v0, int v1, bool valid, const double* lds,
                                                      double (&g)[9]) {
  const ClassArgs& A = G.c;
  constexpr int NA = ncart(LA), NB = ncart(LB), ND = ncart(LD);
  constexpr int LAB = LA + LB, LCD = LC + LD, L = LAB + LCD;
  constexpr int NAB = NA * NB, NS = CD1 - CD0;
  (void)NA;

  const PairRec bp = uload(A.bra + ib);
  const PairRec kp = A.ket[ik];
  const SubShell sA = uload(A.sh + bp.a), sB = uload(A.sh + bp.b), sC = A.sh[kp.a], sD = A.sh[kp.b];
  const int n = V.n;

  // density weights of the lane's integrals (component scales included)
  double w[NAB][NS];
  {
    double deg = valid ? 1.0 : 0.0;
    if (bp.a == bp.b) deg *= 0.5;
    if (kp.a == kp.b) deg *= 0.5;
    if (A.same && ib == ik) deg *= 0.5;
    const int fa = sA.bf_off, fb = sB.bf_off, fc = sC.bf_off, fd = sD.bf_off;
    const double k2 = 2.0 * G.kfac;
    sfor<NS>([&](auto csI) {
      constexpr int cs = decltype(csI)::value, cd = CD0 + cs, ic_ = cd / ND, id_ = cd % ND;
      const int pc = fc + ic_, pd = fd + id_;
      const double scd = deg * A.bscale[sC.gbf + ic_] * A.bscale[sD.gbf + id_];
      const double dcd = 4.0 * V.DJ[(size_t)pc * n + pd];
      sfor<NAB>([&](auto abI) {
        constexpr int ab = decltype(abI)::value, ia = ab / NB, jb = ab % NB;
        const size_t pa = fa + ia, pb = fb + jb;
        const double xa = V.DA[pa * n + pc] * V.DA[pb * n + pd] + V.DA[pa * n + pd] * V.DA[pb * n + pc];
        const double xb = V.DB[pa * n + pc] * V.DB[pb * n + pd] + V.DB[pa * n + pd] * V.DB[pb * n + pc];
        w[ab][cs] = scd * A.bscale[sA.gbf + ia] * A.bscale[sB.gbf + jb] *
                    (V.DJ[pa * n + pb] * dcd - k2 * (xa + xb));
      });
    });
  }

  for (int v = v0; v < v1; ++v) {
    const PrimPair Q = A.pp[kp.pp_off + v * kp.pp_stride];
    Herm1<LC + 1, LD, double> Fx, Fy, Fz;
    herm_coef<LC + 1, LD, double>(Fx, Q.Px - sC.x, Q.Px - sD.x, Q.hp);
    herm_coef<LC + 1, LD, double>(Fy, Q.Py - sC.y, Q.Py - sD.y, Q.hp);
    herm_coef<LC + 1, LD, double>(Fz, Q.Pz - sC.z, Q.Pz - sD.z, Q.hp);
    const double g2 = 2.0 * Q.ea;
    for (int u = 0; u < nppb; ++u) {
      const PrimPair P = uload(A.pp + (bp.pp_off + u * bp.pp_stride));
      const double s = P.p + Q.p;
      const double rs = rsqrt_nr(s);
      if (kTwoPi25 * P.eb * Q.eb * rs < kPrimCutoff) continue;
      const double alpha = P.p * Q.p * (rs * rs);
      const double X = P.Px - Q.Px, Y = P.Py - Q.Py, Z = P.Pz - Q.Pz;
      double F[L + 2];
      boys<L + 1, double>(alpha * (X * X + Y * Y + Z * Z), F, lds);
      double Rt[nherm(L + 1)];
      herm_R<L + 1, double>(Rt, F, alpha, X, Y, Z);
      Herm1<LA + 1, LB, double> Ex, Ey, Ez;
      herm_coef<LA + 1, LB, double>(Ex, P.Px - sA.x, P.Px - sB.x, P.hp);
      herm_coef<LA + 1, LB, double>(Ey, P.Py - sA.y, P.Py - sB.y, P.hp);
      herm_coef<LA + 1, LB, double>(Ez, P.Pz - sA.z, P.Pz - sB.z, P.hp);
      const double a2 = 2.0 * P.ea;
      double gq[9];
      sfor<9>([&](auto I) { gq[I.value] = 0.0; });

      sfor<NS>([&](auto csI) {
        constexpr int cs = decltype(csI)::value;
        constexpr int cd = CD0 + cs;
        constexpr Cart c = cart(LC, cd / ND), d = cart(LD, cd % ND);
        constexpr int kx = c.x + d.x, ky = c.y + d.y, kz = c.z + d.z;
        // W_tuv for t+u+v <= LAB+1 (the P shift needs one order above the bra's)
        double W[LAB + 2][LAB + 2][LAB + 2];
        sfor<LAB + 2>([&](auto tI) {
          constexpr int t = decltype(tI)::value;
          sfor<LAB + 2 - t>([&](auto uI) {
            constexpr int uu = decltype(uI)::value;
            sfor<LAB + 2 - t - uu>([&](auto vI) {
              constexpr int vv = decltype(vI)::value;
              double acc = 0.0;
              sfor<kx + 1>([&](auto TI) {
                constexpr int T_ = decltype(TI)::value;
                sfor<ky + 1>([&](auto UI) {
                  constexpr int U_ = decltype(UI)::value;
                  double wz = 0.0;
                  sfor<kz + 1>([&](auto VI) {
                    constexpr int V_ = decltype(VI)::value;
                    constexpr double sg = ((T_ + U_ + V_) & 1) ? -1.0 : 1.0;
                    wz += sg * Fz.e[c.z][d.z][V_] * Rt[hidx(t + T_, uu + U_, vv + V_)];
                  });
                  acc += Fx.e[c.x][d.x][T_] * Fy.e[c.y][d.y][U_] * wz;
                });
              });
              W[t][uu][vv] = acc;
            });
          });
        });
        // bra: Hermite density H (t+u+v <= LAB) and the A derivative
        double H[LAB + 1][LAB + 1][LAB + 1];
        sfor<LAB + 1>([&](auto tI) {
          sfor<LAB + 1 - tI.value>([&](auto uI) {
            sfor<LAB + 1 - tI.value - uI.value>([&](auto vI) { H[tI.value][uI.value][vI.value] = 0.0; });
          });
        });
        sfor<NAB>([&](auto abI) {
          constexpr int ab = decltype(abI)::value;
          constexpr Cart a = cart(LA, ab / NB), b = cart(LB, ab % NB);
          const double wab = w[ab][cs];
          // H += w E^{ab}
          sfor<a.x + b.x + 1>([&](auto tI) {
            constexpr int t = decltype(tI)::value;
            sfor<a.y + b.y + 1>([&](auto uI) {
              constexpr int uu = decltype(uI)::value;
              const double exy = wab * Ex.e[a.x][b.x][t] * Ey.e[a.y][b.y][uu];
              sfor<a.z + b.z + 1>([&](auto vI) {
                constexpr int vv = decltype(vI)::value;
                H[t][uu][vv] += exy * Ez.e[a.z][b.z][vv];
              });
            });
          });
          // d/dA along each axis: the bra coefficients of that axis raised and lowered
          auto shifted = [&](const auto& E, auto ai, auto bi, auto ti) -> double {
            constexpr int i = decltype(ai)::value, j = decltype(bi)::value, t = decltype(ti)::value;
            double e = a2 * E.e[i + 1][j][t];
            if constexpr (i > 0 && t <= i - 1 + j) e -= double(i) * E.e[i - 1][j][t];
            return e;
          };
          double dx = 0.0, dy = 0.0, dz = 0.0;
          sfor<a.x + b.x + 2>([&](auto tI) {
            constexpr int t = decltype(tI)::value;
            const double ex = shifted(Ex, ic<a.x>{}, ic<b.x>{}, tI);
            sfor<a.y + b.y + 1>([&](auto uI) {
              constexpr int uu = decltype(uI)::value;
              double sz = 0.0;
              sfor<a.z + b.z + 1>([&](auto vI) {
                constexpr int vv = decltype(vI)::value;
                sz += Ez.e[a.z][b.z][vv] * W[t][uu][vv];
              });
              dx += ex * Ey.e[a.y][b.y][uu] * sz;
            });
          });
          sfor<a.y + b.y + 2>([&](auto uI) {
            constexpr int uu = decltype(uI)::value;
            const double ey = shifted(Ey, ic<a.y>{}, ic<b.y>{}, uI);
            sfor<a.x + b.x + 1>([&](auto tI) {
              constexpr int t = decltype(tI)::value;
              double sz = 0.0;
              sfor<a.z + b.z + 1>([&](auto vI) {
                constexpr int vv = decltype(vI)::value;
                sz += Ez.e[a.z][b.z][vv] * W[t][uu][vv];
              });
              dy += Ex.e[a.x][b.x][t] * ey * sz;
            });
          });
          sfor<a.z + b.z + 2>([&](auto vI) {
            constexpr int vv = decltype(vI)::value;
            const double ez = shifted(Ez, ic<a.z>{}, ic<b.z>{}, vI);
            sfor<a.x + b.x + 1>([&](auto tI) {
              constexpr int t = decltype(tI)::value;
              double sy = 0.0;
              sfor<a.y + b.y + 1>([&](auto uI) {
                constexpr int uu = decltype(uI)::value;
                sy += Ey.e[a.y][b.y][uu] * W[t][uu][vv];
              });
              dz += Ex.e[a.x][b.x][t] * sy * ez;
            });
          });
          gq[0] += wab * dx;
          gq[1] += wab * dy;
          gq[2] += wab * dz;
        });
        // P shift and the C derivative, through H
        sfor<LAB + 1>([&](auto tI) {
          constexpr int t = decltype(tI)::value;
          sfor<LAB + 1 - t>([&](auto uI) {
            constexpr int uu = decltype(uI)::value;
            sfor<LAB + 1 - t - uu>([&](auto vI) {
              constexpr int vv = decltype(vI)::value;
              const double h = H[t][uu][vv];
              gq[3] += h * W[t + 1][uu][vv];
              gq[4] += h * W[t][uu + 1][vv];
              gq[5] += h * W[t][uu][vv + 1];
              // ket coefficients with the c power raised and lowered along one axis
              auto kshift = [&](const auto& E, auto ci, auto di, auto Ti) -> double {
                constexpr int i = decltype(ci)::value, j = decltype(di)::value, T_ = decltype(Ti)::value;
                double e = g2 * E.e[i + 1][j][T_];
                if constexpr (i > 0 && T_ <= i - 1 + j) e -= double(i) * E.e[i - 1][j][T_];
                return e;
              };
              double cx = 0.0, cy = 0.0, cz = 0.0;
              sfor<kx + 2>([&](auto TI) {
                constexpr int T_ = decltype(TI)::value;
                sfor<ky + 2>([&](auto UI) {
                  constexpr int U_ = decltype(UI)::value;
                  sfor<kz + 2>([&](auto VI) {
                    constexpr int V_ = decltype(VI)::value;
                    constexpr double sg = ((T_ + U_ + V_) & 1) ? -1.0 : 1.0;
                    constexpr int nx = (T_ <= kx) + (U_ <= ky) + (V_ <= kz);
                    if constexpr (nx >= 2) {
                      const double r = sg * Rt[hidx(t + T_, uu + U_, vv + V_)];
                      if constexpr (U_ <= ky && V_ <= kz)
                        cx += kshift(Fx, ic<c.x>{}, ic<d.x>{}, TI) * Fy.e[c.y][d.y][U_] * Fz.e[c.z][d.z][V_] * r;
                      if constexpr (T_ <= kx && V_ <= kz)
                        cy += Fx.e[c.x][d.x][T_] * kshift(Fy, ic<c.y>{}, ic<d.y>{}, UI) * Fz.e[c.z][d.z][V_] * r;
                      if constexpr (T_ <= kx && U_ <= ky)
                        cz += Fx.e[c.x][d.x][T_] * Fy.e[c.y][d.y][U_] * kshift(Fz, ic<c.z>{}, ic<d.z>{}, VI) * r;
                    }
                  });
                });
              });
              gq[6] += h * cx;
              gq[7] += h * cy;
              gq[8] += h * cz;
            });
          });
        });
      });
      const double pref = kTwoPi25 * P.cc * Q.cc * rs;
      sfor<9>([&](auto I) { g[I.value] += pref * gq[I.value]; });
    }
  }
}

// One launch per class pair; blockIdx.y selects the ket-component slice.
// Grid-stride over wavefronts, a wavefront per tile (as the J/K kernel).
// Per tile a lane keeps the gradient of its ket pair's two centres and of
// the run's first bra shell in registers; the second bra shell changes per
// row and is reduced over the wavefront row by row.
template <int LA, int LB, int LC, int LD, int BS>
__global__ void __launch_bounds__(BS) k_grad(GradArgs G) {
  using SI = GradSlice<LA, LB, LC, LD>;
  constexpr int L = LA + LB + LC + LD;
  extern __shared__ double lds[];
  const ClassArgs& A = G.c;
  boys_load_lds<L + 1>(lds, A.boys);
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
    for (long long wv = (long long)blockIdx.x * wpb + wib; wv < A.total; wv += wstride) {
      const Tile tl = uload(A.tile + wv);
      const Segment sg = uload(A.seg + tl.seg);
      const View V = make_view(A, sg.inst);
      const int j = tl.k0 + lane;  // ket position in the group
      double ga[3] = {0.0, 0.0, 0.0}, gc[3] = {0.0, 0.0, 0.0}, gd[3] = {0.0, 0.0, 0.0};
      bool any = false;
      int sa = -1;
      for (int r = tl.r0; r < tl.r1; ++r) {
        const int ib = uload(A.rbra + r);
        const int lo = uload(A.rlo + r);
        const bool in = j >= lo && j < uload(A.rhi + r);
        const int ik = sg.ket0 + (in ? j : lo);  // idle lanes evaluate a valid pair, weight 0
        const PairRec bp = uload(A.bra + ib);
        sa = bp.a;
        any = any || in;
        const bool valid = in && density_screen_pass(V, bp, A.ket[ik]);
        if (!__any(valid)) continue;  // uniform: the row's bra pair is the wavefront's
        double g[9];
        sfor<9>([&](auto I) { g[I.value] = 0.0; });
        grad_quartet<LA, LB, LC, LD, CD0, CD1>(G, V, ib, ik, sg.nppb, tl.v0, tl.v1, valid, lds, g);
        sfor<3>([&](auto I) {
          constexpr int x = decltype(I)::value;
          ga[x] += g[x];
          gc[x] += g[6 + x];
          gd[x] -= g[3 + x] + g[6 + x];
          const double gb = wave_sum(g[3 + x] - g[x]);
          if (lane == 0) fatomic(&G.gsub[3 * bp.b + x], gb);
        });
      }
      if (sa >= 0) {
        sfor<3>([&](auto I) {
          constexpr int x = decltype(I)::value;
          const double s = wave_sum(ga[x]);
          if (lane == 0) fatomic(&G.gsub[3 * sa + x], s);
        });
      }
      if (any) {
        const PairRec kp = A.ket[sg.ket0 + j];
        sfor<3>([&](auto I) {
          constexpr int x = decltype(I)::value;
          fatomic(&G.gsub[3 * kp.a + x], gc[x]);
          fatomic(&G.gsub[3 * kp.b + x], gd[x]);
        });
      }
    }
  });
}

// ---------------------------------------------------------------------------
// High-angular-momentum classes: one wavefront per contracted quartet, lanes
// over components (the structure of the (dd|dd) J/K kernel).  Per primitive
// quartet, in the wavefront's LDS slice:
//   R_tuv (orders <= L+1) by level recursion; the bra and ket Hermite tables
//   with one extra order on a and c (one axis per lane);
//   H[cd][tuv] = sum_ab w_abcd E^{ab}_tuv                         lane = cd;
//   W[cd][tuv] = sum (-1)^|T| E^{cd}_T R_{tuv+T}, |tuv| <= LAB+1,  and the C
//   derivative sum_cd H[cd][tuv] sum (-1)^|T| Ec^{cd}_T R_{tuv+T}  lane = tuv;
//   the A derivative sum_ab w_abcd sum Ea^{ab}_tuv W[cd][tuv] and the P shift
//   sum_tuv H[cd][tuv] W[cd][tuv + e]                              lane = cd;
// (Ea, Ec: coefficients with a, c raised and lowered along one axis).  The
// nine sums are reduced over the wavefront once per quartet.
constexpr int kGWaves = 2;  // wavefronts per workgroup

// (t, u, v) of compact Hermite index i, at compile time
__host__ __device__ constexpr int kHermT(int i) { return make_herm_tab().t[i]; }
__host__ __device__ constexpr int kHermU(int i) { return make_herm_tab().u[i]; }
__host__ __device__ constexpr int kHermV(int i) { return make_herm_tab().v[i]; }

template <int LA, int LB, int LC, int LD>
struct GWLayout {
  static constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  static constexpr int LAB = LA + LB, LCD = LC + LD, L = LAB + LCD, LT = L + 1;
  static constexpr int NAB = NA * NB, NCD = NC * ND;
  static constexpr int NH = nherm(LT), NHAB = nherm(LAB), NHAB1 = nherm(LAB + 1), NHCD1 = nherm(LCD + 1);
  static constexpr int EB = (LA + 2) * (LB + 1) * (LAB + 2), EK = (LC + 2) * (LD + 1) * (LCD + 2);
  // per-wavefront LDS (doubles)
  static constexpr int oG = 0, oR0 = LT + 1, oR1 = oR0 + NH, oEb = oR1 + NH, oEk = oEb + 3 * EB;
  static constexpr int oH = oEk + 3 * EK, oW = oH + NCD * NHAB, PER_WAVE = oW + NCD * NHAB1;
  // workgroup-shared ints: R recursion step per entry (4 ints), gather
  // offsets [T][lane] of R_{tuv+T} for the lane's bra Hermite index tuv
  static constexpr int oStep = 0, oRoff = 4 * NH, SHARED_INTS = oRoff + NHCD1 * 64;
  static_assert(NHAB1 <= 64 && NCD <= 64, "one lane per component");
  static_assert(LT <= 9, "Hermite index table and Boys cut-offs cover L <= 9");
  __host__ __device__ static constexpr int eb(int ax, int i, int j, int t) {
    return oEb + ax * EB + (i * (LB + 1) + j) * (LAB + 2) + t;
  }
  __host__ __device__ static constexpr int ek(int ax, int i, int j, int t) {
    return oEk + ax * EK + (i * (LD + 1) + j) * (LCD + 2) + t;
  }
};

template <int LA, int LB, int LC, int LD>
__global__ void __launch_bounds__(64 * kGWaves) k_grad_wave(GradArgs G) {
  using W_ = GWLayout<LA, LB, LC, LD>;
  constexpr int NB = W_::NB, ND = W_::ND, NAB = W_::NAB, NCD = W_::NCD;
  constexpr int LAB = W_::LAB, LCD = W_::LCD, LT = W_::LT;
  constexpr int NH = W_::NH, NHAB = W_::NHAB, NHAB1 = W_::NHAB1, NHCD1 = W_::NHCD1;
  const ClassArgs& A = G.c;
  __shared__ double lds_all[kGWaves * W_::PER_WAVE];
  __shared__ int tabs[W_::SHARED_INTS];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  double* lds = lds_all + wave * W_::PER_WAVE;
  double* g = lds + W_::oG;
  int* step = tabs + W_::oStep;
  int* roff = tabs + W_::oRoff;
  for (int e = threadIdx.x; e < NH; e += blockDim.x) {
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
  for (int e = threadIdx.x; e < NHCD1 * 64; e += blockDim.x) {
    const int T = e >> 6, l = e & 63;
    const int h = l < NHAB1 ? l : 0;
    const int t = kHerm.t[h] + kHerm.t[T], u = kHerm.u[h] + kHerm.u[T], v = kHerm.v[h] + kHerm.v[T];
    roff[e] = (t + u + v <= LT) ? hidx(t, u, v) : 0;  // orders above LT are never read
  }
  __syncthreads();  // the only workgroup barrier: the wavefronts run independently from here

  const long long items = A.total * 64;
  for (long long it = (long long)blockIdx.x * kGWaves + wave; it < items; it += (long long)gridDim.x * kGWaves) {
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
      // density weights of column cd = lane
      double wcol[NAB];
      {
        double deg = lane < NCD ? 1.0 : 0.0;
        if (bp.a == bp.b) deg *= 0.5;
        if (kp.a == kp.b) deg *= 0.5;
        if (A.same && ib == ik) deg *= 0.5;
        const int cl = lane < NCD ? lane : 0;
        const int pc = sC.bf_off + cl / ND, pd = sD.bf_off + cl % ND;
        const double scd = deg * A.bscale[sC.gbf + cl / ND] * A.bscale[sD.gbf + cl % ND];
        const double dcd = 4.0 * V.DJ[(size_t)pc * n + pd];
        const double k2 = 2.0 * G.kfac;
        sfor<NAB>([&](auto abI) {
          constexpr int ab = decltype(abI)::value, ia = ab / NB, jb = ab % NB;
          const size_t pa = sA.bf_off + ia, pb = sB.bf_off + jb;
          const double xa = V.DA[pa * n + pc] * V.DA[pb * n + pd] + V.DA[pa * n + pd] * V.DA[pb * n + pc];
          const double xb = V.DB[pa * n + pc] * V.DB[pb * n + pd] + V.DB[pa * n + pd] * V.DB[pb * n + pc];
          wcol[ab] = scd * A.bscale[sA.gbf + ia] * A.bscale[sB.gbf + jb] * (V.DJ[pa * n + pb] * dcd - k2 * (xa + xb));
        });
      }
      double acc[9];
      sfor<9>([&](auto I) { acc[I.value] = 0.0; });

      for (int u = 0; u < nppb; ++u) {
        const PrimPair P = uload(A.pp + (bp.pp_off + u * bp.pp_stride));
        for (int v = v0; v < v1; ++v) {
          const PrimPair Q = uload(A.pp + (kp.pp_off + v * kp.pp_stride));
          const double s = P.p + Q.p;
          const double rs = rsqrt_nr(s);
          if (kTwoPi25 * P.eb * Q.eb * rs < kPrimCutoff) continue;  // uniform
          const double alpha = P.p * Q.p * (rs * rs);
          const double X = P.Px - Q.Px, Y = P.Py - Q.Py, Z = P.Pz - Q.Pz;
          const double T = alpha * (X * X + Y * Y + Z * Z);
          const double pref = kTwoPi25 * P.cc * Q.cc * rs;
          const double a2 = 2.0 * P.ea, g2 = 2.0 * Q.ea;
          // g_m = (-2 alpha)^m F_m(T); the six Hermite tables, one axis per lane
          if (lane <= LT) {
            double pw = 1.0;
            for (int m = 0; m < lane; ++m) pw *= -2.0 * alpha;
            g[lane] = pw * boys_one(lane, T, A.boys);
          } else if (lane >= 16 && lane < 19) {
            const int ax = lane - 16;
            const double pc_ = ax == 0 ? P.Px : (ax == 1 ? P.Py : P.Pz);
            const double ac = ax == 0 ? sA.x : (ax == 1 ? sA.y : sA.z), bc = ax == 0 ? sB.x : (ax == 1 ? sB.y : sB.z);
            herm_axis_to_lds<LA + 1, LB, double>(lds + W_::oEb + ax * W_::EB, pc_ - ac, pc_ - bc, P.hp);
          } else if (lane >= 32 && lane < 35) {
            const int ax = lane - 32;
            const double qc = ax == 0 ? Q.Px : (ax == 1 ? Q.Py : Q.Pz);
            const double cc_ = ax == 0 ? sC.x : (ax == 1 ? sC.y : sC.z), dc = ax == 0 ? sD.x : (ax == 1 ? sD.y : sD.z);
            herm_axis_to_lds<LC + 1, LD, double>(lds + W_::oEk + ax * W_::EK, qc - cc_, qc - dc, Q.hp);
          }
          wsync();
          // R^0_tuv over the levels LT-1 .. 0, alternating buffers so that level 0 lands in R0
#pragma unroll 1
          for (int lev = LT - 1; lev >= 0; --lev) {
            const double* prev = (lev == LT - 1) ? g + LT : (((lev + 1) & 1) ? lds + W_::oR1 : lds + W_::oR0);
            double* next = (lev & 1) ? lds + W_::oR1 : lds + W_::oR0;
            for (int e = lane; e < NH; e += 64) {
              const int s0 = step[4 * e];
              if ((s0 >> 4) <= LT - lev) {
                const int ax = (s0 & 15) - 1;
                const double pa = ax == 0 ? X : (ax == 1 ? Y : Z);
                next[e] = ax < 0 ? g[lev] : pa * prev[step[4 * e + 1]] + double(step[4 * e + 3]) * prev[step[4 * e + 2]];
              }
            }
            wsync();
          }
          const double* R0 = lds + W_::oR0;
          // H[cd][tuv] = sum_ab w E^{ab}_tuv, lane = cd
          if (lane < NCD) {
            double h[NHAB];
            sfor<NHAB>([&](auto I) { h[I.value] = 0.0; });
            sfor<NAB>([&](auto abI) {
              constexpr int ab = decltype(abI)::value;
              constexpr Cart a = cart(LA, ab / NB), b = cart(LB, ab % NB);
              const double wab = wcol[ab];
              sfor<a.x + b.x + 1>([&](auto tI) {
                sfor<a.y + b.y + 1>([&](auto uI) {
                  const double exy = wab * lds[W_::eb(0, a.x, b.x, tI.value)] * lds[W_::eb(1, a.y, b.y, uI.value)];
                  sfor<a.z + b.z + 1>([&](auto vI) {
                    h[hidx(tI.value, uI.value, vI.value)] += exy * lds[W_::eb(2, a.z, b.z, vI.value)];
                  });
                });
              });
            });
            sfor<NHAB>([&](auto I) { lds[W_::oH + lane * NHAB + I.value] = h[I.value]; });
          }
          wsync();
          // W and the C derivative, lane = bra Hermite index
          if (lane < NHAB1) {
            const bool lowr = lane < NHAB;  // order <= LAB: the C derivative too
            double rg[NHCD1];
            sfor<NHCD1>([&](auto TI) { rg[TI.value] = R0[roff[TI.value * 64 + lane]]; });
            double cx = 0.0, cy = 0.0, cz = 0.0;
            sfor<NCD>([&](auto cdI) {
              __builtin_amdgcn_sched_barrier(0);
              constexpr int cd = decltype(cdI)::value;
              constexpr Cart c = cart(LC, cd / ND), d = cart(LD, cd % ND);
              constexpr int kx = c.x + d.x, ky = c.y + d.y, kz = c.z + d.z;
              double w = 0.0;
              sfor<kx + 1>([&](auto TI) {
                sfor<ky + 1>([&](auto UI) {
                  double wz = 0.0;
                  sfor<kz + 1>([&](auto VI) {
                    constexpr double sgn = ((TI.value + UI.value + VI.value) & 1) ? -1.0 : 1.0;
                    wz += sgn * lds[W_::ek(2, c.z, d.z, VI.value)] * rg[hidx(TI.value, UI.value, VI.value)];
                  });
                  w += lds[W_::ek(0, c.x, d.x, TI.value)] * lds[W_::ek(1, c.y, d.y, UI.value)] * wz;
                });
              });
              lds[W_::oW + cd * NHAB1 + lane] = w;
              if (lowr) {
                // coefficients with c raised and lowered along one axis
                auto ksh = [&](auto axI, auto ci, auto di, auto Ti) -> double {
                  constexpr int ax = decltype(axI)::value, i = decltype(ci)::value, jj = decltype(di)::value,
                                t = decltype(Ti)::value;
                  double e = g2 * lds[W_::ek(ax, i + 1, jj, t)];
                  if constexpr (i > 0) e -= double(i) * lds[W_::ek(ax, i - 1, jj, t)];
                  return e;
                };
                double sx = 0.0, sy = 0.0, sz = 0.0;
                sfor<kx + 2>([&](auto TI) {
                  sfor<ky + 2>([&](auto UI) {
                    sfor<kz + 2>([&](auto VI) {
                      constexpr int T_ = TI.value, U_ = UI.value, V_ = VI.value;
                      constexpr double sgn = ((T_ + U_ + V_) & 1) ? -1.0 : 1.0;
                      if constexpr (U_ <= ky && V_ <= kz)
                        sx += sgn * ksh(ic<0>{}, ic<c.x>{}, ic<d.x>{}, TI) * lds[W_::ek(1, c.y, d.y, U_)] *
                              lds[W_::ek(2, c.z, d.z, V_)] * rg[hidx(T_, U_, V_)];
                      if constexpr (T_ <= kx && V_ <= kz)
                        sy += sgn * lds[W_::ek(0, c.x, d.x, T_)] * ksh(ic<1>{}, ic<c.y>{}, ic<d.y>{}, UI) *
                              lds[W_::ek(2, c.z, d.z, V_)] * rg[hidx(T_, U_, V_)];
                      if constexpr (T_ <= kx && U_ <= ky)
                        sz += sgn * lds[W_::ek(0, c.x, d.x, T_)] * lds[W_::ek(1, c.y, d.y, U_)] *
                              ksh(ic<2>{}, ic<c.z>{}, ic<d.z>{}, VI) * rg[hidx(T_, U_, V_)];
                    });
                  });
                });
                const double hv = lds[W_::oH + cd * NHAB + lane];
                cx += hv * sx;
                cy += hv * sy;
                cz += hv * sz;
              }
            });
            acc[6] += pref * cx;
            acc[7] += pref * cy;
            acc[8] += pref * cz;
          }
          wsync();
          // A derivative and P shift, lane = cd
          if (lane < NCD) {
            double wr[NHAB1], hr[NHAB];
            sfor<NHAB1>([&](auto I) { wr[I.value] = lds[W_::oW + lane * NHAB1 + I.value]; });
            sfor<NHAB>([&](auto I) { hr[I.value] = lds[W_::oH + lane * NHAB + I.value]; });
            double px = 0.0, py = 0.0, pz = 0.0;
            sfor<NHAB>([&](auto I) {
              constexpr int t = kHermT(I.value), u = kHermU(I.value), v = kHermV(I.value);
              px += hr[I.value] * wr[hidx(t + 1, u, v)];
              py += hr[I.value] * wr[hidx(t, u + 1, v)];
              pz += hr[I.value] * wr[hidx(t, u, v + 1)];
            });
            double ax_ = 0.0, ay_ = 0.0, az_ = 0.0;
            sfor<NAB>([&](auto abI) {
              __builtin_amdgcn_sched_barrier(0);
              constexpr int ab = decltype(abI)::value;
              constexpr Cart a = cart(LA, ab / NB), b = cart(LB, ab % NB);
              auto bsh = [&](auto axI, auto ai, auto bi, auto ti) -> double {
                constexpr int ax = decltype(axI)::value, i = decltype(ai)::value, jj = decltype(bi)::value,
                              t = decltype(ti)::value;
                double e = a2 * lds[W_::eb(ax, i + 1, jj, t)];
                if constexpr (i > 0) e -= double(i) * lds[W_::eb(ax, i - 1, jj, t)];
                return e;
              };
              double dx = 0.0, dy = 0.0, dz = 0.0;
              sfor<a.x + b.x + 2>([&](auto tI) {
                sfor<a.y + b.y + 2>([&](auto uI) {
                  sfor<a.z + b.z + 2>([&](auto vI) {
                    constexpr int t = tI.value, u = uI.value, v = vI.value;
                    if constexpr (u <= a.y + b.y && v <= a.z + b.z)
                      dx += bsh(ic<0>{}, ic<a.x>{}, ic<b.x>{}, tI) * lds[W_::eb(1, a.y, b.y, u)] *
                            lds[W_::eb(2, a.z, b.z, v)] * wr[hidx(t, u, v)];
                    if constexpr (t <= a.x + b.x && v <= a.z + b.z)
                      dy += lds[W_::eb(0, a.x, b.x, t)] * bsh(ic<1>{}, ic<a.y>{}, ic<b.y>{}, uI) *
                            lds[W_::eb(2, a.z, b.z, v)] * wr[hidx(t, u, v)];
                    if constexpr (t <= a.x + b.x && u <= a.y + b.y)
                      dz += lds[W_::eb(0, a.x, b.x, t)] * lds[W_::eb(1, a.y, b.y, u)] *
                            bsh(ic<2>{}, ic<a.z>{}, ic<b.z>{}, vI) * wr[hidx(t, u, v)];
                  });
                });
              });
              ax_ += wcol[ab] * dx;
              ay_ += wcol[ab] * dy;
              az_ += wcol[ab] * dz;
            });
            acc[0] += pref * ax_;
            acc[1] += pref * ay_;
            acc[2] += pref * az_;
            acc[3] += pref * px;
            acc[4] += pref * py;
            acc[5] += pref * pz;
          }
          wsync();
        }
      }
      // d/dA, d/dA + d/dB, d/dC of the quartet, summed over the wavefront
      double sum[9];
      sfor<9>([&](auto I) { sum[I.value] = wave_sum(acc[I.value]); });
      if (lane == 0) {
        sfor<3>([&](auto I) {
          constexpr int x = decltype(I)::value;
          fatomic(&G.gsub[3 * bp.a + x], sum[x]);
          fatomic(&G.gsub[3 * bp.b + x], sum[3 + x] - sum[x]);
          fatomic(&G.gsub[3 * kp.a + x], sum[6 + x]);
          fatomic(&G.gsub[3 * kp.b + x], -sum[3 + x] - sum[6 + x]);
        });
      }
    }  // rows
  }
}

template <int LA, int LB, int LC, int LD>
constexpr int grad_block() {
  return 256;
}

// Classes of total angular momentum 7 and above take the wavefront kernel:
// the per-thread kernel's registers (R up to L+1, both pairs' raised Hermite
// tables) no longer fit.  Classes of L = 5 and 6 have both kernels, chosen
// at run time (SQJK_GRAD_KERNEL=thread or wave; by default the per-thread
// kernel, which measures 2-4x faster there despite its register spills: the
// wavefront kernel keeps only a few lanes busy for classes with few ket
// components).
template <int LA, int LB, int LC, int LD>
constexpr bool grad_wave() {
  return LA + LB + LC + LD >= 7;
}
template <int LA, int LB, int LC, int LD>
constexpr bool grad_both() {
  return LA + LB + LC + LD == 5 || LA + LB + LC + LD == 6;
}
template <int LA, int LB, int LC, int LD>
constexpr bool grad_thread_default() {
  return LA + LB + LC + LD <= 6;
}

template <int LA, int LB, int LC, int LD>
constexpr size_t grad_lds_bytes() {
  return sizeof(double) * boys_rows<LA + LB + LC + LD + 1>() * kBoysLdsW;
}

}  // namespace sqjk
