// sq_one_e.hip - one-electron matrices T, V, S on the device (the loops of
// src/uhf.c:1307-1334 over GTO_kinetic, GTO_nuclei and GTO_overlap,
// src/matrix.c:35-164 with src/int.c:100-121, 262-291, 370-411).
//
// One lane owns one unique sub-shell pair (la >= lb) of a pair class and
// writes its NA x NB blocks of T, V and S (and the transposed blocks), so no
// two lanes write the same element.  McMurchie-Davidson throughout:
//   1-D overlap    s_ij = E^{ij}_0 sqrt(pi/p)
//   1-D kinetic    t_ij = b(2j+1) s_ij - 2 b^2 s_{i,j+2} - j(j-1)/2 s_{i,j-2}
//   nuclear        -Z_C (2 pi/p) K_ab sum_tuv E_t E_u E_v R_tuv(p, P - C)
// with the primitive prefactor cut of the reference's nai (2 pi/p K_ab below
// 1e-15 drops the primitive pair for every nucleus, src/int.c:392-393).
#include <hip/hip_runtime.h>

#include "sq_one_e.h"

namespace sqjk {

namespace {

constexpr double kPi = 3.14159265358979323846;

template <int LA, int LB>
__global__ void __launch_bounds__(256) k_one_e(OneEArgs A) {
  constexpr int NA = ncart(LA), NB = ncart(LB), L = LA + LB;
  extern __shared__ double lds[];
  boys_load_lds<L>(lds, A.boys);
  __syncthreads();
  const int n = A.n;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < A.np; t += gridDim.x * blockDim.x) {
    const PairRec pr = A.pr[t];
    const SubShell sa = A.sh[pr.a], sb = A.sh[pr.b];
    const double abx = sa.x - sb.x, aby = sa.y - sb.y, abz = sa.z - sb.z;
    const double ab2 = abx * abx + aby * aby + abz * abz;
    double st[NA][NB], tt[NA][NB], vt[NA][NB];
    sfor<NA>([&](auto I) { sfor<NB>([&](auto J) { st[I.value][J.value] = tt[I.value][J.value] = vt[I.value][J.value] = 0.0; }); });
    for (int ka = 0; ka < sa.nprim; ++ka)
      for (int kb = 0; kb < sb.nprim; ++kb) {
        const double al = A.pexp[sa.prim_off + ka], be = A.pexp[sb.prim_off + kb];
        const double p = al + be, rp = 1.0 / p;
        const double K = exp(-al * be * rp * ab2);
        const double cc = A.pcoef[sa.prim_off + ka] * A.pcoef[sb.prim_off + kb];
        const double px = (al * sa.x + be * sb.x) * rp, py = (al * sa.y + be * sb.y) * rp,
                     pz = (al * sa.z + be * sb.z) * rp;
        const double h2p = 0.5 * rp;
        // overlap and kinetic: E^{ij}_0 up to j = lb + 2
        Herm1<LA, LB + 2> Ox, Oy, Oz;
        herm_coef<LA, LB + 2>(Ox, px - sa.x, px - sb.x, h2p);
        herm_coef<LA, LB + 2>(Oy, py - sa.y, py - sb.y, h2p);
        herm_coef<LA, LB + 2>(Oz, pz - sa.z, pz - sb.z, h2p);
        const double so = cc * K * kPi * rp * sqrt(kPi * rp);  // c c' K (pi/p)^{3/2}
        sfor<NA>([&](auto I) {
          constexpr Cart a = cart(LA, I.value);
          sfor<NB>([&](auto J) {
            constexpr Cart b = cart(LB, J.value);
            const double sx = Ox.e[a.x][b.x][0], sy = Oy.e[a.y][b.y][0], sz = Oz.e[a.z][b.z][0];
            auto kin = [&](const auto& E, auto ii, auto jj) {
              constexpr int i = decltype(ii)::value, j = decltype(jj)::value;
              double v = be * (2 * j + 1) * E.e[i][j][0] - 2.0 * be * be * E.e[i][j + 2][0];
              if constexpr (j >= 2) v -= 0.5 * j * (j - 1) * E.e[i][j - 2][0];
              return v;
            };
            const double tx = kin(Ox, ic<a.x>{}, ic<b.x>{}), ty = kin(Oy, ic<a.y>{}, ic<b.y>{}),
                         tz = kin(Oz, ic<a.z>{}, ic<b.z>{});
            st[I.value][J.value] += so * sx * sy * sz;
            tt[I.value][J.value] += so * (tx * sy * sz + sx * ty * sz + sx * sy * tz);
          });
        });
        // nuclear attraction
        const double pre = 2.0 * kPi * rp * K;
        if (pre < kPrimCutoff) continue;
        Herm1<LA, LB> Ex, Ey, Ez;
        herm_coef<LA, LB>(Ex, px - sa.x, px - sb.x, h2p);
        herm_coef<LA, LB>(Ey, py - sa.y, py - sb.y, h2p);
        herm_coef<LA, LB>(Ez, pz - sa.z, pz - sb.z, h2p);
        for (int c = 0; c < A.natom; ++c) {
          const double* at = A.atom + 4 * c;
          const double X = px - at[0], Y = py - at[1], Z = pz - at[2];
          double F[L + 1];
          boys<L>(p * (X * X + Y * Y + Z * Z), F, lds);
          double Rt[nherm(L)];
          herm_R<L>(Rt, F, p, X, Y, Z);
          const double w = -at[3] * pre * cc;
          sfor<NA>([&](auto I) {
            constexpr Cart a = cart(LA, I.value);
            sfor<NB>([&](auto J) {
              constexpr Cart b = cart(LB, J.value);
              double s = 0.0;
              sfor<a.x + b.x + 1>([&](auto T_) {
                sfor<a.y + b.y + 1>([&](auto U_) {
                  sfor<a.z + b.z + 1>([&](auto V_) {
                    s += Ex.e[a.x][b.x][T_.value] * Ey.e[a.y][b.y][U_.value] * Ez.e[a.z][b.z][V_.value] *
                         Rt[hidx(T_.value, U_.value, V_.value)];
                  });
                });
              });
              vt[I.value][J.value] += w * s;
            });
          });
        }
      }
    const int fa = sa.bf_off, fb = sb.bf_off;
    sfor<NA>([&](auto I) {
      sfor<NB>([&](auto J) {
        const int x = fa + I.value, y = fb + J.value;
        if (pr.a == pr.b && x < y) return;  // a diagonal pair writes each element once
        const double sc = A.bscale[x] * A.bscale[y];
        const double sv = sc * st[I.value][J.value], tv = sc * tt[I.value][J.value], vv = sc * vt[I.value][J.value];
        if (A.S) { A.S[(size_t)x * n + y] = sv; A.S[(size_t)y * n + x] = sv; }
        if (A.T) { A.T[(size_t)x * n + y] = tv; A.T[(size_t)y * n + x] = tv; }
        if (A.V) { A.V[(size_t)x * n + y] = vv; A.V[(size_t)y * n + x] = vv; }
      });
    });
  }
}

template <int LA, int LB>
int launch_one(const OneEArgs& a, hipStream_t st) {
  if (a.np == 0) return 0;
  const int grid = (a.np + 255) / 256;
  hipLaunchKernelGGL((k_one_e<LA, LB>), dim3(grid), dim3(256), sizeof(double) * boys_rows<LA + LB>() * kBoysLdsW, st,
                     a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace

int launch_one_electron(int pair_class, const OneEArgs& a, hipStream_t st) {
  switch (pair_class) {
    case 0: return launch_one<0, 0>(a, st);
    case 1: return launch_one<1, 0>(a, st);
    case 2: return launch_one<1, 1>(a, st);
    case 3: return launch_one<2, 0>(a, st);
    case 4: return launch_one<2, 1>(a, st);
    case 5: return launch_one<2, 2>(a, st);
    default: return -1;
  }
}

}  // namespace sqjk
