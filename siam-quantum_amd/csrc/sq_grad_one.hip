// sq_grad_one.hip - one-electron part of the Hartree-Fock energy gradient on
// the device (the loops of src/grad.c:2236-2306 over GTO_grad_overlap,
// GTO_grad_kinetic, GTO_grad_nai and GTO_efi).
//
// For E1 = sum_pq PT_pq (T + V)_pq - sum_pq QT_pq S_pq (QT the energy-weighted
// density) a lane owns one unique sub-shell pair (A, B) and adds
//   dE1/dA = 2 sum_ab [PT_ab d/dA (T + V)_ab - QT_ab d/dA S_ab]
//   dE1/dB = 2 sum_ab [PT_ab d/dB (T + V)_ab - QT_ab d/dB S_ab]   (A != B)
//   dE1/dC = f sum_ab PT_ab d/dC V^C_ab      for every nucleus C, f = 2 (A != B) or 1,
// with d/dA chi_a = 2 alpha chi_a(+1) - a_x chi_a(-1) along each axis, d/dB
// of S and T = -d/dA (they depend on A - B only), and for the attraction of
// nucleus C, V^C = -Z_C (2 pi / p) K sum E_tuv R_tuv(p, P - C): d/dC V^C from
// R_tuv -> -R_{t+1,u,v}, d/dB V^C = -(d/dA + d/dC) V^C.  A diagonal pair
// (A == B) contributes through its ordered components: its S and T terms
// cancel, its V term is the whole derivative on A.
#include <hip/hip_runtime.h>

#include "sq_one_e.h"

namespace sqjk {

namespace {

constexpr double kPi = 3.14159265358979323846;

template <int LA, int LB>
__global__ void __launch_bounds__(256) k_one_e_grad(OneEGradArgs A) {
  constexpr int NA = ncart(LA), NB = ncart(LB), L = LA + LB;
  extern __shared__ double lds[];
  boys_load_lds<L + 1>(lds, A.boys);
  __syncthreads();
  const int n = A.n;
  const int lane = threadIdx.x & 63;
  // every lane runs every iteration (lanes past the end on pair 0 with weight
  // 0), so that the per-nucleus sums can be reduced across the wavefront
  for (long long base = (long long)blockIdx.x * blockDim.x; base < A.np; base += (long long)gridDim.x * blockDim.x) {
    const long long t = base + threadIdx.x;
    const bool valid = t < A.np;
    const PairRec pr = A.pr[valid ? t : 0];
    const SubShell sa = A.sh[pr.a], sb = A.sh[pr.b];
    const bool diag = pr.a == pr.b;
    const double fw = valid ? 1.0 : 0.0;
    const double abx = sa.x - sb.x, aby = sa.y - sb.y, abz = sa.z - sb.z;
    const double ab2 = abx * abx + aby * aby + abz * abz;
    // weights: PT and -QT per component, scaled
    double wp[NA][NB], wq[NA][NB];
    sfor<NA>([&](auto I) {
      sfor<NB>([&](auto J) {
        const int x = sa.bf_off + I.value, y = sb.bf_off + J.value;
        const double sc = fw * A.bscale[x] * A.bscale[y];
        wp[I.value][J.value] = sc * A.PT[(size_t)x * n + y];
        wq[I.value][J.value] = -sc * A.QT[(size_t)x * n + y];
      });
    });
    double gA[3] = {0.0, 0.0, 0.0};  // d/dA of S and T, then of the attraction too
    for (int ka = 0; ka < sa.nprim; ++ka)
      for (int kb = 0; kb < sb.nprim; ++kb) {
        const double al = A.pexp[sa.prim_off + ka], be = A.pexp[sb.prim_off + kb];
        const double p = al + be, rp = 1.0 / p;
        const double K = exp(-al * be * rp * ab2);
        const double cc = A.pcoef[sa.prim_off + ka] * A.pcoef[sb.prim_off + kb];
        const double px = (al * sa.x + be * sb.x) * rp, py = (al * sa.y + be * sb.y) * rp,
                     pz = (al * sa.z + be * sb.z) * rp;
        const double h2p = 0.5 * rp;
        const double a2 = 2.0 * al;
        // overlap and kinetic: E^{ij}_0, i <= la + 1, j <= lb + 2
        Herm1<LA + 1, LB + 2> Ox, Oy, Oz;
        herm_coef<LA + 1, LB + 2>(Ox, px - sa.x, px - sb.x, h2p);
        herm_coef<LA + 1, LB + 2>(Oy, py - sa.y, py - sb.y, h2p);
        herm_coef<LA + 1, LB + 2>(Oz, pz - sa.z, pz - sb.z, h2p);
        const double so = cc * K * kPi * rp * sqrt(kPi * rp);
        // 1-D overlap s(i, j) and kinetic t(i, j) of one axis
        auto s1 = [&](const auto& E, auto ii, auto jj) -> double {
          return E.e[decltype(ii)::value][decltype(jj)::value][0];
        };
        auto t1 = [&](const auto& E, auto ii, auto jj) -> double {
          constexpr int i = decltype(ii)::value, j = decltype(jj)::value;
          double v = be * (2 * j + 1) * E.e[i][j][0] - 2.0 * be * be * E.e[i][j + 2][0];
          if constexpr (j >= 2) v -= 0.5 * j * (j - 1) * E.e[i][j - 2][0];
          return v;
        };
        // derivative along the axis with respect to A of a 1-D function f(i, j)
        auto dA = [&](auto f, const auto& E, auto ii, auto jj) -> double {
          constexpr int i = decltype(ii)::value;
          double v = a2 * f(E, ic<i + 1>{}, jj);
          if constexpr (i > 0) v -= double(i) * f(E, ic<i - 1>{}, jj);
          return v;
        };
        sfor<NA>([&](auto I) {
          constexpr Cart a = cart(LA, I.value);
          sfor<NB>([&](auto J) {
            constexpr Cart b = cart(LB, J.value);
            const double sx = s1(Ox, ic<a.x>{}, ic<b.x>{}), sy = s1(Oy, ic<a.y>{}, ic<b.y>{}),
                         sz = s1(Oz, ic<a.z>{}, ic<b.z>{});
            const double tx = t1(Ox, ic<a.x>{}, ic<b.x>{}), ty = t1(Oy, ic<a.y>{}, ic<b.y>{}),
                         tz = t1(Oz, ic<a.z>{}, ic<b.z>{});
            const double dsx = dA(s1, Ox, ic<a.x>{}, ic<b.x>{}), dsy = dA(s1, Oy, ic<a.y>{}, ic<b.y>{}),
                         dsz = dA(s1, Oz, ic<a.z>{}, ic<b.z>{});
            const double dtx = dA(t1, Ox, ic<a.x>{}, ic<b.x>{}), dty = dA(t1, Oy, ic<a.y>{}, ic<b.y>{}),
                         dtz = dA(t1, Oz, ic<a.z>{}, ic<b.z>{});
            const double ws = so * wq[I.value][J.value], wt = so * wp[I.value][J.value];
            gA[0] += ws * dsx * sy * sz + wt * (dtx * sy * sz + dsx * ty * sz + dsx * sy * tz);
            gA[1] += ws * sx * dsy * sz + wt * (tx * dsy * sz + sx * dty * sz + sx * dsy * tz);
            gA[2] += ws * sx * sy * dsz + wt * (tx * sy * dsz + sx * ty * dsz + sx * sy * dtz);
          });
        });
      }
    // nuclear attraction, nucleus by nucleus (the nucleus loop is uniform across
    // the wavefront, the primitive loops inside it are per lane)
    double gB[3] = {0.0, 0.0, 0.0};  // -(d/dC) parts of d/dB of the attraction
    const double f = diag ? 1.0 : 2.0;
    for (int c = 0; c < A.natom; ++c) {
      const double* at = A.atom + 4 * c;
      double dc[3] = {0.0, 0.0, 0.0};
      for (int ka = 0; ka < sa.nprim; ++ka)
        for (int kb = 0; kb < sb.nprim; ++kb) {
          const double al = A.pexp[sa.prim_off + ka], be = A.pexp[sb.prim_off + kb];
          const double p = al + be, rp = 1.0 / p;
          const double K = exp(-al * be * rp * ab2);
          const double pre = 2.0 * kPi * rp * K;
          if (pre < kPrimCutoff) continue;
          const double cc = A.pcoef[sa.prim_off + ka] * A.pcoef[sb.prim_off + kb];
          const double px = (al * sa.x + be * sb.x) * rp, py = (al * sa.y + be * sb.y) * rp,
                       pz = (al * sa.z + be * sb.z) * rp;
          const double h2p = 0.5 * rp;
          const double a2 = 2.0 * al;
          Herm1<LA + 1, LB> Ex, Ey, Ez;
          herm_coef<LA + 1, LB>(Ex, px - sa.x, px - sb.x, h2p);
          herm_coef<LA + 1, LB>(Ey, py - sa.y, py - sb.y, h2p);
          herm_coef<LA + 1, LB>(Ez, pz - sa.z, pz - sb.z, h2p);
          const double X = px - at[0], Y = py - at[1], Z = pz - at[2];
          double F[L + 2];
          boys<L + 1>(p * (X * X + Y * Y + Z * Z), F, lds);
          double Rt[nherm(L + 1)];
          herm_R<L + 1>(Rt, F, p, X, Y, Z);
          const double w = -at[3] * pre * cc;
          sfor<NA>([&](auto I) {
            constexpr Cart a = cart(LA, I.value);
            sfor<NB>([&](auto J) {
              constexpr Cart b = cart(LB, J.value);
              const double wv = w * wp[I.value][J.value];
              // sum E^{ab}_tuv R_{tuv + e_x}: minus the C derivative
              double rx = 0.0, ry = 0.0, rz = 0.0, ax_ = 0.0, ay_ = 0.0, az_ = 0.0;
              sfor<a.x + b.x + 1>([&](auto T_) {
                sfor<a.y + b.y + 1>([&](auto U_) {
                  sfor<a.z + b.z + 1>([&](auto V_) {
                    constexpr int t = T_.value, u = U_.value, v = V_.value;
                    const double e = Ex.e[a.x][b.x][t] * Ey.e[a.y][b.y][u] * Ez.e[a.z][b.z][v];
                    rx += e * Rt[hidx(t + 1, u, v)];
                    ry += e * Rt[hidx(t, u + 1, v)];
                    rz += e * Rt[hidx(t, u, v + 1)];
                  });
                });
              });
              // d/dA: raised and lowered a along each axis
              sfor<a.x + b.x + 2>([&](auto T_) {
                constexpr int t = T_.value;
                double e = a2 * Ex.e[a.x + 1][b.x][t];
                if constexpr (a.x > 0 && t <= a.x - 1 + b.x) e -= double(a.x) * Ex.e[a.x - 1][b.x][t];
                sfor<a.y + b.y + 1>([&](auto U_) {
                  sfor<a.z + b.z + 1>([&](auto V_) {
                    ax_ += e * Ey.e[a.y][b.y][U_.value] * Ez.e[a.z][b.z][V_.value] * Rt[hidx(t, U_.value, V_.value)];
                  });
                });
              });
              sfor<a.y + b.y + 2>([&](auto U_) {
                constexpr int u = U_.value;
                double e = a2 * Ey.e[a.y + 1][b.y][u];
                if constexpr (a.y > 0 && u <= a.y - 1 + b.y) e -= double(a.y) * Ey.e[a.y - 1][b.y][u];
                sfor<a.x + b.x + 1>([&](auto T_) {
                  sfor<a.z + b.z + 1>([&](auto V_) {
                    ay_ += Ex.e[a.x][b.x][T_.value] * e * Ez.e[a.z][b.z][V_.value] * Rt[hidx(T_.value, u, V_.value)];
                  });
                });
              });
              sfor<a.z + b.z + 2>([&](auto V_) {
                constexpr int v = V_.value;
                double e = a2 * Ez.e[a.z + 1][b.z][v];
                if constexpr (a.z > 0 && v <= a.z - 1 + b.z) e -= double(a.z) * Ez.e[a.z - 1][b.z][v];
                sfor<a.x + b.x + 1>([&](auto T_) {
                  sfor<a.y + b.y + 1>([&](auto U_) {
                    az_ += Ex.e[a.x][b.x][T_.value] * Ey.e[a.y][b.y][U_.value] * e * Rt[hidx(T_.value, U_.value, v)];
                  });
                });
              });
              dc[0] -= wv * rx;
              dc[1] -= wv * ry;
              dc[2] -= wv * rz;
              gA[0] += wv * ax_;
              gA[1] += wv * ay_;
              gA[2] += wv * az_;
            });
          });
        }
      // d/dC of this nucleus, reduced over the wavefront (all lanes are here)
      sfor<3>([&](auto X_) {
        constexpr int x = X_.value;
        const double s = wave_sum(f * dc[x]);
        if (lane == 0) fatomic(&A.gatom[3 * c + x], s);
        gB[x] -= dc[x];
      });
    }
    // d/dA on A; d/dB = -d/dA for S and T, -(d/dA + d/dC) for the attraction
    if (valid)
      sfor<3>([&](auto X_) {
        constexpr int x = X_.value;
        fatomic(&A.gsub[3 * pr.a + x], 2.0 * gA[x]);
        if (!diag) fatomic(&A.gsub[3 * pr.b + x], 2.0 * (gB[x] - gA[x]));
      });
  }
}

template <int LA, int LB>
int launch_one_grad(const OneEGradArgs& a, hipStream_t st) {
  if (a.np == 0) return 0;
  const int grid = (a.np + 255) / 256;
  hipLaunchKernelGGL((k_one_e_grad<LA, LB>), dim3(grid), dim3(256),
                     sizeof(double) * boys_rows<LA + LB + 1>() * kBoysLdsW, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace

int launch_one_electron_grad(int pair_class, const OneEGradArgs& a, hipStream_t st) {
  switch (pair_class) {
    case 0: return launch_one_grad<0, 0>(a, st);
    case 1: return launch_one_grad<1, 0>(a, st);
    case 2: return launch_one_grad<1, 1>(a, st);
    case 3: return launch_one_grad<2, 0>(a, st);
    case 4: return launch_one_grad<2, 1>(a, st);
    case 5: return launch_one_grad<2, 2>(a, st);
    default: return -1;
  }
}

}  // namespace sqjk
