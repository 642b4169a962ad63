// sq_md.h - device-side building blocks of the two-electron kernels.
//
// Integrals are evaluated with the McMurchie-Davidson scheme: Hermite
// expansion coefficients E^{ij}_t per Cartesian axis, the Hermite Coulomb
// tensor R_{tuv} from the Boys function, and the contraction
//   (ab|cd) = 2 pi^{5/2}/(p q sqrt(p+q)) K_ab K_cd
//             sum_{tuv} E^{ab}_{tuv} sum_{TUV} (-1)^{T+U+V} E^{cd}_{TUV} R_{t+T,u+U,v+V}.
// Every loop bound is a template parameter, so the whole primitive-quartet
// evaluation unrolls into straight-line code with all intermediates in
// registers.  The Boys function F_m(T) is read from a table that each
// workgroup keeps in LDS (Taylor expansion about the nearest grid point plus
// downward recursion).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>
#include <utility>

namespace sqmd {

// ---------------------------------------------------------------- basics
template <int V> using ic = std::integral_constant<int, V>;

template <typename F, int... I>
__host__ __device__ __forceinline__ void sfor_impl(F&& f, std::integer_sequence<int, I...>) {
  (f(ic<I>{}), ...);
}
// compile-time for: f(ic<0>), ..., f(ic<N-1>)
template <int N, typename F>
__host__ __device__ __forceinline__ void sfor(F&& f) {
  sfor_impl(f, std::make_integer_sequence<int, N>{});
}

__host__ __device__ constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }
// number of (t,u,v) with t+u+v <= L
__host__ __device__ constexpr int nherm(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }

// Cartesian component order used by the GAMESS reader of the reference
// (src/basis.c:464-498): s; x y z; xx yy zz xy xz yz; f in that file's order.
struct Cart { int x, y, z; };
// Component i of angular momentum l, by arithmetic only (no lookup arrays,
// so that the value is the same in every compilation context).
__host__ __device__ constexpr Cart cart(int l, int i) {
  return l == 0 ? Cart{0, 0, 0}
       : l == 1 ? Cart{i == 0 ? 1 : 0, i == 1 ? 1 : 0, i == 2 ? 1 : 0}
       : l == 2 ? (i < 3 ? Cart{i == 0 ? 2 : 0, i == 1 ? 2 : 0, i == 2 ? 2 : 0}
                         : Cart{i == 5 ? 0 : 1, i == 4 ? 0 : 1, i == 3 ? 0 : 1})
       : (i == 0 ? Cart{3, 0, 0} : i == 1 ? Cart{0, 3, 0} : i == 2 ? Cart{0, 0, 3}
        : i == 3 ? Cart{2, 1, 0} : i == 4 ? Cart{1, 2, 0} : i == 5 ? Cart{2, 0, 1}
        : i == 6 ? Cart{1, 0, 2} : i == 7 ? Cart{0, 2, 1} : i == 8 ? Cart{0, 1, 2}
        : Cart{1, 1, 1});
}

// compact index of (t,u,v), t+u+v <= L, grouped by total order
__host__ __device__ constexpr int hidx(int t, int u, int v) {
  const int n = t + u + v;
  // entries of lower total order, then position inside order n
  const int below = n * (n + 1) * (n + 2) / 6;
  const int m = u + v;           // t = n - m
  return below + m * (m + 1) / 2 + v;
}

constexpr double kTwoPi25 = 34.986836655249725;  // 2 pi^{5/2}
constexpr double kPrimCutoff = 1.0e-15;            // bare prefactor cut

// ---------------------------------------------------------------- Boys
// Table: F_m(k*kBoysStep) for k in [0,kBoysNT), m in [0,kBoysNM), then
// exp(-k*kBoysStep) in column kBoysNM (row stride kBoysCols).
constexpr int kBoysNT = 651;
constexpr double kBoysStep = 0.1;
constexpr double kBoysTMax = (kBoysNT - 1) * kBoysStep;  // 65
constexpr int kBoysNM = 28;
constexpr int kBoysCols = kBoysNM + 1;
constexpr int kBoysTaylor = 9;  // terms of the Taylor expansion
constexpr int kBoysLdsW = kBoysTaylor + 1;  // LDS row: F_L .. F_{L+8}, exp(-T_k)

// Large-T branch of boys<L>: from T >= kBoysTCut[L] - 0.05 on, F_0 =
// sqrt(pi/T)/2 and the upward recursion without the exp(-T) term are within
// 2e-17 relative of F_m for every m <= L (checked against mpmath in steps of
// 0.5), so a class of total angular momentum L needs table rows up to
// kBoysTCut[L] only.  L = 9 is the (dd|dd) energy gradient.
constexpr double kBoysTCut[10] = {36.5, 40.5, 44.0, 47.0, 50.0, 52.5, 55.0, 57.5, 60.0, 62.5};
template <int L>
constexpr int boys_rows() {
  static_assert(L >= 0 && L <= 9, "Boys table cut-offs cover L <= 9");
  return (int)(kBoysTCut[L] / kBoysStep + 0.5) + 1;
}
static_assert(boys_rows<9>() <= kBoysNT, "table shorter than the largest cut-off");

// LDS copy for a class of total angular momentum L: per grid point the
// Taylor coefficients F_L..F_{L+8} and exp(-T_k).
template <int L>
__device__ __forceinline__ void boys_load_lds(double* lds, const double* __restrict__ tab) {
  for (int e = threadIdx.x + threadIdx.y * blockDim.x; e < boys_rows<L>() * kBoysLdsW;
       e += blockDim.x * blockDim.y) {
    const int k = e / kBoysLdsW, j = e - k * kBoysLdsW;
    lds[e] = tab[k * kBoysCols + (j < kBoysTaylor ? L + j : kBoysNM)];
  }
}

// 1/sqrt(s) for s > 0: the hardware estimate refined by two Newton steps
// y <- y (3/2 - s/2 y^2), each of which squares the relative error.
__host__ __device__ __forceinline__ double rsqrt_nr(double s) {
#ifdef __HIP_DEVICE_COMPILE__
  double y = __builtin_amdgcn_rsq(s);
  const double h = 0.5 * s;
  y = y * fma(-h * y, y, 1.5);
  y = y * fma(-h * y, y, 1.5);
  return y;
#else
  return 1.0 / sqrt(s);
#endif
}
__host__ __device__ __forceinline__ float rsqrt_nr(float s) {
#ifdef __HIP_DEVICE_COMPILE__
  float y = __builtin_amdgcn_rsqf(s);  // hardware estimate, then one Newton step
  return y * (1.5f - 0.5f * s * y * y);
#else
  return 1.0f / sqrtf(s);
#endif
}


// F_0..F_L at T, evaluated in precision R (the table itself is FP64).
// T < kBoysTCut[L] - 0.05: F_L by Taylor expansion about the nearest grid point T_k,
// exp(-T) = exp(-T_k) exp(T_k - T) with the second factor by its Taylor
// series (|T_k - T| <= 0.05), then downward recursion.  Larger T: F_0 =
// sqrt(pi/T)/2 from the refined reciprocal square root (erfc(sqrt(T)) <
// 1e-27), and the upward recursion F_{m+1} = (2m+1) F_m / (2T) without the
// exp(-T) term, which changes F_m (m <= L) by less than 2e-17 relative
// there, so no exponential is evaluated on this branch.  FP32 evaluation
// needs fewer Taylor terms for the same relative accuracy.
template <int L, typename R = double>
__host__ __device__ __forceinline__ void boys(R T, R (&F)[L + 1], const double* lds) {
  constexpr int NT = sizeof(R) == 4 ? 5 : kBoysTaylor;  // 0.05^5/5! ~ 2.6e-9
  if (T < R(kBoysTCut[L] - 0.5 * kBoysStep)) {
    const int k = (int)(T * R(1.0 / kBoysStep) + R(0.5));
    const R d = R(k * kBoysStep) - T;  // expansion in (T_k - T)
    const double* row = lds + k * kBoysLdsW;
    // F_L(T) = sum_j F_{L+j}(T_k) (T_k - T)^j / j!
    R s = R(row[NT - 1]);
    sfor<NT - 1>([&](auto jj) {
      constexpr int j = NT - 2 - decltype(jj)::value;
      s = s * (d * R(1.0 / (j + 1))) + R(row[j]);
    });
    F[L] = s;
    if constexpr (L > 0) {
      // exp(d) = sum_j d^j / j!, same number of terms as above
      R x = R(1);
      sfor<NT - 1>([&](auto jj) {
        constexpr int j = NT - 1 - decltype(jj)::value;
        x = x * (d * R(1.0 / j)) + R(1);
      });
      const R e = R(row[kBoysTaylor]) * x, t2 = T + T;
      sfor<L>([&](auto mm) {
        constexpr int m = L - 1 - decltype(mm)::value;
        F[m] = (t2 * F[m + 1] + e) * R(1.0 / (2 * m + 1));
      });
    }
  } else {
    // large T: F_0 = sqrt(pi/T)/2, F_{m+1} = (2m+1) F_m / (2T)
    const R y = rsqrt_nr(T);
    F[0] = R(0.88622692545275801365) * y;
    if constexpr (L > 0) {
      const R h = R(0.5) * y * y;
      sfor<L>([&](auto mm) {
        constexpr int m = decltype(mm)::value;
        F[m + 1] = (R(2 * m + 1) * h) * F[m];
      });
    }
  }
}

// ------------------------------------------------------ Hermite coefficients
// E[i][j][t], i<=LI, j<=LJ, t<=i+j  (unit K_ab factored out)
template <int LI, int LJ, typename R = double>
struct Herm1 {
  R e[LI + 1][LJ + 1][LI + LJ + 1];
};

template <int LI, int LJ, typename R>
__host__ __device__ __forceinline__ void herm_coef(Herm1<LI, LJ, R>& E, R xpa, R xpb, R h2p) {
  E.e[0][0][0] = R(1);
  sfor<LI>([&](auto ii) {
    constexpr int i = decltype(ii)::value;
    sfor<i + 2>([&](auto tt) {
      constexpr int t = decltype(tt)::value;
      R v = R(0);
      if constexpr (t >= 1) v += h2p * E.e[i][0][t - 1];
      if constexpr (t <= i) v += xpa * E.e[i][0][t];
      if constexpr (t + 1 <= i) v += R(t + 1) * E.e[i][0][t + 1];
      E.e[i + 1][0][t] = v;
    });
  });
  sfor<LI + 1>([&](auto ii) {
    constexpr int i = decltype(ii)::value;
    sfor<LJ>([&](auto jj) {
      constexpr int j = decltype(jj)::value;
      sfor<i + j + 2>([&](auto tt) {
        constexpr int t = decltype(tt)::value;
        R v = R(0);
        if constexpr (t >= 1) v += h2p * E.e[i][j][t - 1];
        if constexpr (t <= i + j) v += xpb * E.e[i][j][t];
        if constexpr (t + 1 <= i + j) v += R(t + 1) * E.e[i][j][t + 1];
        E.e[i][j + 1][t] = v;
      });
    });
  });
}

// ------------------------------------------------------ Hermite Coulomb R
// R^{(0)}_{tuv} for t+u+v <= L, built level by level from
// R^{(n)}_{000} = (-2 alpha)^n F_n.
template <int L, typename Rl = double>
__host__ __device__ __forceinline__ void herm_R(Rl (&R)[nherm(L)], const Rl (&F)[L + 1],
                                       Rl alpha, Rl X, Rl Y, Rl Z) {
  Rl lev[L + 1][nherm(L)];
  Rl m2a = Rl(-2) * alpha, pw = Rl(1);
  sfor<L + 1>([&](auto nn) {
    constexpr int n = decltype(nn)::value;
    lev[n][0] = pw * F[n];
    pw *= m2a;
  });
  // level n holds orders up to L-n; build from n = L-1 down to 0
  sfor<L>([&](auto kk) {
    constexpr int n = L - 1 - decltype(kk)::value;
    constexpr int top = L - n;
    sfor<top>([&](auto oo) {
      constexpr int o = decltype(oo)::value + 1;  // total order 1..top
      sfor<o + 1>([&](auto mm) {
        constexpr int m = decltype(mm)::value;     // u+v
        sfor<m + 1>([&](auto vv) {
          constexpr int v = decltype(vv)::value;
          constexpr int u = m - v, t = o - m;
          Rl r;
          if constexpr (t > 0) {
            r = X * lev[n + 1][hidx(t - 1, u, v)];
            if constexpr (t > 1) r += Rl(t - 1) * lev[n + 1][hidx(t - 2, u, v)];
          } else if constexpr (u > 0) {
            r = Y * lev[n + 1][hidx(t, u - 1, v)];
            if constexpr (u > 1) r += Rl(u - 1) * lev[n + 1][hidx(t, u - 2, v)];
          } else {
            r = Z * lev[n + 1][hidx(t, u, v - 1)];
            if constexpr (v > 1) r += Rl(v - 1) * lev[n + 1][hidx(t, u, v - 2)];
          }
          lev[n][hidx(t, u, v)] = r;
        });
      });
    });
  });
  sfor<nherm(L)>([&](auto ii) { R[decltype(ii)::value] = lev[0][decltype(ii)::value]; });
}

}  // namespace sqmd
