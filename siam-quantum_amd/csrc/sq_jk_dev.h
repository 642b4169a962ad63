// sq_jk_dev.h - data layout in HBM and the per-class quartet kernels.
//
// Work unit: one contracted shell quartet (bra pair i, ket pair j) of one
// angular class (LA LB | LC LD), LA>=LB, LC>=LD, (LA,LB) >= (LC,LD).  In the
// per-thread kernel below a thread evaluates a slice of the bra components
// of that quartet for all ket components (one launch per slice) and
// digests its integrals into J/K with atomics.
#pragma once
#include "sq_md.h"

namespace sqjk {

// one pure-angular-momentum shell (a Pople L shell becomes an s and a p)
struct SubShell {
  double x, y, z;
  int l, nprim, prim_off, bf_off;  // first basis function
  int parent;                       // reference shell (screening unit)
  int pad;
};

// unique shell pair, a = higher angular momentum
struct PairRec {
  int a, b;          // sub-shell indices
  int pp_off, npp;   // primitive pairs
  int pa, pb;        // parent reference shells of a, b
  double sch;        // Schwarz bound of the parent pair (block maximum)
};

// primitive pair data, 48 B
struct PrimPair {
  double p;          // alpha_a + alpha_b
  double Px, Py, Pz;
  double cc;         // c_a c_b exp(-alpha_a alpha_b AB^2 / p) / p
  double eb;         //         exp(-alpha_a alpha_b AB^2 / p) / p
};

enum Mode { kRHF = 0, kUHF = 1, kSchwarz = 2 };

struct ClassArgs {
  const SubShell* __restrict__ sh;
  const PairRec* __restrict__ bra;
  const PairRec* __restrict__ ket;
  const PrimPair* __restrict__ pp;
  const long long* __restrict__ off;  // [nbra+1] quartet offsets per bra pair
  const int* __restrict__ kstart;     // [nbra] first ket pair per bra pair (NULL: 0)
  const double* __restrict__ boys;    // Boys table (global)
  const double* __restrict__ DJ;      // density for J (PA+PB)
  const double* __restrict__ DA;      // alpha density for K
  const double* __restrict__ DB;      // beta density for K (UHF)
  const double* __restrict__ PM;      // per reference-shell-pair max |P|
  const double* __restrict__ bscale;  // per basis function component scale
  double* AJ;                         // J half-accumulator (n x n)
  double* AKA;                        // K half-accumulators
  double* AKB;
  double* sch_out;                    // Schwarz mode output (n x n)
  long long total;                    // quartets in this launch
  double fixed_cut;                   // density-weighted screening cut
  int nbra, same, n, nref;
};

using namespace sqmd;

// slice geometry: bra components per thread
template <int LA, int LB, int LC, int LD>
struct SliceInfo {
  static constexpr int NAB = ncart(LA) * ncart(LB);
  static constexpr int NCD = ncart(LC) * ncart(LD);
  static constexpr int SZ = (36 / NCD) > 0 ? ((36 / NCD) < NAB ? (36 / NCD) : NAB) : 1;
  static constexpr int NSLICE = (NAB + SZ - 1) / SZ;
};

__host__ __device__ __forceinline__ void fatomic(double* p, double v) {
#ifdef __HIP_DEVICE_COMPILE__
  unsafeAtomicAdd(p, v);
#else
  *p += v;  // host instantiation (used only to check the kernel arithmetic on the CPU)
#endif
}

// R: precision of the integral evaluation (double, or float for the
// mixed-precision range); digestion and accumulation are always FP64.
template <int LA, int LB, int LC, int LD, int AB0, int AB1, int MODE, typename R = double>
__host__ __device__ __forceinline__ void quartet(const ClassArgs& A, int ib, int ik, const double* lds) {
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  constexpr int LAB = LA + LB, LCD = LC + LD, L = LAB + LCD;
  constexpr int NS = AB1 - AB0, NCD = NC * ND;

  const PairRec bp = A.bra[ib];
  const PairRec kp = A.ket[ik];
  const SubShell sA = A.sh[bp.a], sB = A.sh[bp.b], sC = A.sh[kp.a], sD = A.sh[kp.b];

  R acc[NS][NCD];
  sfor<NS>([&](auto s) { sfor<NCD>([&](auto c) { acc[s][c] = R(0); }); });

  for (int u = 0; u < bp.npp; ++u) {
    const PrimPair P = A.pp[bp.pp_off + u];
    const R h2p = R(0.5 / P.p);
    Herm1<LA, LB, R> Ex, Ey, Ez;
    herm_coef<LA, LB, R>(Ex, R(P.Px - sA.x), R(P.Px - sB.x), h2p);
    herm_coef<LA, LB, R>(Ey, R(P.Py - sA.y), R(P.Py - sB.y), h2p);
    herm_coef<LA, LB, R>(Ez, R(P.Pz - sA.z), R(P.Pz - sB.z), h2p);
    for (int v = 0; v < kp.npp; ++v) {
      const PrimPair Q = A.pp[kp.pp_off + v];
      const double s = P.p + Q.p;
      const double rs = 1.0 / sqrt(s);
      if constexpr (MODE != kSchwarz) {
        if (kTwoPi25 * P.eb * Q.eb * rs < kPrimCutoff) continue;
      }
      const R alpha = R(P.p * Q.p / s);
      const R X = R(P.Px - Q.Px), Y = R(P.Py - Q.Py), Z = R(P.Pz - Q.Pz);
      const R T = alpha * (X * X + Y * Y + Z * Z);
      R F[L + 1];
      boys<L, R>(T, F, lds);
      R Rt[nherm(L)];
      herm_R<L, R>(Rt, F, alpha, X, Y, Z);
      const R h2q = R(0.5 / Q.p);
      Herm1<LC, LD, R> Fx, Fy, Fz;
      herm_coef<LC, LD, R>(Fx, R(Q.Px - sC.x), R(Q.Px - sD.x), h2q);
      herm_coef<LC, LD, R>(Fy, R(Q.Py - sC.y), R(Q.Py - sD.y), h2q);
      herm_coef<LC, LD, R>(Fz, R(Q.Pz - sC.z), R(Q.Pz - sD.z), h2q);
      const R pref = R(kTwoPi25 * P.cc * Q.cc * rs);

      sfor<NCD>([&](auto cdI) {
        constexpr int cd = decltype(cdI)::value;
        constexpr Cart c = cart(LC, cd / ND), d = cart(LD, cd % ND);
        constexpr int kx = c.x + d.x, ky = c.y + d.y, kz = c.z + d.z;
        // W[t][u][v] = sum_{TUV} (-1)^{T+U+V} E^{cd}_{TUV} R_{t+T,u+U,v+V}
        R W[LAB + 1][LAB + 1][LAB + 1];
        sfor<LAB + 1>([&](auto tI) {
          constexpr int t = decltype(tI)::value;
          sfor<LAB + 1 - t>([&](auto uI) {
            constexpr int uu = decltype(uI)::value;
            sfor<LAB + 1 - t - uu>([&](auto vI) {
              constexpr int vv = decltype(vI)::value;
              R w = R(0);
              sfor<kx + 1>([&](auto TI) {
                constexpr int T_ = decltype(TI)::value;
                sfor<ky + 1>([&](auto UI) {
                  constexpr int U_ = decltype(UI)::value;
                  R wz = R(0);
                  sfor<kz + 1>([&](auto VI) {
                    constexpr int V_ = decltype(VI)::value;
                    constexpr R sg = ((T_ + U_ + V_) & 1) ? R(-1) : R(1);
                    wz += sg * Fz.e[c.z][d.z][V_] * Rt[hidx(t + T_, uu + U_, vv + V_)];
                  });
                  w += Fx.e[c.x][d.x][T_] * Fy.e[c.y][d.y][U_] * wz;
                });
              });
              W[t][uu][vv] = w;
            });
          });
        });
        sfor<NS>([&](auto sI) {
          constexpr int ab = AB0 + decltype(sI)::value;
          constexpr Cart a = cart(LA, ab / NB), b = cart(LB, ab % NB);
          R val = R(0);
          sfor<a.x + b.x + 1>([&](auto tI) {
            constexpr int t = decltype(tI)::value;
            sfor<a.y + b.y + 1>([&](auto uI) {
              constexpr int uu = decltype(uI)::value;
              R vz = R(0);
              sfor<a.z + b.z + 1>([&](auto vI) {
                constexpr int vv = decltype(vI)::value;
                vz += Ez.e[a.z][b.z][vv] * W[t][uu][vv];
              });
              val += Ex.e[a.x][b.x][t] * Ey.e[a.y][b.y][uu] * vz;
            });
          });
          acc[decltype(sI)::value][cd] += pref * val;
        });
      });
    }
  }

  const int n = A.n;
  const int fa = sA.bf_off, fb = sB.bf_off, fc = sC.bf_off, fd = sD.bf_off;

  if constexpr (MODE == kSchwarz) {
    // diagonal (ab|ab) of the quartet (i,i): sqrt|.| into both (a,b), (b,a)
    sfor<NS>([&](auto sI) {
      constexpr int ab = AB0 + decltype(sI)::value;
      constexpr int ia = ab / NB, jb = ab % NB;
      const double sc = A.bscale[fa + ia] * A.bscale[fb + jb];
      const double v = sqrt(fabs((double)acc[decltype(sI)::value][ab] * sc * sc));
      A.sch_out[(fa + ia) * n + fb + jb] = v;
      A.sch_out[(fb + jb) * n + fa + ia] = v;
    });
    return;
  } else {
    double deg = 1.0;
    if (bp.a == bp.b) deg *= 0.5;
    if (kp.a == kp.b) deg *= 0.5;
    if (A.same && ib == ik) deg *= 0.5;

    double sc_c[NC], sc_d[ND];
    sfor<NC>([&](auto I) { sc_c[decltype(I)::value] = A.bscale[fc + decltype(I)::value]; });
    sfor<ND>([&](auto I) { sc_d[decltype(I)::value] = A.bscale[fd + decltype(I)::value]; });

    // scaled integrals, FP64 from here on
    double vv_[NS][NCD];
    sfor<NS>([&](auto sI) {
      constexpr int ab = AB0 + decltype(sI)::value;
      const double sab = deg * A.bscale[fa + ab / NB] * A.bscale[fb + ab % NB];
      sfor<NCD>([&](auto cI) {
        constexpr int cd = decltype(cI)::value;
        vv_[decltype(sI)::value][cd] = (double)acc[decltype(sI)::value][cd] * sab * sc_c[cd / ND] * sc_d[cd % ND];
      });
    });

    // Coulomb: A_J[a,b] += 2 sum_cd v D[c,d];  A_J[c,d] += 2 sum_ab v D[a,b]
    double dcd[NCD];
    sfor<NCD>([&](auto cI) {
      constexpr int cd = decltype(cI)::value;
      dcd[cd] = A.DJ[(fc + cd / ND) * n + fd + cd % ND];
    });
    double jcd[NCD];
    sfor<NCD>([&](auto cI) { jcd[decltype(cI)::value] = 0.0; });
    sfor<NS>([&](auto sI) {
      constexpr int ab = AB0 + decltype(sI)::value;
      constexpr int s = decltype(sI)::value;
      const double dab = A.DJ[(fa + ab / NB) * n + fb + ab % NB];
      double jab = 0.0;
      sfor<NCD>([&](auto cI) {
        constexpr int cd = decltype(cI)::value;
        jab += vv_[s][cd] * dcd[cd];
        jcd[cd] += vv_[s][cd] * dab;
      });
      fatomic(&A.AJ[(fa + ab / NB) * n + fb + ab % NB], 2.0 * jab);
    });
    sfor<NCD>([&](auto cI) {
      constexpr int cd = decltype(cI)::value;
      fatomic(&A.AJ[(fc + cd / ND) * n + fd + cd % ND], 2.0 * jcd[cd]);
    });

    // Exchange (per spin): A_K[a,c] += v D[b,d]; A_K[b,c] += v D[a,d];
    //                      A_K[a,d] += v D[b,c]; A_K[b,d] += v D[a,c]
    auto exchange = [&](const double* __restrict__ D, double* AK) {
      constexpr int A0 = AB0 / NB, A1 = (AB1 - 1) / NB;  // bra a-range in slice
      constexpr int NAr = A1 - A0 + 1;
      constexpr int B0 = (NAr > 1) ? 0 : AB0 % NB;
      constexpr int B1 = (NAr > 1) ? NB - 1 : (AB1 - 1) % NB;
      constexpr int NBr = B1 - B0 + 1;
      double kac[NAr][NC], kad[NAr][ND], kbc[NBr][NC], kbd[NBr][ND];
      sfor<NAr>([&](auto I) { sfor<NC>([&](auto J) { kac[I.value][J.value] = 0.0; });
                              sfor<ND>([&](auto J) { kad[I.value][J.value] = 0.0; }); });
      sfor<NBr>([&](auto I) { sfor<NC>([&](auto J) { kbc[I.value][J.value] = 0.0; });
                              sfor<ND>([&](auto J) { kbd[I.value][J.value] = 0.0; }); });
      sfor<NS>([&](auto sI) {
        constexpr int s = decltype(sI)::value;
        constexpr int ab = AB0 + s;
        constexpr int ia = ab / NB, jb = ab % NB;
        const int ga = fa + ia, gb = fb + jb;
        sfor<NC>([&](auto cI) {
          constexpr int ic_ = decltype(cI)::value;
          const int gc = fc + ic_;
          const double dbc = D[gb * n + gc], dac = D[ga * n + gc];
          sfor<ND>([&](auto dI) {
            constexpr int id = decltype(dI)::value;
            const int gd = fd + id;
            const double v = vv_[s][ic_ * ND + id];
            kac[ia - A0][ic_] += v * D[gb * n + gd];
            kbc[jb - B0][ic_] += v * D[ga * n + gd];
            kad[ia - A0][id] += v * dbc;
            kbd[jb - B0][id] += v * dac;
          });
        });
      });
      sfor<NAr>([&](auto I) {
        constexpr int ia = A0 + decltype(I)::value;
        sfor<NC>([&](auto J) { fatomic(&AK[(fa + ia) * n + fc + J.value], kac[I.value][J.value]); });
        sfor<ND>([&](auto J) { fatomic(&AK[(fa + ia) * n + fd + J.value], kad[I.value][J.value]); });
      });
      sfor<NBr>([&](auto I) {
        constexpr int jb = B0 + decltype(I)::value;
        sfor<NC>([&](auto J) { fatomic(&AK[(fb + jb) * n + fc + J.value], kbc[I.value][J.value]); });
        sfor<ND>([&](auto J) { fatomic(&AK[(fb + jb) * n + fd + J.value], kbd[I.value][J.value]); });
      });
    };
    exchange(A.DA, A.AKA);
    if constexpr (MODE == kUHF) exchange(A.DB, A.AKB);
  }
}

// quartet index -> (bra pair, ket pair) through the per-bra offsets
__device__ __forceinline__ int find_bra(const long long* __restrict__ off, int nbra, long long t) {
  int lo = 0, hi = nbra;  // off[lo] <= t < off[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= t) lo = mid; else hi = mid;
  }
  return lo;
}

// One launch per (class, mode, slice S).  Grid-stride over the quartets of
// the class; each workgroup first copies the Boys rows it needs into LDS.
template <int LA, int LB, int LC, int LD, int MODE, int S, typename R = double>
__global__ void __launch_bounds__(256) k_class(ClassArgs A) {
  using SI = SliceInfo<LA, LB, LC, LD>;
  constexpr int L = LA + LB + LC + LD;
  constexpr int AB0 = S * SI::SZ;
  constexpr int AB1 = (AB0 + SI::SZ < SI::NAB) ? AB0 + SI::SZ : SI::NAB;
  extern __shared__ double lds[];
  boys_load_lds<L>(lds, A.boys);
  __syncthreads();
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < A.total; t += stride) {
    int ib, ik;
    if constexpr (MODE == kSchwarz) {
      ib = ik = (int)t;
    } else {
      ib = find_bra(A.off, A.nbra, t);
      ik = (int)(t - A.off[ib]) + (A.kstart ? A.kstart[ib] : 0);
      // density-weighted screening at the reference-shell level
      const PairRec bp = A.bra[ib], kp = A.ket[ik];
      const double* PM = A.PM;
      const int nr = A.nref;
      const double mp = PM[bp.pa * nr + bp.pb] + PM[kp.pa * nr + kp.pb] + PM[bp.pa * nr + kp.pa] +
                        PM[bp.pa * nr + kp.pb] + PM[bp.pb * nr + kp.pa] + PM[bp.pb * nr + kp.pb];
      if (bp.sch * kp.sch * mp < A.fixed_cut) continue;
    }
    quartet<LA, LB, LC, LD, AB0, AB1, MODE, R>(A, ib, ik, lds);
  }
}

template <int LA, int LB, int LC, int LD>
constexpr size_t class_lds_bytes() {
  return sizeof(double) * kBoysNT * (LA + LB + LC + LD + kBoysTaylor);
}

}  // namespace sqjk
