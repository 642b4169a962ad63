// sq_jk_dev.h - data layout in HBM and the per-class quartet kernels.
//
// Work unit: one contracted shell quartet (bra pair i, ket pair j) of one
// angular class (LA LB | LC LD), LA>=LB, LC>=LD, (LA,LB) >= (LC,LD).  In the
// per-thread kernel below a lane evaluates all bra components of that quartet
// for one slice of its ket components (one launch per slice) and digests its
// integrals into J/K with atomics.
//
// Work lists are wavefront-uniform: a launch covers one class pair and one
// precision, and is cut into segments (bra group x ket group, where a group
// holds the pairs of one class with the same number of primitive pairs), so
// all lanes of a wavefront run the same primitive loop counts.  A tile is a
// run of bra rows with the same first shell a and a chunk of 64 consecutive
// ket pairs: lane l keeps ket pair k0 + l for every row of the run, so its J
// and a-side exchange contributions stay in registers across the run.
#pragma once
#include "sq_md.h"

namespace sqjk {

// one pure-angular-momentum shell (a Pople L shell becomes an s and a p)
struct SubShell {
  double x, y, z;
  int l, nprim, prim_off, bf_off;  // first basis function (of its molecule)
  int parent;                       // reference shell (screening unit)
  int gbf;                          // first basis function in the component
                                    // scales (bscale), across a batch
};

// Unique shell pair, a = higher angular momentum.  Primitive pair v of the
// pair is pp[pp_off + v * pp_stride]: the primitive data of a group is stored
// primitive-major, so consecutive pairs of a group read consecutive records.
struct PairRec {
  int a, b;               // sub-shell indices
  int pa, pb;             // parent reference shells of a, b (of their molecule)
  int pp_off, pp_stride;  // primitive pairs
  int npp, inst;          // inst: molecule of the pair in a batch (0 otherwise)
  double sch;             // Schwarz bound of the parent pair (block maximum)
};

// primitive pair data, 64 B
struct __attribute__((aligned(16))) PrimPair {
  double p;          // alpha_a + alpha_b
  double Px, Py, Pz;
  double cc;         // c_a c_b exp(-alpha_a alpha_b AB^2 / p) / p
  double eb;         //         exp(-alpha_a alpha_b AB^2 / p) / p
  double hp;         // 1 / (2p)
  double ea;         // alpha_a (derivatives with respect to the centre of a)
};

// One bra group x ket group block of a launch.
struct Segment {
  int ket0;          // class-list index of the first ket pair of the ket group
  int nppb, nppk;    // primitive pairs of every bra / ket pair of the segment
  int inst;          // molecule of the segment in a batched launch (0 otherwise)
};

// Row r: bra pair rbra[r] (class-list index) with the ket pairs at positions
// [rlo[r], rhi[r]) of its segment's ket group.  A tile covers rows [r0, r1)
// (one bra shell a), ket positions [k0, k0 + 64) and the ket primitive pairs
// [v0, v1): the primitive sum of a highly contracted segment is split over
// several tiles (each digests its partial integrals; J and K are linear in
// them), so that no tile is much longer than the others.
struct Tile {
  int seg, r0, r1, k0, v0, v1;
};

// kStore writes the contracted integrals of each quartet into a dense
// pair-by-pair matrix (the MP2 transform's input) instead of digesting them.
enum Mode { kRHF = 0, kUHF = 1, kSchwarz = 2, kStore = 3 };

// Per-molecule data of a batched launch (one entry per molecule): its
// matrices, screening maxima, half-accumulators and, per ket pair class,
// its exchange staging and the range of its ket pairs in the class list.
struct InstDesc {
  const double* DJ;
  const double* DA;
  const double* DB;
  const double* PM;
  double* AJ;
  double* AKA;
  double* AKB;
  double* SA[6];
  double* SB[6];
  int np[6];          // ket pairs of the molecule in each pair class
  int p0[6];          // first of them in the class list
  int n, nref;
  double fixed_cut;
};

// Load of read-only work-list data through the constant address space: the
// compiler may then assume the memory is not written during the kernel, so a
// wavefront-uniform address (the bra rows and their primitive pairs, tiles,
// segments) becomes one scalar load instead of a per-lane vector load plus
// readfirstlane.  A lane-varying address still compiles to a vector load.
// The record is read as words (16, 8 or 4 bytes by its alignment), since an
// aggregate copy would be lowered back to global-memory loads.
template <class T>
__host__ __device__ __forceinline__ T uload(const T* p) {
#ifdef __HIP_DEVICE_COMPILE__
  using W = typename std::conditional<
      alignof(T) >= 16, unsigned __attribute__((ext_vector_type(4))),
      typename std::conditional<alignof(T) >= 8, unsigned __attribute__((ext_vector_type(2))), unsigned>::type>::type;
  static_assert(sizeof(T) % sizeof(W) == 0, "record size must be a multiple of its load word");
  constexpr int N = sizeof(T) / sizeof(W);
  const __attribute__((address_space(4))) W* q = (const __attribute__((address_space(4))) W*)p;
  W w[N];
#pragma unroll
  for (int i = 0; i < N; ++i) w[i] = q[i];
  T t;
  __builtin_memcpy(&t, w, sizeof(T));
  return t;
#else
  return *p;
#endif
}

struct ClassArgs {
  const SubShell* __restrict__ sh;
  const PairRec* __restrict__ bra;
  const PairRec* __restrict__ ket;
  const PrimPair* __restrict__ pp;
  const Segment* __restrict__ seg;    // segments of this launch
  const Tile* __restrict__ tile;      // tiles of this launch
  const int* __restrict__ rbra;       // rows: bra pair
  const int* __restrict__ rlo;        // rows: first ket position
  const int* __restrict__ rhi;        // rows: end of the ket positions
  const double* __restrict__ boys;    // Boys table (global)
  const double* __restrict__ DJ;      // density for J (PA+PB)
  const double* __restrict__ DA;      // alpha density for K
  const double* __restrict__ DB;      // beta density for K (UHF)
  const double* __restrict__ PM;      // per reference-shell-pair max |P|
  const double* __restrict__ bscale;  // per basis function component scale
  double* AJ;                         // J half-accumulator (n x n), wave kernel
  double* AKA;                        // K half-accumulators, wave kernel
  double* AKB;
  double* sch_out;                    // Schwarz mode output (n x n)
  // Pair-indexed digestion of the per-thread kernel (lanes hold consecutive
  // ket pairs, so these atomics are contiguous across a wavefront):
  double* JpB;                        // [NAB][bra pair] J contributions, bra class
  double* JpK;                        // [NCD][ket pair] J contributions, ket class
  double* SA;                         // [n][NC+ND][ket pair] K staging, alpha
  double* SB;                         // same, beta (UHF)
  int npb, npk;                       // pairs in the bra / ket class
  long long total;                    // tiles (RHF/UHF), pairs (Schwarz)
  double fixed_cut;                   // density-weighted screening cut
  int same, n, nref;
  // batched launch: per-molecule data (nullptr for one molecule) and the
  // ket pair class of the launch
  const InstDesc* inst;
  int ycls;
  // Hermite coefficient tables of the class's primitive pairs (dd launches of
  // one molecule; nullptr otherwise) and the primitive-pair index of entry 0
  const double* herm;
  int herm_base;
  // exchange staging of one molecule held in ket-pair chunks: SA/SB hold the
  // ket pairs [s_k0, s_k0 + s_np) of the class (s_np = 0: the whole class)
  int s_k0, s_np;
};

// The per-molecule part of the arguments, as one tile sees it: from the
// arguments themselves for a single molecule, from the tile's molecule
// entry in a batched launch (wavefront-uniform scalar loads).  S holds the
// exchange staging of the molecule's ket pairs [k0, k0 + npk) of the class.
struct View {
  const double* DJ;
  const double* DA;
  const double* DB;
  const double* PM;
  double* AJ;
  double* AKA;
  double* AKB;
  double* SA;
  double* SB;
  int n, nref, npk, k0;
  double fixed_cut;
};

__host__ __device__ __forceinline__ View make_view(const ClassArgs& A, int inst) {
  View v;
  if (A.inst == nullptr) {
    v.DJ = A.DJ; v.DA = A.DA; v.DB = A.DB; v.PM = A.PM;
    v.AJ = A.AJ; v.AKA = A.AKA; v.AKB = A.AKB; v.SA = A.SA; v.SB = A.SB;
    v.n = A.n; v.nref = A.nref; v.npk = A.s_np > 0 ? A.s_np : A.npk; v.k0 = A.s_k0; v.fixed_cut = A.fixed_cut;
  } else {
    const InstDesc* d = A.inst + inst;
    v.DJ = uload(&d->DJ); v.DA = uload(&d->DA); v.DB = uload(&d->DB); v.PM = uload(&d->PM);
    v.AJ = uload(&d->AJ); v.AKA = uload(&d->AKA); v.AKB = uload(&d->AKB);
    v.SA = uload(&d->SA[A.ycls]); v.SB = uload(&d->SB[A.ycls]);
    v.n = uload(&d->n); v.nref = uload(&d->nref); v.npk = uload(&d->np[A.ycls]); v.k0 = uload(&d->p0[A.ycls]);
    v.fixed_cut = uload(&d->fixed_cut);
  }
  return v;
}

using namespace sqmd;

// Integrals a lane holds (see SliceInfo) and the wavefronts per SIMD the
// kernel is compiled for, per class.  The default (36 integrals, no bound) has
// (dp|ps), (ds|ps) and (pp|pp) need a few registers more than 256 plus
// accumulation registers, one wavefront per SIMD; with 18 integrals ((dp|ps),
// (pp|pp): one ket component less per slice) or a bound of two wavefronts
// they run two wavefronts per SIMD without spills ((ds|ps): 8 bytes), and the
// second wavefront pays for the repeated R and Hermite work of the smaller
// slices ((dp|ps) 26.8 -> 25.2 ms at C60).  (dp|pp) at 18 measured 5 % slower.
template <int LA, int LB, int LC, int LD>
constexpr int slice_budget() {
  return ((LA == 2 && LB == 1 && LC == 1 && LD == 0) || (LA == 1 && LB == 1 && LC == 1 && LD == 1)) ? 18 : 36;
}
template <int LA, int LB, int LC, int LD>
constexpr int class_min_waves() {
  return ((LA == 2 && LB == 1 && LC == 1 && LD == 0) || (LA == 1 && LB == 1 && LC == 1 && LD == 1) ||
          (LA == 2 && LB == 0 && LC == 1 && LD == 0))
             ? 2
             : 1;
}

// Ket-component slices: a lane holds all NAB bra components for SZ ket
// components (at most slice_budget integrals, at least one ket component).
// Slicing the ket side keeps the W intermediates of different slices
// disjoint.
template <int LA, int LB, int LC, int LD>
struct SliceInfo {
  static constexpr int NAB = ncart(LA) * ncart(LB);
  static constexpr int NCD = ncart(LC) * ncart(LD);
  static constexpr int PER = (slice_budget<LA, LB, LC, LD>() / NAB) > 0 ? (slice_budget<LA, LB, LC, LD>() / NAB) : 1;
  static constexpr int NSLICE = (NCD + PER - 1) / PER;
  static constexpr int SZ = (NCD + NSLICE - 1) / NSLICE;
};

__host__ __device__ __forceinline__ void fatomic(double* p, double v) {
#ifdef __HIP_DEVICE_COMPILE__
  unsafeAtomicAdd(p, v);
#else
  *p += v;  // host instantiation (used only to check the kernel arithmetic on the CPU)
#endif
}

// Packed index of the basis-function pair (p, q), p >= q after ordering:
// p (p + 1) / 2 + q.
__host__ __device__ __forceinline__ size_t pair_index(int p, int q) {
  return p >= q ? (size_t)p * (p + 1) / 2 + q : (size_t)q * (q + 1) / 2 + p;
}

// Integral store of kStore: a plain store when the quartet's whole primitive
// sum is in hand, an atomic add of a partial sum otherwise (a segment whose
// ket primitive pairs are split over several tiles).
__host__ __device__ __forceinline__ void eri_put(double* base, size_t off, double v, bool whole) {
  if (whole)
    base[off] = v;
  else
    fatomic(base + off, v);
}

// Lane permutation of a double by a DPP control (both 32-bit halves).
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double read_lane(double v, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                          __builtin_amdgcn_readlane(__double2loint(v), l));
}

// Sum over the 64 lanes of the wavefront (all lanes active), returned to
// every lane: DPP butterflies inside each row of 16 lanes (quad_perm
// [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror, row_mirror), then the
// four row sums through scalar registers.  No LDS traffic.
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_mov<0xB1>(v);
  v += dpp_mov<0x4E>(v);
  v += dpp_mov<0x141>(v);
  v += dpp_mov<0x140>(v);
  return (read_lane(v, 0) + read_lane(v, 16)) + (read_lane(v, 32) + read_lane(v, 48));
}

// Ket c- and d-component ranges covered by the ket-component slice [CD0, CD1).
template <int ND, int CD0, int CD1>
struct KetRange {
  static constexpr int C0 = CD0 / ND, C1 = (CD1 - 1) / ND;
  static constexpr int NCr = C1 - C0 + 1;
  static constexpr int D0 = (NCr > 1) ? 0 : CD0 % ND;
  static constexpr int D1 = (NCr > 1) ? ND - 1 : (CD1 - 1) % ND;
  static constexpr int NDr = D1 - D0 + 1;
};

// Contributions a lane keeps in registers over a run of bra rows with the
// same shell a and the same ket pair: J_cd and the a-side exchange blocks
// (A_K[a,c], A_K[a,d]) per spin.
template <int LA, int LC, int LD, int CD0, int CD1>
struct RunAcc {
  static constexpr int NA = ncart(LA), ND = ncart(LD), NS = CD1 - CD0;
  using KR = KetRange<ND, CD0, CD1>;
  double jcd[NS];
  double kac[2][NA][KR::NCr];
  double kad[2][NA][KR::NDr];
  __host__ __device__ __forceinline__ void zero() {
    sfor<NS>([&](auto I) { jcd[I.value] = 0.0; });
    sfor<2>([&](auto S) {
      sfor<NA>([&](auto I) {
        sfor<KR::NCr>([&](auto J) { kac[S.value][I.value][J.value] = 0.0; });
        sfor<KR::NDr>([&](auto J) { kad[S.value][I.value][J.value] = 0.0; });
      });
    });
  }
};

// Contracted quartet (ib, ik) over the bra primitive pairs [0, nppb) and the
// ket primitive pairs [v0, v1), ket components [CD0, CD1).  R: precision of
// the integral evaluation (double, or float for the mixed-precision range);
// digestion and accumulation are FP64.  All lanes of a wavefront call this
// together; `valid` gates the lane's contributions, so cross-lane reductions
// see every lane.
// RUN: J_cd and the a-side exchange go to `run` instead of memory.
template <int LA, int LB, int LC, int LD, int CD0, int CD1, int MODE, typename R = double, bool RUN = false>
__host__ __device__ __forceinline__ void quartet(const ClassArgs& A, const View& V, int ib, int ik, int nppb, int v0,
                                                 int v1, bool valid, const double* lds,
                                                 RunAcc<LA, LC, LD, CD0, CD1>* run = nullptr) {
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  constexpr int LAB = LA + LB, L = LAB + LC + LD;
  constexpr int NAB = NA * NB, NS = CD1 - CD0;

  const PairRec bp = uload(A.bra + ib);
  const PairRec kp = A.ket[ik];
  const SubShell sA = uload(A.sh + bp.a), sB = uload(A.sh + bp.b), sC = A.sh[kp.a], sD = A.sh[kp.b];

  R acc[NAB][NS];
  sfor<NAB>([&](auto a) { sfor<NS>([&](auto c) { acc[a][c] = R(0); }); });

  // ket primitive outer (per-lane data, loaded once), bra primitive inner
  // (the same record for every lane of a bra-uniform wavefront)
  for (int v = v0; v < v1; ++v) {
    const PrimPair Q = A.pp[kp.pp_off + v * kp.pp_stride];
    const R h2q = R(Q.hp);
    Herm1<LC, LD, R> Fx, Fy, Fz;
    herm_coef<LC, LD, R>(Fx, R(Q.Px - sC.x), R(Q.Px - sD.x), h2q);
    herm_coef<LC, LD, R>(Fy, R(Q.Py - sC.y), R(Q.Py - sD.y), h2q);
    herm_coef<LC, LD, R>(Fz, R(Q.Pz - sC.z), R(Q.Pz - sD.z), h2q);
    for (int u = 0; u < nppb; ++u) {
      const PrimPair P = uload(A.pp + (bp.pp_off + u * bp.pp_stride));
      const double s = P.p + Q.p;
      const double rs = rsqrt_nr(s);
      if constexpr (MODE != kSchwarz) {
        if (kTwoPi25 * P.eb * Q.eb * rs < kPrimCutoff) continue;
      }
      const R alpha = R(P.p * Q.p * (rs * rs));
      const R X = R(P.Px - Q.Px), Y = R(P.Py - Q.Py), Z = R(P.Pz - Q.Pz);
      const R T = alpha * (X * X + Y * Y + Z * Z);
      R F[L + 1];
      boys<L, R>(T, F, lds);
      R Rt[nherm(L)];
      herm_R<L, R>(Rt, F, alpha, X, Y, Z);
      const R h2p = R(P.hp);
      Herm1<LA, LB, R> Ex, Ey, Ez;
      herm_coef<LA, LB, R>(Ex, R(P.Px - sA.x), R(P.Px - sB.x), h2p);
      herm_coef<LA, LB, R>(Ey, R(P.Py - sA.y), R(P.Py - sB.y), h2p);
      herm_coef<LA, LB, R>(Ez, R(P.Pz - sA.z), R(P.Pz - sB.z), h2p);
      const R pref = R(kTwoPi25 * P.cc * Q.cc * rs);

      sfor<NS>([&](auto csI) {
        constexpr int cs = decltype(csI)::value;
        constexpr int cd = CD0 + cs;
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
        sfor<NAB>([&](auto abI) {
          constexpr int ab = decltype(abI)::value;
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
          acc[ab][cs] += pref * val;
        });
      });
    }
  }

  const int n = V.n;
  const int fa = sA.bf_off, fb = sB.bf_off, fc = sC.bf_off, fd = sD.bf_off;
  const int ga = sA.gbf, gb = sB.gbf, gc = sC.gbf, gd = sD.gbf;  // component scale indices

  if constexpr (MODE == kSchwarz) {
    // diagonal (ab|ab) of the quartet (i,i): sqrt|.| into both (a,b), (b,a)
    sfor<NS>([&](auto csI) {
      constexpr int cs = decltype(csI)::value;
      constexpr int ab = CD0 + cs;
      constexpr int ia = ab / NB, jb = ab % NB;
      const double sc = A.bscale[ga + ia] * A.bscale[gb + jb];
      const double v = sqrt(fabs((double)acc[ab][cs] * sc * sc));
      A.sch_out[(fa + ia) * n + fb + jb] = v;
      A.sch_out[(fb + jb) * n + fa + ia] = v;
    });
    return;
  } else if constexpr (MODE == kStore) {
    // (pq|rs) into V[pair(p,q)][pair(r,s)] and V[pair(r,s)][pair(p,q)] of the
    // n(n+1)/2 x n(n+1)/2 matrix at sch_out; a pair of one sub-shell with
    // itself stores p >= q only, a quartet of one pair with itself one of the
    // two positions (its component loop visits both)
    if (!valid) return;
    const size_t ld = (size_t)n * (n + 1) / 2;
    const bool sb = bp.a == bp.b, sk = kp.a == kp.b, diag = A.same && ib == ik;
    const bool whole = v0 == 0 && v1 == kp.npp;
    sfor<NS>([&](auto csI) {
      constexpr int cs = decltype(csI)::value;
      constexpr int cd = CD0 + cs;
      const int r = fc + cd / ND, s = fd + cd % ND;
      if (sk && s > r) return;
      const double scd = A.bscale[gc + cd / ND] * A.bscale[gd + cd % ND];
      const size_t Q = pair_index(r, s);
      sfor<NAB>([&](auto abI) {
        constexpr int ab = decltype(abI)::value;
        const int p = fa + ab / NB, q = fb + ab % NB;
        if (sb && q > p) return;
        const size_t P = pair_index(p, q);
        const double v = (double)acc[ab][cs] * scd * A.bscale[ga + ab / NB] * A.bscale[gb + ab % NB];
        eri_put(A.sch_out, P * ld + Q, v, whole);
        if (!diag) eri_put(A.sch_out, Q * ld + P, v, whole);
      });
    });
    return;
  } else {
    double deg = valid ? 1.0 : 0.0;
    if (bp.a == bp.b) deg *= 0.5;
    if (kp.a == kp.b) deg *= 0.5;
    if (A.same && ib == ik) deg *= 0.5;

    double sc_a[NA], sc_b[NB];
    sfor<NA>([&](auto I) { sc_a[I.value] = A.bscale[ga + I.value]; });
    sfor<NB>([&](auto I) { sc_b[I.value] = A.bscale[gb + I.value]; });

    // scaled integrals, FP64 from here on
    double vv_[NAB][NS];
    sfor<NS>([&](auto csI) {
      constexpr int cs = decltype(csI)::value;
      constexpr int cd = CD0 + cs;
      const double scd = deg * A.bscale[gc + cd / ND] * A.bscale[gd + cd % ND];
      sfor<NAB>([&](auto abI) {
        constexpr int ab = decltype(abI)::value;
        vv_[ab][cs] = (double)acc[ab][cs] * scd * sc_a[ab / NB] * sc_b[ab % NB];
      });
    });

    // Coulomb: A_J[a,b] += 2 sum_cd v D[c,d];  A_J[c,d] += 2 sum_ab v D[a,b]
    double dcd[NS], jcd[NS], jab[NAB];
    sfor<NS>([&](auto csI) {
      constexpr int cd = CD0 + decltype(csI)::value;
      dcd[csI.value] = V.DJ[(fc + cd / ND) * n + fd + cd % ND];
      jcd[csI.value] = 0.0;
    });
    sfor<NAB>([&](auto abI) {
      constexpr int ab = decltype(abI)::value;
      const double dab = V.DJ[(fa + ab / NB) * n + fb + ab % NB];
      double s = 0.0;
      sfor<NS>([&](auto csI) {
        constexpr int cs = decltype(csI)::value;
        s += vv_[ab][cs] * dcd[cs];
        jcd[cs] += vv_[ab][cs] * dab;
      });
      jab[ab] = s;
    });
    bool bra_uniform = false;
#ifdef __HIP_DEVICE_COMPILE__
    // all lanes on one bra pair (the common case): one atomic per J_ab element
    bra_uniform = __all(ib == __builtin_amdgcn_readfirstlane(ib));
    if (bra_uniform) {
      const bool lead = (threadIdx.x & 63) == 0;
      sfor<NAB>([&](auto abI) {
        constexpr int ab = decltype(abI)::value;
        const double s = wave_sum(jab[ab]);
        if (lead) fatomic(&A.JpB[(size_t)ab * A.npb + ib], 2.0 * s);
      });
    }
#endif
    if (valid && !bra_uniform)
      sfor<NAB>([&](auto abI) {
        constexpr int ab = decltype(abI)::value;
        fatomic(&A.JpB[(size_t)ab * A.npb + ib], 2.0 * jab[ab]);
      });
    if constexpr (RUN) {
      sfor<NS>([&](auto csI) { run->jcd[csI.value] += jcd[csI.value]; });  // zero when !valid
    } else if (valid) {
      sfor<NS>([&](auto csI) {
        constexpr int cd = CD0 + decltype(csI)::value;
        fatomic(&A.JpK[(size_t)cd * A.npk + ik], 2.0 * jcd[csI.value]);  // class-list index
      });
    }

    // Exchange (per spin): A_K[a,c] += v D[b,d]; A_K[b,c] += v D[a,d];
    //                      A_K[a,d] += v D[b,c]; A_K[b,d] += v D[a,c]
    // staged as S[row][comp][ket pair] with comp = c component, or NC + d
    // component; a later pass adds S into the rows of A_K.
    // exchange staging of the molecule: its ket pairs [k0, k0 + npk)
    constexpr int NCOMP = NC + ND;
    const size_t npk = (size_t)V.npk;
    const int jk = ik - V.k0;
    auto exchange = [&](const double* __restrict__ D, double* S, auto spinI) {
      using KR = KetRange<ND, CD0, CD1>;
      constexpr int C0 = KR::C0, NCr = KR::NCr, D0 = KR::D0, NDr = KR::NDr;
      double kac[NA][NCr], kad[NA][NDr], kbc[NB][NCr], kbd[NB][NDr];
      sfor<NA>([&](auto I) { sfor<NCr>([&](auto J) { kac[I.value][J.value] = 0.0; });
                             sfor<NDr>([&](auto J) { kad[I.value][J.value] = 0.0; }); });
      sfor<NB>([&](auto I) { sfor<NCr>([&](auto J) { kbc[I.value][J.value] = 0.0; });
                             sfor<NDr>([&](auto J) { kbd[I.value][J.value] = 0.0; }); });
      sfor<NS>([&](auto csI) {
        constexpr int cs = decltype(csI)::value;
        constexpr int cd = CD0 + cs;
        constexpr int ic = cd / ND, id = cd % ND;
        const int gc = fc + ic, gd = fd + id;
        sfor<NAB>([&](auto abI) {
          constexpr int ab = decltype(abI)::value;
          constexpr int ia = ab / NB, jb = ab % NB;
          const int ga = fa + ia, gb = fb + jb;
          const double v = vv_[ab][cs];
          kac[ia][ic - C0] += v * D[gb * n + gd];
          kbc[jb][ic - C0] += v * D[ga * n + gd];
          kad[ia][id - D0] += v * D[gb * n + gc];
          kbd[jb][id - D0] += v * D[ga * n + gc];
        });
      });
      if constexpr (RUN) {
        constexpr int sp = decltype(spinI)::value;
        sfor<NA>([&](auto I) {
          sfor<NCr>([&](auto J) { run->kac[sp][I.value][J.value] += kac[I.value][J.value]; });
          sfor<NDr>([&](auto J) { run->kad[sp][I.value][J.value] += kad[I.value][J.value]; });
        });
      }
      if (!valid) return;
      if constexpr (!RUN) {
        sfor<NA>([&](auto I) {
          const size_t row = (size_t)(fa + I.value) * NCOMP;
          sfor<NCr>([&](auto J) { fatomic(&S[(row + C0 + J.value) * npk + jk], kac[I.value][J.value]); });
          sfor<NDr>([&](auto J) { fatomic(&S[(row + NC + D0 + J.value) * npk + jk], kad[I.value][J.value]); });
        });
      }
      sfor<NB>([&](auto I) {
        const size_t row = (size_t)(fb + I.value) * NCOMP;
        sfor<NCr>([&](auto J) { fatomic(&S[(row + C0 + J.value) * npk + jk], kbc[I.value][J.value]); });
        sfor<NDr>([&](auto J) { fatomic(&S[(row + NC + D0 + J.value) * npk + jk], kbd[I.value][J.value]); });
      });
    };
    exchange(V.DA, V.SA, ic<0>{});
    if constexpr (MODE == kUHF) exchange(V.DB, V.SB, ic<1>{});
  }
}

// Writes a run's register contributions (ket pair ik, bra shell starting at
// basis function fa) to the staging buffers.
template <int LA, int LB, int LC, int LD, int CD0, int CD1, int MODE>
__host__ __device__ __forceinline__ void run_flush(const ClassArgs& A, const View& V,
                                                   const RunAcc<LA, LC, LD, CD0, CD1>& run, int fa, int ik) {
  constexpr int NA = ncart(LA), NC = ncart(LC), ND = ncart(LD), NCOMP = NC + ND;
  using KR = KetRange<ND, CD0, CD1>;
  const size_t npk = (size_t)V.npk;
  const int jk = ik - V.k0;
  sfor<CD1 - CD0>([&](auto csI) {
    constexpr int cd = CD0 + decltype(csI)::value;
    fatomic(&A.JpK[(size_t)cd * A.npk + ik], 2.0 * run.jcd[csI.value]);
  });
  sfor<(MODE == kUHF) ? 2 : 1>([&](auto spI) {
    constexpr int sp = decltype(spI)::value;
    double* S = sp ? V.SB : V.SA;
    sfor<NA>([&](auto I) {
      const size_t row = (size_t)(fa + I.value) * NCOMP;
      sfor<KR::NCr>([&](auto J) { fatomic(&S[(row + KR::C0 + J.value) * npk + jk], run.kac[sp][I.value][J.value]); });
      sfor<KR::NDr>([&](auto J) {
        fatomic(&S[(row + NC + KR::D0 + J.value) * npk + jk], run.kad[sp][I.value][J.value]);
      });
    });
  });
}

// Density-weighted screening at the reference-shell level (the reference's
// six-block sum and test, src/matrix.c:3921-3924, 3936, 3943).
__host__ __device__ __forceinline__ bool density_screen_pass(const View& V, const PairRec& bp, const PairRec& kp) {
  const double* PM = V.PM;
  const int nr = V.nref;
  const double mp = uload(PM + (bp.pa * nr + bp.pb)) + PM[kp.pa * nr + kp.pb] + PM[bp.pa * nr + kp.pa] +
                    PM[bp.pa * nr + kp.pb] + PM[bp.pb * nr + kp.pa] + PM[bp.pb * nr + kp.pb];
  return !(bp.sch * kp.sch * mp < V.fixed_cut);
}

// One launch per (class pair, precision, mode); blockIdx.y selects the
// ket-component slice (a compile-time branch per slice, uniform in the
// workgroup), so the slices of a class fill the device together.
// RHF/UHF: grid-stride over wavefronts; a wavefront takes 64 consecutive
// slots of one segment.  Schwarz: grid-stride over the pairs of the class,
// diagonal quartets.
// BS: workgroup size; the LDS Boys table is shared by the BS/64 wavefronts of
// a workgroup, so classes with few registers per lane use large workgroups.
template <int LA, int LB, int LC, int LD, int MODE, typename R = double, int BS = 256>
__global__ void __launch_bounds__(BS, (class_min_waves<LA, LB, LC, LD>())) k_class(ClassArgs A) {
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
    if constexpr (MODE == kSchwarz) {
      const long long stride = (long long)gridDim.x * blockDim.x;
      for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < A.total; t += stride) {
        const int npp = A.bra[t].npp;
        const View V = make_view(A, 0);
        quartet<LA, LB, LC, LD, CD0, CD1, MODE, R>(A, V, (int)t, (int)t, npp, 0, npp, true, lds);
      }
    } else {
      const int lane = threadIdx.x & 63;
      const int wpb = blockDim.x >> 6;
      const long long wstride = (long long)gridDim.x * wpb;
      const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
      for (long long w = (long long)blockIdx.x * wpb + wib; w < A.total; w += wstride) {
        const Tile tl = uload(A.tile + w);
        const Segment sg = uload(A.seg + tl.seg);
        const View V = make_view(A, sg.inst);
        const int j = tl.k0 + lane;  // ket position in the group
        RunAcc<LA, LC, LD, CD0, CD1> run;
        run.zero();
        bool any = false;
        for (int r = tl.r0; r < tl.r1; ++r) {
          const int ib = uload(A.rbra + r);
          const int lo = uload(A.rlo + r);
          const bool in = j >= lo && j < uload(A.rhi + r);
          const int ik = sg.ket0 + (in ? j : lo);  // idle lanes evaluate a valid pair, weight 0
          any = any || in;
          const bool valid = in && density_screen_pass(V, uload(A.bra + ib), A.ket[ik]);
          if (!__any(valid)) continue;  // the whole wavefront screened out: skip the row (uniform)
          quartet<LA, LB, LC, LD, CD0, CD1, MODE, R, true>(A, V, ib, ik, sg.nppb, tl.v0, tl.v1, valid, lds, &run);
        }
        if (any)
          run_flush<LA, LB, LC, LD, CD0, CD1, MODE>(A, V, run, uload(A.sh + (uload(A.bra + (uload(A.rbra + tl.r0))).a)).bf_off,
                                                    sg.ket0 + j);
      }
    }
  });
}

// Workgroup size of the per-thread kernel: 512 lanes (8 wavefronts share one
// LDS Boys table) for classes holding at most 9 integrals per lane, else 256.
// These kernels keep about 110 SGPRs per wavefront, so a SIMD holds at most 7
// wavefronts: with 512-lane workgroups (2 wavefronts per SIMD each) a CU
// keeps 3 workgroups, 6 wavefronts per SIMD, where 1024-lane workgroups
// would leave it at one workgroup and 4.
template <int LA, int LB, int LC, int LD>
constexpr int class_block() {
  using SI = SliceInfo<LA, LB, LC, LD>;
  return SI::NAB * SI::SZ <= 9 ? 512 : 256;
}

template <int LA, int LB, int LC, int LD>
constexpr size_t class_lds_bytes() {
  return sizeof(double) * boys_rows<LA + LB + LC + LD>() * kBoysLdsW;
}

}  // namespace sqjk
