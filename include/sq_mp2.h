/*
 * sq_mp2.h - closed-shell MP2 correlation energy on the device.
 *
 * Drop-ins for src/mp2.h (mp2_rhf_aqij_Parallel, mp2_rhf_aqij,
 * mp2_uhf_aqij, getFrozenCore_RHF) and the context-level entry points under them.  The
 * integrals come from the J/K context's pair lists and work plan (every
 * shell quartet above the Schwarz cut-off, no density screening), written
 * into a dense matrix over basis-function pairs; the four-index transform is
 * DGEMMs on the matrix cores; the energy sum is one device pass.  Device
 * memory: (n(n+1)/2)^2 + n(n+1)/2 nc nv doubles at once (n basis functions,
 * nc correlated occupied and nv virtual orbitals), e.g. 22.5 GB for
 * (H2O)16/6-31G* (n = 304).  Cartesian s, p, d only, as the J/K path.
 */
#ifndef SQ_MP2_H
#define SQ_MP2_H

#include "sq_jk.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  double e2;                      /* MP2 correlation energy (Hartree) */
  double store_ms;                /* integral evaluation into the pair matrix */
  double half1_ms, half2_ms;      /* first / second half-transform (with the energy sum) */
  double half1_flop, half2_flop;  /* DGEMM flops of the two halves */
  long long quartets;             /* shell quartets evaluated */
  long long device_bytes;         /* pair matrix + half-transformed integrals */
} sqjk_mp2_info;

/* The two-electron integrals of every shell quartet with Schwarz_ab Schwarz_cd
 * >= cutoff as the dense symmetric matrix V[pair(p,q) * m + pair(r,s)] =
 * (pq|rs), m = n(n+1)/2, pair(p,q) = p(p+1)/2 + q for p >= q (host array of
 * m*m doubles; screened quartets are 0).  The integral generator the MP2
 * path runs on; it replaces the (pq|ij) evaluation of mp2_gen_aqij_Quartet
 * (src/matrix.c:4864-5073) before its first quarter-transform. */
int sqjk_eri_pairs(sqjk_ctx *ctx, double cutoff, double *V);

/* MP2 correlation energy of a closed-shell determinant (Szabo-Ostlund 6.74):
 *   E2 = sum_{a,b} sum_{r,s} (ar|bs) [2 (ar|bs) - (as|br)] / (e_a + e_b - e_r - e_s)
 * over the occupied orbitals nCore <= a, b < nOcc and the virtual orbitals
 * nOcc <= r, s < n.  C: orbital-major coefficients C[k*n + i] (n x n, host),
 * e: orbital energies (n).  Shell quartets with Schwarz_ab Schwarz_cd below
 * cutoff are left out (0: every quartet).  Replaces mp2_rhf_aqij_Parallel's
 * work (src/mp2.c:581-927) for all correlated orbitals at once.  info may be
 * NULL; its e2 is also the result.  SQJK_ENOMEM with the sizes when the
 * device cannot hold the pair matrix and the half-transformed integrals. */
int sqjk_mp2_rhf(sqjk_ctx *ctx, int nOcc, int nCore, const double *e, const double *C, double cutoff,
                 sqjk_mp2_info *info);

/* Unrestricted MP2 correlation energy (src/mp2.c:2126-2432):
 *   E2 = sum_{a<b, r<s} [(ar|bs) - (as|br)]^2 / D  over alpha orbitals
 *      + the same over beta orbitals
 *      + sum_{a, r alpha; b, s beta} (ar|bs)^2 / D,   D = e_a + e_b - e_r - e_s,
 * with the first nCore orbitals of each spin frozen.  CA, CB orbital-major,
 * eA, eB their energies.  The pair matrix is evaluated once for both spins;
 * device memory (n(n+1)/2)^2 + n(n+1)/2 (ncA nvA + ncB nvB) doubles. */
int sqjk_mp2_uhf(sqjk_ctx *ctx, int nEA, int nEB, int nCore, const double *eA, const double *eB, const double *CA,
                 const double *CB, double cutoff, sqjk_mp2_info *info);

/* ---- drop-in surface (reference signatures, src/mp2.h) ---- */

/* Frozen-core orbitals: half the noble-gas core electrons of every atom
 * (src/mp2.c:29-67); exits for Z > 54, as the reference. */
int getFrozenCore_RHF(const struct Molecule_t *mol);

/* src/mp2.c:713-927: frozen-core RHF MP2 energy correction, printed in the
 * reference's report format and returned.  All correlated orbitals in one
 * pass on the device (opt->nCPU and opt->maxMem, which the reference uses to
 * cut the orbitals into cycles, are not needed); shell quartets below the
 * Schwarz product 1e-10 are left out.  Exits like the reference for an
 * unrestricted option set, and with the library's message on device errors. */
double mp2_rhf_aqij_Parallel(int nBasis, int nOcc, const double *e, const double *C,
                             const struct GTOBasis_t *gto, const struct Molecule_t *mol,
                             const struct option_t *opt);

/* src/mp2.c:344: the serial variant, the same result. */
double mp2_rhf_aqij(int nBasis, int nOcc, const double *e, const double *C, const struct GTOBasis_t *gto,
                    const struct Molecule_t *mol, const struct option_t *opt);

/* src/mp2.c:2126-2432: frozen-core unrestricted MP2 energy correction,
 * printed in the reference's report format and returned; both spins in one
 * device pass.  Shell quartets below the Schwarz product 1e-10 are left out. */
double mp2_uhf_aqij(int nBasis, int nEA, int nEB, const double *eA, const double *eB, const double *CA,
                    const double *CB, const struct GTOBasis_t *gto, const struct Molecule_t *mol,
                    const struct option_t *opt);

#ifdef __cplusplus
}
#endif

#endif
