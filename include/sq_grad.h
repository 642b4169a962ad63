/*
 * sq_grad.h - energy gradient (forces) of Hartree-Fock on the device.
 *
 * Drop-ins for src/grad.h (uhf_force, GradEri_ShellSet_Parallel,
 * printForce) and the context-level entry points under them.  The
 * two-electron part runs on the J/K context's pair lists and work plan; the
 * one-electron part (overlap against the energy-weighted density, kinetic
 * and attraction, basis-function and operator derivatives) runs on the
 * device pair by pair.  Cartesian s, p, d only (SQJK_EBASIS above d), as the
 * J/K path.
 */
#ifndef SQ_GRAD_H
#define SQ_GRAD_H

#include "sq_jk.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Two-electron energy gradient dE2/dR per atom, ADDED into G[3*atom + xyz]:
 *   E2 = sum_pqij [PT_pq PT_ij / 2 - Kfactor (PA_pi PA_qj + PB_pi PB_qj) / 2] (pq|ij)
 * for the alpha and beta densities PA, PB (host, n x n; RHF: PA = PB = P/2).
 * Shell quartets with Schwarz_ab Schwarz_cd (sum of the six |P| block maxima)
 * below cutoff are skipped; the reference skips Schwarz_ab Schwarz_cd PM_ab
 * PM_cd below GRAD_SCHWARZ_CUTOFF = 1e-10 (src/grad.c:1861-1862), a subset.
 * basis2Atom[p] is the atom of basis function p.  Replaces the work of
 * GradEri_ShellSet_Parallel (src/grad.c:1734-2090) for all childIDs at once. */
int sqjk_grad_eri(sqjk_ctx *ctx, const double *PA, const double *PB, double Kfactor, double cutoff,
                  const int *basis2Atom, int nAtom, double *G);

/* One-electron energy gradient per atom, ADDED into G[3*atom + xyz]:
 * d/dR of sum_pq PT_pq (T + V)_pq - sum_pq QT_pq S_pq with QT the
 * energy-weighted density sum_i e_i C_i C_i^T (src/grad.c:2236-2306: the
 * loops over GTO_grad_overlap, GTO_grad_kinetic, GTO_grad_nai and GTO_efi). */
int sqjk_grad_one_electron(int nBasis, const struct GTOBasis_t *gto, const struct Molecule_t *mol,
                           const double *PT, const double *QT, const int *basis2Atom, int device, double *G);

/* ---- drop-in surface (reference signatures, src/grad.h) ---- */

/* src/grad.h:21-24: prints the force table. */
void printForce(const struct Molecule_t *mol, const double *Fx, const double *Fy, const double *Fz);

/* src/grad.h:26-39 (src/grad.c:2155-2499): Hartree-Fock forces -dE/dR per
 * atom into Fx, Fy, Fz (any may be NULL), from the orbitals CA, CB (one per
 * row) and orbital energies of a converged SCF; uniform field opt->Ex/Ey/Ez
 * included.  Methods other than Hartree-Fock (DFT exchange-correlation
 * forces) are outside this library: it prints an error and exits, as the
 * reference does for unsupported options. */
void uhf_force(int nBasis, struct GTOBasis_t *gto, struct Molecule_t *mol, int nEA, int nEB,
               double *CA, double *CB, double *eA, double *eB, struct option_t *opt,
               double *Fx, double *Fy, double *Fz);

/* src/grad.h:41-52: two-electron gradient per atom, ADDED into Gx, Gy, Gz,
 * for the total and spin densities PT, PS and exchange weight Kfactor.  All
 * work is done by childID 0 (cut-off GRAD_SCHWARZ_CUTOFF = 1e-10); other
 * childIDs add nothing, so a caller's sum over children is unchanged. */
void GradEri_ShellSet_Parallel(int childID, int nCPU, int nBasis, const struct GTOBasis_t *gto,
                               const int *basis2Atom, const double *PT, const double *PS, const double Kfactor,
                               double *Gx, double *Gy, double *Gz);

#ifdef __cplusplus
}
#endif
#endif /* SQ_GRAD_H */
