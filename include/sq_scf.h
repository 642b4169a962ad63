/*
 * sq_scf.h - host-side drop-ins around the J/K engine: SCF driver, density
 * DIIS, eigen-solvers, one-electron integrals, Boys function and input
 * readers.  Each declaration keeps the reference signature it replaces.
 * The J/K builds inside uhf()/rhf() and the eigen-solves run on the HIP
 * device; the one-electron integrals and readers are host C, as in the
 * reference.
 */
#ifndef SQ_SCF_H
#define SQ_SCF_H

#include <stdio.h>
#include "sq_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- SCF (src/uhf.h:93-103, src/uhf.c:1171-1964) ---- */
double uhf(int nBasis, struct GTOBasis_t *gto, struct Molecule_t *mol, int nEA, int nEB,
           double *CA, double *CB, double *eA, double *eB, struct option_t *opt);
/* src/uhf.h:24 (src/uhf.c:77-90): P = sum_{k<nOcc} C_k C_k^T */
void uhf_getDMatrix(int nBasis, int nOcc, double *C, double *P);

/* ---- legacy RHF (src/rhf.h:26-47, src/rhf.c:59-72, 286-548) ---- */
void rhf_getDMatrix(int nBasis, int nOcc, double *C, double *P);
double rhf(int nBasis, struct GTOBasis_t *gto, struct Molecule_t *mol, int nE,
           double *C, double *e, struct option_t *opt);

/* ---- convergence (src/conv.h:22-37, src/conv.c:40-844) ----
 * Density damping and 2-, 3-, 4-point density DIIS; each keeps its own
 * history, and a call (0, 0.0, NULL, NULL) resets it. */
double conv_damping(int nBasis, double alpha, double *PA, double *PB);
double conv_diis2(int nBasis, double alpha, double *PA, double *PB);
double conv_diis3(int nBasis, double alpha, double *PA, double *PB);
double conv_diis4(int nBasis, double alpha, double *PA, double *PB);

/* ---- SCF with its state on the device ----
 * The iteration of uhf() (src/uhf.c:1464-1783) with H, S, the densities,
 * the accumulated J/K, the Fock matrices, the DIIS history (any
 * opt->convMethod) and the orbitals held on the HIP device; per iteration
 * only the energy, the rms density change and the DIIS scalars reach the
 * host.  uhf() runs on it.  A context keeps its state across solves, so a
 * run of geometry steps reuses the J/K context (sqjk_update_geometry) and
 * can start from the previous orbitals (the CACB guess of src/qmd.c:585).
 * No counterpart in the reference, whose state is host memory. */
struct sqjk_ctx;
typedef struct sqscf_ctx sqscf_ctx;
typedef struct {
	int restricted, nEA, nEB;
	double SCFConv;
	int SCFMax;
	double SCFDrag;
	int threeStep;          /* 3-step cut-offs 1e-3/1e-6/1e-9 (SCFACCURACY_3STEP) */
	int convMethod;         /* CONVMETHOD_DAMPING / DIIS2 / DIIS3 / DIIS4 */
	double Econst;          /* nuclear repulsion and field terms of the energy */
	double mixedThreshold;  /* FP32 below this Schwarz product (0: FP64) */
	int eigenEveryIter;     /* orbitals every iteration (else only at the end) */
	/* called after every iteration (may read sqscf_orbitals when
	 * eigenEveryIter is set); NULL for none */
	void (*onIter)(void *user, sqscf_ctx *ctx, int iter, double E, double dE, double rmsDP);
	void *user;
} sqscf_params;
typedef struct {
	double energy, dE, rmsDP, realCutoff, seconds;
	int iterations, converged;
} sqscf_result;
int sqscf_create(int device, sqscf_ctx **out);
/* H = T + V (+ field) and S of the basis, host arrays.  The same basis at a
 * new geometry keeps the J/K context. */
int sqscf_set_system(sqscf_ctx *ctx, int nBasis, const struct GTOBasis_t *gto, const double *H, const double *S);
int sqscf_set_density(sqscf_ctx *ctx, const double *PA, const double *PB);
/* Initial density from the orbitals of the last solve, S-normalised in the
 * current S (SCFGUESS_CACB, src/uhf.c:1358-1364). */
int sqscf_guess_from_orbitals(sqscf_ctx *ctx, int nEA, int nEB);
int sqscf_run(sqscf_ctx *ctx, const sqscf_params *prm, sqscf_result *res);
int sqscf_orbitals(sqscf_ctx *ctx, double *CA, double *CB, double *eA, double *eB);
int sqscf_density(sqscf_ctx *ctx, double *PA, double *PB);
int sqscf_jk_context(sqscf_ctx *ctx, struct sqjk_ctx **jk);
void sqscf_destroy(sqscf_ctx *ctx);
const char *sqscf_errmsg(void);

/* ---- dense linear algebra (src/lin.h:20-33) ---- */
int gen_sym_eigen(int nDim, const double *A, const double *S, double *e, double *C);
int gen_sym_SomeEigen(int nDim, int nEigen, const double *A, const double *S, double *e, double *C);
int linear_solver(int nRow, int nCol, const double *A, double *B);
/* Device eigen-solver behind the two calls above; 0 on success. */
int sqla_gen_eigen(int n, const double *F, const double *S, double *e, double *C);
/* P = sum_{k<nOcc} C_k C_k^T (C orbital-major) as one device DGEMM; 0 on success. */
int sqla_density(int n, int nOcc, const double *C, double *P);
/* Density of the lowest nOcc solutions of F C = S C e straight from F
 * (spectral projection on the device, no eigenvectors).  Returns 0 on
 * success, 1 when the projection does not settle (no gap at nOcc; P is then
 * untouched and the caller solves the eigenproblem), < 0 on device errors.
 * No counterpart in the reference, whose SCF forms P from the orbitals of
 * gen_sym_SomeEigen (src/uhf.c:1703-1710); uhf() uses it for the same P. */
int sqla_density_sp2(int n, int nOcc, const double *F, const double *S, double *P);

/* ---- Boys function (src/fgamma.h:19-24) ---- */
double fgamma_0(double x);
double fgamma(int m, double x);
void fgamma_set(int max, double x, double *fset);

/* ---- primitive one-electron integrals (src/int.h:20-48) ---- */
double overlap(double alpha1, int l1, int m1, int n1, double xa, double ya, double za,
               double alpha2, int l2, int m2, int n2, double xb, double yb, double zb);
double moment(double alpha1, int l1, int m1, int n1, double xa, double ya, double za,
              double alpha2, int l2, int m2, int n2, double xb, double yb, double zb,
              int mx, int my, int mz, double xc, double yc, double zc);
double kinetic(double alpha1, int l1, int m1, int n1, double xa, double ya, double za,
               double alpha2, int l2, int m2, int n2, double xb, double yb, double zb);
double nai(double x1, double y1, double z1, double norm1, int l1, int m1, int n1, double alpha1,
           double x2, double y2, double z2, double norm2, int l2, int m2, int n2, double alpha2,
           double x3, double y3, double z3);
/* primitive ERI with norms (src/int.h:64-73) and without (src/int.h:75-79;
 * the reference's primeERI omits the 2 pi^{5/2} K_ab K_cd/(pq) factor) */
double eri(double xa, double ya, double za, double norma, int la, int ma, int na, double alphaa,
           double xb, double yb, double zb, double normb, int lb, int mb, int nb, double alphab,
           double xc, double yc, double zc, double normc, int lc, int mc, int nc, double alphac,
           double xd, double yd, double zd, double normd, int ld, int md, int nd, double alphad);

/* ---- contracted one-electron matrix elements (src/matrix.h:29-46) ---- */
double GTO_overlap(int i, int j, const struct GTOBasis_t *gto);
double GTO_moment(int i, int j, const struct GTOBasis_t *gto, int mx, int my, int mz,
                  double xc, double yc, double zc);
double GTO_kinetic(int i, int j, const struct GTOBasis_t *gto);
double GTO_nuclei(int i, int j, const struct GTOBasis_t *gto, const struct Molecule_t *mol);

/* ---- input (src/mol.h:31-37, src/basis.h:57-73) ---- */
struct Molecule_t *readMolecule_XYZ(FILE *inFile);
int get_nElectron(struct Molecule_t *mol);
double nuclei_coulomb(const struct Molecule_t *mol);
int get_nEA(struct Molecule_t *mol, int M);
int get_nEB(struct Molecule_t *mol, int M);
struct GTOBasisSet_t *read_GAMESS_BasisSet(FILE *inFile, char *name, int *nBasisSet);
/* ---- checkpoint and density files (src/check.h, src/basis.h) ----
 * Same text formats as the reference, so files move between the two. */
void save_checkpoint(int nBasis, struct GTOBasis_t *gto, struct Molecule_t *mol, double Etot, double dE,
                     double rmsDP, double *CA, double *CB, double *eA, double *eB,
                     struct option_t *opt);                                   /* src/check.c:432 */
struct Molecule_t *load_checkpoint_molecule(const struct option_t *opt);      /* src/check.c:37 */
int load_checkpoint_multiplicity(const struct option_t *opt);                 /* src/check.c:136 */
struct GTOBasis_t *load_checkpoint_basis(const struct option_t *opt, int *nBasis); /* src/check.c:175 */
double load_checkpoint_orbital(int nBasis, double *CA, double *CB, double *eA, double *eB,
                               const struct option_t *opt);                   /* src/check.c:210 */
void guess_checkpoint_orbital(int nBasis, const struct GTOBasis_t *gto, const struct Molecule_t *mol,
                              const struct option_t *opt, const double *S, double *CA,
                              double *CB);                                    /* src/check.c:306 */
int printGTOBasis(FILE *out, int nBasis, struct GTOBasis_t *gto);             /* src/basis.c:37 */
struct GTOBasis_t *readGTOInput(FILE *in, int *nBasis);                       /* src/basis.c:84 */
/* -LDMATRIX / -SVDMATRIX files: n*n values (alpha), then n*n (beta) unless
 * restricted, whitespace separated (src/uhf.c:1414-1462, 1843-1870). */
void sq_load_density(int n, int restricted, const char *path, double *PA, double *PB);
void sq_save_density(int n, const char *path, const double *PA, const double *PB);

struct GTOBasis_t *genBasis(struct Molecule_t *mol, int *nBasis, int dbSize,
                            const struct GTOBasisSet_t *basisDB);

#ifdef __cplusplus
}
#endif
#endif /* SQ_SCF_H */
