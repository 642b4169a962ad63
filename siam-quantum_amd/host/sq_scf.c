/*
 * sq_scf.c - the SCF drivers behind src/uhf.h and src/rhf.h, and the
 * density DIIS of src/conv.h.
 *
 * uhf() keeps the Hartree-Fock semantics of src/uhf.c:1171-1783:
 *   realCutoff = SCFConv/(sA^2+sA sB+sB^2)/50 with sX = sum of all P_X
 *   elements; three-step cut-offs 1e-3/1e-6/1e-9 (or one step); J/K built
 *   on the density change dP and accumulated; F = H + J + K; lowest
 *   eigenpairs; E = E_NN + sum (H + J/2)(PA+PB) + 1/2 sum (KA PA + KB PB);
 *   density DIIS (four points after three damped steps); converged when
 *   |dE| <= SCFConv and rms dP <= SCFConv; 0.0 returned if not converged.
 * The guess and the one-electron set-up are prepared here; the iteration
 * itself runs with all of its matrices on the device (sqscf_*, in
 * csrc/sq_scf_dev.hip), any opt->convMethod (damping, DIIS2/3/4).
 * The one difference from the reference is that every J/K build uses exact
 * integrals: the reference's monopole substitution for far quartets
 * (src/matrix.c:3946-3949) is not applied.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../../include/sq_jk.h"
#include "../../include/sq_scf.h"

#define ALLOC(p, n)                                              \
	do {                                                         \
		(p) = calloc((size_t)(n), sizeof(double));               \
		if (!(p)) {                                              \
			printf("uhf: Error - Cannot allocate memory\n");     \
			exit(EXIT_FAILURE);                                  \
		}                                                        \
	} while (0)

/* P = sum_{k < nOcc} C_k C_k^T as rank-one updates over blocks of 16 rows
 * (the block stays in cache while the orbitals stream past; each element
 * sums its terms in increasing k, as the plain triple loop does) */
void uhf_getDMatrix(int nBasis, int nOcc, double *C, double *P)
{
	const int n = nBasis;
	for (int i0 = 0; i0 < n; i0 += 16) {
		const int i1 = i0 + 16 < n ? i0 + 16 : n;
		memset(P + (size_t)i0 * n, 0, sizeof(double) * (size_t)(i1 - i0) * n);
		for (int k = 0; k < nOcc; k++) {
			const double *c = C + (size_t)k * n;
			for (int i = i0; i < i1; i++) {
				const double ci = c[i];
				double *row = P + (size_t)i * n;
				for (int j = 0; j < n; j++) row[j] += ci * c[j];
			}
		}
	}
}

void rhf_getDMatrix(int nBasis, int nOcc, double *C, double *P)
{
	uhf_getDMatrix(nBasis, nOcc, C, P);
	for (int e = 0; e < nBasis * nBasis; e++) P[e] *= 2.0;
}

/* S-normalise the first nOcc orbitals (src/uhf.c:48-64) */
static void normalize_orbitals(int n, int nOcc, const double *S, double *C)
{
	for (int k = 0; k < nOcc; k++) {
		double s = 0.0;
		for (int i = 0; i < n; i++)
			for (int j = 0; j < n; j++) s += C[k * n + i] * C[k * n + j] * S[i * n + j];
		s = 1.0 / sqrt(s);
		for (int i = 0; i < n; i++) C[k * n + i] *= s;
	}
}

/* ---------------------------------------------------------------- DIIS */
/* Density damping and K-point density DIIS (src/conv.c:40-844).  Each
 * method keeps its own history, as the reference keeps its own statics:
 * inputs P[0..K-1] (P[K-1] the newest stored density) and errors
 * dP[0..K-1].  A call stores the error cur - P[K-1] in dP[K-1], damps
 * (the first K-1 calls, every call for K = 1) or extrapolates with
 * P[a] + dP[a], a < K-1, and cur, then shifts both lists by one and stores
 * the result as the newest input. */
struct diis_hist {
	int it, n;
	double *P[4][2], *dP[4][2];
};
static struct diis_hist g_damp, g_diis2, g_diis3, g_diis4;

static void diis_free(struct diis_hist *h)
{
	for (int s = 0; s < 4; s++)
		for (int k = 0; k < 2; k++) {
			free(h->P[s][k]);
			free(h->dP[s][k]);
			h->P[s][k] = h->dP[s][k] = NULL;
		}
	h->it = 0;
	h->n = 0;
}

static double conv_density(struct diis_hist *h, int K, const char *name, int nBasis, double alpha, double *PA,
                           double *PB)
{
	if (nBasis == 0 && alpha == 0.0 && PA == NULL && PB == NULL) {
		diis_free(h);
		return 0.0;
	}
	if (alpha < 0.0 || alpha > 1.0) {
		printf("%s : invalid range of alpha\n", name);
		exit(-1);
	}
	const size_t nn = (size_t)nBasis * nBasis;
	double *cur[2] = {PA, PB};
	if (h->it == 0) {
		if (h->P[0][0]) {
			printf("%s : detect prior initialized pointers\n", name);
			exit(-1);
		}
		h->n = nBasis;
		for (int s = 0; s < K; s++)
			for (int k = 0; k < 2; k++) {
				h->P[s][k] = calloc(nn, sizeof(double));
				h->dP[s][k] = calloc(nn, sizeof(double));
				if (!h->P[s][k] || !h->dP[s][k]) {
					printf("%s : cannot allocate memory\n", name);
					exit(-1);
				}
			}
		for (int k = 0; k < 2; k++) memcpy(h->P[K - 1][k], cur[k], nn * sizeof(double));
		h->it = 1;
		return 0.0;
	}
	double rms = 0.0;
	for (int k = 0; k < 2; k++)
		for (size_t e = 0; e < nn; e++) {
			const double t = cur[k][e] - h->P[K - 1][k][e];
			rms += t * t;
			h->dP[K - 1][k][e] = t;
		}
	rms = sqrt(rms / nBasis / nBasis / 2.0);
	if (K == 1 || h->it <= K - 1) {
		for (int k = 0; k < 2; k++)
			for (size_t e = 0; e < nn; e++) cur[k][e] = alpha * cur[k][e] + (1.0 - alpha) * h->P[K - 1][k][e];
	} else {
		double B[4][4];
		for (int a = 0; a < K; a++)
			for (int b = a; b < K; b++) {
				double s = 0.0;
				for (int k = 0; k < 2; k++)
					for (size_t e = 0; e < nn; e++) s += h->dP[a][k][e] * h->dP[b][k][e];
				B[a][b] = B[b][a] = s;
			}
		double c[4] = {0, 0, 0, 0};
		if (K == 2) {  /* src/conv.c:243-249 */
			double c1 = (B[0][0] - B[0][1]) / (B[1][1] - B[0][1]);
			double c0 = 1.0 / (1.0 + c1);
			c1 *= c0;
			if (B[1][1] > B[0][0]) { c1 = alpha; c0 = 1.0 - alpha; }
			c[0] = c0; c[1] = c1;
		} else if (K == 3) {  /* src/conv.c:413-428 */
			const double B00 = B[0][0], B11 = B[1][1], B22 = B[2][2], B01 = B[0][1], B02 = B[0][2], B12 = B[1][2];
			const double den = B12 * B12 - B01 * B12 + B02 * B11 + B01 * B22 - B02 * B12 - B11 * B22;
			double c1 = (B02 * B02 + B00 * B12 - B01 * B02 - B00 * B22 + B01 * B22 - B02 * B12) / den;
			double c2 = (B01 * B01 - B00 * B11 - B01 * B02 + B02 * B11 + B12 * B00 - B12 * B01) / den;
			double c0 = 1.0 / (1.0 + c1 + c2);
			c1 *= c0;
			c2 *= c0;
			if (B22 > B11) { c2 = alpha; c1 = 1.0 - alpha; c0 = 0.0; }
			c[0] = c0; c[1] = c1; c[2] = c2;
		} else {  /* src/conv.c:775-795 */
			double A[25], rhs[5] = {0, 0, 0, 0, -1};
			for (int a = 0; a < 5; a++)
				for (int b = 0; b < 5; b++)
					A[a * 5 + b] = (a < 4 && b < 4) ? B[a][b] : ((a == 4 && b == 4) ? 0.0 : -1.0);
			linear_solver(5, 1, A, rhs);
			c[0] = rhs[0]; c[1] = rhs[1]; c[2] = rhs[2]; c[3] = rhs[3];
			if (B[3][3] > B[2][2]) { c[3] = alpha; c[2] = 1.0 - alpha; c[1] = 0.0; c[0] = 0.0; }
		}
		for (int k = 0; k < 2; k++)
			for (size_t e = 0; e < nn; e++) {
				double v = c[K - 1] * cur[k][e];
				for (int a = 0; a < K - 1; a++) v += c[a] * (h->P[a][k][e] + h->dP[a][k][e]);
				cur[k][e] = v;
			}
	}
	for (int k = 0; k < 2; k++) {
		double *t = h->dP[0][k], *u = h->P[0][k];
		for (int s = 0; s + 1 < K; s++) { h->dP[s][k] = h->dP[s + 1][k]; h->P[s][k] = h->P[s + 1][k]; }
		h->dP[K - 1][k] = t;
		h->P[K - 1][k] = u;
		memcpy(h->P[K - 1][k], cur[k], nn * sizeof(double));
	}
	h->it++;
	return rms;
}

double conv_damping(int nBasis, double alpha, double *PA, double *PB)
{
	return conv_density(&g_damp, 1, "conv_damping", nBasis, alpha, PA, PB);
}

double conv_diis2(int nBasis, double alpha, double *PA, double *PB)
{
	return conv_density(&g_diis2, 2, "conv_diis2", nBasis, alpha, PA, PB);
}

double conv_diis3(int nBasis, double alpha, double *PA, double *PB)
{
	return conv_density(&g_diis3, 3, "conv_diis3", nBasis, alpha, PA, PB);
}

double conv_diis4(int nBasis, double alpha, double *PA, double *PB)
{
	return conv_density(&g_diis4, 4, "conv_diis4", nBasis, alpha, PA, PB);
}

static double now_s(void)
{
	struct timespec t;
	clock_gettime(CLOCK_MONOTONIC, &t);
	return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

/* One-electron matrices H = T + V (+ field) and S (src/uhf.c:1307-1345):
 * T, V and S on the device (sqjk_one_electron), the field term here. */
static void one_electron(int n, const struct GTOBasis_t *gto, const struct Molecule_t *mol,
                         const struct option_t *opt, double *H, double *S)
{
	double *V;
	ALLOC(V, (size_t)n * n);
	const int rc = sqjk_one_electron(n, gto, mol, 0, H, V, S);
	if (rc) {
		printf("uhf: Error - device one-electron matrices failed (%d): %s\n", rc, sqjk_errmsg());
		exit(EXIT_FAILURE);
	}
	for (size_t e = 0; e < (size_t)n * n; e++) H[e] += V[e];
	free(V);
	if (opt->Ex != 0.0 || opt->Ey != 0.0 || opt->Ez != 0.0)
		for (int i = 0; i < n; i++)
			for (int j = 0; j < n; j++)
				H[i * n + j] += GTO_moment(i, j, gto, 1, 0, 0, 0.0, 0.0, 0.0) * opt->Ex +
				                GTO_moment(i, j, gto, 0, 1, 0, 0.0, 0.0, 0.0) * opt->Ey +
				                GTO_moment(i, j, gto, 0, 0, 1, 0.0, 0.0, 0.0) * opt->Ez;
}

static void eigen_or_die(int n, int nOcc, const double *F, const double *S, double *e, double *C)
{
	if (nOcc == 0) return;
	int rc = sqla_gen_eigen(n, F, S, e, C);
	if (rc) {
		printf("gen_sym_SomeEigen: Error - device eigensolver failed (%d)\n", rc);
		exit(EXIT_FAILURE);
	}
}

/* -SAVECHECKALL: the checkpoint of every iteration (src/uhf.c:1739-1741) */
struct save_each {
	int n;
	struct GTOBasis_t *gto;
	struct Molecule_t *mol;
	struct option_t *opt;
	double *CA, *CB, *eA, *eB;
};

static void save_each_iter(void *user, sqscf_ctx *ctx, int iter, double E, double dE, double rmsDP)
{
	struct save_each *s = user;
	(void)iter;
	if (sqscf_orbitals(ctx, s->CA, s->CB, s->eA, s->eB)) {
		printf("uhf: Error - %s\n", sqscf_errmsg());
		exit(EXIT_FAILURE);
	}
	save_checkpoint(s->n, s->gto, s->mol, E, dE, rmsDP, s->CA, s->CB, s->eA, s->eB, s->opt);
}

double uhf(int nBasis, struct GTOBasis_t *gto, struct Molecule_t *mol, int nEA, int nEB, double *CA,
           double *CB, double *eA, double *eB, struct option_t *opt)
{
	const int n = nBasis;
	const size_t nn = (size_t)n * n;
	if (opt->restricted && nEA != nEB) {
		printf("uhf - error number of electron in each spin is not the same\n");
		exit(-1);
	}
	if (opt->convMethod < CONVMETHOD_DAMPING || opt->convMethod > CONVMETHOD_DIIS4) {
		printf("uhf - error unknown convergence method %d\n", opt->convMethod);
		exit(-1);
	}
	double *PA, *PB, *H, *S;
	ALLOC(PA, nn); ALLOC(PB, nn); ALLOC(H, nn); ALLOC(S, nn);
	/* SQ_SCF_TIMING=1 prints the wall time of each phase to stderr */
	const int timing = getenv("SQ_SCF_TIMING") != NULL;
	double tp[4] = {0}, t0 = now_s();
	one_electron(n, gto, mol, opt, H, S);
	tp[0] += now_s() - t0;

	switch (opt->SCFGuess) {
	case SCFGUESS_CACB:
		normalize_orbitals(n, nEA, S, CA);
		normalize_orbitals(n, nEB, S, CB);
		uhf_getDMatrix(n, nEA, CA, PA);
		uhf_getDMatrix(n, nEB, CB, PB);
		break;
	case SCFGUESS_CORE:
		eigen_or_die(n, n, H, S, eA, CA);
		memcpy(CB, CA, nn * sizeof(double));
		memcpy(eB, eA, (size_t)n * sizeof(double));
		normalize_orbitals(n, nEA, S, CA);
		normalize_orbitals(n, nEB, S, CB);
		uhf_getDMatrix(n, nEA, CA, PA);
		uhf_getDMatrix(n, nEB, CB, PB);
		break;
	case SCFGUESS_DIAG: {
		double tr = 0.0;
		for (int i = 0; i < n; i++) tr += S[i * n + i];
		for (int i = 0; i < n; i++) {
			PA[i * n + i] = nEA / tr;
			PB[i * n + i] = nEB / tr;
		}
		break;
	}
	case SCFGUESS_CHECK:
		/* orbitals of a checkpoint projected on this basis (src/uhf.c:1404-1411) */
		guess_checkpoint_orbital(n, gto, mol, opt, S, CA, CB);
		normalize_orbitals(n, nEA, S, CA);
		normalize_orbitals(n, nEB, S, CB);
		uhf_getDMatrix(n, nEA, CA, PA);
		uhf_getDMatrix(n, nEB, CB, PB);
		break;
	default:
		printf("uhf - error this build supports the DIAG, CORE, CHECK and CACB guesses\n");
		exit(-1);
	}

	if (opt->loadDMatrix) sq_load_density(n, opt->restricted, opt->DMatrixFile, PA, PB);  /* src/uhf.c:1414-1462 */

	/* the SCF loop with its state on the device; one J/K context gives the
	 * Schwarz bounds and every build */
	t0 = now_s();
	sqscf_ctx *ctx = NULL;
	if (sqscf_create(0, &ctx) || sqscf_set_system(ctx, n, gto, H, S) || sqscf_set_density(ctx, PA, PB)) {
		printf("uhf: Error - device SCF set-up failed: %s\n", sqscf_errmsg());
		exit(EXIT_FAILURE);
	}
	tp[1] += now_s() - t0;
	double Econst = nuclei_coulomb(mol);
	for (int i = 0; i < mol->nAtom; i++)
		Econst -= mol->Z[i] * (mol->x[i] * opt->Ex + mol->y[i] * opt->Ey + mol->z[i] * opt->Ez);
	struct save_each se = {n, gto, mol, opt, CA, CB, eA, eB};
	sqscf_params prm;
	memset(&prm, 0, sizeof(prm));
	prm.restricted = opt->restricted;
	prm.nEA = nEA;
	prm.nEB = nEB;
	prm.SCFConv = opt->SCFConv;
	prm.SCFMax = opt->SCFMax;
	prm.SCFDrag = opt->SCFDrag;
	prm.threeStep = opt->SCFAccuracy == SCFACCURACY_3STEP;
	prm.convMethod = opt->convMethod;
	prm.Econst = Econst;
	prm.mixedThreshold = sqjk_get_dropin_mixed();
	prm.eigenEveryIter = opt->saveCheckAll;
	prm.onIter = opt->saveCheckAll ? save_each_iter : NULL;
	prm.user = &se;
	sqscf_result res;
	t0 = now_s();
	if (sqscf_run(ctx, &prm, &res)) {
		printf("uhf: Error - device SCF failed: %s\n", sqscf_errmsg());
		exit(EXIT_FAILURE);
	}
	tp[2] += now_s() - t0;
	opt->SCFCutoff = res.realCutoff;
	/* orbitals and orbital energies of the last Fock matrices */
	t0 = now_s();
	if (sqscf_orbitals(ctx, CA, CB, eA, eB) || sqscf_density(ctx, PA, PB)) {
		printf("uhf: Error - %s\n", sqscf_errmsg());
		exit(EXIT_FAILURE);
	}
	tp[3] += now_s() - t0;
	sqscf_destroy(ctx);

	/* density and checkpoint files (src/uhf.c:1843-1883) */
	if (opt->saveDMatrix) sq_save_density(n, opt->DMatrixFile, PA, PB);
	if (opt->saveCheck) save_checkpoint(n, gto, mol, res.energy, res.dE, res.rmsDP, CA, CB, eA, eB, opt);
	if (timing)
		fprintf(stderr, "uhf timing [s]: one-electron %.3f  set-up (Schwarz, metric) %.3f  SCF loop %.3f"
		        "  orbitals %.3f  (%d iterations)\n", tp[0], tp[1], tp[2], tp[3], res.iterations);
	free(PA); free(PB); free(H); free(S);
	return res.converged ? res.energy : 0.0;
}

/* Legacy RHF (src/rhf.c:286-548): core-Hamiltonian guess, G = J[P] - K[P]/2
 * with P = 2 sum C C^T and a fixed 1e-10 Schwarz cut, damping with SCFDrag,
 * stop at |dE| <= 1e-6 and rms dP <= 1e-6. */
double rhf(int nBasis, struct GTOBasis_t *gto, struct Molecule_t *mol, int nE, double *C, double *e,
           struct option_t *opt)
{
	const int n = nBasis, nOcc = nE / 2;
	const size_t nn = (size_t)n * n;
	double *H, *S, *P, *prevP, *F, *Ph, *J, *K, *Kb;
	ALLOC(H, nn); ALLOC(S, nn); ALLOC(P, nn); ALLOC(prevP, nn); ALLOC(F, nn); ALLOC(Ph, nn);
	ALLOC(J, nn); ALLOC(K, nn); ALLOC(Kb, nn);
	one_electron(n, gto, mol, opt, H, S);
	eigen_or_die(n, n, H, S, e, C);
	normalize_orbitals(n, n, S, C);
	rhf_getDMatrix(n, nOcc, C, P);
	double *Schwarz = create_Schwarz(n, gto);
	sqjk_ctx *ctx = NULL;
	int rc = sqjk_create(n, gto, Schwarz, 0, &ctx);
	if (rc) {
		printf("rhf - error %d: %s\n", rc, sqjk_errmsg());
		exit(-1);
	}
	const double Enn = nuclei_coulomb(mol);
	double Etot = Enn, dE = 0.0, avgdP = 0.0;
	for (size_t q = 0; q < nn; q++) Etot += 0.5 * (H[q] + H[q]) * P[q];
	int iter = 0;
	do {
		iter++;
		for (size_t q = 0; q < nn; q++) { Ph[q] = 0.5 * P[q]; J[q] = K[q] = Kb[q] = 0.0; }
		rc = sqjk_build(ctx, Ph, Ph, 0.0, 1.0e-10, 1, J, K, Kb);
		if (rc) {
			printf("rhf - error %d: %s\n", rc, sqjk_errmsg());
			exit(-1);
		}
		for (size_t q = 0; q < nn; q++) F[q] = H[q] + J[q] + K[q];
		eigen_or_die(n, n, F, S, e, C);
		normalize_orbitals(n, n, S, C);
		memcpy(prevP, P, nn * sizeof(double));
		rhf_getDMatrix(n, nOcc, C, P);
		avgdP = 0.0;
		for (size_t q = 0; q < nn; q++) avgdP += (P[q] - prevP[q]) * (P[q] - prevP[q]);
		avgdP = sqrt(avgdP / n / n);
		dE = Etot;
		Etot = Enn;
		for (size_t q = 0; q < nn; q++) Etot += 0.5 * (H[q] + F[q]) * P[q];
		dE = Etot - dE;
		for (size_t q = 0; q < nn; q++) P[q] = prevP[q] + opt->SCFDrag * (P[q] - prevP[q]);
		if (iter >= opt->SCFMax) break;
	} while (fabs(dE) > 1.0e-6 || avgdP > 1.0e-6);
	sqjk_destroy(ctx);
	free(Schwarz);
	free(H); free(S); free(P); free(prevP); free(F); free(Ph); free(J); free(K); free(Kb);
	return Etot;
}
