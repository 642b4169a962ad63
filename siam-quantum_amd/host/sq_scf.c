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
 * The J/K builds (uhf_getGMatrix) and the eigen-solves run on the device.
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

void uhf_getDMatrix(int nBasis, int nOcc, double *C, double *P)
{
	for (int i = 0; i < nBasis; i++)
		for (int j = 0; j < nBasis; j++) {
			double s = 0.0;
			for (int k = 0; k < nOcc; k++) s += C[k * nBasis + i] * C[k * nBasis + j];
			P[i * nBasis + j] = s;
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
static struct {
	int it, n;
	double *P[4][2], *dP[4][2];   /* slot 3 = latest stored input */
} g_diis;

static void diis_free(void)
{
	for (int s = 0; s < 4; s++)
		for (int k = 0; k < 2; k++) {
			free(g_diis.P[s][k]);
			free(g_diis.dP[s][k]);
			g_diis.P[s][k] = g_diis.dP[s][k] = NULL;
		}
	g_diis.it = 0;
	g_diis.n = 0;
}

double conv_diis4(int nBasis, double alpha, double *PA, double *PB)
{
	if (nBasis == 0 && alpha == 0.0 && PA == NULL && PB == NULL) {
		diis_free();
		return 0.0;
	}
	if (alpha < 0.0 || alpha > 1.0) {
		printf("conv_diis4 : invalid range of alpha\n");
		exit(-1);
	}
	const size_t nn = (size_t)nBasis * nBasis;
	double *cur[2] = {PA, PB};
	if (g_diis.it == 0) {
		if (g_diis.P[0][0]) {
			printf("conv_diis4 : detect prior initialized pointers\n");
			exit(-1);
		}
		g_diis.n = nBasis;
		for (int s = 0; s < 4; s++)
			for (int k = 0; k < 2; k++) {
				g_diis.P[s][k] = calloc(nn, sizeof(double));
				g_diis.dP[s][k] = calloc(nn, sizeof(double));
				if (!g_diis.P[s][k] || !g_diis.dP[s][k]) {
					printf("conv_diis4 : cannot allocate memory\n");
					exit(-1);
				}
			}
		for (int k = 0; k < 2; k++) memcpy(g_diis.P[3][k], cur[k], nn * sizeof(double));
		g_diis.it = 1;
		return 0.0;
	}
	double rms = 0.0;
	for (int k = 0; k < 2; k++)
		for (size_t e = 0; e < nn; e++) {
			const double t = cur[k][e] - g_diis.P[3][k][e];
			rms += t * t;
			g_diis.dP[3][k][e] = t;
		}
	rms = sqrt(rms / nBasis / nBasis / 2.0);
	if (g_diis.it <= 3) {
		for (int k = 0; k < 2; k++)
			for (size_t e = 0; e < nn; e++) cur[k][e] = alpha * cur[k][e] + (1.0 - alpha) * g_diis.P[3][k][e];
	} else {
		double B[4][4];
		for (int a = 0; a < 4; a++)
			for (int b = a; b < 4; b++) {
				double s = 0.0;
				for (int k = 0; k < 2; k++)
					for (size_t e = 0; e < nn; e++) s += g_diis.dP[a][k][e] * g_diis.dP[b][k][e];
				B[a][b] = B[b][a] = s;
			}
		double A[25], rhs[5] = {0, 0, 0, 0, -1};
		for (int a = 0; a < 5; a++)
			for (int b = 0; b < 5; b++)
				A[a * 5 + b] = (a < 4 && b < 4) ? B[a][b] : ((a == 4 && b == 4) ? 0.0 : -1.0);
		linear_solver(5, 1, A, rhs);
		double c[4] = {rhs[0], rhs[1], rhs[2], rhs[3]};
		if (B[3][3] > B[2][2]) { c[3] = alpha; c[2] = 1.0 - alpha; c[1] = 0.0; c[0] = 0.0; }
		for (int k = 0; k < 2; k++)
			for (size_t e = 0; e < nn; e++)
				cur[k][e] = c[3] * cur[k][e] + c[2] * (g_diis.P[2][k][e] + g_diis.dP[2][k][e]) +
				            c[1] * (g_diis.P[1][k][e] + g_diis.dP[1][k][e]) +
				            c[0] * (g_diis.P[0][k][e] + g_diis.dP[0][k][e]);
	}
	for (int k = 0; k < 2; k++) {
		double *t = g_diis.dP[0][k];
		g_diis.dP[0][k] = g_diis.dP[1][k]; g_diis.dP[1][k] = g_diis.dP[2][k]; g_diis.dP[2][k] = g_diis.dP[3][k];
		g_diis.dP[3][k] = t;
		double *u = g_diis.P[0][k];
		g_diis.P[0][k] = g_diis.P[1][k]; g_diis.P[1][k] = g_diis.P[2][k]; g_diis.P[2][k] = g_diis.P[3][k];
		g_diis.P[3][k] = u;
		memcpy(g_diis.P[3][k], cur[k], nn * sizeof(double));
	}
	g_diis.it++;
	return rms;
}

/* --------------------------------------------------------------- energy */
static double total_energy(int n, const struct Molecule_t *mol, const struct option_t *opt, const double *PA,
                           const double *PB, const double *H, const double *JT, const double *KA, const double *KB)
{
	double E = 0.0;
	for (int i = 0; i < mol->nAtom; i++)
		E -= mol->Z[i] * (mol->x[i] * opt->Ex + mol->y[i] * opt->Ey + mol->z[i] * opt->Ez);
	E += nuclei_coulomb(mol);
	for (size_t e = 0; e < (size_t)n * n; e++)
		E += (H[e] + 0.5 * JT[e]) * (PA[e] + PB[e]) + 0.5 * KA[e] * PA[e] + 0.5 * KB[e] * PB[e];
	return E;
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

static void density_or_die(int n, int nOcc, const double *C, double *P)
{
	int rc = sqla_density(n, nOcc, C, P);
	if (rc) {
		printf("uhf: Error - device density build failed (%d)\n", rc);
		exit(EXIT_FAILURE);
	}
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

/* The density of the lowest nOcc solutions of F C = S C e.  With use_proj it
 * comes straight from F (spectral projection, no orbitals; e and C are not
 * updated); otherwise, or when the projection does not settle (no gap at
 * nOcc), from the orbitals as in the reference (src/uhf.c:1703-1710). */
static void density_from_fock(int n, int nOcc, const double *F, const double *S, double *e, double *C,
                              double *P, int use_proj)
{
	if (use_proj) {
		int rc = sqla_density_sp2(n, nOcc, F, S, P);
		if (rc == 0) return;
		if (rc < 0) {
			printf("uhf: Error - device density projection failed (%d)\n", rc);
			exit(EXIT_FAILURE);
		}
	}
	eigen_or_die(n, nOcc, F, S, e, C);
	density_or_die(n, nOcc, C, P);
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
	double *JT, *dJT, *KA, *dKA, *KB, *dKB, *PA, *dPA, *PB, *dPB, *FA, *FB, *H, *S;
	ALLOC(JT, nn); ALLOC(dJT, nn); ALLOC(KA, nn); ALLOC(dKA, nn); ALLOC(KB, nn); ALLOC(dKB, nn);
	ALLOC(PA, nn); ALLOC(dPA, nn); ALLOC(PB, nn); ALLOC(dPB, nn); ALLOC(FA, nn); ALLOC(FB, nn);
	ALLOC(H, nn); ALLOC(S, nn);
	/* SQ_SCF_TIMING=1 prints the wall time of each phase to stderr */
	const int timing = getenv("SQ_SCF_TIMING") != NULL;
	double tp[7] = {0}, t0 = now_s();
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

	double sA = 0.0, sB = 0.0;
	for (size_t e = 0; e < nn; e++) { sA += PA[e]; sB += PB[e]; }
#define GETCUTOFF(acc) ((acc) / (sA * sA + sA * sB + sB * sB) / 50.0)
	const double realCutoff = GETCUTOFF(opt->SCFConv);
	opt->SCFCutoff = realCutoff;
	t0 = now_s();
	double *Schwarz = create_Schwarz(n, gto);
	tp[1] += now_s() - t0;

	conv_diis4(0, 0.0, NULL, NULL);
	const double gamma = opt->SCFDrag;
	conv_diis4(n, gamma, PA, PB);
	memcpy(dPA, PA, nn * sizeof(double));
	memcpy(dPB, PB, nn * sizeof(double));

	double cutA = 0.0, cutB = 0.0, cutC = 0.0, thisCutoff;
	if (opt->SCFAccuracy == SCFACCURACY_3STEP) {
		cutA = GETCUTOFF(SCFACCURACY_3STEP_A);
		cutB = GETCUTOFF(SCFACCURACY_3STEP_B);
		cutC = GETCUTOFF(SCFACCURACY_3STEP_C);
		thisCutoff = cutA;
	} else {
		thisCutoff = realCutoff;
	}
#undef GETCUTOFF

	double Etot = 0.0, dE = 0.0, avgdP = 0.0;
	int iter = 0, notConverged = 1;
	do {
		iter++;
		t0 = now_s();
		uhf_getGMatrix(n, gto, Schwarz, thisCutoff, dPA, opt->restricted ? dPA : dPB, dJT, dKA, dKB, opt);
		tp[2] += now_s() - t0;
		t0 = now_s();
		for (size_t e = 0; e < nn; e++) {
			JT[e] += dJT[e];
			KA[e] += dKA[e];
			KB[e] += dKB[e];
			dPA[e] = PA[e];
			dPB[e] = PB[e];
			FA[e] = H[e] + JT[e] + KA[e];
			FB[e] = H[e] + JT[e] + KB[e];
		}
		tp[5] += now_s() - t0;
		/* The new densities come straight from FA, FB by spectral
		 * projection; the orbitals and orbital energies are solved once,
		 * after the loop, from the last FA, FB (or every iteration when
		 * -SAVECHECKALL writes them).  The device eigen-solver returns
		 * S-orthonormal orbitals, so the reference's re-normalisation
		 * (src/uhf.c:1705-1706) is a no-op here.  Restricted: KB = KA, so
		 * FB = FA and the beta solve would repeat the alpha one exactly. */
		t0 = now_s();
		density_from_fock(n, nEA, FA, S, eA, CA, PA, !opt->saveCheckAll);
		if (opt->restricted && nEA == nEB) {
			memcpy(PB, PA, nn * sizeof(double));
			if (opt->saveCheckAll) {
				memcpy(CB, CA, nn * sizeof(double));
				memcpy(eB, eA, (size_t)n * sizeof(double));
			}
		} else {
			density_from_fock(n, nEB, FB, S, eB, CB, PB, !opt->saveCheckAll);
		}
		tp[4] += now_s() - t0;
		t0 = now_s();
		dE = Etot;
		Etot = total_energy(n, mol, opt, PA, PB, H, JT, KA, KB);
		dE = Etot - dE;
		avgdP = conv_diis4(n, gamma, PA, PB);
		notConverged = fabs(dE) > opt->SCFConv || avgdP > opt->SCFConv;
		for (size_t e = 0; e < nn; e++) {
			dPA[e] = PA[e] - dPA[e];
			dPB[e] = PB[e] - dPB[e];
		}
		tp[5] += now_s() - t0;
		if (opt->saveCheckAll) save_checkpoint(n, gto, mol, Etot, dE, avgdP, CA, CB, eA, eB, opt);
		if (iter >= opt->SCFMax) break;
		if (opt->SCFAccuracy == SCFACCURACY_3STEP) {
			if (!notConverged) break;
			if ((thisCutoff == cutA && fabs(dE) <= SCFACCURACY_3STEP_A && avgdP <= SCFACCURACY_3STEP_A) ||
			    (thisCutoff == cutB && fabs(dE) <= SCFACCURACY_3STEP_B && avgdP <= SCFACCURACY_3STEP_B) ||
			    (thisCutoff == cutC && fabs(dE) <= SCFACCURACY_3STEP_C && avgdP <= SCFACCURACY_3STEP_C)) {
				if (thisCutoff == cutA) thisCutoff = cutB;
				else if (thisCutoff == cutB) thisCutoff = cutC;
				else if (thisCutoff == cutC) thisCutoff = realCutoff;
			}
		}
	} while (notConverged);

	/* orbitals and orbital energies of the last Fock matrices */
	t0 = now_s();
	eigen_or_die(n, nEA, FA, S, eA, CA);
	if (opt->restricted && nEA == nEB) {
		memcpy(CB, CA, nn * sizeof(double));
		memcpy(eB, eA, (size_t)n * sizeof(double));
	} else {
		eigen_or_die(n, nEB, FB, S, eB, CB);
	}
	tp[3] += now_s() - t0;

	/* density and checkpoint files (src/uhf.c:1843-1883) */
	if (opt->saveDMatrix) sq_save_density(n, opt->DMatrixFile, PA, PB);
	if (opt->saveCheck) save_checkpoint(n, gto, mol, Etot, dE, avgdP, CA, CB, eA, eB, opt);

	/* release the J/K state (the reference resets its statics here) */
	{
		double z = 0.0;
		t0 = now_s();
		GTO_JK_Matrix_Quartet_Parallel(0, 0, &z, &z, gto, &z, 0.0, &z, &z, &z, opt);
		tp[6] += now_s() - t0;
	}
	if (timing)
		fprintf(stderr, "uhf timing [s]: one-electron %.3f  Schwarz %.3f  J/K %.3f  eigen %.3f  density %.3f"
		        "  host %.3f  release %.3f  (%d iterations)\n", tp[0], tp[1], tp[2], tp[3], tp[4], tp[5], tp[6], iter);
	conv_diis4(0, 0.0, NULL, NULL);
	free(Schwarz);
	free(JT); free(dJT); free(KA); free(dKA); free(KB); free(dKB);
	free(PA); free(dPA); free(PB); free(dPB); free(FA); free(FB); free(H); free(S);
	return notConverged ? 0.0 : Etot;
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
