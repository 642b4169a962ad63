/*
 * sq_grad.c - drop-ins for src/grad.h: uhf_force, GradEri_ShellSet_Parallel,
 * printForce.  The integral work runs on the device (sqjk_grad_eri,
 * sqjk_grad_one_electron); this file forms the densities, the
 * energy-weighted density and the nuclear and field terms, as
 * src/grad.c:2155-2499 does around its integral loops.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sq_grad.h"
#include "sq_scf.h"

#define GRAD_SCHWARZ_CUTOFF 1.0e-10 /* src/grad.h:20 */

void uhf_getDMatrix(int nBasis, int nOcc, double *C, double *P);
double moment(double a1, int l1, int m1, int n1, double xa, double ya, double za, double a2, int l2, int m2,
              int n2, double xb, double yb, double zb, int mx, int my, int mz, double xc, double yc, double zc);

static void *xcalloc(size_t n, size_t sz, const char *who)
{
	void *p = calloc(n ? n : 1, sz);
	if (!p) {
		printf("%s - error cannot allocate memory\n", who);
		exit(-1);
	}
	return p;
}

static void die_jk(const char *who, int rc)
{
	printf("%s - error %d: %s\n", who, rc, sqjk_errmsg());
	exit(-1);
}

/* atom of each basis function: the first atom at the function's centre
 * (src/grad.c:2216-2225) */
static int *basis_atoms(int n, const struct GTOBasis_t *g, const struct Molecule_t *mol, const char *who)
{
	int *b2a = xcalloc(n, sizeof(int), who);
	for (int p = 0; p < n; p++) {
		b2a[p] = -1;
		for (int i = 0; i < mol->nAtom; i++)
			if (g[p].x0 == mol->x[i] && g[p].y0 == mol->y[i] && g[p].z0 == mol->z[i]) {
				b2a[p] = i;
				break;
			}
		if (b2a[p] < 0) {
			printf("%s - error basis function %d is not on an atom\n", who, p + 1);
			exit(-1);
		}
	}
	return b2a;
}

void printForce(const struct Molecule_t *mol, const double *Fx, const double *Fy, const double *Fz)
{
	printf("\n"
	       "-------------------------------------------------------------\n"
	       " Center  Atomic                Forces (Hartrees/Bohr)        \n"
	       " Number  Number          Fx             Fy             Fz    \n"
	       "-------------------------------------------------------------\n");
	for (int a = 0; a < mol->nAtom; a++)
		printf("%5d  %5d    %15.9f%15.9f%15.9f\n", a + 1, mol->Z[a], Fx[a], Fy[a], Fz[a]);
	printf("-------------------------------------------------------------\n");
}

/* d/dX_i of the moment integral (i| (x-cx)^mx (y-cy)^my (z-cz)^mz |j), X_i the
 * centre of function i (src/grad.c:166-247) */
static double grad_moment(int i, int j, const struct GTOBasis_t *g, int mx, int my, int mz, int dir)
{
	const int dl = dir == 0, dm = dir == 1, dn = dir == 2;
	const int lo = dl * g[i].l + dm * g[i].m + dn * g[i].n;
	double down = 0.0, up = 0.0;
	for (int a = 0; a < g[i].nContract; a++)
		for (int b = 0; b < g[j].nContract; b++) {
			const double c = g[i].coef[a] * g[j].coef[b] * g[i].norm[a] * g[j].norm[b];
			if (lo > 0)
				down += c * moment(g[i].exp[a], g[i].l - dl, g[i].m - dm, g[i].n - dn, g[i].x0, g[i].y0, g[i].z0,
				                   g[j].exp[b], g[j].l, g[j].m, g[j].n, g[j].x0, g[j].y0, g[j].z0, mx, my, mz,
				                   0.0, 0.0, 0.0);
			up += g[i].exp[a] * c *
			      moment(g[i].exp[a], g[i].l + dl, g[i].m + dm, g[i].n + dn, g[i].x0, g[i].y0, g[i].z0, g[j].exp[b],
			             g[j].l, g[j].m, g[j].n, g[j].x0, g[j].y0, g[j].z0, mx, my, mz, 0.0, 0.0, 0.0);
		}
	return -(down * lo) + up + up;
}

/* the library-held context of the two-electron gradient drop-in: moved to
 * the new centres when only they changed, rebuilt otherwise */
static sqjk_ctx *g_ctx;
static const struct GTOBasis_t *g_ctx_gto;
static int g_ctx_n;

static sqjk_ctx *grad_context(int nBasis, const struct GTOBasis_t *gto, const char *who)
{
	if (g_ctx && (g_ctx_gto != gto || g_ctx_n != nBasis || sqjk_update_geometry(g_ctx, nBasis, gto) != SQJK_OK)) {
		sqjk_destroy(g_ctx);
		g_ctx = NULL;
	}
	if (!g_ctx) {
		int rc = sqjk_create(nBasis, gto, NULL, 0, &g_ctx);
		if (rc) die_jk(who, rc);
		g_ctx_gto = gto;
		g_ctx_n = nBasis;
	}
	return g_ctx;
}

void GradEri_ShellSet_Parallel(int childID, int nCPU, int nBasis, const struct GTOBasis_t *gto,
                               const int *basis2Atom, const double *PT, const double *PS, const double Kfactor,
                               double *Gx, double *Gy, double *Gz)
{
	(void)nCPU;
	if (childID != 0) return;
	int na = 0;
	for (int p = 0; p < nBasis; p++)
		if (basis2Atom[p] + 1 > na) na = basis2Atom[p] + 1;
	const size_t nn = (size_t)nBasis * nBasis;
	double *PA = xcalloc(nn, sizeof(double), "GradEri_ShellSet_Parallel");
	double *PB = xcalloc(nn, sizeof(double), "GradEri_ShellSet_Parallel");
	double *G = xcalloc((size_t)3 * na, sizeof(double), "GradEri_ShellSet_Parallel");
	for (size_t e = 0; e < nn; e++) {
		PA[e] = 0.5 * (PT[e] + PS[e]);
		PB[e] = 0.5 * (PT[e] - PS[e]);
	}
	sqjk_ctx *c = grad_context(nBasis, gto, "GradEri_ShellSet_Parallel");
	int rc = sqjk_grad_eri(c, PA, PB, Kfactor, GRAD_SCHWARZ_CUTOFF, basis2Atom, na, G);
	if (rc) die_jk("GradEri_ShellSet_Parallel", rc);
	for (int a = 0; a < na; a++) {
		Gx[a] += G[3 * a];
		Gy[a] += G[3 * a + 1];
		Gz[a] += G[3 * a + 2];
	}
	free(PA);
	free(PB);
	free(G);
}

void uhf_force(int nBasis, struct GTOBasis_t *gto, struct Molecule_t *mol, int nEA, int nEB, double *CA,
               double *CB, double *eA, double *eB, struct option_t *opt, double *Fx, double *Fy, double *Fz)
{
	const char *who = "uhf_force";
	if (opt->method != METHOD_HF) {
		printf("uhf_force - error only Hartree-Fock forces are supported by this library\n");
		exit(-1);
	}
	printf("                                                             \n"
	       "                                                             \n"
	       "-------------------------------------------------------------\n"
	       "-----                 FORCE CALCULATIONS                -----\n"
	       "-------------------------------------------------------------\n");
	fflush(stdout);
	const int n = nBasis, na = mol->nAtom;
	const size_t nn = (size_t)n * n;
	int *b2a = basis_atoms(n, gto, mol, who);
	double *PA = xcalloc(nn, sizeof(double), who), *PB = xcalloc(nn, sizeof(double), who);
	double *PT = xcalloc(nn, sizeof(double), who), *QT = xcalloc(nn, sizeof(double), who);
	double *G = xcalloc((size_t)3 * na, sizeof(double), who);
	uhf_getDMatrix(n, nEA, CA, PA);
	uhf_getDMatrix(n, nEB, CB, PB);
	for (size_t e = 0; e < nn; e++) PT[e] = PA[e] + PB[e];
	/* energy-weighted density QT = sum_i e_i C_i C_i^T over both spins, as
	 * rank-one updates over blocks of 16 rows (alpha orbitals, then beta, in
	 * increasing i for every element) */
	for (int p0 = 0; p0 < n; p0 += 16) {
		const int p1 = p0 + 16 < n ? p0 + 16 : n;
		for (int s = 0; s < 2; s++) {
			const int no = s ? nEB : nEA;
			const double *Cs = s ? CB : CA, *es = s ? eB : eA;
			for (int i = 0; i < no; i++) {
				const double *c = Cs + (size_t)i * n;
				for (int p = p0; p < p1; p++) {
					const double cp = es[i] * c[p];
					double *row = QT + (size_t)p * n;
					for (int q = 0; q < n; q++) row[q] += cp * c[q];
				}
			}
		}
	}
	/* nuclear repulsion (src/grad.c:43-66) */
	for (int a = 0; a < na; a++)
		for (int b = 0; b < a; b++) {
			const double dx = mol->x[a] - mol->x[b], dy = mol->y[a] - mol->y[b], dz = mol->z[a] - mol->z[b];
			const double r = sqrt(dx * dx + dy * dy + dz * dz);
			const double f = mol->Z[a] * mol->Z[b] / (r * r * r);
			G[3 * a] -= f * dx; G[3 * b] += f * dx;
			G[3 * a + 1] -= f * dy; G[3 * b + 1] += f * dy;
			G[3 * a + 2] -= f * dz; G[3 * b + 2] += f * dz;
		}
	int rc = sqjk_grad_one_electron(n, gto, mol, PT, QT, b2a, 0, G);
	if (rc) die_jk(who, rc);
	/* uniform field (src/grad.c:2308-2327) */
	if (opt->Ex != 0.0 || opt->Ey != 0.0 || opt->Ez != 0.0) {
		const double E3[3] = {opt->Ex, opt->Ey, opt->Ez};
		for (int p = 0; p < n; p++)
			for (int q = 0; q < n; q++)
				for (int c = 0; c < 3; c++)
					for (int d = 0; d < 3; d++)
						if (E3[c] != 0.0)
							G[3 * b2a[p] + d] += 2.0 * grad_moment(p, q, gto, c == 0, c == 1, c == 2, d) *
							                     PT[(size_t)p * n + q] * E3[c];
		for (int a = 0; a < na; a++)
			for (int d = 0; d < 3; d++) G[3 * a + d] -= E3[d] * mol->Z[a];
	}
	/* two-electron part, Hartree-Fock exchange weight 1 */
	sqjk_ctx *c = grad_context(n, gto, who);
	rc = sqjk_grad_eri(c, PA, PB, 1.0, GRAD_SCHWARZ_CUTOFF, b2a, na, G);
	if (rc) die_jk(who, rc);
	/* force = -gradient */
	double *F = xcalloc((size_t)3 * na, sizeof(double), who);
	for (int a = 0; a < na; a++)
		for (int d = 0; d < 3; d++) F[a * 3 + d] = -G[3 * a + d];
	double *fx = xcalloc(na, sizeof(double), who), *fy = xcalloc(na, sizeof(double), who),
	       *fz = xcalloc(na, sizeof(double), who);
	for (int a = 0; a < na; a++) {
		fx[a] = F[3 * a];
		fy[a] = F[3 * a + 1];
		fz[a] = F[3 * a + 2];
	}
	printForce(mol, fx, fy, fz);
	fflush(stdout);
	for (int a = 0; a < na; a++) {
		if (Fx) Fx[a] = fx[a];
		if (Fy) Fy[a] = fy[a];
		if (Fz) Fz[a] = fz[a];
	}
	free(b2a); free(PA); free(PB); free(PT); free(QT); free(G); free(F); free(fx); free(fy); free(fz);
}
