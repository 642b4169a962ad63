/*
 * sq_check.c - checkpoint and density-matrix files of the SCF.
 *
 * The text formats are those of the reference so that files move between
 * the two programs in both directions:
 *   checkpoint  src/check.c:432-570 (save_checkpoint), read back by
 *               src/check.c:37-134, 175-200, 210-300 (keyword search, then
 *               whitespace-separated values); basis block as printed by
 *               src/basis.c:37-75 and read by src/basis.c:84-203;
 *   density     src/uhf.c:1414-1462 (-LDMATRIX, n*n values per spin) and
 *               src/uhf.c:1843-1870 (saved alpha then beta).
 * The orbital guess from a checkpoint (src/check.c:306-430) projects the
 * stored orbitals onto the current basis: C = S^-1 Y C_chk with Y the
 * overlap between the current and the stored basis.
 */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../../include/sq_scf.h"

#define CHECK_FORMAT_VERSION "1.2.5"

static void die(const char *fmt, const char *arg)
{
	printf(fmt, arg);
	printf("\n");
	exit(-1);
}

/* Advance fd past the whitespace-separated token `key`; 0 when found. */
static int seek_token(FILE *fd, const char *key)
{
	char buf[1024];
	while (fscanf(fd, "%1023s", buf) == 1)
		if (strcmp(buf, key) == 0) return 0;
	return -1;
}

static FILE *open_or_die(const char *path, const char *mode, const char *who)
{
	FILE *fd = fopen(path, mode);
	if (fd == NULL) {
		printf("%s - error cannot open file %s\n", who, path);
		exit(-1);
	}
	return fd;
}

int printGTOBasis(FILE *out, int nBasis, struct GTOBasis_t *g)
{
	fprintf(out, "nBasis: %6d\n", nBasis);
	for (int i = 0; i < nBasis; i++) {
		fprintf(out, "%2d%2d%2d %15.8E %15.8E %15.8E %6d\n", g[i].l, g[i].m, g[i].n, g[i].x0, g[i].y0, g[i].z0,
		        g[i].nContract);
		for (int j = 0; j < g[i].nContract; j++)
			fprintf(out, "                       %15.8E %15.8E\n", g[i].exp[j], g[i].coef[j]);
	}
	return 1;
}

struct GTOBasis_t *readGTOInput(FILE *in, int *nBasis)
{
	if (in == NULL || fscanf(in, "%*s %d", nBasis) != 1 || *nBasis < 1)
		die("readGTOInput: Error - %s", "reading the number of basis functions");
	struct GTOBasis_t *g = calloc((size_t)*nBasis, sizeof(struct GTOBasis_t));
	if (g == NULL) die("readGTOInput: Error - %s", "cannot allocate memory");
	for (int i = 0; i < *nBasis; i++) {
		if (fscanf(in, "%d %d %d", &g[i].l, &g[i].m, &g[i].n) != 3 ||
		    fscanf(in, "%lf %lf %lf", &g[i].x0, &g[i].y0, &g[i].z0) != 3 ||
		    fscanf(in, "%d", &g[i].nContract) != 1 || g[i].nContract < 1)
			die("readGTOInput: Error - %s", "reading a basis function header");
		g[i].exp = calloc((size_t)g[i].nContract, sizeof(double));
		g[i].coef = calloc((size_t)g[i].nContract, sizeof(double));
		g[i].norm = calloc((size_t)g[i].nContract, sizeof(double));
		if (!g[i].exp || !g[i].coef || !g[i].norm) die("readGTOInput: Error - %s", "cannot allocate memory");
		for (int j = 0; j < g[i].nContract; j++)
			if (fscanf(in, "%lf %lf", &g[i].exp[j], &g[i].coef[j]) != 2)
				die("readGTOInput: Error - %s", "reading contracted functions");
		/* primitive normalisation 1/sqrt(<g|g>) with the component's own l,m,n */
		for (int j = 0; j < g[i].nContract; j++)
			g[i].norm[j] = 1.0 / sqrt(overlap(g[i].exp[j], g[i].l, g[i].m, g[i].n, g[i].x0, g[i].y0, g[i].z0,
			                                  g[i].exp[j], g[i].l, g[i].m, g[i].n, g[i].x0, g[i].y0, g[i].z0));
	}
	return g;
}

static void free_basis(struct GTOBasis_t *g, int n)
{
	for (int i = 0; i < n; i++) {
		free(g[i].exp);
		free(g[i].coef);
		free(g[i].norm);
	}
	free(g);
}

void save_checkpoint(int nBasis, struct GTOBasis_t *gto, struct Molecule_t *mol, double Etot, double dE,
                     double rmsDP, double *CA, double *CB, double *eA, double *eB, struct option_t *opt)
{
	FILE *fd = fopen(opt->CheckFile, "w");
	if (fd == NULL) {
		printf("save_checkpoint - error cannot open file %s for writing\n", opt->CheckFile);
		return;
	}
	char stamp[64];
	const time_t now = time(NULL);
	strftime(stamp, sizeof stamp, "%a %b %d %H:%M:%S %Y", localtime(&now));
	fprintf(fd, "SIAM_QUANTUM_CHECKPOINT\n");
	fprintf(fd, "FORMAT_VERSION    %s\n", CHECK_FORMAT_VERSION);
	fprintf(fd, "TIME %s\n", stamp);
	fprintf(fd, "SECTION_MOLECULAR_SPECIFICATION_BEGIN\n");
	fprintf(fd, "TOTAL_CHARGE   %d\n", mol->Q);
	fprintf(fd, "XYZ_FORMAT\n");
	fprintf(fd, "%d\n\n", mol->nAtom);
	for (int i = 0; i < mol->nAtom; i++)
		fprintf(fd, "%3d %16.8E %16.8E %16.8E\n", mol->Z[i], mol->x[i], mol->y[i], mol->z[i]);
	fprintf(fd, "SECTION_MOLECULAR_SPECIFICATION_END\n");
	fprintf(fd, "SECTION_BASIS_FUNCTION_BEGIN\n");
	printGTOBasis(fd, nBasis, gto);
	fprintf(fd, "SECTION_BASIS_FUNCTION_END\n");
	fprintf(fd, "SECTION_ELECTRONIC_STRUCTURE_BEGIN\n");
	fprintf(fd, "SPIN_MULTIPLICITY   %d\n", opt->multiplicity);
	fprintf(fd, "SCF_METHOD   HF\n");
	fprintf(fd, "TOTAL_ENERGY   %20.8f\n", Etot);
	fprintf(fd, "ENERGY_CONVERGENCE   %15.3E\n", dE);
	fprintf(fd, "DENSITY_CONVERGENCE   %15.3E\n", rmsDP);
	const struct {
		const char *key;
		const double *v;
		int count;
	} blocks[4] = {{"EIGEN_VALUE_ALPHA", eA, nBasis}, {"EIGEN_VALUE_BETA", eB, nBasis},
	               {"ORBITAL_ALPHA", CA, nBasis * nBasis}, {"ORBITAL_BETA", CB, nBasis * nBasis}};
	for (int b = 0; b < 4; b++) {
		fprintf(fd, "%s", blocks[b].key);
		for (int i = 0; i < blocks[b].count; i++) {
			if (i % 5 == 0) fprintf(fd, "\n");
			fprintf(fd, " %16.8E ", blocks[b].v[i]);
		}
		fprintf(fd, "\n");
	}
	fprintf(fd, "SECTION_ELECTRONIC_STRUCTURE_END\n");
	fclose(fd);
}

struct Molecule_t *load_checkpoint_molecule(const struct option_t *opt)
{
	FILE *fd = open_or_die(opt->CheckFile, "r", "load_checkpoint_molecule");
	int Q, nAtom;
	if (seek_token(fd, "TOTAL_CHARGE") || fscanf(fd, "%d", &Q) != 1)
		die("load_checkpoint_molecule - error %s", "reading TOTAL_CHARGE");
	if (seek_token(fd, "XYZ_FORMAT") || fscanf(fd, "%d", &nAtom) != 1 || nAtom < 1)
		die("load_checkpoint_molecule - error %s", "reading XYZ_FORMAT");
	struct Molecule_t *mol = calloc(1, sizeof(struct Molecule_t));
	if (mol == NULL) die("load_checkpoint_molecule - error %s", "cannot allocate memory");
	mol->nAtom = nAtom;
	mol->Q = Q;
	mol->Z = calloc((size_t)nAtom, sizeof(int));
	mol->x = calloc((size_t)nAtom, sizeof(double));
	mol->y = calloc((size_t)nAtom, sizeof(double));
	mol->z = calloc((size_t)nAtom, sizeof(double));
	if (!mol->Z || !mol->x || !mol->y || !mol->z) die("load_checkpoint_molecule - error %s", "cannot allocate memory");
	for (int i = 0; i < nAtom; i++)
		if (fscanf(fd, "%d %lf %lf %lf", mol->Z + i, mol->x + i, mol->y + i, mol->z + i) != 4)
			die("load_checkpoint_molecule - error %s", "reading atom coordinates");
	fclose(fd);
	return mol;
}

int load_checkpoint_multiplicity(const struct option_t *opt)
{
	FILE *fd = open_or_die(opt->CheckFile, "r", "load_checkpoint_multiplicity");
	int M;
	if (seek_token(fd, "SPIN_MULTIPLICITY") || fscanf(fd, "%d", &M) != 1)
		die("load_checkpoint_multiplicity - error %s", "reading SPIN_MULTIPLICITY");
	fclose(fd);
	return M;
}

struct GTOBasis_t *load_checkpoint_basis(const struct option_t *opt, int *nBasis)
{
	FILE *fd = open_or_die(opt->CheckFile, "r", "load_checkpoint_basis");
	if (seek_token(fd, "SECTION_BASIS_FUNCTION_BEGIN"))
		die("load_checkpoint_basis - error %s", "cannot find SECTION_BASIS_FUNCTION_BEGIN");
	struct GTOBasis_t *g = readGTOInput(fd, nBasis);
	fclose(fd);
	return g;
}

static void read_block(FILE *fd, const char *key, double *v, int count)
{
	if (seek_token(fd, key)) die("load_checkpoint_orbital - error cannot find %s keyword", key);
	for (int i = 0; i < count; i++)
		if (fscanf(fd, "%lf", v + i) != 1) die("load_checkpoint_orbital - error reading %s", key);
}

double load_checkpoint_orbital(int nBasis, double *CA, double *CB, double *eA, double *eB,
                               const struct option_t *opt)
{
	FILE *fd = open_or_die(opt->CheckFile, "r", "load_checkpoint_orbital");
	double Etot;
	if (seek_token(fd, "TOTAL_ENERGY") || fscanf(fd, "%lf", &Etot) != 1)
		die("load_checkpoint_orbital - error %s", "reading TOTAL_ENERGY");
	read_block(fd, "EIGEN_VALUE_ALPHA", eA, nBasis);
	read_block(fd, "EIGEN_VALUE_BETA", eB, nBasis);
	read_block(fd, "ORBITAL_ALPHA", CA, nBasis * nBasis);
	read_block(fd, "ORBITAL_BETA", CB, nBasis * nBasis);
	fclose(fd);
	return Etot;
}

void guess_checkpoint_orbital(int nBasis, const struct GTOBasis_t *gto, const struct Molecule_t *mol,
                              const struct option_t *opt, const double *S, double *CA, double *CB)
{
	struct Molecule_t *cm = load_checkpoint_molecule(opt);
	if (cm->nAtom != mol->nAtom) die("guess_checkpoint_orbital - error %s", "checkpoint has a different number of atoms");
	for (int i = 0; i < cm->nAtom; i++)
		if (cm->Z[i] != mol->Z[i]) die("guess_checkpoint_orbital - error %s", "atom types in checkpoint do not match");
	int nc;
	struct GTOBasis_t *cg = load_checkpoint_basis(opt, &nc);
	if (nc > nBasis) die("guess_checkpoint_orbital - error %s", "cannot use a larger basis as guess");
	double *Y = calloc((size_t)nBasis * nc, sizeof(double)), *B = calloc((size_t)nBasis * nc, sizeof(double));
	double *cA = calloc((size_t)nc * nc, sizeof(double)), *cB = calloc((size_t)nc * nc, sizeof(double));
	double *eA = calloc((size_t)nc, sizeof(double)), *eB = calloc((size_t)nc, sizeof(double));
	if (!Y || !B || !cA || !cB || !eA || !eB) die("guess_checkpoint_orbital - error %s", "cannot allocate memory");
	load_checkpoint_orbital(nc, cA, cB, eA, eB, opt);
	/* Y[i][j] = <current i | stored j> */
	for (int i = 0; i < nBasis; i++)
		for (int j = 0; j < nc; j++) {
			double s = 0.0;
			for (int a = 0; a < gto[i].nContract; a++)
				for (int b = 0; b < cg[j].nContract; b++)
					s += gto[i].coef[a] * cg[j].coef[b] * gto[i].norm[a] * cg[j].norm[b] *
					     overlap(gto[i].exp[a], gto[i].l, gto[i].m, gto[i].n, gto[i].x0, gto[i].y0, gto[i].z0,
					             cg[j].exp[b], cg[j].l, cg[j].m, cg[j].n, cg[j].x0, cg[j].y0, cg[j].z0);
			Y[i * nc + j] = s;
		}
	for (int spin = 0; spin < 2; spin++) {
		const double *cs = spin ? cB : cA;
		double *C = spin ? CB : CA;
		/* B = Y C_chk^T (orbital k of the checkpoint in column k), solve S X = B */
		for (int i = 0; i < nBasis; i++)
			for (int k = 0; k < nc; k++) {
				double s = 0.0;
				for (int m = 0; m < nc; m++) s += Y[i * nc + m] * cs[k * nc + m];
				B[i * nc + k] = s;
			}
		linear_solver(nBasis, nc, S, B);
		for (int k = 0; k < nc; k++)
			for (int i = 0; i < nBasis; i++) C[k * nBasis + i] = B[i * nc + k];
	}
	free(Y); free(B); free(cA); free(cB); free(eA); free(eB);
	free_basis(cg, nc);
	free(cm->Z); free(cm->x); free(cm->y); free(cm->z); free(cm);
}

void sq_load_density(int n, int restricted, const char *path, double *PA, double *PB)
{
	FILE *fd = open_or_die(path, "r", "uhf");
	const size_t nn = (size_t)n * n;
	for (size_t e = 0; e < nn; e++)
		if (fscanf(fd, "%lf", PA + e) != 1) die("uhf - error reading density matrix %s", path);
	if (restricted) {
		memcpy(PB, PA, nn * sizeof(double));
	} else {
		for (size_t e = 0; e < nn; e++)
			if (fscanf(fd, "%lf", PB + e) != 1) die("uhf - error reading beta spin density matrix %s", path);
	}
	fclose(fd);
	printf("Density matrix loaded from %s\n", path);
}

void sq_save_density(int n, const char *path, const double *PA, const double *PB)
{
	FILE *fd = fopen(path, "w");
	if (fd == NULL) {
		printf("uhf - error cannot open file %s for writing\n", path);
		return;
	}
	const size_t nn = (size_t)n * n;
	for (int spin = 0; spin < 2; spin++) {
		const double *P = spin ? PB : PA;
		for (size_t e = 0; e < nn; e++) {
			if (e % 5 == 0) fprintf(fd, "\n");
			fprintf(fd, " %15.8E ", P[e]);
		}
	}
	fclose(fd);
	printf("Density matrix saved to %s\n", path);
}
