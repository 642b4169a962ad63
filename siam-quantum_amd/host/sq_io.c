/*
 * sq_io.c - molecule and basis input behind src/mol.h and src/basis.h.
 *
 * readMolecule_XYZ   <- src/mol.c:92-166  (angstrom -> bohr, charge 0)
 * nuclei_coulomb     <- src/mol.c:196-208
 * get_nElectron/nEA/nEB <- src/mol.c:31-67
 * read_GAMESS_BasisSet <- src/basis.c:316-514 (S/L/P/D/F, 'D' exponents)
 * genBasis           <- src/basis.c:529-623 (file order per atom,
 *                       primitive norms 1/sqrt(<g|g>) per Cartesian component)
 * Error behaviour follows the reference: a message and exit().
 */
#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/sq_scf.h"

static const char *SYM[] = {"", "H", "HE", "LI", "BE", "B", "C", "N", "O", "F", "NE", "NA", "MG",
	"AL", "SI", "P", "S", "CL", "AR", "K", "CA", "SC", "TI", "V", "CR", "MN", "FE", "CO", "NI",
	"CU", "ZN", "GA", "GE", "AS", "SE", "BR", "KR"};
static const char *NAME[] = {"", "HYDROGEN", "HELIUM", "LITHIUM", "BERYLLIUM", "BORON", "CARBON",
	"NITROGEN", "OXYGEN", "FLUORINE", "NEON", "SODIUM", "MAGNESIUM", "ALUMINUM", "SILICON",
	"PHOSPHORUS", "SULFUR", "CHLORINE", "ARGON", "POTASSIUM", "CALCIUM", "SCANDIUM", "TITANIUM",
	"VANADIUM", "CHROMIUM", "MANGANESE", "IRON", "COBALT", "NICKEL", "COPPER", "ZINC", "GALLIUM",
	"GERMANIUM", "ARSENIC", "SELENIUM", "BROMINE", "KRYPTON"};
#define NELEM 36

static int lookup(const char *s, const char **tab)
{
	char u[64];
	size_t k;
	for (k = 0; s[k] && k < 63; k++) u[k] = (char)toupper((unsigned char)s[k]);
	u[k] = 0;
	for (int z = 1; z <= NELEM; z++)
		if (strcmp(u, tab[z]) == 0) return z;
	return 0;
}

struct Molecule_t *readMolecule_XYZ(FILE *f)
{
	char line[512];
	struct Molecule_t *m = calloc(1, sizeof(*m));
	if (!m || !fgets(line, sizeof line, f) || sscanf(line, "%d", &m->nAtom) != 1) {
		printf("readMolecule_XYZ: Error - Reading number of atoms\n");
		exit(EXIT_FAILURE);
	}
	if (!fgets(line, sizeof line, f)) {
		printf("readMolecule_XYZ: Error - Reading comment from XYZ file\n");
		exit(EXIT_FAILURE);
	}
	m->Z = calloc(m->nAtom, sizeof(int));
	m->x = calloc(m->nAtom, sizeof(double));
	m->y = calloc(m->nAtom, sizeof(double));
	m->z = calloc(m->nAtom, sizeof(double));
	for (int i = 0; i < m->nAtom; i++) {
		char sym[64];
		if (fscanf(f, "%63s %lf %lf %lf", sym, &m->x[i], &m->y[i], &m->z[i]) != 4) {
			printf("readMolecule_XYZ: Error - Reading coordinate\n");
			exit(EXIT_FAILURE);
		}
		m->Z[i] = lookup(sym, SYM);
		m->x[i] *= SQ_ANGSTROM2BOHR;
		m->y[i] *= SQ_ANGSTROM2BOHR;
		m->z[i] *= SQ_ANGSTROM2BOHR;
	}
	m->Q = 0;
	return m;
}

int get_nElectron(struct Molecule_t *m)
{
	int n = 0;
	for (int i = 0; i < m->nAtom; i++) n += m->Z[i];
	return n - m->Q;
}

double nuclei_coulomb(const struct Molecule_t *m)
{
	double U = 0.0;
	for (int i = 0; i < m->nAtom; i++)
		for (int j = 0; j < i; j++)
			U += m->Z[i] * m->Z[j] / sqrt((m->x[i] - m->x[j]) * (m->x[i] - m->x[j]) +
			                              (m->y[i] - m->y[j]) * (m->y[i] - m->y[j]) +
			                              (m->z[i] - m->z[j]) * (m->z[i] - m->z[j]));
	return U;
}

int get_nEA(struct Molecule_t *m, int M)
{
	int N = get_nElectron(m);
	if ((N + M - 1) % 2 == 1) {
		printf("get_nEA - error number of electron does not match multiplicity\n");
		exit(-1);
	}
	return (N + M - 1) / 2;
}

int get_nEB(struct Molecule_t *m, int M)
{
	int N = get_nElectron(m);
	if ((N - M + 1) % 2 == 1) {
		printf("get_nEB - error number of electron does not match multiplicity\n");
		exit(-1);
	}
	if (N - M + 1 < 0) {
		printf("get_nEB - error multiplicity cannot larger than number of electron\n");
		exit(-1);
	}
	return (N - M + 1) / 2;
}

static double to_num(char *s)
{
	for (char *c = s; *c; c++)
		if (*c == 'D' || *c == 'd') *c = 'E';
	return strtod(s, NULL);
}

struct GTOBasisSet_t *read_GAMESS_BasisSet(FILE *f, char *name, int *nBasisSet)
{
	static const int comp[5][10][3] = {
		{{0, 0, 0}},
		{{0, 0, 0}},
		{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}},
		{{2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}},
		{{3, 0, 0}, {0, 3, 0}, {0, 0, 3}, {2, 1, 0}, {1, 2, 0}, {2, 0, 1}, {1, 0, 2}, {0, 2, 1}, {0, 1, 2}, {1, 1, 1}}};
	char w[1024];
	int found = 0;
	while (fscanf(f, "%1023s", w) == 1)
		if (strcmp(w, "$DATA") == 0) { found = 1; break; }
	if (!found) {
		printf("read_GAMESS_BasisSet - Error: Cannot find $DATA keyword\n");
		exit(EXIT_FAILURE);
	}
	int cap = 1024, ns = 0, Z = 0;
	struct GTOBasisSet_t *db = calloc(cap, sizeof(*db));
	for (;;) {
		if (fscanf(f, "%1023s", w) != 1) {
			printf("read_GAMESS_BasisSet - Error: File corrupted\n");
			exit(EXIT_FAILURE);
		}
		if (strcmp(w, "$END") == 0) break;
		int z = lookup(w, NAME);
		if (z > 0) { Z = z; continue; }
		int kind = -1;
		if (!strcmp(w, "S")) kind = 0; else if (!strcmp(w, "L")) kind = 1;
		else if (!strcmp(w, "P")) kind = 2; else if (!strcmp(w, "D")) kind = 3;
		else if (!strcmp(w, "F")) kind = 4;
		if (kind < 0) continue;
		int nc;
		if (fscanf(f, "%d", &nc) != 1 || nc < 1 || nc > 1024) {
			printf("read_GAMESS_BasisSet - Error: No contracted function number\n");
			exit(EXIT_FAILURE);
		}
		double *ex = malloc(nc * sizeof(double)), *ca = malloc(nc * sizeof(double)), *cb = malloc(nc * sizeof(double));
		for (int i = 0; i < nc; i++) {
			char s1[256], s2[256], s3[256];
			int ok = (kind == 1) ? (fscanf(f, "%*d %255s %255s %255s", s1, s2, s3) == 3)
			                     : (fscanf(f, "%*d %255s %255s", s1, s2) == 2);
			if (!ok) {
				printf("read_GAMESS_BasisSet - Error: Reading coefficient\n");
				exit(EXIT_FAILURE);
			}
			ex[i] = to_num(s1);
			ca[i] = to_num(s2);
			cb[i] = (kind == 1) ? to_num(s3) : 1.0;
			if (ex[i] <= 0.0 || ca[i] == 0.0 || cb[i] == 0.0) {
				printf("read_GAMESS_BasisSet - Error: Reading parsing coefficient\n");
				exit(EXIT_FAILURE);
			}
		}
		const int ncomp = (kind == 0) ? 1 : (kind == 1) ? 4 : (kind == 2) ? 3 : (kind == 3) ? 6 : 10;
		for (int k = 0; k < ncomp; k++) {
			if (ns >= cap) { cap *= 2; db = realloc(db, cap * sizeof(*db)); }
			struct GTOBasisSet_t *b = &db[ns++];
			memset(b, 0, sizeof(*b));
			if (name) strncpy(b->name, name, SQ_BASIS_NAME - 1);
			b->Z = Z;
			b->nContract = nc;
			const int *c3 = (kind == 1) ? ((k == 0) ? comp[0][0] : comp[2][k - 1]) : comp[kind][k];
			const double *cc = (kind == 1 && k > 0) ? cb : ca;
			b->l = c3[0]; b->m = c3[1]; b->n = c3[2];
			b->exp = malloc(nc * sizeof(double));
			b->coef = malloc(nc * sizeof(double));
			memcpy(b->exp, ex, nc * sizeof(double));
			memcpy(b->coef, cc, nc * sizeof(double));
		}
		free(ex); free(ca); free(cb);
	}
	*nBasisSet = ns;
	return db;
}

struct GTOBasis_t *genBasis(struct Molecule_t *mol, int *nBasis, int dbSize, const struct GTOBasisSet_t *db)
{
	int cap = 0;
	for (int i = 0; i < mol->nAtom; i++) {
		int hit = 0;
		for (int b = 0; b < dbSize; b++)
			if (db[b].Z == mol->Z[i]) { cap++; hit = 1; }
		if (!hit) {
			printf("genBasis: Error - Cannot find atom %d in basis file.\n", mol->Z[i]);
			exit(EXIT_FAILURE);
		}
	}
	struct GTOBasis_t *g = calloc(cap, sizeof(*g));
	int n = 0;
	for (int i = 0; i < mol->nAtom; i++)
		for (int b = 0; b < dbSize; b++) {
			if (db[b].Z != mol->Z[i]) continue;
			struct GTOBasis_t *t = &g[n++];
			t->nContract = db[b].nContract;
			t->l = db[b].l; t->m = db[b].m; t->n = db[b].n;
			t->x0 = mol->x[i]; t->y0 = mol->y[i]; t->z0 = mol->z[i];
			t->exp = malloc(t->nContract * sizeof(double));
			t->coef = malloc(t->nContract * sizeof(double));
			t->norm = malloc(t->nContract * sizeof(double));
			for (int k = 0; k < t->nContract; k++) {
				t->exp[k] = db[b].exp[k];
				t->coef[k] = db[b].coef[k];
				t->norm[k] = 1.0 / sqrt(overlap(t->exp[k], t->l, t->m, t->n, t->x0, t->y0, t->z0,
				                                t->exp[k], t->l, t->m, t->n, t->x0, t->y0, t->z0));
			}
		}
	*nBasis = n;
	return g;
}
