/*
 * sq_int1e.c - host-side scalar integrals behind the reference's int.h,
 * fgamma.h and the one-electron part of matrix.h.
 *
 * These are per-primitive host functions (the reference's own surface for
 * callers such as gradients and properties); the two-electron hot path does
 * not use them.  Methods (own code):
 *   Boys F_m(x): power series for the top order then downward recursion
 *     for x <= 30; erf for F_0 and upward recursion above.
 *   overlap / moment / kinetic: Obara-Saika 1-D recursions.
 *   nai, eri: McMurchie-Davidson Hermite expansion.
 * Replaces src/fgamma.c:139-206, src/int.c:100-121, 219-243, 262-291,
 * 370-411, 880-935 and src/matrix.c:35-164.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/sq_scf.h"

#define PRIM_CUTOFF 1.0e-15 /* src/int.h:19 */
#define HALF_SQRT_PI 0.88622692545275801365
#define MAXL1 8  /* per-axis angular momentum capacity (l1 + l2 + moment) */

/* ------------------------------------------------------------ Boys */
static double boys_series(int m, double x)
{
	double term = 1.0 / (2 * m + 1), sum = 0.0;
	for (int i = 1; i < 400; i++) {
		sum += term;
		term *= 2.0 * x / (2 * m + 2 * i + 1);
		if (term < 1e-17 * sum) break;
	}
	return exp(-x) * sum;
}

void fgamma_set(int max, double x, double *F)
{
	if (x <= 30.0) {
		F[max] = boys_series(max, x);
		const double e = exp(-x);
		for (int m = max - 1; m >= 0; m--) F[m] = (2.0 * x * F[m + 1] + e) / (2 * m + 1);
	} else {
		const double r = sqrt(x), e = exp(-x);
		F[0] = HALF_SQRT_PI * erf(r) / r;
		for (int m = 0; m < max; m++) F[m + 1] = ((2 * m + 1) * F[m] - e) / (2.0 * x);
	}
}

double fgamma_0(double x)
{
	double F[1];
	fgamma_set(0, x, F);
	return F[0];
}

double fgamma(int m, double x)
{
	double F[64];
	if (m > 63) m = 63;
	fgamma_set(m, x, F);
	return F[m];
}

/* ------------------------------------------------ Obara-Saika 1-D tables */
/* S[i][j] = int x_A^i x_B^j exp(-a x_A^2 - b x_B^2) / (sqrt(pi/p) K), i.e.
 * the polynomial part; the caller multiplies by the prefactor. */
static void os_table(int li, int lj, double pa, double pb, double p, double S[MAXL1 + 3][MAXL1 + 3])
{
	const double h = 0.5 / p;
	S[0][0] = 1.0;
	for (int i = 0; i < li; i++)
		S[i + 1][0] = pa * S[i][0] + (i > 0 ? h * i * S[i - 1][0] : 0.0);
	for (int j = 0; j < lj; j++)
		for (int i = 0; i <= li; i++)
			S[i][j + 1] = pb * S[i][j] + h * ((i > 0 ? i * S[i - 1][j] : 0.0) + (j > 0 ? j * S[i][j - 1] : 0.0));
}

double overlap(double a1, int l1, int m1, int n1, double xa, double ya, double za,
               double a2, int l2, int m2, int n2, double xb, double yb, double zb)
{
	const double p = a1 + a2, mu = a1 * a2 / p;
	const double px = (a1 * xa + a2 * xb) / p, py = (a1 * ya + a2 * yb) / p, pz = (a1 * za + a2 * zb) / p;
	const double ab2 = (xa - xb) * (xa - xb) + (ya - yb) * (ya - yb) + (za - zb) * (za - zb);
	double Sx[MAXL1 + 3][MAXL1 + 3], Sy[MAXL1 + 3][MAXL1 + 3], Sz[MAXL1 + 3][MAXL1 + 3];
	os_table(l1, l2, px - xa, px - xb, p, Sx);
	os_table(m1, m2, py - ya, py - yb, p, Sy);
	os_table(n1, n2, pz - za, pz - zb, p, Sz);
	const double pre = pow(M_PI / p, 1.5) * exp(-mu * ab2);
	return pre * Sx[l1][l2] * Sy[m1][m2] * Sz[n1][n2];
}

/* int (x-A)^i (x-B)^j (x-C)^k via (x-C)^k = sum_r C(k,r) (B-C)^(k-r) (x-B)^r */
static double moment_1d(int i, int j, int k, double pa, double pb, double bc, double p)
{
	double S[MAXL1 + 3][MAXL1 + 3];
	os_table(i, j + k, pa, pb, p, S);
	double s = 0.0, binom = 1.0;
	for (int r = 0; r <= k; r++) {
		s += binom * pow(bc, k - r) * S[i][j + r];
		binom = binom * (k - r) / (r + 1);
	}
	return s;
}

double moment(double a1, int l1, int m1, int n1, double xa, double ya, double za,
              double a2, int l2, int m2, int n2, double xb, double yb, double zb,
              int mx, int my, int mz, double xc, double yc, double zc)
{
	const double p = a1 + a2, mu = a1 * a2 / p;
	const double px = (a1 * xa + a2 * xb) / p, py = (a1 * ya + a2 * yb) / p, pz = (a1 * za + a2 * zb) / p;
	const double ab2 = (xa - xb) * (xa - xb) + (ya - yb) * (ya - yb) + (za - zb) * (za - zb);
	const double pre = pow(M_PI / p, 1.5) * exp(-mu * ab2);
	return pre * moment_1d(l1, l2, mx, px - xa, px - xb, xb - xc, p) *
	       moment_1d(m1, m2, my, py - ya, py - yb, yb - yc, p) *
	       moment_1d(n1, n2, mz, pz - za, pz - zb, zb - zc, p);
}

/* -1/2 d^2/dx^2 acting on x_B^j exp(-b x_B^2):
 *   -1/2 [ j(j-1) x^(j-2) - 2b(2j+1) x^j + 4 b^2 x^(j+2) ] */
static double kin_1d(int i, int j, double b, double S[MAXL1 + 3][MAXL1 + 3])
{
	double t = -2.0 * b * (2 * j + 1) * S[i][j] + 4.0 * b * b * S[i][j + 2];
	if (j >= 2) t += j * (j - 1) * S[i][j - 2];
	return -0.5 * t;
}

double kinetic(double a1, int l1, int m1, int n1, double xa, double ya, double za,
               double a2, int l2, int m2, int n2, double xb, double yb, double zb)
{
	const double p = a1 + a2, mu = a1 * a2 / p;
	const double px = (a1 * xa + a2 * xb) / p, py = (a1 * ya + a2 * yb) / p, pz = (a1 * za + a2 * zb) / p;
	const double ab2 = (xa - xb) * (xa - xb) + (ya - yb) * (ya - yb) + (za - zb) * (za - zb);
	double Sx[MAXL1 + 3][MAXL1 + 3], Sy[MAXL1 + 3][MAXL1 + 3], Sz[MAXL1 + 3][MAXL1 + 3];
	os_table(l1, l2 + 2, px - xa, px - xb, p, Sx);
	os_table(m1, m2 + 2, py - ya, py - yb, p, Sy);
	os_table(n1, n2 + 2, pz - za, pz - zb, p, Sz);
	const double pre = pow(M_PI / p, 1.5) * exp(-mu * ab2);
	return pre * (kin_1d(l1, l2, a2, Sx) * Sy[m1][m2] * Sz[n1][n2] +
	              Sx[l1][l2] * kin_1d(m1, m2, a2, Sy) * Sz[n1][n2] +
	              Sx[l1][l2] * Sy[m1][m2] * kin_1d(n1, n2, a2, Sz));
}

/* ------------------------------------------------ McMurchie-Davidson */
#define MD_L 16  /* max total order of R tables (4 shells x l<=4) */

/* E[i][j][t] of one axis, i <= li, j <= lj */
static void md_E(int li, int lj, double pa, double pb, double p, double E[MAXL1 + 1][MAXL1 + 1][2 * MAXL1 + 2])
{
	const double h = 0.5 / p;
	memset(E, 0, sizeof(double) * (MAXL1 + 1) * (MAXL1 + 1) * (2 * MAXL1 + 2));
	E[0][0][0] = 1.0;
	for (int i = 0; i <= li; i++) {
		if (i > 0)
			for (int t = 0; t <= i; t++)
				E[i][0][t] = (t > 0 ? h * E[i - 1][0][t - 1] : 0.0) + pa * E[i - 1][0][t] +
				             (t + 1 <= i - 1 ? (t + 1) * E[i - 1][0][t + 1] : 0.0);
		for (int j = 1; j <= lj; j++)
			for (int t = 0; t <= i + j; t++)
				E[i][j][t] = (t > 0 ? h * E[i][j - 1][t - 1] : 0.0) + pb * E[i][j - 1][t] +
				             (t + 1 <= i + j - 1 ? (t + 1) * E[i][j - 1][t + 1] : 0.0);
	}
}

/* R[t][u][v] = R^(0)_{tuv}(alpha, X, Y, Z) for t+u+v <= L */
static void md_R(int L, double alpha, double X, double Y, double Z, const double *F,
                 double R[MD_L + 1][MD_L + 1][MD_L + 1])
{
	static _Thread_local double W[MD_L + 1][MD_L + 1][MD_L + 1][MD_L + 1];
	double pw = 1.0;
	for (int n = 0; n <= L; n++) {
		W[n][0][0][0] = pw * F[n];
		pw *= -2.0 * alpha;
	}
	for (int n = L - 1; n >= 0; n--)
		for (int o = 1; o <= L - n; o++)
			for (int t = 0; t <= o; t++)
				for (int u = 0; u <= o - t; u++) {
					const int v = o - t - u;
					double r;
					if (t > 0)
						r = X * W[n + 1][t - 1][u][v] + (t > 1 ? (t - 1) * W[n + 1][t - 2][u][v] : 0.0);
					else if (u > 0)
						r = Y * W[n + 1][t][u - 1][v] + (u > 1 ? (u - 1) * W[n + 1][t][u - 2][v] : 0.0);
					else
						r = Z * W[n + 1][t][u][v - 1] + (v > 1 ? (v - 1) * W[n + 1][t][u][v - 2] : 0.0);
					W[n][t][u][v] = r;
				}
	for (int t = 0; t <= L; t++)
		for (int u = 0; u <= L - t; u++)
			for (int v = 0; v <= L - t - u; v++) R[t][u][v] = W[0][t][u][v];
}

/* -(a|1/r_C|b) with the reference's bare-prefactor cutoff */
double nai(double x1, double y1, double z1, double norm1, int l1, int m1, int n1, double a1,
           double x2, double y2, double z2, double norm2, int l2, int m2, int n2, double a2,
           double x3, double y3, double z3)
{
	const double p = a1 + a2;
	const double px = (a1 * x1 + a2 * x2) / p, py = (a1 * y1 + a2 * y2) / p, pz = (a1 * z1 + a2 * z2) / p;
	const double ab2 = (x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2) + (z1 - z2) * (z1 - z2);
	const double pre = norm1 * norm2 * 2.0 * M_PI / p * exp(-a1 * a2 * ab2 / p);
	if (pre < PRIM_CUTOFF) return 0.0;
	static _Thread_local double Ex[MAXL1 + 1][MAXL1 + 1][2 * MAXL1 + 2], Ey[MAXL1 + 1][MAXL1 + 1][2 * MAXL1 + 2],
	    Ez[MAXL1 + 1][MAXL1 + 1][2 * MAXL1 + 2], R[MD_L + 1][MD_L + 1][MD_L + 1];
	md_E(l1, l2, px - x1, px - x2, p, Ex);
	md_E(m1, m2, py - y1, py - y2, p, Ey);
	md_E(n1, n2, pz - z1, pz - z2, p, Ez);
	const int L = l1 + l2 + m1 + m2 + n1 + n2;
	double F[MD_L + 1];
	const double X = px - x3, Y = py - y3, Z = pz - z3;
	fgamma_set(L, p * (X * X + Y * Y + Z * Z), F);
	md_R(L, p, X, Y, Z, F, R);
	double s = 0.0;
	for (int t = 0; t <= l1 + l2; t++)
		for (int u = 0; u <= m1 + m2; u++)
			for (int v = 0; v <= n1 + n2; v++) s += Ex[l1][l2][t] * Ey[m1][m2][u] * Ez[n1][n2][v] * R[t][u][v];
	return -pre * s;
}

double eri(double xa, double ya, double za, double norma, int la, int ma, int na, double aa,
           double xb, double yb, double zb, double normb, int lb, int mb, int nb, double ab,
           double xc, double yc, double zc, double normc, int lc, int mc, int nc, double ac,
           double xd, double yd, double zd, double normd, int ld, int md, int nd, double ad)
{
	const double p = aa + ab, q = ac + ad, s = p + q, alpha = p * q / s;
	const double Px = (aa * xa + ab * xb) / p, Py = (aa * ya + ab * yb) / p, Pz = (aa * za + ab * zb) / p;
	const double Qx = (ac * xc + ad * xd) / q, Qy = (ac * yc + ad * yd) / q, Qz = (ac * zc + ad * zd) / q;
	const double ab2 = (xa - xb) * (xa - xb) + (ya - yb) * (ya - yb) + (za - zb) * (za - zb);
	const double cd2 = (xc - xd) * (xc - xd) + (yc - yd) * (yc - yd) + (zc - zd) * (zc - zd);
	static _Thread_local double E1[3][MAXL1 + 1][MAXL1 + 1][2 * MAXL1 + 2], E2[3][MAXL1 + 1][MAXL1 + 1][2 * MAXL1 + 2],
	    R[MD_L + 1][MD_L + 1][MD_L + 1];
	md_E(la, lb, Px - xa, Px - xb, p, E1[0]);
	md_E(ma, mb, Py - ya, Py - yb, p, E1[1]);
	md_E(na, nb, Pz - za, Pz - zb, p, E1[2]);
	md_E(lc, ld, Qx - xc, Qx - xd, q, E2[0]);
	md_E(mc, md, Qy - yc, Qy - yd, q, E2[1]);
	md_E(nc, nd, Qz - zc, Qz - zd, q, E2[2]);
	const int L = la + lb + lc + ld + ma + mb + mc + md + na + nb + nc + nd;
	const double X = Px - Qx, Y = Py - Qy, Z = Pz - Qz;
	double F[MD_L + 1];
	fgamma_set(L, alpha * (X * X + Y * Y + Z * Z), F);
	md_R(L, alpha, X, Y, Z, F, R);
	double sum = 0.0;
	for (int t = 0; t <= la + lb; t++)
	for (int u = 0; u <= ma + mb; u++)
	for (int v = 0; v <= na + nb; v++) {
		const double eb = E1[0][la][lb][t] * E1[1][ma][mb][u] * E1[2][na][nb][v];
		if (eb == 0.0) continue;
		for (int T = 0; T <= lc + ld; T++)
		for (int U = 0; U <= mc + md; U++)
		for (int V = 0; V <= nc + nd; V++) {
			const double sg = ((T + U + V) & 1) ? -1.0 : 1.0;
			sum += eb * sg * E2[0][lc][ld][T] * E2[1][mc][md][U] * E2[2][nc][nd][V] * R[t + T][u + U][v + V];
		}
	}
	return 2.0 * pow(M_PI, 2.5) / (p * q * sqrt(s)) * exp(-aa * ab / p * ab2 - ac * ad / q * cd2) * sum *
	       norma * normb * normc * normd;
}

/* ------------------------------------------- contracted 1-e elements */
#define CONTRACT(expr)                                                          \
	double s = 0.0;                                                             \
	for (int a = 0; a < g[i].nContract; a++)                                    \
		for (int b = 0; b < g[j].nContract; b++)                                \
			s += g[i].coef[a] * g[j].coef[b] * g[i].norm[a] * g[j].norm[b] * (expr); \
	return s

double GTO_overlap(int i, int j, const struct GTOBasis_t *g)
{
	CONTRACT(overlap(g[i].exp[a], g[i].l, g[i].m, g[i].n, g[i].x0, g[i].y0, g[i].z0,
	                 g[j].exp[b], g[j].l, g[j].m, g[j].n, g[j].x0, g[j].y0, g[j].z0));
}

double GTO_moment(int i, int j, const struct GTOBasis_t *g, int mx, int my, int mz, double xc, double yc, double zc)
{
	CONTRACT(moment(g[i].exp[a], g[i].l, g[i].m, g[i].n, g[i].x0, g[i].y0, g[i].z0,
	                g[j].exp[b], g[j].l, g[j].m, g[j].n, g[j].x0, g[j].y0, g[j].z0, mx, my, mz, xc, yc, zc));
}

double GTO_kinetic(int i, int j, const struct GTOBasis_t *g)
{
	CONTRACT(kinetic(g[i].exp[a], g[i].l, g[i].m, g[i].n, g[i].x0, g[i].y0, g[i].z0,
	                 g[j].exp[b], g[j].l, g[j].m, g[j].n, g[j].x0, g[j].y0, g[j].z0));
}

double GTO_nuclei(int i, int j, const struct GTOBasis_t *g, const struct Molecule_t *mol)
{
	double s = 0.0;
	for (int a = 0; a < g[i].nContract; a++)
		for (int b = 0; b < g[j].nContract; b++)
			for (int k = 0; k < mol->nAtom; k++)
				s += g[i].coef[a] * g[j].coef[b] * g[i].norm[a] * g[j].norm[b] * (double)mol->Z[k] *
				     nai(g[i].x0, g[i].y0, g[i].z0, 1.0, g[i].l, g[i].m, g[i].n, g[i].exp[a],
				         g[j].x0, g[j].y0, g[j].z0, 1.0, g[j].l, g[j].m, g[j].n, g[j].exp[b],
				         mol->x[k], mol->y[k], mol->z[k]);
	return s;
}
