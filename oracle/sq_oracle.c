/*
 * sq_oracle.c - CPU restatement of Siam Quantum's two-electron path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP path is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library never links it.
 *
 * It follows the reference algorithms step by step (citations per
 * function, paths relative to the reference repository):
 *   Boys F_m(T)         src/fgamma.c:27-206
 *   THO 1-D B terms     src/int.c:40-58 (f_coef), 529-634 (get_H),
 *                       647-873 (eriB_1d, generic branch)
 *   primitive ERI       src/int.c:880-935 (eri)
 *   1-e integrals       src/int.c:100-121, 262-291, 303-411
 *   basis / molecule    src/basis.c:316-623, src/mol.c:31-208
 *   shells              src/matrix.c:796-822, 2415-2541
 *   Schwarz             src/matrix.c:250-292, 389-440
 *   multipole approx.   src/multipole.c:33-86, src/matrix.c:2551-2647
 *   quartet integrals   src/matrix.c:2676-2916 (direct per-tuple B; no HRR)
 *   J/K digestion       src/matrix.c:2934-3396 (six PQIJ cases)
 *   J/K driver          src/matrix.c:3703-3986
 *   SCF                 src/uhf.c:48-315, 1171-1783; src/conv.c:606-844
 *   MP2                 src/mp2.c:29-67, 581-712, 2126-2432;
 *                       src/matrix.c:4797-5073
 * How each part is pinned (tests/test_oracle.py unless named):
 *   Boys                mpmath quadrature; recursion invariant; bit-identical
 *                       to the compiled reference (oracle/_ref)
 *   primitive ERI       closed-form (ss|ss); Szabo-Ostlund H2/STO-3G values;
 *                       permutation symmetry; Schwarz inequality; 200 random
 *                       primitives up to d against the compiled reference
 *   1-e integrals       S, T, V against the compiled reference's GTO_overlap,
 *                       GTO_kinetic, GTO_nuclei
 *   J/K + screening     brute-force dense J/K; the compiled reference's
 *                       GTO_JK_Matrix_Quartet_Parallel (exact, screened, with
 *                       monopole substitution)
 *   readers             tests/test_host_c.py
 *   SCF                 GAMESS and the reference's own log energies
 *   MP2                 the compiled reference's mp2_rhf_aqij_Parallel and
 *                       mp2_uhf_aqij at their own cut-off; the textbook
 *                       full-tensor sums over MO and spin orbitals; the
 *                       published H2O/STO-3G MP2 energy; the closed-shell
 *                       identity UMP2 = RMP2 (tests/test_oracle_mp2.py)
 * Parity unpinned: eigenvector phases and the DIIS path of the SCF (only
 * converged energies fix them).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <ctype.h>
#include "../include/sq_types.h"

#define OR_PRIM_CUTOFF 1.0e-15          /* src/int.h:19 */

/* ------------------------------------------------------------------ */
/* Boys function (src/fgamma.c)                                        */
/* ------------------------------------------------------------------ */

/* The reference uses sqrt(pi)/2 truncated to 13 digits
 * (src/fgamma.c:76,148); the oracle keeps the same constant so that the
 * comparison against the compiled reference is tight. */
#define OR_HALF_SQRTPI_TRUNC 0.8862269254528

/* Power series F_m(x) = e^-x sum_i (2x)^i / ((2m+1)(2m+3)...(2m+2i+1)),
 * stopped when a term drops below 1e-15 (src/fgamma.c:27-41). */
static double or_boys_series(int m, double x)
{
	double den = 2.0 * m + 1.0, term = 1.0 / den, acc = 0.0;
	for (int i = 1; i < 100; i++) {
		acc += term;
		term *= 2.0 * x / (den + 2.0 * i);
		if (term < 1.0e-15) break;
	}
	return exp(-x) * acc;
}

/* Asymptotic value minus the incomplete-gamma tail evaluated by Steed's
 * continued fraction (src/fgamma.c:52-80). */
static double or_boys_cf(int m, double x)
{
	double a = m + 0.5, b = x + 1.0 - a, an = 0.0;
	double d = 1.0 / b, h = d, dh = d;
	for (int i = 1; i <= 100; i++) {
		an += a - 2.0 * i + 1.0;
		b += 2.0;
		d = 1.0 / (b + an * d);
		dh *= b * d - 1.0;
		h += dh;
		if (fabs(dh / h) < 1.0e-15) break;
	}
	double lead = 1.0, inv2x = 1.0 / (2.0 * x);
	for (int k = 1; k <= 2 * m - 1; k += 2) lead *= k * inv2x;
	lead *= OR_HALF_SQRTPI_TRUNC / sqrt(x);
	return lead - 0.5 * exp(-x) * h;
}

double or_fgamma_0(double x)
{
	if (x < 0.0005) return 1.0 - x / 3.0 + x * x / 10.0;
	double t = sqrt(x);
	return OR_HALF_SQRTPI_TRUNC * erf(t) / t;
}

double or_fgamma(int m, double x)
{
	return (x < m + 0.5) ? or_boys_series(m, x) : or_boys_cf(m, x);
}

/* F_0..F_max: short Taylor for tiny x; otherwise F_0 from erf, upward
 * recursion while (2i-1)/(2x) <= 1, and downward recursion from a direct
 * F_max for the rest (src/fgamma.c:172-206). */
void or_fgamma_set(int max, double x, double *F)
{
	if (x < 0.0005) {
		for (int i = 0; i <= max; i++)
			F[i] = 1.0 / (2 * i + 1.0) - x / (2 * i + 3.0) + x * x / (4 * i + 10.0);
		return;
	}
	F[0] = or_fgamma_0(x);
	if (max < 1) return;
	double inv2x = 1.0 / (2.0 * x), ex = inv2x * exp(-x);
	int i;
	for (i = 1; i <= max; i++) {
		double b = (2 * i - 1) * inv2x;
		if (b > 1.0) break;
		F[i] = b * F[i - 1] - ex;
	}
	if (max < i) return;
	F[max] = or_fgamma(max, x);
	double twox = 2.0 * x, ex2 = ex * twox;   /* = exp(-x) */
	for (int j = max - 1; j >= i; j--)
		F[j] = (twox * F[j + 1] + ex2) / (2 * j + 1);
}

/* ------------------------------------------------------------------ */
/* THO one-dimensional pieces (src/int.c)                              */
/* ------------------------------------------------------------------ */

static double or_fact(int n)
{
	double f = 1.0;
	for (int i = 2; i <= n; i++) f *= i;
	return f;
}

static double or_binom(int n, int k)
{
	return or_fact(n) / (or_fact(k) * or_fact(n - k));
}

/* coefficients of x^j in (x+a)^l (x+b)^m  (src/int.c:40-58) */
static void or_poly2(double *f, int l, int m, double a, double b)
{
	for (int j = 0; j <= l + m; j++) f[j] = 0.0;
	for (int i = 0; i <= l; i++)
		for (int k = 0; k <= m; k++)
			f[i + k] += or_binom(l, i) * pow(a, l - i) * or_binom(m, k) * pow(b, m - k);
}

/* coefficients of x^j in (x+a)^l (x+b)^m (x+c)^n  (src/int.c:134-166) */
static void or_poly3(double *f, int l, int m, int n, double a, double b, double c)
{
	for (int j = 0; j <= l + m + n; j++) f[j] = 0.0;
	for (int i = 0; i <= l; i++)
		for (int k = 0; k <= m; k++)
			for (int q = 0; q <= n; q++)
				f[i + k + q] += or_binom(l, i) * pow(a, l - i) * or_binom(m, k) *
				                pow(b, m - k) * or_binom(n, q) * pow(c, n - q);
}

/* 1-D overlap factor sum_i f_2i (2i-1)!! (eta/2)^i  (src/int.c:66-85) */
static double or_overlap_1d(int l1, int l2, double pa, double pb, double eta)
{
	double f[32];
	or_poly2(f, l1, l2, pa, pb);
	double s = 0.0, dfac = 1.0, h = 0.5 * eta, hp = 1.0;
	for (int i = 0; 2 * i <= l1 + l2; i++) {
		if (i > 0) dfac *= (2 * i - 1);
		s += f[2 * i] * dfac * hp;
		hp *= h;
	}
	return s;
}

double or_overlap(double a1, int l1, int m1, int n1, double xa, double ya, double za,
                  double a2, int l2, int m2, int n2, double xb, double yb, double zb)
{
	double eta = 1.0 / (a1 + a2);
	double px = (a1 * xa + a2 * xb) * eta, py = (a1 * ya + a2 * yb) * eta,
	       pz = (a1 * za + a2 * zb) * eta;
	double rab2 = (xa - xb) * (xa - xb) + (ya - yb) * (ya - yb) + (za - zb) * (za - zb);
	double wx = or_overlap_1d(l1, l2, px - xa, px - xb, eta);
	double wy = or_overlap_1d(m1, m2, py - ya, py - yb, eta);
	double wz = or_overlap_1d(n1, n2, pz - za, pz - zb, eta);
	return M_PI * eta * sqrt(M_PI * eta) * exp(-a1 * a2 * rab2 * eta) * wx * wy * wz;
}

/* 1-D moment factor with (x - xc)^mx inserted (src/int.c:172-243) */
static double or_moment_1d(int l1, int l2, int mx, double pa, double pb, double pc, double eta)
{
	double f[40];
	or_poly3(f, l1, l2, mx, pa, pb, pc);
	double s = 0.0, dfac = 1.0, h = 0.5 * eta, hp = 1.0;
	for (int i = 0; 2 * i <= l1 + l2 + mx; i++) {
		if (i > 0) dfac *= (2 * i - 1);
		s += f[2 * i] * dfac * hp;
		hp *= h;
	}
	return s;
}

double or_moment(double a1, int l1, int m1, int n1, double xa, double ya, double za,
                 double a2, int l2, int m2, int n2, double xb, double yb, double zb,
                 int mx, int my, int mz, double xc, double yc, double zc)
{
	double eta = 1.0 / (a1 + a2);
	double px = (a1 * xa + a2 * xb) * eta, py = (a1 * ya + a2 * yb) * eta,
	       pz = (a1 * za + a2 * zb) * eta;
	double rab2 = (xa - xb) * (xa - xb) + (ya - yb) * (ya - yb) + (za - zb) * (za - zb);
	double wx = or_moment_1d(l1, l2, mx, px - xa, px - xb, px - xc, eta);
	double wy = or_moment_1d(m1, m2, my, py - ya, py - yb, py - yc, eta);
	double wz = or_moment_1d(n1, n2, mz, pz - za, pz - zb, pz - zc, eta);
	return M_PI * eta * sqrt(M_PI * eta) * exp(-a1 * a2 * rab2 * eta) * wx * wy * wz;
}

/* kinetic energy from shifted overlaps (src/int.c:262-291) */
double or_kinetic(double a1, int l1, int m1, int n1, double xa, double ya, double za,
                  double a2, int l2, int m2, int n2, double xb, double yb, double zb)
{
#define OV(dl, dm, dn) or_overlap(a1, l1, m1, n1, xa, ya, za, a2, l2 + (dl), m2 + (dm), n2 + (dn), xb, yb, zb)
	double t0 = a2 * (2 * (l2 + m2 + n2) + 3) * OV(0, 0, 0);
	double t1 = -2.0 * a2 * a2 * (OV(2, 0, 0) + OV(0, 2, 0) + OV(0, 0, 2));
	double t2 = -0.5 * ((l2 < 2 ? 0.0 : l2 * (l2 - 1) * OV(-2, 0, 0)) +
	                    (m2 < 2 ? 0.0 : m2 * (m2 - 1) * OV(0, -2, 0)) +
	                    (n2 < 2 ? 0.0 : n2 * (n2 - 1) * OV(0, 0, -2)));
#undef OV
	return t0 + t1 + t2;
}

/* A_I of THO eq. 2.18 (src/int.c:320-357) */
static void or_nai_A(double *A, int l1, int l2, double pa, double pb, double cp, double eta)
{
	double f[32];
	or_poly2(f, l1, l2, pa, pb);
	double e4 = eta / 4.0;
	for (int I = 0; I <= l1 + l2; I++) A[I] = 0.0;
	for (int i = 0; i <= l1 + l2; i++)
		for (int r = 0; 2 * r <= i; r++)
			for (int u = 0; 2 * u <= i - 2 * r; u++) {
				int I = i - 2 * r - u;
				double sgn = ((i + u) & 1) ? -1.0 : 1.0;
				A[I] += f[i] * sgn * or_fact(i) / (or_fact(r) * or_fact(u) * or_fact(I - u)) *
				        pow(cp, I - u) * pow(e4, r + u);
			}
}

/* -(a|1/r_C|b), with the bare-prefactor cutoff (src/int.c:370-411) */
double or_nai(double x1, double y1, double z1, double norm1, int l1, int m1, int n1, double a1,
              double x2, double y2, double z2, double norm2, int l2, int m2, int n2, double a2,
              double x3, double y3, double z3)
{
	double eta = 1.0 / (a1 + a2);
	double xp = (a1 * x1 + a2 * x2) * eta, yp = (a1 * y1 + a2 * y2) * eta,
	       zp = (a1 * z1 + a2 * z2) * eta;
	double rab2 = (x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2) + (z1 - z2) * (z1 - z2);
	double rcp2 = (x3 - xp) * (x3 - xp) + (y3 - yp) * (y3 - yp) + (z3 - zp) * (z3 - zp);
	double pre = norm1 * norm2 * 2.0 * M_PI * eta * exp(-a1 * a2 * rab2 * eta);
	if (pre < OR_PRIM_CUTOFF) return 0.0;
	double Ax[32], Ay[32], Az[32], F[64];
	or_nai_A(Ax, l1, l2, xp - x1, xp - x2, xp - x3, eta);
	or_nai_A(Ay, m1, m2, yp - y1, yp - y2, yp - y3, eta);
	or_nai_A(Az, n1, n2, zp - z1, zp - z2, zp - z3, eta);
	or_fgamma_set(l1 + l2 + m1 + m2 + n1 + n2, rcp2 / eta, F);
	double s = 0.0;
	for (int i = 0; i <= l1 + l2; i++)
		for (int j = 0; j <= m1 + m2; j++)
			for (int k = 0; k <= n1 + n2; k++)
				s += Ax[i] * Ay[j] * Az[k] * F[i + j + k];
	return -pre * s;
}

/* H_L = sum_{i-2r=L} f_i i!/(r! L!) (eta/4)^(i-r)  (src/int.c:529-634) */
static void or_tho_H(double *H, const double *f, int l12, double eta)
{
	double e4 = eta / 4.0;
	for (int L = 0; L <= l12; L++) H[L] = 0.0;
	for (int i = 0; i <= l12; i++)
		for (int r = 0; 2 * r <= i; r++) {
			int L = i - 2 * r;
			H[L] += f[i] * or_fact(i) / (or_fact(r) * or_fact(L)) * pow(e4, i - r);
		}
}

/* THO B_I for one axis (src/int.c:647-873, generic branch):
 * B_I = sum_{L,M,u; I=L+M-u} HL_L HM_M (-1)^(u+M) (L+M)!/u!
 *       * QP^(L+M-2u)/(L+M-2u)! * zeta^I                                  */
void or_eriB_1d(double *B, int l1, int l2, int l3, int l4,
                double Px, double Ax, double Bx, double Qx, double Cx, double Dx,
                double eta1, double eta2, double zeta)
{
	double f1[32], f2[32], HL[32], HM[32];
	double QP = Qx - Px;
	or_poly2(f1, l1, l2, Px - Ax, Px - Bx);
	or_poly2(f2, l3, l4, Qx - Cx, Qx - Dx);
	or_tho_H(HL, f1, l1 + l2, eta1);
	or_tho_H(HM, f2, l3 + l4, eta2);
	for (int I = 0; I <= l1 + l2 + l3 + l4; I++) B[I] = 0.0;
	for (int L = 0; L <= l1 + l2; L++)
		for (int M = 0; M <= l3 + l4; M++)
			for (int u = 0; 2 * u <= L + M; u++) {
				int I = L + M - u, k = L + M - 2 * u;
				double sgn = ((u + M) & 1) ? -1.0 : 1.0;
				B[I] += HL[L] * HM[M] * sgn * or_fact(L + M) / or_fact(u) *
				        pow(QP, k) / or_fact(k) * pow(zeta, I);
			}
}

#define OR_TWO_PI_POW2_5 34.986836655249725

/* Primitive ERI (ab|cd) with primitive norms (src/int.c:880-935).
 * If cut != 0 the bare prefactor test of src/int.c:1276-1282 is applied. */
static double or_eri_prim(double xa, double ya, double za, int la, int ma, int na, double aa,
                          double xb, double yb, double zb, int lb, int mb, int nb, double ab,
                          double xc, double yc, double zc, int lc, int mc, int nc, double ac,
                          double xd, double yd, double zd, int ld, int md, int nd, double ad,
                          int cut)
{
	double eta1 = 1.0 / (aa + ab), eta2 = 1.0 / (ac + ad);
	double zeta = 4.0 * (aa + ab) * (ac + ad) / (aa + ab + ac + ad);
	double xp = (aa * xa + ab * xb) * eta1, yp = (aa * ya + ab * yb) * eta1, zp = (aa * za + ab * zb) * eta1;
	double xq = (ac * xc + ad * xd) * eta2, yq = (ac * yc + ad * yd) * eta2, zq = (ac * zc + ad * zd) * eta2;
	double rab2 = (xa - xb) * (xa - xb) + (ya - yb) * (ya - yb) + (za - zb) * (za - zb);
	double rcd2 = (xc - xd) * (xc - xd) + (yc - yd) * (yc - yd) + (zc - zd) * (zc - zd);
	double rpq2 = (xp - xq) * (xp - xq) + (yp - yq) * (yp - yq) + (zp - zq) * (zp - zq);
	double pre = OR_TWO_PI_POW2_5 * eta1 * eta2 / sqrt(aa + ab + ac + ad) *
	             exp(-aa * ab * rab2 * eta1 - ac * ad * rcd2 * eta2);
	if (cut && pre < OR_PRIM_CUTOFF) return 0.0;
	double F[64], Bx[64], By[64], Bz[64];
	int L = la + lb + lc + ld + ma + mb + mc + md + na + nb + nc + nd;
	or_fgamma_set(L, 0.25 * rpq2 * zeta, F);
	or_eriB_1d(Bx, la, lb, lc, ld, xp, xa, xb, xq, xc, xd, eta1, eta2, zeta);
	or_eriB_1d(By, ma, mb, mc, md, yp, ya, yb, yq, yc, yd, eta1, eta2, zeta);
	or_eriB_1d(Bz, na, nb, nc, nd, zp, za, zb, zq, zc, zd, eta1, eta2, zeta);
	double s = 0.0;
	for (int i = 0; i <= la + lb + lc + ld; i++)
		for (int j = 0; j <= ma + mb + mc + md; j++)
			for (int k = 0; k <= na + nb + nc + nd; k++)
				s += Bx[i] * By[j] * Bz[k] * F[i + j + k];
	return pre * s;
}

double or_eri(double xa, double ya, double za, double norma, int la, int ma, int na, double aa,
              double xb, double yb, double zb, double normb, int lb, int mb, int nb, double ab,
              double xc, double yc, double zc, double normc, int lc, int mc, int nc, double ac,
              double xd, double yd, double zd, double normd, int ld, int md, int nd, double ad)
{
	return norma * normb * normc * normd *
	       or_eri_prim(xa, ya, za, la, ma, na, aa, xb, yb, zb, lb, mb, nb, ab,
	                   xc, yc, zc, lc, mc, nc, ac, xd, yd, zd, ld, md, nd, ad, 0);
}

/* ------------------------------------------------------------------ */
/* Contracted one- and two-electron integrals over basis functions     */
/* ------------------------------------------------------------------ */

double or_gto_overlap(int i, int j, const struct GTOBasis_t *g)
{
	double s = 0.0;
	for (int a = 0; a < g[i].nContract; a++)
		for (int b = 0; b < g[j].nContract; b++)
			s += g[i].coef[a] * g[j].coef[b] * g[i].norm[a] * g[j].norm[b] *
			     or_overlap(g[i].exp[a], g[i].l, g[i].m, g[i].n, g[i].x0, g[i].y0, g[i].z0,
			                g[j].exp[b], g[j].l, g[j].m, g[j].n, g[j].x0, g[j].y0, g[j].z0);
	return s;
}

double or_gto_moment(int i, int j, const struct GTOBasis_t *g, int mx, int my, int mz,
                     double xc, double yc, double zc)
{
	double s = 0.0;
	for (int a = 0; a < g[i].nContract; a++)
		for (int b = 0; b < g[j].nContract; b++)
			s += g[i].coef[a] * g[j].coef[b] * g[i].norm[a] * g[j].norm[b] *
			     or_moment(g[i].exp[a], g[i].l, g[i].m, g[i].n, g[i].x0, g[i].y0, g[i].z0,
			               g[j].exp[b], g[j].l, g[j].m, g[j].n, g[j].x0, g[j].y0, g[j].z0,
			               mx, my, mz, xc, yc, zc);
	return s;
}

double or_gto_kinetic(int i, int j, const struct GTOBasis_t *g)
{
	double s = 0.0;
	for (int a = 0; a < g[i].nContract; a++)
		for (int b = 0; b < g[j].nContract; b++)
			s += g[i].coef[a] * g[j].coef[b] * g[i].norm[a] * g[j].norm[b] *
			     or_kinetic(g[i].exp[a], g[i].l, g[i].m, g[i].n, g[i].x0, g[i].y0, g[i].z0,
			                g[j].exp[b], g[j].l, g[j].m, g[j].n, g[j].x0, g[j].y0, g[j].z0);
	return s;
}

double or_gto_nuclei(int i, int j, const struct GTOBasis_t *g, const struct Molecule_t *mol)
{
	double s = 0.0;
	for (int a = 0; a < g[i].nContract; a++)
		for (int b = 0; b < g[j].nContract; b++)
			for (int k = 0; k < mol->nAtom; k++)
				s += g[i].coef[a] * g[j].coef[b] * g[i].norm[a] * g[j].norm[b] * (double)mol->Z[k] *
				     or_nai(g[i].x0, g[i].y0, g[i].z0, 1.0, g[i].l, g[i].m, g[i].n, g[i].exp[a],
				            g[j].x0, g[j].y0, g[j].z0, 1.0, g[j].l, g[j].m, g[j].n, g[j].exp[b],
				            mol->x[k], mol->y[k], mol->z[k]);
	return s;
}

/* contracted (pq|ij) over basis functions, no primitive cutoff
 * (src/matrix.c:250-292) */
double or_contr_eri(const struct GTOBasis_t *g, int p, int q, int i, int j)
{
	double s = 0.0;
	for (int a = 0; a < g[p].nContract; a++)
	for (int b = 0; b < g[q].nContract; b++)
	for (int c = 0; c < g[i].nContract; c++)
	for (int d = 0; d < g[j].nContract; d++)
		s += g[p].coef[a] * g[q].coef[b] * g[i].coef[c] * g[j].coef[d] *
		     or_eri(g[p].x0, g[p].y0, g[p].z0, g[p].norm[a], g[p].l, g[p].m, g[p].n, g[p].exp[a],
		            g[q].x0, g[q].y0, g[q].z0, g[q].norm[b], g[q].l, g[q].m, g[q].n, g[q].exp[b],
		            g[i].x0, g[i].y0, g[i].z0, g[i].norm[c], g[i].l, g[i].m, g[i].n, g[i].exp[c],
		            g[j].x0, g[j].y0, g[j].z0, g[j].norm[d], g[j].l, g[j].m, g[j].n, g[j].exp[d]);
	return s;
}

/* Sch[p,q] = sqrt|(pq|pq)| over all n^2 pairs (src/matrix.c:389-440) */
void or_schwarz(int n, const struct GTOBasis_t *g, double *sch)
{
	for (int p = 0; p < n; p++)
		for (int q = 0; q < n; q++)
			sch[p * n + q] = sqrt(fabs(or_contr_eri(g, p, q, p, q)));
}

/* ------------------------------------------------------------------ */
/* Shells (src/matrix.c:796-822, 2415-2541)                            */
/* ------------------------------------------------------------------ */

#define OR_MAXB 16
#define OR_MAXK 32
typedef struct {
	double x, y, z;
	int nBasis, nContract, maxL, min, max, type;
	int l[OR_MAXB], m[OR_MAXB], n[OR_MAXB];
	double exps[OR_MAXK];
	double coef[OR_MAXK][OR_MAXB];   /* norm * coef per component */
} or_shell;

static int or_same_shell(const struct GTOBasis_t *g, int i, int j)
{
	if (g[i].x0 != g[j].x0 || g[i].y0 != g[j].y0 || g[i].z0 != g[j].z0) return 0;
	if (g[i].nContract != g[j].nContract) return 0;
	for (int k = 0; k < g[i].nContract; k++)
		if (g[i].exp[k] != g[j].exp[k]) return 0;
	return 1;
}

/* group consecutive same-centre same-exponent functions; type by size;
 * stable sort by type (S < L < P < D < F < other) */
static or_shell *or_make_shells(int n, const struct GTOBasis_t *g, int *nsh)
{
	or_shell *sh = calloc(n, sizeof(or_shell));
	int ns = 0;
	for (int p = 0; p < n;) {
		or_shell *s = &sh[ns];
		s->min = p; s->x = g[p].x0; s->y = g[p].y0; s->z = g[p].z0;
		s->nContract = g[p].nContract;
		for (int k = 0; k < g[p].nContract; k++) s->exps[k] = g[p].exp[k];
		int q = p;
		while (q < n && or_same_shell(g, p, q)) {
			int L = g[q].l + g[q].m + g[q].n;
			if (L > s->maxL) s->maxL = L;
			s->max = q;
			s->l[s->nBasis] = g[q].l; s->m[s->nBasis] = g[q].m; s->n[s->nBasis] = g[q].n;
			for (int k = 0; k < g[q].nContract; k++) s->coef[k][s->nBasis] = g[q].norm[k] * g[q].coef[k];
			s->nBasis++;
			q++;
		}
		switch (s->nBasis) {
		case 1: s->type = 0; break;
		case 4: s->type = 1; break;
		case 3: s->type = 2; break;
		case 6: s->type = 3; break;
		case 10: s->type = 4; break;
		default: s->type = 6; break;
		}
		p = q;
		ns++;
	}
	/* stable bubble sort on type */
	for (int a = 0; a < ns; a++)
		for (int b = 0; b + 1 < ns; b++)
			if (sh[b].type > sh[b + 1].type) {
				or_shell t = sh[b]; sh[b] = sh[b + 1]; sh[b + 1] = t;
			}
	*nsh = ns;
	return sh;
}

/* contracted shell-quartet block EE[((p*nQ+q)*nI+i)*nJ+j], with the
 * per-primitive-quartet bare prefactor cutoff (src/matrix.c:2676-2916) */
static void or_quartet(const or_shell *P, const or_shell *Q, const or_shell *I, const or_shell *J, double *EE)
{
	int nEE = P->nBasis * Q->nBasis * I->nBasis * J->nBasis;
	for (int k = 0; k < nEE; k++) EE[k] = 0.0;
	for (int a = 0; a < P->nContract; a++)
	for (int b = 0; b < Q->nContract; b++)
	for (int c = 0; c < I->nContract; c++)
	for (int d = 0; d < J->nContract; d++) {
		int e = 0;
		for (int p = 0; p < P->nBasis; p++)
		for (int q = 0; q < Q->nBasis; q++)
		for (int i = 0; i < I->nBasis; i++)
		for (int j = 0; j < J->nBasis; j++, e++) {
			double v = or_eri_prim(P->x, P->y, P->z, P->l[p], P->m[p], P->n[p], P->exps[a],
			                       Q->x, Q->y, Q->z, Q->l[q], Q->m[q], Q->n[q], Q->exps[b],
			                       I->x, I->y, I->z, I->l[i], I->m[i], I->n[i], I->exps[c],
			                       J->x, J->y, J->z, J->l[j], J->m[j], J->n[j], J->exps[d], 1);
			EE[e] += v * P->coef[a][p] * Q->coef[b][q] * I->coef[c][i] * J->coef[d][j];
		}
	}
}

/* ------------------------------------------------------------------ */
/* Multipole approximation (src/multipole.c:33-86, matrix.c:2551-2647) */
/* ------------------------------------------------------------------ */

typedef struct { double q, x, y, z; } or_mpole;

static or_mpole *or_make_mpole(int n, const struct GTOBasis_t *g)
{
	or_mpole *mp = calloc((size_t)n * n, sizeof(or_mpole));
	for (int i = 0; i < n; i++)
		for (int j = 0; j < n; j++) {
			double q = or_gto_overlap(i, j, g);
			if (fabs(q) < 1.0e-18) q = 0.0;
			double x = 0.5 * (g[i].x0 + g[j].x0), y = 0.5 * (g[i].y0 + g[j].y0),
			       z = 0.5 * (g[i].z0 + g[j].z0);
			double px = or_gto_moment(i, j, g, 1, 0, 0, x, y, z);
			double py = or_gto_moment(i, j, g, 0, 1, 0, x, y, z);
			double pz = or_gto_moment(i, j, g, 0, 0, 1, x, y, z);
			if (q != 0.0) { x += px / q; y += py / q; z += pz / q; }
			mp[i * n + j] = (or_mpole){q, x, y, z};
		}
	return mp;
}

#define OR_ERR_MIN -120
#define OR_ERR_MAX 120
#define OR_ERR_UNKNOWN 121
#define OR_LNBASEI 2.4663034624   /* 1/ln(1.5), src/matrix.c:1743 */

static double or_mpole_value(const or_mpole *m1, const or_mpole *m2)
{
	double r2 = (m1->x - m2->x) * (m1->x - m2->x) + (m1->y - m2->y) * (m1->y - m2->y) +
	            (m1->z - m2->z) * (m1->z - m2->z);
	return (r2 < 1.0) ? 0.0 : m1->q * m2->q * sqrt(1.0 / r2);
}

static signed char or_mpole_err(const or_shell *P, const or_shell *Q, const or_shell *I, const or_shell *J,
                                const or_mpole *mp, int n, const double *EE)
{
	double mx = 0.0;
	int e = 0;
	for (int p = 0; p < P->nBasis; p++)
	for (int q = 0; q < Q->nBasis; q++)
	for (int i = 0; i < I->nBasis; i++)
	for (int j = 0; j < J->nBasis; j++, e++) {
		double d = EE[e] - or_mpole_value(&mp[(P->min + p) * n + Q->min + q], &mp[(I->min + i) * n + J->min + j]);
		if (fabs(d) > mx) mx = fabs(d);
	}
	mx = (mx == 0.0) ? OR_ERR_MIN : log(mx) * OR_LNBASEI;
	if (mx < OR_ERR_MIN) mx = OR_ERR_MIN;
	if (mx > OR_ERR_MAX) mx = OR_ERR_MAX;
	return (signed char)mx;
}

static void or_mpole_quartet(const or_shell *P, const or_shell *Q, const or_shell *I, const or_shell *J,
                             const or_mpole *mp, int n, double *EE)
{
	int e = 0;
	for (int p = 0; p < P->nBasis; p++)
	for (int q = 0; q < Q->nBasis; q++)
	for (int i = 0; i < I->nBasis; i++)
	for (int j = 0; j < J->nBasis; j++, e++)
		EE[e] = or_mpole_value(&mp[(P->min + p) * n + Q->min + q], &mp[(I->min + i) * n + J->min + j]);
}

/* ------------------------------------------------------------------ */
/* J/K digestion (src/matrix.c:2934-3396)                              */
/* ------------------------------------------------------------------ */

enum { OR_ALLDISTINCT = 0, OR_PIQJ = 1, OR_IJPAIR = 2, OR_PQPAIR = 3, OR_2PAIRS = 4, OR_ALLSAME = 5 };

/* Strides of EE for the four shell slots (p,q,i,j). */
typedef struct { const or_shell *s[4]; int st[4]; } or_view;

/* Coulomb block: for (a,b) in slots (A,B), G = f * sum_{c,d} PT[c,d] EE;
 * add to JT[a,b] and, if both, to JT[b,a]. */
static void or_jblock(const or_view *v, int A, int B, int C, int D, double f, int both,
                      int n, double *JT, const double *PT, const double *EE)
{
	const or_shell *SA = v->s[A], *SB = v->s[B], *SC = v->s[C], *SD = v->s[D];
	for (int a = 0; a < SA->nBasis; a++)
	for (int b = 0; b < SB->nBasis; b++) {
		double g = 0.0;
		for (int c = 0; c < SC->nBasis; c++)
		for (int d = 0; d < SD->nBasis; d++) {
			double e = EE[a * v->st[A] + b * v->st[B] + c * v->st[C] + d * v->st[D]];
			g += PT[(SC->min + c) * n + SD->min + d] * e * f;
		}
		JT[(SA->min + a) * n + SB->min + b] += g;
		if (both) JT[(SB->min + b) * n + SA->min + a] += g;
	}
}

/* Exchange block: for (a,c) in slots (A,C), G = sum_{b,d} P[b,d] EE;
 * subtract from K[a,c] and, if both, from K[c,a]. */
static void or_kblock(const or_view *v, int A, int B, int C, int D, int both,
                      int n, double *K, const double *P, const double *EE)
{
	const or_shell *SA = v->s[A], *SB = v->s[B], *SC = v->s[C], *SD = v->s[D];
	for (int a = 0; a < SA->nBasis; a++)
	for (int c = 0; c < SC->nBasis; c++) {
		double g = 0.0;
		for (int b = 0; b < SB->nBasis; b++)
		for (int d = 0; d < SD->nBasis; d++) {
			double e = EE[a * v->st[A] + b * v->st[B] + c * v->st[C] + d * v->st[D]];
			g += P[(SB->min + b) * n + SD->min + d] * e;
		}
		K[(SA->min + a) * n + SC->min + c] -= g;
		if (both) K[(SC->min + c) * n + SA->min + a] -= g;
	}
}

/* slot indices: 0=P 1=Q 2=I 3=J; EE index layout (p,q,i,j) */
static void or_store(const or_shell *P, const or_shell *Q, const or_shell *I, const or_shell *J,
                     int type, int restricted, int n, double *JT, double *KA, double *KB,
                     const double *PT, const double *PA, const double *PB, const double *EE)
{
	or_view v = {{P, Q, I, J}, {Q->nBasis * I->nBasis * J->nBasis, I->nBasis * J->nBasis, J->nBasis, 1}};
	int nk = restricted ? 1 : 2;
	double *K[2] = {KA, KB};
	const double *PK[2] = {PA, PB};
	switch (type) {
	case OR_ALLSAME:
		or_jblock(&v, 0, 1, 2, 3, 1.0, 0, n, JT, PT, EE);
		for (int s = 0; s < nk; s++) or_kblock(&v, 0, 1, 2, 3, 0, n, K[s], PK[s], EE);
		break;
	case OR_2PAIRS:
		or_jblock(&v, 0, 1, 2, 3, 1.0, 0, n, JT, PT, EE);
		or_jblock(&v, 2, 3, 0, 1, 1.0, 0, n, JT, PT, EE);
		for (int s = 0; s < nk; s++) or_kblock(&v, 0, 1, 2, 3, 1, n, K[s], PK[s], EE);
		break;
	case OR_PQPAIR:
		or_jblock(&v, 0, 1, 2, 3, 2.0, 0, n, JT, PT, EE);
		or_jblock(&v, 2, 3, 0, 1, 1.0, 1, n, JT, PT, EE);
		for (int s = 0; s < nk; s++) {
			or_kblock(&v, 0, 1, 2, 3, 1, n, K[s], PK[s], EE);
			or_kblock(&v, 0, 1, 3, 2, 1, n, K[s], PK[s], EE);
		}
		break;
	case OR_IJPAIR:
		or_jblock(&v, 0, 1, 2, 3, 1.0, 1, n, JT, PT, EE);
		or_jblock(&v, 2, 3, 0, 1, 2.0, 0, n, JT, PT, EE);
		for (int s = 0; s < nk; s++) {
			or_kblock(&v, 0, 1, 2, 3, 1, n, K[s], PK[s], EE);
			or_kblock(&v, 1, 0, 2, 3, 1, n, K[s], PK[s], EE);
		}
		break;
	case OR_PIQJ:
		or_jblock(&v, 0, 1, 2, 3, 2.0, 1, n, JT, PT, EE);
		for (int s = 0; s < nk; s++) {
			or_kblock(&v, 0, 1, 2, 3, 0, n, K[s], PK[s], EE);
			or_kblock(&v, 0, 1, 3, 2, 0, n, K[s], PK[s], EE);
			or_kblock(&v, 1, 0, 2, 3, 0, n, K[s], PK[s], EE);
			or_kblock(&v, 1, 0, 3, 2, 0, n, K[s], PK[s], EE);
		}
		break;
	default: /* OR_ALLDISTINCT */
		or_jblock(&v, 0, 1, 2, 3, 2.0, 1, n, JT, PT, EE);
		or_jblock(&v, 2, 3, 0, 1, 2.0, 1, n, JT, PT, EE);
		for (int s = 0; s < nk; s++) {
			or_kblock(&v, 0, 1, 2, 3, 1, n, K[s], PK[s], EE);
			or_kblock(&v, 0, 1, 3, 2, 1, n, K[s], PK[s], EE);
			or_kblock(&v, 1, 0, 2, 3, 1, n, K[s], PK[s], EE);
			or_kblock(&v, 1, 0, 3, 2, 1, n, K[s], PK[s], EE);
		}
		break;
	}
}

/* ------------------------------------------------------------------ */
/* J/K driver with the reference's static state held in a context      */
/* (src/matrix.c:3703-3986)                                            */
/* ------------------------------------------------------------------ */

typedef struct {
	int n, nsh, first;
	or_shell *sh;
	double *sch, *PM, *PT;
	or_mpole *mp;
	signed char *er;
	double *bank;
	double memcut;
	int use_mpole;   /* 1: follow the reference's monopole substitution */
	int maxMem;      /* MB for the integral bank (0: no bank)           */
} or_jk;

or_jk *or_jk_create(int use_mpole, int maxMem)
{
	or_jk *c = calloc(1, sizeof(or_jk));
	c->first = 1; c->use_mpole = use_mpole; c->maxMem = maxMem;
	return c;
}

void or_jk_reset(or_jk *c)
{
	free(c->sh); free(c->sch); free(c->PM); free(c->PT); free(c->mp); free(c->er); free(c->bank);
	int um = c->use_mpole, mm = c->maxMem;
	memset(c, 0, sizeof(*c));
	c->first = 1; c->use_mpole = um; c->maxMem = mm;
}

void or_jk_destroy(or_jk *c) { or_jk_reset(c); free(c); }

#define OR_LOOP_BEGIN                                               \
	for (int P = 0; P < ns; P++)                                      \
	for (int Q = 0; Q <= P; Q++)                                      \
	for (int I = 0; I <= P; I++) {                                    \
		int Nmax = (I == P) ? Q : I;                                  \
		for (int J = 0; J <= Nmax; J++) {
#define OR_LOOP_END }}

static double or_memcut(const or_jk *c, double fixedCutoff)
{
	int ns = c->nsh;
	const or_shell *sh = c->sh;
	/* a zero start would never grow (the reference loops forever there) */
	double cut = fixedCutoff > 1.0e-300 ? fixedCutoff : 1.0e-300;
	unsigned long maxEE = (unsigned long)c->maxMem * 1000000UL / sizeof(double);
	int coarse = 1;
	for (;;) {
		unsigned long cnt = 0;
		int over = 0;
		OR_LOOP_BEGIN
			if (over) continue;
			if (c->sch[P * ns + Q] * c->sch[I * ns + J] < cut) continue;
			cnt += (unsigned long)sh[P].nBasis * sh[Q].nBasis * sh[I].nBasis * sh[J].nBasis;
			if (cnt > maxEE) over = 1;
		OR_LOOP_END
		if (over) { cut *= coarse ? 2.0 : 1.05; continue; }
		if (coarse) { coarse = 0; cut /= 2.0; continue; }
		break;
	}
	if (cut < fixedCutoff) cut = fixedCutoff;
	return cut;
}

/* Accumulates JT += J, KA -= K_alpha, KB -= K_beta and symmetrises
 * lower -> upper, exactly as the reference entry does. */
void or_jk_build(or_jk *c, int n, const double *PA, const double *PB, const struct GTOBasis_t *g,
                 const double *schb, double fixedCutoff, double SCFCutoff, int restricted,
                 double *JT, double *KA, double *KB)
{
	static double EE[50625];
	if (c->first) {
		c->n = n;
		c->sh = or_make_shells(n, g, &c->nsh);
		int ns = c->nsh;
		c->mp = or_make_mpole(n, g);
		c->sch = calloc((size_t)ns * ns, sizeof(double));
		c->PM = calloc((size_t)ns * ns, sizeof(double));
		c->PT = calloc((size_t)n * n, sizeof(double));
		for (int P = 0; P < ns; P++)
			for (int Q = 0; Q < ns; Q++) {
				double r = 0.0;
				for (int p = c->sh[P].min; p <= c->sh[P].max; p++)
					for (int q = c->sh[Q].min; q <= c->sh[Q].max; q++)
						if (r < schb[p * n + q]) r = schb[p * n + q];
				c->sch[P * ns + Q] = r;
			}
		unsigned long ner = 0;
		OR_LOOP_BEGIN
			if (c->sch[P * ns + Q] * c->sch[I * ns + J] < SCFCutoff) continue;
			ner++;
		OR_LOOP_END
		c->er = malloc(ner + 1);
		memset(c->er, OR_ERR_UNKNOWN, ner + 1);
		c->memcut = (c->maxMem > 0) ? or_memcut(c, fixedCutoff) : INFINITY;
		unsigned long nbank = 0;
		if (c->maxMem > 0) {
			OR_LOOP_BEGIN
				if (c->sch[P * ns + Q] * c->sch[I * ns + J] < c->memcut) continue;
				nbank += (unsigned long)c->sh[P].nBasis * c->sh[Q].nBasis * c->sh[I].nBasis * c->sh[J].nBasis;
			OR_LOOP_END
		}
		c->bank = malloc((nbank + 1) * sizeof(double));
	}
	int ns = c->nsh;
	or_shell *sh = c->sh;
	for (int k = 0; k < n * n; k++) c->PT[k] = PA[k] + PB[k];
	for (int P = 0; P < ns; P++)
		for (int Q = 0; Q < ns; Q++) {
			double r = 0.0;
			for (int p = sh[P].min; p <= sh[P].max; p++)
				for (int q = sh[Q].min; q <= sh[Q].max; q++) {
					double t = fabs(c->PT[p * n + q]);
					if (r < t) r = t;
					t = fabs(PA[p * n + q]); if (r < t) r = t;
					t = fabs(PB[p * n + q]); if (r < t) r = t;
				}
			c->PM[P * ns + Q] = r;
		}
	signed char *er = c->er;
	double *bank = c->bank;
	OR_LOOP_BEGIN
		double r = c->sch[P * ns + Q] * c->sch[I * ns + J];
		if (r < SCFCutoff) continue;
		signed char erExp = *er++;
		if (r < fixedCutoff) continue;
		double *PMm = c->PM;
		double maxP = PMm[P * ns + Q] + PMm[I * ns + J] + PMm[P * ns + I] + PMm[P * ns + J] +
		              PMm[Q * ns + I] + PMm[Q * ns + J];
		int nEE = sh[P].nBasis * sh[Q].nBasis * sh[I].nBasis * sh[J].nBasis;
		if (r >= c->memcut) {
			if (c->first) {
				or_quartet(&sh[P], &sh[Q], &sh[I], &sh[J], EE);
				memcpy(bank, EE, sizeof(double) * nEE);
				bank += nEE;
			} else {
				if (r * maxP < fixedCutoff) { bank += nEE; continue; }
				memcpy(EE, bank, sizeof(double) * nEE);
				bank += nEE;
			}
		} else {
			if (r * maxP < fixedCutoff) continue;
			if (c->use_mpole && erExp != OR_ERR_UNKNOWN && pow(1.5, erExp) * maxP < fixedCutoff)
				or_mpole_quartet(&sh[P], &sh[Q], &sh[I], &sh[J], c->mp, n, EE);
			else
				or_quartet(&sh[P], &sh[Q], &sh[I], &sh[J], EE);
		}
		if (erExp == OR_ERR_UNKNOWN)
			er[-1] = or_mpole_err(&sh[P], &sh[Q], &sh[I], &sh[J], c->mp, n, EE);
		int type;
		if (P == Q && I == J && P == I) type = OR_ALLSAME;
		else if (P == Q && I == J) type = OR_2PAIRS;
		else if (P == Q) type = OR_PQPAIR;
		else if (I == J) type = OR_IJPAIR;
		else if (P == I && Q == J) type = OR_PIQJ;
		else type = OR_ALLDISTINCT;
		or_store(&sh[P], &sh[Q], &sh[I], &sh[J], type, restricted, n, JT, KA, KB, c->PT, PA, PB, EE);
	OR_LOOP_END
	if (restricted) memcpy(KB, KA, sizeof(double) * n * n);
	for (int p = 0; p < n; p++)
		for (int q = 0; q < p; q++) {
			JT[q * n + p] = JT[p * n + q];
			KA[q * n + p] = KA[p * n + q];
			KB[q * n + p] = KB[p * n + q];
		}
	c->first = 0;
}

/* Brute-force dense J/K from all n^4 contracted integrals: the semantics
 * of GTO_JK_Matrix_NoSymm (src/matrix.c:5098-5183) for small n. */
void or_jk_dense(int n, const struct GTOBasis_t *g, const double *PA, const double *PB,
                 double *JT, double *KA, double *KB)
{
	for (int k = 0; k < n * n; k++) JT[k] = KA[k] = KB[k] = 0.0;
	for (int p = 0; p < n; p++)
	for (int q = 0; q < n; q++)
	for (int i = 0; i < n; i++)
	for (int j = 0; j < n; j++) {
		double v = or_contr_eri(g, p, q, i, j);
		JT[p * n + q] += v * (PA[i * n + j] + PB[i * n + j]);
		KA[p * n + i] -= v * PA[q * n + j];
		KB[p * n + i] -= v * PB[q * n + j];
	}
}

/* ------------------------------------------------------------------ */
/* Molecule and basis input (src/mol.c, src/basis.c, src/util.c)       */
/* ------------------------------------------------------------------ */

static const char *OR_SYM[] = {"", "H", "HE", "LI", "BE", "B", "C", "N", "O", "F", "NE",
	"NA", "MG", "AL", "SI", "P", "S", "CL", "AR", "K", "CA", "SC", "TI", "V", "CR", "MN",
	"FE", "CO", "NI", "CU", "ZN", "GA", "GE", "AS", "SE", "BR", "KR"};
static const char *OR_NAME[] = {"", "HYDROGEN", "HELIUM", "LITHIUM", "BERYLLIUM", "BORON",
	"CARBON", "NITROGEN", "OXYGEN", "FLUORINE", "NEON", "SODIUM", "MAGNESIUM", "ALUMINUM",
	"SILICON", "PHOSPHORUS", "SULFUR", "CHLORINE", "ARGON", "POTASSIUM", "CALCIUM",
	"SCANDIUM", "TITANIUM", "VANADIUM", "CHROMIUM", "MANGANESE", "IRON", "COBALT", "NICKEL",
	"COPPER", "ZINC", "GALLIUM", "GERMANIUM", "ARSENIC", "SELENIUM", "BROMINE", "KRYPTON"};
#define OR_NELEM 36

static int or_lookup(const char *s, const char **tab)
{
	char u[64];
	size_t k;
	for (k = 0; s[k] && k < 63; k++) u[k] = (char)toupper((unsigned char)s[k]);
	u[k] = 0;
	for (int z = 1; z <= OR_NELEM; z++)
		if (strcmp(u, tab[z]) == 0) return z;
	return 0;
}

/* XYZ (angstrom) -> bohr; charge set to 0 (src/mol.c:92-166) */
struct Molecule_t *or_read_xyz(const char *path)
{
	FILE *f = fopen(path, "r");
	if (!f) return NULL;
	char line[512];
	struct Molecule_t *m = calloc(1, sizeof(*m));
	if (!fgets(line, sizeof line, f) || sscanf(line, "%d", &m->nAtom) != 1) { fclose(f); free(m); return NULL; }
	if (!fgets(line, sizeof line, f)) { fclose(f); free(m); return NULL; }
	m->Z = calloc(m->nAtom, sizeof(int));
	m->x = calloc(m->nAtom, sizeof(double));
	m->y = calloc(m->nAtom, sizeof(double));
	m->z = calloc(m->nAtom, sizeof(double));
	for (int i = 0; i < m->nAtom; i++) {
		char sym[64];
		if (fscanf(f, "%63s %lf %lf %lf", sym, &m->x[i], &m->y[i], &m->z[i]) != 4) { fclose(f); return NULL; }
		m->Z[i] = or_lookup(sym, OR_SYM);
		m->x[i] *= SQ_ANGSTROM2BOHR; m->y[i] *= SQ_ANGSTROM2BOHR; m->z[i] *= SQ_ANGSTROM2BOHR;
	}
	fclose(f);
	return m;
}

double or_nuclei_coulomb(const struct Molecule_t *m)
{
	double U = 0.0;
	for (int i = 0; i < m->nAtom; i++)
		for (int j = 0; j < i; j++)
			U += m->Z[i] * m->Z[j] / sqrt((m->x[i] - m->x[j]) * (m->x[i] - m->x[j]) +
			                              (m->y[i] - m->y[j]) * (m->y[i] - m->y[j]) +
			                              (m->z[i] - m->z[j]) * (m->z[i] - m->z[j]));
	return U;
}

/* GAMESS $DATA reader (src/basis.c:316-514): element long names, then
 * S/L/P/D/F blocks; 'D' exponent markers accepted. */
struct GTOBasisSet_t *or_read_basis(const char *path, int *nset)
{
	FILE *f = fopen(path, "r");
	if (!f) return NULL;
	char w[1024];
	int found = 0;
	while (fscanf(f, "%1023s", w) == 1)
		if (strcmp(w, "$DATA") == 0) { found = 1; break; }
	if (!found) { fclose(f); return NULL; }
	int cap = 4096, ns = 0, Z = 0;
	struct GTOBasisSet_t *db = calloc(cap, sizeof(*db));
	static const int comp[5][10][3] = {
		{{0, 0, 0}},
		{{0, 0, 0}},
		{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}},
		{{2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}},
		{{3, 0, 0}, {0, 3, 0}, {0, 0, 3}, {2, 1, 0}, {1, 2, 0}, {2, 0, 1}, {1, 0, 2}, {0, 2, 1}, {0, 1, 2}, {1, 1, 1}}};
	while (fscanf(f, "%1023s", w) == 1 && strcmp(w, "$END") != 0) {
		int z = or_lookup(w, OR_NAME);
		if (z > 0) { Z = z; continue; }
		int kind = -1;
		if (!strcmp(w, "S")) kind = 0; else if (!strcmp(w, "L")) kind = 1;
		else if (!strcmp(w, "P")) kind = 2; else if (!strcmp(w, "D")) kind = 3;
		else if (!strcmp(w, "F")) kind = 4;
		if (kind < 0) continue;
		int nc;
		if (fscanf(f, "%d", &nc) != 1) break;
		double ex[64], ca[64], cb[64];
		for (int i = 0; i < nc; i++) {
			char s1[128], s2[128], s3[128];
			if (kind == 1) {
				if (fscanf(f, "%*d %127s %127s %127s", s1, s2, s3) != 3) goto fail;
			} else {
				if (fscanf(f, "%*d %127s %127s", s1, s2) != 2) goto fail;
				s3[0] = '0'; s3[1] = 0;
			}
			for (char *c = s1; *c; c++) if (*c == 'D') *c = 'E';
			for (char *c = s2; *c; c++) if (*c == 'D') *c = 'E';
			for (char *c = s3; *c; c++) if (*c == 'D') *c = 'E';
			ex[i] = strtod(s1, NULL); ca[i] = strtod(s2, NULL); cb[i] = strtod(s3, NULL);
		}
		int ncomp = (kind == 0) ? 1 : (kind == 1) ? 4 : (kind == 2) ? 3 : (kind == 3) ? 6 : 10;
		for (int k = 0; k < ncomp; k++) {
			if (ns >= cap) { cap *= 2; db = realloc(db, cap * sizeof(*db)); }
			struct GTOBasisSet_t *b = &db[ns++];
			memset(b, 0, sizeof(*b));
			b->Z = Z; b->nContract = nc;
			const int *c3;
			const double *cc;
			if (kind == 1) { c3 = (k == 0) ? comp[0][0] : comp[2][k - 1]; cc = (k == 0) ? ca : cb; }
			else { c3 = comp[kind][k]; cc = ca; }
			b->l = c3[0]; b->m = c3[1]; b->n = c3[2];
			b->exp = malloc(nc * sizeof(double));
			b->coef = malloc(nc * sizeof(double));
			for (int i = 0; i < nc; i++) { b->exp[i] = ex[i]; b->coef[i] = cc[i]; }
		}
	}
	fclose(f);
	*nset = ns;
	return db;
fail:
	fclose(f);
	return NULL;
}

/* expand per atom in file order; primitive norms 1/sqrt(<g|g>) for each
 * Cartesian component (src/basis.c:529-623) */
struct GTOBasis_t *or_gen_basis(const struct Molecule_t *m, int nset, const struct GTOBasisSet_t *db, int *nb)
{
	int cap = 0;
	for (int i = 0; i < m->nAtom; i++)
		for (int b = 0; b < nset; b++) if (db[b].Z == m->Z[i]) cap++;
	struct GTOBasis_t *g = calloc(cap ? cap : 1, sizeof(*g));
	int n = 0;
	for (int i = 0; i < m->nAtom; i++) {
		int any = 0;
		for (int b = 0; b < nset; b++) {
			if (db[b].Z != m->Z[i]) continue;
			any = 1;
			struct GTOBasis_t *t = &g[n++];
			t->nContract = db[b].nContract;
			t->l = db[b].l; t->m = db[b].m; t->n = db[b].n;
			t->x0 = m->x[i]; t->y0 = m->y[i]; t->z0 = m->z[i];
			t->exp = malloc(t->nContract * sizeof(double));
			t->coef = malloc(t->nContract * sizeof(double));
			t->norm = malloc(t->nContract * sizeof(double));
			for (int k = 0; k < t->nContract; k++) {
				t->exp[k] = db[b].exp[k];
				t->coef[k] = db[b].coef[k];
				double s = or_overlap(t->exp[k], t->l, t->m, t->n, t->x0, t->y0, t->z0,
				                      t->exp[k], t->l, t->m, t->n, t->x0, t->y0, t->z0);
				t->norm[k] = 1.0 / sqrt(s);
			}
		}
		if (!any) { *nb = -m->Z[i]; return g; }
	}
	*nb = n;
	return g;
}

/* ------------------------------------------------------------------ */
/* Dense linear algebra for the SCF (own code; LAPACK semantics of     */
/* src/lin.c:66-391)                                                   */
/* ------------------------------------------------------------------ */

/* cyclic Jacobi on a symmetric n x n matrix A (destroyed); eigenvalues
 * ascending in w, eigenvector k in V[k*n + i] */
static void or_jacobi(int n, double *A, double *w, double *V)
{
	double *Z = calloc((size_t)n * n, sizeof(double));
	for (int i = 0; i < n; i++) Z[i * n + i] = 1.0;
	for (int sweep = 0; sweep < 100; sweep++) {
		double off = 0.0, tot = 0.0;
		for (int i = 0; i < n; i++)
			for (int j = 0; j < n; j++) {
				double a = A[i * n + j] * A[i * n + j];
				tot += a;
				if (i != j) off += a;
			}
		if (off <= 1e-30 * tot || off == 0.0) break;
		for (int p = 0; p < n; p++)
			for (int q = p + 1; q < n; q++) {
				double apq = A[p * n + q];
				if (fabs(apq) < 1e-300) continue;
				double th = (A[q * n + q] - A[p * n + p]) / (2.0 * apq);
				double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
				double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
				for (int k = 0; k < n; k++) {
					double akp = A[k * n + p], akq = A[k * n + q];
					A[k * n + p] = c * akp - s * akq;
					A[k * n + q] = s * akp + c * akq;
				}
				for (int k = 0; k < n; k++) {
					double apk = A[p * n + k], aqk = A[q * n + k];
					A[p * n + k] = c * apk - s * aqk;
					A[q * n + k] = s * apk + c * aqk;
				}
				for (int k = 0; k < n; k++) {
					double zkp = Z[k * n + p], zkq = Z[k * n + q];
					Z[k * n + p] = c * zkp - s * zkq;
					Z[k * n + q] = s * zkp + c * zkq;
				}
			}
	}
	int *idx = malloc(n * sizeof(int));
	for (int i = 0; i < n; i++) idx[i] = i;
	for (int i = 0; i < n; i++)
		for (int j = i + 1; j < n; j++)
			if (A[idx[j] * n + idx[j]] < A[idx[i] * n + idx[i]]) { int t = idx[i]; idx[i] = idx[j]; idx[j] = t; }
	for (int k = 0; k < n; k++) {
		w[k] = A[idx[k] * n + idx[k]];
		for (int i = 0; i < n; i++) V[k * n + i] = Z[i * n + idx[k]];
	}
	free(idx); free(Z);
}

/* F C = S C e via Cholesky S = L L^T; all eigenpairs; C[k*n+i] */
int or_gen_eigen(int n, const double *F, const double *S, double *e, double *C)
{
	double *L = calloc((size_t)n * n, sizeof(double));
	for (int j = 0; j < n; j++) {
		double d = S[j * n + j];
		for (int k = 0; k < j; k++) d -= L[j * n + k] * L[j * n + k];
		if (d <= 0.0) { free(L); return 0; }
		L[j * n + j] = sqrt(d);
		for (int i = j + 1; i < n; i++) {
			double s = S[i * n + j];
			for (int k = 0; k < j; k++) s -= L[i * n + k] * L[j * n + k];
			L[i * n + j] = s / L[j * n + j];
		}
	}
	/* A = L^-1 F L^-T */
	double *T = malloc((size_t)n * n * sizeof(double)), *A = malloc((size_t)n * n * sizeof(double));
	for (int c = 0; c < n; c++)          /* T = L^-1 F, column by column */
		for (int i = 0; i < n; i++) {
			double s = F[i * n + c];
			for (int k = 0; k < i; k++) s -= L[i * n + k] * T[k * n + c];
			T[i * n + c] = s / L[i * n + i];
		}
	for (int r = 0; r < n; r++)          /* A = T L^-T, row by row */
		for (int i = 0; i < n; i++) {
			double s = T[r * n + i];
			for (int k = 0; k < i; k++) s -= L[i * n + k] * A[r * n + k];
			A[r * n + i] = s / L[i * n + i];
		}
	for (int i = 0; i < n; i++)
		for (int j = 0; j < i; j++) { double s = 0.5 * (A[i * n + j] + A[j * n + i]); A[i * n + j] = A[j * n + i] = s; }
	double *Y = malloc((size_t)n * n * sizeof(double));
	or_jacobi(n, A, e, Y);
	for (int k = 0; k < n; k++)          /* C_k = L^-T y_k */
		for (int i = n - 1; i >= 0; i--) {
			double s = Y[k * n + i];
			for (int j = i + 1; j < n; j++) s -= L[j * n + i] * C[k * n + j];
			C[k * n + i] = s / L[i * n + i];
		}
	free(L); free(T); free(A); free(Y);
	return 1;
}

/* Gaussian elimination with partial pivoting, row-major A (n x n) */
static int or_solve(int n, double *A, double *b)
{
	for (int c = 0; c < n; c++) {
		int p = c;
		for (int r = c + 1; r < n; r++) if (fabs(A[r * n + c]) > fabs(A[p * n + c])) p = r;
		if (A[p * n + c] == 0.0) return 0;
		if (p != c) {
			for (int k = 0; k < n; k++) { double t = A[c * n + k]; A[c * n + k] = A[p * n + k]; A[p * n + k] = t; }
			double t = b[c]; b[c] = b[p]; b[p] = t;
		}
		for (int r = c + 1; r < n; r++) {
			double f = A[r * n + c] / A[c * n + c];
			for (int k = c; k < n; k++) A[r * n + k] -= f * A[c * n + k];
			b[r] -= f * b[c];
		}
	}
	for (int r = n - 1; r >= 0; r--) {
		double s = b[r];
		for (int k = r + 1; k < n; k++) s -= A[r * n + k] * b[k];
		b[r] = s / A[r * n + r];
	}
	return 1;
}

/* ------------------------------------------------------------------ */
/* Density DIIS, four points (src/conv.c:606-844)                      */
/* ------------------------------------------------------------------ */

typedef struct {
	int it, n;
	double *P[4][2], *dP[4][2];    /* slot 3 = most recent input */
} or_diis;

static void or_diis_free(or_diis *d)
{
	for (int s = 0; s < 4; s++)
		for (int k = 0; k < 2; k++) { free(d->P[s][k]); free(d->dP[s][k]); d->P[s][k] = d->dP[s][k] = NULL; }
	d->it = 0;
}

static double or_diis4(or_diis *d, int n, double alpha, double *PA, double *PB)
{
	size_t nn = (size_t)n * n;
	double *cur[2] = {PA, PB};
	if (d->it == 0) {
		d->n = n;
		for (int s = 0; s < 4; s++)
			for (int k = 0; k < 2; k++) { d->P[s][k] = calloc(nn, sizeof(double)); d->dP[s][k] = calloc(nn, sizeof(double)); }
		for (int k = 0; k < 2; k++) memcpy(d->P[3][k], cur[k], nn * sizeof(double));
		d->it = 1;
		return 0.0;
	}
	double rms = 0.0;
	for (int k = 0; k < 2; k++)
		for (size_t e = 0; e < nn; e++) {
			double t = cur[k][e] - d->P[3][k][e];
			rms += t * t;
			d->dP[3][k][e] = t;
		}
	rms = sqrt(rms / n / n / 2.0);
	if (d->it <= 3) {
		for (int k = 0; k < 2; k++)
			for (size_t e = 0; e < nn; e++) cur[k][e] = alpha * cur[k][e] + (1.0 - alpha) * d->P[3][k][e];
	} else {
		double B[4][4] = {{0}};
		for (int a = 0; a < 4; a++)
			for (int b = a; b < 4; b++) {
				double s = 0.0;
				for (int k = 0; k < 2; k++)
					for (size_t e = 0; e < nn; e++) s += d->dP[a][k][e] * d->dP[b][k][e];
				B[a][b] = B[b][a] = s;
			}
		double A[25], rhs[5] = {0, 0, 0, 0, -1};
		for (int a = 0; a < 5; a++)
			for (int b = 0; b < 5; b++)
				A[a * 5 + b] = (a < 4 && b < 4) ? B[a][b] : ((a == 4 && b == 4) ? 0.0 : -1.0);
		or_solve(5, A, rhs);
		double c[4] = {rhs[0], rhs[1], rhs[2], rhs[3]};
		if (B[3][3] > B[2][2]) { c[3] = alpha; c[2] = 1.0 - alpha; c[1] = 0.0; c[0] = 0.0; }
		for (int k = 0; k < 2; k++)
			for (size_t e = 0; e < nn; e++)
				cur[k][e] = c[3] * cur[k][e] + c[2] * (d->P[2][k][e] + d->dP[2][k][e]) +
				            c[1] * (d->P[1][k][e] + d->dP[1][k][e]) + c[0] * (d->P[0][k][e] + d->dP[0][k][e]);
	}
	for (int k = 0; k < 2; k++) {
		double *t = d->dP[0][k]; d->dP[0][k] = d->dP[1][k]; d->dP[1][k] = d->dP[2][k]; d->dP[2][k] = d->dP[3][k]; d->dP[3][k] = t;
		double *u = d->P[0][k]; d->P[0][k] = d->P[1][k]; d->P[1][k] = d->P[2][k]; d->P[2][k] = d->P[3][k]; d->P[3][k] = u;
		memcpy(d->P[3][k], cur[k], nn * sizeof(double));
	}
	d->it++;
	return rms;
}

/* ------------------------------------------------------------------ */
/* SCF (src/uhf.c:1171-1783), Hartree-Fock branch, DIAG or CORE guess  */
/* ------------------------------------------------------------------ */

/* Returns E (0.0 if not converged).  iters receives the iteration count. */
double or_uhf(int n, const struct GTOBasis_t *g, const struct Molecule_t *mol, int nEA, int nEB,
              int restricted, double SCFConv, int SCFMax, double drag, int accuracy3step,
              int guess, int use_mpole, int maxMem, double *CA, double *CB, double *eA, double *eB,
              int *iters)
{
	size_t nn = (size_t)n * n;
	double *H = calloc(nn, 8), *S = calloc(nn, 8), *JT = calloc(nn, 8), *KA = calloc(nn, 8), *KB = calloc(nn, 8);
	double *dJ = calloc(nn, 8), *dKA = calloc(nn, 8), *dKB = calloc(nn, 8), *PA = calloc(nn, 8), *PB = calloc(nn, 8);
	double *dPA = calloc(nn, 8), *dPB = calloc(nn, 8), *FA = calloc(nn, 8), *FB = calloc(nn, 8), *sch = calloc(nn, 8);
	for (int i = 0; i < n; i++)
		for (int j = 0; j <= i; j++) {
			double t = or_gto_kinetic(i, j, g) + or_gto_nuclei(i, j, g, mol);
			H[i * n + j] = H[j * n + i] = t;
			double s = or_gto_overlap(i, j, g);
			S[i * n + j] = S[j * n + i] = s;
		}
	if (guess == SCFGUESS_CORE) {
		or_gen_eigen(n, H, S, eA, CA);
		memcpy(CB, CA, nn * 8);
		for (int i = 0; i < n; i++)
			for (int j = 0; j < n; j++) {
				double a = 0, b = 0;
				for (int k = 0; k < nEA; k++) a += CA[k * n + i] * CA[k * n + j];
				for (int k = 0; k < nEB; k++) b += CB[k * n + i] * CB[k * n + j];
				PA[i * n + j] = a; PB[i * n + j] = b;
			}
	} else {
		double tr = 0.0;
		for (int i = 0; i < n; i++) tr += S[i * n + i];
		for (int i = 0; i < n; i++) { PA[i * n + i] = nEA / tr; PB[i * n + i] = nEB / tr; }
	}
	double sA = 0.0, sB = 0.0;
	for (size_t e = 0; e < nn; e++) { sA += PA[e]; sB += PB[e]; }
#define OR_CUT(acc) ((acc) / (sA * sA + sA * sB + sB * sB) / 50.0)
	double realCut = OR_CUT(SCFConv);
	or_schwarz(n, g, sch);
	or_diis dd = {0};
	or_diis4(&dd, n, drag, PA, PB);
	memcpy(dPA, PA, nn * 8); memcpy(dPB, PB, nn * 8);
	double cA = OR_CUT(SCFACCURACY_3STEP_A), cB = OR_CUT(SCFACCURACY_3STEP_B), cC = OR_CUT(SCFACCURACY_3STEP_C);
	double cur = accuracy3step ? cA : realCut;
	or_jk *jk = or_jk_create(use_mpole, maxMem);
	double E = 0.0, dE = 0.0, rms = 0.0;
	int it = 0, notconv = 1;
	double Enn = or_nuclei_coulomb(mol);
	do {
		it++;
		for (size_t e = 0; e < nn; e++) dJ[e] = dKA[e] = dKB[e] = 0.0;
		or_jk_build(jk, n, dPA, restricted ? dPA : dPB, g, sch, cur, realCut, restricted, dJ, dKA, dKB);
		for (size_t e = 0; e < nn; e++) {
			JT[e] += dJ[e]; KA[e] += dKA[e]; KB[e] += dKB[e];
			dPA[e] = PA[e]; dPB[e] = PB[e];
			FA[e] = H[e] + JT[e] + KA[e];
			FB[e] = H[e] + JT[e] + KB[e];
		}
		or_gen_eigen(n, FA, S, eA, CA);
		or_gen_eigen(n, FB, S, eB, CB);
		for (int i = 0; i < n; i++)
			for (int j = 0; j < n; j++) {
				double a = 0, b = 0;
				for (int k = 0; k < nEA; k++) a += CA[k * n + i] * CA[k * n + j];
				for (int k = 0; k < nEB; k++) b += CB[k * n + i] * CB[k * n + j];
				PA[i * n + j] = a; PB[i * n + j] = b;
			}
		double Eold = E;
		E = Enn;
		for (size_t e = 0; e < nn; e++)
			E += (H[e] + 0.5 * JT[e]) * (PA[e] + PB[e]) + 0.5 * KA[e] * PA[e] + 0.5 * KB[e] * PB[e];
		dE = E - Eold;
		rms = or_diis4(&dd, n, drag, PA, PB);
		notconv = fabs(dE) > SCFConv || rms > SCFConv;
		for (size_t e = 0; e < nn; e++) { dPA[e] = PA[e] - dPA[e]; dPB[e] = PB[e] - dPB[e]; }
		if (it >= SCFMax) break;
		if (accuracy3step) {
			if (!notconv) break;
			if ((cur == cA && fabs(dE) <= SCFACCURACY_3STEP_A && rms <= SCFACCURACY_3STEP_A) ||
			    (cur == cB && fabs(dE) <= SCFACCURACY_3STEP_B && rms <= SCFACCURACY_3STEP_B) ||
			    (cur == cC && fabs(dE) <= SCFACCURACY_3STEP_C && rms <= SCFACCURACY_3STEP_C)) {
				if (cur == cA) cur = cB; else if (cur == cB) cur = cC; else if (cur == cC) cur = realCut;
			}
		}
	} while (notconv);
#undef OR_CUT
	or_jk_destroy(jk);
	or_diis_free(&dd);
	if (iters) *iters = it;
	free(H); free(S); free(JT); free(KA); free(KB); free(dJ); free(dKA); free(dKB);
	free(PA); free(PB); free(dPA); free(dPB); free(FA); free(FB); free(sch);
	return notconv ? 0.0 : E;
}

/* ------------------------------------------------------------------ */
/* Forces (src/grad.c)                                                 */
/* ------------------------------------------------------------------ */

#define OR_GRAD_CUTOFF 1.0e-10          /* GRAD_SCHWARZ_CUTOFF, src/grad.h:20 */

/* nuclear repulsion gradient, added into G (src/grad.c:43-66) */
void or_grad_vnn(const struct Molecule_t *m, double *Gx, double *Gy, double *Gz)
{
	for (int a = 0; a < m->nAtom; a++)
		for (int b = 0; b < a; b++) {
			double dx = m->x[a] - m->x[b], dy = m->y[a] - m->y[b], dz = m->z[a] - m->z[b];
			double r = sqrt(dx * dx + dy * dy + dz * dz);
			double f = m->Z[a] * m->Z[b] / (r * r * r);
			Gx[a] -= f * dx; Gx[b] += f * dx;
			Gy[a] -= f * dy; Gy[b] += f * dy;
			Gz[a] -= f * dz; Gz[b] += f * dz;
		}
}

/* one primitive pair of a one-electron operator with the angular powers of
 * function i replaced by (l, m, n) */
enum { OR_1E_S, OR_1E_T, OR_1E_V, OR_1E_M };
typedef struct { const struct Molecule_t *mol; int mx, my, mz; double cx, cy, cz; } or_1e_op;

static double or_1e_prim(int kind, const or_1e_op *op, const struct GTOBasis_t *g,
                         int i, int a, int l, int m, int n, int j, int b)
{
	const struct GTOBasis_t *A = g + i, *B = g + j;
	switch (kind) {
	case OR_1E_S:
		return or_overlap(A->exp[a], l, m, n, A->x0, A->y0, A->z0, B->exp[b], B->l, B->m, B->n, B->x0, B->y0, B->z0);
	case OR_1E_T:
		return or_kinetic(A->exp[a], l, m, n, A->x0, A->y0, A->z0, B->exp[b], B->l, B->m, B->n, B->x0, B->y0, B->z0);
	case OR_1E_M:
		return or_moment(A->exp[a], l, m, n, A->x0, A->y0, A->z0, B->exp[b], B->l, B->m, B->n, B->x0, B->y0, B->z0,
		                 op->mx, op->my, op->mz, op->cx, op->cy, op->cz);
	default: {
		double s = 0.0;
		for (int k = 0; k < op->mol->nAtom; k++)
			s += (double)op->mol->Z[k] *
			     or_nai(A->x0, A->y0, A->z0, 1.0, l, m, n, A->exp[a], B->x0, B->y0, B->z0, 1.0, B->l, B->m, B->n,
			            B->exp[b], op->mol->x[k], op->mol->y[k], op->mol->z[k]);
		return s;
	}
	}
}

/* derivative with respect to the centre of function i along dir (0,1,2) of
 * a contracted one-electron integral: d/dX_i chi_i = 2 alpha chi_i(l+1) -
 * l chi_i(l-1) (src/grad.c:89-160 overlap, 166-247 moment, 248-315 kinetic,
 * 321-400 nuclear attraction, which differentiates the basis function only) */
static double or_grad_1e(int kind, const or_1e_op *op, const struct GTOBasis_t *g, int i, int j, int dir)
{
	if (dir < 0 || dir > 2) { printf("or_grad_1e: error unrecognized direction flag\n"); exit(-1); }
	int dl = dir == 0, dm = dir == 1, dn = dir == 2;
	int lo = dl * g[i].l + dm * g[i].m + dn * g[i].n;
	double down = 0.0, up = 0.0;
	for (int a = 0; a < g[i].nContract; a++)
		for (int b = 0; b < g[j].nContract; b++) {
			double c = g[i].coef[a] * g[j].coef[b] * g[i].norm[a] * g[j].norm[b];
			if (lo > 0) down += c * or_1e_prim(kind, op, g, i, a, g[i].l - dl, g[i].m - dm, g[i].n - dn, j, b);
			up += g[i].exp[a] * c * or_1e_prim(kind, op, g, i, a, g[i].l + dl, g[i].m + dm, g[i].n + dn, j, b);
		}
	return -(down * lo) + up + up;
}

double or_grad_overlap(int i, int j, const struct GTOBasis_t *g, int dir)
{
	or_1e_op op = {0};
	return or_grad_1e(OR_1E_S, &op, g, i, j, dir);
}

double or_grad_kinetic(int i, int j, const struct GTOBasis_t *g, int dir)
{
	or_1e_op op = {0};
	return or_grad_1e(OR_1E_T, &op, g, i, j, dir);
}

double or_grad_nai(int i, int j, const struct GTOBasis_t *g, const struct Molecule_t *mol, int dir)
{
	or_1e_op op = {0};
	op.mol = mol;
	return or_grad_1e(OR_1E_V, &op, g, i, j, dir);
}

double or_grad_moment(int i, int j, const struct GTOBasis_t *g, int mx, int my, int mz,
                      double cx, double cy, double cz, int dir)
{
	or_1e_op op = {0};
	op.mx = mx; op.my = my; op.mz = mz; op.cx = cx; op.cy = cy; op.cz = cz;
	return or_grad_1e(OR_1E_M, &op, g, i, j, dir);
}

/* Electric-field integral at C between contracted functions i and j
 * (src/matrix.c:207-240 over src/int.c:469-525).  The attraction integral
 * depends on A - C and B - C only, so its derivative with respect to C is
 * minus the sum of its derivatives with respect to A and B.  The field
 * integral is e = d/dC nai(i, j; C), with nai = -(i|1/|r-C||j). */
void or_gto_efi(int i, int j, const struct GTOBasis_t *g, double xc, double yc, double zc,
                double *ex, double *ey, double *ez)
{
	struct Molecule_t one;
	int Z = 1;
	double x = xc, y = yc, z = zc;
	one.nAtom = 1; one.Q = 0; one.Z = &Z; one.x = &x; one.y = &y; one.z = &z;
	or_1e_op op = {0};
	op.mol = &one;
	double e[3];
	for (int d = 0; d < 3; d++)
		e[d] = -(or_grad_1e(OR_1E_V, &op, g, i, j, d) + or_grad_1e(OR_1E_V, &op, g, j, i, d));
	*ex = e[0]; *ey = e[1]; *ez = e[2];
}

/* the shells of the gradient loop: runs of consecutive functions with the
 * same centre and exponents, no reordering (src/grad.c:982-1074) */
static int or_grad_shells(int n, const struct GTOBasis_t *g, int *map, int *maxL)
{
	int ns = 1;
	map[0] = 0;
	maxL[0] = g[0].l + g[0].m + g[0].n;
	for (int k = 1; k < n; k++) {
		if (or_same_shell(g, k, map[ns - 1])) {
			int L = g[k].l + g[k].m + g[k].n;
			if (L > maxL[ns - 1]) maxL[ns - 1] = L;
		} else {
			map[ns] = k;
			maxL[ns] = g[k].l + g[k].m + g[k].n;
			ns++;
		}
	}
	map[ns] = n;
	return ns;
}

/* 1-D THO B arrays of one primitive quartet, computed on first use for the
 * angular powers (l1, l2, l3, l4) of one axis */
#define OR_GB_MAXL 4
typedef struct {
	double P[3], A[3], B[3], Q[3], Cc[3], D[3], eta1, eta2, zeta;
	unsigned char have[3][OR_GB_MAXL][OR_GB_MAXL][OR_GB_MAXL][OR_GB_MAXL];
	double Bv[3][OR_GB_MAXL][OR_GB_MAXL][OR_GB_MAXL][OR_GB_MAXL][4 * OR_GB_MAXL];
} or_bcache;

static const double *or_bget(or_bcache *c, int ax, int l1, int l2, int l3, int l4)
{
	double *B = c->Bv[ax][l1][l2][l3][l4];
	if (!c->have[ax][l1][l2][l3][l4]) {
		or_eriB_1d(B, l1, l2, l3, l4, c->P[ax], c->A[ax], c->B[ax], c->Q[ax], c->Cc[ax], c->D[ax],
		           c->eta1, c->eta2, c->zeta);
		c->have[ax][l1][l2][l3][l4] = 1;
	}
	return B;
}

/* (pq|ij) of one primitive quartet without the prefactor, for powers
 * a[4][3] = (l,m,n) of p, q, i, j */
static double or_bsum(or_bcache *c, const int a[4][3], const double *F)
{
	const double *Bx = or_bget(c, 0, a[0][0], a[1][0], a[2][0], a[3][0]);
	const double *By = or_bget(c, 1, a[0][1], a[1][1], a[2][1], a[3][1]);
	const double *Bz = or_bget(c, 2, a[0][2], a[1][2], a[2][2], a[3][2]);
	int mx = a[0][0] + a[1][0] + a[2][0] + a[3][0];
	int my = a[0][1] + a[1][1] + a[2][1] + a[3][1];
	int mz = a[0][2] + a[1][2] + a[2][2] + a[3][2];
	double s = 0.0;
	for (int x = 0; x <= mx; x++)
		for (int y = 0; y <= my; y++)
			for (int z = 0; z <= mz; z++)
				s += Bx[x] * By[y] * Bz[z] * F[x + y + z];
	return s;
}

/* Two-electron part of the energy gradient, added into G per atom
 * (src/grad.c:1734-2090, GradEri_ShellSet_Parallel): shell pairs pqS with
 * pqS % nCPU == childID; shell quartets skipped when
 * Sch_pq Sch_ij PM_pq PM_ij < 1e-10 (PM: max |PT| in the shell pair); per
 * distinct basis quartet the density weight summed over its distinct
 * permutations of 0.5 PT_pq PT_ij - Kfactor/4 (PT_pi PT_qj + PS_pi PS_qj);
 * per primitive quartet the bare-prefactor cut 1e-15 and the per-term cut
 * |weight * prefactor * coefficients| < 1e-10; derivatives with respect to
 * the centres of p, q and i from raised and lowered powers, j's by
 * translational invariance. */
void or_grad_eri_cut(int childID, int nCPU, int n, const struct GTOBasis_t *g, const int *basis2Atom,
                     const double *PT, const double *PS, double Kfactor, double cutoff,
                     double *Gx, double *Gy, double *Gz)
{
	double *sch = malloc(sizeof(double) * n * n);
	or_schwarz(n, g, sch);
	int *map = malloc(sizeof(int) * (n + 1)), *maxL = malloc(sizeof(int) * (n + 1));
	int ns = or_grad_shells(n, g, map, maxL);
	for (int k = 0; k < ns; k++)
		if (maxL[k] > OR_GB_MAXL - 2 || map[k + 1] - map[k] > 16) {
			printf("or_grad_eri: shells above d are not restated\n");
			exit(-1);
		}
	double *SS = calloc((size_t)ns * ns, sizeof(double)), *PM = calloc((size_t)ns * ns, sizeof(double));
	for (int pS = 0; pS < ns; pS++)
		for (int qS = 0; qS < ns; qS++) {
			double r = 0.0, d = 0.0;
			for (int p = map[pS]; p < map[pS + 1]; p++)
				for (int q = map[qS]; q < map[qS + 1] && q < p + 1; q++) {
					if (sch[p * n + q] > r) r = sch[p * n + q];
					if (fabs(PT[p * n + q]) > d) d = fabs(PT[p * n + q]);
				}
			SS[pS * ns + qS] = r;
			PM[pS * ns + qS] = d;
		}
	double *gx = calloc(n, sizeof(double)), *gy = calloc(n, sizeof(double)), *gz = calloc(n, sizeof(double));
	double *gv[3] = {gx, gy, gz};
	double *wt = malloc(sizeof(double) * 16 * 16 * 16 * 16);
	or_bcache *bc = malloc(sizeof(or_bcache));
	double kc = 0.25 * Kfactor;
#define OR_W(a, b, c, d) (0.5 * PT[(a) * n + (b)] * PT[(c) * n + (d)] - \
                          kc * (PT[(a) * n + (c)] * PT[(b) * n + (d)] + PS[(a) * n + (c)] * PS[(b) * n + (d)]))
	for (int pqS = 0, pS = 0; pS < ns; pS++)
	for (int qS = 0; qS <= pS; qS++, pqS++) {
		if (pqS % nCPU != childID) continue;
		for (int iS = 0; iS <= pS; iS++)
		for (int jS = 0; jS <= pS; jS++) {
			if (SS[pS * ns + qS] * SS[iS * ns + jS] * PM[pS * ns + qS] * PM[iS * ns + jS] < cutoff) continue;
			/* density weights of the distinct basis quartets, in loop order */
			int nq = 0;
			for (int p = map[pS]; p < map[pS + 1]; p++)
			for (int q = map[qS]; q < map[qS + 1] && q <= p; q++)
			for (int i = map[iS]; i < map[iS + 1] && i <= p; i++)
			for (int j = map[jS]; j < map[jS + 1]; j++) {
				int lim = (i == p) ? q + 1 : i + 1;
				if (j >= lim) break;
				double w;
				if (p == q && i == j && p == i) w = OR_W(p, q, i, j);
				else if (p == q && i == j) w = OR_W(p, q, i, j) + OR_W(i, j, p, q);
				else if (p == q) w = OR_W(p, q, i, j) + OR_W(p, q, j, i) + OR_W(i, j, p, q) + OR_W(j, i, p, q);
				else if (i == j) w = OR_W(p, q, i, j) + OR_W(q, p, i, j) + OR_W(i, j, p, q) + OR_W(i, j, q, p);
				else if (p == i && q == j) w = OR_W(p, q, i, j) + OR_W(q, p, i, j) + OR_W(p, q, j, i) + OR_W(q, p, j, i);
				else w = OR_W(p, q, i, j) + OR_W(q, p, i, j) + OR_W(p, q, j, i) + OR_W(q, p, j, i) +
				         OR_W(i, j, p, q) + OR_W(j, i, p, q) + OR_W(i, j, q, p) + OR_W(j, i, q, p);
				wt[nq++] = w;
			}
			const struct GTOBasis_t *G0 = g + map[pS], *G1 = g + map[qS], *G2 = g + map[iS], *G3 = g + map[jS];
			int Lsum = maxL[pS] + maxL[qS] + maxL[iS] + maxL[jS] + 2;
			for (int a = 0; a < G0->nContract; a++)
			for (int b = 0; b < G1->nContract; b++)
			for (int c = 0; c < G2->nContract; c++)
			for (int d = 0; d < G3->nContract; d++) {
				double aa = G0->exp[a], ab = G1->exp[b], ac = G2->exp[c], ad = G3->exp[d];
				double eta1 = 1.0 / (aa + ab), eta2 = 1.0 / (ac + ad);
				double Ac[3] = {G0->x0, G0->y0, G0->z0}, Bc[3] = {G1->x0, G1->y0, G1->z0};
				double Cc[3] = {G2->x0, G2->y0, G2->z0}, Dc[3] = {G3->x0, G3->y0, G3->z0};
				double rab2 = 0, rcd2 = 0, rpq2 = 0;
				for (int x = 0; x < 3; x++) {
					bc->A[x] = Ac[x]; bc->B[x] = Bc[x]; bc->Cc[x] = Cc[x]; bc->D[x] = Dc[x];
					bc->P[x] = (aa * Ac[x] + ab * Bc[x]) * eta1;
					bc->Q[x] = (ac * Cc[x] + ad * Dc[x]) * eta2;
					rab2 += (Ac[x] - Bc[x]) * (Ac[x] - Bc[x]);
					rcd2 += (Cc[x] - Dc[x]) * (Cc[x] - Dc[x]);
					rpq2 += (bc->P[x] - bc->Q[x]) * (bc->P[x] - bc->Q[x]);
				}
				bc->eta1 = eta1; bc->eta2 = eta2;
				bc->zeta = 4.0 * (aa + ab) * (ac + ad) / (aa + ab + ac + ad);
				double t = OR_TWO_PI_POW2_5 * eta1 * eta2 / sqrt(aa + ab + ac + ad) *
				           exp(-aa * ab * rab2 * eta1 - ac * ad * rcd2 * eta2);
				if (t < OR_PRIM_CUTOFF) continue;
				double F[64];
				or_fgamma_set(Lsum, 0.25 * rpq2 * bc->zeta, F);
				memset(bc->have, 0, sizeof(bc->have));
				int k = 0;
				for (int p = map[pS]; p < map[pS + 1]; p++)
				for (int q = map[qS]; q < map[qS + 1] && q <= p; q++)
				for (int i = map[iS]; i < map[iS + 1] && i <= p; i++)
				for (int j = map[jS]; j < map[jS + 1]; j++) {
					int lim = (i == p) ? q + 1 : i + 1;
					if (j >= lim) break;
					double w = wt[k++];
					double r = t * g[p].coef[a] * g[p].norm[a] * g[q].coef[b] * g[q].norm[b] *
					           g[i].coef[c] * g[i].norm[c] * g[j].coef[d] * g[j].norm[d];
					if (fabs(w * r) < cutoff) continue;
					int fn[4] = {p, q, i, j};
					double ex[3] = {aa, ab, ac};
					for (int ax = 0; ax < 3; ax++) {
						double dd[3];
						for (int ctr = 0; ctr < 3; ctr++) {
							int pw[4][3];
							for (int f = 0; f < 4; f++) {
								pw[f][0] = g[fn[f]].l; pw[f][1] = g[fn[f]].m; pw[f][2] = g[fn[f]].n;
							}
							int lo = pw[ctr][ax];
							pw[ctr][ax] = lo + 1;
							double v = 2.0 * ex[ctr] * or_bsum(bc, pw, F);
							if (lo > 0) {
								pw[ctr][ax] = lo - 1;
								v -= lo * or_bsum(bc, pw, F);
							}
							dd[ctr] = v * r;
						}
						gv[ax][p] += w * dd[0];
						gv[ax][q] += w * dd[1];
						gv[ax][i] += w * dd[2];
						gv[ax][j] -= w * (dd[0] + dd[1] + dd[2]);
					}
				}
			}
		}
	}
#undef OR_W
	for (int p = 0; p < n; p++) {
		Gx[basis2Atom[p]] += gx[p];
		Gy[basis2Atom[p]] += gy[p];
		Gz[basis2Atom[p]] += gz[p];
	}
	free(sch); free(map); free(maxL); free(SS); free(PM); free(gx); free(gy); free(gz); free(wt); free(bc);
}

/* the reference's cut-off GRAD_SCHWARZ_CUTOFF; or_grad_eri_cut with cutoff 0
 * is the exact derivative of the two-electron energy */
void or_grad_eri(int childID, int nCPU, int n, const struct GTOBasis_t *g, const int *basis2Atom,
                 const double *PT, const double *PS, double Kfactor, double *Gx, double *Gy, double *Gz)
{
	or_grad_eri_cut(childID, nCPU, n, g, basis2Atom, PT, PS, Kfactor, OR_GRAD_CUTOFF, Gx, Gy, Gz);
}

/* Hartree-Fock forces F = -dE/dR per atom (src/grad.c:2155-2420 without the
 * DFT branch): nuclear repulsion, overlap against the energy-weighted
 * density QT = sum_occ e_i C_i C_i^T, kinetic and attraction (basis-function
 * and operator derivatives) against PT, the field terms when (Ex, Ey, Ez) is
 * not zero, and or_grad_eri with exchange weight Kfactor.  CA/CB hold one
 * orbital per row (C[i*n + p]). */
void or_uhf_force(int n, const struct GTOBasis_t *g, const struct Molecule_t *mol, int nEA, int nEB,
                  const double *CA, const double *CB, const double *eA, const double *eB,
                  double Ex, double Ey, double Ez, double Kfactor, double *Fx, double *Fy, double *Fz)
{
	int na = mol->nAtom;
	int *b2a = malloc(sizeof(int) * n);
	for (int p = 0; p < n; p++)
		for (int i = 0; i < na; i++)
			if (g[p].x0 == mol->x[i] && g[p].y0 == mol->y[i] && g[p].z0 == mol->z[i]) { b2a[p] = i; break; }
	double *PA = calloc((size_t)n * n, sizeof(double)), *PB = calloc((size_t)n * n, sizeof(double));
	double *PT = malloc(sizeof(double) * n * n), *PS = malloc(sizeof(double) * n * n);
	double *QT = calloc((size_t)n * n, sizeof(double));
	for (int p = 0; p < n; p++)
		for (int q = 0; q < n; q++) {
			for (int i = 0; i < nEA; i++) {
				PA[p * n + q] += CA[i * n + p] * CA[i * n + q];
				QT[p * n + q] += eA[i] * CA[i * n + p] * CA[i * n + q];
			}
			for (int i = 0; i < nEB; i++) {
				PB[p * n + q] += CB[i * n + p] * CB[i * n + q];
				QT[p * n + q] += eB[i] * CB[i * n + p] * CB[i * n + q];
			}
			PT[p * n + q] = PA[p * n + q] + PB[p * n + q];
			PS[p * n + q] = PA[p * n + q] - PB[p * n + q];
		}
	double *f[3];
	for (int d = 0; d < 3; d++) f[d] = calloc(na, sizeof(double));
	or_grad_vnn(mol, f[0], f[1], f[2]);
	for (int p = 0; p < n; p++)
		for (int q = 0; q < n; q++)
			for (int d = 0; d < 3; d++) {
				f[d][b2a[p]] -= or_grad_overlap(p, q, g, d) * QT[p * n + q];
				f[d][b2a[q]] -= or_grad_overlap(q, p, g, d) * QT[p * n + q];
				f[d][b2a[p]] += or_grad_kinetic(p, q, g, d) * PT[p * n + q];
				f[d][b2a[q]] += or_grad_kinetic(q, p, g, d) * PT[p * n + q];
				f[d][b2a[p]] += 2.0 * or_grad_nai(p, q, g, mol, d) * PT[p * n + q];
			}
	for (int p = 0; p < n; p++)
		for (int q = 0; q <= p; q++)
			for (int i = 0; i < na; i++) {
				double e[3];
				or_gto_efi(p, q, g, mol->x[i], mol->y[i], mol->z[i], e, e + 1, e + 2);
				double w = (p != q ? 2.0 : 1.0) * mol->Z[i] * PT[p * n + q];
				for (int d = 0; d < 3; d++) f[d][i] += e[d] * w;
			}
	if (Ex != 0.0 || Ey != 0.0 || Ez != 0.0) {
		double E3[3] = {Ex, Ey, Ez};
		for (int p = 0; p < n; p++)
			for (int q = 0; q < n; q++)
				for (int c = 0; c < 3; c++)
					for (int d = 0; d < 3; d++)
						f[d][b2a[p]] += 2.0 * or_grad_moment(p, q, g, c == 0, c == 1, c == 2, 0.0, 0.0, 0.0, d) *
						                PT[p * n + q] * E3[c];
		for (int i = 0; i < na; i++)
			for (int d = 0; d < 3; d++) f[d][i] -= E3[d] * mol->Z[i];
	}
	or_grad_eri(0, 1, n, g, b2a, PT, PS, Kfactor, f[0], f[1], f[2]);
	for (int i = 0; i < na; i++) { Fx[i] = -f[0][i]; Fy[i] = -f[1][i]; Fz[i] = -f[2][i]; }
	for (int d = 0; d < 3; d++) free(f[d]);
	free(b2a); free(PA); free(PB); free(PT); free(PS); free(QT);
}

/* convenience for tests: free what or_gen_basis / or_read_xyz made */
void or_free_basis(int n, struct GTOBasis_t *g)
{
	for (int i = 0; i < n; i++) { free(g[i].exp); free(g[i].coef); free(g[i].norm); }
	free(g);
}

/* ------------------------------------------------------------------ */
/* MP2 (src/mp2.c:29-67, 581-712; src/matrix.c:4797-5073)             */
/* ------------------------------------------------------------------ */

/* frozen-core orbitals: half the noble-gas core electrons per atom
 * (src/mp2.c:29-67) */
int or_mp2_frozen_core(const struct Molecule_t *m)
{
	int core = 0;
	for (int a = 0; a < m->nAtom; a++) {
		const int z = m->Z[a];
		if (z > 54) return -1;
		core += z <= 2 ? 0 : z <= 10 ? 2 : z <= 18 ? 10 : z <= 36 ? 18 : 36;
	}
	return core / 2;
}

/* the dense pair matrix V[pair(p,q) m + pair(r,s)] = (pq|rs), m = n(n+1)/2,
 * pair(p,q) = p(p+1)/2 + q for p >= q: every contracted integral, no
 * screening (the layout the device integral store writes) */
void or_eri_pairs(int n, const struct GTOBasis_t *g, double *V)
{
	const long m = (long)n * (n + 1) / 2;
	for (int p = 0, P = 0; p < n; p++)
		for (int q = 0; q <= p; q++, P++)
			for (int r = 0, Q = 0; r < n && Q <= P; r++)
				for (int s = 0; s <= r && Q <= P; s++, Q++) {
					const double v = or_contr_eri(g, p, q, r, s);
					V[P * m + Q] = v;
					V[Q * m + P] = v;
				}
}

/* One ordered tuple (P,Q,I,J) of the integral EE into the (aq|ij) bank of
 * mp2_gen_aqij_Quartet (its aqij macro, src/matrix.c:4797-4802): pairs
 * I >= J that the bank keeps receive C[P, a] EE in row a, column Q. */
static void or_aqij_add(double *bank, const long *slot, int n, int nCorr, int nCore, const double *C, int P,
                        int Q, int I, int J, double EE)
{
	if (J > I || slot[I * n + J] < 0) return;
	double *b = bank + slot[I * n + J] * (long)nCorr * n;
	for (int a = 0; a < nCorr; a++) b[(long)a * n + Q] += C[(long)(nCore + a) * n + P] * EE;
}

/* The (aq|ij) bank of mp2_gen_aqij_Quartet for the correlated orbitals
 * [nCore, nCore + nCorr) of C (src/matrix.c:4864-5073, one cycle, no
 * multipole substitution): rows slot[i*n+j] * nCorr + a of n doubles, for
 * the kept pairs i >= j (basis Schwarz >= cut); canonical shell quartets with
 * shell Schwarz product >= cut, integrals |EE| > cut, each unique integral
 * added for the distinct orderings of its shell-quartet type. */
static double *or_mp2_bank(int n, int nCorr, int nCore, const double *C, const struct GTOBasis_t *g,
                           const double *sch, double cut, const long *slot, long nEE)
{
	double *bank = calloc((size_t)(nEE > 0 ? nEE : 1) * nCorr * n, sizeof(double));
	int ns = 0;
	or_shell *sh = or_make_shells(n, g, &ns);
	double *ssh = calloc((size_t)ns * ns, sizeof(double));
	for (int P = 0; P < ns; P++)
		for (int Q = 0; Q < ns; Q++) {
			double mx = 0.0;
			for (int p = sh[P].min; p <= sh[P].max; p++)
				for (int q = sh[Q].min; q <= sh[Q].max; q++)
					if (mx < sch[p * n + q]) mx = sch[p * n + q];
			ssh[P * ns + Q] = mx;
		}
	double *EE = malloc(sizeof(double) * OR_MAXB * OR_MAXB * OR_MAXB * OR_MAXB);
	for (int P = 0; P < ns; P++)
		for (int Q = 0; Q <= P; Q++)
			for (int I = 0; I <= P; I++)
				for (int J = 0; J <= (I == P ? Q : I); J++) {
					if (ssh[P * ns + Q] * ssh[I * ns + J] < cut) continue;
					or_quartet(sh + P, sh + Q, sh + I, sh + J, EE);
					/* orderings of the type: bit k of the mask takes the k-th of
					 * (pq|ij) (qp|ij) (pq|ji) (qp|ji) (ij|pq) (ji|pq) (ij|qp) (ji|qp) */
					int mask;
					if (P == Q && I == J && P == I) mask = 0x01;        /* one shell */
					else if (P == Q && I == J) mask = 0x11;             /* two diagonal pairs */
					else if (P == Q) mask = 0x35;                       /* bra diagonal */
					else if (I == J) mask = 0x53;                       /* ket diagonal */
					else if (P == I && Q == J) mask = 0x0F;             /* bra pair = ket pair */
					else mask = 0xFF;
					const double *v = EE;
					for (int p = sh[P].min; p <= sh[P].max; p++)
						for (int q = sh[Q].min; q <= sh[Q].max; q++)
							for (int i = sh[I].min; i <= sh[I].max; i++)
								for (int j = sh[J].min; j <= sh[J].max; j++, v++) {
									if (!(fabs(*v) > cut)) continue;
									const int t[8][4] = {{p, q, i, j}, {q, p, i, j}, {p, q, j, i}, {q, p, j, i},
									                     {i, j, p, q}, {j, i, p, q}, {i, j, q, p}, {j, i, q, p}};
									for (int k = 0; k < 8; k++)
										if (mask >> k & 1)
											or_aqij_add(bank, slot, n, nCorr, nCore, C, t[k][0], t[k][1], t[k][2],
											            t[k][3], *v);
								}
				}
	free(EE); free(sh); free(ssh);
	return bank;
}

/* (ar|bj) for one bank row a and occupied orbital cb (coefficients of b):
 * (aq|bj) = sum_i C_bi (aq|ij) over |Sch_ij C_bi| > cut, then
 * arbj[r*n + j] = sum_q Cv[r][q] (aq|bj) for the nVir virtual orbitals Cv
 * (src/mp2.c:653-671, 2240-2262) */
static void or_mp2_arbj(int n, const double *bank, int nCorr, int a, const long *slot, const double *cb,
                        const double *sch, double cut, const double *Cv, int nVir, double *aqbj, double *arbj)
{
	memset(aqbj, 0, sizeof(double) * n * n);
	for (int i = 0; i < n; i++)
		for (int j = 0; j < n; j++) {
			if (slot[i * n + j] < 0) continue;
			const double cbi = cb[i];
			if (!(fabs(sch[i * n + j] * cbi) > cut)) continue;
			const double *src = bank + (slot[i * n + j] * nCorr + a) * (long)n;
			for (int q = 0; q < n; q++) aqbj[j * n + q] += cbi * src[q];
		}
	for (int r = 0; r < nVir; r++)
		for (int j = 0; j < n; j++) {
			double s = 0.0;
			for (int q = 0; q < n; q++) s += Cv[(long)r * n + q] * aqbj[j * n + q];
			arbj[r * n + j] = s;
		}
}

static long *or_mp2_slots(int n, const double *sch, double cut, long *nEE)
{
	long *slot = malloc(sizeof(long) * n * n), k = 0;
	for (int i = 0; i < n; i++)
		for (int j = 0; j <= i; j++) {
			slot[i * n + j] = sch[i * n + j] < cut ? -1 : k++;
			slot[j * n + i] = slot[i * n + j];
		}
	*nEE = k;
	return slot;
}

/* Closed-shell MP2 energy by the reference's aqij algorithm for all
 * correlated orbitals in one cycle (src/mp2.c:581-712): per a >= b the
 * (ar|bs), (as|br) and the r >= s sum of
 *   2 [(ar|bs)^2 + (as|br)^2 - (ar|bs)(as|br)] / (e_a + e_b - e_r - e_s),
 * halved for r = s, doubled for a != b.  cut = 0 drops nothing: the exact
 * MP2 energy.  sch: basis Schwarz (n x n). */
double or_mp2_rhf(int n, int nOcc, int nCore, const double *e, const double *C, const struct GTOBasis_t *g,
                  const double *sch, double cut)
{
	const int nCorr = nOcc - nCore, nVir = n - nOcc;
	if (nCorr <= 0 || nVir <= 0) return 0.0;
	long nEE;
	long *slot = or_mp2_slots(n, sch, cut, &nEE);
	double *bank = or_mp2_bank(n, nCorr, nCore, C, g, sch, cut, slot, nEE);
	double *aqbj = malloc(sizeof(double) * n * n), *arbj = malloc(sizeof(double) * nVir * n);
	const double *eCor = e + nCore, *eVir = e + nOcc, *Cv = C + (long)nOcc * n;
	double sum = 0.0;
	for (int a = 0; a < nCorr; a++)
		for (int b = 0; b <= a; b++) {
			or_mp2_arbj(n, bank, nCorr, a, slot, C + (long)(b + nCore) * n, sch, cut, Cv, nVir, aqbj, arbj);
			for (int r = 0; r < nVir; r++)
				for (int s = 0; s <= r; s++) {
					double arbs = 0.0, asbr = 0.0;
					for (int j = 0; j < n; j++) {
						arbs += Cv[(long)s * n + j] * arbj[r * n + j];
						asbr += Cv[(long)r * n + j] * arbj[s * n + j];
					}
					double c = (arbs * arbs + asbr * asbr - arbs * asbr) * 2.0 / (eCor[a] + eCor[b] - eVir[r] - eVir[s]);
					if (r == s) c *= 0.5;
					if (a != b) c *= 2.0;
					sum += c;
				}
		}
	free(aqbj); free(arbj); free(bank); free(slot);
	return sum;
}

/* Unrestricted MP2 energy by the reference's algorithm (src/mp2.c:2126-2432,
 * each spin's bank in one cycle, no multipole substitution): for each spin
 * the a < b, r < s sum of [(ar|bs) - (as|br)]^2 / D over its own orbitals,
 * and the alpha-beta sum of (ar|bs)^2 / D with a, r alpha and b, s beta. */
double or_mp2_uhf(int n, int nEA, int nEB, int nCore, const double *eA, const double *eB, const double *CA,
                  const double *CB, const struct GTOBasis_t *g, const double *sch, double cut)
{
	const int nOcc[2] = {nEA, nEB};
	const double *e[2] = {eA, eB}, *C[2] = {CA, CB};
	long nEE;
	long *slot = or_mp2_slots(n, sch, cut, &nEE);
	double *aqbj = malloc(sizeof(double) * n * n), *arbj = malloc(sizeof(double) * n * n);
	double sum = 0.0;
	for (int sp = 0; sp < 2; sp++) {
		const int nCorr = nOcc[sp] - nCore, nVir = n - nOcc[sp];
		if (nCorr <= 0 || nVir <= 0) continue;
		double *bank = or_mp2_bank(n, nCorr, nCore, C[sp], g, sch, cut, slot, nEE);
		const double *eCor = e[sp] + nCore, *eVir = e[sp] + nOcc[sp], *Cv = C[sp] + (long)nOcc[sp] * n;
		for (int a = 0; a < nCorr; a++) {
			/* same spin, b > a, r < s */
			for (int b = a + 1; b < nCorr; b++) {
				or_mp2_arbj(n, bank, nCorr, a, slot, C[sp] + (long)(b + nCore) * n, sch, cut, Cv, nVir, aqbj, arbj);
				for (int r = 0; r < nVir; r++)
					for (int s = r + 1; s < nVir; s++) {
						double arbs = 0.0, asbr = 0.0;
						for (int j = 0; j < n; j++) {
							arbs += Cv[(long)s * n + j] * arbj[r * n + j];
							asbr += Cv[(long)r * n + j] * arbj[s * n + j];
						}
						const double d = arbs - asbr;
						sum += d * d / (eCor[a] + eCor[b] - eVir[r] - eVir[s]);
					}
			}
			/* opposite spin, from the alpha bank only */
			if (sp == 0) {
				const int nCorrB = nEB - nCore, nVirB = n - nEB;
				const double *eCorB = eB + nCore, *eVirB = eB + nEB, *CvB = CB + (long)nEB * n;
				for (int b = 0; b < nCorrB; b++) {
					or_mp2_arbj(n, bank, nCorr, a, slot, CB + (long)(b + nCore) * n, sch, cut, Cv, nVir, aqbj, arbj);
					for (int r = 0; r < nVir; r++)
						for (int s = 0; s < nVirB; s++) {
							double arbs = 0.0;
							for (int j = 0; j < n; j++) arbs += CvB[(long)s * n + j] * arbj[r * n + j];
							sum += arbs * arbs / (eCor[a] + eCorB[b] - eVir[r] - eVirB[s]);
						}
				}
			}
		}
		free(bank);
	}
	free(aqbj); free(arbj); free(slot);
	return sum;
}
