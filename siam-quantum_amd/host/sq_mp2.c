/*
 * sq_mp2.c - drop-ins for src/mp2.h: mp2_rhf_aqij_Parallel, mp2_rhf_aqij,
 * mp2_uhf_aqij, getFrozenCore_RHF.  The integrals, the four-index transform and the energy
 * sum run on the device (sqjk_mp2_rhf); this file counts the frozen core and
 * prints the reference's report around it (src/mp2.c:713-927, 2126-2432).
 */
#include <stdio.h>
#include <stdlib.h>

#include "sq_mp2.h"

#define MP2_DROPIN_CUTOFF 1.0e-10 /* Schwarz product below which shell quartets are left out */
#define HARTREE_TO_EV 27.2113957  /* src/util.h:23 */

/* core electrons per element up to Xe: the noble-gas core below each
 * element's valence shell (He, Ne, Ar, Kr cores), as src/mp2.c:29-35 */
static int noble_gas_core(int z)
{
	if (z <= 2) return 0;
	if (z <= 10) return 2;
	if (z <= 18) return 10;
	if (z <= 36) return 18;
	return 36;
}

int getFrozenCore_RHF(const struct Molecule_t *mol)
{
	int core = 0;
	for (int a = 0; a < mol->nAtom; a++) {
		if (mol->Z[a] > 54) {
			printf("getFrozenCore_RHF - error no available noble gas atom info.\n");
			exit(-1);
		}
		core += noble_gas_core(mol->Z[a]);
	}
	return core / 2;
}

double mp2_rhf_aqij_Parallel(int nBasis, int nOcc, const double *e, const double *C,
                             const struct GTOBasis_t *gto, const struct Molecule_t *mol,
                             const struct option_t *opt)
{
	if (opt->restricted == 0) {
		printf("mp2_rhf_aqij - error only restricted MP2 is supported\n");
		exit(-1);
	}
	const int nCore = getFrozenCore_RHF(mol);
	printf("                                                             \n"
	       "                                                             \n"
	       "-------------------------------------------------------------\n"
	       "-----               MP2 ENERGY CORRECTION               -----\n"
	       "-------------------------------------------------------------\n");
	printf("The number of excluded core orbitals:          %4d\n", nCore);
	printf("The number of included occupied orbitals:      %4d\n", nOcc - nCore);
	printf("The number of included virtual  orbitals:      %4d\n", nBasis - nOcc);
	fflush(stdout);
	if (nCore > nOcc) {
		printf("mp2_rhf_aqij_Parallel - error more frozen-core orbitals (%d) than occupied orbitals (%d)\n", nCore,
		       nOcc);
		exit(-1);
	}
	sqjk_ctx *ctx = NULL;
	int rc = sqjk_create(nBasis, gto, NULL, 0, &ctx);
	sqjk_mp2_info info;
	if (rc == SQJK_OK) rc = sqjk_mp2_rhf(ctx, nOcc, nCore, e, C, MP2_DROPIN_CUTOFF, &info);
	if (ctx) sqjk_destroy(ctx);
	if (rc != SQJK_OK) {
		printf("mp2_rhf_aqij_Parallel - error %d: %s\n", rc, sqjk_errmsg());
		exit(-1);
	}
	printf("Device memory for (pq|ij) and (ar|ij):       %8.1f MB\n", info.device_bytes * 1.0e-6);
	printf("\n"
	       "MP2 Energy Correction is %12.6f Hartree\n"
	       "MP2 Energy Correction is %12.6f eV\n",
	       info.e2, HARTREE_TO_EV * info.e2);
	printf("-------------------------------------------------------------\n");
	fflush(stdout);
	return info.e2;
}

double mp2_rhf_aqij(int nBasis, int nOcc, const double *e, const double *C, const struct GTOBasis_t *gto,
                    const struct Molecule_t *mol, const struct option_t *opt)
{
	return mp2_rhf_aqij_Parallel(nBasis, nOcc, e, C, gto, mol, opt);
}

double mp2_uhf_aqij(int nBasis, int nEA, int nEB, const double *eA, const double *eB, const double *CA,
                    const double *CB, const struct GTOBasis_t *gto, const struct Molecule_t *mol,
                    const struct option_t *opt)
{
	(void)opt;
	const int nCore = getFrozenCore_RHF(mol);
	printf("                                                             \n"
	       "                                                             \n"
	       "-------------------------------------------------------------\n"
	       "-----        UNRESTRICTED MP2 ENERGY CORRECTION         -----\n"
	       "-------------------------------------------------------------\n");
	printf("The number of excluded core orbitals:      %4d %4d\n", nCore, nCore);
	printf("The number of included occupied orbitals:  %4d %4d\n", nEA - nCore, nEB - nCore);
	printf("The number of included virtual  orbitals:  %4d %4d\n", nBasis - nEA, nBasis - nEB);
	fflush(stdout);
	if (nCore > nEA || nCore > nEB) {
		printf("mp2_uhf_aqij - error more frozen-core orbitals (%d) than occupied orbitals (%d %d)\n", nCore, nEA,
		       nEB);
		exit(-1);
	}
	sqjk_ctx *ctx = NULL;
	int rc = sqjk_create(nBasis, gto, NULL, 0, &ctx);
	sqjk_mp2_info info;
	if (rc == SQJK_OK) rc = sqjk_mp2_uhf(ctx, nEA, nEB, nCore, eA, eB, CA, CB, MP2_DROPIN_CUTOFF, &info);
	if (ctx) sqjk_destroy(ctx);
	if (rc != SQJK_OK) {
		printf("mp2_uhf_aqij - error %d: %s\n", rc, sqjk_errmsg());
		exit(-1);
	}
	printf("Device memory for (pq|ij) and (ar|ij):       %8.1f MB\n", info.device_bytes * 1.0e-6);
	printf("\n"
	       "MP2 Energy Correction is %12.6f Hartree\n"
	       "                         %12.6f eV\n",
	       info.e2, HARTREE_TO_EV * info.e2);
	printf("-------------------------------------------------------------\n");
	fflush(stdout);
	return info.e2;
}
