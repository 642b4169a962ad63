/*
 * sq_serve.h - serving SCF solves on one HIP device.
 *
 * A request is a molecule with a run of geometry steps (as the reference's
 * geometry optimisation and molecular dynamics loops produce them,
 * src/optimize.c:283-340, src/qmd.c:564-667), its charge and multiplicity
 * (RHF when 1, UHF otherwise) and an arrival time.  sqsrv_run serves every
 * submitted request: it admits requests in arrival order while one of
 * `slots` is free and the summed device footprint of the requests in flight
 * (sqjk_footprint_bytes with the SCF) stays within memBudget, and runs the
 * admitted ones side by side, each on its own stream with its own
 * device-resident SCF state.  The first step of a request starts from the
 * DIAG guess, later steps from the previous step's orbitals (CACB), and the
 * J/K context follows the geometry.  Every step is one SCF with DIIS4 and
 * the given convergence settings.  No counterpart in the reference, which
 * runs one molecule per process.
 */
#ifndef SQ_SERVE_H
#define SQ_SERVE_H

#include "sq_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sqsrv sqsrv;

typedef struct {
	int device;
	int slots;              /* requests in flight at once (1..64) */
	double memBudget;       /* bytes of device memory for the requests in flight */
	double mixedThreshold;  /* FP32 integrals below this Schwarz product (0: FP64) */
	double SCFConv;
	int SCFMax;
	int threeStep;          /* 3-step cut-offs 1e-3/1e-6/1e-9 */
	double SCFDrag;
} sqsrv_config;

typedef struct {
	int nstep, converged_steps, nBasis, iterations;
	double arrival, start, finish;  /* seconds from the start of sqsrv_run */
	double footprint;               /* bytes admitted for the request */
	int peak_in_flight;             /* of the whole run */
	double peak_memory;             /* of the whole run, bytes */
} sqsrv_stats;

/* db: basis set records as read by read_GAMESS_BasisSet (kept by the caller
 * for the life of the server). */
int sqsrv_create(const sqsrv_config *cfg, int dbSize, const struct GTOBasisSet_t *db, sqsrv **out);
/* xyz: nStep geometries of nAtom atoms, bohr, [step][atom][x y z].  Returns
 * the request id (>= 0) or an error code (< 0). */
int sqsrv_submit(sqsrv *s, int nAtom, const int *Z, int nStep, const double *xyz, int charge, int multiplicity,
                 double arrival);
/* Serves every submitted request; returns when all are done. */
int sqsrv_run(sqsrv *s);
/* energies[nstep] (0.0 where a step did not converge), iterations[nstep] */
int sqsrv_result(const sqsrv *s, int id, double *energies, int *iterations, sqsrv_stats *st);
void sqsrv_destroy(sqsrv *s);
const char *sqsrv_errmsg(void);

#ifdef __cplusplus
}
#endif
#endif /* SQ_SERVE_H */
