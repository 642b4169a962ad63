/*
 * sq_jk.h - C-ABI of the MI355X two-electron engine.
 *
 * Two layers:
 *  1. Drop-in entry points with the exact signatures of the reference's J/K
 *     surface, so a caller of the reference links against this library
 *     unchanged:
 *       GTO_JK_Matrix_Quartet_Parallel  <- src/matrix.h:155-166 (src/matrix.c:3773-3986)
 *       uhf_getGMatrix                  <- src/uhf.h:26-33      (src/uhf.c:112-202)
 *       create_Schwarz                  <- src/matrix.h:69-71   (src/matrix.c:389-440)
 *  2. A re-entrant context API (sqjk_*) that the drop-ins are built on.  It
 *     holds per-molecule state explicitly instead of in statics, so several
 *     molecules can be served side by side.
 *
 * All matrices are row-major n x n double arrays in host memory.  Densities
 * are assumed symmetric (as in the reference).  Every call runs the integral
 * and digestion work on the HIP device; there is no CPU fallback: when no
 * device is usable the drop-ins print an error and exit(-1), matching the
 * reference's error behaviour, and the sqjk_* calls return SQJK_ENODEV.
 */
#ifndef SQ_JK_H
#define SQ_JK_H

#include "sq_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- return codes ---- */
#define SQJK_OK        0
#define SQJK_ENODEV   -1  /* no usable HIP device                        */
#define SQJK_EHIP     -2  /* a HIP runtime call failed                    */
#define SQJK_EBASIS   -3  /* basis not representable (see sqjk_errmsg)   */
#define SQJK_EARG     -4  /* invalid argument                             */
#define SQJK_ENOMEM   -5  /* device allocation failed                     */

typedef struct sqjk_ctx sqjk_ctx;

/* Per-build statistics, filled by sqjk_build / sqjk_last_stats. */
typedef struct {
	double kernel_ms;      /* integral+digestion kernels, device time      */
	double total_ms;       /* whole build incl. uploads/downloads (host)   */
	long long quartets;    /* contracted shell quartets launched           */
	long long quartets_fp32; /* of which evaluated in FP32 (mixed precision)  */
	int launches;          /* kernel launches in this build                */
} sqjk_stats;

/* Create a context for one basis.  schwarz_basis: optional n x n matrix of
 * sqrt|(pq|pq)| (as from create_Schwarz); NULL computes it on the device.
 * device: HIP device ordinal.  Returns SQJK_OK and *out on success. */
int sqjk_create(int nBasis, const struct GTOBasis_t *gto, const double *schwarz_basis,
                int device, sqjk_ctx **out);

/* Device memory a J/K context for this basis takes (pair and primitive
 * data, matrices, exchange staging for RHF or UHF), plus that of a
 * device-resident SCF when with_scf is set; an estimate from the shell and
 * pair counts, computed on the host without touching the device.  Used to
 * admit served requests by HBM footprint.  Negative on error. */
long long sqjk_footprint_bytes(int nBasis, const struct GTOBasis_t *gto, int uhf, int with_scf);

/* Moves the context to a new geometry of the same basis (same functions,
 * exponents and coefficients; only the centres differ), as in the
 * geometry steps of an optimisation or molecular dynamics run
 * (src/qmd.c:564-667, src/optimize.c:283-340): primitive pair data and the
 * Schwarz bounds are recomputed on the device, while the shells, pair
 * lists, device buffers and stream of the context are kept.  Returns
 * SQJK_EARG if the basis (nBasis functions) differs in anything but its
 * centres. */
int sqjk_update_geometry(sqjk_ctx *ctx, int nBasis, const struct GTOBasis_t *gto);

/* One J/K build on density PA (alpha) and PB (beta):
 *   JT += J[PA+PB],  KA -= K[PA],  KB -= K[PB]   (accumulating, as the
 * reference does), then KB := KA when restricted, then upper triangle :=
 * lower triangle of JT, KA, KB.
 * Screening follows the reference: shell quartets with
 * Sch_PQ*Sch_IJ < max(scfCutoff, fixedCutoff), or with
 * Sch_PQ*Sch_IJ*(sum of six block-max |P|) < fixedCutoff, are skipped, and
 * primitive quartets with a bare prefactor below 1e-15 are dropped. */
int sqjk_build(sqjk_ctx *ctx, const double *PA, const double *PB, double fixedCutoff,
               double scfCutoff, int restricted, double *JT, double *KA, double *KB);

/* Device-resident variant of sqjk_build for callers that keep their matrices
 * in HIP device memory (an SCF that stays on the device, batched and served
 * solves): dPA, dPB are n x n device densities; dJ, dKA, dKB (device) are
 * OVERWRITTEN with the full symmetric J[PA+PB], -K[PA], -K[PB] (dKB := dKA
 * when restricted), no accumulation.  Same screening as sqjk_build.  The
 * work is queued on the context's stream and the call returns without
 * waiting; sqjk_sync waits for it.  The inputs must stay valid and the
 * outputs untouched until then.  No counterpart in the reference. */
int sqjk_build_device(sqjk_ctx *ctx, const double *dPA, const double *dPB, double fixedCutoff,
                      double scfCutoff, int restricted, double *dJ, double *dKA, double *dKB);
/* Allocation-free variant of sqjk_build_device, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same inputs, outputs, screening
 * and results.  The work plan and the digestion buffers live in a scratch
 * array of SQJK_NOALLOC_SCRATCH_BYTES that the library holds on each GPU it
 * runs on for the life of the process; the context's own buffers (made by
 * sqjk_create) hold the densities and accumulators.  The call makes no
 * device or pinned-host allocation (it rebuilds the context's host work
 * plan when the cut-offs change).  A build that needs more scratch than
 * that returns SQJK_ENOMEM with the size in sqjk_errmsg.  Every call on one
 * GPU shares that GPU's scratch, so calls that overlap on one GPU (several
 * contexts' streams, several host threads) race: finish one (sqjk_sync)
 * before the next starts.  Per-class profiling is off for these builds.
 * No counterpart in the reference; it exists for the callers above. */
#define SQJK_NOALLOC_SCRATCH_BYTES (64ull << 20)
int sqjk_build_device_noalloc(sqjk_ctx *ctx, const double *dPA, const double *dPB, double fixedCutoff,
                              double scfCutoff, int restricted, double *dJ, double *dKA, double *dKB);

/* Waits for the context's queued work (and reads its timing). */
int sqjk_sync(sqjk_ctx *ctx);

/* Mixed precision: shell quartets whose Schwarz product Sch_PQ*Sch_IJ is
 * below `threshold` are evaluated in FP32 (Boys, Hermite and contraction)
 * in the angular classes of at most 81 integrals per quartet; larger ones
 * and all other classes stay FP64, and all digestion and accumulation is
 * FP64.
 * threshold 0 (the default) keeps every quartet in FP64.  No counterpart in
 * the reference (north-star mixed-precision mode). */
int sqjk_set_mixed(sqjk_ctx *ctx, double threshold);

/* Schwarz matrix sqrt|(pq|pq)| (n x n) as held by the context. */
int sqjk_schwarz(sqjk_ctx *ctx, double *out);

/* Statistics of the last build. */
int sqjk_last_stats(const sqjk_ctx *ctx, sqjk_stats *st);

/* Work of the last build per angular class (21 classes for l <= 2):
 * lq[4k..4k+3] = (la,lb,lc,ld), quartets[k] = contracted shell quartets
 * launched (after Schwarz screening, before the density screen and the
 * primitive cut-off), prim_quartets[k] = their primitive quartets.  Any
 * output pointer may be NULL.  Returns the number of classes. */
int sqjk_class_stats(const sqjk_ctx *ctx, int *lq, long long *quartets, long long *prim_quartets);

/* Work of the last build that survives all screening, per angular class
 * (same class order as sqjk_class_stats): contracted shell quartets that pass
 * the density-weighted test, and their primitive quartets whose bare
 * prefactor passes the 1e-15 cut, split by the precision they were
 * evaluated in.  Counted on the device with the last build's density and
 * cut-offs (one extra pass over the work plan; no integrals).  Any output
 * pointer may be NULL.  Returns the number of classes, or an error code.
 * No counterpart in the reference (measurement support). */
int sqjk_class_work(sqjk_ctx *ctx, long long *quartets, long long *prim_quartets,
                    long long *quartets_fp32, long long *prim_quartets_fp32);

/* Work-plan tuning per class pair (arrays of 6 x 6 ints indexed
 * [bra pair class][ket pair class], pair classes ss, ps, pp, ds, dp, dd; any
 * may be NULL to keep the current values; all NULL restores the defaults):
 *   rows          bra rows per tile (1..64),
 *   split_budget  integral-primitive products per lane and tile above which
 *                 a segment's ket primitive pairs are split over tiles,
 *   grid_over     workgroups per launch as a multiple of those the device
 *                 holds at once (0: the kernel's default).
 * Results do not depend on them; the autotuner (sqamd/tune.py) picks them.
 * No counterpart in the reference. */
int sqjk_set_tuning(sqjk_ctx *ctx, const int *rows, const int *split_budget, const int *grid_over);
int sqjk_get_tuning(const sqjk_ctx *ctx, int *rows, int *split_budget, int *grid_over);

/* ---- batches: many molecules' J/K builds as one set of launches ----
 * A batch joins existing contexts (which must outlive it; their geometry
 * must not change while it exists).  Per (class pair, precision, spin mode)
 * one launch covers the quartets of every molecule; the whole build is
 * captured into a HIP graph on its first graph build and replayed while
 * the cut-offs stay the same.  Densities and results live in the batch's
 * device buffers.  restricted[i] selects RHF (K_B := K_A) or UHF per
 * molecule.  No counterpart in the reference, whose J/K state is static. */
typedef struct sqjk_batch sqjk_batch;
int sqjk_batch_create(int count, sqjk_ctx *const *ctxs, const int *restricted, sqjk_batch **out);
/* density of molecule i (n_i x n_i; host or device memory) */
int sqjk_batch_set_density(sqjk_batch *b, int i, const double *PA, const double *PB, int on_device);
/* cut-offs of molecule i (i = -1: every molecule), as in sqjk_build */
int sqjk_batch_set_cutoffs(sqjk_batch *b, int i, double fixedCutoff, double scfCutoff);
int sqjk_batch_set_mixed(sqjk_batch *b, double threshold);
int sqjk_batch_set_tuning(sqjk_batch *b, const int *rows, const int *split_budget, const int *grid_over);
/* per-class event timing of eager builds */
int sqjk_batch_set_profile(sqjk_batch *b, int on);
/* One build of every molecule: J = J[PA+PB], KA = -K[PA], KB = -K[PB]
 * (full symmetric), overwritten.  use_graph: replay the captured graph
 * (captured after an eager build on the first call of a plan). */
int sqjk_batch_build(sqjk_batch *b, int use_graph);
int sqjk_batch_result(sqjk_batch *b, int i, double *J, double *KA, double *KB, int on_device);
/* device time of the last build, its screened quartets and class launches,
 * and (profile mode, eager) per-class ms in sqjk_class_stats order */
int sqjk_batch_stats(const sqjk_batch *b, double *kernel_ms, long long *quartets, int *launches,
                     double *class_ms);
void sqjk_batch_destroy(sqjk_batch *b);

/* Per-class kernel timing.  With `on` set, every later sqjk_build records a
 * pair of HIP events around each class launch group on the context's stream
 * (all ket-component slices, FP64 and FP32 launches of one class) and
 * sqjk_class_times returns their durations in ms, in the class order of
 * sqjk_class_stats.  Off by default; no counterpart in the reference
 * (measurement support). */
int sqjk_set_profile(sqjk_ctx *ctx, int on);
int sqjk_class_times(const sqjk_ctx *ctx, double *kernel_ms);

/* One-electron matrices on the device: kinetic T, nuclear attraction V and
 * overlap S (n x n, row-major; any output may be NULL).  Replaces the three
 * element loops over GTO_kinetic, GTO_nuclei and GTO_overlap of
 * src/uhf.c:1307-1334 (same integrals and the same 1e-15 primitive cut of
 * src/int.c:392-393); the electric-field term is left to the caller. */
int sqjk_one_electron(int nBasis, const struct GTOBasis_t *gto, const struct Molecule_t *mol, int device,
                      double *T, double *V, double *S);

/* Boys function F_0..F_L (L <= 9) at nT points T >= 0, evaluated on the
 * device by the integral kernels' own code (LDS table with Taylor expansion
 * and downward recursion below the per-L cut-off, sqrt(pi/T)/2 with upward
 * recursion above it), in FP32 when fp32 is set.  F[i*(L+1) + m] = F_m(T_i).
 * Same quantity as the reference's fgamma_set (src/fgamma.c:172-206); this
 * entry exists so that the device evaluation can be checked on its own. */
int sqjk_boys(int L, int nT, const double *T, int fp32, int device, double *F);

/* Number of basis functions, reference shells and pure-l sub-shells. */
int sqjk_dims(const sqjk_ctx *ctx, int *nBasis, int *nShell, int *nSubShell);

void sqjk_destroy(sqjk_ctx *ctx);

/* Text of the last error on this thread. */
const char *sqjk_errmsg(void);

/* ---- drop-in surface (reference signatures) ---- */

/* src/matrix.h:155-166.  Same accumulate / sign / restricted / symmetrise
 * semantics; the per-molecule state lives in one library-held context that
 * is built on the first call and released by a call with nBasis == 0, as in
 * the reference (src/matrix.c:3817-3828).  With opt->nCPU > 1 all work is
 * done by childID 0 and other childIDs add nothing, so the caller's sum over
 * children is unchanged.  Reads opt->SCFCutoff, opt->restricted, opt->nCPU. */
void GTO_JK_Matrix_Quartet_Parallel(int childID, int nBasis, const double *PA, const double *PB,
                                    const struct GTOBasis_t *gto, const double *schwarz_basis,
                                    double fixedCutoff, double *JT, double *KA, double *KB,
                                    struct option_t *opt);

/* Mixed-precision threshold used by the drop-in entry points (and so by
 * uhf()); 0 (default) is the plain FP64 path.  No counterpart in the
 * reference. */
void sqjk_set_dropin_mixed(double threshold);
double sqjk_get_dropin_mixed(void);

/* src/uhf.h:26-33: zero JT, KA, KB, then one J/K build. */
void uhf_getGMatrix(int nBasis, const struct GTOBasis_t *gto, const double *Schwarz,
                    double cutoff, const double *PA, const double *PB,
                    double *JT, double *KA, double *KB, struct option_t *opt);

/* src/matrix.h:69-71: malloc'd n x n Schwarz matrix, computed on the device. */
double *create_Schwarz(int nBasis, const struct GTOBasis_t *gto);

#ifdef __cplusplus
}
#endif
#endif /* SQ_JK_H */
