/*
 * sq_types.h - data structures shared by the drop-in C surface.
 *
 * These three structs cross the C-ABI boundary unchanged, so their field
 * order, types and sizes are fixed by the reference's own headers:
 *   struct GTOBasis_t  <- src/basis.h:21-32
 *   struct Molecule_t  <- src/mol.h:19-26
 *   struct option_t    <- src/option.h:19-180
 * A caller compiled against the reference headers can pass its pointers to
 * this library as-is.  struct option_t has two layouts in the reference
 * (with and without -DLIBXC, src/option.h:46-51,118-123); define
 * SQ_OPTION_LIBXC_LAYOUT before including this header to match a caller
 * that was built with -DLIBXC.
 */
#ifndef SQ_TYPES_H
#define SQ_TYPES_H

#ifdef __cplusplus
extern "C" {
#endif

/* One contracted Cartesian Gaussian basis function (src/basis.h:21-32).
 * Value: sum_k coef[k]*norm[k]*x^l y^m z^n exp(-exp[k] r^2), centred at
 * (x0,y0,z0) in bohr.  norm[k] is the primitive normaliser of this very
 * Cartesian component; the contraction itself is not renormalised. */
struct GTOBasis_t {
	int     nContract;   /* number of primitives                       */
	int     l;           /* power of x                                 */
	int     m;           /* power of y                                 */
	int     n;           /* power of z                                 */
	double  x0;          /* centre, bohr                               */
	double  y0;
	double  z0;
	double *coef;        /* [nContract] contraction coefficients       */
	double *exp;         /* [nContract] exponents                      */
	double *norm;        /* [nContract] primitive normalisation        */
};

/* Nuclear framework (src/mol.h:19-26); coordinates in bohr. */
struct Molecule_t {
	int     nAtom;
	int     Q;           /* total molecular charge                     */
	int    *Z;
	double *x;
	double *y;
	double *z;
};

/* GAMESS-format basis record for one element (src/basis.h:40-49). */
#define SQ_BASIS_NAME 256
struct GTOBasisSet_t {
	char    name[SQ_BASIS_NAME];
	int     Z;
	int     nContract;
	int     l, m, n;
	double *exp;
	double *coef;
};

/* Run options (src/option.h:19-180).  Only SCFConv, SCFDrag, SCFCutoff,
 * SCFMax, convMethod, SCFAccuracy, SCFGuess, maxMem, restricted, nCPU,
 * molCharge and multiplicity are read by this library; every other field
 * is carried only so that the layout matches. */
struct option_t {
	int    molCharge;
	int    multiplicity;
	int    outVolumeType;
#ifdef SQ_OPTION_LIBXC_LAYOUT
	int    volumeFunc1Id;
	int    volumeFunc2Id;
	double volume_hyb_hf_coef;
#endif
	int    outDensityType;
	int    outWhichMO;
	int    outVolume;
	int    outVolumeGrid;
	double outVolumeCut;
	int    outVolumeFormat;
	int    outGAUSSIAN;
	int    outGAMESS;
	int    outLine;
	double lineDEF[6];
	double lineSPAN[2];
	int    lineNP;
	double SCFConv;
	double SCFDrag;
	double SCFCutoff;
	int    SCFMax;
	int    convMethod;
	int    SCFAccuracy;
	int    SCFGuess;
	int    maxMem;
	int    loadDMatrix;
	int    saveDMatrix;
	int    saveCheck;
	int    saveCheckAll;
	int    loadCheck;
	int    opt;
	int    optMax;
	int    restricted;
	int    method;
	int    whichHF2xDFT;
#ifdef SQ_OPTION_LIBXC_LAYOUT
	int    func1Id;
	int    func2Id;
	double hyb_hf_coef;
#endif
	int    whichGrid;
	int    nradial;
	int    nlabedev;
	int    nradialInit;
	int    nlabedevInit;
	int    force;
	int    hessian;
	char   DMatrixFile[256];
	char   CheckFile[256];
	int    nCPU;
	char   prefixStr[256];
	int    MECP;
	int    mecpMax;
	int    mecpMA;
	int    mecpMB;
	char   DMatrixFileA[256];
	char   DMatrixFileB[256];
	char   CheckFileA[256];
	char   CheckFileB[256];
	char   gaussEXE[256];
	char   gaussINA[256];
	char   gaussINB[256];
	int    excite;
	double Ex, Ey, Ez;
	int    qmd;
	double initTemp;
	double keepTemp;
	char   initVel[256];
	char   traj[256];
	double qmdFreq;
	double tStep;
	int    maxQMDStep;
};

/* Option values used on this path (src/option.h:67-90). */
#define CONVMETHOD_DAMPING   0
#define CONVMETHOD_DIIS2     1
#define CONVMETHOD_DIIS3     2
#define CONVMETHOD_DIIS4     3
#define SCFACCURACY_3STEP    0
#define SCFACCURACY_1STEP    1
#define SCFACCURACY_3STEP_A  1.0E-3
#define SCFACCURACY_3STEP_B  1.0E-6
#define SCFACCURACY_3STEP_C  1.0E-9
#define SCFGUESS_DIAG        0
#define SCFGUESS_CORE        1
#define SCFGUESS_CHECK       2
#define SCFGUESS_CACB        3
#define METHOD_HF            0

/* Physical constants (src/util.h:20-21). */
#define SQ_ANGSTROM2BOHR 1.889725988579
#define SQ_BOHR2ANGSTROM 0.529177249000

#ifdef __cplusplus
}
#endif
#endif /* SQ_TYPES_H */
