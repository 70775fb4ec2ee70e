/*
 * alis.h -- public C API of the MI355X FDTD engine.
 *
 * Source-compatible with the reference API (src/alis.h:7-8 -> src/alis_c.h:1-334 and
 * src/alis_m.h:18-66): the same type names, struct members (in the same order, since
 * user programs fill dom/res/sur/matter/object with positional initialisers), macros,
 * function names and function-pointer identities (Ex, Ppol, png, Sine ...), so the
 * reference's example programs compile against this header unchanged.  Struct world and
 * struct phaser carry extra private members after the reference ones; user code never
 * names them.
 *
 * Library sources define ALIS_LIBRARY before including this header so that the material
 * database and colour maps (ordinary definitions, instantiated in the user translation
 * unit exactly as with the reference) are not emitted into the library.
 */
#ifndef ALIS_H
#define ALIS_H

#include <stdio.h>
#include <stdlib.h>
#include <stdarg.h>
#include <math.h>
#include <values.h>
#include <ieee754.h>
#ifdef __cplusplus
/* C++ translation units (the HIP engine) only need the layout: float _Complex is the
 * compiler extension with the same representation as C99 float complex. */
#define ALIS_CPLX(t) t _Complex
#else
#include <complex.h>
#define ALIS_CPLX(t) t complex
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Boundary codes (src/alis_c.h:15-21). */
#define NIL 7
#define BBC 6
#define HBC 5
#define PBC 4
#define PML 3
#define PEC 2
#define SYM 1
#define MAXPOLES 6

#define INF FLT_MAX
#define PI 3.14159265358979323846f

#define sq(x) ((x)*(x))
#define n(n) ((matter){(n)*(n)})
#define nk(n,k,f) ((matter){{(n)*(n)-(k)*(k)},{0,2*n*k*((float)f)}})
#define png(cmap,max) png,cmap,(double)max
#define objects (object[])
#define MIN(x,y) ((x)<(y)?(x):(y))
#define MAX(x,y) ((x)>(y)?(x):(y))
/* calloc with a shifted base so that p[m..m+n-1] is valid; aborts on exhaustion. */
#define MALLOC(p,m,n) if ((p)=calloc((n),sizeof(*(p)))) (p)-=(m); else printf("Insufficient memory!\n"), exit(1)
/* Vararg terminators: a quiet NaN whose low mantissa word carries a small integer. */
#define INT2NAN(i) ((union ieee754_double){.ieee.exponent=-1,.ieee.mantissa1=i}.d)
#define NAN2INT(n) ((union ieee754_double){.ieee.exponent=-1}.ieee.exponent==(union ieee754_double){.d=n}.ieee.exponent&&!(union ieee754_double){.d=n}.ieee.mantissa0?(union ieee754_double){.d=n}.ieee.mantissa1:0)

typedef struct world *world;
typedef struct matter matter;
typedef struct object object;
typedef struct source *source;
typedef struct tfsf *tfsf;
typedef struct slice *slice;
typedef struct phaser *phaser;
typedef float field(world, int, int, int);

/* Resolution: time step, cell sizes, eV scale / sub-pixel count, non-uniform grid,
 * optional grid functions (src/alis_c.h:50-70). */
typedef struct {
	float dt;
	struct { float x, y, z; } d;
	struct { float eV; int subN; } spec;
	struct { struct { int m; float Min, Max; } x, y, z; } nug;
	struct {
		float (*itox)(world, float);
		float (*jtoy)(world, float);
		float (*ktoz)(world, float);
	} gridFunction;
} res;

/* Domain extents in nm (src/alis_c.h:72-76). */
typedef struct {
	struct { float Min, Max; } x, y, z;
} dom;

/* Surfaces (boundary codes or periods) and CPML parameters (src/alis_c.h:78-86). */
typedef struct {
	struct { float MinSur, MaxSur; } x, y, z;
	struct { int Num; float SigmaMax, KappaMax, AlphaMax, SigmaPower, AlphaPower; } pml;
} sur;

/* A vector field: three components, CPML auxiliary slabs, ADE polarisation currents. */
typedef struct {
	float ***x, ***y, ***z;
	float ***yMinPx, ***zMinPx, ***xMinPy, ***zMinPy, ***xMinPz, ***yMinPz;
	float ***yMaxPx, ***zMaxPx, ***xMaxPy, ***zMaxPy, ***xMaxPz, ***yMaxPz;
	float *****Jx, *****Jy, *****Jz, *****Kx, *****Ky, *****Kz;
	float t;
} vfield;

/* Update coefficients of one vector field (src/alis_c.h:96-107). */
typedef struct {
	float *dx, *dy, *dz, *dtdx, *dtdy, *dtdz;
	float *xP, *yP, *zP, *xPy, *xPz, *yPz, *yPx, *zPx, *zPy;
	float ***x, ***y, ***z;
	float ****Jx, ****Jy, ****Jz;
	float *ax, *ay, *az;
	float **JJx, **JJy, **JJz, **JKx, **JKy, **JKz, **JEx, **JEy, **JEz;
	float **KJx, **KJy, **KJz, **KKx, **KKy, **KKz, **KEx, **KEy, **KEz;
	int N, *NJ, *NK, *iMin, *iMax, *jMin, *jMax, *kMin, *kMax;
	object *O;
	matter *M;
} coeffs;

struct world {
	dom Dom;
	res Res;
	sur Sur;
	float t, dt, dx, dy, dz, eV, f;
	float xMin, xMax, yMin, yMax, zMin, zMax;
	float xMIN, xMAX, yMIN, yMAX, zMIN, zMAX;
	float coskx, sinkx, cosky, sinky, coskz, sinkz, coskyp, coskym, sinkyp, sinkym;
	float xSize, ySize, zSize, xyArea, xzArea, yzArea;
	int iMin, iMax, iNum, jMin, jMax, jNum, kMin, kMax, kNum;
	int iMIN, iMAX, iNUM, jMIN, jMAX, jNUM, kMIN, kMAX, kNUM;
	int xMinSur, xMaxSur, yMinSur, yMaxSur, zMinSur, zMaxSur;
	int subN, complexField, sourceType;
	unsigned long int n, N, T;
	int (*xtoi)(world, float), (*ytoj)(world, float), (*ztok)(world, float);
	float (*itox)(world, float), (*jtoy)(world, float), (*ktoz)(world, float);
	char ID[128];
	vfield *E, *H;
	coeffs *CE, *CH;
	source SE, SH;
	tfsf S;
	/* ---- private engine state (not part of the reference struct) ---- */
	void *dev;          /* committed device world (alis_dev_world), or 0      */
	int devPrec;        /* 0: FP32 fields, 1: FP64 fields                      */
	int devDftPrec;     /* 0: FP32 phaser accumulators, 1: FP64                */
	int hostStale;      /* host mirrors of E/H/J/K older than the device copy  */
	int isAux;          /* 1 for a TF/SF auxiliary world                       */
	world parent;       /* owning world of an auxiliary world                  */
	int expandX, expandY; /* row aliasing of a TF/SF auxiliary world          */
	void *probes;       /* the boxes get() reads (host/residency.c), or 0      */
	void *nodes;        /* nearest nodes already found (host/grid.c), or 0     */
};

world createWorld(dom, res, sur, char*, ...);
world cloneWorld(world);
void rebootWorld(world W);
void deleteWorld(world W);
void deleteCloneWorld(world W);
float ***makeField(int, int, int, int, int, int);
ALIS_CPLX(float) ***makeComplexField(int, int, int, int, int, int);
#define removeField(F,i,j,k) free(&F[i][j][k]),free(&F[i][j]),free(&F[i]),F=0
void *h5read(int*, char*, ...);
void h5write(int*, float*, char*, ...);

/* Objects, shapes and materials (src/alis_c.h:144-176). */
typedef int shape(object, float, float, float);

struct matter {
	struct { float x, y, z; } e;
	struct { float w, r, f, g; } P[MAXPOLES];
	float d[3][6];
};

struct object {
	shape *S;
	float f[3][3];
	void *p;
};

void putObjectArray(world, matter*, object*);
void putObjects(world, ...);
#define putObjects(...) ((putObjects)(__VA_ARGS__,(object){0}))
void removeObjects(world);
matter LT(matter, float);
matter nkFile(char*, float, float);
shape Box, Ball, Diamond;
shape RodX, RodY, RodZ, DRodX, DRodY, DRodZ, BiconeX, BiconeY, BiconeZ;
shape HRodXY, HRodXZ, HRodYZ, HRodYX, HRodZX, HRodZY;
shape Translation, Rotation;
shape Combination, Intersection, Difference;
shape Lattice, DLattice, CLatticeX, CLatticeY, CLatticeZ;
shape ImportContour;

/* Time stepping (src/alis_c.h:179-188). */
void copyBoundary(float***, float***, float, float, int, int, int, int, int, int, int, int, int);
#define copyBoundary(W,F,x,...) ((copyBoundary)(W->F[0].x,W->complexField?W->F[1].x:0,__VA_ARGS__))
void updateE(world);
void updateH(world);
void nonLinearSHG(world, world, float);
void nonLinearTWM(world, world, world, float);
int timer(long int, long int, char*, ...);
#define timer(...) ((timer)(__VA_ARGS__,""))

/* Fields, slices and output (src/alis_c.h:191-248). */
typedef void (*method)(world, slice, float, char*, char*, unsigned long int*);
typedef unsigned long int colormap[];
struct slice {
	int iMin, iMax, jMin, jMax, kMin, kMax;
	int iMIN, iMAX, iNUM, jMIN, jMAX, jNUM, kMIN, kMAX, kNUM;
	float ***C, ***F;
	float ***(*interpolate)(world, slice, float***);
};

void exec(char*, ...);
void writeTxt(world, char*, ...);
void writeRow(world, char*, ...);
void writeSpectrum(world, int, float, float, char*, ...);
#define writeTxt(...) ((writeTxt)(__VA_ARGS__,""))
#define writeRow(...) ((writeRow)(__VA_ARGS__,INT2NAN(1)))
#define writeSpectrum(...) ((writeSpectrum)(__VA_ARGS__,INT2NAN(1)))
slice createSlice(world, float, float, float, float, float, float);
#define createSliceX(W,yCut,zCut) createSlice(W, W->xMin, W->xMax, yCut, yCut, zCut, zCut)
#define createSliceY(W,xCut,zCut) createSlice(W, xCut, xCut, W->yMin, W->yMax, zCut, zCut)
#define createSliceZ(W,xCut,yCut) createSlice(W, xCut, xCut, yCut, yCut, W->zMin, W->zMax)
#define createSliceXY(W,zCut) createSlice(W, W->xMin, W->xMax, W->yMin, W->yMax, zCut, zCut)
#define createSliceXZ(W,yCut) createSlice(W, W->xMin, W->xMax, yCut, yCut, W->zMin, W->zMax)
#define createSliceYZ(W,xCut) createSlice(W, xCut, xCut, W->yMin, W->yMax, W->zMin, W->zMax)
void deleteSlice(slice);
void (h5)(world, slice, float, char*, char*, colormap);
void (txt)(world, slice, float, char*, char*, colormap);
void (png)(world, slice, float, char*, char*, colormap);
void sliceSnap(world, field, slice, char*, char*, ...);
#define sliceSnap(W,F,S,...) ((sliceSnap)(W, F, S, #F, #S, __VA_ARGS__))
void sliceTimeAvg(world, field, slice, char*, char*, ...);
#define sliceTimeAvg(W,F,S,...) ((sliceTimeAvg)(W, F, S, #F, #S, __VA_ARGS__))
void sliceFreqDom(world, field, slice, char*, char*, int, float, ...);
#define sliceFreqDom(W,F,S,...) ((sliceFreqDom)(W, F, S, #F, #S, __VA_ARGS__))
float sliceMax(world, field, slice);
float sliceSum(world, field, slice);
float worldMax(world, field);
float worldSum(world, field);
float get(world, field, float, float, float);
float objectAbsorption(world, object);
float poyntingX(world W, float, float, float, float, float, ...);
float poyntingY(world W, float, float, float, float, float, ...);
float poyntingZ(world W, float, float, float, float, float, ...);
float poyntingOut(world, float, float, float, float, float, float);
#define poyntingX(...) ((poyntingX)(__VA_ARGS__,-INF,INF,-INF,INF))
#define poyntingY(...) ((poyntingY)(__VA_ARGS__,-INF,INF,-INF,INF))
#define poyntingZ(...) ((poyntingZ)(__VA_ARGS__,-INF,INF,-INF,INF))
#define poyntingIn(...) (-poyntingOut(__VA_ARGS__))
field RI, LogRI, ContourRI, Contour;
field raweEffx, raweEffy, raweEffz, eEff;
field rawEx, rawEy, rawEz, rawHx, rawHy, rawHz, rawJx, rawJy, rawJz;
field Ex, Ey, Ez, Hx, Hy, Hz, iEx, iEy, iEz, iHx, iHy, iHz, Jx, Jy, Jz;
field ExHy, ExHz, EyHz, EyHx, EzHx, EzHy, Sx, Sy, Sz;
field JE, EE, UE, HH, UH, U, LogJE, LogEE, LogUE, LogHH, LogUH, LogU;
field ScEx, ScEy, ScEz, ScEE, SoEx, SoEy, SoEz, SoEE;
field DivE, JxHy, JxHz, JyHz, JyHx, JzHx, JzHy, Fx, Fy, Fz;
field Txx, Txy, Txz, Tyx, Tyy, Tyz, Tzx, Tzy, Tzz;

/* Phasers and far field (src/alis_c.h:251-299). */
typedef void projmap(int, int, int, int, float[2]);

typedef struct {
	float A;
	ALIS_CPLX(float) *Ex, *Ey, *Ez, *Hx, *Hy, *Hz;
} tfield;

typedef struct {
	ALIS_CPLX(float) ***Ex, ***Ey, ***Ez, ***Hx, ***Hy, ***Hz;
} cfield;

struct phaser {
	world W;
	int iMin, iMax, jMin, jMax, kMin, kMax, kMIN, xNUM, yNUM;
	unsigned long int N;
	float f, A, *dE, *dH;
	ALIS_CPLX(float) expikx, expiky, *ex, *ez;
	cfield xMin, xMax, yMin, yMax, zMin, zMax;
	/* ---- private engine state ---- */
	int count;          /* phasers in this array (set on the first element)     */
	int devId;          /* index of the array in the device world, -1 if none   */
	int hostStale;      /* cfield host copies older than the device accumulators */
	double *view;       /* double copy of the accumulators (see phasers.c), or 0  */
	int viewValid;      /* view matches the accumulators                          */
	double *faceFlux;   /* [frequency][face] net flux terms (see phasers.c), or 0  */
	int faceFluxValid;  /* faceFlux matches the accumulators                      */
};

phaser createPhaser(world, float, float, float, float, float, float, float, ...);
#define createPhaser(...) ((createPhaser)(__VA_ARGS__,-INF,INF,-INF,INF,-INF,INF))
phaser createPhasers(world, float, float, float, float, float, float, float, float, float, ...);
#define createPhasers(...) ((createPhasers)(__VA_ARGS__,-INF,INF,-INF,INF,-INF,INF))
phaser choosePhaser(phaser, float);
void updatePhaser(world, phaser);
#define updatePhasers(...) updatePhaser(__VA_ARGS__)
void deletePhaser(phaser);
#define deletePhasers(...) deletePhaser(__VA_ARGS__)
tfield Ppol(phaser, float, float);
tfield Spol(phaser, float, float);
tfield Unpol(phaser, float, float);
float phaserPoynting(phaser);
float phaserPoyntingIn(phaser);
float phaserPoyntingOut(phaser);
ALIS_CPLX(float) farField(phaser, tfield(phaser,float,float), float, float);
float farFieldI(phaser, tfield(phaser,float,float), float, float);
float farFieldFlux(phaser, tfield(phaser,float,float), float, float, float, float, float, ...);
#define farFieldFlux(...) ((farFieldFlux)(__VA_ARGS__,0,0))
void farFieldTheta(world, phaser, float, char*, ...);
void farFieldPhi(world, phaser, float, char*, ...);
void farFieldProfile(world, phaser, projmap, int, int, ...);
void writePoyntingSpectrum(world, char*, ...);
#define writePoyntingSpectrum(...) ((writePoyntingSpectrum)(__VA_ARGS__,0))
void writeFarFieldSpectrum(world, tfield(phaser,float,float), float, float, char*, ...);
#define writeFarFieldSpectrum(...) ((writeFarFieldSpectrum)(__VA_ARGS__,0))
projmap CylindericalMap, AzimuthalMap, NorthernAzimuthalMap, SouthernAzimuthalMap;

/* Sources (src/alis_c.h:302-334). */
typedef float waveform(world, source, float, float);
struct source {
	waveform *waveForm;
	float wavelength, supplement, resolution, amplitude, phase, frequency, deviation, peaktime;
	float ***Amplitude, ***Phase, ***F;
	int iMin, iMax, jMin, jMax, kMin, kMax, nMin, nMax;
	unsigned long int N;
	source S;
};
struct tfsf {
	int iMin, iMax, jMin, jMax, kMin, kMax;
	int iMIN, iMAX, jMIN, jMAX, kMIN, kMAX;
	world W;
	tfsf S;
};

waveform Sine, Pulse, Band;
tfsf createTFSF(world, dom, sur);
void pointDipole(world, field, float, float, float, waveform, float, float, float, float, ...);
void guidedWaveX(world W, char *file, field F, float x, waveform waveForm, float wavelength, float supplement, float amplitude, float phase, ...);
void guidedWaveY(world W, char *file, field F, float x, waveform waveForm, float wavelength, float supplement, float amplitude, float phase, ...);
void guidedWaveZ(world W, char *file, field F, float x, waveform waveForm, float wavelength, float supplement, float amplitude, float phase, ...);
void planewave(world, dom, tfield(phaser,float,float), float, float, waveform, float, float, float, float, ...);
void focusedBeam(world, dom, field, float, float, float, float, waveform, float, float, float, float, float, ...);
void removeSources(world W);
#define pointDipole(...) ((pointDipole)(__VA_ARGS__,0,0,0,0,0))
#define guidedWaveX(...) ((guidedWaveX)(__VA_ARGS__,0,0,0,0,0))
#define guidedWaveY(...) ((guidedWaveY)(__VA_ARGS__,0,0,0,0,0))
#define guidedWaveZ(...) ((guidedWaveZ)(__VA_ARGS__,0,0,0,0,0))
#define planewave(...) ((planewave)(__VA_ARGS__,0,0,0,0,0))
#define focusedBeam(...) ((focusedBeam)(__VA_ARGS__,0,0,0,0,0,0))

/* ---- engine extensions (no counterpart in the reference) ---- */
/* Field/accumulator precision of a world: fieldBits and dftBits are 32 or 64.  Must be
 * called before the first time step; the default comes from ALIS_PRECISION
 * ("fp32" = FP32 fields + FP64 DFT, "fp64" = FP64 fields + FP64 DFT, "ref" = FP32
 * fields + FP32 DFT as in the reference), else "fp32". */
void worldPrecision(world W, int fieldBits, int dftBits);
/* Copy every device-resident quantity of W (fields, ADE state, phasers) to the host
 * arrays of the reference layout.  Library observers call it themselves. */
void syncWorld(world W);
/* One E (H) half step of n worlds and their TF/SF auxiliary worlds in shared launches; the
 * same results as updateE (updateH) on each world in turn. */
void updateWorldsE(world *W, int n);
void updateWorldsH(world *W, int n);
/* One DFT sample of the phaser arrays P[0..n-1] (P[a] recording world W[a]) in one launch;
 * the same as updatePhaser(W[a], P[a]) on each. */
void updatePhaserArrays(world *W, phaser *P, int n);

#ifdef __cplusplus
}
#endif

#ifndef ALIS_LIBRARY
/*
 * Material database and colour maps, defined in the user translation unit.
 * Matter format {{ex, ey, ez}, {pole, ...}} with poles {0,sigma} (conductivity),
 * {w,gamma} (Drude), {w,gamma,f} (Lorentz), {w,gamma,f,g} (critical point); energies in eV.
 * Parameter values are those of the reference database (src/alis_m.h:18-66).
 */
matter PM = {0};
matter Air = {1};
matter Vacuum = {1};

#ifdef NUV
/* 250-750 nm fits */
matter Ag = {{1.9201},{{9.3937,0.0302},{4.2333,1.2373,1.3399,1.6141},{3.9548,0.259,-0.1356,0.2961},{0.8394,0.5873,2.652,6.1529}}};
matter Au = {{2.03},{{8.4306,0.0},{4.3848,1.2748,-0.2723,-3.7849},{4.7338,2.0464,5.3396,4.7241},{2.6147,0.3187,0.371,0.723}}};
matter Cu = {{1.7411},{{8.8195,0.0835},{4.6887,0.7044,0.3576,0.4646},{3.4555,2.8361,7.7571,-0.1944},{2.1607,0.2625,0.0964,1.0257}}};
matter Al = {{1.0},{{13.4862,0.1406},{2.24,1.4861,3.8125,-0.5311},{1.8272,0.3877,3.8946,2.3493},{1.5717,0.2272,9.7451,-1.7878}}};
matter Cr = {{1.0009},{{9.1925,0.7681},{3.4893,0.4922,0.7387,-0.0827},{0.9432,0.2803,-21.3216,29.9879},{2.4741,0.6165,2.6444,-5.3032}}};
matter Ni = {{2.04},{{7.8334,0.5169},{4.3024,1.0858,2.3911,1.602},{1.3054,0.4506,8.9446,6.5275},{3.3426,1.7382,1.3531,-3.329}}};
matter Ti = {{1.0},{{12.3768,5.5499},{3.9759,0.8038,-1.0395,-0.3514},{2.1995,1.048,-3.7293,0.9582},{1.0376,0.2488,-0.596,6.2558}}};
matter Pd = {{1.7139},{{9.3495,0.7156},{4.7399,0.7234,0.2268,0.0691},{2.0356,3.4652,6.1216,-2.2174},{1.3945,0.7049,7.316,5.9221}}};
matter Pt = {{1.9043},{{11.5026,1.6472},{1.2596,1.3057,5.818,14.2589}}};
matter Ge = {{1.0},{{3.3888,1.0394,6.1945,3.7488},{4.0446,1.7946,5.5177,-2.0495},{4.3299,0.2043,1.1731,-0.9613},{2.269,0.2808,3.7623,1.9046},{2.1174,0.0858,0.5609,0.0901}}};
matter Si = {{1.0},{{4.2785,0.2217,2.9926,-0.7684},{5.6927,1.5593,2.3925,-1.2208},{3.6614,0.4785,4.1976,2.2951},{3.3756,0.1172,1.3635,1.0095}}};
matter aSi = {{1.0},{{2.9963,0.7273,2.1465,1.9077},{2.083,1.0256,0.434,5.1983},{4.3391,1.9704,3.9252,-1.6418},{3.5877,1.0243,5.027,-1.4322}}};
#else
/* 400-2000 nm fits */
matter Ag = {{4.0323},{{9.1864,0.021}}};
matter Au = {{2.1452},{{8.4146,0.0255},{2.6203,0.2147,0.2082,0.3059},{3.1543,1.1662,4.1353,2.5005}}};
matter Cu = {{4.0934},{{9.0117,0.0943},{3.1915,1.534,3.6969,-0.0729},{2.1831,0.2548,0.2499,0.914}}};
matter Al = {{1.0},{{13.3237,0.1128},{1.492,0.2014,8.2647,4.9317},{0.0844,1.2677,23.8069,-4.9835},{1.2054,1.0353,21.7077,-5.7556}}};
matter Cr = {{1.0},{{4.9087,0.0465},{2.4765,0.4536,1.5602,-2.1909},{0.3726,0.1348,21.4832,36.8394},{1.2778,1.0631,49.801,4.9223}}};
matter Ni = {{1.0},{{5.5189,0.0658},{1.6545,0.6339,0.6792,-3.2872},{0.3545,0.1956,57.2374,79.7729}}};
matter Ti = {{1.28},{{12.6869,4.8641},{0.4422,0.0729,-19.4166,-3.9574},{1.1044,0.9335,-12.1464,8.0836},{1.5619,0.5214,3.0698,-3.7725}}};
matter Pd = {{1.5512},{{6.049,0.0379},{0.3874,0.2534,87.1642,5.9413},{0.6749,1.6453,87.1631,5.9296},{0.6006,0.7056,87.1633,5.9317}}};
matter Pt = {{1.0513},{{6.5658,0.2092},{1.1371,1.1336,13.527,-0.9232},{0.6331,0.1574,0.3107,19.5815},{0.7853,0.4354,15.1059,-0.6342},{0.4794,0.0923,-5.5467,26.6143}}};
matter Ge = {{1.6259},{{2.0327,0.2169,-0.6161,2.889},{3.5657,0.5248,5.8672,1.3512},{2.097,0.6267,10.4417,2.3849}}};
matter Si = {{3.7334},{{3.6745,0.1452,8.0789,0.6725}}};
matter aSi = {{1.6976},{{3.59,1.087,11.3535,-2.5534},{2.1662,1.1161,-1.0961,5.9546}}};
#endif
/* 400-1040 nm fits */
matter oCdS = {{1.0},{{5.0323,0.7745,2.3657,0.1689},{2.7075,0.4865,0.2835,0.4341},{2.4458,0.0664,0.0553,0.0341},{6.1838,7.2264,1.4749,-0.1979}}};
matter eCdS = {{1.0},{{4.6077,0.7854,2.5833,0.4566},{2.4621,0.0784,0.0674,0.0377},{2.7499,0.4034,0.2438,0.2882},{5.2255,7.1767,1.43,0.1797}}};

/* Colour maps: entry 0 is the number of colours (negative: symmetric range). */
colormap gray= { 2, 0xFFFFFF, 0x000000};
colormap red = { 3, 0xFFFFFF, 0xFF0000, 0x000000};
colormap blue= { 3, 0xFFFFFF, 0x0000FF, 0x000000};
colormap dkbr= {-5, 0x002244, 0x0022FF, 0xFFFFFF, 0xFF2200, 0x442200};
colormap dkrb= {-5, 0x442200, 0xFF2200, 0xFFFFFF, 0x0022FF, 0x002244};
colormap hsv = { 7, 0xFF0000, 0xFFFF00, 0x00FF00, 0x00FFFF, 0x0000FF, 0xFF00FF, 0xFF0000};
colormap hot = { 9, 0x000000, 0x550000, 0xAA0000, 0xFF0000, 0xFF5500, 0xFFAA00, 0xFFFF00, 0xFFFF80, 0xFFFFFF};
colormap jet = {16, 0x0000C0, 0x0000FF, 0x0040FF, 0x0080FF, 0x00C0FF, 0x00FFFF, 0x40FFFF, 0x80FFC0,
                    0xC0FF80, 0xFFFF40, 0xFFFF00, 0xFFC000, 0xFF8000, 0xFF4000, 0xFF0000, 0xC00000};
#endif /* ALIS_LIBRARY */

#endif /* ALIS_H */
