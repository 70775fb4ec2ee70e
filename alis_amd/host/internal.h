/*
 * Internal declarations shared by the host library sources (not installed).
 */
#ifndef ALIS_INTERNAL_H
#define ALIS_INTERNAL_H

#define ALIS_LIBRARY
#include "alis.h"
#include "alis_dev.h"

/* grid.c */
/* One axis of a world (0 x, 1 y, 2 z): pointers to its members in the reference's struct. */
typedef struct {
	int *Min, *Max, *Num, *MIN, *MAX, *NUM, *minSur, *maxSur;
	float *lo, *hi, *LO, *HI, *size, *step, *cos, *sin;
	float (**coord)(world, float);
	int (**node)(world, float);
} axis;
axis axisOf(world W, int a);
float ***makeFieldFilled(int iMin, int iMax, int jMin, int jMax, int kMin, int kMax, float v);
void resetField3(float ***f, int iMin, int iMax, int jMin, int jMax, int kMin, int kMax);
void makeAuxFieldArrays(world W, vfield *F, coeffs *C);

/* media.c */
int shapeBounds(const object *O, float b[6]);   /* basic shapes: widened bounds, else 0 */
int dispersiveMatter(matter M);   /* 2: nonlinear tensor, 1: poles, 0: plain dielectric */
float complex polePermittivity(matter M, float w, int p);

/* engine.c: device residency of a world */
int alisPrecisionDefault(int *fieldBits, int *dftBits);
void alisCommit(world W);          /* create the device world if absent                */
void alisInvalidate(world W);      /* setup changed: bring state home, drop the device */
void alisDropDevice(world W);      /* host arrays are newer: drop the device copy      */
void alisSyncBox(world W, int i0, int i1, int j0, int j1, int k0, int k1);
void alisSyncArrays(world W, int i0, int i1, int j0, int j1, int k0, int k1, int mask);   /* the arrays in mask over a box */
void alisProbe(world W, int i0, int i1, int j0, int j1, int k0, int k1, int mask);   /* get(): box copy of the arrays in mask, batched per half step */
void alisStepped(void);            /* device fields changed: probe copies are stale */
void alisSyncPhaser(phaser P);
void alisPhaserDevice(phaser B);   /* register the array B on the device if it is not */
alis_dev_batch *alisBatch(world *L, int n);
void alisSyncAll(void);            /* every live world and phaser (timer() hook)        */
void alisRegisterWorld(world W);
void alisForgetWorld(world W);
void alisRegisterPhaser(phaser P);
void alisForgetPhaser(phaser P);
void alisFatal(const char *what, int rc);
phaser phaserBase(phaser P);     /* first element of the array holding P, or 0 */

/* observers.c */
/* format a name from a copy of *ap and advance *ap past the arguments its format names */
void alisFormatName(char *out, size_t size, const char *format, va_list *ap);
void alisForgetTexts(world W);    /* close the text outputs of a world */
void alisForgetSpectra(world W);  /* drop the writeSpectrum records of a world */

/* describe.c */
void describeWorld(world W, alis_dev_world_desc *d);
void freeWorldDesc(alis_dev_world_desc *d);
void describePhaser(phaser B, alis_dev_phaser_desc *d);

/* sources.c */
void sourceHalfStep(world W, source S, float t, int h);  /* evaluate and inject a chain */

/* fields.c: field identity helpers used by sources and probes */
int fieldComponent(field *F, int *isH, int *m, int *axis);
int rawComponent(field *F, int *isH, int *m, int *axis);

#endif
