/*
 * Double-precision access to phaser accumulators (shared by phasers.c and farfield.c).
 */
#ifndef ALIS_PHASERVIEW_H
#define ALIS_PHASERVIEW_H
#include "internal.h"

/* Faces present (order xMin xMax yMin yMax zMin zMax), their first cell in the concatenated
 * cell list and the length of their fast axis. */
typedef struct {
	int on[6];
	long off[6];
	int n2[6];
	long ncell;
} phaserGeom;

void phaserFaces(phaser B, phaserGeom *g);
const double *phaserDoubles(phaser B, phaserGeom *g);
void neumaierAdd(double *sum, double *comp, double x);
double phaserFluxD(phaser P);
source normalisingSource(world W);

#endif
