/*
 * Benchmark workloads written against the public alis.h API only, so the same file builds
 * against this library (alis_amd/libalis_workloads.so) and against the reference's own C
 * sources (oracle/Makefile, for the CPU baseline).
 *
 *   nanoblock  TF/SF scattering by a 60x60x40 nm gold block (Drude pole + two Lorentz
 *              pairs, eps_inf 1.5) on an n=1.5 substrate: dx = 2 nm, dt = 1 nm/c (Courant
 *              0.5), 16-cell CPML on all faces, x-polarised plane wave with a 400-900 nm
 *              band waveform, a phaser box of 32 frequencies around the TF/SF box.  `cells`
 *              is the grid edge including CPML (256 for the 256^3 workload).
 *   dipole     Ez point dipole (600 nm pulse) in vacuum: dx = 10 nm, Courant 0.5, 10-cell
 *              CPML, 8 fixed probes; `cells` = grid edge including CPML (64).
 *
 * One step is updateE + updateH (+ updatePhaser for nanoblock), src/update.c:275-341.
 */
#include "alis.h"
#include <string.h>
#include <time.h>

typedef struct {
	world W;
	phaser P;
	int probes;
	float pts[8][3];
	double probeSum;
} wl;

/* Domain extents giving `cells` grid nodes including two CPML layers of pml+1 cells:
 * nodes iMin..iMax = -(d-2)/2 .. d/2 with d = cells - 2(pml+1), src/world.c:229-234. */
static void extent(int cells, int pml, float dx, float *lo, float *hi)
{
	int d = cells - 2*(pml+1);
	*lo = -dx * ((d-2)/2);
	*hi = dx * (d/2);
}

void *wl_create(const char *name, int cells, int fieldBits, int dftBits)
{
	wl *c = calloc(1, sizeof *c);
	if (!strcmp(name, "nanoblock")) {
		float lo, hi;
		extent(cells, 16, 2, &lo, &hi);
		dom D = {{lo, hi}, {lo, hi}, {lo, hi}};
		res R = {1, {2}, {1240}};
		sur S = {{PML}, {PML}, {PML}, {16}};
		c->W = createWorld(D, R, S, "");
		matter Gold = {{1.5}, {{9, 0.07}, {2.6, 0.3, 0.5}, {3.5, 0.6, 0.8}}};
		object Block = {Box, {{-30, 30}, {-30, 30}, {0, 40}}};
		object Substrate = {Box, {{-INF, INF}, {-INF, INF}, {-INF, 0}}};
		putObjects(c->W, Gold, Block, n(1.5), Substrate, Air);
		dom DTF = {{-50, 50}, {-50, 50}, {-30, 60}};
		planewave(c->W, DTF, Ppol, 0, 0, Band, 400, 900);
		c->P = createPhasers(c->W, 400, 900, 16.5, -70, 70, -70, 70, -50, 80);
	}
	else if (!strcmp(name, "dipole")) {
		float lo, hi;
		extent(cells, 10, 10, &lo, &hi);
		dom D = {{lo, hi}, {lo, hi}, {lo, hi}};
		res R = {5, {10}, {1240}};
		sur S = {{PML}, {PML}, {PML}, {10}};
		c->W = createWorld(D, R, S, "");
		pointDipole(c->W, Ez, 0, 0, 0, Pulse, 600, 240);
		float pts[8][3] = {{0,0,50}, {50,0,0}, {0,-50,0}, {80,80,0}, {-100,30,60}, {40,-120,-40}, {150,150,150}, {-170,0,-90}};
		memcpy(c->pts, pts, sizeof pts);
		c->probes = 8;
	}
	else {
		free(c);
		return 0;
	}
#ifndef ALIS_REF
	worldPrecision(c->W, fieldBits, dftBits);
#else
	(void)fieldBits, (void)dftBits;
#endif
	return c;
}

void wl_step(void *ctx, int n)
{
	wl *c = ctx;
	for (int s = 0; s < n; s++) {
		updateE(c->W);
		updateH(c->W);
		if (c->P) updatePhaser(c->W, c->P);
		for (int p = 0; p < c->probes; p++) {
			c->probeSum += get(c->W, Ez, c->pts[p][0], c->pts[p][1], c->pts[p][2]);
			c->probeSum += get(c->W, Hy, c->pts[p][0], c->pts[p][1], c->pts[p][2]);
		}
	}
}

/* Grid nodes of the main world including CPML (the cell count of one step). */
long wl_cells(void *ctx)
{
	wl *c = ctx;
	return (long)c->W->iNUM * c->W->jNUM * c->W->kNUM;
}

/* Phaser frequencies and surface cells (for the DFT byte count). */
int wl_phaser_info(void *ctx, long *surfaceCells)
{
	wl *c = ctx;
	*surfaceCells = 0;
	if (!c->P) return 0;
	phaser P = c->P;
	long ni = P->iMax-P->iMin+1, nj = P->jMax-P->jMin+1, nk = P->kMax-P->kMin+1;
	world W = c->W;
	*surfaceCells = (W->xMinSur==PML) * nj*nk + (W->xMaxSur==PML) * nj*nk + (W->yMinSur==PML) * ni*nk
	              + (W->yMaxSur==PML) * ni*nk + (W->zMinSur==PML) * ni*nj + (W->zMaxSur==PML) * ni*nj;
	int n = 0;
	while (P[n].f) n++;
	return n;
}

/* CPML layer thickness (Num+1 cells) per face, 0 where the face is not PML. */
int wl_geometry(void *ctx, int *num, int *pml)
{
	wl *c = ctx;
	world W = c->W;
	num[0] = W->iNUM, num[1] = W->jNUM, num[2] = W->kNUM;
	pml[0] = W->xMinSur==PML ? W->iMin-W->iMIN : 0, pml[1] = W->xMaxSur==PML ? W->iMAX-W->iMax : 0;
	pml[2] = W->yMinSur==PML ? W->jMin-W->jMIN : 0, pml[3] = W->yMaxSur==PML ? W->jMAX-W->jMax : 0;
	pml[4] = W->zMinSur==PML ? W->kMin-W->kMIN : 0, pml[5] = W->zMaxSur==PML ? W->kMAX-W->kMax : 0;
	return W->CE->N;
}

/* Sum of squares of every E and H component over the grid (sequential, in double), plus
 * the probe samples: equal for two runs exactly when their states agree to the bit (up to
 * cancellation that a state difference cannot plausibly produce). */
double wl_checksum(void *ctx)
{
	wl *c = ctx;
#ifndef ALIS_REF
	syncWorld(c->W);
#endif
	double s = c->probeSum;
	world W = c->W;
	float ***F[6] = { W->E->x, W->E->y, W->E->z, W->H->x, W->H->y, W->H->z };
	for (int a = 0; a < 6; a++)
		for (int i = W->iMIN; i <= W->iMAX; i++)
			for (int j = W->jMIN; j <= W->jMAX; j++)
				for (int k = W->kMIN; k <= W->kMAX; k++)
					s += (double)F[a][i][j][k] * F[a][i][j][k];
	return s;
}

void wl_free(void *ctx)
{
	wl *c = ctx;
	if (c->P) deletePhaser(c->P);
	deleteWorld(c->W);
	free(c);
}

#ifndef ALIS_REF
/* Device world of a committed workload (after its first step), for kernel timing. */
void *wl_device(void *ctx)
{
	return ((wl *)ctx)->W->dev;
}
#endif

#ifdef WL_MAIN
/* CPU timing program: wl NAME CELLS WARMUP STEPS -> "seconds_per_step cells threads". */
#include <omp.h>
#include <unistd.h>
int main(int argc, char **argv)
{
	if (argc < 5) return fprintf(stderr, "usage: %s NAME CELLS WARMUP STEPS\n", argv[0]), 2;
	/* the library's progress lines go to /dev/null, the result line to the original stdout */
	fflush(stdout);
	FILE *keep = fdopen(dup(1), "w");
	if (!keep || !freopen("/dev/null", "w", stdout)) return 1;
	void *c = wl_create(argv[1], atoi(argv[2]), 32, 32);
	if (!c) return 2;
	wl_step(c, atoi(argv[3]));
	struct timespec a, b;
	clock_gettime(CLOCK_MONOTONIC, &a);
	int steps = atoi(argv[4]);
	wl_step(c, steps);
	clock_gettime(CLOCK_MONOTONIC, &b);
	double sec = (b.tv_sec - a.tv_sec) + 1e-9 * (b.tv_nsec - a.tv_nsec);
	fprintf(keep, "%.9g %ld %d %.17g\n", sec / steps, wl_cells(c), omp_get_max_threads(), wl_checksum(c));
	fclose(keep);
	return 0;
}
#endif
