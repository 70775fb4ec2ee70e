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
 *   particles  the nanoblock grid with 64 gold spheres (radius 10-16 nm, pitch 80 nm) in
 *              glass instead of the block, TF/SF plane wave, no phaser: many dispersive
 *              objects.
 *   dipole     Ez point dipole (600 nm pulse) in vacuum: dx = 10 nm, Courant 0.5, 10-cell
 *              CPML, 8 fixed probes; `cells` = grid edge including CPML (64).
 *   kpoints    a band sweep: `cells` worlds (k-points along Gamma-X-M), each a photonic
 *              crystal slab cell of 64 x 64 x 31 cells plus an 8-cell CPML above and below
 *              (dx = 10 nm; Bloch-periodic x and y with period 640 nm = 2a, a = 320 nm; an
 *              eps = 12 slab 0.5a thick with air holes of radius 0.3a on the square lattice
 *              of period a; complex fields), 8 point dipoles per world at positions and
 *              phases drawn from a splitmix64 stream (seed 2024), and 8 probes per world at
 *              positions from the same stream (Hz, Ez alternately) whose values feed one
 *              writeSpectrum per world every step: 256 frequency bins over 145.5-2000 nm
 *              for an 8000-step run.  All worlds step together with
 *              updateWorldsE/updateWorldsH.
 *   metasurface an 8 x 8 array of TiO2 pillars (eps 6.25, 600 nm tall, radii drawn uniformly
 *              from 60-120 nm by a splitmix64 stream, seed 2024) on glass (eps 2.25) in a
 *              periodic cell of 8 x 400 nm: dx = 5 nm, Courant 0.5, periodic x and y, 16-cell
 *              CPML in z, normal-incidence x-polarised plane wave with a 400-900 nm band
 *              waveform on a TF/SF box, one 550 nm phaser on the z faces.  `cells` is the
 *              x/y period in cells (640: 641 x 641 x 320 nodes including CPML, z edge half
 *              of it); the pitch stays 400 nm, so the array has cells/80 pillars a side.
 *
 * One step is updateE + updateH (+ updatePhaser for nanoblock; the probes and their spectra
 * for dipole and kpoints), src/update.c:275-341.  Every probe value read is kept
 * (wl_trace).
 */
#include "alis.h"
#include <string.h>
#include <time.h>
#include <stdint.h>
#ifndef ALIS_REF
#include "alis_dev.h"
#endif

typedef struct {
	world W;
	phaser P;
	int probes;
	float pts[8][3];
	double probeSum;
	double *trace;              /* every probe value read, in order */
	long ntrace, captrace;
	float (*kprobe)[8][3];      /* kpoints: probe nodes per world */
	int nw;                     /* worlds stepped together (W = Ws[0]) */
	world *Ws;
	object *parts;              /* objects the world refers to after putObjects */
} wl;

static uint64_t splitmix(uint64_t *s)
{
	uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
	return z ^ (z >> 31);
}

static double uniform(uint64_t *s) { return (splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

/* Record one probe value: it enters the checksum and the trace. */
static float keep(wl *c, float v)
{
	if (c->ntrace == c->captrace) {
		c->captrace = c->captrace ? 2 * c->captrace : 4096;
		c->trace = realloc(c->trace, c->captrace * sizeof(double));
	}
	c->trace[c->ntrace++] = v;
	c->probeSum += v;
	return v;
}

/* One k-point cell of the band sweep: q-th of nq points along Gamma-X-M. */
static world kpointWorld(int q, int nq, uint64_t *seed)
{
	const float a = 320, L = 2 * a;
	float t = (q + 0.5f) / nq * 2, kx, ky;                 /* path parameter 0..2 */
	if (t < 1) kx = 0.5f * t, ky = 0;
	else kx = 0.5f, ky = 0.5f * (t - 1);
	dom D = {{-L/2, L/2}, {-L/2, L/2}, {-150, 160}};     /* period L: node iMAX is node iMIN */
	res R = {5, {10}, {1240}};
	sur S = {{BBC, ky >= 0 && kx > 0 ? a/kx : INF}, {ky > 0 ? BBC : PBC, ky > 0 ? a/ky : INF}, {PML}, {8}};
	world W = createWorld(D, R, S, "");
	object Slab = {Difference, {2}, objects {
		{Box, {{-INF, INF}, {-INF, INF}, {-80, 80}}},
		{Lattice, {{a}, {a}}, objects {{RodZ, {{-96, 96}, {-96, 96}, {-INF, INF}}}}},
	}};
	putObjects(W, n(sqrtf(12)), Slab, Air);
	for (int d = 0; d < 8; d++) {
		float x = (float)(uniform(seed) * (L - 40) - L/2 + 20), y = (float)(uniform(seed) * (L - 40) - L/2 + 20);
		float z = (float)(uniform(seed) * 120 - 60), ph = (float)(uniform(seed) * 360);
		pointDipole(W, d % 2 ? Ez : Hz, x, y, z, Band, 1000, 2500, 1, ph);
	}
	return W;
}

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
	else if (!strcmp(name, "particles")) {
		/* 64 gold spheres (Drude + 2 Lorentz pairs) of radius 10-16 nm on a 4 x 4 x 4 lattice
		 * of pitch 80 nm in glass, TF/SF plane wave: the nanoblock grid with many dispersive
		 * objects */
		float lo, hi;
		extent(cells, 16, 2, &lo, &hi);
		dom D = {{lo, hi}, {lo, hi}, {lo, hi}};
		res R = {1, {2}, {1240}};
		sur S = {{PML}, {PML}, {PML}, {16}};
		c->W = createWorld(D, R, S, "");
		matter Gold = {{1.5}, {{9, 0.07}, {2.6, 0.3, 0.5}, {3.5, 0.6, 0.8}}};
		matter M[65];
		object O[65];
		int n = 0;
		for (int a = 0; a < 4; a++)
			for (int b = 0; b < 4; b++)
				for (int z = 0; z < 4; z++, n++) {
					const float x = -120 + 80 * a, y = -120 + 80 * b, h = -120 + 80 * z, r = 10 + 2 * (n % 4);
					M[n] = Gold;
					O[n] = (object){Ball, {{x - r, x + r}, {y - r, y + r}, {h - r, h + r}}};
				}
		M[n] = n(1.5);
		O[n] = (object){0};
		putObjectArray(c->W, M, O);
		dom DTF = {{-170, 170}, {-170, 170}, {-170, 170}};
		planewave(c->W, DTF, Ppol, 0, 0, Band, 400, 900);
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
	else if (!strcmp(name, "metasurface")) {
		const float dx = 5, half = dx * cells / 2;
		const int nz = cells / 2, d = nz - 2 * 17;            /* z nodes without the CPML */
		dom D = {{-half, half}, {-half, half}, {-dx * ((d - 2) / 2), dx * (d / 2)}};
		res R = {2.5, {dx}, {1240}};
		sur S = {{PBC}, {PBC}, {PML}, {16}};
		c->W = createWorld(D, R, S, "");
		/* one object per pillar (small bounding boxes keep the sub-cell averaging local),
		 * then the glass (a layer reaching outside the TF/SF box comes after the objects
		 * inside it, which the TF/SF background leaves out), then the air background */
		const int side = cells / 80, np = side * side;
		object *O = calloc(np + 2, sizeof(object));
		matter *M = calloc(np + 2, sizeof(matter));
		uint64_t seed = 2024;
		for (int a = 0; a < np; a++) {
			const float x = -half + 400 * (a % side) + 200, y = -half + 400 * (a / side) + 200;
			const float r = (float)(60 + 60 * uniform(&seed));
			M[a] = n(2.5), O[a] = (object) {RodZ, {{x - r, x + r}, {y - r, y + r}, {0, 600}}};
		}
		M[np] = n(1.5), O[np] = (object) {Box, {{-INF, INF}, {-INF, INF}, {-INF, 0}}};
		M[np+1] = Air;
		putObjectArray(c->W, M, O);
		free(M);
		c->parts = O;               /* the TF/SF background is cut from the same objects */
		dom DTF = {{-half, half}, {-half, half}, {-40, 640}};
		planewave(c->W, DTF, Ppol, 0, 0, Band, 400, 900);
		c->P = createPhaser(c->W, 550, -INF, INF, -INF, INF, -80, 670);
	}
	else if (!strcmp(name, "kpoints")) {
		uint64_t seed = 2024;
		c->nw = cells;
		c->Ws = calloc(cells, sizeof(world));
		for (int q = 0; q < cells; q++) c->Ws[q] = kpointWorld(q, cells, &seed);
		c->W = c->Ws[0];
		c->kprobe = calloc(cells, sizeof *c->kprobe);
		for (int q = 0; q < cells; q++)
			for (int p = 0; p < 8; p++) {
				c->kprobe[q][p][0] = (float)(uniform(&seed) * 600 - 300);
				c->kprobe[q][p][1] = (float)(uniform(&seed) * 600 - 300);
				c->kprobe[q][p][2] = (float)(uniform(&seed) * 200 - 100);
			}
	}
	else {
		free(c);
		return 0;
	}
	if (!c->nw) c->nw = 1, c->Ws = calloc(1, sizeof(world)), c->Ws[0] = c->W;
#ifndef ALIS_REF
	for (int q = 0; q < c->nw; q++) worldPrecision(c->Ws[q], fieldBits, dftBits);
#else
	(void)fieldBits, (void)dftBits;
#endif
	return c;
}

void wl_step(void *ctx, int n)
{
	wl *c = ctx;
	for (int s = 0; s < n; s++) {
		if (c->nw > 1) {
#ifndef ALIS_REF
			updateWorldsE(c->Ws, c->nw);
			updateWorldsH(c->Ws, c->nw);
#else
			for (int q = 0; q < c->nw; q++) updateE(c->Ws[q]), updateH(c->Ws[q]);
#endif
		}
		else {
			updateE(c->W);
			updateH(c->W);
		}
		if (c->P) updatePhaser(c->W, c->P);
		for (int p = 0; p < c->probes; p++) {
			keep(c, get(c->W, Ez, c->pts[p][0], c->pts[p][1], c->pts[p][2]));
			keep(c, get(c->W, Hy, c->pts[p][0], c->pts[p][1], c->pts[p][2]));
		}
		for (int q = 0; c->kprobe && q < c->nw; q++) {
			world W = c->Ws[q];
			float v[8];
			for (int p = 0; p < 8; p++) v[p] = keep(c, get(W, p % 2 ? Ez : Hz, c->kprobe[q][p][0], c->kprobe[q][p][1], c->kprobe[q][p][2]));
			writeSpectrum(W, 8000, 145.5, 2000, "kpoints-spectrum-%d", q, v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]);
		}
	}
}

/* Probe values read by the last wl_step calls (in order) into out[0..n-1]; returns how many
 * were read in total. */
long wl_trace(void *ctx, double *out, long n)
{
	wl *c = ctx;
	for (long a = 0; a < n && a < c->ntrace; a++) out[a] = c->trace[a];
	return c->ntrace;
}

#ifndef ALIS_REF
long alisSpectrumRead(world W, const char *name, double *out, long n);

/* Raw sums of world q's probe spectrum (kpoints): the source row, then the 8 probes, 256 bins
 * each as (re, im); returns the doubles written. */
long wl_spectrum(void *ctx, int q, double *out, long n)
{
	wl *c = ctx;
	char name[64];
	if (!c->kprobe || q < 0 || q >= c->nw) return 0;
	snprintf(name, sizeof name, "kpoints-spectrum-%d", q);
	return alisSpectrumRead(c->Ws[q], name, out, n);
}

/* One field array of world q as doubles over the ghost-inclusive grid (alis_dev_read_field):
 * h 0 E / 1 H, part m, axis a; returns the doubles written. */
long wl_field(void *ctx, int q, int h, int m, int a, double *out, long n)
{
	wl *c = ctx;
	if (q < 0 || q >= c->nw || !c->Ws[q]->dev) return 0;
	world W = c->Ws[q];
	long per = (long)(W->iNUM + 2) * (W->jNUM + 2) * (W->kNUM + 2);
	if (n < per || m > W->complexField) return 0;
	return alis_dev_read_field(W->dev, h, m, a, out) ? 0 : per;
}
#endif

/* Grid nodes of the stepped worlds including CPML (the cell count of one step; a complex
 * cell of a Bloch world counts once). */
long wl_cells(void *ctx)
{
	wl *c = ctx;
	long n = 0;
	for (int q = 0; q < c->nw; q++) n += (long)c->Ws[q]->iNUM * c->Ws[q]->jNUM * c->Ws[q]->kNUM;
	return n;
}

/* Worlds stepped together and whether they carry complex (Bloch) fields. */
int wl_worlds(void *ctx, int *complexField)
{
	wl *c = ctx;
	*complexField = c->W->complexField;
	return c->nw;
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

/* Net outward flux of every frequency of the workload's phaser (phaserPoyntingOut,
 * src/phaser.c:228-245) into out[0..n-1]; returns the frequencies written. */
int wl_flux(void *ctx, double *out, int n)
{
	wl *c = ctx;
	int m = 0;
	if (!c->P) return 0;
	while (m < n && c->P[m].f) out[m] = phaserPoyntingOut(c->P + m), m++;
	return m;
}

/* Grid nodes and CPML layer thickness (Num+1 cells) per face, 0 where the face is not PML;
 * returns 1 when the E update reads 1/eps (objects were put, src/update.c:109 vs :114). */
int wl_geometry(void *ctx, int *num, int *pml)
{
	wl *c = ctx;
	world W = c->W;
	num[0] = W->iNUM, num[1] = W->jNUM, num[2] = W->kNUM;
	pml[0] = W->xMinSur==PML ? W->iMin-W->iMIN : 0, pml[1] = W->xMaxSur==PML ? W->iMAX-W->iMax : 0;
	pml[2] = W->yMinSur==PML ? W->jMin-W->jMIN : 0, pml[3] = W->yMaxSur==PML ? W->jMAX-W->jMax : 0;
	pml[4] = W->zMinSur==PML ? W->kMin-W->kMIN : 0, pml[5] = W->zMaxSur==PML ? W->kMAX-W->kMax : 0;
	return W->CE->x != 0;
}

/* Sum of squares of every E and H component over the grid (sequential, in double), plus
 * the probe samples: equal for two runs exactly when their states agree to the bit (up to
 * cancellation that a state difference cannot plausibly produce). */
double wl_checksum(void *ctx)
{
	wl *c = ctx;
#ifndef ALIS_REF
#endif
	double s = c->probeSum;
	for (int q = 0; q < c->nw; q++) {
		world W = c->Ws[q];
#ifndef ALIS_REF
		syncWorld(W);
#endif
		for (int m = 0; m <= W->complexField; m++) {
			float ***F[6] = { W->E[m].x, W->E[m].y, W->E[m].z, W->H[m].x, W->H[m].y, W->H[m].z };
			for (int a = 0; a < 6; a++)
				for (int i = W->iMIN; i <= W->iMAX; i++)
					for (int j = W->jMIN; j <= W->jMAX; j++)
						for (int k = W->kMIN; k <= W->kMAX; k++)
							s += (double)F[a][i][j][k] * F[a][i][j][k];
		}
	}
	return s;
}

/* Far-field images of the workload's phaser over an X x Y cylindrical (theta, phi) map
 * (src/farfld.c:525-594), written as text files named <dir>/ff-*.txt, or with an empty dir
 * evaluated without writing files (their maximum enters the checksum). */
#ifndef ALIS_REF
/* An image writer that keeps the maps' maximum only: the map is timed without its files,
 * like the reference's evaluations below. */
static double mapMax;
static void keepMax(world W, slice S, float Max, char *F, char *name, unsigned long int *cmap)
{
	(void)W, (void)S, (void)F, (void)name, (void)cmap;
	mapMax = Max;
}
#endif

void wl_farfield(void *ctx, int X, int Y, const char *dir)
{
	wl *c = ctx;
	if (!c->P) return;
#ifndef ALIS_REF
	if (dir && dir[0]) farFieldProfile(c->W, c->P, CylindericalMap, X, Y, txt, "%s/ff-", dir);
	else {
		farFieldProfile(c->W, c->P, CylindericalMap, X, Y, keepMax, "ff-");
		c->probeSum += mapMax;
	}
#else
	/* the reference build has no image writers (slices.c needs libpng): the same two
	 * evaluations per direction that its farFieldProfile makes (src/farfld.c:553-556) */
	(void)dir;
	double sum = 0;
	farField(c->P, Ppol, 0, 0);     /* sets up the coefficients once, before the threads read them */
	#pragma omp parallel for schedule(dynamic) reduction(+:sum)
	for (long xy = 0; xy < (long)X * Y; xy++) {
		float ang[2];
		CylindericalMap(xy / Y, xy % Y, X, Y, ang);
		float complex p = farField(c->P, Ppol, ang[0], ang[1]), s = farField(c->P, Spol, ang[0], ang[1]);
		sum += crealf(p * conjf(p)) + crealf(s * conjf(s));
	}
	c->probeSum += sum;
#endif
}

void wl_free(void *ctx)
{
	wl *c = ctx;
	if (c->P) deletePhaser(c->P);
	for (int q = 0; q < c->nw; q++) deleteWorld(c->Ws[q]);
	free(c->Ws);
	free(c->parts);
	free(c->trace);
	free(c->kprobe);
	free(c);
}

#ifndef ALIS_REF
/* Device world of a committed workload (after its first step), for kernel timing. */
void *wl_device(void *ctx)
{
	return ((wl *)ctx)->W->dev;
}

/* Device world of the q-th world stepped together. */
void *wl_device_of(void *ctx, int q)
{
	wl *c = ctx;
	return q >= 0 && q < c->nw ? c->Ws[q]->dev : 0;
}
#endif

#ifdef WL_MAIN
/* CPU timing program: wl NAME CELLS WARMUP STEPS -> "seconds_per_step cells threads". */
#include <omp.h>
#include <unistd.h>
int main(int argc, char **argv)
{
	if (argc < 5) return fprintf(stderr, "usage: %s NAME CELLS WARMUP STEPS [X Y DIR]\n", argv[0]), 2;
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
	double ff = 0;
	if (argc >= 8) {
		/* far-field map of X x Y directions after the steps: seconds per direction */
		const int X = atoi(argv[5]), Y = atoi(argv[6]);
		clock_gettime(CLOCK_MONOTONIC, &a);
		wl_farfield(c, X, Y, argv[7]);
		clock_gettime(CLOCK_MONOTONIC, &b);
		ff = ((b.tv_sec - a.tv_sec) + 1e-9 * (b.tv_nsec - a.tv_nsec)) / ((double)X * Y);
	}
	fprintf(keep, "%.9g %ld %d %.17g %.9g\n", sec / (steps ? steps : 1), wl_cells(c), omp_get_max_threads(), wl_checksum(c), ff);
	fclose(keep);
	return 0;
}
#endif
