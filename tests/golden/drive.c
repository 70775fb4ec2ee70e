/*
 * Scenario driver for parity tests.  Written against the public alis.h API only, so the same
 * file compiles against the reference library (-DALIS_REF, built from the reference sources
 * by oracle/Makefile) and against this engine.  Each scenario sets up a small world, steps
 * it, and prints "key value" lines: FNV-1a hashes and sums of squares of setup arrays and of
 * the final state, probe traces, phaser accumulators, fluxes and far-field samples.
 *
 * usage (program): drive SCENARIO [STEPS [OUTPATH]]   or   drive --text SCENARIO STEPS DIR
 * usage (library): alis_drive(SCENARIO, STEPS, OUTPATH) writes the same lines to OUTPATH.
 */
#include "alis.h"
#include <stdint.h>
#include <string.h>

static FILE *OUT;
static world LAST;      /* the (first) scenario's world and phaser array, for the raw dumps */
static phaser LASTP;
#define MAXSCEN 64
static world RUNW[MAXSCEN];   /* everything a run created, freed by releaseRun */
static phaser RUNP[MAXSCEN];
static world RUNPUMP[MAXSCEN];  /* their pump worlds */
static char OUTDIR[512] = ".";   /* where scenarios write their spectrum files */
static int RUNSTEPS;            /* steps of the current run */
static const char *RUNNAME[MAXSCEN];
static int NRUN;

/* Free the worlds of the last run.  Against the reference library they are left alive: its
 * deletePhaser passes the array's end sentinel, not its start, to free() and aborts
 * (src/phaser.c:135-138). */
static void releaseRun(void)
{
#ifndef ALIS_REF
	for (int a = 0; a < NRUN; a++) {
		if (RUNP[a]) deletePhaser(RUNP[a]);
		deleteWorld(RUNW[a]);
		if (RUNPUMP[a]) deleteWorld(RUNPUMP[a]);
	}
#endif
	NRUN = 0;
	LAST = 0, LASTP = 0;
}

static uint64_t fnv(uint64_t h, const void *p, size_t n)
{
	const unsigned char *b = p;
	for (size_t a = 0; a < n; a++) h = (h ^ b[a]) * 1099511628211ULL;
	return h;
}

/* Hash and sum of squares of a host array block (float or complex float). */
static void putBlock(const char *key, const void *p, long n, int cplx)
{
	if (!p) return;
	const float *f = p;
	long m = cplx ? 2*n : n;
	double s = 0;
	for (long a = 0; a < m; a++) s += (double)f[a] * f[a];
	fprintf(OUT, "%s %016llx %.17g\n", key, (unsigned long long)fnv(1469598103934665603ULL, p, m*sizeof(float)), s);
}

static void put3(const char *key, float ***F, int i0, int i1, int j0, int j1, int k0, int k1)
{
	if (!F) return;
	putBlock(key, &F[i0][j0][k0], (long)(i1-i0+1)*(j1-j0+1)*(k1-k0+1), 0);
}

static void put1(const char *key, float *a, int i0, int i1)
{
	if (!a) return;
	putBlock(key, &a[i0], i1-i0+1, 0);
}

static void sync(world W)
{
#ifndef ALIS_REF
	syncWorld(W);
#else
	(void)W;
#endif
}

/* Setup arrays: grid coefficients, CPML profiles, 1/eps, ADE tables, source amplitudes. */
static void dumpSetup(const char *tag, world W)
{
	char key[256];
	fprintf(OUT, "%s.grid %d %d %d %d %d %d %d %d %d %d %d %d %d\n", tag, W->iMIN, W->iMAX, W->jMIN, W->jMAX, W->kMIN, W->kMAX,
		W->iMin, W->iMax, W->jMin, W->jMax, W->kMin, W->kMax, W->complexField);
	fprintf(OUT, "%s.phase %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g\n", tag, W->coskx, W->sinkx, W->cosky, W->sinky,
		W->coskz, W->sinkz, W->coskyp, W->coskym, W->sinkyp, W->sinkym);
	fprintf(OUT, "%s.clock %lu %lu %.9g %.9g\n", tag, W->N, W->T, W->f, W->dt);
	for (int h = 0; h < 2; h++) {
		coeffs *C = h ? W->CH : W->CE;
		const char *c = h ? "H" : "E";
		snprintf(key, sizeof key, "%s.dt%s.x", tag, c); put1(key, C->dtdx, W->iMIN, W->iMAX);
		snprintf(key, sizeof key, "%s.dt%s.y", tag, c); put1(key, C->dtdy, W->jMIN, W->jMAX);
		snprintf(key, sizeof key, "%s.dt%s.z", tag, c); put1(key, C->dtdz, W->kMIN, W->kMAX);
		snprintf(key, sizeof key, "%s.cpml%s.x", tag, c); if (C->xP) put1(key, C->xP, W->iMIN, W->iMAX), put1(key, C->xPy, W->iMIN, W->iMAX), put1(key, C->xPz, W->iMIN, W->iMAX);
		snprintf(key, sizeof key, "%s.cpml%s.y", tag, c); if (C->yP) put1(key, C->yP, W->jMIN, W->jMAX), put1(key, C->yPz, W->jMIN, W->jMAX), put1(key, C->yPx, W->jMIN, W->jMAX);
		snprintf(key, sizeof key, "%s.cpml%s.z", tag, c); if (C->zP) put1(key, C->zP, W->kMIN, W->kMAX), put1(key, C->zPx, W->kMIN, W->kMAX), put1(key, C->zPy, W->kMIN, W->kMAX);
	}
	coeffs *C = W->CE;
	snprintf(key, sizeof key, "%s.eps.x", tag); put3(key, C->x, W->iMIN, W->iMAX+1, W->jMIN, W->jMAX, W->kMIN, W->kMAX);
	snprintf(key, sizeof key, "%s.eps.y", tag); put3(key, C->y, W->iMIN, W->iMAX, W->jMIN, W->jMAX+1, W->kMIN, W->kMAX);
	snprintf(key, sizeof key, "%s.eps.z", tag); put3(key, C->z, W->iMIN, W->iMAX, W->jMIN, W->jMAX, W->kMIN, W->kMAX+1);
	for (int n = 0; n < C->N; n++) {
		fprintf(OUT, "%s.obj%d %d %d %d %d %d %d %d %d %.9g %.9g %.9g\n", tag, n, C->iMin[n], C->iMax[n], C->jMin[n], C->jMax[n],
			C->kMin[n], C->kMax[n], C->NJ[n], C->NK[n], C->ax[n], C->ay[n], C->az[n]);
		for (int p = 0; p < C->NK[n]; p++) {
			fprintf(OUT, "%s.obj%d.pole%d %.9g %.9g %.9g %.9g %.9g %.9g\n", tag, n, p, C->JJx[n][p], C->JJy[n][p], C->JJz[n][p], C->JEx[n][p], C->JEy[n][p], C->JEz[n][p]);
			if (p >= C->NJ[n]) fprintf(OUT, "%s.obj%d.pair%d %.9g %.9g %.9g %.9g %.9g %.9g\n", tag, n, p, C->JKx[n][p], C->KJx[n][p], C->KKx[n][p], C->KEx[n][p], C->KEy[n][p], C->KEz[n][p]);
		}
		snprintf(key, sizeof key, "%s.obj%d.fill.x", tag, n); put3(key, C->Jx[n], C->iMin[n], C->iMax[n], C->jMin[n], C->jMax[n], C->kMin[n], C->kMax[n]);
		snprintf(key, sizeof key, "%s.obj%d.fill.y", tag, n); put3(key, C->Jy[n], C->iMin[n], C->iMax[n], C->jMin[n], C->jMax[n], C->kMin[n], C->kMax[n]);
		snprintf(key, sizeof key, "%s.obj%d.fill.z", tag, n); put3(key, C->Jz[n], C->iMin[n], C->iMax[n], C->jMin[n], C->jMax[n], C->kMin[n], C->kMax[n]);
	}
	int s = 0;
	for (int h = 0; h < 2; h++)
		for (source S = h ? W->SH : W->SE; S; S = S->S, s++) {
			fprintf(OUT, "%s.src%d %d %d %d %d %d %d %lu %.9g %.9g %.9g %.9g %.9g %.9g\n", tag, s, S->iMin, S->iMax, S->jMin, S->jMax, S->kMin, S->kMax,
				S->N, S->amplitude, S->phase, S->frequency, S->deviation, S->peaktime, S->resolution);
			snprintf(key, sizeof key, "%s.src%d.amp", tag, s); put3(key, S->Amplitude, S->iMin, S->iMax, S->jMin, S->jMax, S->kMin, S->kMax);
			snprintf(key, sizeof key, "%s.src%d.phase", tag, s); put3(key, S->Phase, S->iMin, S->iMax, S->jMin, S->jMax, S->kMin, S->kMax);
		}
	int t = 0;
	for (tfsf T = W->S; T; T = T->S, t++) {
		fprintf(OUT, "%s.tfsf%d %d %d %d %d %d %d %d %d %d %d %d %d\n", tag, t, T->iMin, T->iMax, T->jMin, T->jMax, T->kMin, T->kMax,
			T->iMIN, T->iMAX, T->jMIN, T->jMAX, T->kMIN, T->kMAX);
		char sub[64];
		snprintf(sub, sizeof sub, "%s.aux%d", tag, t);
		dumpSetup(sub, T->W);
	}
}

/* Final state: fields with ghosts, CPML slabs, ADE currents; TF/SF auxiliary worlds too. */
static void dumpState(const char *tag, world W)
{
	char key[256];
	sync(W);
	const char *an = "xyz";
	for (int h = 0; h < 2; h++) {
		vfield *V = h ? W->H : W->E;
		for (int m = 0; m <= W->complexField; m++) {
			float ***c[3] = { V[m].x, V[m].y, V[m].z };
			for (int a = 0; a < 3; a++) {
				snprintf(key, sizeof key, "%s.%s%d.%c", tag, h ? "H" : "E", m, an[a]);
				putBlock(key, &c[a][W->iMIN-1][W->jMIN-1][W->kMIN-1], (long)(W->iNUM+2)*(W->jNUM+2)*(W->kNUM+2), 0);
			}
			int Off = h;
			float ***sl[12] = { V[m].xMinPy, V[m].xMinPz, V[m].yMinPz, V[m].yMinPx, V[m].zMinPx, V[m].zMinPy,
			                    V[m].xMaxPy, V[m].xMaxPz, V[m].yMaxPz, V[m].yMaxPx, V[m].zMaxPx, V[m].zMaxPy };
			for (int f = 0; f < 12; f++) {
				if (!sl[f]) continue;
				int face = f / 2, lo, hi;
				int i0 = W->iMIN, i1 = W->iMAX, j0 = W->jMIN, j1 = W->jMAX, k0 = W->kMIN, k1 = W->kMAX;
				if (face < 3) lo = (face==0 ? W->iMIN : face==1 ? W->jMIN : W->kMIN) + Off, hi = (face==0 ? W->iMin : face==1 ? W->jMin : W->kMin) + Off - 2;
				else lo = (face==3 ? W->iMax : face==4 ? W->jMax : W->kMax) + 2, hi = face==3 ? W->iMAX : face==4 ? W->jMAX : W->kMAX;
				if (face % 3 == 0) i0 = lo, i1 = hi;
				if (face % 3 == 1) j0 = lo, j1 = hi;
				if (face % 3 == 2) k0 = lo, k1 = hi;
				snprintf(key, sizeof key, "%s.%s%d.psi%d", tag, h ? "H" : "E", m, f);
				put3(key, sl[f], i0, i1, j0, j1, k0, k1);
			}
		}
	}
	coeffs *C = W->CE;
	for (int n = 0; n < C->N; n++)
		for (int m = 0; m <= W->complexField; m++)
			for (int p = 0; p < C->NK[n]; p++) {
				float ****J[3] = { W->E[m].Jx[n], W->E[m].Jy[n], W->E[m].Jz[n] };
				for (int a = 0; a < 3; a++) {
					snprintf(key, sizeof key, "%s.J%d.%d.%d.%c", tag, n, m, p, an[a]);
					put3(key, J[a][p], C->iMin[n], C->iMax[n], C->jMin[n], C->jMax[n], C->kMin[n], C->kMax[n]);
				}
				if (p >= C->NJ[n]) {
					float ****K[3] = { W->E[m].Kx[n], W->E[m].Ky[n], W->E[m].Kz[n] };
					for (int a = 0; a < 3; a++) {
						snprintf(key, sizeof key, "%s.K%d.%d.%d.%c", tag, n, m, p, an[a]);
						put3(key, K[a][p], C->iMin[n], C->iMax[n], C->jMin[n], C->jMax[n], C->kMin[n], C->kMax[n]);
					}
				}
			}
	int t = 0;
	for (tfsf T = W->S; T; T = T->S, t++) {
		char sub[64];
		snprintf(sub, sizeof sub, "%s.aux%d", tag, t);
		dumpState(sub, T->W);
	}
	fprintf(OUT, "%s.time %lu %.9g %.9g %.9g\n", tag, W->n, W->t, W->E->t, W->H->t);
}

static void dumpPhasers(const char *tag, phaser P, int farfield)
{
	char key[256];
	for (int q = 0; P[q].f; q++) {
		phaser Q = &P[q];
		fprintf(OUT, "%s.ph%d %.9g %.9g %lu %d %d %d %d %d %d\n", tag, q, Q->f, Q->A, Q->N, Q->iMin, Q->iMax, Q->jMin, Q->jMax, Q->kMin, Q->kMax);
		cfield *F[6] = { &Q->xMin, &Q->xMax, &Q->yMin, &Q->yMax, &Q->zMin, &Q->zMax };
		const char *fn[6] = { "xMin", "xMax", "yMin", "yMax", "zMin", "zMax" };
		for (int f = 0; f < 6; f++) {
			float complex ***c[6] = { F[f]->Ex, F[f]->Ey, F[f]->Ez, F[f]->Hx, F[f]->Hy, F[f]->Hz };
			const char *cn[6] = { "Ex", "Ey", "Ez", "Hx", "Hy", "Hz" };
			int i = f == 1 ? Q->iMax : Q->iMin, j = f == 3 ? Q->jMax : Q->jMin, k = f == 5 ? Q->kMax : Q->kMin;
			int ni = f < 2 ? 1 : Q->iMax-Q->iMin+1, nj = (f == 2 || f == 3) ? 1 : Q->jMax-Q->jMin+1, nk = f >= 4 ? 1 : Q->kMax-Q->kMin+1;
			for (int a = 0; a < 6; a++) {
				if (!c[a]) continue;
				snprintf(key, sizeof key, "%s.ph%d.%s.%s", tag, q, fn[f], cn[a]);
				putBlock(key, &c[a][i][j][k], (long)ni*nj*nk, 1);
			}
		}
		fprintf(OUT, "%s.ph%d.flux %.9g\n", tag, q, phaserPoyntingOut(Q));
		if (farfield) {
			float th[5] = { 0, 20, 45, 130, 180 }, ph[5] = { 0, 30, 0, 75, 0 };
			for (int a = 0; a < 5; a++) {
				float complex p = farField(Q, Ppol, th[a], ph[a]), s = farField(Q, Spol, th[a], ph[a]);
				fprintf(OUT, "%s.ph%d.ff%d %.9g %.9g %.9g %.9g\n", tag, q, a, crealf(p), cimagf(p), crealf(s), cimagf(s));
			}
		}
	}
}

/* Probe trace: hash of every sampled value plus the last sample of each probe. */
typedef struct { uint64_t h; double s; float last[16]; int n; } trace;

static void sample(trace *T, world W, field *F[], float pts[][3], int n)
{
	for (int a = 0; a < n; a++) {
		float v = get(W, F[a], pts[a][0], pts[a][1], pts[a][2]);
		T->h = fnv(T->h ? T->h : 1469598103934665603ULL, &v, sizeof v);
		T->s += (double)v * v;
		T->last[a] = v;
	}
	T->n = n;
}

static void putTrace(const char *tag, trace *T)
{
	fprintf(OUT, "%s.trace %016llx %.17g", tag, (unsigned long long)T->h, T->s);
	for (int a = 0; a < T->n; a++) fprintf(OUT, " %.9g", T->last[a]);
	fprintf(OUT, "\n");
}

/* ---- scenarios ---- */

/* A scenario: its world, phaser array (or none) and probes. */
typedef struct {
	const char *name;
	world W;
	phaser P;
	int np;
	field *F[8];
	float pts[8][3];
	trace T;
	int kq;                     /* k-point scenarios: their index, whose spectrum they write */
	float spec[3][3];           /* and the nodes of their spectrum's probes */
	object absorbers[4];        /* dispersive objects whose absorbed power is recorded */
	int nabs;
	double poyntingRun;         /* poyntingOut of the observer box every 25 steps, summed */
	world pump;                 /* a pump world coupled into W by nonLinearSHG, or 0 */
	float alt;
} scen;
static void dipoleSetup(scen *c)
{
	dom D = {{60}, {60}, {60}};
	res R = {5, {10}, {1240}};
	sur S = {{PML}, {PML}, {PML}, {8}};
	world W = createWorld(D, R, S, "");
	pointDipole(W, Ez, 0, 0, 0, Pulse, 600, 300);
	pointDipole(W, Hx, 20, -10, 30, Sine, 500, 2, 0.5);
	field *F[6] = { Ez, Ez, Hy, Ex, Hz, Ey };
	float pts[6][3] = {{0,0,0}, {30,-20,10}, {10,10,-40}, {-50,40,0}, {20,20,20}, {-30,0,50}};
	c->W = W, c->P = 0, c->np = 6;
	memcpy(c->F, F, sizeof F);
	memcpy(c->pts, pts, sizeof pts);
}
static void symbandSetup(scen *c)
{
	dom D = {{150}, {150}, {150}};
	res R = {10, {20}, {1240}, {{2, 60}, {2, 60}, {2, 60}}};
	sur S = {{SYM, PML}, {SYM, PML}, {SYM, PML}, {8}};
	world W = createWorld(D, R, S, "");
	pointDipole(W, Ex, 0, 0, 0, Band, 400, 800);
	phaser P = createPhasers(W, 400, 800, 100);
	field *F[4] = { Ex, Ex, Hy, Ez };
	float pts[4][3] = {{0,0,100}, {-40,60,0}, {20,-20,80}, {100,0,0}};
	c->W = W, c->P = P, c->np = 4;
	memcpy(c->F, F, sizeof F);
	memcpy(c->pts, pts, sizeof pts);
}
static void nanoblockSetup(scen *c)
{
	dom DTF = {{36}, {36}, {36}};
	dom DSF = {{48}, {48}, {48}};
	res R = {2, {4}, {1240}};
	sur S = {{PML}, {PML}, {PML}, {8}};
	world W = createWorld(DSF, R, S, "");
	matter Gold = {{1.5}, {{9, 0.07}, {2.6, 0.3, 0.5}, {3.5, 0.6, 0.8}}};
	object Block = {Box, {{-16, 16}, {-16, 16}, {-8, 8}}};
	object Substrate = {Box, {{-INF, INF}, {-INF, INF}, {-INF, -8}}};
	putObjects(W, Gold, Block, n(1.5), Substrate, Air);
	planewave(W, DTF, Ppol, 0, 0, Pulse, 600, 400);
	phaser P = createPhasers(W, 500, 700, 100, -42, 42, -42, 42, -42, 42);
	c->absorbers[c->nabs++] = Block;
	field *F[5] = { Ex, Ex, Hy, Ez, Ey };
	float pts[5][3] = {{0,0,12}, {0,0,0}, {10,-6,20}, {-20,4,-20}, {40,40,40}};
	c->W = W, c->P = P, c->np = 5;
	memcpy(c->F, F, sizeof F);
	memcpy(c->pts, pts, sizeof pts);
}
/* 18 metal spheres (radii 6-10 nm) on a 3 x 3 x 2 lattice of pitch 24 x 24 x 20 nm,
 * alternately a Drude + 2-Lorentz gold and a Drude silver, in glass, lit by a TF/SF plane
 * wave: many dispersive objects, each cell inside at most one of them. */
static void particlesSetup(scen *c)
{
	dom DTF = {{40}, {40}, {24}};
	dom DSF = {{52}, {52}, {36}};
	res R = {2, {4}, {1240}};
	sur S = {{PML}, {PML}, {PML}, {8}};
	world W = createWorld(DSF, R, S, "");
	matter Gold = {{1.5}, {{9, 0.07}, {2.6, 0.3, 0.5}, {3.5, 0.6, 0.8}}};
	matter Silver = {{1.2}, {{8.9, 0.04}}};
	matter M[19];
	object O[19];
	int n = 0;
	for (int a = 0; a < 3; a++)
		for (int b = 0; b < 3; b++)
			for (int z = 0; z < 2; z++, n++) {
				const float x = -24 + 24 * a, y = -24 + 24 * b, h = -10 + 20 * z, r = 6 + 2 * (n % 3);
				M[n] = n % 2 ? Silver : Gold;
				O[n] = (object){Ball, {{x - r, x + r}, {y - r, y + r}, {h - r, h + r}}};
			}
	M[n] = n(1.5);
	O[n] = (object){0};
	putObjectArray(W, M, O);
	planewave(W, DTF, Ppol, 0, 0, Pulse, 600, 400);
	phaser P = createPhasers(W, 500, 700, 100, -46, 46, -46, 46, -30, 30);
	c->absorbers[c->nabs++] = O[4];
	c->absorbers[c->nabs++] = O[13];
	field *F[4] = { Ex, Ez, Hy, Ex };
	float pts[4][3] = {{0,0,10}, {24,-24,-10}, {12,12,0}, {-30,20,25}};
	c->W = W, c->P = P, c->np = 4;
	memcpy(c->F, F, sizeof F);
	memcpy(c->pts, pts, sizeof pts);
}
/* Second-harmonic generation: a pump world with a 900 nm Ex dipole and a signal world on the
 * same grid, both holding a slab of a nonlinear crystal (a Drude pole and a full d tensor);
 * every step couples the pump's E into the signal (nonLinearSHG), whose only field is what
 * the coupling puts there. */
static void shgSetup(scen *c)
{
	dom D = {{40}, {40}, {40}};
	res R = {2, {4}, {1240}};
	sur S = {{PML}, {PML}, {PML}, {8}};
	matter Crystal = {{4.8}, {{2.0, 0.1}},
	                  {{0.004, 0.001, 0.002, 0.0005, 0.003, -0.002}, {-0.002, 0.004, 0.001, 0.003, 0.0007, 0.001},
	                   {0.003, 0.003, 0.008, -0.001, 0.002, 0.0015}}};
	object Slab = {Box, {{-24, 24}, {-20, 28}, {-12, 16}}};
	world WP = createWorld(D, R, S, "");
	putObjects(WP, Crystal, Slab, Air);
	pointDipole(WP, Ex, 2, -2, 4, Pulse, 900, 300);
	world WS = createWorld(D, R, S, "");
	putObjects(WS, Crystal, Slab, Air);
	c->W = WS, c->pump = WP, c->alt = 1;
	c->absorbers[c->nabs++] = Slab;
	field *F[4] = { Ex, Ey, Ez, Hy };
	float pts[4][3] = {{0,0,0}, {8,-4,8}, {-12,12,-4}, {20,0,8}};
	c->np = 4;
	memcpy(c->F, F, sizeof F);
	memcpy(c->pts, pts, sizeof pts);
}
static void obliqueSetup(scen *c)
{
	dom DTF = {{50}, {50}, {30}};
	dom DSF = {{50}, {50}, {50}};
	res R = {5, {10}, {1240}};
	sur S = {{BBC}, {SYM}, {PML}, {8}};
	world W = createWorld(DSF, R, S, "");
	object Block = {Lattice, {{100}, {100}}, objects {{Box, {{0, 30}, {-20, 20}, {-20, 20}}}}};
	object Substrate = {Box, {{-INF, INF}, {-INF, INF}, {-INF, -20}}};
	putObjects(W, n(2), Block, n(1.5), Substrate, Air);
	planewave(W, DTF, Ppol, 35, 0, Sine, 530, 3);
	phaser P = createPhaser(W, 530, -300, 300, -300, 300);
	field *F[4] = { Ex, iEx, Hy, iHz };
	float pts[4][3] = {{0,0,40}, {20,10,0}, {-30,0,-40}, {10,-20,20}};
	c->W = W, c->P = P, c->np = 4;
	memcpy(c->F, F, sizeof F);
	memcpy(c->pts, pts, sizeof pts);
}
/* A periodic cell of 2 x 2 dielectric pillars of different radii on glass under a
 * normal-incidence plane wave (the metasurface geometry at small size): one object per
 * pillar through putObjectArray, the glass after them. */
static object pillarObjects[6];
static matter pillarMatter[6];
static void pillarsSetup(scen *c)
{
	dom D = {{-200, 200}, {-200, 200}, {-160, 400}};
	res R = {5, {10}, {1240}};
	sur S = {{PBC}, {PBC}, {PML}, {8}};
	world W = createWorld(D, R, S, "");
	const float r[4] = {32, 47, 55, 38};
	for (int a = 0; a < 4; a++) {
		float x = -100 + 200 * (a % 2), y = -100 + 200 * (a / 2);
		pillarMatter[a] = n(2.5);
		pillarObjects[a] = (object) {RodZ, {{x - r[a], x + r[a]}, {y - r[a], y + r[a]}, {0, 240}}};
	}
	pillarMatter[4] = n(1.5), pillarObjects[4] = (object) {Box, {{-INF, INF}, {-INF, INF}, {-INF, 0}}};
	pillarMatter[5] = Air, pillarObjects[5] = (object) {0};
	putObjectArray(W, pillarMatter, pillarObjects);
	dom DTF = {{-200, 200}, {-200, 200}, {-40, 300}};
	planewave(W, DTF, Ppol, 0, 0, Band, 400, 900);
	phaser P = createPhasers(W, 500, 700, 100, -INF, INF, -INF, INF, -80, 340);
	field *F[4] = { Ex, Ey, Hy, Ez };
	float pts[4][3] = {{0,0,100}, {-100,-100,120}, {60,-30,-60}, {100,100,320}};
	c->W = W, c->P = P, c->np = 4;
	memcpy(c->F, F, sizeof F);
	memcpy(c->pts, pts, sizeof pts);
}
static void periodicSetup(scen *c)
{
	dom D = {{40}, {40}, {-40, 80}};
	res R = {5, {10}, {1240}};
	sur S = {{SYM, PBC}, {SYM, PBC}, {PEC, PML}, {8}};
	world W = createWorld(D, R, S, "");
	matter InGaN = nk(2.46, 0.045, 1240/450);
	object Slab = {Box, {{-INF, INF}, {-INF, INF}, {0, 40}}};
	object Mirror = {Box, {{-INF, INF}, {-INF, INF}, {-INF, 0}}};
	object Layer = {Box, {{-INF, INF}, {-INF, INF}, {10, 20}}};
	putObjects(W, InGaN, Layer, n(2.46), Slab, Ag, Mirror, n(1.4));
	pointDipole(W, Ex, 0, 0, 20, Pulse, 450, 100);
	field *F[4] = { Ex, Ey, Hy, Hz };
	float pts[4][3] = {{0,0,20}, {20,10,40}, {-10,30,60}, {30,-30,5}};
	c->W = W, c->P = 0, c->np = 4;
	memcpy(c->F, F, sizeof F);
	memcpy(c->pts, pts, sizeof pts);
}
static void hexagonalSetup(scen *c)
{
	float kx = 0.2, ky = 1;
	dom D = {{50}, {0, 50*sqrtf(3)}, {60}};
	res R = {5, {10, 10*sqrtf(3)/2, 10}, {1240}};
	sur S = {{BBC, 100*3/kx}, {HBC, 100*sqrtf(3)/ky}, {SYM, PML}, {8}};
	world W = createWorld(D, R, S, "");
	object Slab = {Difference, {2}, objects {
		{Box, {{-INF, INF}, {-INF, INF}, {-20, 20}}},
		{DLattice, {{100}, {100*sqrtf(3)}}, objects {{RodZ, {{-30, 30}, {-30, 30}, {-INF, INF}}}}},
	}};
	putObjects(W, n(3.5), Slab, Air);
	pointDipole(W, iHz, 20, 10*sqrtf(3), 0, Band, 750, 2000);
	field *F[3] = { Hz, iHz, Ex };
	float pts[3][3] = {{20, 0, 0}, {-20, 30, 10}, {0, 60, 30}};
	c->W = W, c->P = 0, c->np = 3;
	memcpy(c->F, F, sizeof F);
	memcpy(c->pts, pts, sizeof pts);
}
static void multipoleSetup(scen *c)
{
	dom D = {{50}, {50}, {50}};
	res R = {2.5, {5}, {1240}};
	sur S = {{PML}, {PML}, {PML}, {8}};
	world W = createWorld(D, R, S, "");
	object A = {Ball, {{-30, 0}, {-20, 10}, {-15, 15}}};
	object B = {Box, {{-5, 25}, {-10, 20}, {-10, 10}}};
	object C = {RodY, {{10, 30}, {-40, 40}, {-25, -5}}};
	object Dd = {Diamond, {{-40, -10}, {10, 40}, {0, 30}}};
	matter Aniso = {{2, 3, 4}, {{5, 0.1, 1, 0}, {6, 0.2, 1.2, 0}, {4, 0.3, 0.8, 0}}};
	putObjects(W, Au, A, Ag, B, Si, C, Aniso, Dd, Air);
	c->absorbers[c->nabs++] = A, c->absorbers[c->nabs++] = B, c->absorbers[c->nabs++] = Dd;
	pointDipole(W, Ey, 0, 0, 0, Pulse, 500, 300);
	pointDipole(W, Ez, 10, -10, 5, Pulse, 550, 200, 2, 45);
	field *F[4] = { Ey, Ez, Hx, Jy };
	float pts[4][3] = {{0,0,0}, {10,-10,5}, {-20,20,10}, {15,5,0}};
	c->W = W, c->P = 0, c->np = 4;
	memcpy(c->F, F, sizeof F);
	memcpy(c->pts, pts, sizeof pts);
}
/* The q-th of NKP k-points along Gamma-X-M of a small photonic-crystal slab (the band
 * sweep at small size): a cell of period 2a = 160 nm (a = 80 nm, dx = 10 nm) with
 * Bloch-periodic x and y, an eps = 12 slab 40 nm thick with air holes of radius 0.3a on the
 * square lattice of period a, 8-cell CPML above and below, complex fields.  Two point
 * dipoles (Pulse 300 nm) at positions and phases drawn from a splitmix64 stream seeded by
 * q; three probes feed writeSpectrum over 100-800 nm, written on the run's last step. */
#define NKP 40
static uint64_t splitmix(uint64_t *s)
{
	uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
	return z ^ (z >> 31);
}
static float unitDraw(uint64_t *s) { return (float)((splitmix(s) >> 11) * (1.0 / 9007199254740992.0)); }
static void kpointSetup(scen *c, int q)
{
	const float a = 80, L = 2 * a;
	float t = (q + 0.5f) / NKP * 2, kx, ky;
	if (t < 1) kx = 0.5f * t, ky = 0;
	else kx = 0.5f, ky = 0.5f * (t - 1);
	dom D = {{-L/2, L/2}, {-L/2, L/2}, {-60, 60}};
	res R = {5, {10}, {1240}};
	sur S = {{BBC, a/kx}, {ky > 0 ? BBC : PBC, ky > 0 ? a/ky : INF}, {PML}, {8}};
	world W = createWorld(D, R, S, "");
	object Slab = {Difference, {2}, objects {
		{Box, {{-INF, INF}, {-INF, INF}, {-20, 20}}},
		{Lattice, {{a}, {a}}, objects {{RodZ, {{-24, 24}, {-24, 24}, {-INF, INF}}}}},
	}};
	putObjects(W, n(sqrtf(12)), Slab, Air);
	uint64_t seed = 2024 + q;
	for (int d = 0; d < 2; d++) {
		float x = unitDraw(&seed) * 120 - 60, y = unitDraw(&seed) * 120 - 60, z = unitDraw(&seed) * 30 - 15;
		pointDipole(W, d ? Ez : Hz, x, y, z, Pulse, 300, 200, 1, unitDraw(&seed) * 360);
	}
	field *F[3] = { Hz, iEz, Ex };
	float pts[3][3] = {{20, -30, 0}, {-40, 10, 30}, {50, 50, -10}};
	float spec[3][3] = {{30, 30, 0}, {-30, 0, 10}, {0, -50, -30}};
	c->W = W, c->P = 0, c->np = 3, c->kq = q;
	memcpy(c->F, F, sizeof F);
	memcpy(c->pts, pts, sizeof pts);
	memcpy(c->spec, spec, sizeof spec);
}
static int setupScenario(scen *c, const char *name)
{
	memset(c, 0, sizeof *c);
	c->name = name;
	c->kq = -1;
	if (!strncmp(name, "kp", 2) && name[2] >= '0' && name[2] <= '9') {
		int q = atoi(name + 2);
		if (q < 0 || q >= NKP) return fprintf(stderr, "k-point %d outside 0..%d\n", q, NKP - 1), 2;
		return kpointSetup(c, q), 0;
	}
	if (!strcmp(name, "dipole")) return dipoleSetup(c), 0;
	if (!strcmp(name, "symband")) return symbandSetup(c), 0;
	if (!strcmp(name, "nanoblock")) return nanoblockSetup(c), 0;
	if (!strcmp(name, "oblique")) return obliqueSetup(c), 0;
	if (!strcmp(name, "periodic")) return periodicSetup(c), 0;
	if (!strcmp(name, "pillars")) return pillarsSetup(c), 0;
	if (!strcmp(name, "hexagonal")) return hexagonalSetup(c), 0;
	if (!strcmp(name, "multipole")) return multipoleSetup(c), 0;
	if (!strcmp(name, "particles")) return particlesSetup(c), 0;
	if (!strcmp(name, "shg")) return shgSetup(c), 0;
	return fprintf(stderr, "unknown scenario %s\n", name), 2;
}

/* The observer box: the middle half of the domain on every axis. */
static float boxOut(world W)
{
	return poyntingOut(W, 0.5f * W->xMin, 0.5f * W->xMax, 0.5f * W->yMin, 0.5f * W->yMax, 0.5f * W->zMin, 0.5f * W->zMax);
}

static void stepScenario(scen *c, int phaserDone)
{
	if (c->P && !phaserDone) updatePhaser(c->W, c->P);
	sample(&c->T, c->W, c->F, c->pts, c->np);
	if (c->kq < 0 && c->W->n % 25 == 0) c->poyntingRun += boxOut(c->W);
	if (c->kq >= 0) {
		float (*p)[3] = c->spec;
		writeSpectrum(c->W, RUNSTEPS, 100, 800, "%s/kp%d-spectrum", OUTDIR, c->kq, get(c->W, Hz, p[0][0], p[0][1], p[0][2]),
			get(c->W, iHz, p[1][0], p[1][1], p[1][2]), get(c->W, Ez, p[2][0], p[2][1], p[2][2]));
	}
}

/* The spectrum file a k-point scenario wrote on its last step: hash of its bytes and the sum
 * of its numbers. */
static void putSpectrum(scen *c)
{
	char path[700];
	snprintf(path, sizeof path, "%s/kp%d-spectrum.txt", OUTDIR, c->kq);
	FILE *f = fopen(path, "rb");
	if (!f) {
		fprintf(OUT, "%s.spectrum missing\n", c->name);
		return;
	}
	static char text[1 << 16];
	size_t n = fread(text, 1, sizeof text - 1, f);
	fclose(f);
	text[n] = 0;
	double sum = 0;
	for (char *p = text, *e; *p; p = e) {
		double v = strtod(p, &e);
		if (e == p) e = p + 1;
		else sum += v;
	}
	fprintf(OUT, "%s.spectrum %016llx %.9g\n", c->name, (unsigned long long)fnv(1469598103934665603ULL, text, n), sum);
}

/* The far field is evaluated from stepped worlds only (on the device, which a zero-step
 * setup check does not have). */
static void dumpScenario(scen *c, int stepped)
{
	if (c->P) sync(c->W);
	putTrace(c->name, &c->T);
	if (c->kq >= 0 && stepped) putSpectrum(c);
	if (c->kq < 0 && stepped) {
		/* observers at the last step: planes through the observer box and the box's net flux;
		 * the power absorbed by each recorded dispersive object */
		world W = c->W;
		fprintf(OUT, "%s.poynting %.9g %.9g %.9g %.9g %.9g\n", c->name,
			poyntingX(W, 0.3f * W->xMax, 0.5f * W->yMin, 0.5f * W->yMax, 0.5f * W->zMin, 0.5f * W->zMax),
			poyntingY(W, 0.3f * W->yMin, 0.5f * W->xMin, 0.5f * W->xMax, 0.5f * W->zMin, 0.5f * W->zMax),
			poyntingZ(W, 0.3f * W->zMax, 0.5f * W->xMin, 0.5f * W->xMax, 0.5f * W->yMin, 0.5f * W->yMax),
			boxOut(W), c->poyntingRun);
		for (int a = 0; a < c->nabs; a++) fprintf(OUT, "%s.absorb%d %.9g\n", c->name, a, objectAbsorption(W, c->absorbers[a]));
	}
	dumpState(c->name, c->W);
	if (c->P) dumpPhasers(c->name, c->P, stepped);
}

/* Run the scenarios of a comma-separated list.  One scenario steps with updateE/updateH;
 * several step together with updateWorldsE/updateWorldsH (shared launches). */
static int run(const char *names, int steps)
{
	static scen c[MAXSCEN];
	int n = 0;
	/* "kpoints" stands for the whole k-point sweep kp0,...,kp39 */
	static char buf[4096];
	buf[0] = 0;
	char tmp[1024];
	snprintf(tmp, sizeof tmp, "%s", names);
	for (char *tok = strtok(tmp, ","); tok; tok = strtok(0, ",")) {
		size_t l = strlen(buf);
		if (!strcmp(tok, "kpoints"))
			for (int q = 0; q < NKP; q++) l += snprintf(buf + l, sizeof buf - l, "%skp%d", l ? "," : "", q);
		else snprintf(buf + l, sizeof buf - l, "%s%s", l ? "," : "", tok);
	}
	RUNSTEPS = steps;
	for (char *tok = strtok(buf, ","); tok; tok = strtok(0, ",")) {
		if (n == MAXSCEN) return fprintf(stderr, "too many scenarios\n"), 2;
		if (setupScenario(&c[n], tok)) return 2;
		static char label[MAXSCEN][32];
		snprintf(label[n], sizeof label[n], "%s", tok);
		c[n].name = label[n];
		dumpSetup(c[n].name, c[n].W);
		n++;
	}
	world Ws[MAXSCEN];
	for (int a = 0; a < n; a++) Ws[a] = c[a].W;
	for (int s = 1; s <= steps; s++) {
		for (int a = 0; a < n; a++)
			if (c[a].pump) updateE(c[a].pump), updateH(c[a].pump);
		if (n == 1) {
			updateE(Ws[0]);
			updateH(Ws[0]);
		}
		int phasersDone = 0;
		if (n > 1) {
#ifndef ALIS_REF
			/* several scenarios: shared launches for the half steps and the phasers */
			updateWorldsE(Ws, n);
			updateWorldsH(Ws, n);
			world PW[MAXSCEN];
			phaser PP[MAXSCEN];
			int np = 0;
			for (int a = 0; a < n; a++) if (c[a].P) PW[np] = c[a].W, PP[np++] = c[a].P;
			updatePhaserArrays(PW, PP, np);
			phasersDone = 1;
#else
			for (int a = 0; a < n; a++) updateE(Ws[a]), updateH(Ws[a]);
#endif
		}
		for (int a = 0; a < n; a++)
			if (c[a].pump) nonLinearSHG(c[a].W, c[a].pump, c[a].alt);
		for (int a = 0; a < n; a++) stepScenario(&c[a], phasersDone);
	}
	for (int a = 0; a < n; a++) dumpScenario(&c[a], steps > 0);
	for (int a = 0; a < n; a++) RUNW[a] = c[a].W, RUNP[a] = c[a].P, RUNPUMP[a] = c[a].pump, RUNNAME[a] = c[a].name;
	NRUN = n;
	LAST = c[0].W, LASTP = c[0].P;
	return 0;
}

/* Spectrum files go next to the output file. */
static void outDir(const char *path)
{
	snprintf(OUTDIR, sizeof OUTDIR, "%s", path);
	char *slash = strrchr(OUTDIR, '/');
	if (slash) *slash = 0;
	else snprintf(OUTDIR, sizeof OUTDIR, ".");
}

int alis_drive(const char *name, int steps, const char *path)
{
	outDir(path);
	OUT = fopen(path, "w");
	if (!OUT) return 1;
	int rc = run(name, steps);
	fclose(OUT);
	OUT = 0;
	releaseRun();
	return rc;
}

/* Nearest-node lookups (xtoi, ytoj, ztok) through the per-world cache: 3n coordinates --
 * random ones, nodes and midpoints between nodes -- read once in order and again in reverse
 * (after evictions and slot collisions), against a second world that meets each coordinate
 * for the first time.  Returns the number of differing indices. */
int alis_drive_nodes(int n)
{
	dom D = {{-200, 300}, {-100, 100}, {-50, 250}};
	res R = {1, {3.7}, {1240}};
	sur S = {{PML}, {PML}, {PML}, {8}};
	world A = createWorld(D, R, S, ""), B = createWorld(D, R, S, "");
	int (*fa[3])(world, float) = {A->xtoi, A->ytoj, A->ztok}, (*fb[3])(world, float) = {B->xtoi, B->ytoj, B->ztok};
	float (*coord[3])(world, float) = {A->itox, A->jtoy, A->ktoz};
	float *x = malloc(3 * (size_t)n * sizeof(float));
	unsigned long long seed = 7;
	for (int a = 0; a < 3; a++)
		for (int q = 0; q < n; q++) {
			seed = seed * 6364136223846793005ULL + 1442695040888963407ULL;
			const float u = (float)(seed >> 40) / 16777216.0f;
			const int i = (int)(u * 60);
			x[a * n + q] = q % 3 == 0 ? -260 + 620 * u : q % 3 == 1 ? coord[a](A, i) : 0.5f * (coord[a](A, i) + coord[a](A, i + 1));
		}
	int bad = 0;
	for (int a = 0; a < 3; a++)
		for (int q = 0; q < n; q++) (void)fa[a](A, x[a * n + q]);
	for (int a = 0; a < 3; a++)
		for (int q = n - 1; q >= 0; q--)
			if (fa[a](A, x[a * n + q]) != fb[a](B, x[a * n + q])) bad++;
	free(x);
	deleteWorld(A), deleteWorld(B);
	return bad;
}

/* Text outputs of a scenario with a phaser array after `steps` steps, written into dir:
 * writeRow and writeTxt lines, a row call with no numbers (which writes no file), the
 * Poynting and far-field spectra of the array (writePoyntingSpectrum with the array twice,
 * writeFarFieldSpectrum at theta 30, phi 45) and the far-field cuts of its first frequency
 * (farFieldTheta at phi 0, farFieldPhi at theta 60).  Over this library it also writes a
 * text with a '%' and one longer than 1 KiB, which the reference does not write as given. */
int alis_drive_text(const char *name, int steps, const char *dir)
{
	OUT = fopen("/dev/null", "w");
	int rc = run(name, steps);
	fclose(OUT);
	OUT = 0;
	if (rc || !LAST || !LASTP) return rc ? rc : 3;
	world W = LAST;
	phaser P = LASTP;
	writeRow(W, "%s/rows", dir, 1.5, 2.0, -3.25e-7);
	writeRow(W, "%s/rows", dir, (double)steps, 1e30);
	writeRow(W, "%s/empty", dir);
	writeTxt(W, "%s/notes-%d", dir, 3, "%s %d %g\n", "steps", steps, 0.125);
	writePoyntingSpectrum(W, "%s/poynting", dir, P, P);
	writeFarFieldSpectrum(W, Ppol, 30, 45, "%s/farfield", dir, P);
	farFieldTheta(W, P, 0, "%s/theta", dir);
	farFieldPhi(W, P, 60, "%s/phi", dir);
#ifndef ALIS_REF
	char longText[3001];
	memset(longText, 'x', 3000);
	longText[3000] = 0;
	writeTxt(W, "%s/percent", dir, "%d%%\n", 100);
	writeTxt(W, "%s/long", dir, "%s\n", longText);
#endif
	fflush(0);
	releaseRun();
	return 0;
}

#ifndef ALIS_REF
#include "alis_dev.h"
/* Raw sums of a writeSpectrum record (alis_amd/host/observers.c). */
long alisSpectrumRead(world W, const char *name, double *out, long n);

/* Run scenarios and write the final device state as doubles: for every world of the run, for
 * E then H, every component of the real (and imaginary) part over the ghost-inclusive grid;
 * then the first world's phaser accumulators if it has phasers; then the raw sums of every
 * k-point world's spectrum (source row first, then each probe, bins as re/im).  Header: 8
 * longs (iNUM, jNUM, kNUM, complexField of the first world, field doubles, phaser doubles,
 * spectrum doubles, worlds). */
int alis_drive_raw(const char *name, int steps, const char *path)
{
	outDir(path);
	OUT = fopen("/dev/null", "w");
	int rc = run(name, steps);
	fclose(OUT);
	OUT = 0;
	if (rc || !LAST || !LAST->dev) return rc ? rc : 3;
	FILE *f = fopen(path, "wb");
	if (!f) return 1;
	world W = LAST;
	long nf = 0;
	for (int a = 0; a < NRUN; a++)
		nf += 6L * (RUNW[a]->iNUM+2) * (RUNW[a]->jNUM+2) * (RUNW[a]->kNUM+2) * (1 + RUNW[a]->complexField);
	long np = 0;
	int pid = -1;
	if (LASTP) {
		phaser B = LASTP;
		pid = B->devId;
		if (pid >= 0) np = alis_dev_phaser_size(W->dev, pid);
	}
	static double spec[1 << 16];
	long ns = 0;
	for (int a = 0; a < NRUN; a++) {
		int q = -1;
		if (sscanf(RUNNAME[a], "kp%d", &q) != 1) continue;
		char nm[700];
		snprintf(nm, sizeof nm, "%s/kp%d-spectrum", OUTDIR, q);
		long got = alisSpectrumRead(RUNW[a], nm, spec + ns, (long)(sizeof spec / sizeof spec[0]) - ns);
		if (got <= 0) rc = 5;
		ns += got > 0 ? got : 0;
	}
	long head[8] = { W->iNUM, W->jNUM, W->kNUM, W->complexField, nf, np > 0 ? np : 0, ns, NRUN };
	fwrite(head, sizeof head, 1, f);
	for (int a = 0; a < NRUN; a++) {
		world V = RUNW[a];
		long per = (long)(V->iNUM+2) * (V->jNUM+2) * (V->kNUM+2);
		double *buf = malloc(per * sizeof(double));
		for (int h = 0; h < 2; h++)
			for (int m = 0; m <= V->complexField; m++)
				for (int x = 0; x < 3; x++) {
					if (alis_dev_read_field(V->dev, h, m, x, buf)) rc = 4;
					fwrite(buf, sizeof(double), per, f);
				}
		free(buf);
	}
	if (np > 0) {
		double *pb = malloc(np * sizeof(double));
		if (alis_dev_phaser_read(W->dev, pid, pb, np)) rc = 4;
		fwrite(pb, sizeof(double), np, f);
		free(pb);
	}
	fwrite(spec, sizeof(double), ns, f);
	fclose(f);
	releaseRun();
	return rc;
}

/* Closed-form check of the Yee update: a vacuum box periodic in x, y and z (node iMAX is
 * node iMIN, period iNUM-1) holding Ez = cos(2 pi (m i/(iNUM-1) + p j/(jNUM-1))) and H = 0.
 * This is an eigenvector of the discrete curl-curl, so every sample obeys the leapfrog
 * recurrence E^{n+1} + E^{n-1} = r E^n with r = 2 - 4 sum_a (dtE/da)(dtH/da) sin^2(k_a d_a/2),
 * for n >= 2 counted from the sample the mode is written at,
 * i.e. sin^2(w dt/2)/dt^2 = sum_a sin^2(k_a d_a/2)/d_a^2.  Returns r in *expect and the
 * recurrence ratio at every step n = 2..steps-1 (where |E^n| > 0.05) in ratio[]; the
 * return value is the number of ratios written. */
int alis_physics_dispersion(int cells, int m, int p, int steps, int fp64, double *ratio, double *expect)
{
	dom D = {{0, 10.0f * (cells-1)}, {0, 10.0f * (cells-1)}, {0, 40}};
	res R = {5, {10}, {1240}};
	sur S = {{PBC}, {PBC}, {PBC}};
	world W = createWorld(D, R, S, "");
	worldPrecision(W, fp64 ? 64 : 32, 64);
	const int ni = W->iNUM - 1, nj = W->jNUM - 1;
	const long per = (long)(W->iNUM+2) * (W->jNUM+2) * (W->kNUM+2);
	double *e = malloc((steps + 1) * sizeof(double)), *buf = malloc(per * sizeof(double));
	/* the first half step commits the (zero) world; the mode is then written in double */
	updateE(W);
	long c0 = 0;
	for (int i = W->iMIN-1; i <= W->iMAX+1; i++)
		for (int j = W->jMIN-1; j <= W->jMAX+1; j++)
			for (int k = W->kMIN-1; k <= W->kMAX+1; k++)
				buf[c0++] = cos(2 * M_PI * ((double)m * (i - W->iMIN) / ni + (double)p * (j - W->jMIN) / nj));
	alis_dev_write_field(W->dev, 0, 0, 2, buf);
	const int i0 = W->iMIN + 1, j0 = W->jMIN, k0 = W->kMIN;
	const double sx = sin(M_PI * m / ni), sy = sin(M_PI * p / nj);
	*expect = 2 - 4 * (double)W->CE->dtdx[i0] * W->CH->dtdx[i0] * sx * sx - 4 * (double)W->CE->dtdy[j0] * W->CH->dtdy[j0] * sy * sy;
	const long at = ((long)(i0 - W->iMIN + 1) * (W->jNUM+2) + (j0 - W->jMIN + 1)) * (W->kNUM+2) + (k0 - W->kMIN + 1);
	e[1] = buf[at];
	for (int n = 2; n <= steps; n++) {
		updateH(W);
		updateE(W);
		alis_dev_read_field(W->dev, 0, 0, 2, buf);
		e[n] = buf[at];
	}
	int c = 0;
	for (int n = 2; n < steps; n++)
		if (fabs(e[n]) > 0.05) ratio[c++] = (e[n+1] + e[n-1]) / e[n];
	free(e), free(buf);
	deleteWorld(W);
	return c;
}

/* Run a scenario and keep its worlds alive (for tests that call the engine ABI directly):
 * the device world of its first world, the device index of that world's phaser array and
 * the array's frequency count.  alis_drive_release frees them. */
int alis_drive_hold(const char *name, int steps, void **dev, int *pid, int *nf)
{
	OUT = fopen("/dev/null", "w");
	int rc = run(name, steps);
	fclose(OUT);
	OUT = 0;
	if (rc || !LAST || !LAST->dev || !LASTP) return rc ? rc : 3;
	if (LASTP->devId < 0) return 4;
	*dev = LAST->dev;
	*pid = LASTP->devId;
	*nf = LASTP->count;
	return 0;
}

void alis_drive_release(void)
{
	releaseRun();
}

/* Run a scenario and record slices of its first world into dir at its last steps: snapshots
 * of Ez and Hy on the plane z = 10 (text) and of Ex on the line through the centre along x
 * (text), an Ez image (PNG), the average of Hy over the last 10 steps and the 500 nm phasor of
 * Ez over the last 20 steps (text, 12 frames); returns 0. */
int alis_drive_slices(const char *name, int steps, const char *dir)
{
	OUT = fopen("/dev/null", "w");
	int rc = run(name, steps > 20 ? steps - 20 : 0);
	fclose(OUT);
	OUT = 0;
	if (rc || !LAST) return rc ? rc : 3;
	world W = LAST;
	slice XY = createSliceXY(W, 10), X = createSliceX(W, 0, 0);
	for (int s = 0; s < 20; s++) {
		updateE(W);
		updateH(W);
		sliceTimeAvg(W, Hy, XY, 10, txt, "%s/avg-%%", dir);
		sliceFreqDom(W, Ez, XY, 20, 500, txt, "%s/phasor-%%", dir);
	}
	sliceSnap(W, Ez, XY, 0, txt, "%s/snap-%%", dir);
	sliceSnap(W, Hy, XY, 0, txt, "%s/snap-%%", dir);
	sliceSnap(W, Ex, X, 0, txt, "%s/line-%%", dir);
	sliceSnap(W, Ez, XY, 0, png(dkbr, 0), "%s/image-%%", dir);
	deleteSlice(XY), deleteSlice(X);
	releaseRun();
	return 0;
}

/* The host library's residency control (alis_amd/host/residency.c), which brings a world's
 * state and its phasers' accumulators home and drops the device copy, and its double
 * precision net flux behind phaserPoyntingOut (alis_amd/host/phasers.c). */
void alisInvalidate(world W);
double phaserFluxD(phaser P);

/* The double net flux (phaserFluxD, phaserPoyntingOut before its rounding to float) of a
 * phaser array whose x and y faces have rows of 181 cells (longer than the 64 lane sums of
 * the engine's flux terms): an 88 x 88 x 408 nm vacuum world (dx 2 nm, Courant 0.5, 6-cell CPML) with an
 * Ex dipole, phaser box +-16 x +-16 x +-180 nm, 5 wavelengths 500-700 nm.  After `steps`
 * steps, first with the accumulators on the device (flux sums from the engine), then after
 * alisInvalidate brought them home (the same sums on the host): out[2q], out[2q + 1].
 * Returns the number of frequencies (at most maxq). */
int alis_drive_flux_paths(int steps, double *out, int maxq)
{
	dom D = {{-44, 44}, {-44, 44}, {-204, 204}};
	res R = {1, {2}, {1240}};
	sur S = {{PML}, {PML}, {PML}, {6}};
	world W = createWorld(D, R, S, "");
	pointDipole(W, Ex, 2, -2, 10, Band, 400, 800);
	phaser P = createPhasers(W, 500, 700, 50, -16, 16, -16, 16, -180, 180);
	for (int s = 0; s < steps; s++) {
		updateE(W);
		updateH(W);
		updatePhaser(W, P);
	}
	int n = 0;
	for (; n < maxq && P[n].f; n++) out[2 * n] = phaserFluxD(P + n);
	alisInvalidate(W);
	for (int q = 0; q < n; q++) out[2 * q + 1] = phaserFluxD(P + q);
	deletePhaser(P);
	deleteWorld(W);
	return n;
}

/* The host library's double-precision far field (farFieldD, alis_amd/host/farfield.c). */
double complex farFieldD(phaser P, tfield pol(phaser,float,float), double theta, double phi);

/* Run a scenario with phasers and evaluate the far field of every frequency of its phaser
 * array in double precision at the n directions (theta[a], phi[a]): out[(q*n + a)*4 + ...]
 * = P re, P im, S re, S im.  Returns the number of frequencies written (at most maxq), or a
 * negative code. */
int alis_drive_farfield(const char *name, int steps, int n, const double *theta, const double *phi, int maxq, double *out)
{
	OUT = fopen("/dev/null", "w");
	int rc = run(name, steps);
	fclose(OUT);
	OUT = 0;
	if (rc || !LASTP) return rc ? -rc : -3;
	int nq = 0;
	for (phaser Q = LASTP; Q->f && nq < maxq; Q++, nq++)
		for (int a = 0; a < n; a++) {
			double complex p = farFieldD(Q, Ppol, theta[a], phi[a]), s = farFieldD(Q, Spol, theta[a], phi[a]);
			double *o = out + ((long)nq * n + a) * 4;
			o[0] = creal(p), o[1] = cimag(p), o[2] = creal(s), o[3] = cimag(s);
		}
	releaseRun();
	return nq;
}
#endif

#ifdef DRIVE_MAIN
int main(int argc, char **argv)
{
	if (argc < 2) return fprintf(stderr, "usage: %s SCENARIO [STEPS [OUTPATH]] | --text SCENARIO STEPS DIR\n", argv[0]), 2;
	if (argc > 4 && !strcmp(argv[1], "--text")) return alis_drive_text(argv[2], atoi(argv[3]), argv[4]);
	if (argc > 3) return alis_drive(argv[1], atoi(argv[2]), argv[3]);
	OUT = stdout;
	setvbuf(stdout, 0, _IOLBF, 0);
	return run(argv[1], argc > 2 ? atoi(argv[2]) : 60);
}
#endif
