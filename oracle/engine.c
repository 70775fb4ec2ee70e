/*
 * THE ORACLE -- test infrastructure only.  Nothing in the product (alis_amd/) links, loads
 * or calls this file; the tests, the entry point's smoke check and the benchmark's CPU
 * baseline are its only users.
 *
 * A CPU restatement of the reference's half-step arithmetic behind the engine ABI
 * (include/alis_dev.h): the same calls the HIP engine serves, computed with plain
 * sequential loops that follow src/update.c:48-341 and src/phaser.c:83-110 line by line,
 * for float (bitwise the reference's own float code: same operation order, no
 * contraction) and for double (the FP64 oracle: the same loops on doubles, coefficients
 * widened from the float setup).  Linked under the host library (alis_amd/host) it runs
 * whole alis.h programs on the CPU.
 *
 * Pinning: the float instantiation is checked bit for bit against the reference library
 * compiled from its own sources (oracle/Makefile -> oracle/_ref) on the scenarios of
 * tests/golden/drive.c, whose outputs are committed under tests/golden/.  The double
 * instantiation runs the same code with T = double; it is further checked by closed-form
 * properties in tests/test_oracle_physics.py (discrete dispersion of a periodic box mode,
 * energy conservation in a closed PEC cavity, convergence of the float path towards it).
 *
 * FP64 phaser accumulators add with Neumaier compensation, folded into the sums before every
 * read, as the engine does; tests/test_phaser_compensation.py pins it with a sample stream
 * whose plain double sum loses every term but the first.
 *
 * Every access goes through a bounds-checked index map that aborts on a miss.
 */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "alis_dev.h"

#define MIN(a, b) ((a) < (b) ? (a) : (b))
#define MAX(a, b) ((a) > (b) ? (a) : (b))

/* Boundary codes of the reference (src/alis_c.h:15-21). */
enum { SYM = 1, PEC = 2, PML = 3, PBC = 4, HBC = 5, BBC = 6 };

typedef struct alis_dev_world world;

/* A 3-D array over [i0..i1][j0..j1][k0..k1], k fastest.  Worlds that stand in for a
 * TF/SF line (expandX/expandY) alias every row above iTop/jTop to row iTop+alt/jTop+alt,
 * as expandFieldX/Y do with row pointers (src/source.c:223-248). */
typedef struct {
	void *p;
	int i0, i1, j0, j1, k0, k1;
	long si, sj;
	int exX, exY, iTop, jTop, altX, altY;
	const char *what;
} arr;

typedef struct {
	int live, nf, dft64;
	int iMin, iMax, jMin, jMax, kMin, kMax;
	int nface, kind[6];
	long off[6], ncell;
	double *accD, *cmpD;      /* FP64 accumulators and their Neumaier compensations */
	float *accF;
	float **cf;
} phs;

struct alis_dev_world {
	alis_dev_world_desc W;
	alis_dev_ade *obj;
	alis_dev_src *src[2];
	alis_dev_tfsf *tf;
	world **aux;
	int fp64, dft64, hasEps;
	arr F[2][2][3];                        /* [E,H][m][axis] */
	arr slab[2][2][6][2];                  /* [E,H][m][face][a,b] */
	arr (*J)[2][3][ALIS_DEV_MAXPOLES];     /* [object][m][axis][pole] */
	arr (*K)[2][3][ALIS_DEV_MAXPOLES];
	phs *ph;
	int nph;
	long bytes;
};

static char errbuf[256];

static int fail(const char *what)
{
	snprintf(errbuf, sizeof errbuf, "%s", what);
	return -1;
}

const char *alis_dev_error(void) { return errbuf; }

static void oob(const arr *a, int i, int j, int k)
{
	fprintf(stderr, "oracle: index (%d,%d,%d) outside %s [%d..%d][%d..%d][%d..%d]\n", i, j, k, a->what,
		a->i0, a->i1, a->j0, a->j1, a->k0, a->k1);
	abort();
}

static inline long aix(const arr *a, int i, int j, int k)
{
	if (a->exX && i > a->iTop) i = a->iTop + a->altX;
	if (a->exY && j > a->jTop) j = a->jTop + a->altY;
	if (i < a->i0 || i > a->i1 || j < a->j0 || j > a->j1 || k < a->k0 || k > a->k1) oob(a, i, j, k);
	return (i - a->i0) * a->si + (j - a->j0) * a->sj + (k - a->k0);
}

static long acount(const arr *a)
{
	return (long)(a->i1 - a->i0 + 1) * (a->j1 - a->j0 + 1) * (a->k1 - a->k0 + 1);
}

static void amake(world *D, arr *a, int i0, int i1, int j0, int j1, int k0, int k1, const char *what)
{
	memset(a, 0, sizeof *a);
	a->i0 = i0, a->i1 = i1, a->j0 = j0, a->j1 = j1, a->k0 = k0, a->k1 = k1;
	a->sj = k1 - k0 + 1;
	a->si = a->sj * (j1 - j0 + 1);
	a->what = what;
	const size_t es = D->fp64 ? sizeof(double) : sizeof(float);
	const long n = acount(a);
	a->p = calloc(n > 0 ? n : 1, es);
	D->bytes += n * es;
}

/* Host float*** <-> state.  The host array covers the same box (aliased rows of an
 * expanded world hold the same values, so each direction is consistent). */
static void aload(world *D, arr *a, float ***h)
{
	for (int i = a->i0; i <= a->i1; i++)
		for (int j = a->j0; j <= a->j1; j++)
			for (int k = a->k0; k <= a->k1; k++) {
				if (D->fp64) ((double *)a->p)[aix(a, i, j, k)] = h[i][j][k];
				else ((float *)a->p)[aix(a, i, j, k)] = h[i][j][k];
			}
}

static void astore(world *D, arr *a, float ***h, int i0, int i1, int j0, int j1, int k0, int k1)
{
	i0 = MAX(i0, a->i0), i1 = MIN(i1, a->i1), j0 = MAX(j0, a->j0), j1 = MIN(j1, a->j1), k0 = MAX(k0, a->k0), k1 = MIN(k1, a->k1);
	for (int i = i0; i <= i1; i++)
		for (int j = j0; j <= j1; j++)
			for (int k = k0; k <= k1; k++)
				h[i][j][k] = D->fp64 ? (float)((double *)a->p)[aix(a, i, j, k)] : ((float *)a->p)[aix(a, i, j, k)];
}

#define EPS(D, c, i, j, k) ((D)->W.eps[c][i][j][k])

/* Where a pump value of the three-wave coupling sits (ops.h): at the component's own
 * position, or the mean of four values of another component. */
enum { TW_OWN = -1, TW_YX, TW_ZX, TW_ZY, TW_XY, TW_XZ, TW_YZ };

#define T float
#define SUF(name) name##_f
#define V(a, i, j, k) (((float *)(a)->p)[aix((a), (i), (j), (k))])
#include "ops.h"
#undef T
#undef SUF
#undef V

#define T double
#define SUF(name) name##_d
#define V(a, i, j, k) (((double *)(a)->p)[aix((a), (i), (j), (k))])
#include "ops.h"
#undef T
#undef SUF
#undef V

/* Aliasing steps of expandFieldX / expandFieldY per component E.x E.y E.z H.x H.y H.z
 * (src/source.c:293-312). */
static const int altX[2][3] = {{1, 0, 0}, {0, 1, 1}}, altY[2][3] = {{0, 1, 0}, {1, 0, 1}};

int alis_dev_commit(const alis_dev_world_desc *Wd, int fp64, int dft64, alis_dev_world **out)
{
	world *D = calloc(1, sizeof *D);
	D->W = *Wd;
	D->fp64 = fp64, D->dft64 = dft64;
	const alis_dev_world_desc *W = &D->W;
	D->obj = calloc(W->nobj ? W->nobj : 1, sizeof *D->obj);
	memcpy(D->obj, Wd->obj, W->nobj * sizeof *D->obj);
	D->W.obj = D->obj;
	for (int h = 0; h < 2; h++) {
		D->src[h] = calloc(W->nsrc[h] ? W->nsrc[h] : 1, sizeof *D->src[h]);
		memcpy(D->src[h], Wd->src[h], W->nsrc[h] * sizeof *D->src[h]);
		D->W.src[h] = D->src[h];
	}
	D->tf = calloc(W->ntfsf ? W->ntfsf : 1, sizeof *D->tf);
	memcpy(D->tf, Wd->tfsf, W->ntfsf * sizeof *D->tf);
	D->W.tfsf = D->tf;
	D->aux = calloc(W->ntfsf ? W->ntfsf : 1, sizeof *D->aux);
	D->hasEps = W->eps[0] && W->eps[1] && W->eps[2];

	/* fields with the ghost layer (src/world.c:164-189) */
	for (int h = 0; h < 2; h++)
		for (int m = 0; m <= W->complexField; m++)
			for (int a = 0; a < 3; a++) {
				arr *A = &D->F[h][m][a];
				amake(D, A, W->iMIN-1, W->iMAX+1, W->jMIN-1, W->jMAX+1, W->kMIN-1, W->kMAX+1, h ? "H field" : "E field");
				A->exX = W->expandX, A->iTop = W->iMAX, A->altX = altX[h][a];
				A->exY = W->expandY, A->jTop = W->jMAX, A->altY = altY[h][a];
				aload(D, A, W->field[h][m][a]);
			}
	/* CPML slabs over the loop ranges of src/update.c:119-124 (Off 0 for E, 1 for H) */
	for (int h = 0; h < 2; h++)
		for (int f = 0; f < 6; f++) {
			if (!W->slab[h][0][f][0]) continue;
			const int ax = f % 3, Off = h;
			int lo, hi;
			if (f < 3) lo = (ax == 0 ? W->iMIN : ax == 1 ? W->jMIN : W->kMIN) + Off, hi = (ax == 0 ? W->iMin : ax == 1 ? W->jMin : W->kMin) + Off - 2;
			else lo = (ax == 0 ? W->iMax : ax == 1 ? W->jMax : W->kMax) + 2, hi = ax == 0 ? W->iMAX : ax == 1 ? W->jMAX : W->kMAX;
			int i0 = W->iMIN, i1 = W->iMAX, j0 = W->jMIN, j1 = W->jMAX, k0 = W->kMIN, k1 = W->kMAX;
			if (ax == 0) i0 = lo, i1 = hi;
			if (ax == 1) j0 = lo, j1 = hi;
			if (ax == 2) k0 = lo, k1 = hi;
			for (int m = 0; m <= W->complexField; m++)
				for (int s = 0; s < 2; s++) {
					amake(D, &D->slab[h][m][f][s], i0, i1, j0, j1, k0, k1, "CPML slab");
					aload(D, &D->slab[h][m][f][s], W->slab[h][m][f][s]);
				}
		}
	/* ADE currents over each object box (src/object.c:178-226) */
	D->J = calloc(W->nobj ? W->nobj : 1, sizeof *D->J);
	D->K = calloc(W->nobj ? W->nobj : 1, sizeof *D->K);
	for (int n = 0; n < W->nobj; n++) {
		const alis_dev_ade *o = &D->obj[n];
		if (o->nk > ALIS_DEV_MAXPOLES) { free(D); return fail("too many poles"); }
		for (int m = 0; m <= W->complexField; m++)
			for (int c = 0; c < 3; c++)
				for (int p = 0; p < o->nk; p++) {
					amake(D, &D->J[n][m][c][p], o->i0, o->i1, o->j0, o->j1, o->k0, o->k1, "ADE J");
					aload(D, &D->J[n][m][c][p], o->J[m][c][p]);
					if (p >= o->nj) {
						amake(D, &D->K[n][m][c][p], o->i0, o->i1, o->j0, o->j1, o->k0, o->k1, "ADE K");
						aload(D, &D->K[n][m][c][p], o->K[m][c][p]);
					}
				}
	}
	*out = D;
	return 0;
}

int alis_dev_link_tfsf(alis_dev_world *D, int s, alis_dev_world *aux)
{
	if (s < 0 || s >= D->W.ntfsf) return fail("no such TF/SF box");
	D->aux[s] = aux;
	return 0;
}

static void freeArr(arr *a) { free(a->p); a->p = 0; }

void alis_dev_release(alis_dev_world *D)
{
	if (!D) return;
	for (int h = 0; h < 2; h++)
		for (int m = 0; m < 2; m++) {
			for (int a = 0; a < 3; a++) freeArr(&D->F[h][m][a]);
			for (int f = 0; f < 6; f++) freeArr(&D->slab[h][m][f][0]), freeArr(&D->slab[h][m][f][1]);
		}
	for (int n = 0; n < D->W.nobj; n++)
		for (int m = 0; m < 2; m++)
			for (int c = 0; c < 3; c++)
				for (int p = 0; p < ALIS_DEV_MAXPOLES; p++) freeArr(&D->J[n][m][c][p]), freeArr(&D->K[n][m][c][p]);
	for (int q = 0; q < D->nph; q++) free(D->ph[q].accD), free(D->ph[q].cmpD), free(D->ph[q].accF), free(D->ph[q].cf);
	free(D->ph);
	free(D->J), free(D->K), free(D->obj), free(D->src[0]), free(D->src[1]), free(D->tf), free(D->aux);
	free(D);
}

/* updateE / updateH core (src/update.c:277-280, 304-307). */
int alis_dev_core(alis_dev_world *D, int h)
{
	for (int m = 0; m <= D->W.complexField; m++) {
		if (D->fp64) {
			if (!h) objectsE_d(D, m);
			curl_d(D, h, m);
			if (h) objectsH_d(D, m);
		}
		else {
			if (!h) objectsE_f(D, m);
			curl_f(D, h, m);
			if (h) objectsH_f(D, m);
		}
	}
	return 0;
}

int alis_dev_strips(alis_dev_world *D, int s, int h)
{
	if (s < 0 || s >= D->W.ntfsf || !D->aux[s]) return fail("TF/SF box without its auxiliary world");
	if (D->aux[s]->fp64 != D->fp64) return fail("TF/SF auxiliary world of another precision");
	if (D->fp64) strips_d(D, s, h);
	else strips_f(D, s, h);
	return 0;
}

int alis_dev_inject(alis_dev_world *D, int h, int index, float wave, const float *cellWave, long ncell)
{
	if (h < 0 || h > 1 || index < 0 || index >= D->W.nsrc[h]) return fail("no such source");
	const alis_dev_src *s = &D->src[h][index];
	if (s->m < 0 || s->m > D->W.complexField || s->axis < 0) return fail("source targets no field array");
	const long cells = (long)(s->i1 - s->i0 + 1) * (s->j1 - s->j0 + 1) * (s->k1 - s->k0 + 1);
	if (cellWave && ncell != cells) return fail("per-cell wave of the wrong size");
	if (D->fp64) inject_d(D, h, index, wave, cellWave);
	else inject_f(D, h, index, wave, cellWave);
	return 0;
}

int alis_dev_ghosts(alis_dev_world *D, int h)
{
	if (D->fp64) ghosts_d(D, h);
	else ghosts_f(D, h);
	return 0;
}

int alis_dev_sync(alis_dev_world *D, int i0, int i1, int j0, int j1, int k0, int k1)
{
	const alis_dev_world_desc *W = &D->W;
	const int whole = i0 > i1;
	if (whole) i0 = j0 = k0 = -(1 << 30), i1 = j1 = k1 = 1 << 30;
	for (int h = 0; h < 2; h++)
		for (int m = 0; m <= W->complexField; m++)
			for (int a = 0; a < 3; a++) astore(D, &D->F[h][m][a], W->field[h][m][a], i0, i1, j0, j1, k0, k1);
	for (int n = 0; n < W->nobj; n++) {
		const alis_dev_ade *o = &D->obj[n];
		if (!whole && (o->i1 < i0 || o->i0 > i1 || o->j1 < j0 || o->j0 > j1 || o->k1 < k0 || o->k0 > k1)) continue;
		for (int m = 0; m <= W->complexField; m++)
			for (int c = 0; c < 3; c++)
				for (int p = 0; p < o->nk; p++) {
					arr *J = &D->J[n][m][c][p];
					astore(D, J, o->J[m][c][p], J->i0, J->i1, J->j0, J->j1, J->k0, J->k1);
					if (p >= o->nj) {
						arr *K = &D->K[n][m][c][p];
						astore(D, K, o->K[m][c][p], K->i0, K->i1, K->j0, K->j1, K->k0, K->k1);
					}
				}
	}
	if (!whole) return 0;
	for (int h = 0; h < 2; h++)
		for (int m = 0; m <= W->complexField; m++)
			for (int f = 0; f < 6; f++)
				for (int s = 0; s < 2; s++) {
					arr *S = &D->slab[h][m][f][s];
					if (S->p) astore(D, S, W->slab[h][m][f][s], S->i0, S->i1, S->j0, S->j1, S->k0, S->k1);
				}
	return 0;
}

int alis_dev_read_field(alis_dev_world *D, int h, int m, int axis, double *out)
{
	if (h < 0 || h > 1 || m < 0 || m > D->W.complexField || axis < 0 || axis > 2) return fail("no such field array");
	arr *A = &D->F[h][m][axis];
	long c = 0;
	for (int i = A->i0; i <= A->i1; i++)
		for (int j = A->j0; j <= A->j1; j++)
			for (int k = A->k0; k <= A->k1; k++)
				out[c++] = D->fp64 ? ((double *)A->p)[aix(A, i, j, k)] : ((float *)A->p)[aix(A, i, j, k)];
	return 0;
}

int alis_dev_write_field(alis_dev_world *D, int h, int m, int axis, const double *in)
{
	if (h < 0 || h > 1 || m < 0 || m > D->W.complexField || axis < 0 || axis > 2) return fail("no such field array");
	arr *A = &D->F[h][m][axis];
	long c = 0;
	for (int i = A->i0; i <= A->i1; i++)
		for (int j = A->j0; j <= A->j1; j++)
			for (int k = A->k0; k <= A->k1; k++, c++) {
				if (D->fp64) ((double *)A->p)[aix(A, i, j, k)] = in[c];
				else ((float *)A->p)[aix(A, i, j, k)] = (float)in[c];
			}
	return 0;
}

/* ---- phasers (src/phaser.c:21-110) ---- */

int alis_dev_phaser_add(alis_dev_world *D, const alis_dev_phaser_desc *P, int *id)
{
	const alis_dev_world_desc *W = &D->W;
	if (P->iMin < W->iMIN || P->iMax > W->iMAX || P->jMin < W->jMIN || P->jMax > W->jMAX || P->kMin < W->kMIN || P->kMax > W->kMAX ||
	    P->iMin > P->iMax || P->jMin > P->jMax || P->kMin > P->kMax)
		return fail("phaser box outside the grid");
	D->ph = realloc(D->ph, (D->nph + 1) * sizeof *D->ph);
	phs *Q = &D->ph[D->nph];
	memset(Q, 0, sizeof *Q);
	Q->nf = P->nf, Q->dft64 = D->dft64;
	Q->iMin = P->iMin, Q->iMax = P->iMax, Q->jMin = P->jMin, Q->jMax = P->jMax, Q->kMin = P->kMin, Q->kMax = P->kMax;
	Q->cf = malloc((size_t)Q->nf * 24 * sizeof(float *));
	memcpy(Q->cf, P->cf, (size_t)Q->nf * 24 * sizeof(float *));
	const long ni = Q->iMax-Q->iMin+1, nj = Q->jMax-Q->jMin+1, nk = Q->kMax-Q->kMin+1;
	for (int kind = 0; kind < 6; kind++) {
		if (!Q->cf[kind*4]) continue;
		Q->kind[Q->nface] = kind;
		Q->off[Q->nface] = Q->ncell;
		Q->ncell += kind < 2 ? nj*nk : kind < 4 ? ni*nk : ni*nj;
		Q->nface++;
	}
	const long n = (long)Q->nf * 4 * Q->ncell * 2;
	if (Q->dft64) Q->accD = calloc(n ? n : 1, sizeof(double)), Q->cmpD = calloc(n ? n : 1, sizeof(double));
	else Q->accF = calloc(n ? n : 1, sizeof(float));
	for (int q = 0; q < Q->nf; q++)
		for (int f = 0; f < Q->nface; f++) {
			const long cells = (f + 1 < Q->nface ? Q->off[f+1] : Q->ncell) - Q->off[f];
			for (int comp = 0; comp < 4; comp++) {
				const float *src = Q->cf[((size_t)q*6 + Q->kind[f])*4 + comp];
				const long base = 2*(((long)q*4 + comp)*Q->ncell + Q->off[f]);
				for (long c = 0; c < 2*cells; c++) {
					if (Q->dft64) Q->accD[base + c] = src[c];
					else Q->accF[base + c] = src[c];
				}
			}
		}
	Q->live = 1;
	*id = D->nph++;
	return 0;
}

static phs *getPhaser(alis_dev_world *D, int id)
{
	return id >= 0 && id < D->nph && D->ph[id].live ? &D->ph[id] : 0;
}

/* Fold the compensations of FP64 accumulators into their sums (before every read): where a
 * compensation is not zero, acc += cmp and cmp = 0. */
static void settle(phs *Q)
{
	if (!Q->dft64) return;
	const long n = (long)Q->nf * 4 * Q->ncell * 2;
	for (long c = 0; c < n; c++)
		if (Q->cmpD[c] != 0) Q->accD[c] += Q->cmpD[c], Q->cmpD[c] = 0;
}

int alis_dev_phaser_update(alis_dev_world *D, int id, int q0, int nq, const double *expE, const double *expH)
{
	phs *P = getPhaser(D, id);
	if (!P) return fail("no such phaser");
	if (q0 < 0 || nq < 0 || q0 + nq > P->nf) return fail("phaser frequency range out of bounds");
	if (D->fp64) phaserStep_d(D, P, q0, nq, expE, expH);
	else phaserStep_f(D, P, q0, nq, expE, expH);
	return 0;
}

int alis_dev_sync_boxes(alis_dev_world *const *D, const int *box, int n)
{
	/* every array inside the box: a superset of the selected ones */
	for (int e = 0; e < n; e++) {
		const int *b = box + 7 * e;
		if (alis_dev_sync(D[e], b[0], b[1], b[2], b[3], b[4], b[5])) return -1;
	}
	return 0;
}

int alis_dev_phaser_update_many(alis_dev_world *const *D, const int *id, const int *q0, const int *nq,
                                const double *const *expE, const double *const *expH, int n)
{
	for (int e = 0; e < n; e++) if (alis_dev_phaser_update(D[e], id[e], q0[e], nq[e], expE[e], expH[e])) return -1;
	return 0;
}

int alis_dev_phaser_sync(alis_dev_world *D, int id)
{
	phs *Q = getPhaser(D, id);
	if (!Q) return fail("no such phaser");
	settle(Q);
	for (int q = 0; q < Q->nf; q++)
		for (int f = 0; f < Q->nface; f++) {
			const long cells = (f + 1 < Q->nface ? Q->off[f+1] : Q->ncell) - Q->off[f];
			for (int comp = 0; comp < 4; comp++) {
				float *dst = Q->cf[((size_t)q*6 + Q->kind[f])*4 + comp];
				const long base = 2*(((long)q*4 + comp)*Q->ncell + Q->off[f]);
				for (long c = 0; c < 2*cells; c++) dst[c] = Q->dft64 ? (float)Q->accD[base + c] : Q->accF[base + c];
			}
		}
	return 0;
}

long alis_dev_phaser_size(alis_dev_world *D, int id)
{
	phs *Q = getPhaser(D, id);
	return Q ? (long)Q->nf * 4 * Q->ncell * 2 : -1;
}

int alis_dev_phaser_read(alis_dev_world *D, int id, double *out, long n)
{
	phs *Q = getPhaser(D, id);
	const long need = alis_dev_phaser_size(D, id);
	if (!Q || n < need) return fail("phaser read buffer too small");
	settle(Q);
	for (long c = 0; c < need; c++) out[c] = Q->dft64 ? Q->accD[c] : Q->accF[c];
	return 0;
}

/* The flux's inner trapezoid sums (src/phaser.c:142-223) in the term form of the device
 * entry: each output adds its cells in the stated order, so the two agree to the bit. */
int alis_dev_phaser_flux_sums(alis_dev_world *D, int id, int q, int nterm, const int *terms,
                              const double *w, long nw, double *out, long nout)
{
	phs *Q = getPhaser(D, id);
	if (!Q) return fail("no such phaser");
	if (q < 0 || q >= Q->nf || nterm < 1 || nterm > 12 || nw < 1 || nout < 1) return fail("flux sums: bad frequency, term count or sizes");
	for (int n = 0; n < nterm; n++) {
		const int *u = terms + 8 * n;
		const long na = u[2], nb = u[3], len = u[0] == 0 ? na * 64 : nb, wl = u[0] == 0 ? nb : na;
		if ((u[0] != 0 && u[0] != 1) || na < 1 || nb < 1 || u[1] < 0 || u[1] + na * nb > Q->ncell || u[4] < 0 || u[4] > 3
		    || u[5] < 0 || u[5] > 3 || u[6] < 0 || u[6] + wl > nw || u[7] < 0 || u[7] + len > nout)
			return fail("flux sums: term outside the phaser faces or the buffers");
	}
	settle(Q);
	for (long c = 0; c < nout; c++) out[c] = 0;
	for (int n = 0; n < nterm; n++) {
		const int *u = terms + 8 * n;
		const long na = u[2], nb = u[3];
		const long bx = 2 * (((long)q * 4 + u[4]) * Q->ncell + u[1]), by = 2 * (((long)q * 4 + u[5]) * Q->ncell + u[1]);
		double *o = out + u[7];
		if (u[0] == 0) for (long c = 0; c < na * 64; c++) o[c] = 0;
		else for (long c = 0; c < nb; c++) o[c] = 0;
		for (long a = 0; a < na; a++)
			for (long b = 0; b < nb; b++) {
				const long px = bx + 2 * (a * nb + b), py = by + 2 * (a * nb + b);
				const double xr = Q->dft64 ? Q->accD[px] : Q->accF[px], xi = Q->dft64 ? Q->accD[px + 1] : Q->accF[px + 1];
				const double yr = Q->dft64 ? Q->accD[py] : Q->accF[py], yi = Q->dft64 ? Q->accD[py + 1] : Q->accF[py + 1];
				if (u[0] == 0) o[a * 64 + b % 64] += w[u[6] + b] * (xr * yr + xi * yi);
				else o[b] += w[u[6] + a] * (xr * yr + xi * yi);
			}
	}
	return 0;
}

/* The allocation-free form of the device entry: the same sums, refused beyond the size of
 * the device's scratch array (outputs, weights and term table in 524288 doubles). */
int alis_dev_phaser_flux_sums_range(alis_dev_world *D, int id, int q0, int nq, int nterm, const int *terms,
                                    const double *w, long nw, double *out, long nout)
{
	phs *Q = getPhaser(D, id);
	if (!Q) return fail("no such phaser");
	if (q0 < 0 || nq < 1 || q0 + nq > Q->nf) return fail("flux sums: bad frequency range, term count or sizes");
	for (int q = q0; q < q0 + nq; q++)
		if (alis_dev_phaser_flux_sums(D, id, q, nterm, terms, w, nw, out + (long)(q - q0) * nout, nout)) return -1;
	return 0;
}

int alis_dev_phaser_flux_sums_noalloc(alis_dev_world *D, int id, int q, int nterm, const int *terms,
                                      const double *w, long nw, double *out, long nout)
{
	if (nterm >= 1 && nout + nw + (9L * nterm + 1) / 2 > (1L << 19)) return fail("flux sums without allocation: request exceeds the scratch array");
	return alis_dev_phaser_flux_sums(D, id, q, nterm, terms, w, nw, out, nout);
}

int alis_dev_phaser_remove(alis_dev_world *D, int id)
{
	phs *Q = getPhaser(D, id);
	if (!Q) return fail("no such phaser");
	free(Q->accD), free(Q->cmpD), free(Q->accF), free(Q->cf);
	Q->accD = 0, Q->cmpD = 0, Q->accF = 0, Q->cf = 0, Q->live = 0;
	return 0;
}

/* ---- near-to-far surface sums (src/farfld.c:248-459) ----
 * For each column the loops of farFieldX/Y/Z: an inner sum along the face's last axis and an
 * outer sum (compensated) along its first, of F times the two weight vectors. */

typedef double complex dcx;

/* In-plane phase weight of cell p (0..np-1) of an axis: E widths at nodes with the two end
 * cells halved and the mirror image subtracted (half = 0), or H widths at half nodes, the
 * first cell left out and the mirror image added (half = 1) (src/farfld.c:262-268,282-288). */
static dcx ffPhase(int half, const double *d, const double *pos, int p, int np, double k, int sym, double cosk)
{
	double w = half ? (p ? 1 : 0) : (p == 0 ? 0.5 : 0) + (p == np-1 ? 0.5 : 0) + (p > 0 && p < np-1 ? 1 : 0);
	dcx v = d[p] * cexp(I * k * pos[p]);
	if (sym) v += (half ? 1 : -1) * conj(v) * cosk;
	return w * v;
}

/* Profile weight of cell v of the box along z for T component t of polarisation pol. */
static dcx ffProfile(int half, const alis_dev_ff_geom *g, const alis_dev_ff_col *c, int pol, int t, int k0, int v, int nk)
{
	const int src = c->src[pol][t];
	const double fac = c->fac[pol][t];
	if (src < 0 || !fac || (half && !v)) return 0;
	const double *T = g->T + 2 * ((long)c->slot[pol] * 6 + src) * g->nz;
	const int k = k0 + v;
#define TV(kk) (T[2*((kk)-g->kLo)] + I * T[2*((kk)-g->kLo)+1])
	dcx z;
	if (half) z = g->dzH[v] * (TV(k) + (g->symZ ? TV(1-k) * g->coskz : 0));
	else z = ((v == 0 ? 0.5 : 0) + (v == nk-1 ? 0.5 : 0) + (v > 0 && v < nk-1 ? 1 : 0)) * g->dzE[v] * (TV(k) - (g->symZ ? TV(-k) * g->coskz : 0));
#undef TV
	return fac * z;
}

static void ffAdd(dcx *s, dcx *c, dcx x)
{
	double sr = creal(*s), si = cimag(*s), cr = creal(*c), ci = cimag(*c);
	double t = sr + creal(x);
	cr += fabs(sr) >= fabs(creal(x)) ? (sr - t) + creal(x) : (creal(x) - t) + sr;
	sr = t;
	t = si + cimag(x);
	ci += fabs(si) >= fabs(cimag(x)) ? (si - t) + cimag(x) : (cimag(x) - t) + si;
	si = t;
	*s = sr + I * si, *c = cr + I * ci;
}

int alis_dev_farfield(alis_dev_world *D, int id, int q, const alis_dev_ff_geom *g, int n,
                      const alis_dev_ff_col *cols, double *S)
{
	phs *Q = getPhaser(D, id);
	if (!Q) return fail("no such phaser");
	if (q < 0 || q >= Q->nf) return fail("phaser frequency out of range");
	const int ni = Q->iMax-Q->iMin+1, nj = Q->jMax-Q->jMin+1, nk = Q->kMax-Q->kMin+1;
	int side = 0;
	for (int f = 0; f < Q->nface; f++) side |= Q->kind[f] < 4;
	if (side) {
		const int lo = g->symZ ? MIN(Q->kMin, -Q->kMax) : Q->kMin, hi = g->symZ ? MAX(Q->kMax, 1 - Q->kMin) : Q->kMax;
		if (!g->T || g->kLo > lo || g->kLo + g->nz - 1 < hi) return fail("far field: profiles do not cover the phaser box");
		for (int c = 0; c < n; c++)
			for (int p = 0; p < 2; p++) {
				if (cols[c].slot[p] < 0 || cols[c].slot[p] >= g->nslot) return fail("far field: profile slot out of range");
				for (int t = 0; t < 6; t++) if (cols[c].src[p][t] >= 6) return fail("far field: no such profile component");
			}
	}
	settle(Q);
	memset(S, 0, (size_t)n * 96 * sizeof(double));
	/* T components paired with the face components (E1 E2 H1 H2) of x and y faces */
	static const int tX[4] = {5, 4, 2, 1}, tY[4] = {3, 5, 0, 2};
	for (int col = 0; col < n; col++) {
		const alis_dev_ff_col *c = &cols[col];
		for (int f = 0; f < Q->nface; f++) {
			const int kind = Q->kind[f];
			for (int comp = 0; comp < 4; comp++) {
				const long base = 2 * (((long)q*4 + comp) * Q->ncell + Q->off[f]);
#define FV(u, w, nw) (Q->dft64 ? Q->accD[base + 2*((long)(u)*(nw)+(w))] + I * Q->accD[base + 2*((long)(u)*(nw)+(w))+1] \
                               : Q->accF[base + 2*((long)(u)*(nw)+(w))] + I * Q->accF[base + 2*((long)(u)*(nw)+(w))+1])
				const int blockA = comp == 0 || comp == 3;
				for (int pol = 0; pol < (kind < 4 ? 2 : 1); pol++) {
					dcx s = 0, e = 0;
					if (kind >= 4) {          /* F[i][j]: Y (node) with X (half) or X (node) with Y (half) */
						for (int i = 0; i < ni; i++) {
							dcx r = 0;
							for (int j = 0; j < nj; j++)
								r += FV(i, j, nj) * (blockA ? ffPhase(0, g->dyE, g->yN, j, nj, c->ky, g->symY, g->cosky)
								                            : ffPhase(1, g->dyH, g->yH, j, nj, c->ky, g->symY, g->cosky));
							ffAdd(&s, &e, r * (blockA ? ffPhase(1, g->dxH, g->xH, i, ni, c->kx, g->symX, g->coskx)
							                          : ffPhase(0, g->dxE, g->xN, i, ni, c->kx, g->symX, g->coskx)));
						}
					}
					else if (kind < 2) {      /* F[j][k]: Y (half) with Z_N, or Y (node) with Z_H */
						for (int j = 0; j < nj; j++) {
							dcx r = 0;
							for (int k = 0; k < nk; k++)
								r += FV(j, k, nk) * ffProfile(!blockA, g, c, pol, tX[comp], Q->kMin, k, nk);
							ffAdd(&s, &e, r * (blockA ? ffPhase(1, g->dyH, g->yH, j, nj, c->ky, g->symY, g->cosky)
							                          : ffPhase(0, g->dyE, g->yN, j, nj, c->ky, g->symY, g->cosky)));
						}
					}
					else {                    /* F[i][k]: X (node) with Z_H, or X (half) with Z_N */
						for (int i = 0; i < ni; i++) {
							dcx r = 0;
							for (int k = 0; k < nk; k++)
								r += FV(i, k, nk) * ffProfile(blockA, g, c, pol, tY[comp], Q->kMin, k, nk);
							ffAdd(&s, &e, r * (blockA ? ffPhase(0, g->dxE, g->xN, i, ni, c->kx, g->symX, g->coskx)
							                          : ffPhase(1, g->dxH, g->xH, i, ni, c->kx, g->symX, g->coskx)));
						}
					}
#undef FV
					s += e;
					double *o = S + 2 * ((((long)col * 6 + kind) * 4 + comp) * 2 + pol);
					o[0] = creal(s), o[1] = cimag(s);
				}
			}
		}
	}
	return 0;
}

int alis_dev_plane_sums(alis_dev_world *D, int nterm, const int *terms, const double *w, long nw, double *out, long nout)
{
	if (nterm < 1 || nw < 1 || nout < 1) return fail("plane sums: no terms or no buffers");
	const alis_dev_world_desc *W = &D->W;
	const int lo[3] = {W->iMIN - 1, W->jMIN - 1, W->kMIN - 1}, hi[3] = {W->iMAX + 1, W->jMAX + 1, W->kMAX + 1};
	for (int n = 0; n < nterm; n++) {
		const int *u = terms + 11 * n;
		const int nrm = u[0], row = u[4], inner = 3 - u[0] - u[4];
		if (nrm < 0 || nrm > 2 || row < 0 || row > 2 || row == nrm || u[2] < 0 || u[2] > 2 || u[3] < 0 || u[3] > 2 || u[6] < u[5] || u[8] < u[7]
		    || u[1] < lo[nrm] || u[1] + 1 > hi[nrm] || u[5] < lo[row] || u[6] > hi[row] || u[7] < lo[inner] || u[8] > hi[inner]
		    || u[9] < 0 || u[9] + (u[8] - u[7] + 1) > nw || u[10] < 0 || u[10] + (u[6] - u[5] + 1) > nout)
			return fail("plane sums: term outside the grid or the buffers");
	}
	for (long c = 0; c < nout; c++) out[c] = 0;
	for (int n = 0; n < nterm; n++) {
		if (D->fp64) planeSums_d(D, terms + 11 * n, w, out);
		else planeSums_f(D, terms + 11 * n, w, out);
	}
	return 0;
}

int alis_dev_absorption(alis_dev_world *D, int n, int i0, int i1, int j0, int j1, int k0, int k1,
                        const double *dzE, const double *dzH, double *out)
{
	if (n < 0 || n >= D->W.nobj) return fail("absorption: no such dispersive object");
	if (!D->hasEps) return fail("absorption: the world has no permittivity arrays");
	const alis_dev_ade *o = &D->obj[n];
	const int ci0 = MAX(D->W.iMIN, o->i0), ci1 = MIN(D->W.iMAX, o->i1), cj0 = MAX(D->W.jMIN, o->j0), cj1 = MIN(D->W.jMAX, o->j1);
	const int ck0 = MAX(D->W.kMIN, o->k0), ck1 = MIN(D->W.kMAX, o->k1);
	if (i0 > i1 || j0 > j1 || k0 > k1 || i0 < ci0 || i1 > ci1 || j0 < cj0 || j1 > cj1 || k0 < ck0 || k1 > ck1)
		return fail("absorption: columns outside the object's box");
	const int b[6] = {i0, i1, j0, j1, k0, k1};
	if (D->fp64) absorption_d(D, n, b, dzE, dzH, out);
	else absorption_f(D, n, b, dzE, dzH, out);
	return 0;
}

/* Second-order coupling (src/update.c:344-411) between worlds of one grid; the signal's own
 * E never feeds the coupling (it is refused as a pump). */
int alis_dev_three_wave(alis_dev_world *S, alis_dev_world *P1, alis_dev_world *P2, int n, const float *d, float alt)
{
	if (S == P1 || S == P2) return fail("three-wave coupling: the signal world is one of the pumps");
	for (int p = 0; p < 2; p++) {
		const alis_dev_world *P = p ? P2 : P1;
		if (P->fp64 != S->fp64 || P->W.complexField != S->W.complexField || P->W.iMIN != S->W.iMIN || P->W.iMAX != S->W.iMAX
		    || P->W.jMIN != S->W.jMIN || P->W.jMAX != S->W.jMAX || P->W.kMIN != S->W.kMIN || P->W.kMAX != S->W.kMAX)
			return fail("three-wave coupling: the pump and signal worlds differ in grid or precision");
	}
	if (n < 0 || n >= S->W.nobj || !S->hasEps) return fail("three-wave coupling: no such dispersive object");
	const alis_dev_ade *o = &S->obj[n];
	const int b[6] = {o->i0, o->i1 - 1, o->j0, o->j1 - 1, o->k0, o->k1 - 1};
	if (b[0] < S->W.iMIN || b[1] > S->W.iMAX || b[2] < S->W.jMIN || b[3] > S->W.jMAX || b[4] < S->W.kMIN || b[5] > S->W.kMAX)
		return fail("three-wave coupling: object box outside the grid");
	if (S->fp64) threeWave_d(S, P1, P2, n, d, alt, b);
	else threeWave_f(S, P1, P2, n, d, alt, b);
	return 0;
}

/* ---- probe spectra (src/output.c:177-230) ----
 * The running DFT of writeSpectrum in double with compensated (Neumaier) summation: every
 * sample is added to every bin as it arrives, in call order (the device stages samples and
 * adds them later in the same order). */
struct alis_dev_spectrum {
	int np, nb, nTop;
	double *sum, *cmp;          /* [probe][bin][re, im] */
};

static void neumaier(double *s, double *c, double x)
{
	const double t = *s + x;
	*c += fabs(*s) >= fabs(x) ? (*s - t) + x : (x - t) + *s;
	*s = t;
}

int alis_dev_spectrum_create(int nprobe, int nbin, int nTop, alis_dev_spectrum **out)
{
	if (nprobe < 1 || nbin < 1) return fail("spectrum: no probes or no frequency bins");
	alis_dev_spectrum *S = calloc(1, sizeof *S);
	S->np = nprobe, S->nb = nbin, S->nTop = nTop;
	S->sum = calloc(2L * nprobe * nbin, sizeof(double));
	S->cmp = calloc(2L * nprobe * nbin, sizeof(double));
	*out = S;
	return 0;
}

int alis_dev_spectrum_add(alis_dev_spectrum *S, double tau, const double *v)
{
	for (int b = 0; b < S->nb; b++) {
		const double ph = 2 * M_PI * tau * (S->nTop - b), cs = cos(ph), sn = sin(ph);
		for (int p = 0; p < S->np; p++) {
			const long e = 2 * ((long)p * S->nb + b);
			neumaier(&S->sum[e], &S->cmp[e], v[p] * cs);
			neumaier(&S->sum[e + 1], &S->cmp[e + 1], v[p] * sn);
		}
	}
	return 0;
}

int alis_dev_spectrum_read(alis_dev_spectrum *S, double *out)
{
	for (long e = 0; e < 2L * S->np * S->nb; e++) out[e] = S->sum[e] + S->cmp[e];
	return 0;
}

void alis_dev_spectrum_release(alis_dev_spectrum *S)
{
	if (!S) return;
	free(S->sum), free(S->cmp), free(S);
}

int alis_dev_synchronize(alis_dev_world *D) { (void)D; return 0; }
int alis_dev_timing(alis_dev_world *D, int enable) { (void)D, (void)enable; return 0; }
int alis_dev_timing_read(alis_dev_world *D, int h, double *ms, long *launches) { (void)D, (void)h; *ms = 0, *launches = 0; return 0; }
long alis_dev_bytes(alis_dev_world *D) { return D->bytes; }
double alis_dev_eps_bytes(const alis_dev_world *D) { return D->hasEps ? 3.0 * (D->fp64 ? 8 : 4) : 0.0; }

/* ---- batches: the worlds one after the other (the oracle applies strips and sources as
 * they are recorded, so a batch half step is core of every world, then their ghosts) ---- */
struct alis_dev_batch {
	int n;
	alis_dev_world **W;
};

int alis_dev_batch_create(alis_dev_world *const *D, int n, alis_dev_batch **out)
{
	if (n < 1) return fail("empty batch");
	alis_dev_batch *B = calloc(1, sizeof *B);
	B->n = n;
	B->W = malloc(n * sizeof *B->W);
	memcpy(B->W, D, n * sizeof *B->W);
	*out = B;
	return 0;
}

int alis_dev_batch_core(alis_dev_batch *B, int h)
{
	for (int a = 0; a < B->n; a++) if (alis_dev_core(B->W[a], h)) return -1;
	return 0;
}

int alis_dev_batch_post(alis_dev_batch *B, int h)
{
	for (int a = 0; a < B->n; a++) if (alis_dev_ghosts(B->W[a], h)) return -1;
	return 0;
}

void alis_dev_batch_release(alis_dev_batch *B)
{
	if (!B) return;
	free(B->W);
	free(B);
}
