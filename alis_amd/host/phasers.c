/*
 * Phasers: surface running DFTs on the faces of a box (src/phaser.c:21-295).  The
 * accumulators live on the device; flux integrals read them in double precision and sum
 * with Neumaier compensation.
 */
#include <string.h>
#include "internal.h"
#include "phaserview.h"

/* Source used for spectral normalisation (src/phaser.c:25-26): the E or else the H source
 * whose duration is W's step count, looked up in W and then in W's first TF/SF auxiliary
 * world. */
source normalisingSource(world W)
{
	world look[2] = {W, W->S ? W->S->W : 0};
	for (int a = 0; a < 2 && look[a]; a++) {
		if (look[a]->SE && look[a]->SE->N == W->N) return look[a]->SE;
		if (look[a]->SH && look[a]->SH->N == W->N) return look[a]->SH;
	}
	return 0;
}

static cfield *faceOf(phaser P, int kind);

/* The tangential components a face normal to `axis` records, as E1, E2, H1, H2
 * (src/phaser.c:59-64): x faces Ey Ez Hy Hz, y faces Ez Ex Hz Hx, z faces Ex Ey Hx Hy. */
static void tangential(cfield *F, int axis, float complex ****c[4])
{
	float complex ****e[3] = {&F->Ex, &F->Ey, &F->Ez}, ****h[3] = {&F->Hx, &F->Hy, &F->Hz};
	const int a1 = (axis + 1) % 3, a2 = (axis + 2) % 3;
	c[0] = e[a1], c[1] = e[a2], c[2] = h[a1], c[3] = h[a2];
}

/* The face arrays of face `kind` (xMin xMax yMin yMax zMin zMax) over the box of Q with the
 * face's axis fixed at its lower or upper index. */
static void makeFace(phaser Q, int kind)
{
	const int axis = kind / 2;
	int lo[3] = {Q->iMin, Q->jMin, Q->kMin}, hi[3] = {Q->iMax, Q->jMax, Q->kMax};
	if (kind % 2) lo[axis] = hi[axis];
	else hi[axis] = lo[axis];
	float complex ****c[4];
	tangential(faceOf(Q, kind), axis, c);
	for (int q = 0; q < 4; q++) *c[q] = makeComplexField(lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]);
}

/* 1 / |sum over steps s = 1 .. N-1 of the source waveform at s dt times e^{2 pi i f s dt}|
 * (float complex, as src/phaser.c:38-42), per unit of wavelength for line and plane
 * sources (:43-44). */
static float inverseSpectrum(world W, source S, float f)
{
	float complex sum = 0;
	const float complex step = 2 * PI * I * f * W->dt;
	for (int s = 1; s < W->N; s++) sum += S->waveForm(W, S, W->dt * s, S->phase) * cexpf(step * s);
	float A = 1 / cabsf(sum);
	if (W->sourceType == 2) A /= sqrtf(f * S->wavelength);
	if (W->sourceType == 3) A /= f * S->wavelength;
	return A;
}

/* Phasers at wavelengths LMin..LMax (frequencies equally spaced in 1/L by dL) on the faces
 * of a box; the array ends with an f = 0 entry (src/phaser.c:21-66).  A normalises the
 * accumulated fields to a unit source; a box that spans whole periods of a periodic x or y
 * axis records how many (xNUM, yNUM) and the Bloch phase per period.  Faces exist where the
 * world's surface beyond them is a PML, and a lower face only when the box has depth. */
phaser (createPhasers)(world W, float LMin, float LMax, float dL, float xMin, float xMax, float yMin, float yMax, float zMin, float zMax, ...)
{
	const int nMin = floor(LMin/dL), nMax = ceil(LMax/dL), count = nMax - nMin + 1;
	const float boxLo[3] = {xMin, yMin, zMin}, boxHi[3] = {xMax, yMax, zMax};
	source S = normalisingSource(W);
	phaser P;
	MALLOC(P, 0, count + 1);
	for (int n = 0; n < count; n++) {
		phaser Q = P + n;
		Q->W = W;
		Q->N = nMin * nMax * dL / W->dt;
		Q->f = (nMax-n) / (dL * nMin * nMax);
		Q->A = 2.0 / (W->N ? W->N : P[0].N);
		Q->devId = -1;
		int *lo[3] = {&Q->iMin, &Q->jMin, &Q->kMin}, *hi[3] = {&Q->iMax, &Q->jMax, &Q->kMax};
		for (int a = 0; a < 3; a++) {
			const axis X = axisOf(W, a);
			*lo[a] = MAX(*X.Min, (*X.node)(W, boxLo[a]));
			*hi[a] = MIN(*X.Max, (*X.node)(W, boxHi[a]));
		}
		if (S && W->N) Q->A = inverseSpectrum(W, S, Q->f);
		for (int a = 0; a < 2; a++) {
			const axis X = axisOf(W, a);
			const int periodic = *X.maxSur == SYM || *X.maxSur == PBC || *X.maxSur == BBC;
			if (!(boxHi[a] - boxLo[a] < INF && boxLo[a] < *X.lo && *X.hi < boxHi[a] && periodic)) continue;
			*(a ? &Q->yNUM : &Q->xNUM) = floor((boxHi[a] - boxLo[a]) / (*X.hi - *X.lo) + 0.5);
			*(a ? &Q->expiky : &Q->expikx) = *X.maxSur == BBC ? *X.cos + I * *X.sin : 1;
		}
		for (int kind = 0; kind < 6; kind++) {
			const axis X = axisOf(W, kind / 2);
			const int upper = kind % 2, flat = *lo[kind / 2] == *hi[kind / 2];
			if ((upper ? *X.maxSur : *X.minSur) == PML && (upper || !flat)) makeFace(Q, kind);
		}
	}
	P[0].count = count;
	P[count].devId = -1;
	alisRegisterPhaser(P);
	return P;
}

phaser (createPhaser)(world W, float L, float xMin, float xMax, float yMin, float yMax, float zMin, float zMax, ...)
{
	return (createPhasers)(W, L, L, L, xMin, xMax, yMin, yMax, zMin, zMax);
}

/* Phaser of the array nearest to `wavelength` in frequency (src/phaser.c:75-80). */
phaser choosePhaser(phaser P, float wavelength)
{
	do P++;
	while (1/wavelength < P->f);
	return (1/wavelength - P->f < (P-1)->f - 1/wavelength) ? P : P-1;
}

/* Make the phaser array B device-resident (its world committed, accumulators uploaded from
 * the host cfields when it is not registered yet). */
void alisPhaserDevice(phaser B)
{
	alisCommit(B->W);
	if (B->devId < 0) {
		alis_dev_phaser_desc d;
		describePhaser(B, &d);
		int rc = alis_dev_phaser_add(B->W->dev, &d, &B->devId);
		free(d.cf);
		alisFatal("phaser registration", rc);
		B->faceFluxValid = 0;
	}
}

/* Phase factors e^{2 pi i f t} of the frequencies of P (to the end of its array) at the E
 * and H times of W (src/phaser.c:97-98), as re/im pairs in eE/eH: in the reference's float
 * arithmetic for FP32 accumulators, in double (with the double pi) for FP64 accumulators.
 * updatePhaser accumulates E at W->E->t and the face-averaged H at W->H->t (src/phaser.c:83-110). */
static void phaserFactors(world W, phaser P, int nq, double *eE, double *eH)
{
	for (int q = 0; q < nq; q++) {
		if (W->devDftPrec) {
			double aE = 2 * M_PI * (double)P[q].f * (double)W->E->t, aH = 2 * M_PI * (double)P[q].f * (double)W->H->t;
			eE[2*q] = cos(aE), eE[2*q+1] = sin(aE);
			eH[2*q] = cos(aH), eH[2*q+1] = sin(aH);
		}
		else {
			float complex e = cexpf(2 * PI * I * P[q].f * W->E->t);
			float complex h = cexpf(2 * PI * I * P[q].f * W->H->t);
			eE[2*q] = crealf(e), eE[2*q+1] = cimagf(e);
			eH[2*q] = crealf(h), eH[2*q+1] = cimagf(h);
		}
	}
}

static phaser checkedBase(world W, phaser P, const char *who)
{
	phaser B = phaserBase(P);
	if (!B) printf("%s: unknown phaser\n", who), exit(1);
	if (B->W != W) printf("%s: the phaser belongs to another world\n", who), exit(1);
	alisPhaserDevice(B);
	return B;
}

void updatePhaser(world W, phaser P)
{
	phaser B = checkedBase(W, P, "updatePhaser");
	int q0 = (int)(P - B), nq = B->count - q0;
	static double *eE, *eH;
	static int cap;
	if (nq > cap) cap = nq, eE = realloc(eE, 2*cap*sizeof(double)), eH = realloc(eH, 2*cap*sizeof(double));
	phaserFactors(W, P, nq, eE, eH);
	alisFatal("phaser update", alis_dev_phaser_update(W->dev, B->devId, q0, nq, eE, eH));
	B->hostStale = 1;
	B->viewValid = B->faceFluxValid = 0;
}

/* updatePhaser(W[a], P[a]) for a = 0..n-1 in one launch (the phasers' worlds share their
 * precision; they join one stream). */
void updatePhaserArrays(world *W, phaser *P, int n)
{
	if (n < 1) return;
	alis_dev_world **D = malloc(n * sizeof *D);
	int *id = malloc(3 * n * sizeof(int)), *q0 = id + n, *nq = q0 + n;
	double **eE = malloc(2 * n * sizeof(double *)), **eH = eE + n;
	for (int a = 0; a < n; a++) {
		phaser B = checkedBase(W[a], P[a], "updatePhaserArrays");
		D[a] = W[a]->dev, id[a] = B->devId, q0[a] = (int)(P[a] - B), nq[a] = B->count - q0[a];
		eE[a] = malloc(4 * (nq[a] ? nq[a] : 1) * sizeof(double)), eH[a] = eE[a] + 2 * nq[a];
		phaserFactors(W[a], P[a], nq[a], eE[a], eH[a]);
		B->hostStale = 1;
		B->viewValid = B->faceFluxValid = 0;
	}
	alisFatal("phaser update", alis_dev_phaser_update_many(D, id, q0, nq, (const double *const *)eE, (const double *const *)eH, n));
	for (int a = 0; a < n; a++) free(eE[a]);
	free(eE), free(id), free(D);
}

void deletePhaser(phaser P)
{
	phaser B = P;
	if (B->devId >= 0 && B->W->dev) alis_dev_phaser_remove(B->W->dev, B->devId);
	alisForgetPhaser(B);
	free(B->view), free(B->faceFlux);
	for (; P->f; P++) {
		cfield *F[6] = { &P->xMin, &P->xMax, &P->yMin, &P->yMax, &P->zMin, &P->zMax };
		for (int a = 0; a < 6; a++) {
			if (F[a]->Ex) removeField(F[a]->Ex, P->iMin, a == 3 ? P->jMax : P->jMin, a == 5 ? P->kMax : P->kMin);
			if (F[a]->Ey) removeField(F[a]->Ey, a == 1 ? P->iMax : P->iMin, a == 3 ? P->jMax : P->jMin, a == 5 ? P->kMax : P->kMin);
			if (F[a]->Ez) removeField(F[a]->Ez, a == 1 ? P->iMax : P->iMin, a == 3 ? P->jMax : P->jMin, P->kMin);
			if (F[a]->Hx) removeField(F[a]->Hx, P->iMin, a == 3 ? P->jMax : P->jMin, a == 5 ? P->kMax : P->kMin);
			if (F[a]->Hy) removeField(F[a]->Hy, a == 1 ? P->iMax : P->iMin, a == 3 ? P->jMax : P->jMin, a == 5 ? P->kMax : P->kMin);
			if (F[a]->Hz) removeField(F[a]->Hz, a == 1 ? P->iMax : P->iMin, a == 3 ? P->jMax : P->jMin, P->kMin);
		}
		if (P->dE) free(&P->dE[P->kMIN]), free(&P->dH[P->kMIN]), free(&P->ex[P->kMIN]), free(&P->ez[P->kMIN]);
	}
	free(B);
}


/* ---- double-precision view of the accumulators ---- */

static cfield *faceOf(phaser P, int kind)
{
	cfield *F[6] = { &P->xMin, &P->xMax, &P->yMin, &P->yMax, &P->zMin, &P->zMax };
	return F[kind];
}

void phaserFaces(phaser B, phaserGeom *g)
{
	int ni = B->iMax - B->iMin + 1, nj = B->jMax - B->jMin + 1, nk = B->kMax - B->kMin + 1;
	long off = 0;
	for (int kind = 0; kind < 6; kind++) {
		cfield *F = faceOf(B, kind);
		g->on[kind] = F->Ex || F->Ey || F->Ez;
		g->off[kind] = off;
		g->n2[kind] = kind < 4 ? nk : nj;
		if (g->on[kind]) off += kind < 2 ? (long)nj * nk : kind < 4 ? (long)ni * nk : (long)ni * nj;
	}
	g->ncell = off;
}

/* Accumulators of the whole array of B as doubles in the device read layout. */
const double *phaserDoubles(phaser B, phaserGeom *g)
{
	phaserFaces(B, g);
	long need = (long)B->count * 4 * g->ncell * 2;
	if (B->viewValid && B->view) return B->view;
	if (!B->view) B->view = calloc(need ? need : 1, sizeof(double));
	if (B->devId >= 0 && B->W->dev) {
		alisFatal("phaser read", alis_dev_phaser_read(B->W->dev, B->devId, B->view, need));
	}
	else {
		for (int q = 0; q < B->count; q++) {
			phaser P = B + q;
			for (int kind = 0; kind < 6; kind++) {
				if (!g->on[kind]) continue;
				cfield *F = faceOf(P, kind);
				float complex ***c[4];
				if (kind < 2) c[0] = F->Ey, c[1] = F->Ez, c[2] = F->Hy, c[3] = F->Hz;
				else if (kind < 4) c[0] = F->Ez, c[1] = F->Ex, c[2] = F->Hz, c[3] = F->Hx;
				else c[0] = F->Ex, c[1] = F->Ey, c[2] = F->Hx, c[3] = F->Hy;
				int i = kind == 1 ? P->iMax : P->iMin, j = kind == 3 ? P->jMax : P->jMin, k = kind == 5 ? P->kMax : P->kMin;
				long cells = kind < 2 ? (long)(P->jMax-P->jMin+1)*(P->kMax-P->kMin+1) : kind < 4 ? (long)(P->iMax-P->iMin+1)*(P->kMax-P->kMin+1) : (long)(P->iMax-P->iMin+1)*(P->jMax-P->jMin+1);
				for (int comp = 0; comp < 4; comp++) {
					float complex *src = &c[comp][i][j][k];
					double *dst = B->view + 2 * (((long)q * 4 + comp) * g->ncell + g->off[kind]);
					for (long a = 0; a < cells; a++) dst[2*a] = crealf(src[a]), dst[2*a+1] = cimagf(src[a]);
				}
			}
		}
	}
	B->viewValid = 1;
	return B->view;
}

void neumaierAdd(double *sum, double *comp, double x)
{
	double t = *sum + x;
	if (fabs(*sum) >= fabs(x)) *comp += (*sum - t) + x;
	else *comp += (x - t) + *sum;
	*sum = t;
}


/* Trapezoid weights of the E nodes lo..hi along one axis (half weight at both ends). */
static void trapezoid(double *w, const float *d, int lo, int hi)
{
	for (int n = lo; n <= hi; n++) w[n - lo] = (n == lo || n == hi) ? 0.5 * d[n] : d[n];
}

static double rowSum(const double *rows, int a)
{
	double r = 0;
	for (int l = 0; l < 64; l++) r += rows[(long)a * 64 + l];
	return r;
}

/* A face's flux from its inner trapezoid sums (see faceFluxes): rows holds 64 partial sums
 * per row, cols one sum per column; the outer sums weight them by the H cell widths. */
static double faceFluxSums(phaser P, int kind, const double *rows, const double *cols)
{
	world W = P->W;
	double s = 0, c = 0;
	if (kind < 2) {          /* a = j, b = k: E1 x H2 along k (rows), E2 x H1 along j (columns) */
		for (int j = P->jMin+1; j <= P->jMax; j++) neumaierAdd(&s, &c, W->CH->dy[j] * rowSum(rows, j - P->jMin));
		for (int k = P->kMin+1; k <= P->kMax; k++) neumaierAdd(&s, &c, W->CH->dz[k] * -cols[k - P->kMin]);
	}
	else if (kind < 4) {     /* a = i, b = k: E1 x H2 along i (columns), E2 x H1 along k (rows) */
		for (int k = P->kMin+1; k <= P->kMax; k++) neumaierAdd(&s, &c, W->CH->dz[k] * cols[k - P->kMin]);
		for (int i = P->iMin+1; i <= P->iMax; i++) neumaierAdd(&s, &c, W->CH->dx[i] * -rowSum(rows, i - P->iMin));
	}
	else {                   /* a = i, b = j: E1 x H2 along j (rows), E2 x H1 along i (columns) */
		for (int i = P->iMin+1; i <= P->iMax; i++) neumaierAdd(&s, &c, W->CH->dx[i] * rowSum(rows, i - P->iMin));
		for (int j = P->jMin+1; j <= P->jMax; j++) neumaierAdd(&s, &c, W->CH->dy[j] * -cols[j - P->jMin]);
	}
	return s + c;
}

/* The inner sums of the flux terms on host accumulators v (the device read layout), in the
 * form and order of the engine's alis_dev_phaser_flux_sums: 64 in-order lane sums per row
 * term, in-order column sums, each cell adding w * (xr yr + xi yi). */
static void fluxSumsHost(const double *v, const phaserGeom *g, int q, int terms[][8], int nterm, const double *w,
                         double *out, long nout)
{
	memset(out, 0, nout * sizeof *out);
	for (int n = 0; n < nterm; n++) {
		const int *u = terms[n];
		const long na = u[2], nb = u[3];
		const double *x = v + 2 * (((long)q * 4 + u[4]) * g->ncell + u[1]);
		const double *y = v + 2 * (((long)q * 4 + u[5]) * g->ncell + u[1]);
		double *o = out + u[7];
		for (long a = 0; a < na; a++)
			for (long b = 0; b < nb; b++) {
				const double *p = x + 2 * (a * nb + b), *r = y + 2 * (a * nb + b);
				const double t = p[0] * r[0] + p[1] * r[1];
				if (u[0] == 0) o[a * 64 + b % 64] += w[u[6] + b] * t;
				else o[b] += w[u[6] + a] * t;
			}
	}
}

/* The six faces' fluxes (src/phaser.c:142-223): trapezoid sums over each face's E nodes
 * weighted by the E and H cell widths, as one row term (inner axis = the face's fast axis)
 * and one column term per face.  With host accumulators v: frequency q of P, the inner sums
 * computed here.  Otherwise every frequency of the array, face[6 n + kind] for frequency n,
 * the inner sums from one engine launch, which reads the accumulators once and returns
 * about 64 + 1 doubles per face row and frequency.  Either way the same bits. */
static void faceFluxes(phaser P, phaser B, const phaserGeom *g, int q, const double *v, double *face)
{
	world W = P->W;
	const int ni = B->iMax-B->iMin+1, nj = B->jMax-B->jMin+1, nk = B->kMax-B->kMin+1;
	double *w = malloc((size_t)(ni + nj + nk) * sizeof(double));
	const int wx = 0, wy = ni, wz = ni + nj;
	trapezoid(w + wx, W->CE->dx, B->iMin, B->iMax);
	trapezoid(w + wy, W->CE->dy, B->jMin, B->jMax);
	trapezoid(w + wz, W->CE->dz, B->kMin, B->kMax);
	int terms[12][8], nterm = 0;
	long rowOff[6] = {0}, colOff[6] = {0}, nout = 0;
	for (int kind = 0; kind < 6; kind++) {
		face[kind] = 0;
		if (!g->on[kind]) continue;
		const int na = kind < 4 ? (kind < 2 ? nj : ni) : ni, nb = g->n2[kind], off = (int)g->off[kind];
		/* row term: E1 x H2 on x and z faces, E2 x H1 on y faces; column term the other */
		const int rx = kind < 2 || kind >= 4 ? 0 : 1, ry = kind < 2 || kind >= 4 ? 3 : 2;
		const int rw = kind < 4 ? wz : wy, cw = kind < 2 ? wy : wx;
		rowOff[kind] = nout, nout += (long)na * 64;
		colOff[kind] = nout, nout += nb;
		int *r = terms[nterm++], *t = terms[nterm++];
		r[0] = 0, r[1] = off, r[2] = na, r[3] = nb, r[4] = rx, r[5] = ry, r[6] = rw, r[7] = (int)rowOff[kind];
		t[0] = 1, t[1] = off, t[2] = na, t[3] = nb, t[4] = rx ? 0 : 1, t[5] = rx ? 3 : 2, t[6] = cw, t[7] = (int)colOff[kind];
	}
	const int nq = v ? 1 : B->count;
	for (int n = 0; n < nq; n++)
		for (int kind = 0; kind < 6; kind++) face[6 * n + kind] = 0;
	if (!nterm) { free(w); return; }
	double *out = malloc((size_t)nq * nout * sizeof(double));
	if (v) fluxSumsHost(v, g, q, terms, nterm, w, out, nout);
	else alisFatal("phaser flux", alis_dev_phaser_flux_sums_range(W->dev, B->devId, 0, nq, nterm, &terms[0][0], w, ni + nj + nk, out, nout));
	for (int n = 0; n < nq; n++)
		for (int kind = 0; kind < 6; kind++)
			if (g->on[kind]) face[6 * n + kind] = faceFluxSums(P, kind, out + (long)n * nout + rowOff[kind], out + (long)n * nout + colOff[kind]);
	free(out);
	free(w);
}

/* Net outward flux of phaser P normalised by A^2; 2-D worlds divide by the unit length of
 * the flat axis, symmetry planes double the result (src/phaser.c:228-245). */
double phaserFluxD(phaser P)
{
	phaser B = phaserBase(P);
	if (!B) return 0;
	phaserGeom g;
	int q = (int)(P - B);
	world W = P->W;
	double face[6];
	if (B->devId >= 0 && W->dev) {
		/* every frequency of the array at once, kept until the accumulators change */
		if (!B->faceFluxValid) {
			phaserFaces(B, &g);
			if (!B->faceFlux) B->faceFlux = malloc((size_t)6 * B->count * sizeof(double));
			faceFluxes(P, B, &g, q, 0, B->faceFlux);
			B->faceFluxValid = 1;
		}
		memcpy(face, B->faceFlux + 6 * q, sizeof face);
	}
	else {
		const double *v = phaserDoubles(B, &g);
		faceFluxes(P, B, &g, q, v, face);
	}
	double s = 0, c = 0;
	neumaierAdd(&s, &c, face[1]);
	neumaierAdd(&s, &c, -face[0]);
	neumaierAdd(&s, &c, face[3]);
	neumaierAdd(&s, &c, -face[2]);
	neumaierAdd(&s, &c, face[5]);
	neumaierAdd(&s, &c, -face[4]);
	double Sum = s + c;
	if (W->Dom.x.Min==W->Dom.x.Max) Sum /= W->CH->dx[P->iMax]+W->CH->dx[P->iMax-1];
	if (W->Dom.y.Min==W->Dom.y.Max) Sum /= W->CH->dy[P->jMax]+W->CH->dy[P->jMax-1];
	if (W->Dom.z.Min==W->Dom.z.Max) Sum /= W->CH->dz[P->kMax]+W->CH->dz[P->kMax-1];
	if (W->xMinSur==SYM) Sum *= 2;
	if (W->yMinSur==SYM) Sum *= 2;
	if (W->zMinSur==SYM) Sum *= 2;
	return Sum * (double)P->A * (double)P->A;
}

float phaserPoyntingOut(phaser P) { return phaserFluxD(P); }
float phaserPoynting(phaser P) { return fabsf(phaserPoyntingOut(P)); }
float phaserPoyntingIn(phaser P) { return -phaserPoyntingOut(P); }


