/*
 * Observers (src/output.c:13-101, 177-445): probes, world maxima and integrals, probe
 * spectra, Poynting fluxes through planes and absorbed power.  Probes read host mirrors
 * refreshed over a small box; fluxes and absorption are summed on the device.
 */
#include <string.h>
#include <sys/stat.h>
#include "internal.h"
#include "phaserview.h"

/* Value of field F at the grid node nearest to (x, y, z) (src/output.c:50-53). */
float get(world W, field F, float x, float y, float z)
{
	int i = W->xtoi(W, x), j = W->ytoj(W, y), k = W->ztok(W, z);
	if (W->dev && W->hostStale) {
		int readsAux = F==SoEx || F==SoEy || F==SoEz || F==SoEE || F==ScEx || F==ScEy || F==ScEz || F==ScEE;
		if (readsAux || W->isAux) syncWorld(W);
		else {
			/* the node, mirrored across a symmetry plane when its index is negative (as the
			 * accessors do): the component accessors read one array at the node and its upper
			 * neighbours; the others read within two cells of it, in any array */
			int ci = i < 0 && W->xMinSur==SYM ? -i : i, cj = j < 0 && W->yMinSur==SYM ? -j : j, ck = k < 0 && W->zMinSur==SYM ? -k : k;
			int isH, m, axis;
			if (fieldComponent(F, &isH, &m, &axis) || rawComponent(F, &isH, &m, &axis))
				alisProbe(W, ci, ci+1, cj, cj+1, ck, ck+1, 1 << (isH * 6 + m * 3 + axis));
			else alisProbe(W, ci-2, ci+2, cj-2, cj+2, ck-2, ck+2, ALIS_DEV_PROBE_ALL);
		}
	}
	return F(W, i, j, k);
}

float worldMax(world W, field F)
{
	syncWorld(W);
	float Max = 0;
	#pragma omp parallel for reduction(max:Max)
	for (int i = W->iMin; i < W->iMax; i++) for (int j = W->jMin; j < W->jMax; j++) for (int k = W->kMin; k < W->kMax; k++) {
		float f = fabsf(F(W, i, j, k));
		if (Max < f) Max = f;
	}
	return Max;
}

/* Trapezoid integral of F over the domain (src/output.c:72-101), in double. */
float worldSum(world W, field F)
{
	syncWorld(W);
	double Sum = 0;
	#pragma omp parallel for reduction(+:Sum)
	for (int i = W->iMin; i <= W->iMax; i++) {
		double wi = (i == W->iMin || i == W->iMax) ? 0.5 * W->CE->dx[i] : W->CE->dx[i];
		double sy = 0;
		for (int j = W->jMin; j <= W->jMax; j++) {
			double wj = (j == W->jMin || j == W->jMax) ? 0.5 * W->CE->dy[j] : W->CE->dy[j];
			double sz = 0;
			for (int k = W->kMin; k <= W->kMax; k++) {
				double wk = (k == W->kMin || k == W->kMax) ? 0.5 * W->CE->dz[k] : W->CE->dz[k];
				sz += wk * F(W, i, j, k);
			}
			sy += wj * sz;
		}
		Sum += wi * sy;
	}
	if (W->xMinSur==SYM) Sum *= 2;
	if (W->yMinSur==SYM) Sum *= 2;
	if (W->zMinSur==SYM) Sum *= 2;
	return Sum;
}


/*
 * Probe spectra (writeSpectrum, src/output.c:177-230): the running DFT of the values a
 * program passes each call, over bins n = 0 .. nMax-nMin of the wavelengths
 * dL nMin nMax / (nMax - n), normalised by the spectrum of the world's source, written as
 * "wavelength |S/S0|... arg(S/S0)/pi..." on the N-th call.  One record per (world, name).
 * In the reference precision the sums are the reference's float complex ones, term for
 * term; in the FP32/FP64 precisions the sample sums run on the device in FP64 with
 * compensated summation (alis_dev_spectrum), and the source spectrum in double on the host.
 */
typedef struct spectrum {
	world W;
	char *name;                 /* the name as called, before W->ID is prefixed */
	unsigned long calls;
	int M, nMin, nMax;          /* probes (values per call) and the bin range */
	float dL;
	int f64;
	float complex *acc;         /* reference precision: [1+M][bins], row 0 the source */
	double complex *norm;       /* FP32/FP64: the source's spectrum per bin */
	alis_dev_spectrum *dev;     /* FP32/FP64: the probe sums */
	struct spectrum *next;
} spectrum;

static spectrum *spectra;

/* Drop the spectra of a world that goes away (a later world may reuse its address). */
void alisForgetSpectra(world W)
{
	for (spectrum **p = &spectra; *p;) {
		spectrum *q = *p;
		if (q->W != W) { p = &q->next; continue; }
		*p = q->next;
		alis_dev_spectrum_release(q->dev);
		free(q->name), free(q->acc), free(q->norm), free(q);
	}
}

static spectrum *spectrumOf(world W, const char *name)
{
	for (spectrum *q = spectra; q; q = q->next) if (q->W == W && !strcmp(q->name, name)) return q;
	return 0;
}

/* A new record: bins, the source spectrum and the accumulators of W's precision. */
static spectrum *spectrumNew(world W, const char *name, int N, float LMin, float LMax, int M)
{
	spectrum *q = calloc(1, sizeof *q);
	q->W = W, q->name = strdup(name), q->M = M;
	q->dL = LMin * LMax / (N * W->dt);
	q->nMin = floor(LMin/q->dL), q->nMax = ceil(LMax/q->dL);
	q->f64 = W->devDftPrec != 0;
	const int nF = 1 + q->nMax - q->nMin;
	const float dL = q->dL;
	const int nMin = q->nMin, nMax = q->nMax;
	source S = normalisingSource(W);
	if (!q->f64) {
		q->acc = calloc((size_t)(1 + M) * nF, sizeof(float complex));
		for (int n = 0; S && W->N && n < nF; n++) {
			float f = (nMax-n) / (dL * nMin * nMax);
			float complex s0 = 0;
			for (int m = 1; m < W->N; m++) s0 += S->waveForm(W, S, W->dt*m, S->phase) * cexpf(2*PI*I*f*W->dt*m);
			if (cabsf(s0) < 1E-3) s0 = 0;
			else if (W->sourceType == 2) s0 *= sqrtf(f * S->wavelength);
			else if (W->sourceType == 3) s0 *= f * S->wavelength;
			q->acc[n] = s0;
		}
	}
	else {
		q->norm = calloc(nF, sizeof(double complex));
		for (int n = 0; S && W->N && n < nF; n++) {
			const double f = (nMax - n) / ((double)dL * nMin * nMax);
			double complex s0 = 0;
			for (int m = 1; m < W->N; m++) s0 += (double)S->waveForm(W, S, W->dt*m, S->phase) * cexp(2 * M_PI * I * f * ((double)W->dt * m));
			if (cabs(s0) < 1E-3) s0 = 0;
			else if (W->sourceType == 2) s0 *= sqrt(f * S->wavelength);
			else if (W->sourceType == 3) s0 *= f * S->wavelength;
			q->norm[n] = s0;
		}
		alisFatal("spectrum", alis_dev_spectrum_create(M, nF, nMax, &q->dev));
	}
	q->next = spectra, spectra = q;
	return q;
}

/* Sums per bin of probe p (0: the source's spectrum), reference precision or FP64. */
static double complex spectrumAt(spectrum *q, const double *dev, int p, int n)
{
	const int nF = 1 + q->nMax - q->nMin;
	if (!q->f64) return q->acc[(size_t)p * nF + n];
	if (!p) return q->norm[n];
	return dev[2 * ((size_t)(p-1) * nF + n)] + I * dev[2 * ((size_t)(p-1) * nF + n) + 1];
}

static void spectrumWrite(spectrum *q, const char *path)
{
	const int nF = 1 + q->nMax - q->nMin, M = q->M;
	double *dev = 0;
	if (q->f64) {
		dev = malloc(2 * sizeof(double) * M * nF);
		alisFatal("spectrum read", alis_dev_spectrum_read(q->dev, dev));
	}
	char file[1100];
	snprintf(file, sizeof file, "%s.txt", path);
	FILE *fp = fopen(file, "w");
	for (int n = 0; fp && n < nF; n++) {
		fprintf(fp, "%g", (q->dL * q->nMin * q->nMax) / (q->nMax-n));
		for (int part = 0; part < 2; part++)
			for (int p = 1; p <= M; p++) {
				if (!q->f64) {
					float complex s0 = q->acc[n], r = s0 ? q->acc[(size_t)p * nF + n] / s0 : 0;
					fprintf(fp, "\t%g", s0 ? (part ? cargf(r)/PI : cabsf(r)) : 0);
				}
				else {
					double complex s0 = q->norm[n], r = s0 ? spectrumAt(q, dev, p, n) / s0 : 0;
					fprintf(fp, "\t%g", s0 ? (part ? carg(r)/M_PI : cabs(r)) : 0);
				}
			}
		fprintf(fp, "\n");
	}
	if (fp) fclose(fp);
	free(dev);
}

/* Raw sums of the spectrum (W, name): for p = 0 (the source's spectrum) and every probe,
 * every bin as (re, im) doubles; returns the doubles written (0 if there is no such
 * spectrum or n is too small).  For tests; not part of the reference's API. */
long alisSpectrumRead(world W, const char *name, double *out, long n)
{
	spectrum *q = spectrumOf(W, name);
	if (!q) return 0;
	const int nF = 1 + q->nMax - q->nMin;
	const long need = 2L * (1 + q->M) * nF;
	if (n < need) return 0;
	double *dev = 0;
	if (q->f64) {
		dev = malloc(2 * sizeof(double) * q->M * nF);
		alisFatal("spectrum read", alis_dev_spectrum_read(q->dev, dev));
	}
	for (int p = 0; p <= q->M; p++)
		for (int b = 0; b < nF; b++) {
			double complex v = spectrumAt(q, dev, p, b);
			out[2 * ((long)p * nF + b)] = creal(v), out[2 * ((long)p * nF + b) + 1] = cimag(v);
		}
	free(dev);
	return need;
}

void (writeSpectrum)(world W, int N, float LMin, float LMax, char *format, ...)
{
	char name[1024];
	va_list ap;
	va_start(ap, format);
	alisFormatName(name, sizeof name, format, &ap);
	int M = 0;
	double v[256];
	for (double d = va_arg(ap, double); !NAN2INT(d); d = va_arg(ap, double)) {
		if (M == 256) {
			fprintf(stderr, "writeSpectrum: more than 256 values per call\n");
			exit(1);
		}
		v[M++] = d;
	}
	va_end(ap);
	spectrum *q = spectrumOf(W, name);
	if (!q) {
		q = spectrumNew(W, name, N, LMin, LMax, M);
		if (name[0]=='/') mkdir(W->ID, 0755);
		char dir[1100];
		snprintf(dir, sizeof dir, "%s%s", W->ID, name);
		if (dir[0] && dir[strlen(dir)-1]=='/') mkdir(dir, 0755);
	}
	if (M != q->M) {
		fprintf(stderr, "writeSpectrum: %d values per call, %d before\n", M, q->M);
		exit(1);
	}
	q->calls++;
	const int nF = 1 + q->nMax - q->nMin;
	if (!q->f64) {
		float t = W->t / (q->dL * q->nMin * q->nMax);
		for (int m = 1; m <= M; m++)
			for (int n = 0; n < nF; n++) q->acc[(size_t)m * nF + n] += (float) v[m-1] * cexpf(2*PI*I*t*(q->nMax-n));
	}
	else alisFatal("spectrum", alis_dev_spectrum_add(q->dev, (double)W->t / ((double)q->dL * q->nMin * q->nMax), v));
	if (q->calls == (unsigned long)N) {
		char path[1100];
		snprintf(path, sizeof path, "%s%s", W->ID, name);
		spectrumWrite(q, path);
	}
}


/*
 * Instantaneous Poynting flux through a plane (poyntingX/Y/Z, src/output.c:233-332).  For the
 * plane normal to axis a at index c, with in-plane axes b = a+1 and d = a+2 (cyclic), the flux
 * is the trapezoid sum of E_b (H_d + H_d') over rows along b (from the second index on),
 * weighted by the H widths of b, minus that of E_d (H_b + H_b') over rows along d; H' is the H
 * value one cell further along a.  The row sums run on the device (alis_dev_plane_sums), so
 * no copy of the world is needed; the rows are added here in order, in double.
 */
typedef struct {
	int MIN, MAX, Min, Max, sur;
	float domMin, domMax;
	float *dE, *dH;             /* cell widths of the E and H grids */
	int (*toIndex)(world, float);
} axisView;

static axisView viewOf(world W, int a)
{
	if (a == 0) return (axisView) {W->iMIN, W->iMAX, W->iMin, W->iMax, W->xMinSur, W->Dom.x.Min, W->Dom.x.Max, W->CE->dx, W->CH->dx, W->xtoi};
	if (a == 1) return (axisView) {W->jMIN, W->jMAX, W->jMin, W->jMax, W->yMinSur, W->Dom.y.Min, W->Dom.y.Max, W->CE->dy, W->CH->dy, W->ytoj};
	return (axisView) {W->kMIN, W->kMAX, W->kMin, W->kMax, W->zMinSur, W->Dom.z.Min, W->Dom.z.Max, W->CE->dz, W->CH->dz, W->ztok};
}

/* The index range of [lo, hi] on an axis: one cell wider on each side for a single point. */
static void indexRange(world W, const axisView *v, float lo, float hi, int *from, int *to)
{
	*from = lo == hi ? MAX(v->Min, v->toIndex(W, lo) - 1) : MAX(v->Min, v->toIndex(W, lo));
	*to = lo == hi ? MIN(v->Max, v->toIndex(W, hi) + 1) : MIN(v->Max, v->toIndex(W, hi));
}

static float planeFlux(world W, int a, float pos, float lo1, float hi1, float lo2, float hi2)
{
	alisCommit(W);
	const int b = (a + 1) % 3, d = (a + 2) % 3;
	const axisView A = viewOf(W, a), B = viewOf(W, b), Dd = viewOf(W, d);
	const int c = A.toIndex(W, pos < 0 && A.sur==SYM ? -pos : pos);
	/* the transverse bounds arrive in (x, y, z) order of the two remaining axes */
	const float loB = b < d ? lo1 : lo2, hiB = b < d ? hi1 : hi2, loD = b < d ? lo2 : lo1, hiD = b < d ? hi2 : hi1;
	int b0, b1, d0, d1;
	indexRange(W, &B, loB, hiB, &b0, &b1);
	indexRange(W, &Dd, loD, hiD, &d0, &d1);
	const int nb = b1 - b0 + 1, nd = d1 - d0 + 1;
	double Sum = 0;
	if (nb > 0 && nd > 0) {
		/* inner weights: E widths, halved at both ends; term 1 runs inner over d, term 2 over b */
		double *w = malloc((nb + nd) * sizeof(double)), *out = malloc((nb + nd) * sizeof(double));
		for (int q = d0; q <= d1; q++) w[q - d0] = (q == d0 || q == d1) ? 0.5 * Dd.dE[q] : Dd.dE[q];
		for (int q = b0; q <= b1; q++) w[nd + q - b0] = (q == b0 || q == b1) ? 0.5 * B.dE[q] : B.dE[q];
		int terms[2 * 11], nt = 0;
		if (nb > 1) {
			const int t[11] = {a, c, b, d, b, b0 + 1, b1, d0, d1, 0, 0};
			memcpy(terms + 11 * nt++, t, sizeof t);
		}
		if (nd > 1) {
			const int t[11] = {a, c, d, b, d, d0 + 1, d1, b0, b1, nd, nb};
			memcpy(terms + 11 * nt++, t, sizeof t);
		}
		if (nt) alisFatal("plane sums", alis_dev_plane_sums(W->dev, nt, terms, w, nb + nd, out, nb + nd));
		for (int r = b0 + 1; r <= b1 && nb > 1; r++) Sum += B.dH[r] * out[r - b0 - 1];
		for (int r = d0 + 1; r <= d1 && nd > 1; r++) Sum += Dd.dH[r] * -out[nb + r - d0 - 1];
		free(w), free(out);
	}
	/* a single point or a flat domain on a transverse axis: per unit length (axes in x, y, z order) */
	for (int q = 0; q < 2; q++) {
		const int onB = (q == 0) == (b < d);
		const axisView *V = onB ? &B : &Dd;
		const int top = onB ? b1 : d1;
		if ((onB ? loB == hiB : loD == hiD) || V->domMin == V->domMax) Sum /= V->dH[top] + V->dH[top-1];
	}
	if (B.sur==SYM) Sum *= 2;
	if (Dd.sur==SYM) Sum *= 2;
	return (pos < 0 && A.sur==SYM ? -0.5 : 0.5) * Sum;
}

float (poyntingX)(world W, float x, float yMin, float yMax, float zMin, float zMax, ...) { return planeFlux(W, 0, x, yMin, yMax, zMin, zMax); }
float (poyntingY)(world W, float y, float xMin, float xMax, float zMin, float zMax, ...) { return planeFlux(W, 1, y, xMin, xMax, zMin, zMax); }
float (poyntingZ)(world W, float z, float xMin, float xMax, float yMin, float yMax, ...) { return planeFlux(W, 2, z, xMin, xMax, yMin, yMax); }

/* Net outward flux through the faces of a box; an axis whose box spans a Bloch-periodic world
 * is skipped, and a box reaching a symmetry plane counts its far face twice (src/output.c:335-356). */
float poyntingOut(world W, float xMin, float xMax, float yMin, float yMax, float zMin, float zMax)
{
	const float lo[3] = {xMin, yMin, zMin}, hi[3] = {xMax, yMax, zMax};
	const float wlo[3] = {W->xMin, W->yMin, W->zMin}, whi[3] = {W->xMax, W->yMax, W->zMax};
	float S[3] = {0, 0, 0};
	for (int a = 0; a < 3; a++) {
		const int sur = viewOf(W, a).sur;
		if (!(lo[a] > wlo[a] || hi[a] < whi[a] || sur != BBC)) continue;
		const int b = (a + 1) % 3, d = (a + 2) % 3, u = b < d ? b : d, v = b < d ? d : b;
		S[a] = planeFlux(W, a, hi[a], lo[u], hi[u], lo[v], hi[v]);
		if (lo[a] < 0 && sur==SYM) S[a] *= 2;
		else S[a] -= planeFlux(W, a, lo[a], lo[u], hi[u], lo[v], hi[v]);
	}
	return S[0] + S[1] + S[2];
}


/*
 * Power absorbed by a dispersive object (objectAbsorption, src/output.c:375-435): E . J over
 * the object's box (J: the object's conductive and pole currents over 1/eps,
 * src/output.c:25-47), with the trapezoid corrections where the box reaches the grid's outer
 * planes.  The column sums over z run on the device (alis_dev_absorption); the columns are
 * weighted and added here in order, in double.
 */
/* The weighted power of one column from its device sums s (three sums, the three terms of
 * its first and of its last cell): the first cell's terms come off at the grid's bottom plane
 * and the last cell's x and y terms at its top plane (src/output.c:387-397). */
static double columnPower(world W, const double *s, int i, int j, int k0, int k1, double x, double y, double z)
{
	double xs = s[0], ys = s[1], zs = s[2];
	if (k0 == W->kMIN) xs -= s[3] * 0.5, ys -= s[4] * 0.5, zs -= s[5];
	if (k1 == W->kMAX) xs -= s[6] * 0.5, ys -= s[7] * 0.5;
	coeffs *C = W->CE;
	return x * xs * W->CH->dx[i] * C->dy[j] + y * ys * C->dx[i] * W->CH->dy[j] + z * zs * C->dx[i] * C->dy[j];
}

static int sameObject(object A, object B)
{
	if (A.S != B.S || A.p != B.p) return 0;
	for (int a = 0; a < 3; a++) for (int b = 0; b < 3; b++) if (A.f[a][b] != B.f[a][b]) return 0;
	return 1;
}

float objectAbsorption(world W, object O)
{
	alisCommit(W);
	coeffs *C = W->CE;
	int n = 0;
	for (int o = 0; !sameObject(O, C->O[o]); o++) if (dispersiveMatter(C->M[o])) n++;
	const int i0 = MAX(W->iMIN, C->iMin[n]), i1 = MIN(W->iMAX, C->iMax[n]);
	const int j0 = MAX(W->jMIN, C->jMin[n]), j1 = MIN(W->jMAX, C->jMax[n]);
	const int k0 = MAX(W->kMIN, C->kMin[n]), k1 = MIN(W->kMAX, C->kMax[n]);
	double Sum = 0;
	if (i0 <= i1 && j0 <= j1 && k0 <= k1) {
		const int nj = j1 - j0 + 1, nk = k1 - k0 + 1;
		double *dz = malloc(2 * nk * sizeof(double)), *col = malloc(9 * sizeof(double) * (i1 - i0 + 1) * nj);
		for (int k = k0; k <= k1; k++) dz[k - k0] = C->dz[k], dz[nk + k - k0] = W->CH->dz[k];
		alisFatal("absorption", alis_dev_absorption(W->dev, n, i0, i1, j0, j1, k0, k1, dz, dz + nk, col));
		/* one column with the trapezoid weights x, y, z of its three components */
		#define COLUMN(i, j, x, y, z) columnPower(W, col + 9 * ((long)((i) - i0) * nj + (j) - j0), i, j, k0, k1, x, y, z)
		for (int i = i0; i <= i1; i++) {
			for (int j = j0; j <= j1; j++) Sum += COLUMN(i, j, 1, 1, 1);
			if (j0 == W->jMIN) Sum += COLUMN(i, W->jMIN, -0.5, -1, -0.5);
			if (j1 == W->jMAX) Sum += COLUMN(i, W->jMAX, -0.5, 0, -0.5);
		}
		if (i0 == W->iMIN) {
			for (int j = j0; j <= j1; j++) Sum += COLUMN(W->iMIN, j, -1, -0.5, -0.5);
			if (j0 == W->jMIN) Sum += COLUMN(W->iMIN, W->jMIN, 0.5, 0.5, 0.25);
			if (j1 == W->jMAX) Sum += COLUMN(W->iMIN, W->jMAX, 0.5, 0.0, 0.25);
		}
		if (i1 == W->iMAX) {
			for (int j = j0; j <= j1; j++) Sum += COLUMN(W->iMAX, j, 0, -0.5, -0.5);
			if (j0 == W->jMIN) Sum += COLUMN(W->iMAX, W->jMIN, 0.0, 0.5, 0.25);
			if (j1 == W->jMAX) Sum += COLUMN(W->iMAX, W->jMAX, 0.0, 0.0, 0.25);
		}
		#undef COLUMN
		free(dz), free(col);
	}
	if (W->Dom.x.Min==W->Dom.x.Max) Sum /= W->CH->dx[W->iMax]+W->CH->dx[W->iMax-1];
	if (W->Dom.y.Min==W->Dom.y.Max) Sum /= W->CH->dy[W->jMax]+W->CH->dy[W->jMax-1];
	if (W->Dom.z.Min==W->Dom.z.Max) Sum /= W->CH->dz[W->kMax]+W->CH->dz[W->kMax-1];
	if (W->xMinSur==SYM) Sum *= 2;
	if (W->yMinSur==SYM) Sum *= 2;
	if (W->zMinSur==SYM) Sum *= 2;
	return Sum / W->dt;
}

void exec(char *format, ...)
{
	char str[4096];
	va_list ap;
	va_start(ap, format);
	vsnprintf(str, sizeof str, format, ap);
	va_end(ap);
	if (system(str)) {}
}
