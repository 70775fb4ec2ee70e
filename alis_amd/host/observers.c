/*
 * Observers and text output (src/output.c:13-445): probes, world maxima/integrals, per-file
 * text queues, probe spectra, Poynting fluxes through planes and absorbed power.  They read
 * the host arrays, which are refreshed from the device first (a small box for get()).
 */
#include <string.h>
#include <sys/stat.h>
#include <sys/time.h>
#include "internal.h"
#include "phaserview.h"

/* Output file queue keyed by (world, name) (src/output.c:104-128). */
typedef struct queue *queue;
struct queue {
	world W;
	char *ID;
	FILE *File;
	unsigned long int N;
	unsigned long long int timeStamp;
	queue Q;
};

/* Find or create the queue of (W, name); on return `name` holds W->ID + name.  Creates the
 * output directory when the name starts with '/' or ends with '/'. */
static queue findQueue(world W, char *name, queue *head)
{
	if (!*head) {
		MALLOC(*head, 0, 1);
	}
	queue Q = *head;
	while (Q->Q && (W != Q->W || strcmp(name, Q->ID))) Q = Q->Q;
	if (!Q->Q) {
		Q->W = W;
		MALLOC(Q->ID, 0, 1+strlen(name));
		strcpy(Q->ID, name);
		MALLOC(Q->Q, 0, 1);
		if (Q->ID[0]=='/') mkdir(W->ID, 0755);
	}
	char full[1024];
	snprintf(full, sizeof full, "%s%s", W->ID, Q->ID);
	strcpy(name, full);
	if (Q->N == 0 && name[0] && name[strlen(name)-1]=='/') mkdir(name, 0755);
	Q->N++;
	return Q;
}

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


/* Append text to file W->ID + name.txt; the file is flushed every 5 minutes and at exit
 * (src/output.c:131-154).  The first format names the file, the second the text. */
void (writeTxt)(world W, char *format, ...)
{
	static queue head;
	char str[4096], filename[1024];
	va_list ap;
	va_start(ap, format);
	vsnprintf(filename, sizeof filename, format, ap);
	format = va_arg(ap, char*);
	vsnprintf(str, sizeof str, format, ap);
	va_end(ap);
	queue Q = findQueue(W, filename, &head);
	struct timeval now;
	gettimeofday(&now, 0);
	unsigned long long int time = now.tv_sec * 1000000LL + now.tv_usec;
	if (!Q->timeStamp) {
		strcat(filename, ".txt");
		Q->File = fopen(filename, "w");
		Q->timeStamp = time;
	}
	if (Q->File) fputs(str, Q->File);
	if (time - Q->timeStamp >= 300*1000000LL) {
		if (Q->File) fflush(Q->File);
		Q->timeStamp = time;
	}
}

/* One tab-separated row of doubles terminated by the INT2NAN(1) sentinel (src/output.c:157-174). */
void (writeRow)(world W, char *format, ...)
{
	char str[4096], filename[1024];
	va_list ap;
	va_start(ap, format);
	vsnprintf(filename, sizeof filename, format, ap);
	double f = va_arg(ap, double);
	if (!NAN2INT(f)) {
		int n = snprintf(str, sizeof str, "%g", f);
		f = va_arg(ap, double);
		while (!NAN2INT(f)) {
			n += snprintf(str+n, sizeof str - n, "\t%g", f);
			f = va_arg(ap, double);
		}
		snprintf(str+n, sizeof str - n, "\n");
		(writeTxt)(W, filename, "%s", str, "");
	}
	va_end(ap);
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
	vsnprintf(name, sizeof name, format, ap);
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


/* Instantaneous Poynting flux through a plane normal to x (y, z) at x, within the given
 * transverse bounds; H is averaged onto the E plane (src/output.c:233-332). */
float (poyntingX)(world W, float x, float yMin, float yMax, float zMin, float zMax, ...)
{
	syncWorld(W);
	int i = W->xtoi(W, x<0 && W->xMinSur==SYM ? -x : x);
	int jMin = yMin==yMax ? MAX(W->jMin, W->ytoj(W, yMin)-1) : MAX(W->jMin, W->ytoj(W, yMin));
	int jMax = yMin==yMax ? MIN(W->jMax, W->ytoj(W, yMax)+1) : MIN(W->jMax, W->ytoj(W, yMax));
	int kMin = zMin==zMax ? MAX(W->kMin, W->ztok(W, zMin)-1) : MAX(W->kMin, W->ztok(W, zMin));
	int kMax = zMin==zMax ? MIN(W->kMax, W->ztok(W, zMax)+1) : MIN(W->kMax, W->ztok(W, zMax));
	double Sum = 0;
	for (int j = jMin+1; j <= jMax; j++) {
		double s = 0;
		for (int k = kMin; k <= kMax; k++) {
			double w = (k == kMin || k == kMax) ? 0.5 * W->CE->dz[k] : W->CE->dz[k];
			s += w * W->E->y[i][j][k] * (W->H->z[i][j][k] + W->H->z[i+1][j][k]);
		}
		Sum += W->CH->dy[j] * s;
	}
	for (int k = kMin+1; k <= kMax; k++) {
		double s = 0;
		for (int j = jMin; j <= jMax; j++) {
			double w = (j == jMin || j == jMax) ? 0.5 * W->CE->dy[j] : W->CE->dy[j];
			s -= w * W->E->z[i][j][k] * (W->H->y[i][j][k] + W->H->y[i+1][j][k]);
		}
		Sum += W->CH->dz[k] * s;
	}
	if (yMin==yMax || W->Dom.y.Min==W->Dom.y.Max) Sum /= W->CH->dy[jMax]+W->CH->dy[jMax-1];
	if (zMin==zMax || W->Dom.z.Min==W->Dom.z.Max) Sum /= W->CH->dz[kMax]+W->CH->dz[kMax-1];
	if (W->yMinSur==SYM) Sum *= 2;
	if (W->zMinSur==SYM) Sum *= 2;
	return (x<0 && W->xMinSur==SYM ? -0.5 : 0.5) * Sum;
}

float (poyntingY)(world W, float y, float xMin, float xMax, float zMin, float zMax, ...)
{
	syncWorld(W);
	int j = W->ytoj(W, y<0 && W->yMinSur==SYM ? -y : y);
	int iMin = xMin==xMax ? MAX(W->iMin, W->xtoi(W, xMin)-1) : MAX(W->iMin, W->xtoi(W, xMin));
	int iMax = xMin==xMax ? MIN(W->iMax, W->xtoi(W, xMax)+1) : MIN(W->iMax, W->xtoi(W, xMax));
	int kMin = zMin==zMax ? MAX(W->kMin, W->ztok(W, zMin)-1) : MAX(W->kMin, W->ztok(W, zMin));
	int kMax = zMin==zMax ? MIN(W->kMax, W->ztok(W, zMax)+1) : MIN(W->kMax, W->ztok(W, zMax));
	double Sum = 0;
	for (int k = kMin+1; k <= kMax; k++) {
		double s = 0;
		for (int i = iMin; i <= iMax; i++) {
			double w = (i == iMin || i == iMax) ? 0.5 * W->CE->dx[i] : W->CE->dx[i];
			s += w * W->E->z[i][j][k] * (W->H->x[i][j][k] + W->H->x[i][j+1][k]);
		}
		Sum += W->CH->dz[k] * s;
	}
	for (int i = iMin+1; i <= iMax; i++) {
		double s = 0;
		for (int k = kMin; k <= kMax; k++) {
			double w = (k == kMin || k == kMax) ? 0.5 * W->CE->dz[k] : W->CE->dz[k];
			s -= w * W->E->x[i][j][k] * (W->H->z[i][j][k] + W->H->z[i][j+1][k]);
		}
		Sum += W->CH->dx[i] * s;
	}
	if (xMin==xMax || W->Dom.x.Min==W->Dom.x.Max) Sum /= W->CH->dx[iMax]+W->CH->dx[iMax-1];
	if (zMin==zMax || W->Dom.z.Min==W->Dom.z.Max) Sum /= W->CH->dz[kMax]+W->CH->dz[kMax-1];
	if (W->xMinSur==SYM) Sum *= 2;
	if (W->zMinSur==SYM) Sum *= 2;
	return (y<0 && W->yMinSur==SYM ? -0.5 : 0.5) * Sum;
}

float (poyntingZ)(world W, float z, float xMin, float xMax, float yMin, float yMax, ...)
{
	syncWorld(W);
	int k = W->ztok(W, z<0 && W->zMinSur==SYM ? -z : z);
	int iMin = xMin==xMax ? MAX(W->iMin, W->xtoi(W, xMin)-1) : MAX(W->iMin, W->xtoi(W, xMin));
	int iMax = xMin==xMax ? MIN(W->iMax, W->xtoi(W, xMax)+1) : MIN(W->iMax, W->xtoi(W, xMax));
	int jMin = yMin==yMax ? MAX(W->jMin, W->ytoj(W, yMin)-1) : MAX(W->jMin, W->ytoj(W, yMin));
	int jMax = yMin==yMax ? MIN(W->jMax, W->ytoj(W, yMax)+1) : MIN(W->jMax, W->ytoj(W, yMax));
	double Sum = 0;
	for (int i = iMin+1; i <= iMax; i++) {
		double s = 0;
		for (int j = jMin; j <= jMax; j++) {
			double w = (j == jMin || j == jMax) ? 0.5 * W->CE->dy[j] : W->CE->dy[j];
			s += w * W->E->x[i][j][k] * (W->H->y[i][j][k] + W->H->y[i][j][k+1]);
		}
		Sum += W->CH->dx[i] * s;
	}
	for (int j = jMin+1; j <= jMax; j++) {
		double s = 0;
		for (int i = iMin; i <= iMax; i++) {
			double w = (i == iMin || i == iMax) ? 0.5 * W->CE->dx[i] : W->CE->dx[i];
			s -= w * W->E->y[i][j][k] * (W->H->x[i][j][k] + W->H->x[i][j][k+1]);
		}
		Sum += W->CH->dy[j] * s;
	}
	if (xMin==xMax || W->Dom.x.Min==W->Dom.x.Max) Sum /= W->CH->dx[iMax]+W->CH->dx[iMax-1];
	if (yMin==yMax || W->Dom.y.Min==W->Dom.y.Max) Sum /= W->CH->dy[jMax]+W->CH->dy[jMax-1];
	if (W->xMinSur==SYM) Sum *= 2;
	if (W->yMinSur==SYM) Sum *= 2;
	return (z<0 && W->zMinSur==SYM ? -0.5 : 0.5) * Sum;
}

/* Net outward flux through the faces of a box; periodic axes spanning the world are
 * skipped (src/output.c:335-356). */
float poyntingOut(world W, float xMin, float xMax, float yMin, float yMax, float zMin, float zMax)
{
	float Sx = 0, Sy = 0, Sz = 0;
	if (xMin > W->xMin || xMax < W->xMax || W->xMinSur != BBC) {
		Sx = (poyntingX)(W, xMax, yMin, yMax, zMin, zMax);
		if (xMin < 0 && W->xMinSur==SYM) Sx *= 2;
		else Sx -= (poyntingX)(W, xMin, yMin, yMax, zMin, zMax);
	}
	if (yMin > W->yMin || yMax < W->yMax || W->yMinSur != BBC) {
		Sy = (poyntingY)(W, yMax, xMin, xMax, zMin, zMax);
		if (yMin < 0 && W->yMinSur==SYM) Sy *= 2;
		else Sy -= (poyntingY)(W, yMin, xMin, xMax, zMin, zMax);
	}
	if (zMin > W->zMin || zMax < W->zMax || W->zMinSur != BBC) {
		Sz = (poyntingZ)(W, zMax, xMin, xMax, yMin, yMax);
		if (zMin < 0 && W->zMinSur==SYM) Sz *= 2;
		else Sz -= (poyntingZ)(W, zMin, xMin, xMax, yMin, yMax);
	}
	return Sx + Sy + Sz;
}


/* Polarisation current of dispersive object n at an E position: conductive and pole
 * currents over 1/eps (src/output.c:25-47). */
static float objCurrent(world W, int q, int n, int i, int j, int k)
{
	coeffs *C = W->CE;
	float a = q == 0 ? C->ax[n] : q == 1 ? C->ay[n] : C->az[n];
	float ***fill = q == 0 ? C->Jx[n] : q == 1 ? C->Jy[n] : C->Jz[n];
	float ***E = q == 0 ? W->E->x : q == 1 ? W->E->y : W->E->z;
	float ****J = q == 0 ? W->E->Jx[n] : q == 1 ? W->E->Jy[n] : W->E->Jz[n];
	float ***Cq = q == 0 ? C->x : q == 1 ? C->y : C->z;
	float f = 0;
	if (a) f += 2 * fill[i][j][k] * E[i][j][k] / (1 + 1 / a);
	for (int p = 0; p < C->NK[n]; p++) if (J[p]) f += J[p][i][j][k];
	return f / Cq[i][j][k];
}

/* E.J of object n on the column (i, j) with boundary weights x, y, z (src/output.c:375-398). */
static double columnAbsorption(world W, int n, int i, int j, double x, double y, double z)
{
	coeffs *C = W->CE;
	double xs = 0, ys = 0, zs = 0;
	int kMinJ = MAX(W->kMIN, C->kMin[n]), kMaxJ = MIN(W->kMAX, C->kMax[n]);
	for (int k = kMinJ; k <= kMaxJ; k++) {
		if (C->x[i][j][k]) xs += C->dz[k] * W->E->x[i][j][k] * objCurrent(W, 0, n, i, j, k);
		if (C->y[i][j][k]) ys += C->dz[k] * W->E->y[i][j][k] * objCurrent(W, 1, n, i, j, k);
		if (C->z[i][j][k]) zs += W->CH->dz[k] * W->E->z[i][j][k] * objCurrent(W, 2, n, i, j, k);
	}
	if (kMinJ == W->kMIN) {
		if (C->x[i][j][W->kMIN]) xs -= C->dz[W->kMIN] * W->E->x[i][j][W->kMIN] * objCurrent(W, 0, n, i, j, W->kMIN) * 0.5;
		if (C->y[i][j][W->kMIN]) ys -= C->dz[W->kMIN] * W->E->y[i][j][W->kMIN] * objCurrent(W, 1, n, i, j, W->kMIN) * 0.5;
		if (C->z[i][j][W->kMIN]) zs -= W->CH->dz[W->kMIN] * W->E->z[i][j][W->kMIN] * objCurrent(W, 2, n, i, j, W->kMIN);
	}
	if (kMaxJ == W->kMAX) {
		if (C->x[i][j][W->kMAX]) xs -= C->dz[W->kMAX] * W->E->x[i][j][W->kMAX] * objCurrent(W, 0, n, i, j, W->kMAX) * 0.5;
		if (C->y[i][j][W->kMAX]) ys -= C->dz[W->kMAX] * W->E->y[i][j][W->kMAX] * objCurrent(W, 1, n, i, j, W->kMAX) * 0.5;
	}
	return x * xs * W->CH->dx[i] * C->dy[j] + y * ys * C->dx[i] * W->CH->dy[j] + z * zs * C->dx[i] * C->dy[j];
}

static int sameObject(object A, object B)
{
	if (A.S != B.S || A.p != B.p) return 0;
	for (int a = 0; a < 3; a++) for (int b = 0; b < 3; b++) if (A.f[a][b] != B.f[a][b]) return 0;
	return 1;
}

/* Power absorbed by the dispersive object O (E.J over its box, src/output.c:401-435). */
float objectAbsorption(world W, object O)
{
	syncWorld(W);
	int n = 0;
	for (int o = 0; !sameObject(O, W->CE->O[o]); o++) if (dispersiveMatter(W->CE->M[o])) n++;
	coeffs *C = W->CE;
	int iMinJ = MAX(W->iMIN, C->iMin[n]), iMaxJ = MIN(W->iMAX, C->iMax[n]);
	int jMinJ = MAX(W->jMIN, C->jMin[n]), jMaxJ = MIN(W->jMAX, C->jMax[n]);
	double Sum = 0;
	#pragma omp parallel for reduction(+:Sum)
	for (int i = iMinJ; i <= iMaxJ; i++) {
		for (int j = jMinJ; j <= jMaxJ; j++) Sum += columnAbsorption(W, n, i, j, 1, 1, 1);
		if (jMinJ == W->jMIN) Sum += columnAbsorption(W, n, i, W->jMIN, -0.5, -1, -0.5);
		if (jMaxJ == W->jMAX) Sum += columnAbsorption(W, n, i, W->jMAX, -0.5, 0, -0.5);
	}
	if (iMinJ == W->iMIN) {
		for (int j = jMinJ; j <= jMaxJ; j++) Sum += columnAbsorption(W, n, W->iMIN, j, -1, -0.5, -0.5);
		if (jMinJ == W->jMIN) Sum += columnAbsorption(W, n, W->iMIN, W->jMIN, 0.5, 0.5, 0.25);
		if (jMaxJ == W->jMAX) Sum += columnAbsorption(W, n, W->iMIN, W->jMAX, 0.5, 0.0, 0.25);
	}
	if (iMaxJ == W->iMAX) {
		for (int j = jMinJ; j <= jMaxJ; j++) Sum += columnAbsorption(W, n, W->iMAX, j, 0, -0.5, -0.5);
		if (jMinJ == W->jMIN) Sum += columnAbsorption(W, n, W->iMAX, W->jMIN, 0.0, 0.5, 0.25);
		if (jMaxJ == W->jMAX) Sum += columnAbsorption(W, n, W->iMAX, W->jMAX, 0.0, 0.0, 0.25);
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
