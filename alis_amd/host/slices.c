/*
 * Slices of the world and their writers (src/slices.c:1-427): snapshots, time averages and
 * single-frequency phasors on 1-D/2-D cuts, written as text, PNG (own encoder on zlib) or
 * the array container of h5write.  Evaluated on the host arrays after a device sync.
 */
#include <string.h>
#include <limits.h>
#include <sys/stat.h>
#include <zlib.h>
#include "internal.h"

typedef struct squeue *squeue;
struct squeue {
	world W;
	char *ID;
	float max, Max, ***F, ***G;
	long int N;
	method M;
	squeue Q;
};

/* Resample a non-uniform slice onto the uniform grid of spacing dx, dy, dz by piecewise
 * linear interpolation along each non-degenerate axis (src/slices.c:25-64). */
static float ***resampleUniform(world W, slice S, float ***F)
{
	float ***T, ***G = makeField(S->iMIN, S->iMAX, S->jMIN, S->jMAX, S->kMIN, S->kMAX);
	if (S->iMin != S->iMIN || S->iMax != S->iMAX) {
		T = F, F = G, G = T;
		for (int i = S->iMin; i < S->iMax; i++) {
			float x0 = W->itox(W, i), x1 = W->itox(W, i+1);
			int n0 = x0/W->dx+(i>0?1:0), n1 = x1/W->dx-(i<0?1:0);
			for (int j = S->jMin; j <= S->jMax; j++) for (int n = n0; n <= n1; n++) for (int k = S->kMin; k <= S->kMax; k++)
				F[n][j][k] = (G[i][j][k] * (x1 - n*W->dx) + G[i+1][j][k] * (n*W->dx - x0)) / (x1 - x0);
		}
	}
	if (S->jMin != S->jMIN || S->jMax != S->jMAX) {
		T = F, F = G, G = T;
		for (int j = S->jMin; j < S->jMax; j++) {
			float y0 = W->jtoy(W, j), y1 = W->jtoy(W, j+1);
			int n0 = y0/W->dy+(j>0?1:0), n1 = y1/W->dy-(j<0?1:0);
			for (int i = S->iMIN; i <= S->iMAX; i++) for (int n = n0; n <= n1; n++) for (int k = S->kMin; k <= S->kMax; k++)
				F[i][n][k] = (G[i][j][k] * (y1 - n*W->dy) + G[i][j+1][k] * (n*W->dy - y0)) / (y1 - y0);
		}
	}
	if (S->kMin != S->kMIN || S->kMax != S->kMAX) {
		T = F, F = G, G = T;
		for (int k = S->kMin; k < S->kMax; k++) {
			float z0 = W->ktoz(W, k), z1 = W->ktoz(W, k+1);
			int n0 = z0/W->dz+(k>0?1:0), n1 = z1/W->dz-(k<0?1:0);
			for (int i = S->iMIN; i <= S->iMAX; i++) for (int j = S->jMIN; j <= S->jMAX; j++) for (int n = n0; n <= n1; n++)
				F[i][j][n] = (G[i][j][k] * (z1 - n*W->dz) + G[i][j][k+1] * (n*W->dz - z0)) / (z1 - z0);
		}
	}
	removeField(G, S->iMIN, S->jMIN, S->kMIN);
	return F;
}

/* Cut of the world between the given coordinates; 1-D and 2-D cuts of worlds with objects
 * carry a contour overlay of the refractive index (src/slices.c:67-117). */
slice createSlice(world W, float xMin, float xMax, float yMin, float yMax, float zMin, float zMax)
{
	slice S;
	MALLOC(S, 0, 1);
	S->iMin = W->xtoi(W, xMin), S->iMax = W->xtoi(W, xMax);
	S->jMin = W->ytoj(W, yMin), S->jMax = W->ytoj(W, yMax);
	S->kMin = W->ztok(W, zMin), S->kMax = W->ztok(W, zMax);
	if (W->Res.nug.x.m>1 || W->Res.nug.y.m>1 || W->Res.nug.z.m>1) S->interpolate = resampleUniform;

	if (S->iMax > W->iMAX) S->iMax = W->iMAX;
	if (S->jMax > W->jMAX) S->jMax = W->jMAX;
	if (S->kMax > W->kMAX) S->kMax = W->kMAX;
	if (S->iMin < W->iMIN && W->xMinSur!=SYM) S->iMin = W->iMIN;
	if (S->jMin < W->jMIN && W->yMinSur!=SYM) S->jMin = W->jMIN;
	if (S->kMin < W->kMIN && W->zMinSur!=SYM) S->kMin = W->kMIN;
	if (S->iMin <-W->iMAX && W->xMinSur==SYM) S->iMin =-W->iMAX;
	if (S->jMin <-W->jMAX && W->yMinSur==SYM) S->jMin =-W->jMAX;
	if (S->kMin <-W->kMAX && W->zMinSur==SYM) S->kMin =-W->kMAX;
	S->iMIN = (S->iMin==S->iMax || !S->interpolate) ? S->iMin : MIN(S->iMin, W->itox(W, S->iMin)/W->dx);
	S->jMIN = (S->jMin==S->jMax || !S->interpolate) ? S->jMin : MIN(S->jMin, W->jtoy(W, S->jMin)/W->dy);
	S->kMIN = (S->kMin==S->kMax || !S->interpolate) ? S->kMin : MIN(S->kMin, W->ktoz(W, S->kMin)/W->dz);
	S->iMAX = (S->iMin==S->iMax || !S->interpolate) ? S->iMax : MAX(S->iMax, W->itox(W, S->iMax)/W->dx);
	S->jMAX = (S->jMin==S->jMax || !S->interpolate) ? S->jMax : MAX(S->jMax, W->jtoy(W, S->jMax)/W->dy);
	S->kMAX = (S->kMin==S->kMax || !S->interpolate) ? S->kMax : MAX(S->kMax, W->ktoz(W, S->kMax)/W->dz);
	S->iNUM = 1-S->iMIN+S->iMAX;
	S->jNUM = 1-S->jMIN+S->jMAX;
	S->kNUM = 1-S->kMIN+S->kMAX;
	S->F = makeField(S->iMIN, S->iMAX, S->jMIN, S->jMAX, S->kMIN, S->kMAX);

	if (!W->CE->x || !W->CE->y || !W->CE->z || (S->iMin!=S->iMax && S->jMin!=S->jMax && S->kMin!=S->kMax)) return S;

	float Max = 0;
	S->C = makeField(S->iMIN, S->iMAX, S->jMIN, S->jMAX, S->kMIN, S->kMAX);
	for (int i = S->iMIN; i <= S->iMAX; i++) for (int j = S->jMIN; j <= S->jMAX; j++) for (int k = S->kMIN; k <= S->kMAX; k++)
		if (Max < (S->C[i][j][k] = ContourRI(W, i, j, k))) Max = S->C[i][j][k];
	if (Max)
		for (int i = S->iMIN; i <= S->iMAX; i++) for (int j = S->jMIN; j <= S->jMAX; j++) for (int k = S->kMIN; k <= S->kMAX; k++)
			S->C[i][j][k] = 0.5*S->C[i][j][k]/Max;
	return S;
}

void deleteSlice(slice S)
{
	if (S->F) removeField(S->F, S->iMIN, S->jMIN, S->kMIN);
	if (S->C) removeField(S->C, S->iMIN, S->jMIN, S->kMIN);
	free(S);
}


void (h5)(world W, slice S, float Max, char *F, char *filename, unsigned long int *cmap)
{
	(void)Max, (void)F, (void)cmap;
	int dim[3] = {S->iMin==S->iMax?0:S->iNUM, S->jMin==S->jMax?0:S->jNUM, S->kMin==S->kMax?0:S->kNUM};
	if (S->interpolate) S->F = S->interpolate(W, S, S->F);
	h5write(dim, &S->F[S->iMIN][S->jMIN][S->kMIN], "%s", filename);
}

void (txt)(world W, slice S, float Max, char *F, char *filename, unsigned long int *cmap)
{
	(void)Max, (void)F, (void)cmap;
	int d = 0, dim[3] = {0};
	char name[1100];
	snprintf(name, sizeof name, "%s.txt", filename);
	if (S->iMin != S->iMax) dim[d++] = S->iNUM;
	if (S->jMin != S->jMax) dim[d++] = S->jNUM;
	if (S->kMin != S->kMax) dim[d++] = S->kNUM;
	if (d==3) printf("Can't write 3D slice to txt file.\n"), exit(0);
	FILE *fp = fopen(name, "w");
	if (!fp) return;
	if (d==1) {
		if (S->iMin != S->iMax) for (int i = S->iMin; i <= S->iMax; i++) fprintf(fp, "%f\t%e\n", W->itox(W, i), S->F[i][S->jMin][S->kMin]);
		if (S->jMin != S->jMax) for (int j = S->jMin; j <= S->jMax; j++) fprintf(fp, "%f\t%e\n", W->jtoy(W, j), S->F[S->iMin][j][S->kMin]);
		if (S->kMin != S->kMax) for (int k = S->kMin; k <= S->kMax; k++) fprintf(fp, "%f\t%e\n", W->ktoz(W, k), S->F[S->iMin][S->jMin][k]);
	}
	else {
		if (S->interpolate) S->F = S->interpolate(W, S, S->F);
		float *v = &S->F[S->iMIN][S->jMIN][S->kMIN];
		for (int j = dim[1]-1; j >= 0; j--) {
			fprintf(fp, "%e", v[j]);
			for (int i = 1; i < dim[0]; i++) fprintf(fp, "\t%e", v[(long)dim[1]*i+j]);
			fprintf(fp, "\n");
		}
	}
	fclose(fp);
}

/* Minimal PNG (8-bit RGB, one IDAT, filter 0) written with zlib. */
static void pngChunk(FILE *fp, const char *type, const unsigned char *data, unsigned len)
{
	unsigned char b[4] = { len >> 24, len >> 16, len >> 8, len };
	fwrite(b, 1, 4, fp);
	fwrite(type, 1, 4, fp);
	if (len) fwrite(data, 1, len, fp);
	uLong crc = crc32(0L, (const Bytef *)type, 4);
	if (len) crc = crc32(crc, data, len);
	unsigned char c[4] = { crc >> 24, crc >> 16, crc >> 8, crc };
	fwrite(c, 1, 4, fp);
}

static void writePng(const char *name, int width, int height, const unsigned char *rgb)
{
	FILE *fp = fopen(name, "wb");
	if (!fp) return;
	static const unsigned char sig[8] = { 137, 80, 78, 71, 13, 10, 26, 10 };
	fwrite(sig, 1, 8, fp);
	unsigned char ihdr[13] = { width >> 24, width >> 16, width >> 8, width, height >> 24, height >> 16, height >> 8, height, 8, 2, 0, 0, 0 };
	pngChunk(fp, "IHDR", ihdr, 13);
	uLong rawLen = (uLong)height * (3*(uLong)width + 1);
	unsigned char *raw = malloc(rawLen);
	for (long r = 0; r < height; r++) {
		raw[r*(3*(long)width+1)] = 0;
		memcpy(raw + r*(3*(long)width+1) + 1, rgb + r*3*(long)width, 3*(long)width);
	}
	uLongf zLen = compressBound(rawLen);
	unsigned char *z = malloc(zLen);
	compress2(z, &zLen, raw, rawLen, Z_DEFAULT_COMPRESSION);
	pngChunk(fp, "IDAT", z, zLen);
	pngChunk(fp, "IEND", 0, 0);
	fclose(fp);
	free(raw), free(z);
}

/* Colour-mapped image of a 1-D (64 rows tall) or 2-D slice; fields whose name starts with
 * E/H/J/U/S/i are darkened by the contour overlay; "Log" maps span 3 decades
 * (src/slices.c:155-211). */
void (png)(world W, slice S, float Max, char *F, char *filename, unsigned long int *cmap)
{
	(void)W;
	int d = 0, dim[3] = {1,1,1};
	char name[1100];
	snprintf(name, sizeof name, "%s.png", filename);
	float Min = 0;
	if (cmap[0] > LONG_MAX) Min = -Max;
	if (F[0]=='L' && F[1]=='o' && F[2]=='g' && F[3]!='R' && F[4]!='I') Min = Max-3;
	if (S->iMin != S->iMax) dim[d++] = S->iNUM;
	if (S->jMin != S->jMax) dim[d++] = S->jNUM;
	if (S->kMin != S->kMax) dim[d++] = S->kNUM;
	if (d==1) dim[2] = 64;
	if (d==3) printf("Can't write 3D slice to png file.\n"), exit(0);
	if (S->interpolate) S->F = S->interpolate(W, S, S->F);

	int colours = abs((long)cmap[0]);
	unsigned char *img = malloc((size_t)3*dim[0]*dim[1]*dim[2]), *row = malloc((size_t)3*dim[0]);
	float *v = &S->F[S->iMIN][S->jMIN][S->kMIN], *c = S->C ? &S->C[S->iMIN][S->jMIN][S->kMIN] : 0;
	int shade = F[0]=='E' || F[0]=='H' || F[0]=='J' || F[0]=='U' || F[0]=='S' || F[0]=='i';
	long out = 0;
	for (int j = dim[1]-1; j >= 0; j--) {
		for (int i = 0; i < dim[0]; i++) {
			long at = (long)dim[1]*i+j;
			float C = !c ? 1 : 1-c[at];
			float f = (Max-Min) ? MAX(0, MIN(colours-1, (colours-1)*(v[at]-Min)/(Max-Min))) : 0;
			int n = f;
			if (!shade) C = 1;
			if (n > colours-2) n = colours-2;
			row[3*i]   = C * ((1-f+n) * (cmap[n+1]>>0x10&0xFF) + (f-n) * (cmap[n+2]>>0x10&0xFF));
			row[3*i+1] = C * ((1-f+n) * (cmap[n+1]>>0x08&0xFF) + (f-n) * (cmap[n+2]>>0x08&0xFF));
			row[3*i+2] = C * ((1-f+n) * (cmap[n+1]>>0x00&0xFF) + (f-n) * (cmap[n+2]>>0x00&0xFF));
		}
		for (int r = 0; r < dim[2]; r++, out += 3*dim[0]) memcpy(img + out, row, 3*dim[0]);
	}
	writePng(name, dim[0], dim[1]*dim[2], img);
	free(img), free(row);
}


static squeue findSliceQueue(world W, char *name, method M, squeue *head)
{
	if (!*head) {
		MALLOC(*head, 0, 1);
	}
	squeue Q = *head;
	while (Q->Q && (W != Q->W || M != Q->M || strcmp(name, Q->ID))) Q = Q->Q;
	int fresh = !Q->Q;
	if (fresh) {
		Q->W = W, Q->M = M;
		MALLOC(Q->ID, 0, 1+strlen(name));
		strcpy(Q->ID, name);
		MALLOC(Q->Q, 0, 1);
	}
	char full[1024];
	snprintf(full, sizeof full, "%s%s", W->ID, Q->ID);
	strcpy(name, full);
	if (fresh) {
		if (Q->ID[0]=='/') mkdir(W->ID, 0755);
		if (name[0] && name[strlen(name)-1]=='/') mkdir(name, 0755);
	}
	Q->N++;
	return Q;
}

/* Replace the first '%' of name with "<field>-<slice>". */
static void tagName(char *name, const char *f, const char *s)
{
	char *p = strchr(name, '%'), rest[1024];
	if (!p) return;
	strcpy(rest, p+1);
	sprintf(p, "%s-%s%s", f, s, rest);
}

static int staticField(field F)
{
	return F==RI || F==LogRI || F==ContourRI || F==Contour || F==raweEffx || F==raweEffy || F==raweEffz || F==eEff;
}

/* Snapshot of F on S every T/N steps (N per period; 0 writes once); Max < 0 rescales to the
 * running maximum over -Max periods (src/slices.c:242-279). */
void (sliceSnap)(world W, field F, slice S, char *f, char *s, ...)
{
	static squeue head;
	char filename[1024], num[32];
	va_list ap;
	va_start(ap, s);
	int N = staticField(F) ? 0 : va_arg(ap, int);
	method M = va_arg(ap, method);
	unsigned long int *cmap = (M == png) ? va_arg(ap, unsigned long int *) : 0;
	float Max = (M == png) ? va_arg(ap, double) : 1;
	int step = (!N || N>W->T) ? 1 : W->T/N;
	int skip = Max<0 ? ((int)(-Max*W->T/step))*step : 0;
	char *format = va_arg(ap, char*);
	vsnprintf(filename, sizeof filename, format, ap);
	va_end(ap);
	tagName(filename, f, s);
	squeue Q = findSliceQueue(W, filename, M, &head);
	if (N) snprintf(num, sizeof num, "%05ld", (Q->N-skip)/step), strcat(filename, num);

	if (Q->N % step == 0) {
		if (!staticField(F)) syncWorld(W);
		float max = 0;
		for (int i = S->iMin; i <= S->iMax; i++) for (int j = S->jMin; j <= S->jMax; j++) for (int k = S->kMin; k <= S->kMax; k++)
			if (max < fabsf(S->F[i][j][k] = F(W, i, j, k))) max = fabsf(S->F[i][j][k]);
		if (!Max) Max = max;
		else if (Max<0) {
			Max = Q->Max;
			if (Q->max < max) Q->max = max;
			if (Q->N % skip == 0) Q->Max = Q->max, Q->max = 0;
		}
		if (Q->N > skip) M(W, S, Max?Max:1, f, filename, cmap);
	}
}

/* Average of F over every N calls (src/slices.c:282-320). */
void (sliceTimeAvg)(world W, field F, slice S, char *f, char *s, ...)
{
	static squeue head;
	char filename[1024], num[32];
	va_list ap;
	va_start(ap, s);
	int N = staticField(F) ? 0 : va_arg(ap, int);
	method M = va_arg(ap, method);
	unsigned long int *cmap = (M == png) ? va_arg(ap, unsigned long int *) : 0;
	float Max = (M == png) ? va_arg(ap, double) : 1;
	char *format = va_arg(ap, char*);
	vsnprintf(filename, sizeof filename, format, ap);
	va_end(ap);
	tagName(filename, f, s);
	squeue Q = findSliceQueue(W, filename, M, &head);
	if (N) snprintf(num, sizeof num, "%05ld", Q->N / N), strcat(filename, num);
	if (Q->N == 1) Q->F = makeField(S->iMin, S->iMax, S->jMin, S->jMax, S->kMin, S->kMax);
	syncWorld(W);
	for (int i = S->iMin; i <= S->iMax; i++) for (int j = S->jMin; j <= S->jMax; j++) for (int k = S->kMin; k <= S->kMax; k++)
		Q->F[i][j][k] += F(W, i, j, k);
	if (N && Q->N % N == 0) {
		float max = 0;
		for (int i = S->iMin; i <= S->iMax; i++) for (int j = S->jMin; j <= S->jMax; j++) for (int k = S->kMin; k <= S->kMax; k++)
			if (max < fabsf(S->F[i][j][k] = Q->F[i][j][k] / N)) max = fabsf(S->F[i][j][k]);
		if (Max<=0) Max = max;
		M(W, S, Max?Max:1, f, filename, cmap);
		for (int i = S->iMin; i <= S->iMax; i++) for (int j = S->jMin; j <= S->jMax; j++) for (int k = S->kMin; k <= S->kMax; k++)
			Q->F[i][j][k] = 0;
	}
}

/* Single-frequency phasor of F at wavelength L accumulated over N calls, written as 12
 * phase frames (src/slices.c:323-379). */
void (sliceFreqDom)(world W, field F, slice S, char *f, char *s, int N, float L, ...)
{
	field *allowed[30] = {Ex, Ey, Ez, Hx, Hy, Hz, iEx, iEy, iEz, iHx, iHy, iHz, rawEx, rawEy, rawEz, rawHx, rawHy, rawHz,
	                      Jx, Jy, Jz, rawJx, rawJy, rawJz, ScEx, ScEy, ScEz, SoEx, SoEy, SoEz};
	int n = 0;
	while (n < 30 && F != allowed[n]) n++;
	if (n == 30) printf("\nThe %s field is not supported in sliceFreqDom().\n", f), exit(0);

	static squeue head;
	char filename[1024], frame[1100];
	va_list ap;
	va_start(ap, L);
	method M = va_arg(ap, method);
	unsigned long int *cmap = (M == png) ? va_arg(ap, unsigned long int *) : 0;
	float Max = (M == png) ? va_arg(ap, double) : 1;
	char *format = va_arg(ap, char*);
	vsnprintf(filename, sizeof filename, format, ap);
	va_end(ap);
	tagName(filename, f, s);
	squeue Q = findSliceQueue(W, filename, M, &head);
	if (Q->N == 1) {
		Q->F = makeField(S->iMin, S->iMax, S->jMin, S->jMax, S->kMin, S->kMax);
		Q->G = makeField(S->iMin, S->iMax, S->jMin, S->jMax, S->kMin, S->kMax);
	}
	syncWorld(W);
	float re = cosf(2 * PI * Q->N * W->dt / L);
	float im = sinf(2 * PI * Q->N * W->dt / L);
	for (int i = S->iMin; i <= S->iMax; i++) for (int j = S->jMin; j <= S->jMax; j++) for (int k = S->kMin; k <= S->kMax; k++) {
		Q->F[i][j][k] += F(W, i, j, k) * re;
		Q->G[i][j][k] += F(W, i, j, k) * im;
	}
	if (N && Q->N % N == 0) {
		if (Max <= 0) {
			float max = 0;
			for (int i = S->iMin; i <= S->iMax; i++) for (int j = S->jMin; j <= S->jMax; j++) for (int k = S->kMin; k <= S->kMax; k++)
				if (max < sq(Q->F[i][j][k])+sq(Q->G[i][j][k])) max = sq(Q->F[i][j][k])+sq(Q->G[i][j][k]);
			if (Max < sqrtf(max) * 2 / Q->N) Max = sqrtf(max) * 2 / Q->N;
		}
		for (int p = 0; p < 12; p++) {
			for (int i = S->iMin; i <= S->iMax; i++) for (int j = S->jMin; j <= S->jMax; j++) for (int k = S->kMin; k <= S->kMax; k++)
				S->F[i][j][k] = (Q->F[i][j][k] * cosf(p * PI / 6) + Q->G[i][j][k] * sinf(p * PI / 6)) * 2 / Q->N;
			snprintf(frame, sizeof frame, "%s%02d", filename, p+1);
			M(W, S, Max, f, frame, cmap);
		}
	}
}

float sliceMax(world W, field F, slice S)
{
	if (!staticField(F)) syncWorld(W);
	float Max = 0;
	for (int i = S->iMin; i <= S->iMax; i++) for (int j = S->jMin; j <= S->jMax; j++) for (int k = S->kMin; k <= S->kMax; k++) {
		float f = fabsf(F(W, i, j, k));
		if (Max < f) Max = f;
	}
	return Max;
}

/* Trapezoid integral of F over the slice, per unit length along degenerate axes
 * (src/slices.c:382-426). */
float sliceSum(world W, field F, slice S)
{
	if (!staticField(F)) syncWorld(W);
	double Sum = 0;
	for (int i = S->iMin; i <= S->iMax; i++) {
		double wi = (S->iMin == S->iMax) ? W->CE->dx[i] : (i == S->iMin || i == S->iMax) ? 0.5 * W->CE->dx[i] : W->CE->dx[i];
		double sy = 0;
		for (int j = S->jMin; j <= S->jMax; j++) {
			double wj = (S->jMin == S->jMax) ? W->CE->dy[j] : (j == S->jMin || j == S->jMax) ? 0.5 * W->CE->dy[j] : W->CE->dy[j];
			double sz = 0;
			for (int k = S->kMin; k <= S->kMax; k++) {
				double wk = (S->kMin == S->kMax) ? W->CE->dz[k] : (k == S->kMin || k == S->kMax) ? 0.5 * W->CE->dz[k] : W->CE->dz[k];
				sz += wk * F(W, i, j, k);
			}
			sy += wj * sz;
		}
		Sum += wi * sy;
	}
	if (S->iMin==S->iMax) Sum /= W->CE->dx[S->iMin];
	if (S->jMin==S->jMax) Sum /= W->CE->dy[S->jMin];
	if (S->kMin==S->kMax) Sum /= W->CE->dz[S->kMin];
	return Sum;
}
