/*
 * Slices (src/slices.c:1-427): cuts of a world -- a node, a line or a plane of nodes --
 * sampled with a field accessor and written as text, PNG (an encoder of its own over zlib)
 * or the array container of h5write.  A cut is sampled from host mirrors refreshed over its
 * own box only, in the arrays the accessor reads, not from a copy of the whole world.  Worlds
 * on graded grids are resampled onto the uniform grid before they are written.
 */
#include <string.h>
#include <limits.h>
#include <sys/stat.h>
#include <zlib.h>
#include "internal.h"

/* Accessors of the permittivity, which read no device state. */
static int staticField(field F)
{
	return F==RI || F==LogRI || F==ContourRI || F==Contour || F==raweEffx || F==raweEffy || F==raweEffz || F==eEff;
}

/* Refresh the host mirrors that accessor F reads on cut S: the cut's box, mirrored across a
 * symmetry plane and two cells wider (the accessors reach one cell beyond the node), in the
 * component array of a component accessor or in every array otherwise. */
static void refreshCut(world W, field F, slice S)
{
	if (staticField(F) || !W->dev || !W->hostStale) return;
	const int from[3] = {S->iMin, S->jMin, S->kMin}, to[3] = {S->iMax, S->jMax, S->kMax};
	int lo[3], hi[3];
	for (int a = 0; a < 3; a++) {
		const int sym = *axisOf(W, a).minSur == SYM && from[a] < 0;
		lo[a] = (sym ? 0 : from[a]) - 2;
		hi[a] = (sym ? MAX(-from[a], to[a]) : to[a]) + 2;
	}
	int isH, m, axis, mask = ALIS_DEV_PROBE_ALL;
	if (fieldComponent(F, &isH, &m, &axis) || rawComponent(F, &isH, &m, &axis)) mask = 1 << (isH * 6 + m * 3 + axis);
	alisSyncArrays(W, lo[0], hi[0], lo[1], hi[1], lo[2], hi[2], mask);
}

/* F over the cut into S->F; returns the largest magnitude. */
static float sampleCut(world W, field F, slice S)
{
	refreshCut(W, F, S);
	float max = 0;
	for (int i = S->iMin; i <= S->iMax; i++) for (int j = S->jMin; j <= S->jMax; j++) for (int k = S->kMin; k <= S->kMax; k++) {
		const float v = S->F[i][j][k] = F(W, i, j, k);
		if (max < fabsf(v)) max = fabsf(v);
	}
	return max;
}


/* ---- geometry ---- */

/* Linear resampling along axis a of a cut on a graded grid: the value at the uniform node n
 * (position n d) from the two graded nodes around it (src/slices.c:25-64).  Axes before a
 * are already resampled (their uniform range); axes after a keep the cut's node range. */
static void resampleAxis(world W, slice S, float ***src, float ***dst, int a)
{
	const axis X = axisOf(W, a);
	const int cut[3][2] = {{S->iMin, S->iMax}, {S->jMin, S->jMax}, {S->kMin, S->kMax}};
	const int all[3][2] = {{S->iMIN, S->iMAX}, {S->jMIN, S->jMAX}, {S->kMIN, S->kMAX}};
	const int p = a == 0 ? 1 : 0, q = a == 2 ? 1 : 2;
	const int *rp = p < a ? all[p] : cut[p], *rq = q < a ? all[q] : cut[q];
	const float d = *X.step;
	for (int u = cut[a][0]; u < cut[a][1]; u++) {
		const float c0 = (*X.coord)(W, u), c1 = (*X.coord)(W, u+1);
		const int n0 = c0/d + (u > 0 ? 1 : 0), n1 = c1/d - (u < 0 ? 1 : 0);
		for (int s = rp[0]; s <= rp[1]; s++)
			for (int t = rq[0]; t <= rq[1]; t++)
				for (int n = n0; n <= n1; n++) {
					int g0[3], g1[3], o[3];
					g0[a] = u, g1[a] = u + 1, o[a] = n;
					g0[p] = g1[p] = o[p] = s;
					g0[q] = g1[q] = o[q] = t;
					dst[o[0]][o[1]][o[2]] = (src[g0[0]][g0[1]][g0[2]] * (c1 - n*d) + src[g1[0]][g1[1]][g1[2]] * (n*d - c0)) / (c1 - c0);
				}
	}
}

static float ***resampleUniform(world W, slice S, float ***F)
{
	float ***G = makeField(S->iMIN, S->iMAX, S->jMIN, S->jMAX, S->kMIN, S->kMAX);
	const int graded[3] = {S->iMin != S->iMIN || S->iMax != S->iMAX, S->jMin != S->jMIN || S->jMax != S->jMAX, S->kMin != S->kMIN || S->kMax != S->kMAX};
	for (int a = 0; a < 3; a++) {
		if (!graded[a]) continue;
		float ***t = F;
		F = G, G = t;
		resampleAxis(W, S, G, F, a);
	}
	removeField(G, S->iMIN, S->jMIN, S->kMIN);
	return F;
}

/* Cut between the given coordinates, clipped to the grid (mirrored across a symmetry plane);
 * on a graded grid its storage spans the uniform nodes too.  Lines and planes of worlds with
 * objects carry a contour of the refractive index, scaled to 0..0.5 (src/slices.c:67-117). */
slice createSlice(world W, float xMin, float xMax, float yMin, float yMax, float zMin, float zMax)
{
	slice S;
	MALLOC(S, 0, 1);
	if (W->Res.nug.x.m > 1 || W->Res.nug.y.m > 1 || W->Res.nug.z.m > 1) S->interpolate = resampleUniform;
	const float lo[3] = {xMin, yMin, zMin}, hi[3] = {xMax, yMax, zMax};
	int *cMin[3] = {&S->iMin, &S->jMin, &S->kMin}, *cMax[3] = {&S->iMax, &S->jMax, &S->kMax};
	int *sMIN[3] = {&S->iMIN, &S->jMIN, &S->kMIN}, *sMAX[3] = {&S->iMAX, &S->jMAX, &S->kMAX}, *sNUM[3] = {&S->iNUM, &S->jNUM, &S->kNUM};
	for (int a = 0; a < 3; a++) {
		const axis X = axisOf(W, a);
		*cMin[a] = (*X.node)(W, lo[a]), *cMax[a] = (*X.node)(W, hi[a]);
		if (*cMax[a] > *X.MAX) *cMax[a] = *X.MAX;
		if (*X.minSur != SYM && *cMin[a] < *X.MIN) *cMin[a] = *X.MIN;
		if (*X.minSur == SYM && *cMin[a] < -*X.MAX) *cMin[a] = -*X.MAX;
		const int plain = *cMin[a] == *cMax[a] || !S->interpolate;
		*sMIN[a] = plain ? *cMin[a] : MIN(*cMin[a], (*X.coord)(W, *cMin[a]) / *X.step);
		*sMAX[a] = plain ? *cMax[a] : MAX(*cMax[a], (*X.coord)(W, *cMax[a]) / *X.step);
		*sNUM[a] = 1 - *sMIN[a] + *sMAX[a];
	}
	S->F = makeField(S->iMIN, S->iMAX, S->jMIN, S->jMAX, S->kMIN, S->kMAX);
	const int solid = S->iMin != S->iMax && S->jMin != S->jMax && S->kMin != S->kMax;
	if (!W->CE->x || !W->CE->y || !W->CE->z || solid) return S;
	S->C = makeField(S->iMIN, S->iMAX, S->jMIN, S->jMAX, S->kMIN, S->kMAX);
	float *c = &S->C[S->iMIN][S->jMIN][S->kMIN], Max = 0;
	const long n = (long)S->iNUM * S->jNUM * S->kNUM;
	long e = 0;
	for (int i = S->iMIN; i <= S->iMAX; i++) for (int j = S->jMIN; j <= S->jMAX; j++) for (int k = S->kMIN; k <= S->kMAX; k++, e++)
		if (Max < (c[e] = ContourRI(W, i, j, k))) Max = c[e];
	for (e = 0; Max && e < n; e++) c[e] = 0.5*c[e]/Max;
	return S;
}

void deleteSlice(slice S)
{
	if (S->F) removeField(S->F, S->iMIN, S->jMIN, S->kMIN);
	if (S->C) removeField(S->C, S->iMIN, S->jMIN, S->kMIN);
	free(S);
}


/* ---- writers ---- */

/* The extended axes of a cut: their count and node counts, in x, y, z order. */
static int cutAxes(slice S, int axes[3], int n[3])
{
	const int ext[3] = {S->iMin != S->iMax, S->jMin != S->jMax, S->kMin != S->kMax}, num[3] = {S->iNUM, S->jNUM, S->kNUM};
	int d = 0;
	for (int a = 0; a < 3; a++) if (ext[a]) axes[d] = a, n[d++] = num[a];
	return d;
}

void (h5)(world W, slice S, float Max, char *F, char *filename, unsigned long int *cmap)
{
	(void)Max, (void)F, (void)cmap;
	if (S->interpolate) S->F = S->interpolate(W, S, S->F);
	int dim[3] = {S->iMin == S->iMax ? 0 : S->iNUM, S->jMin == S->jMax ? 0 : S->jNUM, S->kMin == S->kMax ? 0 : S->kNUM};
	h5write(dim, &S->F[S->iMIN][S->jMIN][S->kMIN], "%s", filename);
}

/* A line as "coordinate<TAB>value" rows; a plane as rows of the second axis from its last
 * node down, each a tab-separated run along the first axis (src/slices.c:128-152). */
void (txt)(world W, slice S, float Max, char *F, char *filename, unsigned long int *cmap)
{
	(void)Max, (void)F, (void)cmap;
	int axes[3], n[3];
	const int d = cutAxes(S, axes, n);
	if (d == 3) printf("Can't write 3D slice to txt file.\n"), exit(0);
	char name[1100];
	snprintf(name, sizeof name, "%s.txt", filename);
	FILE *fp = fopen(name, "w");
	if (!fp) return;
	if (d == 1) {
		const axis X = axisOf(W, axes[0]);
		const int first[3] = {S->iMin, S->jMin, S->kMin}, last[3] = {S->iMax, S->jMax, S->kMax};
		for (int u = first[axes[0]]; u <= last[axes[0]]; u++) {
			int c[3] = {S->iMin, S->jMin, S->kMin};
			c[axes[0]] = u;
			fprintf(fp, "%f\t%e\n", (*X.coord)(W, u), S->F[c[0]][c[1]][c[2]]);
		}
	}
	else if (d == 2) {
		if (S->interpolate) S->F = S->interpolate(W, S, S->F);
		const float *v = &S->F[S->iMIN][S->jMIN][S->kMIN];
		for (int r = n[1] - 1; r >= 0; r--) {
			for (int u = 0; u < n[0]; u++) fprintf(fp, u ? "\t%e" : "%e", v[(long)n[1] * u + r]);
			fputc('\n', fp);
		}
	}
	fclose(fp);
}

/* PNG (8-bit RGB, one IDAT, filter 0) over zlib. */
static void pngChunk(FILE *fp, const char *type, const unsigned char *data, unsigned len)
{
	const unsigned char b[4] = { len >> 24, len >> 16, len >> 8, len };
	fwrite(b, 1, 4, fp);
	fwrite(type, 1, 4, fp);
	if (len) fwrite(data, 1, len, fp);
	uLong crc = crc32(0L, (const Bytef *)type, 4);
	if (len) crc = crc32(crc, data, len);
	const unsigned char c[4] = { crc >> 24, crc >> 16, crc >> 8, crc };
	fwrite(c, 1, 4, fp);
}

static void writePng(const char *name, int width, int height, const unsigned char *rgb)
{
	FILE *fp = fopen(name, "wb");
	if (!fp) return;
	static const unsigned char sig[8] = { 137, 80, 78, 71, 13, 10, 26, 10 };
	fwrite(sig, 1, 8, fp);
	const unsigned char ihdr[13] = { width >> 24, width >> 16, width >> 8, width, height >> 24, height >> 16, height >> 8, height, 8, 2, 0, 0, 0 };
	pngChunk(fp, "IHDR", ihdr, 13);
	const uLong stride = 3*(uLong)width + 1, rawLen = (uLong)height * stride;
	unsigned char *raw = malloc(rawLen);
	for (long r = 0; r < height; r++) raw[r*stride] = 0, memcpy(raw + r*stride + 1, rgb + r*3*(long)width, 3*(long)width);
	uLongf zLen = compressBound(rawLen);
	unsigned char *z = malloc(zLen);
	compress2(z, &zLen, raw, rawLen, Z_DEFAULT_COMPRESSION);
	pngChunk(fp, "IDAT", z, zLen);
	pngChunk(fp, "IEND", 0, 0);
	fclose(fp);
	free(raw), free(z);
}

/* Colour of value v on colour map cmap (cmap[0] colours, linear between neighbours) over
 * [lo, hi], dimmed by the factor shade (src/slices.c:185-202). */
static void colourOf(const unsigned long int *cmap, float lo, float hi, float v, float shade, unsigned char *rgb)
{
	const int colours = labs((long)cmap[0]);
	const float f = (hi-lo) ? MAX(0, MIN(colours-1, (colours-1)*(v-lo)/(hi-lo))) : 0;
	int n = f;
	if (n > colours-2) n = colours-2;
	for (int c = 0; c < 3; c++) {
		const int sh = 16 - 8*c;
		rgb[c] = shade * ((1-f+n) * (cmap[n+1]>>sh & 0xFF) + (f-n) * (cmap[n+2]>>sh & 0xFF));
	}
}

/* A plane as an image (rows of the second axis from its last node down), a line as a band
 * 64 rows tall; a symmetric colour map spans [-Max, Max], "Log" fields three decades below
 * Max; fields of E, H, J, U, S and i names are dimmed by the cut's contour. */
void (png)(world W, slice S, float Max, char *F, char *filename, unsigned long int *cmap)
{
	int axes[3], n[3] = {1, 1, 1};
	const int d = cutAxes(S, axes, n);
	if (d == 3) printf("Can't write 3D slice to png file.\n"), exit(0);
	if (d < 2) n[1] = 1;
	const int rows = d == 1 ? 64 : 1;
	if (S->interpolate) S->F = S->interpolate(W, S, S->F);
	float lo = 0;
	if (cmap[0] > LONG_MAX) lo = -Max;
	if (F[0]=='L' && F[1]=='o' && F[2]=='g' && F[3]!='R' && F[4]!='I') lo = Max-3;
	const int dim = F[0]=='E' || F[0]=='H' || F[0]=='J' || F[0]=='U' || F[0]=='S' || F[0]=='i';
	const float *v = &S->F[S->iMIN][S->jMIN][S->kMIN], *c = S->C ? &S->C[S->iMIN][S->jMIN][S->kMIN] : 0;
	unsigned char *img = malloc((size_t)3*n[0]*n[1]*rows);
	long out = 0;
	for (int r = n[1]-1; r >= 0; r--) {
		unsigned char *row = img + out;
		for (int u = 0; u < n[0]; u++) {
			const long at = (long)n[1]*u + r;
			colourOf(cmap, lo, Max, v[at], dim && c ? 1-c[at] : 1, row + 3*u);
		}
		out += 3*n[0];
		for (int t = 1; t < rows; t++, out += 3*n[0]) memcpy(img + out, row, 3*n[0]);
	}
	char name[1100];
	snprintf(name, sizeof name, "%s.png", filename);
	writePng(name, n[0], n[1]*rows, img);
	free(img);
}


/* ---- slice observers ---- */

/* A record per (world, method, file name), counting its calls. */
typedef struct record {
	world W;
	method M;
	char *name;
	long calls;
	float runMax, shownMax;     /* sliceSnap: running maximum and the one in use */
	float ***A, ***B;           /* sliceTimeAvg: sum; sliceFreqDom: cosine and sine sums */
	struct record *next;
} record;

/* The record of (W, M, file) with the first '%' of the file name replaced by
 * "<field>-<slice>", counted; the output directories are made on first use. */
static record *recordOf(record **list, world W, method M, char *file, const char *f, const char *s)
{
	char *p = strchr(file, '%');
	if (p) {
		char rest[1024];
		snprintf(rest, sizeof rest, "%s", p + 1);
		snprintf(p, 1024 - (p - file), "%s-%s%s", f, s, rest);
	}
	record *q = *list;
	while (q && (q->W != W || q->M != M || strcmp(q->name, file))) q = q->next;
	if (!q) {
		q = calloc(1, sizeof *q);
		q->W = W, q->M = M, q->name = strdup(file);
		q->next = *list, *list = q;
		if (file[0]=='/') mkdir(W->ID, 0755);
		char full[1100];
		snprintf(full, sizeof full, "%s%s", W->ID, file);
		if (full[0] && full[strlen(full)-1]=='/') mkdir(full, 0755);
	}
	char full[1100];
	snprintf(full, sizeof full, "%s%s", W->ID, q->name);
	snprintf(file, 1024, "%s", full);
	q->calls++;
	return q;
}

/* The method and scale arguments of the slice observers: method, then for png the colour
 * map and the scale, then the file name format and its arguments. */
#define METHOD_ARGS(ap, M, cmap, Max, file) do { \
	M = va_arg(ap, method); \
	cmap = M == png ? va_arg(ap, unsigned long int *) : 0; \
	Max = M == png ? va_arg(ap, double) : 1; \
	const char *fmt_ = va_arg(ap, char*); \
	vsnprintf(file, 1024, fmt_, ap); } while (0)

/* Snapshots of F on S, N per period (0: one per call); Max 0 scales each frame to its own
 * maximum, Max < 0 to the largest magnitude of the previous -Max periods, whose frames are
 * not written (src/slices.c:242-279). */
void (sliceSnap)(world W, field F, slice S, char *f, char *s, ...)
{
	static record *list;
	char file[1024];
	method M;
	unsigned long int *cmap;
	float Max;
	va_list ap;
	va_start(ap, s);
	const int N = staticField(F) ? 0 : va_arg(ap, int);
	METHOD_ARGS(ap, M, cmap, Max, file);
	va_end(ap);
	const int step = (!N || N > W->T) ? 1 : W->T/N;
	const int skip = Max < 0 ? ((int)(-Max*W->T/step))*step : 0;
	record *q = recordOf(&list, W, M, file, f, s);
	if (N) snprintf(file + strlen(file), 32, "%05ld", (q->calls-skip)/step);
	if (q->calls % step) return;
	const float max = sampleCut(W, F, S);
	if (!Max) Max = max;
	else if (Max < 0) {
		Max = q->shownMax;
		if (q->runMax < max) q->runMax = max;
		if (skip && q->calls % skip == 0) q->shownMax = q->runMax, q->runMax = 0;
	}
	if (q->calls > skip) M(W, S, Max ? Max : 1, f, file, cmap);
}

/* The average of F on S over every N calls (src/slices.c:282-320). */
void (sliceTimeAvg)(world W, field F, slice S, char *f, char *s, ...)
{
	static record *list;
	char file[1024];
	method M;
	unsigned long int *cmap;
	float Max;
	va_list ap;
	va_start(ap, s);
	const int N = staticField(F) ? 0 : va_arg(ap, int);
	METHOD_ARGS(ap, M, cmap, Max, file);
	va_end(ap);
	record *q = recordOf(&list, W, M, file, f, s);
	if (N) snprintf(file + strlen(file), 32, "%05ld", q->calls / N);
	if (!q->A) q->A = makeField(S->iMin, S->iMax, S->jMin, S->jMax, S->kMin, S->kMax);
	refreshCut(W, F, S);
	for (int i = S->iMin; i <= S->iMax; i++) for (int j = S->jMin; j <= S->jMax; j++) for (int k = S->kMin; k <= S->kMax; k++)
		q->A[i][j][k] += F(W, i, j, k);
	if (!N || q->calls % N) return;
	float max = 0;
	for (int i = S->iMin; i <= S->iMax; i++) for (int j = S->jMin; j <= S->jMax; j++) for (int k = S->kMin; k <= S->kMax; k++) {
		S->F[i][j][k] = q->A[i][j][k] / N;
		if (max < fabsf(S->F[i][j][k])) max = fabsf(S->F[i][j][k]);
		q->A[i][j][k] = 0;
	}
	if (Max <= 0) Max = max;
	M(W, S, Max ? Max : 1, f, file, cmap);
}

/* The phasor of F on S at wavelength L accumulated over every N calls, written as 12 frames
 * of one period (src/slices.c:323-379); component, current and source/scattered field
 * accessors only. */
void (sliceFreqDom)(world W, field F, slice S, char *f, char *s, int N, float L, ...)
{
	field *allowed[30] = {Ex, Ey, Ez, Hx, Hy, Hz, iEx, iEy, iEz, iHx, iHy, iHz, rawEx, rawEy, rawEz, rawHx, rawHy, rawHz,
	                      Jx, Jy, Jz, rawJx, rawJy, rawJz, ScEx, ScEy, ScEz, SoEx, SoEy, SoEz};
	int known = 0;
	for (int a = 0; a < 30; a++) known |= F == allowed[a];
	if (!known) printf("\nThe %s field is not supported in sliceFreqDom().\n", f), exit(0);
	static record *list;
	char file[1024], frame[1100];
	method M;
	unsigned long int *cmap;
	float Max;
	va_list ap;
	va_start(ap, L);
	METHOD_ARGS(ap, M, cmap, Max, file);
	va_end(ap);
	record *q = recordOf(&list, W, M, file, f, s);
	if (!q->A) {
		q->A = makeField(S->iMin, S->iMax, S->jMin, S->jMax, S->kMin, S->kMax);
		q->B = makeField(S->iMin, S->iMax, S->jMin, S->jMax, S->kMin, S->kMax);
	}
	if (F==ScEx || F==ScEy || F==ScEz || F==SoEx || F==SoEy || F==SoEz) syncWorld(W);
	else refreshCut(W, F, S);
	const float re = cosf(2 * PI * q->calls * W->dt / L), im = sinf(2 * PI * q->calls * W->dt / L);
	for (int i = S->iMin; i <= S->iMax; i++) for (int j = S->jMin; j <= S->jMax; j++) for (int k = S->kMin; k <= S->kMax; k++) {
		q->A[i][j][k] += F(W, i, j, k) * re;
		q->B[i][j][k] += F(W, i, j, k) * im;
	}
	if (!N || q->calls % N) return;
	if (Max <= 0) {
		float max = 0;
		for (int i = S->iMin; i <= S->iMax; i++) for (int j = S->jMin; j <= S->jMax; j++) for (int k = S->kMin; k <= S->kMax; k++)
			max = MAX(max, sq(q->A[i][j][k]) + sq(q->B[i][j][k]));
		if (Max < sqrtf(max) * 2 / q->calls) Max = sqrtf(max) * 2 / q->calls;
	}
	for (int p = 0; p < 12; p++) {
		for (int i = S->iMin; i <= S->iMax; i++) for (int j = S->jMin; j <= S->jMax; j++) for (int k = S->kMin; k <= S->kMax; k++)
			S->F[i][j][k] = (q->A[i][j][k] * cosf(p * PI / 6) + q->B[i][j][k] * sinf(p * PI / 6)) * 2 / q->calls;
		snprintf(frame, sizeof frame, "%s%02d", file, p+1);
		M(W, S, Max, f, frame, cmap);
	}
}

float sliceMax(world W, field F, slice S)
{
	refreshCut(W, F, S);
	float Max = 0;
	for (int i = S->iMin; i <= S->iMax; i++) for (int j = S->jMin; j <= S->jMax; j++) for (int k = S->kMin; k <= S->kMax; k++)
		Max = MAX(Max, fabsf(F(W, i, j, k)));
	return Max;
}

/* Trapezoid integral of F over the cut, in double, per unit length along its flat axes
 * (src/slices.c:382-426). */
float sliceSum(world W, field F, slice S)
{
	refreshCut(W, F, S);
	const int first[3] = {S->iMin, S->jMin, S->kMin}, last[3] = {S->iMax, S->jMax, S->kMax};
	float *width[3] = {W->CE->dx, W->CE->dy, W->CE->dz};
	double Sum = 0;
	for (int i = S->iMin; i <= S->iMax; i++) {
		double sy = 0;
		for (int j = S->jMin; j <= S->jMax; j++) {
			double sz = 0;
			for (int k = S->kMin; k <= S->kMax; k++) {
				const double wk = (first[2] != last[2] && (k == first[2] || k == last[2])) ? 0.5 * width[2][k] : width[2][k];
				sz += wk * F(W, i, j, k);
			}
			sy += ((first[1] != last[1] && (j == first[1] || j == last[1])) ? 0.5 * width[1][j] : width[1][j]) * sz;
		}
		Sum += ((first[0] != last[0] && (i == first[0] || i == last[0])) ? 0.5 * width[0][i] : width[0][i]) * sy;
	}
	for (int a = 0; a < 3; a++) if (first[a] == last[a]) Sum /= width[a][first[a]];
	return Sum;
}
