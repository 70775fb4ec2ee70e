/*
 * Objects and materials: voxel bounds, sub-pixel fill factors, inverse permittivity of the
 * E grid and the Drude / Lorentz (critical-point) ADE coefficients of dispersive objects.
 * Follows src/object.c:23-549; coefficient expressions keep the reference's float / float
 * complex / double types so that the arrays are bitwise equal to the reference's.
 */
#include <string.h>
#include "internal.h"

typedef struct { int i0, i1, j0, j1, k0, k1; } box3;

/* 2 when a nonlinear tensor is given, 1 when the matter has poles or a conductivity,
 * 0 for a plain dielectric (src/object.c:282-287). */
int dispersiveMatter(matter M)
{
	for (int a = 0; a < 3; a++) for (int b = 0; b < 6; b++) if (M.d[a][b]) return 2;
	if (M.P[0].w || M.P[0].r) return 1;
	return 0;
}

/* Pole p as {w, gamma, f, g} in eV: gamma_c = gamma + i w and the residue term
 * sigma_c = i (f w - g gamma_c) of the complex-conjugate pole pair (src/object.c:23-31). */
static float complex poleGamma(matter M, int p) { return M.P[p].r + I * M.P[p].w; }
static float complex poleSigma(matter M, int p) { return I * (M.P[p].f * M.P[p].w - M.P[p].g * poleGamma(M, p)); }

/* Recursion coefficients of the trapezoidal ADE (src/object.c:33-51): Drude current
 * J' = CJa J + CJb E, pole pair J' = CKa J + CKb E (complex, split into a real J/K pair).
 * pi*gamma*dt/eV is gamma_rad*dt/2 since energies are in eV and dt in nm. */
static float drudeDecay(world W, matter M, int p)
{
	return (1 - PI * M.P[p].r * W->dt / W->eV) / (1 + PI * M.P[p].r * W->dt / W->eV);
}

static float drudeDrive(world W, matter M, int p)
{
	return sq(2 * PI * M.P[p].w * W->dt / W->eV) / (1 + PI * M.P[p].r * W->dt / W->eV);
}

static float complex pairDecay(world W, matter M, int p)
{
	return (1 - PI * poleGamma(M, p) * W->dt / W->eV) / (1 + PI * poleGamma(M, p) * W->dt / W->eV);
}

static float complex pairDrive(world W, matter M, int p)
{
	return - poleGamma(M, p) * poleSigma(M, p) * sq(2 * PI * W->dt / W->eV) / (1 + PI * poleGamma(M, p) * W->dt / W->eV);
}

/* Static conductivity of component q: {0, sigma} poles plus the g*w terms of critical
 * points, in units of eV (src/object.c:53-62).  Anisotropic matter stores poles in
 * x/y/z triplets. */
static float conductivity(world W, matter M, int q)
{
	int m = M.e.y || M.e.z ? 1 : 0;
	float s = 0;
	for (int p = 0; p < MAXPOLES; p += 1+2*m) {
		if (!M.P[p+q*m].w && M.P[p+q*m].r) s += M.P[p+q*m].r;
		if (M.P[p+q*m].g) s += M.P[p+q*m].g * M.P[p+q*m].w;
	}
	return s / W->eV;
}

/* Effective high-frequency permittivity of component q including the conductivity term
 * of the semi-implicit update (src/object.c:64-67). */
static float permittivity(world W, matter M, int q)
{
	return ((q==2 && M.e.z) ? M.e.z : (q==1 && M.e.y) ? M.e.y : M.e.x) + PI * W->dt * conductivity(W, M, q);
}

/* Susceptibility of pole p at angular energy w (eV), used by the far-field layer model
 * (src/farfld.c:63-71). */
float complex polePermittivity(matter M, float w, int p)
{
	if (M.P[p].f || M.P[p].g)
		return 0.5
			*((M.P[p].f * M.P[p].w - M.P[p].g * M.P[p].r + I * M.P[p].g * M.P[p].w) / (M.P[p].w + w + I * M.P[p].r)
			+ (M.P[p].f * M.P[p].w - M.P[p].g * M.P[p].r - I * M.P[p].g * M.P[p].w) / (M.P[p].w - w - I * M.P[p].r));
	if (M.P[p].w) return - sq(M.P[p].w) / (sq(w) + I * w * M.P[p].r);
	return -I * M.P[p].r / w;
}


/* 1 when grid point (i,j,k) (fractional indices allowed) lies in object o and in none of
 * the objects before it (earlier objects take precedence; src/object.c:70-75). */
/* Candidate objects of a point, for the inside queries of putObjectArray: the x-y plane
 * of the grid in NB x NB bins; an object with basic-shape bounds is listed in the bins
 * those bounds meet, any other object in every bin.  Lists are in object order.  A query
 * still evaluates the shapes themselves, so the answers are those of a scan over all
 * objects; only objects that cannot contain the point are skipped. */
#define NB 64
typedef struct {
	float x0, y0, sx, sy;       /* bin origin and inverse widths */
	int start[NB*NB+1], *list;
	float (*b)[6];              /* per object: bounds, or infinite */
} objIndex;

static objIndex *buildIndex(world W, object *O, int N)
{
	objIndex *X = calloc(1, sizeof *X);
	const float x0 = W->itox(W, W->iMIN - 1), x1 = W->itox(W, W->iMAX + 1);
	const float y0 = W->jtoy(W, W->jMIN - 1), y1 = W->jtoy(W, W->jMAX + 1);
	X->x0 = x0, X->y0 = y0;
	X->sx = x1 > x0 ? NB / (x1 - x0) : 0, X->sy = y1 > y0 ? NB / (y1 - y0) : 0;
	MALLOC(X->b, 0, N ? N : 1);
	int (*range)[4];
	MALLOC(range, 0, N ? N : 1);
	long total = 0;
	for (int o = 0; o < N; o++) {
		float *b = X->b[o];
		if (!shapeBounds(&O[o], b)) for (int a = 0; a < 6; a++) b[a] = a & 1 ? INF : -INF;
		float f[4] = { (b[0] - x0) * X->sx, (b[1] - x0) * X->sx, (b[2] - y0) * X->sy, (b[3] - y0) * X->sy };
		for (int a = 0; a < 4; a++) range[o][a] = f[a] < 0 ? 0 : f[a] > NB - 1 ? NB - 1 : (int)f[a];
		total += (long)(range[o][1] - range[o][0] + 1) * (range[o][3] - range[o][2] + 1);
	}
	int *count = calloc(NB * NB, sizeof(int));
	for (int o = 0; o < N; o++)
		for (int u = range[o][0]; u <= range[o][1]; u++) for (int v = range[o][2]; v <= range[o][3]; v++) count[u*NB+v]++;
	for (int c = 0; c < NB * NB; c++) X->start[c+1] = X->start[c] + count[c];
	MALLOC(X->list, 0, total ? total : 1);
	memset(count, 0, NB * NB * sizeof(int));
	for (int o = 0; o < N; o++)
		for (int u = range[o][0]; u <= range[o][1]; u++) for (int v = range[o][2]; v <= range[o][3]; v++)
			X->list[X->start[u*NB+v] + count[u*NB+v]++] = o;
	free(count), free(range);
	return X;
}

static void freeIndex(objIndex *X)
{
	if (!X) return;
	free(X->list), free(X->b), free(X);
}

static inline int binOf(const objIndex *X, float x, float y)
{
	float u = (x - X->x0) * X->sx, v = (y - X->y0) * X->sy;
	int a = u < 0 ? 0 : u > NB - 1 ? NB - 1 : (int)u, b = v < 0 ? 0 : v > NB - 1 ? NB - 1 : (int)v;
	return a * NB + b;
}

static inline int maybeInside(const float *b, float x, float y, float z)
{
	return b[0] <= x && x <= b[1] && b[2] <= y && y <= b[3] && b[4] <= z && z <= b[5];
}

/* The index of the putObjectArray call in progress (read by its parallel loops). */
static objIndex *IDX;

/* 1 if the point (grid coordinates) lies in object o and in no earlier object. */
static float ownedBy(world W, object *O, int o, float i, float j, float k)
{
	float x = W->itox(W, i), y = W->jtoy(W, j), z = W->ktoz(W, k);
	if (!O[o].S(O[o], x, y, z)) return 0;
	if (!IDX) {
		while (o--) if (O[o].S(O[o], x, y, z)) return 0;
		return 1;
	}
	const int c = binOf(IDX, x, y);
	for (int q = IDX->start[c]; q < IDX->start[c+1]; q++) {
		const int p = IDX->list[q];
		if (p >= o) break;
		if (maybeInside(IDX->b[p], x, y, z) && O[p].S(O[p], x, y, z)) return 0;
	}
	return 1;
}

/* Volume fraction of object o in the cell around (i,j,k) (src/object.c:78-121).  With
 * subN >= 2: if the centre and the 8 corners at +-(subN-1)/(2 subN) agree the cell is
 * uniform, otherwise the subN^3 sub-sample average.  subN < 0 selects the edge rule: an
 * empty point counts as filled when both ends of a cell edge through it are filled. */
static float fillFactor(world W, object *O, int o, float i, float j, float k)
{
	float n = ownedBy(W, O, o, i, j, k);
	if (W->subN < 2) {
		if (n || W->subN >= 0) return n;
		if (floorf(i) < ceilf(i)) {
			if (ownedBy(W, O, o, i-0.5, j-0.5, k) && ownedBy(W, O, o, i+0.5, j-0.5, k)) return 1;
			if (ownedBy(W, O, o, i-0.5, j+0.5, k) && ownedBy(W, O, o, i+0.5, j+0.5, k)) return 1;
			if (ownedBy(W, O, o, i-0.5, j, k-0.5) && ownedBy(W, O, o, i+0.5, j, k-0.5)) return 1;
			if (ownedBy(W, O, o, i-0.5, j, k+0.5) && ownedBy(W, O, o, i+0.5, j, k+0.5)) return 1;
		}
		else if (floorf(j) < ceilf(j)) {
			if (ownedBy(W, O, o, i-0.5, j-0.5, k) && ownedBy(W, O, o, i-0.5, j+0.5, k)) return 1;
			if (ownedBy(W, O, o, i+0.5, j-0.5, k) && ownedBy(W, O, o, i+0.5, j+0.5, k)) return 1;
			if (ownedBy(W, O, o, i, j-0.5, k-0.5) && ownedBy(W, O, o, i, j+0.5, k-0.5)) return 1;
			if (ownedBy(W, O, o, i, j-0.5, k+0.5) && ownedBy(W, O, o, i, j+0.5, k+0.5)) return 1;
		}
		else if (floorf(k) < ceilf(k)) {
			if (ownedBy(W, O, o, i-0.5, j, k-0.5) && ownedBy(W, O, o, i-0.5, j, k+0.5)) return 1;
			if (ownedBy(W, O, o, i+0.5, j, k-0.5) && ownedBy(W, O, o, i+0.5, j, k+0.5)) return 1;
			if (ownedBy(W, O, o, i, j-0.5, k-0.5) && ownedBy(W, O, o, i, j-0.5, k+0.5)) return 1;
			if (ownedBy(W, O, o, i, j+0.5, k-0.5) && ownedBy(W, O, o, i, j+0.5, k+0.5)) return 1;
		}
		return 0;
	}

	float h = (W->subN-1.0) / (W->subN*2.0);
	int uniform = 1;
	for (int c = 0; c < 8 && uniform; c++)
		if (ownedBy(W, O, o, c&4 ? i+h : i-h, c&2 ? j+h : j-h, c&1 ? k+h : k-h) != n) uniform = 0;
	if (uniform) return n;

	float portion = 0;
	for (int di = 1-W->subN; di < W->subN; di += 2)
		for (int dj = 1-W->subN; dj < W->subN; dj += 2)
			for (int dk = 1-W->subN; dk < W->subN; dk += 2)
				portion += ownedBy(W, O, o, i+0.5*di/W->subN, j+0.5*dj/W->subN, k+0.5*dk/W->subN);
	return portion / (W->subN * W->subN * W->subN);
}


/* Bounding boxes of the grid nodes owned by each object, widened by -1/+2 cells and
 * clipped to the grid (src/object.c:229-279).  A node belongs to the first object that
 * contains it. */
static box3 *objectBounds(world W, object *O, int N)
{
	box3 *B;
	MALLOC(B, 0, N);
	for (int o = 0; o < N; o++) B[o] = (box3){ W->iMAX, W->iMIN, W->jMAX, W->jMIN, W->kMAX, W->kMIN };

	#pragma omp parallel
	{
		box3 *b;
		MALLOC(b, 0, N);
		for (int o = 0; o < N; o++) b[o] = (box3){ W->iMAX, W->iMIN, W->jMAX, W->jMIN, W->kMAX, W->kMIN };
		#pragma omp for schedule(dynamic)
		for (int i = W->iMIN; i <= W->iMAX; i++) for (int j = W->jMIN; j <= W->jMAX; j++) for (int k = W->kMIN; k <= W->kMAX; k++) {
			float x = W->itox(W, i), y = W->jtoy(W, j), z = W->ktoz(W, k);
			int o = N;
			if (IDX) {
				/* the first object containing the node, among the candidates of its bin */
				const int c = binOf(IDX, x, y);
				for (int q = IDX->start[c]; q < IDX->start[c+1]; q++) {
					const int p = IDX->list[q];
					if (maybeInside(IDX->b[p], x, y, z) && O[p].S(O[p], x, y, z)) { o = p; break; }
				}
			}
			else {
				o = 0;
				while (O[o].S && !O[o].S(O[o], x, y, z)) o++;
			}
			if (o < N) {
				b[o].i0 = MIN(i, b[o].i0), b[o].i1 = MAX(i, b[o].i1);
				b[o].j0 = MIN(j, b[o].j0), b[o].j1 = MAX(j, b[o].j1);
				b[o].k0 = MIN(k, b[o].k0), b[o].k1 = MAX(k, b[o].k1);
			}
		}
		#pragma omp critical
		for (int o = 0; o < N; o++) {
			B[o].i0 = MIN(b[o].i0, B[o].i0), B[o].i1 = MAX(b[o].i1, B[o].i1);
			B[o].j0 = MIN(b[o].j0, B[o].j0), B[o].j1 = MAX(b[o].j1, B[o].j1);
			B[o].k0 = MIN(b[o].k0, B[o].k0), B[o].k1 = MAX(b[o].k1, B[o].k1);
		}
		free(b);
	}
	for (int o = 0; o < N; o++) {
		if (B[o].i0 > B[o].i1 || B[o].j0 > B[o].j1 || B[o].k0 > B[o].k1) {
			printf("\rVolume of object %d/%d is zero!\n", o+1, N);
			exit(0);
		}
		B[o].i0 = MAX(B[o].i0-1, W->iMIN), B[o].i1 = MIN(B[o].i1+2, W->iMAX);
		B[o].j0 = MAX(B[o].j0-1, W->jMIN), B[o].j1 = MIN(B[o].j1+2, W->jMAX);
		B[o].k0 = MAX(B[o].k0-1, W->kMIN), B[o].k1 = MIN(B[o].k1+2, W->kMAX);
	}
	return B;
}


/* Inverse permittivity arrays (background value everywhere) and the per-object tables
 * (src/object.c:124-175).  The x array has one extra node along x, likewise y and z. */
static void makeMediumArrays(world W, vfield *F, coeffs *C, float ex, float ey, float ez)
{
	for (int m = 0; m <= W->complexField; m++) {
		MALLOC(F[m].Jx, 0, C->N); MALLOC(F[m].Jy, 0, C->N); MALLOC(F[m].Jz, 0, C->N);
		MALLOC(F[m].Kx, 0, C->N); MALLOC(F[m].Ky, 0, C->N); MALLOC(F[m].Kz, 0, C->N);
	}
	MALLOC(C->NJ, 0, C->N); MALLOC(C->NK, 0, C->N);
	MALLOC(C->iMin, 0, C->N); MALLOC(C->iMax, 0, C->N);
	MALLOC(C->jMin, 0, C->N); MALLOC(C->jMax, 0, C->N);
	MALLOC(C->kMin, 0, C->N); MALLOC(C->kMax, 0, C->N);
	MALLOC(C->Jx, 0, C->N); MALLOC(C->Jy, 0, C->N); MALLOC(C->Jz, 0, C->N);
	MALLOC(C->ax, 0, C->N); MALLOC(C->ay, 0, C->N); MALLOC(C->az, 0, C->N);
	MALLOC(C->JJx, 0, C->N); MALLOC(C->JJy, 0, C->N); MALLOC(C->JJz, 0, C->N);
	MALLOC(C->JKx, 0, C->N); MALLOC(C->JKy, 0, C->N); MALLOC(C->JKz, 0, C->N);
	MALLOC(C->JEx, 0, C->N); MALLOC(C->JEy, 0, C->N); MALLOC(C->JEz, 0, C->N);
	MALLOC(C->KJx, 0, C->N); MALLOC(C->KJy, 0, C->N); MALLOC(C->KJz, 0, C->N);
	MALLOC(C->KKx, 0, C->N); MALLOC(C->KKy, 0, C->N); MALLOC(C->KKz, 0, C->N);
	MALLOC(C->KEx, 0, C->N); MALLOC(C->KEy, 0, C->N); MALLOC(C->KEz, 0, C->N);

	C->x = makeFieldFilled(W->iMIN, W->iMAX+1, W->jMIN, W->jMAX, W->kMIN, W->kMAX, 1/ex);
	C->y = makeFieldFilled(W->iMIN, W->iMAX, W->jMIN, W->jMAX+1, W->kMIN, W->kMAX, 1/ey);
	C->z = makeFieldFilled(W->iMIN, W->iMAX, W->jMIN, W->jMAX, W->kMIN, W->kMAX+1, 1/ez);
}


/* Storage of dispersive object n: box (one extra node on a side that touches the grid
 * maximum), pole coefficient tables, fill arrays and the J/K currents (src/object.c:178-226). */
static void makeObjectTables(world W, vfield *F, coeffs *C, int n, box3 B)
{
	C->iMin[n] = B.i0, C->iMax[n] = B.i1 + (B.i1 == W->iMAX ? 1 : 0);
	C->jMin[n] = B.j0, C->jMax[n] = B.j1 + (B.j1 == W->jMAX ? 1 : 0);
	C->kMin[n] = B.k0, C->kMax[n] = B.k1 + (B.k1 == W->kMAX ? 1 : 0);
	int nj = C->NJ[n], nk = C->NK[n];

	MALLOC(C->JJx[n], 0, nk); MALLOC(C->JJy[n], 0, nk); MALLOC(C->JJz[n], 0, nk);
	MALLOC(C->JKx[n], 0, nk); MALLOC(C->JKy[n], 0, nk); MALLOC(C->JKz[n], 0, nk);
	MALLOC(C->JEx[n], 0, nk); MALLOC(C->JEy[n], 0, nk); MALLOC(C->JEz[n], 0, nk);
	MALLOC(C->KJx[n], nj, nk-nj); MALLOC(C->KJy[n], nj, nk-nj); MALLOC(C->KJz[n], nj, nk-nj);
	MALLOC(C->KKx[n], nj, nk-nj); MALLOC(C->KKy[n], nj, nk-nj); MALLOC(C->KKz[n], nj, nk-nj);
	MALLOC(C->KEx[n], nj, nk-nj); MALLOC(C->KEy[n], nj, nk-nj); MALLOC(C->KEz[n], nj, nk-nj);

	C->Jx[n] = makeField(C->iMin[n], C->iMax[n], C->jMin[n], C->jMax[n], C->kMin[n], C->kMax[n]);
	C->Jy[n] = makeField(C->iMin[n], C->iMax[n], C->jMin[n], C->jMax[n], C->kMin[n], C->kMax[n]);
	C->Jz[n] = makeField(C->iMin[n], C->iMax[n], C->jMin[n], C->jMax[n], C->kMin[n], C->kMax[n]);

	for (int m = 0; m <= W->complexField; m++) {
		MALLOC(F[m].Jx[n], 0, nk); MALLOC(F[m].Jy[n], 0, nk); MALLOC(F[m].Jz[n], 0, nk);
		MALLOC(F[m].Kx[n], nj, nk-nj); MALLOC(F[m].Ky[n], nj, nk-nj); MALLOC(F[m].Kz[n], nj, nk-nj);
		for (int p = 0; p < nk; p++) {
			F[m].Jx[n][p] = makeField(C->iMin[n], C->iMax[n], C->jMin[n], C->jMax[n], C->kMin[n], C->kMax[n]);
			F[m].Jy[n][p] = makeField(C->iMin[n], C->iMax[n], C->jMin[n], C->jMax[n], C->kMin[n], C->kMax[n]);
			F[m].Jz[n][p] = makeField(C->iMin[n], C->iMax[n], C->jMin[n], C->jMax[n], C->kMin[n], C->kMax[n]);
		}
		for (int p = nj; p < nk; p++) {
			F[m].Kx[n][p] = makeField(C->iMin[n], C->iMax[n], C->jMin[n], C->jMax[n], C->kMin[n], C->kMax[n]);
			F[m].Ky[n][p] = makeField(C->iMin[n], C->iMax[n], C->jMin[n], C->jMax[n], C->kMin[n], C->kMax[n]);
			F[m].Kz[n][p] = makeField(C->iMin[n], C->iMax[n], C->jMin[n], C->jMax[n], C->kMin[n], C->kMax[n]);
		}
	}
}


/* Fill factors on the staggered E positions, conductivity factors and per-pole recursion
 * coefficients of dispersive object n = matter M[o] (src/object.c:290-336).  A Lorentz or
 * critical-point pair J' = a J + b E with complex a, b is carried as real J and K:
 *   J' = Re a J - Im a K + Re b E,   K' = KJ J' + KK K + KE E. */
static void setObjectCoeffs(world W, coeffs *C, matter *M, object *O, int n, int o, int m)
{
	#pragma omp parallel for schedule(dynamic)
	for (int i = C->iMin[n]; i <= C->iMax[n]; i++)
		for (int j = C->jMin[n]; j <= C->jMax[n]; j++)
			for (int k = C->kMin[n]; k <= C->kMax[n]; k++) {
				C->Jx[n][i][j][k] = fillFactor(W, O, o, i-0.5, j, k);
				C->Jy[n][i][j][k] = fillFactor(W, O, o, i, j-0.5, k);
				C->Jz[n][i][j][k] = fillFactor(W, O, o, i, j, k-0.5);
			}

	C->ax[n] = PI * W->dt * conductivity(W, M[o], 0) / M[o].e.x;
	C->ay[n] = PI * W->dt * conductivity(W, M[o], 1) / (M[o].e.y ? M[o].e.y : M[o].e.x);
	C->az[n] = PI * W->dt * conductivity(W, M[o], 2) / (M[o].e.z ? M[o].e.z : M[o].e.x);

	float *jj[3] = { C->JJx[n], C->JJy[n], C->JJz[n] }, *jk[3] = { C->JKx[n], C->JKy[n], C->JKz[n] };
	float *je[3] = { C->JEx[n], C->JEy[n], C->JEz[n] }, *kj[3] = { C->KJx[n], C->KJy[n], C->KJz[n] };
	float *kk[3] = { C->KKx[n], C->KKy[n], C->KKz[n] }, *ke[3] = { C->KEx[n], C->KEy[n], C->KEz[n] };

	for (int p = 0, d = 0, l = C->NJ[n]; p < MAXPOLES; p += 1+2*m) {
		if (!M[o].P[p].w) continue;
		if (!M[o].P[p].f && !M[o].P[p].g) {
			for (int q = 0; q < 3; q++) {
				jj[q][d] = drudeDecay(W, M[o], p+q*m);
				je[q][d] = drudeDrive(W, M[o], p+q*m) / permittivity(W, M[o], q);
			}
			d++;
		}
		else {
			for (int q = 0; q < 3; q++) {
				jj[q][l] = creal(pairDecay(W, M[o], p+q*m));
				jk[q][l] =-cimag(pairDecay(W, M[o], p+q*m));
				je[q][l] = creal(pairDrive(W, M[o], p+q*m)) / permittivity(W, M[o], q);
				kj[q][l] =-jk[q][l] / jj[q][l];
				kk[q][l] = jk[q][l] * jk[q][l] / jj[q][l] + jj[q][l];
				ke[q][l] = je[q][l] * jk[q][l] / jj[q][l] + cimag(pairDrive(W, M[o], p+q*m)) / permittivity(W, M[o], q);
			}
			l++;
		}
	}
}


/* Put objects O[0..N-1] of matter M[0..N-1] on the background matter M[N]; earlier objects
 * take precedence (src/object.c:339-405).  Dispersive objects get ADE tables and their
 * fill factor is stored; plain dielectrics only modify the inverse permittivity. */
void putObjectArray(world W, matter *M, object *O)
{
	alisInvalidate(W);
	removeObjects(W);
	int N;
	for (N = 0; O[N].S; N++) if (dispersiveMatter(M[N])) W->CE->N++;
	printf("\rPreparing to put %d objects.\n", N);
	W->subN = W->Res.spec.subN = W->Res.spec.subN ? W->Res.spec.subN : W->CE->N ? 1 : 12;
	IDX = buildIndex(W, O, N);
	box3 *B = objectBounds(W, O, N);
	coeffs *C = W->CE;
	makeMediumArrays(W, W->E, C, permittivity(W, M[N], 0), permittivity(W, M[N], 1), permittivity(W, M[N], 2));
	MALLOC(C->O, 0, N+1);
	MALLOC(C->M, 0, N+1);
	C->M[N] = M[N];

	for (int o = N-1, n = C->N; o >= 0; o--) {
		C->O[o] = O[o];
		C->M[o] = M[o];
		float cx = permittivity(W, M[o], 0) - permittivity(W, M[N], 0);
		float cy = permittivity(W, M[o], 1) - permittivity(W, M[N], 1);
		float cz = permittivity(W, M[o], 2) - permittivity(W, M[N], 2);
		printf("\rPutting objects: %d left.", o+1);
		fflush(stdout);
		box3 b = B[o];

		if (dispersiveMatter(M[o])) {
			n--;
			int m = M[o].e.y || M[o].e.z ? 1 : 0;
			for (int p = 0; p < MAXPOLES; p++) {
				if (M[o].P[p].w && !M[o].P[p].f && !M[o].P[p].g) C->NJ[n]++;
				if (M[o].P[p].w) C->NK[n]++;
			}
			if (m) C->NJ[n] /= 3, C->NK[n] /= 3;
			makeObjectTables(W, W->E, C, n, b);
			setObjectCoeffs(W, C, M, O, n, o, m);

			#pragma omp parallel for
			for (int i = b.i0; i <= b.i1; i++) for (int j = b.j0; j <= b.j1; j++) for (int k = b.k0; k <= b.k1; k++) {
				C->x[i][j][k] = 1 / (1 / C->x[i][j][k] + cx * C->Jx[n][i][j][k]);
				C->y[i][j][k] = 1 / (1 / C->y[i][j][k] + cy * C->Jy[n][i][j][k]);
				C->z[i][j][k] = 1 / (1 / C->z[i][j][k] + cz * C->Jz[n][i][j][k]);
			}
			if (b.i1 == W->iMAX) for (int j = b.j0; j <= b.j1; j++) for (int k = b.k0; k <= b.k1; k++)
				C->x[b.i1+1][j][k] = 1 / (1 / C->x[b.i1+1][j][k] + cx * C->Jx[n][b.i1+1][j][k]);
			if (b.j1 == W->jMAX) for (int i = b.i0; i <= b.i1; i++) for (int k = b.k0; k <= b.k1; k++)
				C->y[i][b.j1+1][k] = 1 / (1 / C->y[i][b.j1+1][k] + cy * C->Jy[n][i][b.j1+1][k]);
			if (b.k1 == W->kMAX) for (int i = b.i0; i <= b.i1; i++) for (int j = b.j0; j <= b.j1; j++)
				C->z[i][j][b.k1+1] = 1 / (1 / C->z[i][j][b.k1+1] + cz * C->Jz[n][i][j][b.k1+1]);
		}
		else {
			#pragma omp parallel for schedule(dynamic)
			for (int i = b.i0; i <= b.i1; i++) for (int j = b.j0; j <= b.j1; j++) for (int k = b.k0; k <= b.k1; k++) {
				C->x[i][j][k] = 1 / (1 / C->x[i][j][k] + cx * fillFactor(W, O, o, i-0.5, j, k));
				C->y[i][j][k] = 1 / (1 / C->y[i][j][k] + cy * fillFactor(W, O, o, i, j-0.5, k));
				C->z[i][j][k] = 1 / (1 / C->z[i][j][k] + cz * fillFactor(W, O, o, i, j, k-0.5));
			}
			if (b.i1 == W->iMAX) for (int j = b.j0; j <= b.j1; j++) for (int k = b.k0; k <= b.k1; k++)
				C->x[b.i1+1][j][k] = 1 / (1 / C->x[b.i1+1][j][k] + cx * fillFactor(W, O, o, b.i1+0.5, j, k));
			if (b.j1 == W->jMAX) for (int i = b.i0; i <= b.i1; i++) for (int k = b.k0; k <= b.k1; k++)
				C->y[i][b.j1+1][k] = 1 / (1 / C->y[i][b.j1+1][k] + cy * fillFactor(W, O, o, i, b.j1+0.5, k));
			if (b.k1 == W->kMAX) for (int i = b.i0; i <= b.i1; i++) for (int j = b.j0; j <= b.j1; j++)
				C->z[i][j][b.k1+1] = 1 / (1 / C->z[i][j][b.k1+1] + cz * fillFactor(W, O, o, i, j, b.k1+0.5));
		}
	}
	free(B);
	freeIndex(IDX);
	IDX = 0;
	printf("\rPutting objects: %d done.\n", N);
	fflush(stdout);
}


void (putObjects)(world W, ...)
{
	int N = 0;
	va_list ap;
	va_start(ap, W);
	while (va_arg(ap, matter), va_arg(ap, object).S) N++;
	va_end(ap);

	matter *M;
	object *O;
	MALLOC(M, 0, N+1);
	MALLOC(O, 0, N+1);
	va_start(ap, W);
	for (int o = 0; o <= N; o++) {
		M[o] = va_arg(ap, matter);
		O[o] = va_arg(ap, object);
	}
	va_end(ap);
	putObjectArray(W, M, O);
	free(M);
	free(O);
}


static void freeObjectState(world W, vfield *F, coeffs *C)
{
	for (int m = 0; m <= W->complexField; m++) {
		for (int n = 0; n < C->N; n++) {
			for (int p = 0; p < C->NK[n]; p++) {
				removeField(F[m].Jx[n][p], C->iMin[n], C->jMin[n], C->kMin[n]);
				removeField(F[m].Jy[n][p], C->iMin[n], C->jMin[n], C->kMin[n]);
				removeField(F[m].Jz[n][p], C->iMin[n], C->jMin[n], C->kMin[n]);
			}
			for (int p = C->NJ[n]; p < C->NK[n]; p++) {
				removeField(F[m].Kx[n][p], C->iMin[n], C->jMin[n], C->kMin[n]);
				removeField(F[m].Ky[n][p], C->iMin[n], C->jMin[n], C->kMin[n]);
				removeField(F[m].Kz[n][p], C->iMin[n], C->jMin[n], C->kMin[n]);
			}
			free(F[m].Jx[n]); free(F[m].Kx[n]+C->NJ[n]);
			free(F[m].Jy[n]); free(F[m].Ky[n]+C->NJ[n]);
			free(F[m].Jz[n]); free(F[m].Kz[n]+C->NJ[n]);
		}
		free(F[m].Jx); free(F[m].Kx);
		free(F[m].Jy); free(F[m].Ky);
		free(F[m].Jz); free(F[m].Kz);
		F[m].Jx = F[m].Jy = F[m].Jz = F[m].Kx = F[m].Ky = F[m].Kz = 0;
	}
}


static void freeObjectCoeffs(world W, coeffs *C)
{
	for (int n = 0; n < C->N; n++) {
		removeField(C->Jx[n], C->iMin[n], C->jMin[n], C->kMin[n]);
		removeField(C->Jy[n], C->iMin[n], C->jMin[n], C->kMin[n]);
		removeField(C->Jz[n], C->iMin[n], C->jMin[n], C->kMin[n]);
		free(C->JJx[n]); free(C->JKx[n]); free(C->JEx[n]);
		free(C->JJy[n]); free(C->JKy[n]); free(C->JEy[n]);
		free(C->JJz[n]); free(C->JKz[n]); free(C->JEz[n]);
		free(C->KJx[n]+C->NJ[n]); free(C->KKx[n]+C->NJ[n]); free(C->KEx[n]+C->NJ[n]);
		free(C->KJy[n]+C->NJ[n]); free(C->KKy[n]+C->NJ[n]); free(C->KEy[n]+C->NJ[n]);
		free(C->KJz[n]+C->NJ[n]); free(C->KKz[n]+C->NJ[n]); free(C->KEz[n]+C->NJ[n]);
	}
	removeField(C->x, W->iMIN, W->jMIN, W->kMIN);
	removeField(C->y, W->iMIN, W->jMIN, W->kMIN);
	removeField(C->z, W->iMIN, W->jMIN, W->kMIN);
	free(C->NJ); free(C->NK);
	free(C->iMin); free(C->iMax); free(C->jMin); free(C->jMax); free(C->kMin); free(C->kMax);
	free(C->Jx); free(C->Jy); free(C->Jz);
	free(C->ax); free(C->ay); free(C->az);
	free(C->JJx); free(C->JKx); free(C->JEx); free(C->KJx); free(C->KKx); free(C->KEx);
	free(C->JJy); free(C->JKy); free(C->JEy); free(C->KJy); free(C->KKy); free(C->KEy);
	free(C->JJz); free(C->JKz); free(C->JEz); free(C->KJz); free(C->KKz); free(C->KEz);
	free(C->O); free(C->M);
	C->O = 0, C->M = 0, C->N = 0;
}


void removeObjects(world W)
{
	if (W->CE->O && W->CE->M) {
		alisInvalidate(W);
		freeObjectState(W, W->E, W->CE);
		freeObjectCoeffs(W, W->CE);
	}
}


/* Temperature scaling of the damping terms (src/object.c:500-507). */
matter LT(matter M, float T)
{
	for (int p = 0; p < MAXPOLES; p++) M.P[p].r *= T, M.P[p].g *= T;
	return M;
}


/* Matter from a tabulated (lambda, n, k) file, linearly interpolated at wavelength l,
 * as a dielectric with a conductivity pole at photon energy f (src/object.c:510-549). */
matter nkFile(char *name, float l, float f)
{
	int N = 0, cap = 0;
	char line[1024];
	float n = 0, k = 0, *la = 0, *na = 0, *ka = 0;
	FILE *fp = fopen(name, "r");
	if (!fp) {
		printf("The (n,k) file not found.\n");
		exit(0);
	}
	while (fgets(line, 256, fp)) {
		if (N >= cap) {
			cap += 256;
			la = realloc(la, cap*sizeof(float));
			na = realloc(na, cap*sizeof(float));
			ka = realloc(ka, cap*sizeof(float));
		}
		char *s = line;
		la[N] = atof(s);
		s = strchr(s,'\t') ? strchr(s,'\t')+1 : strchr(s,' ') ? strchr(s,' ')+1 : strchr(s,',')+1;
		na[N] = atof(s);
		s = strchr(s,'\t') ? strchr(s,'\t')+1 : strchr(s,' ') ? strchr(s,' ')+1 : strchr(s,',')+1;
		ka[N] = atof(s);
		N++;
	}
	fclose(fp);
	int i = 1;
	while (i < N-1 && l > la[i]) i++;
	n = (na[i-1]*(la[i]-l) + na[i]*(l-la[i-1])) / (la[i]-la[i-1]);
	k = (ka[i-1]*(la[i]-l) + ka[i]*(l-la[i-1])) / (la[i]-la[i-1]);
	free(la); free(na); free(ka);
	return (matter) {{n*n-k*k}, {{0, 2*n*k*((float)f)}}};
}
