/*
 * Near-to-far transformation by reciprocity in a z-layered background (src/farfld.c:14-662).
 * For a direction (theta, phi) and polarisation, a discrete plane wave T(z) of the layered
 * grid is obtained by a 1-D recursion (Ppol/Spol); the far field is the overlap of T with
 * the phaser's surface DFT fields on the PML-bounded faces, with separable in-plane phase
 * factors.  Evaluated here in double precision on double accumulators.
 */
#include <string.h>
#include <sys/stat.h>
#include "internal.h"
#include "phaserview.h"

typedef double complex dcx;

/* Projections of image pixels to (theta, phi) in degrees (src/farfld.c:14-41). */
void CylindericalMap(int x, int y, int X, int Y, float angle[2])
{
	angle[0] = 180.0 - 180.0 * y / Y;
	angle[1] = 360.0 * x / X;
}

void NorthernAzimuthalMap(int x, int y, int X, int Y, float angle[2])
{
	angle[0] = 180 * sqrtf(sq(x/(X-1.0)-0.5) + sq(y/(Y-1.0)-0.5));
	angle[1] = 180 * atan2f(y/(Y-1.0)-0.5, x/(X-1.0)-0.5) / PI;
	if (angle[0] > 90) angle[0] = INF;
}

void SouthernAzimuthalMap(int x, int y, int X, int Y, float angle[2])
{
	angle[0] = 180 - 180 * sqrtf(sq(x/(X-1.0)-0.5) + sq(y/(Y-1.0)-0.5));
	angle[1] = 180 * atan2f(y/(Y-1.0)-0.5, x/(X-1.0)-0.5) / PI;
	if (angle[0] < 90) angle[0] = INF;
}

void AzimuthalMap(int x, int y, int X, int Y, float angle[2])
{
	if (x < (X+1)/2) NorthernAzimuthalMap(x, y, (X+1)/2, Y, angle);
	if (x > X/2) SouthernAzimuthalMap(x-X/2, y, (X+1)/2, Y, angle);
}

/* Rotation of a direction by (a, b) about the axis at polar angle theta (src/farfld.c:44-52). */
static float rotateA(float theta, float a, float b)
{
	return acosf(cosf(a*PI/180)*cosf(theta*PI/180)-(sinf(a*PI/180)*cosf(b*PI/180))*sinf(theta*PI/180))*180/PI;
}

static float rotateB(float theta, float a, float b)
{
	return atan2f(sinf(a*PI/180)*sinf(b*PI/180), (sinf(a*PI/180)*cosf(b*PI/180))*cosf(theta*PI/180)+cosf(a*PI/180)*sinf(theta*PI/180))*180/PI;
}


/* Permittivity profile along z at the column (iMax, jMax) of the phaser box, with the
 * dispersive poles evaluated at the phaser frequency, and the cell widths; mirrored below
 * a z symmetry plane (src/farfld.c:74-117). */
static void farfieldCoeffs(phaser P, world W)
{
	P->kMIN = (W->zMinSur==SYM) ? -P->kMax : P->kMin;
	MALLOC(P->dE, P->kMIN, 2-P->kMIN+P->kMax);
	MALLOC(P->dH, P->kMIN, 2-P->kMIN+P->kMax);
	MALLOC(P->ex, P->kMIN, 2-P->kMIN+P->kMax);
	MALLOC(P->ez, P->kMIN, 2-P->kMIN+P->kMax);
	for (int k = P->kMin; k <= P->kMax+1; k++) {
		int kc = MIN(k, W->kMAX);
		P->dE[k] = 1 / W->CE->dtdz[kc];
		P->dH[k] = 1 / W->CH->dtdz[kc];
		P->ex[k] = W->CE->x ? 1/W->CE->x[P->iMax][P->jMax][k] : 1;
		P->ez[k] = W->CE->z ? 1/W->CE->z[P->iMax][P->jMax][k] : 1;
		for (int n = 0, o = 0; n < W->CE->N; n++, o++) {
			while (!dispersiveMatter(W->CE->M[o])) o++;
			if (W->CE->iMin[n]<=P->iMax && P->iMax<=W->CE->iMax[n] && W->CE->jMin[n]<=P->jMax && P->jMax<=W->CE->jMax[n] &&
			    W->CE->kMin[n]<=k && k<=W->CE->kMax[n]) {
				matter M = W->CE->M[o];
				int m = M.e.y || M.e.z ? 1 : 0;
				P->ex[k] += -W->CE->ax[n] * M.e.x * W->CE->Jx[n][P->iMax][P->jMax][k];
				P->ez[k] += -W->CE->az[n] * (M.e.z ? M.e.z : M.e.x) * W->CE->Jz[n][P->iMax][P->jMax][k];
				for (int p = 0; p < MAXPOLES; p += 1+2*m) {
					P->ex[k] += polePermittivity(M, W->eV * P->f, p+0*m) * W->CE->Jx[n][P->iMax][P->jMax][k];
					P->ez[k] += polePermittivity(M, W->eV * P->f, p+2*m) * W->CE->Jz[n][P->iMax][P->jMax][k];
				}
			}
		}
	}
	if (W->zMinSur==SYM)
		for (int k = P->kMIN; k < P->kMin; k++) {
			P->dE[k] = P->dE[-k];
			P->dH[k] = P->dH[1-k];
			P->ex[k] = P->ex[-k];
			P->ez[k] = P->ez[1-k];
		}
}

/* 1-D plane-wave profile over k = kMIN..kMax+1 (double). */
typedef struct { double A; dcx *Ex, *Ey, *Ez, *Hx, *Hy, *Hz; } profile;

static dcx *newProfile(phaser P) { dcx *a = calloc(P->kMax + 2 - P->kMIN, sizeof(dcx)); return a - P->kMIN; }
static void freeProfile(phaser P, dcx *a) { if (a) free(a + P->kMIN); }
static void freeProfiles(phaser P, profile *T)
{
	freeProfile(P, T->Ex), freeProfile(P, T->Ey), freeProfile(P, T->Ez);
	freeProfile(P, T->Hx), freeProfile(P, T->Hy), freeProfile(P, T->Hz);
}

/* TM (p) plane wave leaving the layered grid in direction (theta, phi): upward recursion
 * from the bottom for theta < 90 (downward from the top otherwise) of the discrete
 * E_x/H_y pair, rescaled whenever |H| exceeds 1e10; A normalises the outgoing amplitude
 * at the far side (src/farfld.c:126-185). */
static profile pPolD(phaser P, double theta, double phi)
{
	profile T = {0};
	const int k0 = P->kMIN, k1 = P->kMax;
	const double sn = sin(M_PI * P->f * P->W->dt);
	double sinth = 0, sin2th = 0;
	T.Ex = newProfile(P), T.Hy = newProfile(P);
	dcx *Ex = T.Ex, *Hy = T.Hy;
#define EX(k) ((dcx)P->ex[k])
#define EZ(k) ((dcx)P->ez[k])
	if (theta < 90 || 270 < theta) {
		sinth = sin(theta*M_PI/180) * sqrt(creal(EX(k1))), sin2th = sq(sinth);
		Hy[k0] = -1;
		Ex[k0] = csqrt((1-sin2th/EZ(k0)) / EX(k0));
		Ex[k0] *= cexp(-I * casin(sn * csqrt(EX(k0)-sin2th) * P->dE[k0]));
		for (int k = k0; k <= k1; k++) {
			Hy[k+1] = Hy[k] + 2 * I * sn * P->dE[k] * Ex[k] * EX(k);
			Ex[k+1] = Ex[k] + 2 * I * sn * P->dH[k+1] * Hy[k+1] * (1-sin2th/EZ(k+1));
			if (cabs(Hy[k+1]) > 1E10) for (int kk = k0; kk <= k+1; kk++) Ex[kk] /= 1E10, Hy[kk] /= 1E10;
		}
		T.A = sqrt(sqrt(creal(EX(k1))));
		T.A /= cabs(0.5*(Hy[k1]+Hy[k1+1]) / sqrt(creal(EX(k1))) - Ex[k1] / cos(theta*M_PI/180));
	}
	if (90 < theta && theta < 270) {
		sinth = sin(theta*M_PI/180) * sqrt(creal(EX(k0))), sin2th = sq(sinth);
		Hy[k1+1] = 1;
		Ex[k1+1] = csqrt((1-sin2th/EZ(k1+1)) / EX(k1+1));
		Ex[k1+1] *= cexp(I * casin(sn * csqrt(EX(k1+1)-sin2th) * P->dE[k1+1]));
		for (int k = k1; k >= k0; k--) {
			Ex[k] = Ex[k+1] - 2 * I * sn * P->dH[k+1] * Hy[k+1] * (1-sin2th/EZ(k+1));
			Hy[k] = Hy[k+1] - 2 * I * sn * P->dE[k] * Ex[k] * EX(k);
			if (cabs(Hy[k+1]) > 1E10) for (int kk = k1; kk >= k; kk--) Ex[kk] /= 1E10, Hy[kk] /= 1E10;
		}
		T.A = sqrt(sqrt(creal(EX(k0))));
		T.A /= cabs(0.5*(Hy[k0]+Hy[k0+1]) / sqrt(creal(EX(k0))) - Ex[k0] / cos(theta*M_PI/180));
	}
	if (theta) {
		T.Ez = newProfile(P);
		for (int k = k0; k <= k1+1; k++) T.Ez[k] = Hy[k] * sinth / EZ(k);
	}
	if (phi) {
		double sp = sin(phi*M_PI/180), cp = cos(phi*M_PI/180);
		T.Hx = newProfile(P), T.Ey = newProfile(P);
		for (int k = k0; k <= k1+1; k++) {
			T.Ey[k] = Ex[k] * sp;
			Ex[k] = Ex[k] * cp;
			T.Hx[k] = -Hy[k] * sp;
			Hy[k] = Hy[k] * cp;
		}
	}
	return T;
}

/* TE (s) counterpart with the E_y/H_x pair (src/farfld.c:188-245). */
static profile sPolD(phaser P, double theta, double phi)
{
	profile T = {0};
	const int k0 = P->kMIN, k1 = P->kMax;
	const double sn = sin(M_PI * P->f * P->W->dt);
	double sinth = 0, sin2th = 0;
	T.Hx = newProfile(P), T.Ey = newProfile(P);
	dcx *Ey = T.Ey, *Hx = T.Hx;
	if (theta < 90 || 270 < theta) {
		sinth = sin(theta*M_PI/180) * sqrt(creal(EX(k1))), sin2th = sq(sinth);
		Hx[k0] = csqrt(EX(k0) - sin2th);
		Ey[k0] = cexp(-I * casin(sn * csqrt(EX(k0)-sin2th) * P->dE[k0]));
		for (int k = k0; k <= k1; k++) {
			Hx[k+1] = Hx[k] - 2 * I * sn * P->dE[k] * Ey[k] * (EX(k) - sin2th);
			Ey[k+1] = Ey[k] - 2 * I * sn * P->dH[k+1] * Hx[k+1];
			if (cabs(Hx[k+1]) > 1E10) for (int kk = k0; kk <= k+1; kk++) Ey[kk] /= 1E10, Hx[kk] /= 1E10;
		}
		T.A = sqrt(sqrt(creal(EX(k1))));
		T.A /= cabs(Ey[k1] + 0.5*(Hx[k1]+Hx[k1+1]) / (sqrt(creal(EX(k1))) * cos(theta*M_PI/180)));
	}
	if (90 < theta && theta < 270) {
		sinth = sin(theta*M_PI/180) * sqrt(creal(EX(k0))), sin2th = sq(sinth);
		Ey[k1+1] = cexp(I * casin(sn * csqrt(EX(k1+1)-sin2th) * P->dE[k1+1]));
		Hx[k1+1] = -csqrt(EX(k1+1) - sin2th);
		for (int k = k1; k >= k0; k--) {
			Ey[k] = Ey[k+1] + 2 * I * sn * P->dH[k+1] * Hx[k+1];
			Hx[k] = Hx[k+1] + 2 * I * sn * P->dE[k] * Ey[k] * (EX(k) - sin2th);
			if (cabs(Hx[k+1]) > 1E10) for (int kk = k1; kk >= k; kk--) Ey[kk] /= 1E10, Hx[kk] /= 1E10;
		}
		T.A = sqrt(sqrt(creal(EX(k0))));
		T.A /= cabs(Ey[k0] + 0.5*(Hx[k0]+Hx[k0+1]) / (sqrt(creal(EX(k0))) * cos(theta*M_PI/180)));
	}
	if (theta) {
		T.Hz = newProfile(P);
		for (int k = k0; k <= k1+1; k++) T.Hz[k] = -Ey[k] * sinth;
	}
	if (phi) {
		double sp = sin(phi*M_PI/180), cp = cos(phi*M_PI/180);
		T.Ex = newProfile(P), T.Hy = newProfile(P);
		for (int k = k0; k <= k1+1; k++) {
			T.Ex[k] = -Ey[k] * sp;
			Ey[k] = Ey[k] * cp;
			T.Hy[k] = Hx[k] * sp;
			Hx[k] = Hx[k] * cp;
		}
	}
	return T;
#undef EX
#undef EZ
}

/* Float profiles of the public API from the double ones. */
static float complex *toFloatProfile(phaser P, dcx *a)
{
	if (!a) return 0;
	float complex *f;
	MALLOC(f, P->kMIN, 2-P->kMIN+P->kMax);
	for (int k = P->kMIN; k <= P->kMax+1; k++) f[k] = a[k];
	return f;
}

static tfield publicProfile(phaser P, profile T)
{
	tfield F = {0};
	F.A = T.A;
	F.Ex = toFloatProfile(P, T.Ex), F.Ey = toFloatProfile(P, T.Ey), F.Ez = toFloatProfile(P, T.Ez);
	F.Hx = toFloatProfile(P, T.Hx), F.Hy = toFloatProfile(P, T.Hy), F.Hz = toFloatProfile(P, T.Hz);
	freeProfiles(P, &T);
	return F;
}

tfield Unpol(phaser P, float theta, float phi) { (void)P, (void)theta, (void)phi; return (tfield) {0}; }
tfield Ppol(phaser P, float theta, float phi) { if (!P->dE) farfieldCoeffs(P, P->W); return publicProfile(P, pPolD(P, theta, phi)); }
tfield Spol(phaser P, float theta, float phi) { if (!P->dE) farfieldCoeffs(P, P->W); return publicProfile(P, sPolD(P, theta, phi)); }

static dcx *widen(phaser P, float complex *a)
{
	if (!a) return 0;
	dcx *d = newProfile(P);
	for (int k = P->kMIN; k <= P->kMax+1; k++) d[k] = a[k];
	free(&a[P->kMIN]);
	return d;
}

static profile profileFor(phaser P, tfield pol(phaser,float,float), double theta, double phi)
{
	if (pol == Ppol) return pPolD(P, theta, phi);
	if (pol == Spol) return sPolD(P, theta, phi);
	tfield F = pol(P, theta, phi);
	profile T = { F.A, widen(P, F.Ex), widen(P, F.Ey), widen(P, F.Ez), widen(P, F.Hx), widen(P, F.Hy), widen(P, F.Hz) };
	return T;
}


/* Surface sums of one face of each orientation (src/farfld.c:248-459).  v/g/q address the
 * double accumulators (components E1 E2 H1 H2 of the face). */
typedef struct { const double *v; const phaserGeom *g; int q; } acc;

/* Zeroed complex array indexed lo..hi; release with free(a + lo). */
static dcx *dcxSpan(int lo, int hi) { dcx *a = calloc(hi >= lo ? hi-lo+1 : 1, sizeof(dcx)); return a - lo; }

static dcx F4(const acc *a, int kind, int comp, int u, int w) { return faceValue(a->v, a->g, a->q, kind, comp, u, w); }

static void addC(dcx *s, dcx *c, dcx x)
{
	double sr = creal(*s), cr = creal(*c), si = cimag(*s), ci = cimag(*c);
	neumaierAdd(&sr, &cr, creal(x));
	neumaierAdd(&si, &ci, cimag(x));
	*s = sr + I * si, *c = cr + I * ci;
}

static dcx faceX(phaser P, const acc *a, int kind, int i, profile *T, double kx, double ky)
{
	world W = P->W;
	if (!a->g->on[kind]) return 0;
	int j0 = P->jMin, j1 = P->jMax, k0 = P->kMin, k1 = P->kMax;
	dcx *Y = dcxSpan(j0, j1), *Z = dcxSpan(k0, k1);
	dcx EyHz = 0, HzEy = 0, EzHy = 0, HyEz = 0, c = 0;
	if (T->Hz || T->Ey) {
		for (int j = j0+1; j <= j1; j++) Y[j] = W->CH->dy[j] * cexp(I*ky*W->jtoy(W, j-0.5));
		if (W->yMinSur==SYM) for (int j = j0+1; j <= j1; j++) Y[j] += conj(Y[j]) * W->cosky;
	}
	for (int pass = 0; pass < 2; pass++) {
		dcx *Tp = pass ? T->Ey : T->Hz;
		if (!Tp) continue;
		for (int k = k0; k <= k1; k++) Z[k] = W->CE->dz[k] * Tp[k];
		if (W->zMinSur==SYM) for (int k = k0; k <= k1; k++) Z[k] -= W->CE->dz[k] * Tp[-k] * W->coskz;
		int comp = pass ? 3 : 0;   /* E1 = Ey with T.Hz, H2 = Hz with T.Ey */
		dcx s = 0; c = 0;
		for (int j = j0+1; j <= j1; j++) {
			dcx r = 0.5 * (F4(a, kind, comp, j-j0, 0) * Z[k0] + F4(a, kind, comp, j-j0, k1-k0) * Z[k1]);
			for (int k = k0+1; k < k1; k++) r += F4(a, kind, comp, j-j0, k-k0) * Z[k];
			addC(&s, &c, Y[j] * r);
		}
		if (pass) HzEy = s + c; else EyHz = s + c;
	}
	if (T->Hy || T->Ez) {
		for (int j = j0; j <= j1; j++) Y[j] = W->CE->dy[j] * cexp(I*ky*W->jtoy(W, j));
		if (W->yMinSur==SYM) for (int j = j0; j <= j1; j++) Y[j] -= conj(Y[j]) * W->cosky;
	}
	for (int pass = 0; pass < 2; pass++) {
		dcx *Tp = pass ? T->Ez : T->Hy;
		if (!Tp) continue;
		for (int k = k0+1; k <= k1; k++) Z[k] = W->CH->dz[k] * Tp[k];
		if (W->zMinSur==SYM) for (int k = k0+1; k <= k1; k++) Z[k] += W->CH->dz[k] * Tp[1-k] * W->coskz;
		int comp = pass ? 2 : 1;   /* E2 = Ez with T.Hy, H1 = Hy with T.Ez */
		dcx s = 0; c = 0;
		for (int k = k0+1; k <= k1; k++) {
			dcx r = 0.5 * (F4(a, kind, comp, 0, k-k0) * Y[j0] + F4(a, kind, comp, j1-j0, k-k0) * Y[j1]);
			for (int j = j0+1; j < j1; j++) r += F4(a, kind, comp, j-j0, k-k0) * Y[j];
			addC(&s, &c, Z[k] * r);
		}
		if (pass) HyEz = s + c; else EzHy = s + c;
	}
	free(Y + j0), free(Z + k0);
	double x = W->itox(W, i);
	dcx Sum = (EyHz-HzEy-EzHy+HyEz) * cexp(I*kx*x);
	if (W->xMinSur==SYM) Sum += (EyHz+HzEy-EzHy-HyEz) * cexp(-I*kx*x) * W->coskx;
	return Sum;
}

static dcx faceY(phaser P, const acc *a, int kind, int j, profile *T, double kx, double ky)
{
	world W = P->W;
	if (!a->g->on[kind]) return 0;
	int i0 = P->iMin, i1 = P->iMax, k0 = P->kMin, k1 = P->kMax;
	dcx *X = dcxSpan(i0, i1), *Z = dcxSpan(k0, k1);
	dcx EzHx = 0, HxEz = 0, ExHz = 0, HzEx = 0, c = 0;
	if (T->Hx || T->Ez) {
		for (int i = i0; i <= i1; i++) X[i] = W->CE->dx[i] * cexp(I*kx*W->itox(W, i));
		if (W->xMinSur==SYM) for (int i = i0; i <= i1; i++) X[i] -= conj(X[i]) * W->coskx;
	}
	for (int pass = 0; pass < 2; pass++) {
		dcx *Tp = pass ? T->Ez : T->Hx;
		if (!Tp) continue;
		for (int k = k0+1; k <= k1; k++) Z[k] = W->CH->dz[k] * Tp[k];
		if (W->zMinSur==SYM) for (int k = k0+1; k <= k1; k++) Z[k] += W->CH->dz[k] * Tp[1-k] * W->coskz;
		int comp = pass ? 3 : 0;   /* E1 = Ez with T.Hx, H2 = Hx with T.Ez */
		dcx s = 0; c = 0;
		for (int k = k0+1; k <= k1; k++) {
			dcx r = 0.5 * (F4(a, kind, comp, 0, k-k0) * X[i0] + F4(a, kind, comp, i1-i0, k-k0) * X[i1]);
			for (int i = i0+1; i < i1; i++) r += F4(a, kind, comp, i-i0, k-k0) * X[i];
			addC(&s, &c, Z[k] * r);
		}
		if (pass) HxEz = s + c; else EzHx = s + c;
	}
	if (T->Hz || T->Ex) {
		for (int i = i0+1; i <= i1; i++) X[i] = W->CH->dx[i] * cexp(I*kx*W->itox(W, i-0.5));
		if (W->xMinSur==SYM) for (int i = i0+1; i <= i1; i++) X[i] += conj(X[i]) * W->coskx;
	}
	for (int pass = 0; pass < 2; pass++) {
		dcx *Tp = pass ? T->Ex : T->Hz;
		if (!Tp) continue;
		for (int k = k0; k <= k1; k++) Z[k] = W->CE->dz[k] * Tp[k];
		if (W->zMinSur==SYM) for (int k = k0; k <= k1; k++) Z[k] -= W->CE->dz[k] * Tp[-k] * W->coskz;
		int comp = pass ? 2 : 1;   /* E2 = Ex with T.Hz, H1 = Hz with T.Ex */
		dcx s = 0; c = 0;
		for (int i = i0+1; i <= i1; i++) {
			dcx r = 0.5 * (F4(a, kind, comp, i-i0, 0) * Z[k0] + F4(a, kind, comp, i-i0, k1-k0) * Z[k1]);
			for (int k = k0+1; k < k1; k++) r += F4(a, kind, comp, i-i0, k-k0) * Z[k];
			addC(&s, &c, X[i] * r);
		}
		if (pass) HzEx = s + c; else ExHz = s + c;
	}
	free(X + i0), free(Z + k0);
	double y = W->jtoy(W, j);
	dcx Sum = (EzHx-HxEz-ExHz+HzEx) * cexp(I*ky*y);
	if (W->yMinSur==SYM) Sum += (EzHx+HxEz-ExHz-HzEx) * cexp(-I*ky*y) * W->cosky;
	return Sum;
}

static dcx faceZ(phaser P, const acc *a, int kind, int k, profile *T, double kx, double ky)
{
	world W = P->W;
	if (!a->g->on[kind]) return 0;
	int i0 = P->iMin, i1 = P->iMax, j0 = P->jMin, j1 = P->jMax;
	dcx *X = dcxSpan(i0, i1), *Y = dcxSpan(j0, j1);
	dcx Sum = 0, c = 0;
	int sym = W->zMinSur==SYM;
	if (T->Hy || T->Ex) {
		for (int i = i0+1; i <= i1; i++) X[i] = W->CH->dx[i] * cexp(I*kx*W->itox(W, i-0.5));
		if (W->xMinSur==SYM) for (int i = i0+1; i <= i1; i++) X[i] += conj(X[i]) * W->coskx;
		for (int j = j0; j <= j1; j++) Y[j] = W->CE->dy[j] * cexp(I*ky*W->jtoy(W, j));
		if (W->yMinSur==SYM) for (int j = j0; j <= j1; j++) Y[j] -= conj(Y[j]) * W->cosky;
		for (int pass = 0; pass < 2; pass++) {
			dcx *Tp = pass ? T->Ex : T->Hy;
			if (!Tp) continue;
			int comp = pass ? 3 : 0;   /* E1 = Ex with T.Hy, H2 = Hy with T.Ex */
			dcx s = 0; c = 0;
			for (int i = i0+1; i <= i1; i++) {
				dcx r = 0.5 * (F4(a, kind, comp, i-i0, 0) * Y[j0] + F4(a, kind, comp, i-i0, j1-j0) * Y[j1]);
				for (int j = j0+1; j < j1; j++) r += F4(a, kind, comp, i-i0, j-j0) * Y[j];
				addC(&s, &c, X[i] * r);
			}
			dcx S = s + c;
			if (!pass) {
				Sum += S * 0.5 * (Tp[k]+Tp[k+1]);
				if (sym) Sum += S * 0.5 * (Tp[1-k]+Tp[-k]) * W->coskz;
			}
			else {
				Sum -= S * Tp[k];
				if (sym) Sum += S * Tp[-k] * W->coskz;
			}
		}
	}
	if (T->Hx || T->Ey) {
		for (int i = i0; i <= i1; i++) X[i] = W->CE->dx[i] * cexp(I*kx*W->itox(W, i));
		if (W->xMinSur==SYM) for (int i = i0; i <= i1; i++) X[i] -= conj(X[i]) * W->coskx;
		for (int j = j0+1; j <= j1; j++) Y[j] = W->CH->dy[j] * cexp(I*ky*W->jtoy(W, j-0.5));
		if (W->yMinSur==SYM) for (int j = j0+1; j <= j1; j++) Y[j] += conj(Y[j]) * W->cosky;
		for (int pass = 0; pass < 2; pass++) {
			dcx *Tp = pass ? T->Ey : T->Hx;
			if (!Tp) continue;
			int comp = pass ? 2 : 1;   /* E2 = Ey with T.Hx, H1 = Hx with T.Ey */
			dcx s = 0; c = 0;
			for (int j = j0+1; j <= j1; j++) {
				dcx r = 0.5 * (F4(a, kind, comp, 0, j-j0) * X[i0] + F4(a, kind, comp, i1-i0, j-j0) * X[i1]);
				for (int i = i0+1; i < i1; i++) r += F4(a, kind, comp, i-i0, j-j0) * X[i];
				addC(&s, &c, Y[j] * r);
			}
			dcx S = s + c;
			if (!pass) {
				Sum -= S * 0.5 * (Tp[k]+Tp[k+1]);
				if (sym) Sum -= S * 0.5 * (Tp[1-k]+Tp[-k]) * W->coskz;
			}
			else {
				Sum += S * Tp[k];
				if (sym) Sum -= S * Tp[-k] * W->coskz;
			}
		}
	}
	free(X + i0), free(Y + j0);
	return Sum;
}

/* Complex far-field amplitude of phaser P for polarisation pol at (theta, phi) in degrees,
 * including the periodic array factor (src/farfld.c:462-514). */
dcx farFieldD(phaser P, tfield pol(phaser,float,float), double theta, double phi)
{
	if (!P->dE) farfieldCoeffs(P, P->W);
	if (theta == INF || phi == INF) return 0;
	while (theta < 0) theta += 360;
	while (theta >= 360) theta -= 360;
	while (phi < 0) phi += 360;
	while (phi >= 360) phi -= 360;
	if ( 89 < theta && theta <  91) return 0.5 * (farFieldD(P, pol,  89, phi) + farFieldD(P, pol,  91, phi));
	if (269 < theta && theta < 271) return 0.5 * (farFieldD(P, pol, 269, phi) + farFieldD(P, pol, 271, phi));
	double sinth = 0;
	if (theta < 90 || 270 < theta) {
		if (cimagf(P->ex[P->kMax])) return 0;
		sinth = sin(theta*M_PI/180) * sqrt(crealf(P->ex[P->kMax]));
	}
	if (90 < theta && theta < 270) {
		if (cimagf(P->ex[P->kMIN])) return 0;
		sinth = sin(theta*M_PI/180) * sqrt(crealf(P->ex[P->kMIN]));
	}
	phaser B = phaserBase(P);
	if (!B) return 0;
	phaserGeom g;
	acc a = { phaserDoubles(B, &g), &g, (int)(P - B) };
	profile T = profileFor(P, pol, theta, phi);
	double kx = -2 * M_PI * P->f * sinth * cos(phi*M_PI/180);
	double ky = -2 * M_PI * P->f * sinth * sin(phi*M_PI/180);
	world W = P->W;
	dcx Sum = 0;
	if (W->xMinSur==PML) Sum -= faceX(P, &a, 0, P->iMin, &T, kx, ky);
	if (W->xMaxSur==PML) Sum += faceX(P, &a, 1, P->iMax, &T, kx, ky);
	if (W->yMinSur==PML) Sum -= faceY(P, &a, 2, P->jMin, &T, kx, ky);
	if (W->yMaxSur==PML) Sum += faceY(P, &a, 3, P->jMax, &T, kx, ky);
	if (W->zMinSur==PML) Sum -= faceZ(P, &a, 4, P->kMin, &T, kx, ky);
	if (W->zMaxSur==PML) Sum += faceZ(P, &a, 5, P->kMax, &T, kx, ky);
	freeProfiles(P, &T);
	if (P->xNUM) {
		dcx s = 0, q = 1, p = P->expikx * cexp(I*kx*(W->xMax-W->xMin));
		for (int x = 0; x < P->xNUM; x++) s += (q *= p);
		Sum *= s / P->xNUM;
	}
	if (P->yNUM) {
		dcx s = 0, q = 1, p = P->expiky * cexp(I*ky*(W->yMax-W->yMin));
		for (int y = 0; y < P->yNUM; y++) s += (q *= p);
		Sum *= s / P->yNUM;
	}
	return Sum * T.A * P->A * P->f;
}

float complex farField(phaser P, tfield pol(phaser,float,float), float theta, float phi)
{
	return farFieldD(P, pol, theta, phi);
}

double farFieldID(phaser P, tfield pol(phaser,float,float), double theta, double phi)
{
	if (pol == Unpol) return farFieldID(P, Ppol, theta, phi) + farFieldID(P, Spol, theta, phi);
	dcx f = farFieldD(P, pol, theta, phi);
	return creal(conj(f) * f);
}

float farFieldI(phaser P, tfield pol(phaser,float,float), float theta, float phi)
{
	return farFieldID(P, pol, theta, phi);
}

/* Prepare P for concurrent evaluation (coefficients and the accumulator view). */
static void farFieldReady(phaser P)
{
	if (!P->dE) farfieldCoeffs(P, P->W);
	phaser B = phaserBase(P);
	phaserGeom g;
	if (B) phaserDoubles(B, &g);
}

/* Far-field images over a projection map: X, Y, circular and linear polarisations
 * (src/farfld.c:525-594). */
void farFieldProfile(world W, phaser P, projmap projMap, int X, int Y, ...)
{
	char fieldname[4], filename[1024], fullname[2048];
	va_list ap;
	va_start(ap, Y);
	method M = va_arg(ap, method);
	unsigned long int *cmap = (M == png) ? va_arg(ap, unsigned long int *) : 0;
	float Max = (M == png) ? va_arg(ap, double) : 1;
	char *format = va_arg(ap, char*);
	vsnprintf(filename, sizeof filename, format, ap);
	va_end(ap);
	snprintf(fullname, sizeof fullname, "%s%s", W->ID, filename);
	strcpy(fieldname, strstr(filename, "Log") ? "Log" : "");
	if (filename[0]=='/') mkdir(W->ID, 0755);
	if (fullname[strlen(fullname)-1]=='/') mkdir(fullname, 0755);

	slice S;
	MALLOC(S, 0, 1);
	*S = (struct slice) {0, 0, 0, X-1, 0, Y-1, 0, 0, 0, 0, X-1, X, 0, Y-1, Y};
	float ***maps[7];
	for (int a = 0; a < 7; a++) maps[a] = makeField(S->iMIN, S->iMAX, S->jMIN, S->jMAX, S->kMIN, S->kMAX);
	float ***Up = maps[0], ***Pp = maps[1], ***Sp = maps[2], ***Lp = maps[3], ***Rp = maps[4], ***Xp = maps[5], ***Yp = maps[6];

	int N = 0;
	float angle[2];
	for (int x = 0; x < X; x++) for (int y = 0; y < Y; y++) {
		projMap(x, y, X, Y, angle);
		if (angle[0] != INF) N++;
	}
	printf("Performing Near-to-Far-Field Transformation.\n");
	farFieldReady(P);
	int logScale = fieldname[0] != 0;
	float max = 0;
	#pragma omp parallel for schedule(dynamic) reduction(max:max)
	for (long xy = 0; xy < (long)X*Y; xy++) {
		int x = xy / Y, y = xy % Y;
		float ang[2];
		projMap(x, y, X, Y, ang);
		dcx p = farFieldD(P, Ppol, ang[0], ang[1]);
		dcx s = farFieldD(P, Spol, ang[0], ang[1]);
		double c = cos(ang[1]*M_PI/180), sn = sin(ang[1]*M_PI/180);
		double vx = creal(conj(p*c-s*sn)*(p*c-s*sn)), vy = creal(conj(s*c+p*sn)*(s*c+p*sn));
		double vl = sq(creal(p)+cimag(s))+sq(cimag(p)-creal(s)), vr = sq(creal(p)-cimag(s))+sq(cimag(p)+creal(s));
		double vp = creal(conj(p)*p), vs = creal(conj(s)*s);
		if (logScale) {
			Xp[0][x][y] = log10(vx), Yp[0][x][y] = log10(vy), Lp[0][x][y] = log10(vl), Rp[0][x][y] = log10(vr);
			Pp[0][x][y] = log10(vp), Sp[0][x][y] = log10(vs), Up[0][x][y] = log10(vp+vs);
		}
		else {
			Xp[0][x][y] = vx, Yp[0][x][y] = vy, Lp[0][x][y] = vl, Rp[0][x][y] = vr;
			Pp[0][x][y] = vp, Sp[0][x][y] = vs, Up[0][x][y] = Pp[0][x][y]+Sp[0][x][y];
		}
		max = MAX(max, Up[0][x][y]);
	}
	if (Max <= 0) Max = max;
	const char *suffix[7] = { "Unpol", "P-pol", "S-pol", "CL-pol", "CR-pol", "X-pol", "Y-pol" };
	const int order[7] = { 5, 6, 3, 4, 1, 2, 0 };
	for (int a = 0; a < 7; a++) {
		int m = order[a];
		S->F = maps[m];
		snprintf(filename, sizeof filename, "%s%s", fullname, suffix[m]);
		M(W, S, Max, fieldname, filename, cmap);
		if (m) removeField(maps[m], 0, 0, 0);
	}
	deleteSlice(S);
}

void farFieldTheta(world W, phaser P, float phi, char *format, ...)
{
	char filename[1024];
	va_list ap;
	va_start(ap, format);
	vsnprintf(filename, sizeof filename, format, ap);
	va_end(ap);
	farFieldReady(P);
	for (float theta = 0; theta < 360; theta += 1) {
		dcx p = farFieldD(P, Ppol, theta, phi), s = farFieldD(P, Spol, theta, phi);
		(writeRow)(W, filename, (double)theta, sq(cabs(p))+sq(cabs(s)), sq(cabs(p)), sq(cabs(s)), INT2NAN(1));
	}
}

void farFieldPhi(world W, phaser P, float theta, char *format, ...)
{
	char filename[1024];
	va_list ap;
	va_start(ap, format);
	vsnprintf(filename, sizeof filename, format, ap);
	va_end(ap);
	farFieldReady(P);
	for (float phi = 0; phi < 360; phi += 1) {
		dcx p = farFieldD(P, Ppol, theta, phi), s = farFieldD(P, Spol, theta, phi);
		(writeRow)(W, filename, (double)phi, sq(cabs(p))+sq(cabs(s)), sq(cabs(p)), sq(cabs(s)), INT2NAN(1));
	}
}

/* Power radiated into the cone of half angle aMax (or the ring aMin..aMax) around
 * (theta, phi): rings of width delta, 6n azimuthal samples on ring n (capped at 60
 * degrees), polarisation-resolved weights (src/farfld.c:627-662). */
float (farFieldFlux)(phaser P, tfield pol(phaser,float,float), float theta, float phi, float aMin, float aMax, float delta, ...)
{
	if (!delta) delta = 1;
	if (!aMax) aMax = aMin, aMin = 0;
	farFieldReady(P);
	double Sum = 0;
	int nMin = floor(aMin/delta+0.5);
	int nMax = ceil(aMax/delta-0.5);
	if (!nMin) {
		Sum = 2 * M_PI * (cos(aMin*M_PI/180) - cos(MIN(aMax,0.5*delta)*M_PI/180)) * farFieldID(P, pol, theta, phi);
		aMin = 0.5 * delta;
		nMin++;
	}
	double ringSum = 0;
	if (nMax) {
		#pragma omp parallel for schedule(dynamic) reduction(+:ringSum)
		for (int n = nMin; n <= nMax; n++) {
			int N = n < floor(60/delta+0.5) ? n*6 : floor(60/delta+0.5)*6;
			double dth = cos(MAX(aMin,delta*(n-0.5))*M_PI/180) - cos(MIN(aMax,delta*(n+0.5))*M_PI/180);
			double ring = 0;
			for (int m = 0; m < N; m++) {
				float a = rotateA(theta, delta*n, 360.0*m/N);
				float b = rotateB(theta, delta*n, 360.0*m/N)+phi;
				float c = rotateB(-a, theta, phi-b)+180-360.0*m/N;
				dcx p = farFieldD(P, Ppol, a, b), s = farFieldD(P, Spol, a, b);
				double pp = sq(cabs(p * cos(c*M_PI/180) - s * sin(c*M_PI/180)));
				double ss = sq(cabs(s * cos(c*M_PI/180) + p * sin(c*M_PI/180)));
				if (pol == Ppol) ring += dth * (pp * (M_PI/N + sin(M_PI/N)) + ss * (M_PI/N - sin(M_PI/N)));
				if (pol == Spol) ring += dth * (pp * (M_PI/N - sin(M_PI/N)) + ss * (M_PI/N + sin(M_PI/N)));
				if (pol == Unpol) ring += 2 * M_PI/N * dth * (pp + ss);
			}
			ringSum += ring;
		}
	}
	return Sum + ringSum;
}
