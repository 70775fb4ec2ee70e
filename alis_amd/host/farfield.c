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


/*
 * Far field at many directions (src/farfld.c:462-514 for each).  The directions become
 * columns -- a polar angle within a degree of 90 or 270 is the mean of the two angles one
 * degree either side -- and the plane-wave profiles T are computed once per distinct polar
 * angle (the azimuth only rotates their components).  The engine evaluates the surface sums
 * of every column (alis_dev_farfield: the face integrals of src/farfld.c:248-459 without the
 * factors that are constant over a face); the far field of each column is assembled here:
 * face signs, the in-plane phase of the x and y faces, the T values at the z faces, the
 * periodic array factor and the normalisation.
 */
enum { TEX, TEY, TEZ, THX, THY, THZ };

typedef struct {
	int n, cap, nz;
	double *key;        /* polar angle of a pair of standard slots */
	double *A;          /* per slot: T.A */
	dcx *T;             /* per slot: 6 components x nz, k = kMIN .. kMax+1 */
} ffSlots;

static int ffNewSlot(ffSlots *L)
{
	if (L->n == L->cap) {
		L->cap = L->cap ? 2 * L->cap : 64;
		L->key = realloc(L->key, L->cap * sizeof(double));
		L->A = realloc(L->A, L->cap * sizeof(double));
		L->T = realloc(L->T, (size_t)L->cap * 6 * L->nz * sizeof(dcx));
	}
	memset(L->T + (size_t)L->n * 6 * L->nz, 0, 6 * L->nz * sizeof(dcx));
	L->key[L->n] = NAN, L->A[L->n] = 0;
	return L->n++;
}

static void ffStore(phaser P, ffSlots *L, int slot, int comp, const dcx *a)
{
	if (a) memcpy(L->T + ((size_t)slot * 6 + comp) * L->nz, a + P->kMIN, L->nz * sizeof(dcx));
}

/* Far field of P at directions (theta[a], phi[a]) in degrees: for pol Ppol or Spol both
 * polarisations (out0 P, out1 S), for another profile function its own (out0). */
static void farFieldMany(phaser P, tfield pol(phaser,float,float), int n, const double *theta, const double *phi,
                         dcx *out0, dcx *out1)
{
	world W = P->W;
	const int standard = pol == Ppol || pol == Spol;
	for (int a = 0; a < n; a++) {
		out0[a] = 0;
		if (out1) out1[a] = 0;
	}
	phaser B = phaserBase(P);
	if (!B || !n) return;
	if (!P->dE) farfieldCoeffs(P, W);
	alisPhaserDevice(B);

	/* columns */
	typedef struct { double theta, phi, w; int out; } dir;
	dir *D = malloc(2 * n * sizeof *D);
	alis_dev_ff_col *C = calloc(2 * n, sizeof *C);
	int m = 0;
	for (int a = 0; a < n; a++) {
		double th = theta[a], ph = phi[a];
		if (th == INF || ph == INF) continue;
		while (th < 0) th += 360;
		while (th >= 360) th -= 360;
		while (ph < 0) ph += 360;
		while (ph >= 360) ph -= 360;
		double ths[2] = {th, th}, w = 1;
		int nt = 1;
		if (89 < th && th < 91) ths[0] = 89, ths[1] = 91, nt = 2, w = 0.5;
		if (269 < th && th < 271) ths[0] = 269, ths[1] = 271, nt = 2, w = 0.5;
		for (int t = 0; t < nt; t++) {
			double tt = ths[t], sinth;
			if (tt < 90 || 270 < tt) {
				if (cimagf(P->ex[P->kMax])) continue;
				sinth = sin(tt*M_PI/180) * sqrt(crealf(P->ex[P->kMax]));
			}
			else {
				if (cimagf(P->ex[P->kMIN])) continue;
				sinth = sin(tt*M_PI/180) * sqrt(crealf(P->ex[P->kMIN]));
			}
			D[m] = (dir) {tt, ph, w, a};
			C[m].kx = -2 * M_PI * P->f * sinth * cos(ph*M_PI/180);
			C[m].ky = -2 * M_PI * P->f * sinth * sin(ph*M_PI/180);
			m++;
		}
	}

	/* profiles: a P and an S slot per distinct polar angle, or one slot per column */
	ffSlots L = {0, 0, P->kMax + 2 - P->kMIN, 0, 0, 0};
	if (standard) {
		int *slotOf = malloc((m ? m : 1) * sizeof(int));
		/* open-addressing table from polar angle to slot pair */
		int hcap = 16;
		while (hcap < 4 * m) hcap *= 2;
		int *hash = malloc(hcap * sizeof(int));
		for (int h = 0; h < hcap; h++) hash[h] = -1;
		for (int c = 0; c < m; c++) {
			unsigned long long bits;
			memcpy(&bits, &D[c].theta, sizeof bits);
			int h = (int)((bits * 0x9e3779b97f4a7c15ULL) >> 40) & (hcap - 1);
			while (hash[h] >= 0 && L.key[hash[h]] != D[c].theta) h = (h + 1) & (hcap - 1);
			if (hash[h] < 0) {
				ffNewSlot(&L), ffNewSlot(&L);
				hash[h] = L.n - 2;
				L.key[L.n - 2] = D[c].theta;
			}
			slotOf[c] = hash[h];
		}
		free(hash);
		#pragma omp parallel for schedule(dynamic)
		for (int s = 0; s < L.n; s += 2) {
			profile Tp = pPolD(P, L.key[s], 0), Ts = sPolD(P, L.key[s], 0);
			ffStore(P, &L, s, TEX, Tp.Ex), ffStore(P, &L, s, THY, Tp.Hy), ffStore(P, &L, s, TEZ, Tp.Ez);
			ffStore(P, &L, s+1, TEY, Ts.Ey), ffStore(P, &L, s+1, THX, Ts.Hx), ffStore(P, &L, s+1, THZ, Ts.Hz);
			L.A[s] = Tp.A, L.A[s+1] = Ts.A;
			freeProfiles(P, &Tp), freeProfiles(P, &Ts);
		}
		for (int c = 0; c < m; c++) {
			alis_dev_ff_col *q = &C[c];
			const int s = slotOf[c];
			const double sp = sin(D[c].phi*M_PI/180), cp = cos(D[c].phi*M_PI/180);
			const int rot = D[c].phi != 0, hasZ = D[c].theta != 0;
			q->slot[0] = s, q->slot[1] = s + 1;
			for (int t = 0; t < 6; t++) q->src[0][t] = q->src[1][t] = -1;
			/* P: Ex Ey from Ex0, Hx Hy from Hy0, Ez0 (src/farfld.c:169-184) */
			q->src[0][TEX] = TEX, q->fac[0][TEX] = rot ? cp : 1;
			q->src[0][THY] = THY, q->fac[0][THY] = rot ? cp : 1;
			if (rot) q->src[0][TEY] = TEX, q->fac[0][TEY] = sp, q->src[0][THX] = THY, q->fac[0][THX] = -sp;
			if (hasZ) q->src[0][TEZ] = TEZ, q->fac[0][TEZ] = 1;
			/* S: Ex Ey from Ey0, Hx Hy from Hx0, Hz0 (src/farfld.c:229-244) */
			q->src[1][TEY] = TEY, q->fac[1][TEY] = rot ? cp : 1;
			q->src[1][THX] = THX, q->fac[1][THX] = rot ? cp : 1;
			if (rot) q->src[1][TEX] = TEY, q->fac[1][TEX] = -sp, q->src[1][THY] = THX, q->fac[1][THY] = sp;
			if (hasZ) q->src[1][THZ] = THZ, q->fac[1][THZ] = 1;
		}
		free(slotOf);
	}
	else {
		for (int c = 0; c < m; c++) {
			const int s = ffNewSlot(&L);
			profile T = profileFor(P, pol, D[c].theta, D[c].phi);
			dcx *comp[6] = {T.Ex, T.Ey, T.Ez, T.Hx, T.Hy, T.Hz};
			for (int t = 0; t < 6; t++) {
				ffStore(P, &L, s, t, comp[t]);
				C[c].src[0][t] = comp[t] ? t : -1, C[c].fac[0][t] = 1;
				C[c].src[1][t] = -1;
			}
			L.A[s] = T.A;
			C[c].slot[0] = C[c].slot[1] = s;
			freeProfiles(P, &T);
		}
	}

	/* the engine's surface sums */
	phaserGeom g;
	phaserFaces(B, &g);
	const int ni = P->iMax-P->iMin+1, nj = P->jMax-P->jMin+1, nk = P->kMax-P->kMin+1;
	double *tab = malloc((4 * (ni + nj) + 2 * nk) * sizeof(double));
	double *dxE = tab, *dxH = dxE + ni, *xN = dxH + ni, *xH = xN + ni;
	double *dyE = xH + ni, *dyH = dyE + nj, *yN = dyH + nj, *yH = yN + nj;
	double *dzE = yH + nj, *dzH = dzE + nk;
	for (int i = P->iMin; i <= P->iMax; i++) {
		const int a = i - P->iMin;
		dxE[a] = W->CE->dx[i], dxH[a] = W->CH->dx[i], xN[a] = W->itox(W, i), xH[a] = W->itox(W, i-0.5);
	}
	for (int j = P->jMin; j <= P->jMax; j++) {
		const int a = j - P->jMin;
		dyE[a] = W->CE->dy[j], dyH[a] = W->CH->dy[j], yN[a] = W->jtoy(W, j), yH[a] = W->jtoy(W, j-0.5);
	}
	for (int k = P->kMin; k <= P->kMax; k++) dzE[k-P->kMin] = W->CE->dz[k], dzH[k-P->kMin] = W->CH->dz[k];
	const int sideFaces = g.on[0] || g.on[1] || g.on[2] || g.on[3];
	alis_dev_ff_geom G = {dxE, dxH, xN, xH, dyE, dyH, yN, yH, dzE, dzH,
		W->xMinSur==SYM, W->yMinSur==SYM, W->zMinSur==SYM, W->coskx, W->cosky, W->coskz,
		sideFaces ? L.n : 0, P->kMIN, L.nz, sideFaces ? (const double *)L.T : 0};
	double *S = malloc(((size_t)m * 96 + 1) * sizeof(double));
	if (m) alisFatal("far field", alis_dev_farfield(W->dev, B->devId, (int)(P - B), &G, m, C, S));

	/* assembly */
	const int sym = W->zMinSur==SYM;
	for (int c = 0; c < m; c++) {
		const double *s = S + (size_t)c * 96;
		const double kx = C[c].kx, ky = C[c].ky;
		for (int p = 0; p < (standard ? 2 : 1); p++) {
			/* sums of face `kind`; those of z faces do not involve T and are stored once */
#define SV(kind, comp) (s[2*(((kind)*4+(comp))*2+((kind) < 4 ? p : 0))] + I * s[2*(((kind)*4+(comp))*2+((kind) < 4 ? p : 0))+1])
			const alis_dev_ff_col *q = &C[c];
			/* component t of this column's profile at absolute k */
#define TV(t, kk) (q->src[p][t] < 0 ? (dcx)0 : q->fac[p][t] * L.T[((size_t)q->slot[p] * 6 + q->src[p][t]) * L.nz + (kk) - P->kMIN])
			dcx Sum = 0;
			for (int kind = 0; kind < 2; kind++) {
				if ((kind ? W->xMaxSur : W->xMinSur) != PML) continue;
				const double x = W->itox(W, kind ? P->iMax : P->iMin);
				dcx v = (SV(kind,0) - SV(kind,3) - SV(kind,1) + SV(kind,2)) * cexp(I*kx*x);
				if (W->xMinSur==SYM) v += (SV(kind,0) + SV(kind,3) - SV(kind,1) - SV(kind,2)) * cexp(-I*kx*x) * W->coskx;
				Sum += kind ? v : -v;
			}
			for (int kind = 2; kind < 4; kind++) {
				if ((kind == 2 ? W->yMinSur : W->yMaxSur) != PML) continue;
				const double y = W->jtoy(W, kind == 3 ? P->jMax : P->jMin);
				dcx v = (SV(kind,0) - SV(kind,3) - SV(kind,1) + SV(kind,2)) * cexp(I*ky*y);
				if (W->yMinSur==SYM) v += (SV(kind,0) + SV(kind,3) - SV(kind,1) - SV(kind,2)) * cexp(-I*ky*y) * W->cosky;
				Sum += kind == 3 ? v : -v;
			}
			for (int kind = 4; kind < 6; kind++) {
				if ((kind == 4 ? W->zMinSur : W->zMaxSur) != PML || !g.on[kind]) continue;
				const int k = kind == 5 ? P->kMax : P->kMin;
				dcx v = SV(kind,0) * 0.5 * (TV(THY,k) + TV(THY,k+1));
				if (sym) v += SV(kind,0) * 0.5 * (TV(THY,1-k) + TV(THY,-k)) * W->coskz;
				v -= SV(kind,3) * TV(TEX,k);
				if (sym) v += SV(kind,3) * TV(TEX,-k) * W->coskz;
				v -= SV(kind,1) * 0.5 * (TV(THX,k) + TV(THX,k+1));
				if (sym) v -= SV(kind,1) * 0.5 * (TV(THX,1-k) + TV(THX,-k)) * W->coskz;
				v += SV(kind,2) * TV(TEY,k);
				if (sym) v -= SV(kind,2) * TV(TEY,-k) * W->coskz;
				Sum += kind == 5 ? v : -v;
			}
#undef TV
#undef SV
			if (P->xNUM) {
				dcx a = 0, r = 1, e = P->expikx * cexp(I*kx*(W->xMax-W->xMin));
				for (int x = 0; x < P->xNUM; x++) a += (r *= e);
				Sum *= a / P->xNUM;
			}
			if (P->yNUM) {
				dcx a = 0, r = 1, e = P->expiky * cexp(I*ky*(W->yMax-W->yMin));
				for (int y = 0; y < P->yNUM; y++) a += (r *= e);
				Sum *= a / P->yNUM;
			}
			dcx v = Sum * L.A[q->slot[p]] * P->A * P->f;
			if (p) out1[D[c].out] += D[c].w * v;
			else out0[D[c].out] += D[c].w * v;
		}
	}
	free(S), free(tab), free(D), free(C);
	free(L.key), free(L.A), free(L.T);
}

/* Complex far-field amplitude of phaser P for polarisation pol at (theta, phi) in degrees,
 * including the periodic array factor (src/farfld.c:462-514). */
dcx farFieldD(phaser P, tfield pol(phaser,float,float), double theta, double phi)
{
	dcx p, s;
	farFieldMany(P, pol, 1, &theta, &phi, &p, &s);
	return pol == Spol ? s : p;
}

float complex farField(phaser P, tfield pol(phaser,float,float), float theta, float phi)
{
	return farFieldD(P, pol, theta, phi);
}

double farFieldID(phaser P, tfield pol(phaser,float,float), double theta, double phi)
{
	if (pol == Unpol) {
		dcx p, s;
		farFieldMany(P, Ppol, 1, &theta, &phi, &p, &s);
		return creal(conj(p) * p) + creal(conj(s) * s);
	}
	dcx f = farFieldD(P, pol, theta, phi);
	return creal(conj(f) * f);
}

float farFieldI(phaser P, tfield pol(phaser,float,float), float theta, float phi)
{
	return farFieldID(P, pol, theta, phi);
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
	const long XY = (long)X * Y;
	double *th = malloc(2 * XY * sizeof(double)), *ph = th + XY;
	dcx *pv = malloc(2 * XY * sizeof(dcx)), *sv = pv + XY;
	for (long xy = 0; xy < XY; xy++) {
		float ang[2];
		projMap(xy / Y, xy % Y, X, Y, ang);
		th[xy] = ang[0], ph[xy] = ang[1];
	}
	farFieldMany(P, Ppol, XY, th, ph, pv, sv);
	int logScale = fieldname[0] != 0;
	float max = 0;
	#pragma omp parallel for schedule(static) reduction(max:max)
	for (long xy = 0; xy < XY; xy++) {
		int x = xy / Y, y = xy % Y;
		float ang[2] = {th[xy], ph[xy]};
		dcx p = pv[xy], s = sv[xy];
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
	free(th), free(pv);
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
	double th[360], ph[360];
	dcx p[360], s[360];
	int n = 0;
	for (float theta = 0; theta < 360; theta += 1) th[n] = theta, ph[n] = phi, n++;
	farFieldMany(P, Ppol, n, th, ph, p, s);
	for (int a = 0; a < n; a++)
		(writeRow)(W, filename, th[a], sq(cabs(p[a]))+sq(cabs(s[a])), sq(cabs(p[a])), sq(cabs(s[a])), INT2NAN(1));
}

void farFieldPhi(world W, phaser P, float theta, char *format, ...)
{
	char filename[1024];
	va_list ap;
	va_start(ap, format);
	vsnprintf(filename, sizeof filename, format, ap);
	va_end(ap);
	double th[360], ph[360];
	dcx p[360], s[360];
	int n = 0;
	for (float phi = 0; phi < 360; phi += 1) th[n] = theta, ph[n] = phi, n++;
	farFieldMany(P, Ppol, n, th, ph, p, s);
	for (int a = 0; a < n; a++)
		(writeRow)(W, filename, ph[a], sq(cabs(p[a]))+sq(cabs(s[a])), sq(cabs(p[a])), sq(cabs(s[a])), INT2NAN(1));
}

/* Power radiated into the cone of half angle aMax (or the ring aMin..aMax) around
 * (theta, phi): rings of width delta, 6n azimuthal samples on ring n (capped at 60
 * degrees), polarisation-resolved weights (src/farfld.c:627-662). */
float (farFieldFlux)(phaser P, tfield pol(phaser,float,float), float theta, float phi, float aMin, float aMax, float delta, ...)
{
	if (!delta) delta = 1;
	if (!aMax) aMax = aMin, aMin = 0;
	double Sum = 0;
	int nMin = floor(aMin/delta+0.5);
	int nMax = ceil(aMax/delta-0.5);
	if (!nMin) {
		Sum = 2 * M_PI * (cos(aMin*M_PI/180) - cos(MIN(aMax,0.5*delta)*M_PI/180)) * farFieldID(P, pol, theta, phi);
		aMin = 0.5 * delta;
		nMin++;
	}
	double ringSum = 0;
	if (nMax && nMax >= nMin) {
		/* every sample of every ring in one evaluation */
		long total = 0;
		for (int n = nMin; n <= nMax; n++) total += n < floor(60/delta+0.5) ? n*6 : floor(60/delta+0.5)*6;
		double *ab = malloc(3 * total * sizeof(double) + 1);
		dcx *pv = malloc(2 * total * sizeof(dcx) + 1), *sv = pv + total;
		long at = 0;
		for (int n = nMin; n <= nMax; n++) {
			int N = n < floor(60/delta+0.5) ? n*6 : floor(60/delta+0.5)*6;
			for (int m = 0; m < N; m++, at++) {
				float a = rotateA(theta, delta*n, 360.0*m/N);
				float b = rotateB(theta, delta*n, 360.0*m/N)+phi;
				ab[at] = a, ab[total+at] = b;
				ab[2*total+at] = rotateB(-a, theta, phi-b)+180-360.0*m/N;
			}
		}
		farFieldMany(P, Ppol, total, ab, ab + total, pv, sv);
		at = 0;
		for (int n = nMin; n <= nMax; n++) {
			int N = n < floor(60/delta+0.5) ? n*6 : floor(60/delta+0.5)*6;
			double dth = cos(MAX(aMin,delta*(n-0.5))*M_PI/180) - cos(MIN(aMax,delta*(n+0.5))*M_PI/180);
			double ring = 0;
			for (int m = 0; m < N; m++, at++) {
				float c = ab[2*total+at];
				dcx p = pv[at], s = sv[at];
				double pp = sq(cabs(p * cos(c*M_PI/180) - s * sin(c*M_PI/180)));
				double ss = sq(cabs(s * cos(c*M_PI/180) + p * sin(c*M_PI/180)));
				if (pol == Ppol) ring += dth * (pp * (M_PI/N + sin(M_PI/N)) + ss * (M_PI/N - sin(M_PI/N)));
				if (pol == Spol) ring += dth * (pp * (M_PI/N - sin(M_PI/N)) + ss * (M_PI/N + sin(M_PI/N)));
				if (pol == Unpol) ring += 2 * M_PI/N * dth * (pp + ss);
			}
			ringSum += ring;
		}
		free(ab), free(pv);
	}
	return Sum + ringSum;
}
