/*
 * Field accessors float F(world, i, j, k) on the host arrays (src/fields.c:10-791): node
 * averages of the staggered components, raw components, current densities, energy,
 * Poynting, force and stress densities, source and scattered fields.  A negative index on
 * a symmetry plane is mirrored with the parity of the component (cos = +-1 set by the
 * sources).  The arithmetic matches the reference's float expressions.
 */
#include "internal.h"

/* Identity of the basic E/H accessors: E or H, real or imaginary part, axis. */
int fieldComponent(field *F, int *isH, int *m, int *axis)
{
	field *tab[12] = { Ex, Ey, Ez, Hx, Hy, Hz, iEx, iEy, iEz, iHx, iHy, iHz };
	for (int a = 0; a < 12; a++)
		if (F == tab[a]) {
			*isH = (a % 6) >= 3, *m = a >= 6, *axis = a % 3;
			return 1;
		}
	return 0;
}

/* The same for the single-sample accessors rawEx .. rawHz (real part). */
int rawComponent(field *F, int *isH, int *m, int *axis)
{
	field *tab[6] = { rawEx, rawEy, rawEz, rawHx, rawHy, rawHz };
	for (int a = 0; a < 6; a++)
		if (F == tab[a]) {
			*isH = a >= 3, *m = 0, *axis = a % 3;
			return 1;
		}
	return 0;
}

/* Pole susceptibility and its frequency derivative, real parts (src/fields.c:17-35). */
static float poleFF(matter M, float w, int p)
{
	if (M.P[p].f || M.P[p].g)
		return 0.5
			*((M.P[p].f * M.P[p].w - M.P[p].g * M.P[p].r + I * M.P[p].g * M.P[p].w) / (M.P[p].w + w + I * M.P[p].r)
			+ (M.P[p].f * M.P[p].w - M.P[p].g * M.P[p].r - I * M.P[p].g * M.P[p].w) / (M.P[p].w - w - I * M.P[p].r));
	if (M.P[p].w) return - sq(M.P[p].w) / (sq(w) + I * w * M.P[p].r);
	return 0;
}

static float dPoleFF(matter M, float w, int p)
{
	if (M.P[p].f || M.P[p].g)
		return 0.5
			*((M.P[p].f * M.P[p].w - M.P[p].g * M.P[p].r + I * M.P[p].g * M.P[p].w) / sq(M.P[p].w - w - I * M.P[p].r)
			- (M.P[p].f * M.P[p].w - M.P[p].g * M.P[p].r - I * M.P[p].g * M.P[p].w) / sq(M.P[p].w + w + I * M.P[p].r));
	if (M.P[p].w) return (2 * w + I * M.P[p].r) * sq(M.P[p].w) / sq(sq(w) + I * w * M.P[p].r);
	return 0;
}

/* Effective (energy) permittivity of component q at a staggered E position: inverse of the
 * stored 1/eps, minus the conductivity term, plus Re[d(w chi)/dw] of every pole at the
 * world frequency W->f (src/fields.c:40-93). */
static float effPermittivity(world W, int q, int i, int j, int k)
{
	float ***C = q == 0 ? W->CE->x : q == 1 ? W->CE->y : W->CE->z;
	if (!C) return 1;
	if (i<0 && W->xMinSur==SYM) i = q == 0 ? 1-i : -i;
	if (j<0 && W->yMinSur==SYM) j = q == 1 ? 1-j : -j;
	if (k<0 && W->zMinSur==SYM) k = q == 2 ? 1-k : -k;
	float f = 1 / C[i][j][k];
	for (int n = 0, o = 0; n < W->CE->N; n++, o++) {
		while (!dispersiveMatter(W->CE->M[o])) o++;
		if (W->CE->iMin[n]<=i && i<=W->CE->iMax[n] && W->CE->jMin[n]<=j && j<=W->CE->jMax[n] && W->CE->kMin[n]<=k && k<=W->CE->kMax[n]) {
			matter M = W->CE->M[o];
			int m = M.e.y || M.e.z ? 1 : 0;
			float ***fill = q == 0 ? W->CE->Jx[n] : q == 1 ? W->CE->Jy[n] : W->CE->Jz[n];
			float a = q == 0 ? W->CE->ax[n] : q == 1 ? W->CE->ay[n] : W->CE->az[n];
			float e = q == 0 ? M.e.x : q == 1 ? (M.e.y ? M.e.y : M.e.x) : (M.e.z ? M.e.z : M.e.x);
			f -= a * e * fill[i][j][k];
			for (int p = 0; p < MAXPOLES; p += 1+2*m)
				f += (poleFF(M, W->f*W->eV, p+q*m) + W->f*W->eV * dPoleFF(M, W->f*W->eV, p+q*m)) * fill[i][j][k];
		}
	}
	return f;
}

float raweEffx(world W, int i, int j, int k) { return effPermittivity(W, 0, i, j, k); }
float raweEffy(world W, int i, int j, int k) { return effPermittivity(W, 1, i, j, k); }
float raweEffz(world W, int i, int j, int k) { return effPermittivity(W, 2, i, j, k); }

#define MIRROR() do { \
	if (i<0 && W->xMinSur==SYM) i=-i; \
	if (j<0 && W->yMinSur==SYM) j=-j; \
	if (k<0 && W->zMinSur==SYM) k=-k; } while (0)

float eEff(world W, int i, int j, int k)
{
	MIRROR();
	return (raweEffx(W,i,j,k) + raweEffx(W,i+1,j,k) + raweEffy(W,i,j,k) + raweEffy(W,i,j+1,k) + raweEffz(W,i,j,k) + raweEffz(W,i,j,k+1)) / 6;
}

float RI(world W, int i, int j, int k) { MIRROR(); return sqrtf(fabsf(eEff(W,i,j,k))); }

float ContourRI(world W, int i, int j, int k)
{
	i = W->xtoi(W, i*W->dx), j = W->ytoj(W, j*W->dy), k = W->ztok(W, k*W->dz);
	MIRROR();
	float x = log10f(1+sqrtf(fabsf(raweEffx(W, i+1, j, k)))) - log10f(1+sqrtf(fabsf(raweEffx(W, i, j, k))));
	float y = log10f(1+sqrtf(fabsf(raweEffy(W, i, j+1, k)))) - log10f(1+sqrtf(fabsf(raweEffy(W, i, j, k))));
	float z = log10f(1+sqrtf(fabsf(raweEffz(W, i, j, k+1)))) - log10f(1+sqrtf(fabsf(raweEffz(W, i, j, k))));
	return sqrt(x*x+y*y+z*z);
}

float LogRI(world W, int i, int j, int k) { return log10f(RI(W, i, j, k)); }
float Contour(world W, int i, int j, int k) { return ContourRI(W, i, j, k) ? 1 : 0; }


/* Mirror (i,j,k) across symmetry planes with parity signs sx, sy, sz (+1 or -1) times the
 * plane's cos; p starts at the accessor's weight. */
#define PARITY(p0, sx, sy, sz) float p = (p0); \
	if (i<0 && W->xMinSur==SYM) i=-i, p*=(sx)*W->coskx; \
	if (j<0 && W->yMinSur==SYM) j=-j, p*=(sy)*W->cosky; \
	if (k<0 && W->zMinSur==SYM) k=-k, p*=(sz)*W->coskz

float rawEx(world W, int i, int j, int k) { PARITY(1, 1,-1,-1); return p * W->E->x[i][j][k]; }
float rawEy(world W, int i, int j, int k) { PARITY(1,-1, 1,-1); return p * W->E->y[i][j][k]; }
float rawEz(world W, int i, int j, int k) { PARITY(1,-1,-1, 1); return p * W->E->z[i][j][k]; }
float rawHx(world W, int i, int j, int k) { PARITY(1,-1, 1, 1); return p * W->H->x[i][j][k]; }
float rawHy(world W, int i, int j, int k) { PARITY(1, 1,-1, 1); return p * W->H->y[i][j][k]; }
float rawHz(world W, int i, int j, int k) { PARITY(1, 1, 1,-1); return p * W->H->z[i][j][k]; }

float Ex(world W, int i, int j, int k) { PARITY(0.5, 1,-1,-1); return p * (W->E->x[i][j][k] + W->E->x[i+1][j][k]); }
float Ey(world W, int i, int j, int k) { PARITY(0.5,-1, 1,-1); return p * (W->E->y[i][j][k] + W->E->y[i][j+1][k]); }
float Ez(world W, int i, int j, int k) { PARITY(0.5,-1,-1, 1); return p * (W->E->z[i][j][k] + W->E->z[i][j][k+1]); }

float iEx(world W, int i, int j, int k)
{
	if (!W->complexField) return 0;
	PARITY(0.5, 1,-1,-1); return p * (W->E[1].x[i][j][k] + W->E[1].x[i+1][j][k]);
}
float iEy(world W, int i, int j, int k)
{
	if (!W->complexField) return 0;
	PARITY(0.5,-1, 1,-1); return p * (W->E[1].y[i][j][k] + W->E[1].y[i][j+1][k]);
}
float iEz(world W, int i, int j, int k)
{
	if (!W->complexField) return 0;
	PARITY(0.5,-1,-1, 1); return p * (W->E[1].z[i][j][k] + W->E[1].z[i][j][k+1]);
}

float Hx(world W, int i, int j, int k)
{
	PARITY(0.25,-1, 1, 1);
	return p * (W->H->x[i][j][k] + W->H->x[i][j+1][k] + W->H->x[i][j][k+1] + W->H->x[i][j+1][k+1]);
}
float Hy(world W, int i, int j, int k)
{
	PARITY(0.25, 1,-1, 1);
	return p * (W->H->y[i][j][k] + W->H->y[i][j][k+1] + W->H->y[i+1][j][k] + W->H->y[i+1][j][k+1]);
}
float Hz(world W, int i, int j, int k)
{
	PARITY(0.25, 1, 1,-1);
	return p * (W->H->z[i][j][k] + W->H->z[i+1][j][k] + W->H->z[i][j+1][k] + W->H->z[i+1][j+1][k]);
}
float iHx(world W, int i, int j, int k)
{
	if (!W->complexField) return 0;
	PARITY(0.25,-1, 1, 1);
	return p * (W->H[1].x[i][j][k] + W->H[1].x[i][j+1][k] + W->H[1].x[i][j][k+1] + W->H[1].x[i][j+1][k+1]);
}
float iHy(world W, int i, int j, int k)
{
	if (!W->complexField) return 0;
	PARITY(0.25, 1,-1, 1);
	return p * (W->H[1].y[i][j][k] + W->H[1].y[i][j][k+1] + W->H[1].y[i+1][j][k] + W->H[1].y[i+1][j][k+1]);
}
float iHz(world W, int i, int j, int k)
{
	if (!W->complexField) return 0;
	PARITY(0.25, 1, 1,-1);
	return p * (W->H[1].z[i][j][k] + W->H[1].z[i+1][j][k] + W->H[1].z[i][j+1][k] + W->H[1].z[i+1][j+1][k]);
}


/* Total current density at a staggered E position: curl H minus the vacuum displacement
 * term, the dispersive currents and minus the injected source term (src/fields.c:294-384). */
static float rawCurrent(world W, int q, int i, int j, int k)
{
	float f = 0, p = 1;
	if (i<0 && W->xMinSur==SYM) i = q == 0 ? 1-i : -i, p *= (q == 0 ? 1 : -1) * W->coskx;
	if (j<0 && W->yMinSur==SYM) j = q == 1 ? 1-j : -j, p *= (q == 1 ? 1 : -1) * W->cosky;
	if (k<0 && W->zMinSur==SYM) k = q == 2 ? 1-k : -k, p *= (q == 2 ? 1 : -1) * W->coskz;
	coeffs *C = W->CE;
	float ***Cq = q == 0 ? C->x : q == 1 ? C->y : C->z;
	float ***Eq = q == 0 ? W->E->x : q == 1 ? W->E->y : W->E->z;
	if (Cq) {
		float curl;
		if (q == 0) curl = C->dtdz[k] * (W->H->y[i][j][k] - W->H->y[i][j][k+1]) + C->dtdy[j] * (W->H->z[i][j+1][k] - W->H->z[i][j][k]);
		else if (q == 1) curl = C->dtdx[i] * (W->H->z[i][j][k] - W->H->z[i+1][j][k]) + C->dtdz[k] * (W->H->x[i][j][k+1] - W->H->x[i][j][k]);
		else curl = C->dtdy[j] * (W->H->x[i][j][k] - W->H->x[i][j+1][k]) + C->dtdx[i] * (W->H->y[i+1][j][k] - W->H->y[i][j][k]);
		f += curl * (1 - Cq[i][j][k]);
		for (int n = 0; n < C->N; n++) {
			int in = C->iMin[n]<=i && (q == 0 ? i<=C->iMax[n] : i<C->iMax[n]) &&
			         C->jMin[n]<=j && (q == 1 ? j<=C->jMax[n] : j<C->jMax[n]) &&
			         C->kMin[n]<=k && (q == 2 ? k<=C->kMax[n] : k<C->kMax[n]);
			if (!in) continue;
			float a = q == 0 ? C->ax[n] : q == 1 ? C->ay[n] : C->az[n];
			float ***fill = q == 0 ? C->Jx[n] : q == 1 ? C->Jy[n] : C->Jz[n];
			float ****J = q == 0 ? W->E->Jx[n] : q == 1 ? W->E->Jy[n] : W->E->Jz[n];
			if (a) f += 2 * fill[i][j][k] * Eq[i][j][k] / (1 + 1 / a);
			for (int pp = 0; pp < C->NK[n]; pp++) if (J[pp]) f += J[pp][i][j][k];
		}
	}
	for (source S = W->SE; S; S = S->S)
		if (S->F == Eq && W->n < S->N && S->iMin <= i && i <= S->iMax && S->jMin <= j && j <= S->jMax && S->kMin <= k && k <= S->kMax)
			f -= (S->Amplitude ? S->Amplitude[i][j][k] : S->amplitude) * S->waveForm(W, S, W->H->t, S->Phase ? S->Phase[i][j][k] : S->phase);
	return p * f / W->dt;
}

float rawJx(world W, int i, int j, int k) { return rawCurrent(W, 0, i, j, k); }
float rawJy(world W, int i, int j, int k) { return rawCurrent(W, 1, i, j, k); }
float rawJz(world W, int i, int j, int k) { return rawCurrent(W, 2, i, j, k); }
float Jx(world W, int i, int j, int k) { return 0.5 * (rawJx(W,i,j,k) + rawJx(W,i+1,j,k)); }
float Jy(world W, int i, int j, int k) { return 0.5 * (rawJy(W,i,j,k) + rawJy(W,i,j+1,k)); }
float Jz(world W, int i, int j, int k) { return 0.5 * (rawJz(W,i,j,k) + rawJz(W,i,j,k+1)); }


/* Products of one E and one H component at the node (src/fields.c:397-454). */
#define SIGNS(sx, sy, sz) float p = 0.25; \
	if (i<0 && W->xMinSur==SYM) i=-i, p*=(sx); \
	if (j<0 && W->yMinSur==SYM) j=-j, p*=(sy); \
	if (k<0 && W->zMinSur==SYM) k=-k, p*=(sz)

float ExHy(world W, int i, int j, int k)
{
	SIGNS(1, 1,-1);
	return p * (W->E->x[i][j][k] * (W->H->y[i][j][k] + W->H->y[i][j][k+1]) + W->E->x[i+1][j][k] * (W->H->y[i+1][j][k] + W->H->y[i+1][j][k+1]));
}
float ExHz(world W, int i, int j, int k)
{
	SIGNS(1,-1, 1);
	return p * (W->E->x[i][j][k] * (W->H->z[i][j][k] + W->H->z[i][j+1][k]) + W->E->x[i+1][j][k] * (W->H->z[i+1][j][k] + W->H->z[i+1][j+1][k]));
}
float EyHz(world W, int i, int j, int k)
{
	SIGNS(-1, 1, 1);
	return p * (W->E->y[i][j][k] * (W->H->z[i][j][k] + W->H->z[i+1][j][k]) + W->E->y[i][j+1][k] * (W->H->z[i][j+1][k] + W->H->z[i+1][j+1][k]));
}
float EyHx(world W, int i, int j, int k)
{
	SIGNS(1, 1,-1);
	return p * (W->E->y[i][j][k] * (W->H->x[i][j][k] + W->H->x[i][j][k+1]) + W->E->y[i][j+1][k] * (W->H->x[i][j+1][k] + W->H->x[i][j+1][k+1]));
}
float EzHx(world W, int i, int j, int k)
{
	SIGNS(1,-1, 1);
	return p * (W->E->z[i][j][k] * (W->H->x[i][j][k] + W->H->x[i][j+1][k]) + W->E->z[i][j][k+1] * (W->H->x[i][j][k+1] + W->H->x[i][j+1][k+1]));
}
float EzHy(world W, int i, int j, int k)
{
	SIGNS(-1, 1, 1);
	return p * (W->E->z[i][j][k] * (W->H->y[i][j][k] + W->H->y[i+1][j][k]) + W->E->z[i][j][k+1] * (W->H->y[i][j][k+1] + W->H->y[i+1][j][k+1]));
}

float JE(world W, int i, int j, int k)
{
	float f = 0;
	MIRROR();
	f += W->E->x[i][j][k] * rawJx(W,i,j,k) + W->E->x[i+1][j][k] * rawJx(W,i+1,j,k);
	f += W->E->y[i][j][k] * rawJy(W,i,j,k) + W->E->y[i][j+1][k] * rawJy(W,i,j+1,k);
	f += W->E->z[i][j][k] * rawJz(W,i,j,k) + W->E->z[i][j][k+1] * rawJz(W,i,j,k+1);
	return 0.5 * f;
}

float EE(world W, int i, int j, int k)
{
	float f = 0;
	MIRROR();
	f += sq(W->E->x[i][j][k]) + sq(W->E->x[i+1][j][k]);
	f += sq(W->E->y[i][j][k]) + sq(W->E->y[i][j+1][k]);
	f += sq(W->E->z[i][j][k]) + sq(W->E->z[i][j][k+1]);
	return 0.5 * f;
}

float UE(world W, int i, int j, int k)
{
	float f = 0;
	MIRROR();
	f += sq(W->E->x[i][j][k]) * raweEffx(W, i, j, k) + sq(W->E->x[i+1][j][k]) * raweEffx(W, i+1, j, k);
	f += sq(W->E->y[i][j][k]) * raweEffy(W, i, j, k) + sq(W->E->y[i][j+1][k]) * raweEffy(W, i, j+1, k);
	f += sq(W->E->z[i][j][k]) * raweEffz(W, i, j, k) + sq(W->E->z[i][j][k+1]) * raweEffz(W, i, j, k+1);
	return 0.25 * f;
}

/* Sums of squares of the four H samples around the node. */
static float hxSq(world W, int i, int j, int k) { return sq(W->H->x[i][j][k]) + sq(W->H->x[i][j+1][k]) + sq(W->H->x[i][j][k+1]) + sq(W->H->x[i][j+1][k+1]); }
static float hySq(world W, int i, int j, int k) { return sq(W->H->y[i][j][k]) + sq(W->H->y[i][j][k+1]) + sq(W->H->y[i+1][j][k]) + sq(W->H->y[i+1][j][k+1]); }
static float hzSq(world W, int i, int j, int k) { return sq(W->H->z[i][j][k]) + sq(W->H->z[i+1][j][k]) + sq(W->H->z[i][j+1][k]) + sq(W->H->z[i+1][j+1][k]); }
static float exSq(world W, int i, int j, int k) { return sq(W->E->x[i][j][k]) + sq(W->E->x[i+1][j][k]); }
static float eySq(world W, int i, int j, int k) { return sq(W->E->y[i][j][k]) + sq(W->E->y[i][j+1][k]); }
static float ezSq(world W, int i, int j, int k) { return sq(W->E->z[i][j][k]) + sq(W->E->z[i][j][k+1]); }

float HH(world W, int i, int j, int k)
{
	float f = 0;
	MIRROR();
	f += hxSq(W, i, j, k);
	f += hySq(W, i, j, k);
	f += hzSq(W, i, j, k);
	return 0.25 * f;
}

float Sx(world W, int i, int j, int k) { return EyHz(W,i,j,k) - EzHy(W,i,j,k); }
float Sy(world W, int i, int j, int k) { return EzHx(W,i,j,k) - ExHz(W,i,j,k); }
float Sz(world W, int i, int j, int k) { return ExHy(W,i,j,k) - EyHx(W,i,j,k); }
float UH(world W, int i, int j, int k) { return 0.5*HH(W, i, j, k); }
float U(world W, int i, int j, int k) { return UE(W, i, j, k)+UH(W, i, j, k); }
float LogJE(world W, int i, int j, int k) { return log10f(JE(W, i, j, k)); }
float LogEE(world W, int i, int j, int k) { return log10f(EE(W, i, j, k)); }
float LogUE(world W, int i, int j, int k) { return log10f(UE(W, i, j, k)); }
float LogHH(world W, int i, int j, int k) { return log10f(HH(W, i, j, k)); }
float LogUH(world W, int i, int j, int k) { return log10f(HH(W, i, j, k)); }
float LogU(world W, int i, int j, int k) { return log10f(U(W, i, j, k)); }


/* Products of a current and an H component (force density terms, src/fields.c:476-532). */
float JxHy(world W, int i, int j, int k)
{
	SIGNS(1, 1,-1);
	return p * (rawJx(W,i,j,k) * (W->H->y[i][j][k] + W->H->y[i][j][k+1]) + rawJx(W,i+1,j,k) * (W->H->y[i+1][j][k] + W->H->y[i+1][j][k+1]));
}
float JxHz(world W, int i, int j, int k)
{
	SIGNS(1,-1, 1);
	return p * (rawJx(W,i,j,k) * (W->H->z[i][j][k] + W->H->z[i][j+1][k]) + rawJx(W,i+1,j,k) * (W->H->z[i+1][j][k] + W->H->z[i+1][j+1][k]));
}
float JyHz(world W, int i, int j, int k)
{
	SIGNS(-1, 1, 1);
	return p * (rawJy(W,i,j,k) * (W->H->z[i][j][k] + W->H->z[i+1][j][k]) + rawJy(W,i,j+1,k) * (W->H->z[i][j+1][k] + W->H->z[i+1][j+1][k]));
}
float JyHx(world W, int i, int j, int k)
{
	SIGNS(1, 1,-1);
	return p * (rawJy(W,i,j,k) * (W->H->x[i][j][k] + W->H->x[i][j][k+1]) + rawJy(W,i,j+1,k) * (W->H->x[i][j+1][k] + W->H->x[i][j+1][k+1]));
}
float JzHx(world W, int i, int j, int k)
{
	SIGNS(1,-1, 1);
	return p * (rawJz(W,i,j,k) * (W->H->x[i][j][k] + W->H->x[i][j+1][k]) + rawJz(W,i,j,k+1) * (W->H->x[i][j][k+1] + W->H->x[i][j+1][k+1]));
}
float JzHy(world W, int i, int j, int k)
{
	SIGNS(-1, 1, 1);
	return p * (rawJz(W,i,j,k) * (W->H->y[i][j][k] + W->H->y[i+1][j][k]) + rawJz(W,i,j,k+1) * (W->H->y[i][j][k+1] + W->H->y[i+1][j][k+1]));
}

float DivE(world W, int i, int j, int k)
{
	PARITY(1,-1,-1,-1);
	return ((W->E->x[i+1][j][k] - W->E->x[i][j][k]) * W->CE->dtdx[i] +
	        (W->E->y[i][j+1][k] - W->E->y[i][j][k]) * W->CE->dtdy[j] +
	        (W->E->z[i][j][k+1] - W->E->z[i][j][k]) * W->CE->dtdz[k]) * p / W->dt;
}

/* Maxwell stress diagonal: (sa*Hx^2 + sb*Hy^2 + sc*Hz^2)/2 + (sa*Ex^2 + ...), quarter-weighted
 * (src/fields.c:547-594). */
static float stressDiag(world W, int i, int j, int k, int a)
{
	float f = 0;
	MIRROR();
	float h[3] = { hxSq(W, i, j, k), hySq(W, i, j, k), hzSq(W, i, j, k) };
	float e[3] = { exSq(W, i, j, k), eySq(W, i, j, k), ezSq(W, i, j, k) };
	for (int c = 0; c < 3; c++) f = c == a ? f + h[c] : f - h[c];
	f *= 0.5;
	for (int c = 0; c < 3; c++) f = c == a ? f + e[c] : f - e[c];
	return 0.25 * f;
}
float Txx(world W, int i, int j, int k) { return stressDiag(W, i, j, k, 0); }
float Tyy(world W, int i, int j, int k) { return stressDiag(W, i, j, k, 1); }
float Tzz(world W, int i, int j, int k) { return stressDiag(W, i, j, k, 2); }

/* Off-diagonal stress: node averages of E_a E_b + H_a H_b (src/fields.c:597-684). */
static float eAvg2(world W, int c, int i, int j, int k)
{
	return c == 0 ? W->E->x[i][j][k] + W->E->x[i+1][j][k] : c == 1 ? W->E->y[i][j][k] + W->E->y[i][j+1][k] : W->E->z[i][j][k] + W->E->z[i][j][k+1];
}
static float hAvg4(world W, int c, int i, int j, int k)
{
	return c == 0 ? W->H->x[i][j][k] + W->H->x[i][j+1][k] + W->H->x[i][j][k+1] + W->H->x[i][j+1][k+1]
	     : c == 1 ? W->H->y[i][j][k] + W->H->y[i][j][k+1] + W->H->y[i+1][j][k] + W->H->y[i+1][j][k+1]
	              : W->H->z[i][j][k] + W->H->z[i+1][j][k] + W->H->z[i][j+1][k] + W->H->z[i+1][j+1][k];
}
/* first/second are in the order the reference multiplies them. */
static float stressOff(world W, int i, int j, int k, int first, int second, float sx, float sy, float sz)
{
	float f = 0;
	SIGNS(sx, sy, sz);
	f += eAvg2(W, first, i, j, k) * eAvg2(W, second, i, j, k);
	f += hAvg4(W, first, i, j, k) * hAvg4(W, second, i, j, k) * 0.25;
	return p * f;
}
float Txy(world W, int i, int j, int k) { return stressOff(W, i, j, k, 0, 1, -1, -1,  1); }
float Tyx(world W, int i, int j, int k) { return stressOff(W, i, j, k, 0, 1, -1, -1,  1); }
float Txz(world W, int i, int j, int k) { return stressOff(W, i, j, k, 2, 0, -1,  1, -1); }
float Tzx(world W, int i, int j, int k) { return stressOff(W, i, j, k, 2, 0, -1,  1, -1); }
float Tyz(world W, int i, int j, int k) { return stressOff(W, i, j, k, 1, 2,  1, -1, -1); }
float Tzy(world W, int i, int j, int k) { return stressOff(W, i, j, k, 1, 2,  1, -1, -1); }

float Fx(world W, int i, int j, int k) { return DivE(W, i, j, k) * Ex(W, i, j, k) + (JyHz(W, i, j, k) - JzHy(W, i, j, k)); }
float Fy(world W, int i, int j, int k) { return DivE(W, i, j, k) * Ey(W, i, j, k) + (JzHx(W, i, j, k) - JxHz(W, i, j, k)); }
float Fz(world W, int i, int j, int k) { return DivE(W, i, j, k) * Ez(W, i, j, k) + (JxHy(W, i, j, k) - JyHx(W, i, j, k)); }


/* Incident (source) field: sum over TF/SF auxiliary worlds (src/fields.c:689-737). */
float SoEx(world W, int i, int j, int k)
{
	float f = 0;
	PARITY(0.5, 1,-1,-1);
	for (tfsf S = W->S; S; S = S->S) f += S->W->E->x[i][j][k] + S->W->E->x[i+1][j][k];
	return p * f;
}
float SoEy(world W, int i, int j, int k)
{
	float f = 0;
	PARITY(0.5,-1, 1,-1);
	for (tfsf S = W->S; S; S = S->S) f += S->W->E->y[i][j][k] + S->W->E->y[i][j+1][k];
	return p * f;
}
float SoEz(world W, int i, int j, int k)
{
	float f = 0;
	PARITY(0.5,-1,-1, 1);
	for (tfsf S = W->S; S; S = S->S) f += S->W->E->z[i][j][k] + S->W->E->z[i][j][k+1];
	return p * f;
}
float SoEE(world W, int i, int j, int k) { return sq(SoEx(W,i,j,k))+sq(SoEy(W,i,j,k))+sq(SoEz(W,i,j,k)); }

/* Scattered field: total field minus the incident field inside each TF box
 * (src/fields.c:740-791). */
float ScEx(world W, int i, int j, int k)
{
	PARITY(0.5, 1,-1,-1);
	float f = W->E->x[i][j][k] + W->E->x[i+1][j][k];
	for (tfsf S = W->S; S; S = S->S)
		if (S->jMin <= j && j <= S->jMax && S->kMin <= k && k <= S->kMax) {
			if (S->iMIN <= i && i <= S->iMAX) f -= S->W->E->x[i][j][k];
			if (S->iMIN <= i+1 && i+1 <= S->iMAX) f -= S->W->E->x[i+1][j][k];
		}
	return p * f;
}
float ScEy(world W, int i, int j, int k)
{
	PARITY(0.5,-1, 1,-1);
	float f = W->E->y[i][j][k] + W->E->y[i][j+1][k];
	for (tfsf S = W->S; S; S = S->S)
		if (S->iMin <= i && i <= S->iMax && S->kMin <= k && k <= S->kMax) {
			if (S->jMIN <= j && j <= S->jMAX) f -= S->W->E->y[i][j][k];
			if (S->jMIN <= j+1 && j+1 <= S->jMAX) f -= S->W->E->y[i][j+1][k];
		}
	return p * f;
}
float ScEz(world W, int i, int j, int k)
{
	PARITY(0.5,-1,-1, 1);
	float f = W->E->z[i][j][k] + W->E->z[i][j][k+1];
	for (tfsf S = W->S; S; S = S->S)
		if (S->iMin <= i && i <= S->iMax && S->jMin <= j && j <= S->jMax) {
			if (S->kMIN <= k && k <= S->kMAX) f -= S->W->E->z[i][j][k];
			if (S->kMIN <= k+1 && k+1 <= S->kMAX) f -= S->W->E->z[i][j][k+1];
		}
	return p * f;
}
float ScEE(world W, int i, int j, int k) { return sq(ScEx(W,i,j,k))+sq(ScEy(W,i,j,k))+sq(ScEz(W,i,j,k)); }
