/*
 * Plain-C descriptions of worlds and phaser arrays for the engine ABI (alis_dev.h).  The
 * tables are only needed during the engine call that receives them; the arrays they point
 * to belong to the world.
 */
#include <string.h>
#include "internal.h"

static float ***slabArray(vfield *V, int f, int b)
{
	switch (f) {
	case 0: return b ? V->xMinPz : V->xMinPy;
	case 1: return b ? V->yMinPx : V->yMinPz;
	case 2: return b ? V->zMinPy : V->zMinPx;
	case 3: return b ? V->xMaxPz : V->xMaxPy;
	case 4: return b ? V->yMaxPx : V->yMaxPz;
	default: return b ? V->zMaxPy : V->zMaxPx;
	}
}

/* Fill *d for W; obj/src/tfsf tables are malloc'd and released by freeWorldDesc. */
void describeWorld(world W, alis_dev_world_desc *d)
{
	memset(d, 0, sizeof *d);
	d->iMIN = W->iMIN, d->iMAX = W->iMAX, d->jMIN = W->jMIN, d->jMAX = W->jMAX, d->kMIN = W->kMIN, d->kMAX = W->kMAX;
	d->iMin = W->iMin, d->iMax = W->iMax, d->jMin = W->jMin, d->jMax = W->jMax, d->kMin = W->kMin, d->kMax = W->kMax;
	d->iNUM = W->iNUM, d->jNUM = W->jNUM, d->kNUM = W->kNUM;
	d->xMinSur = W->xMinSur, d->xMaxSur = W->xMaxSur, d->yMinSur = W->yMinSur;
	d->yMaxSur = W->yMaxSur, d->zMinSur = W->zMinSur, d->zMaxSur = W->zMaxSur;
	d->complexField = W->complexField;
	d->coskx = W->coskx, d->sinkx = W->sinkx, d->cosky = W->cosky, d->sinky = W->sinky, d->coskz = W->coskz, d->sinkz = W->sinkz;
	d->coskyp = W->coskyp, d->coskym = W->coskym, d->sinkyp = W->sinkyp, d->sinkym = W->sinkym;
	for (int h = 0; h < 2; h++) {
		coeffs *C = h ? W->CH : W->CE;
		d->dt[h][0] = C->dtdx + W->iMIN, d->dt[h][1] = C->dtdy + W->jMIN, d->dt[h][2] = C->dtdz + W->kMIN;
		if (C->xP) d->cpml[h][0][0] = C->xP + W->iMIN, d->cpml[h][0][1] = C->xPy + W->iMIN, d->cpml[h][0][2] = C->xPz + W->iMIN;
		if (C->yP) d->cpml[h][1][0] = C->yP + W->jMIN, d->cpml[h][1][1] = C->yPz + W->jMIN, d->cpml[h][1][2] = C->yPx + W->jMIN;
		if (C->zP) d->cpml[h][2][0] = C->zP + W->kMIN, d->cpml[h][2][1] = C->zPx + W->kMIN, d->cpml[h][2][2] = C->zPy + W->kMIN;
		vfield *V = h ? W->H : W->E;
		for (int m = 0; m <= W->complexField; m++) {
			d->field[h][m][0] = V[m].x, d->field[h][m][1] = V[m].y, d->field[h][m][2] = V[m].z;
			for (int f = 0; f < 6; f++) d->slab[h][m][f][0] = slabArray(&V[m], f, 0), d->slab[h][m][f][1] = slabArray(&V[m], f, 1);
		}
	}
	coeffs *C = W->CE;
	d->eps[0] = C->x, d->eps[1] = C->y, d->eps[2] = C->z;

	alis_dev_ade *obj = calloc(C->N ? C->N : 1, sizeof *obj);
	for (int n = 0; n < C->N; n++) {
		alis_dev_ade *o = &obj[n];
		o->i0 = C->iMin[n], o->i1 = C->iMax[n], o->j0 = C->jMin[n], o->j1 = C->jMax[n], o->k0 = C->kMin[n], o->k1 = C->kMax[n];
		o->nj = C->NJ[n], o->nk = C->NK[n];
		o->a[0] = C->ax[n], o->a[1] = C->ay[n], o->a[2] = C->az[n];
		float **JJ[3] = { C->JJx, C->JJy, C->JJz }, **JK[3] = { C->JKx, C->JKy, C->JKz }, **JE[3] = { C->JEx, C->JEy, C->JEz };
		float **KJ[3] = { C->KJx, C->KJy, C->KJz }, **KK[3] = { C->KKx, C->KKy, C->KKz }, **KE[3] = { C->KEx, C->KEy, C->KEz };
		float ****fill[3] = { C->Jx, C->Jy, C->Jz };
		for (int c = 0; c < 3; c++) {
			o->fill[c] = fill[c][n];
			for (int p = 0; p < o->nk && p < ALIS_DEV_MAXPOLES; p++) {
				o->JJ[c][p] = JJ[c][n][p], o->JE[c][p] = JE[c][n][p];
				if (p >= o->nj) o->JK[c][p] = JK[c][n][p], o->KJ[c][p] = KJ[c][n][p], o->KK[c][p] = KK[c][n][p], o->KE[c][p] = KE[c][n][p];
			}
		}
		for (int m = 0; m <= W->complexField; m++) {
			float *****J[3] = { W->E[m].Jx, W->E[m].Jy, W->E[m].Jz }, *****K[3] = { W->E[m].Kx, W->E[m].Ky, W->E[m].Kz };
			for (int c = 0; c < 3; c++)
				for (int p = 0; p < o->nk && p < ALIS_DEV_MAXPOLES; p++) {
					o->J[m][c][p] = J[c][n][p];
					if (p >= o->nj) o->K[m][c][p] = K[c][n][p];
				}
		}
	}
	d->nobj = C->N, d->obj = obj;

	for (int h = 0; h < 2; h++) {
		int count = 0;
		for (source S = h ? W->SH : W->SE; S; S = S->S) count++;
		alis_dev_src *src = calloc(count ? count : 1, sizeof *src);
		int q = 0;
		for (source S = h ? W->SH : W->SE; S; S = S->S, q++) {
			alis_dev_src *s = &src[q];
			s->h = h, s->m = -1, s->axis = -1;
			vfield *V = h ? W->H : W->E;
			for (int m = 0; m <= W->complexField; m++) {
				if (S->F == V[m].x) s->m = m, s->axis = 0;
				if (S->F == V[m].y) s->m = m, s->axis = 1;
				if (S->F == V[m].z) s->m = m, s->axis = 2;
			}
			s->i0 = S->iMin, s->i1 = S->iMax, s->j0 = S->jMin, s->j1 = S->jMax, s->k0 = S->kMin, s->k1 = S->kMax;
			s->amplitude = S->amplitude;
			s->Amplitude = S->Amplitude;
			s->cellPhase = S->Phase != 0;
		}
		d->nsrc[h] = count, d->src[h] = src;
	}

	int nt = 0;
	for (tfsf S = W->S; S; S = S->S) nt++;
	alis_dev_tfsf *tf = calloc(nt ? nt : 1, sizeof *tf);
	nt = 0;
	for (tfsf S = W->S; S; S = S->S, nt++)
		tf[nt] = (alis_dev_tfsf){ S->iMin, S->iMax, S->jMin, S->jMax, S->kMin, S->kMax, S->iMIN, S->iMAX, S->jMIN, S->jMAX, S->kMIN, S->kMAX };
	d->ntfsf = nt, d->tfsf = tf;
	d->expandX = W->expandX, d->expandY = W->expandY;
}

void freeWorldDesc(alis_dev_world_desc *d)
{
	free((void *)d->obj);
	free((void *)d->src[0]);
	free((void *)d->src[1]);
	free((void *)d->tfsf);
}

/* Host cfield starts of a createPhasers array; released by the caller with free(d->cf). */
void describePhaser(phaser B, alis_dev_phaser_desc *d)
{
	d->nf = B->count;
	d->iMin = B->iMin, d->iMax = B->iMax, d->jMin = B->jMin, d->jMax = B->jMax, d->kMin = B->kMin, d->kMax = B->kMax;
	d->cf = calloc((size_t)d->nf * 24, sizeof(float *));
	for (int q = 0; q < d->nf; q++) {
		phaser P = B + q;
		cfield *F[6] = { &P->xMin, &P->xMax, &P->yMin, &P->yMax, &P->zMin, &P->zMax };
		for (int kind = 0; kind < 6; kind++) {
			float complex ***c[4];
			if (kind < 2) c[0] = F[kind]->Ey, c[1] = F[kind]->Ez, c[2] = F[kind]->Hy, c[3] = F[kind]->Hz;
			else if (kind < 4) c[0] = F[kind]->Ez, c[1] = F[kind]->Ex, c[2] = F[kind]->Hz, c[3] = F[kind]->Hx;
			else c[0] = F[kind]->Ex, c[1] = F[kind]->Ey, c[2] = F[kind]->Hx, c[3] = F[kind]->Hy;
			if (!c[0]) continue;
			int i = kind == 1 ? P->iMax : P->iMin, j = kind == 3 ? P->jMax : P->jMin, k = kind == 5 ? P->kMax : P->kMin;
			for (int comp = 0; comp < 4; comp++) d->cf[((size_t)q * 6 + kind) * 4 + comp] = (float *)&c[comp][i][j][k];
		}
	}
}
