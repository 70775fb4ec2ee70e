/*
 * Half-step arithmetic of the oracle, instantiated for T = float and T = double by
 * engine.c (SUF(name) appends the type suffix).  Plain C loops in the reference's order;
 * every function cites the reference lines it restates.  Loops whose cells are independent
 * (each iteration writes only its own cell and reads only arrays the loop does not write:
 * the curl, the CPML faces, the object terms) run their outer index on OpenMP threads,
 * which leaves every cell's arithmetic and order as in the sequential loop.
 */

/* ---- curl and CPML: src/update.c:48-125 ---- */

/* updatePx/Py/Pz (src/update.c:48-84) for face f (0..5 = xMin yMin zMin xMax yMax zMax). */
static void SUF(cpmlFace)(world *D, int h, int m, int f, int alt, int hasC)
{
	const alis_dev_world_desc *W = &D->W;
	arr *F = D->F[h][m], *G = D->F[1-h][m];
	arr *pa = &D->slab[h][m][f][0], *pb = &D->slab[h][m][f][1];
	const int ax = f % 3;
	const float *b = W->cpml[h][ax][0], *ca = W->cpml[h][ax][1], *cb = W->cpml[h][ax][2];
	const int MIN = ax == 0 ? W->iMIN : ax == 1 ? W->jMIN : W->kMIN;
	int lo, hi;
	if (f < 3) lo = MIN - (alt-1)/2, hi = (ax == 0 ? W->iMin : ax == 1 ? W->jMin : W->kMin) - (alt+3)/2;
	else lo = (ax == 0 ? W->iMax : ax == 1 ? W->jMax : W->kMax) + 2, hi = ax == 0 ? W->iMAX : ax == 1 ? W->jMAX : W->kMAX;
	int i0 = W->iMIN, i1 = W->iMAX, j0 = W->jMIN, j1 = W->jMAX, k0 = W->kMIN, k1 = W->kMAX;
	if (ax == 0) i0 = lo, i1 = hi;
	if (ax == 1) j0 = lo, j1 = hi;
	if (ax == 2) k0 = lo, k1 = hi;
	/* target components and the source components of their curl terms */
	const int ta = (ax+1) % 3, tb = (ax+2) % 3;        /* x: (y,z)  y: (z,x)  z: (x,y) */
	const int di = ax == 0 ? alt : 0, dj = ax == 1 ? alt : 0, dk = ax == 2 ? alt : 0;
	#pragma omp parallel for schedule(static)
	for (int i = i0; i <= i1; i++)
		for (int j = j0; j <= j1; j++)
			for (int k = k0; k <= k1; k++) {
				const int c = ax == 0 ? i : ax == 1 ? j : k;
				const T bb = (T)b[c-MIN], cA = (T)ca[c-MIN], cB = (T)cb[c-MIN];
				T *Pa = &V(pa, i, j, k), *Pb = &V(pb, i, j, k);
				/* Py uses G.z (x face), Pz uses G.x (y face), Px uses G.y (z face) */
				*Pa = bb * *Pa + cA * (V(&G[tb], i, j, k) - V(&G[tb], i+di, j+dj, k+dk));
				if (hasC) V(&F[ta], i, j, k) += *Pa * (T)EPS(D, ta, i, j, k);
				else V(&F[ta], i, j, k) += *Pa;
				*Pb = bb * *Pb + cB * (V(&G[ta], i, j, k) - V(&G[ta], i+di, j+dj, k+dk));
				if (hasC) V(&F[tb], i, j, k) += *Pb * (T)EPS(D, tb, i, j, k);
				else V(&F[tb], i, j, k) += *Pb;
			}
}

/* updateF (src/update.c:87-125) without the object terms: curl, then the CPML faces. */
static void SUF(curl)(world *D, int h, int m)
{
	const alis_dev_world_desc *W = &D->W;
	const int alt = h ? -1 : 1, hasC = !h && D->hasEps;
	arr *F = D->F[h][m], *G = D->F[1-h][m];
	const float *dx = W->dt[h][0], *dy = W->dt[h][1], *dz = W->dt[h][2];
	#pragma omp parallel for schedule(static)
	for (int i = W->iMIN; i <= W->iMAX; i++)
		for (int j = W->jMIN; j <= W->jMAX; j++)
			for (int k = W->kMIN; k <= W->kMAX; k++) {
				const T tx = (T)dx[i-W->iMIN], ty = (T)dy[j-W->jMIN], tz = (T)dz[k-W->kMIN];
				T ux = tz * (V(&G[1], i, j, k) - V(&G[1], i, j, k+alt)) + ty * (V(&G[2], i, j+alt, k) - V(&G[2], i, j, k));
				T uy = tx * (V(&G[2], i, j, k) - V(&G[2], i+alt, j, k)) + tz * (V(&G[0], i, j, k+alt) - V(&G[0], i, j, k));
				T uz = ty * (V(&G[0], i, j, k) - V(&G[0], i, j+alt, k)) + tx * (V(&G[1], i+alt, j, k) - V(&G[1], i, j, k));
				if (hasC) {
					V(&F[0], i, j, k) += ux * (T)EPS(D, 0, i, j, k);
					V(&F[1], i, j, k) += uy * (T)EPS(D, 1, i, j, k);
					V(&F[2], i, j, k) += uz * (T)EPS(D, 2, i, j, k);
				}
				else {
					V(&F[0], i, j, k) += ux;
					V(&F[1], i, j, k) += uy;
					V(&F[2], i, j, k) += uz;
				}
			}
	/* face order of src/update.c:119-124 */
	const int present[6] = { W->iMIN < W->iMin, W->jMIN < W->jMin, W->kMIN < W->kMin, W->iMAX > W->iMax, W->jMAX > W->jMax, W->kMAX > W->kMax };
	for (int f = 0; f < 6; f++)
		if (present[f] && W->cpml[h][f%3][0]) SUF(cpmlFace)(D, h, m, f, alt, hasC);
}

/* ---- dispersive objects: src/update.c:89-108 (E) and 128-158 (H side) ---- */

static void SUF(objectBox)(world *D, int n, int *b)
{
	const alis_dev_world_desc *W = &D->W;
	const alis_dev_ade *o = &D->obj[n];
	b[0] = MAX(W->iMIN, o->i0), b[1] = MIN(W->iMAX, o->i1);
	b[2] = MAX(W->jMIN, o->j0), b[3] = MIN(W->jMAX, o->j1);
	b[4] = MAX(W->kMIN, o->k0), b[5] = MIN(W->kMAX, o->k1);
}

/* Conductivity scaling of every object, then -sum_p J_p of every object (src/update.c:89-108). */
static void SUF(objectsE)(world *D, int m)
{
	arr *F = D->F[0][m];
	for (int n = 0; n < D->W.nobj; n++) {
		const alis_dev_ade *o = &D->obj[n];
		if (!o->a[0] && !o->a[1] && !o->a[2]) continue;
		int b[6];
		SUF(objectBox)(D, n, b);
		for (int i = b[0]; i <= b[1]; i++) for (int j = b[2]; j <= b[3]; j++) for (int k = b[4]; k <= b[5]; k++)
			for (int c = 0; c < 3; c++) {
				const T a = (T)o->a[c], fl = (T)o->fill[c][i][j][k];
				V(&F[c], i, j, k) *= (1 - a * fl) / (1 + a * fl);
			}
	}
	for (int n = 0; n < D->W.nobj; n++) {
		const alis_dev_ade *o = &D->obj[n];
		int b[6];
		SUF(objectBox)(D, n, b);
		for (int p = 0; p < o->nk; p++)
			for (int i = b[0]; i <= b[1]; i++) for (int j = b[2]; j <= b[3]; j++) for (int k = b[4]; k <= b[5]; k++)
				for (int c = 0; c < 3; c++) V(&F[c], i, j, k) -= V(&D->J[n][m][c][p], i, j, k);
	}
}

/* updateJ on the new E (src/update.c:128-158): Drude poles, then the J/K pairs. */
static void SUF(objectsH)(world *D, int m)
{
	arr *E = D->F[0][m];
	for (int n = 0; n < D->W.nobj; n++) {
		const alis_dev_ade *o = &D->obj[n];
		int b[6];
		SUF(objectBox)(D, n, b);
		for (int c = 0; c < 3; c++) {
			for (int p = 0; p < o->nj; p++)
				for (int i = b[0]; i <= b[1]; i++) for (int j = b[2]; j <= b[3]; j++) for (int k = b[4]; k <= b[5]; k++) {
					T *J = &V(&D->J[n][m][c][p], i, j, k);
					*J = (T)o->JJ[c][p] * *J + (T)o->JE[c][p] * (T)o->fill[c][i][j][k] * V(&E[c], i, j, k);
				}
		}
		for (int c = 0; c < 3; c++)
			for (int p = o->nj; p < o->nk; p++)
				for (int i = b[0]; i <= b[1]; i++) for (int j = b[2]; j <= b[3]; j++) for (int k = b[4]; k <= b[5]; k++) {
					T *J = &V(&D->J[n][m][c][p], i, j, k), *K = &V(&D->K[n][m][c][p], i, j, k);
					const T fe = (T)o->fill[c][i][j][k];
					*J = (T)o->JJ[c][p] * *J + (T)o->JK[c][p] * *K + (T)o->JE[c][p] * fe * V(&E[c], i, j, k);
					*K = (T)o->KJ[c][p] * *J + (T)o->KK[c][p] * *K + (T)o->KE[c][p] * fe * V(&E[c], i, j, k);
				}
	}
}

/* ---- TF/SF face strips: update() and updateTFSFE/H (src/update.c:194-272) ---- */

/* F[i][j][k] += dtds * G_aux[i+di][j+dj][k+dk] * C[i][j][k] over a box. */
static void SUF(strip)(world *D, arr *F, arr *G, int c, float dtds, int i0, int i1, int j0, int j1, int k0, int k1, int di, int dj, int dk)
{
	for (int i = i0; i <= i1; i++) for (int j = j0; j <= j1; j++) for (int k = k0; k <= k1; k++) {
		if (c >= 0) V(F, i, j, k) += (T)dtds * V(G, i+di, j+dj, k+dk) * (T)EPS(D, c, i, j, k);
		else V(F, i, j, k) += (T)dtds * V(G, i+di, j+dj, k+dk);
	}
}

static void SUF(strips)(world *D, int s, int h)
{
	const alis_dev_world_desc *W = &D->W;
	const alis_dev_tfsf *S = &D->tf[s];
	world *A = D->aux[s];
	const float *dx = W->dt[h][0], *dy = W->dt[h][1], *dz = W->dt[h][2];
#define DX(i) dx[(i)-W->iMIN]
#define DY(j) dy[(j)-W->jMIN]
#define DZ(k) dz[(k)-W->kMIN]
	const int ce = h ? -1 : 1;   /* eps factors only on E */
#define CC(c) (ce > 0 && D->hasEps ? (c) : -1)
	for (int n = 0; n <= W->complexField; n++) {
		arr *F = D->F[h][n], *G = A->F[1-h][n];
		if (!h) {
			if (S->iMin > W->iMIN) {
				SUF(strip)(D, &F[2], &G[1], CC(2), -DX(S->iMin), S->iMin, S->iMin, S->jMin, S->jMax, S->kMIN, S->kMAX, 0, 0, 0);
				SUF(strip)(D, &F[1], &G[2], CC(1), +DX(S->iMin), S->iMin, S->iMin, S->jMIN, S->jMAX, S->kMin, S->kMax, 0, 0, 0);
			}
			if (S->jMin > W->jMIN) {
				SUF(strip)(D, &F[0], &G[2], CC(0), -DY(S->jMin), S->iMIN, S->iMAX, S->jMin, S->jMin, S->kMin, S->kMax, 0, 0, 0);
				SUF(strip)(D, &F[2], &G[0], CC(2), +DY(S->jMin), S->iMin, S->iMax, S->jMin, S->jMin, S->kMIN, S->kMAX, 0, 0, 0);
			}
			if (S->kMin > W->kMIN) {
				SUF(strip)(D, &F[1], &G[0], CC(1), -DZ(S->kMin), S->iMin, S->iMax, S->jMIN, S->jMAX, S->kMin, S->kMin, 0, 0, 0);
				SUF(strip)(D, &F[0], &G[1], CC(0), +DZ(S->kMin), S->iMIN, S->iMAX, S->jMin, S->jMax, S->kMin, S->kMin, 0, 0, 0);
			}
			if (S->iMax < W->iMAX) {
				SUF(strip)(D, &F[2], &G[1], CC(2), +DX(S->iMax), S->iMax, S->iMax, S->jMin, S->jMax, S->kMIN, S->kMAX, 1, 0, 0);
				SUF(strip)(D, &F[1], &G[2], CC(1), -DX(S->iMax), S->iMax, S->iMax, S->jMIN, S->jMAX, S->kMin, S->kMax, 1, 0, 0);
			}
			if (S->jMax < W->jMAX) {
				SUF(strip)(D, &F[0], &G[2], CC(0), +DY(S->jMax), S->iMIN, S->iMAX, S->jMax, S->jMax, S->kMin, S->kMax, 0, 1, 0);
				SUF(strip)(D, &F[2], &G[0], CC(2), -DY(S->jMax), S->iMin, S->iMax, S->jMax, S->jMax, S->kMIN, S->kMAX, 0, 1, 0);
			}
			if (S->kMax < W->kMAX) {
				SUF(strip)(D, &F[1], &G[0], CC(1), +DZ(S->kMax), S->iMin, S->iMax, S->jMIN, S->jMAX, S->kMax, S->kMax, 0, 0, 1);
				SUF(strip)(D, &F[0], &G[1], CC(0), -DZ(S->kMax), S->iMIN, S->iMAX, S->jMin, S->jMax, S->kMax, S->kMax, 0, 0, 1);
			}
		}
		else {
			if (S->iMin > W->iMIN) {
				SUF(strip)(D, &F[2], &G[1], -1, +DX(S->iMin), S->iMin, S->iMin, S->jMIN, S->jMAX, S->kMin, S->kMax, 0, 0, 0);
				SUF(strip)(D, &F[1], &G[2], -1, -DX(S->iMin), S->iMin, S->iMin, S->jMin, S->jMax, S->kMIN, S->kMAX, 0, 0, 0);
			}
			if (S->jMin > W->jMIN) {
				SUF(strip)(D, &F[0], &G[2], -1, +DY(S->jMin), S->iMin, S->iMax, S->jMin, S->jMin, S->kMIN, S->kMAX, 0, 0, 0);
				SUF(strip)(D, &F[2], &G[0], -1, -DY(S->jMin), S->iMIN, S->iMAX, S->jMin, S->jMin, S->kMin, S->kMax, 0, 0, 0);
			}
			if (S->kMin > W->kMIN) {
				SUF(strip)(D, &F[1], &G[0], -1, +DZ(S->kMin), S->iMIN, S->iMAX, S->jMin, S->jMax, S->kMin, S->kMin, 0, 0, 0);
				SUF(strip)(D, &F[0], &G[1], -1, -DZ(S->kMin), S->iMin, S->iMax, S->jMIN, S->jMAX, S->kMin, S->kMin, 0, 0, 0);
			}
			if (S->iMax < W->iMAX) {
				SUF(strip)(D, &F[2], &G[1], -1, -DX(S->iMax+1), S->iMax+1, S->iMax+1, S->jMIN, S->jMAX, S->kMin, S->kMax, -1, 0, 0);
				SUF(strip)(D, &F[1], &G[2], -1, +DX(S->iMax+1), S->iMax+1, S->iMax+1, S->jMin, S->jMax, S->kMIN, S->kMAX, -1, 0, 0);
			}
			if (S->jMax < W->jMAX) {
				SUF(strip)(D, &F[0], &G[2], -1, -DY(S->jMax+1), S->iMin, S->iMax, S->jMax+1, S->jMax+1, S->kMIN, S->kMAX, 0, -1, 0);
				SUF(strip)(D, &F[2], &G[0], -1, +DY(S->jMax+1), S->iMIN, S->iMAX, S->jMax+1, S->jMax+1, S->kMin, S->kMax, 0, -1, 0);
			}
			if (S->kMax < W->kMAX) {
				SUF(strip)(D, &F[1], &G[0], -1, -DZ(S->kMax+1), S->iMIN, S->iMAX, S->jMin, S->jMax, S->kMax+1, S->kMax+1, 0, 0, -1);
				SUF(strip)(D, &F[0], &G[1], -1, +DZ(S->kMax+1), S->iMin, S->iMax, S->jMIN, S->jMAX, S->kMax+1, S->kMax+1, 0, 0, -1);
			}
		}
	}
#undef CC
#undef DX
#undef DY
#undef DZ
}

/* ---- ghost layers: copyBoundary (src/update.c:182-191) ---- */

static void SUF(copyB)(world *D, int h, int axis, float cosk, float sink, int i0, int i1, int j0, int j1, int k0, int k1, int di, int dj, int dk)
{
	arr *F = &D->F[h][0][axis], *G = D->W.complexField ? &D->F[h][1][axis] : 0;
	for (int i = i0; i <= i1; i++) for (int j = j0; j <= j1; j++) for (int k = k0; k <= k1; k++) {
		if (G) {
			const T fs = V(F, i+di, j+dj, k+dk), gs = V(G, i+di, j+dj, k+dk);
			V(F, i, j, k) = (T)cosk * fs + (T)sink * gs;
			V(G, i, j, k) = (T)cosk * gs - (T)sink * fs;
		}
		else V(F, i, j, k) = (T)cosk * V(F, i+di, j+dj, k+dk);
	}
}

/* The ghost schedules at the end of updateE (src/update.c:284-297) and updateH (:311-339). */
static void SUF(ghosts)(world *D, int h)
{
	const alis_dev_world_desc *W = &D->W;
	const int iAlt = (W->xMinSur==PBC || W->xMinSur==BBC) ? W->iNUM-1 : 1;
	const int jAlt = (W->yMinSur==PBC || W->yMinSur==BBC) ? W->jNUM-1 : 1;
	const int kAlt = (W->zMinSur==PBC || W->zMinSur==BBC) ? W->kNUM-1 : 1;
	enum { X, Y, Z };
	if (!h) {
		SUF(copyB)(D, 0, X, W->coskx, -W->sinkx, W->iMAX+1, W->iMAX+1, W->jMIN, W->jMAX, W->kMIN, W->kMAX, -iAlt, 0, 0);
		SUF(copyB)(D, 0, Z, W->coskz, -W->sinkz, W->iMIN, W->iMAX, W->jMIN, W->jMAX, W->kMAX+1, W->kMAX+1, 0, 0, -kAlt);
		if (W->yMinSur==HBC) {
			SUF(copyB)(D, 0, Y, W->coskym, -W->sinkym, W->iMIN, 0, W->jMAX+1, W->jMAX+1, W->kMIN, W->kMAX, W->iMAX, 1-W->jNUM, 0);
			SUF(copyB)(D, 0, Y, W->coskyp, -W->sinkyp, 1, W->iMAX+1, W->jMAX+1, W->jMAX+1, W->kMIN, W->kMAX, W->iMIN, 1-W->jNUM, 0);
		}
		else SUF(copyB)(D, 0, Y, W->cosky, -W->sinky, W->iMIN, W->iMAX, W->jMAX+1, W->jMAX+1, W->kMIN, W->kMAX, 0, -jAlt, 0);
		return;
	}
	const float cxm = W->xMaxSur==PEC ? 1 : W->coskx, czm = W->zMaxSur==PEC ? 1 : W->coskz, cym = W->yMaxSur==PEC ? 1 : W->cosky;
	SUF(copyB)(D, 1, Y, cxm, -W->sinkx, W->iMAX+1, W->iMAX+1, W->jMIN, W->jMAX+1, W->kMIN, W->kMAX+1, -iAlt, 0, 0);
	SUF(copyB)(D, 1, Z, cxm, -W->sinkx, W->iMAX+1, W->iMAX+1, W->jMIN, W->jMAX+1, W->kMIN, W->kMAX+1, -iAlt, 0, 0);
	SUF(copyB)(D, 1, X, czm, -W->sinkz, W->iMIN, W->iMAX+1, W->jMIN, W->jMAX+1, W->kMAX+1, W->kMAX+1, 0, 0, -kAlt);
	SUF(copyB)(D, 1, Y, czm, -W->sinkz, W->iMIN, W->iMAX+1, W->jMIN, W->jMAX+1, W->kMAX+1, W->kMAX+1, 0, 0, -kAlt);
	SUF(copyB)(D, 1, Y, W->coskx, W->sinkx, W->iMIN, W->iMIN, W->jMIN, W->jMAX+1, W->kMIN, W->kMAX+1, iAlt, 0, 0);
	SUF(copyB)(D, 1, Z, W->coskx, W->sinkx, W->iMIN, W->iMIN, W->jMIN, W->jMAX+1, W->kMIN, W->kMAX+1, iAlt, 0, 0);
	SUF(copyB)(D, 1, X, W->coskz, W->sinkz, W->iMIN, W->iMAX+1, W->jMIN, W->jMAX+1, W->kMIN, W->kMIN, 0, 0, kAlt);
	SUF(copyB)(D, 1, Y, W->coskz, W->sinkz, W->iMIN, W->iMAX+1, W->jMIN, W->jMAX+1, W->kMIN, W->kMIN, 0, 0, kAlt);
	if (W->yMinSur==HBC) {
		SUF(copyB)(D, 1, X, W->coskyp, W->sinkyp, W->iMIN, 0, W->jMIN, W->jMIN, W->kMIN, W->kMAX, W->iMAX, W->jNUM-1, 0);
		SUF(copyB)(D, 1, Z, W->coskyp, W->sinkyp, W->iMIN, 0, W->jMIN, W->jMIN, W->kMIN, W->kMAX, W->iMAX, W->jNUM-1, 0);
		SUF(copyB)(D, 1, X, W->coskym, -W->sinkym, W->iMIN, 0, W->jMAX+1, W->jMAX+1, W->kMIN, W->kMAX, W->iMAX, 1-W->jNUM, 0);
		SUF(copyB)(D, 1, Z, W->coskym, -W->sinkym, W->iMIN, 0, W->jMAX+1, W->jMAX+1, W->kMIN, W->kMAX, W->iMAX, 1-W->jNUM, 0);
		SUF(copyB)(D, 1, X, W->coskym, W->sinkym, 1, W->iMAX+1, W->jMIN, W->jMIN, W->kMIN, W->kMAX, W->iMIN, W->jNUM-1, 0);
		SUF(copyB)(D, 1, Z, W->coskym, W->sinkym, 1, W->iMAX+1, W->jMIN, W->jMIN, W->kMIN, W->kMAX, W->iMIN, W->jNUM-1, 0);
		SUF(copyB)(D, 1, X, W->coskyp, -W->sinkyp, 1, W->iMAX+1, W->jMAX+1, W->jMAX+1, W->kMIN, W->kMAX, W->iMIN, 1-W->jNUM, 0);
		SUF(copyB)(D, 1, Z, W->coskyp, -W->sinkyp, 1, W->iMAX+1, W->jMAX+1, W->jMAX+1, W->kMIN, W->kMAX, W->iMIN, 1-W->jNUM, 0);
	}
	else {
		SUF(copyB)(D, 1, Z, cym, -W->sinky, W->iMIN, W->iMAX+1, W->jMAX+1, W->jMAX+1, W->kMIN, W->kMAX+1, 0, -jAlt, 0);
		SUF(copyB)(D, 1, X, cym, -W->sinky, W->iMIN, W->iMAX+1, W->jMAX+1, W->jMAX+1, W->kMIN, W->kMAX+1, 0, -jAlt, 0);
		SUF(copyB)(D, 1, Z, W->cosky, W->sinky, W->iMIN, W->iMAX+1, W->jMIN, W->jMIN, W->kMIN, W->kMAX+1, 0, jAlt, 0);
		SUF(copyB)(D, 1, X, W->cosky, W->sinky, W->iMIN, W->iMAX+1, W->jMIN, W->jMIN, W->kMIN, W->kMAX+1, 0, jAlt, 0);
	}
}

/* ---- sources: updateS (src/update.c:161-179), the waveform value given by the caller ---- */

static void SUF(inject)(world *D, int h, int index, float wave, const float *cellWave)
{
	const alis_dev_src *s = &D->src[h][index];
	arr *F = &D->F[h][s->m][s->axis];
	long c = 0;
	for (int i = s->i0; i <= s->i1; i++) for (int j = s->j0; j <= s->j1; j++) for (int k = s->k0; k <= s->k1; k++, c++) {
		const T amp = s->Amplitude ? (T)s->Amplitude[i][j][k] : (T)s->amplitude;
		const T w = cellWave ? (T)cellWave[c] : (T)wave;
		V(F, i, j, k) += amp * w;
	}
}

/* ---- phaser running DFT: updatePhaser (src/phaser.c:83-110), one sample ---- */

static void SUF(phaserStep)(world *D, phs *P, int q0, int nq, const double *eE, const double *eH)
{
	for (int f = 0; f < P->nface; f++) {
		const int kind = P->kind[f];
		const int a1 = kind < 2 ? 1 : kind < 4 ? 2 : 0, a2 = kind < 2 ? 2 : kind < 4 ? 0 : 1;
		const int di = kind < 2, dj = kind == 2 || kind == 3, dk = kind >= 4;
		const int fixed = kind == 0 ? P->iMin : kind == 1 ? P->iMax : kind == 2 ? P->jMin : kind == 3 ? P->jMax : kind == 4 ? P->kMin : P->kMax;
		const int nu = kind < 2 ? P->jMax-P->jMin+1 : P->iMax-P->iMin+1;
		const int nv = kind < 4 ? P->kMax-P->kMin+1 : P->jMax-P->jMin+1;
		arr *E = D->F[0][0], *H = D->F[1][0];
		for (int u = 0; u < nu; u++)
			for (int v = 0; v < nv; v++) {
				int i, j, k;
				if (kind < 2) i = fixed, j = P->jMin + u, k = P->kMin + v;
				else if (kind < 4) i = P->iMin + u, j = fixed, k = P->kMin + v;
				else i = P->iMin + u, j = P->jMin + v, k = fixed;
				const T e1 = V(&E[a1], i, j, k), e2 = V(&E[a2], i, j, k);
				const T h1 = V(&H[a1], i, j, k) + V(&H[a1], i+di, j+dj, k+dk);
				const T h2 = V(&H[a2], i, j, k) + V(&H[a2], i+di, j+dj, k+dk);
				const long cell = P->off[f] + (long)u * nv + v;
				for (int q = q0; q < q0 + nq; q++) {
					const double *xe = eE + 2*(q-q0), *xh = eH + 2*(q-q0);
					if (P->dft64) {
						/* double accumulation with Neumaier compensation */
						const long o = 2*((long)q * 4 * P->ncell + cell), m = 2*P->ncell;
						const double y[8] = { xe[0] * (double)e1, xe[1] * (double)e1, xe[0] * (double)e2, xe[1] * (double)e2,
						                      xh[0] * (double)h1 * 0.5, xh[1] * (double)h1 * 0.5, xh[0] * (double)h2 * 0.5, xh[1] * (double)h2 * 0.5 };
						for (int c = 0; c < 8; c++) {
							double *s = P->accD + o + (c / 2) * m + c % 2, *r = P->cmpD + o + (c / 2) * m + c % 2;
							const double t = *s + y[c];
							*r += fabs(*s) >= fabs(y[c]) ? (*s - t) + y[c] : (y[c] - t) + *s;
							*s = t;
						}
					}
					else {
						/* float complex accumulation as in the reference (exp factor rounded to float) */
						float *acc = P->accF + 2*((long)q * 4 * P->ncell);
						float *p0 = acc + 2*cell, *p1 = acc + 2*(P->ncell + cell), *p2 = acc + 2*(2*P->ncell + cell), *p3 = acc + 2*(3*P->ncell + cell);
						const float er = (float)xe[0], ei = (float)xe[1], hr = (float)xh[0], hi = (float)xh[1];
						p0[0] = p0[0] + er * (float)e1, p0[1] = p0[1] + ei * (float)e1;
						p1[0] = p1[0] + er * (float)e2, p1[1] = p1[1] + ei * (float)e2;
						p2[0] = p2[0] + hr * (float)h1 * 0.5f, p2[1] = p2[1] + hi * (float)h1 * 0.5f;
						p3[0] = p3[0] + hr * (float)h2 * 0.5f, p3[1] = p3[1] + hi * (float)h2 * 0.5f;
					}
				}
			}
	}
}

/* ---- observers: src/output.c:233-332 (Poynting flux through a plane) and :375-398
 * (absorbed power); the device entries' sums, restated with the same loops ---- */

/* Term (11 ints) of alis_dev_plane_sums: one row sum per row, inner cells in order. */
static void SUF(planeSums)(world *D, const int *u, const double *w, double *out)
{
	arr *E = D->F[0][0], *H = D->F[1][0];
	const int nrm = u[0], row = u[4], inner = 3 - u[0] - u[4];
	for (int r = u[5]; r <= u[6]; r++) {
		double sum = 0;
		for (int s = u[7]; s <= u[8]; s++) {
			int c[3], d[3];
			c[nrm] = u[1], c[row] = r, c[inner] = s;
			d[0] = c[0], d[1] = c[1], d[2] = c[2];
			d[nrm]++;
			const T e = V(&E[u[2]], c[0], c[1], c[2]);
			const T hh = V(&H[u[3]], c[0], c[1], c[2]) + V(&H[u[3]], d[0], d[1], d[2]);
			sum += w[u[9] + s - u[7]] * (double)e * (double)hh;
		}
		out[u[10] + r - u[5]] = sum;
	}
}

/* Column sums of object n's absorbed power over the box b (i0 i1 j0 j1 k0 k1). */
static void SUF(absorption)(world *D, int n, const int *b, const double *dzE, const double *dzH, double *out)
{
	const alis_dev_ade *o = &D->obj[n];
	arr *E = D->F[0][0];
	long col = 0;
	for (int i = b[0]; i <= b[1]; i++)
		for (int j = b[2]; j <= b[3]; j++, col++) {
			double sum[3] = {0, 0, 0}, first[3] = {0, 0, 0}, last[3] = {0, 0, 0};
			for (int k = b[4]; k <= b[5]; k++)
				for (int q = 0; q < 3; q++) {
					const T c = (T)EPS(D, q, i, j, k);
					if (c == 0) continue;
					const T e = V(&E[q], i, j, k), a = (T)o->a[q];
					T f = 0;
					if (a != 0) f += (T)2 * (T)o->fill[q][i][j][k] * e / ((T)1 + (T)1 / a);
					for (int p = 0; p < o->nk; p++) f += V(&D->J[n][0][q][p], i, j, k);
					const T d = (T)(q < 2 ? dzE[k - b[4]] : dzH[k - b[4]]);
					const double term = (double)(d * e * (f / c));
					sum[q] += term;
					if (k == b[4]) first[q] = term;
					if (k == b[5]) last[q] = term;
				}
			double *r = out + col * 9;
			for (int q = 0; q < 3; q++) r[q] = sum[q], r[3 + q] = first[q], r[6 + q] = last[q];
		}
}

/* ---- second-order (three-wave) coupling: nonLinearSHG / nonLinearTWM (src/update.c:344-411) ----
 * Over the object's box without its upper plane, every nonzero d[c][q] adds its terms to the
 * signal's E component c in the reference's order.  A pump value is either at the
 * component's own position or the mean of four values (0.25 times their sum taken in T);
 * with T = float a product of two own-position values stays float and any product with a
 * mean is double, as the reference's expressions promote them. */
static double SUF(twMean)(arr *G, int kind, int i, int j, int k)
{
	/* component, then the four (di, dj, dk) */
	static const int tab[6][13] = {
		{1, 0,0,0, -1,0,0, 0,1,0, -1,1,0},  {2, 0,0,0, -1,0,0, 0,0,1, -1,0,1},
		{2, 0,0,0, 0,-1,0, 0,0,1, 0,-1,1},  {0, 0,0,0, 0,-1,0, 1,0,0, 1,-1,0},
		{0, 0,0,0, 0,0,-1, 1,0,0, 1,0,-1},  {1, 0,0,0, 0,0,-1, 0,1,0, 0,1,-1}};
	const int *t = tab[kind];
	arr *A = &G[t[0]];
	T s = V(A, i + t[1], j + t[2], k + t[3]);
	for (int q = 1; q < 4; q++) s = s + V(A, i + t[1 + 3 * q], j + t[2 + 3 * q], k + t[3 + 3 * q]);
	return 0.25 * s;
}

static void SUF(threeWave)(world *S, world *P1, world *P2, int n, const float *d, float alt, const int *b)
{
	/* per component c: (d index, kind of the first pump's value, kind of the second's) */
	static const int terms[3][9][3] = {
		{{0, TW_OWN, TW_OWN}, {1, TW_YX, TW_YX}, {2, TW_ZX, TW_ZX}, {3, TW_YX, TW_ZX}, {3, TW_ZX, TW_YX},
		 {4, TW_ZX, TW_OWN}, {4, TW_OWN, TW_ZX}, {5, TW_OWN, TW_YX}, {5, TW_YX, TW_OWN}},
		{{0, TW_XY, TW_XY}, {1, TW_OWN, TW_OWN}, {2, TW_ZY, TW_ZY}, {3, TW_OWN, TW_ZY}, {3, TW_ZY, TW_OWN},
		 {4, TW_ZY, TW_XY}, {4, TW_XY, TW_ZY}, {5, TW_XY, TW_OWN}, {5, TW_OWN, TW_XY}},
		{{0, TW_XZ, TW_XZ}, {1, TW_YZ, TW_YZ}, {2, TW_OWN, TW_OWN}, {3, TW_YZ, TW_OWN}, {3, TW_OWN, TW_YZ},
		 {4, TW_OWN, TW_XZ}, {4, TW_XZ, TW_OWN}, {5, TW_XZ, TW_YZ}, {5, TW_YZ, TW_XZ}}};
	const alis_dev_ade *o = &S->obj[n];
	for (int m = 0; m <= S->W.complexField; m++) {
		arr *F = S->F[0][m], *A = P1->F[0][m], *B = P2->F[0][m];
		for (int i = b[0]; i <= b[1]; i++)
			for (int j = b[2]; j <= b[3]; j++)
				for (int k = b[4]; k <= b[5]; k++)
					for (int c = 0; c < 3; c++)
						for (int t = 0; t < 9; t++) {
							const int *u = terms[c][t];
							const float dq = d[c * 6 + u[0]];
							if (!dq) continue;
							const T coef = (T)(2 * alt) * (T)dq * (T)EPS(S, c, i, j, k) * (T)o->fill[c][i][j][k];
							T *f = &V(&F[c], i, j, k);
							if (u[1] == TW_OWN && u[2] == TW_OWN) {
								const T g = V(&A[c], i, j, k) * V(&B[c], i, j, k);
								*f = *f + coef * g;
							}
							else {
								const double ga = u[1] == TW_OWN ? (double)V(&A[c], i, j, k) : SUF(twMean)(A, u[1], i, j, k);
								const double gb = u[2] == TW_OWN ? (double)V(&B[c], i, j, k) : SUF(twMean)(B, u[2], i, j, k);
								*f = (T)((double)*f + (double)coef * (ga * gb));
							}
						}
	}
}
