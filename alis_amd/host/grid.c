/*
 * World construction (the reference's createWorld family, src/world.c:12-466), organised
 * around the engine's view of a world: three axes, each with its index extents, boundary
 * codes, coordinate map and per-cell step coefficients, and a field set described by one
 * table of its arrays -- the components with a ghost layer, the two CPML auxiliary arrays of
 * each PML face, the polarisation currents of each dispersive object -- that a visitor
 * allocates, clears or releases.  The arithmetic the reference fixes is kept as it is
 * written there (nearest-node search, graded coordinates, cell widths, the CPML profile,
 * extents and Bloch phases, with their float/double promotions): the setup hashes of the
 * parity fixtures pin it bit for bit.
 */
#include <stdint.h>
#include <string.h>
#include "internal.h"

/* ---- coordinate maps ---- */

/* Node index nearest to coordinate x for a monotone node -> coordinate map: sixteen halving
 * steps from node 0, then the nearer of the two neighbours (src/world.c:12-25; the tie
 * behaviour of this search is part of the setup the fixtures pin). */
static int nearestNode(world W, float (*coord)(world, float), float x)
{
	int i = 0;
	for (int b = 15; b >= 0; b--) {
		const float c = coord(W, i);
		if (x > c) i += 1u << b;
		if (x < c) i -= 1u << b;
	}
	const float dist = fabsf(x - coord(W, i));
	if (fabsf(x - coord(W, i + 1)) < dist) i++;
	if (fabsf(x - coord(W, i - 1)) < dist) i--;
	return i;
}

/* Graded axis (src/world.c:27-42): spacing ds inside the zone [lo, hi] (half a cell beyond
 * its end nodes), growing by a ratio r per cell outside it until the spacing is m ds after n
 * cells, constant after that.  With L = 10m - m/2 - 1/2 (the grown stretch spans L ds),
 * n = round(L ln m / (m - 1)) and r solves r^n - L ln r = 1, by Newton's method from 10/9.
 * The float and double steps are the reference's. */
static float gradedCoord(int m, float lo, float hi, float ds, float i)
{
	const float d = 10;
	const double L = d*m - 0.5*m - 0.5;
	const float iLo = lo/ds - 0.5, iHi = hi/ds + 0.5;
	const int n = floor(L*logf(m)/(m-1) + 0.5);
	float r = d/(d-1), prev;
	do prev = r, r = prev - (powf(prev,n) - L*logf(prev) - 1)/(n*powf(prev,n-1) - L/prev);
	while (fabsf(r-prev) > 1E-6);
	if (i >= iHi+n) return ds * (iHi + L + (i-iHi-n) * m);
	if (i <= iLo-n) return ds * (iLo - L + (i-iLo+n) * m);
	if (i > iHi) return ds * (iHi + (powf(r,i-iHi)-1) / logf(r));
	if (i < iLo) return ds * (iLo - (powf(r,iLo-i)-1) / logf(r));
	return ds * i;
}

static float uniformX(world W, float i) { return W->dx*i; }
static float uniformY(world W, float j) { return W->dy*j; }
static float uniformZ(world W, float k) { return W->dz*k; }
static float gradedX(world W, float i) { return gradedCoord(W->Res.nug.x.m, W->Res.nug.x.Min, W->Res.nug.x.Max, W->dx, i); }
static float gradedY(world W, float j) { return gradedCoord(W->Res.nug.y.m, W->Res.nug.y.Min, W->Res.nug.y.Max, W->dy, j); }
static float gradedZ(world W, float k) { return gradedCoord(W->Res.nug.z.m, W->Res.nug.z.Min, W->Res.nug.z.Max, W->dz, k); }
/* Nearest nodes already found, per axis: programs read the same probe points every step, and
 * the search calls the coordinate function 19 times (80 ns a coordinate on the k-point worlds).
 * A world's grid is fixed when it is created, so an entry holds for its lifetime.  Entries
 * pack the coordinate's bits and the index into one 64-bit word, written and read whole, so a
 * reader on another thread sees a whole entry or an old one (0: empty). */
#define NODECACHE 64
static int cachedNode(world W, int a, float x, float (*coord)(world, float))
{
	uint64_t (*c)[NODECACHE] = W->nodes;
	if (!c) c = W->nodes = calloc(3 * NODECACHE, sizeof(uint64_t));
	uint32_t b;
	memcpy(&b, &x, sizeof b);
	volatile uint64_t *e = &c[a][(b * 2654435761u) >> 26];
	const uint64_t v = *e;
	if (v && (uint32_t)(v >> 32) == b) return (int)(uint32_t)v;
	const int i = nearestNode(W, coord, x);
	*e = (uint64_t)b << 32 | (uint32_t)i;
	return i;
}

static int nodeX(world W, float x) { return cachedNode(W, 0, x, W->itox); }
static int nodeY(world W, float y) { return cachedNode(W, 1, y, W->jtoy); }
static int nodeZ(world W, float z) { return cachedNode(W, 2, z, W->ktoz); }

axis axisOf(world W, int a)
{
	if (a == 0) return (axis) {&W->iMin, &W->iMax, &W->iNum, &W->iMIN, &W->iMAX, &W->iNUM, &W->xMinSur, &W->xMaxSur,
		&W->xMin, &W->xMax, &W->xMIN, &W->xMAX, &W->xSize, &W->dx, &W->coskx, &W->sinkx, &W->itox, &W->xtoi};
	if (a == 1) return (axis) {&W->jMin, &W->jMax, &W->jNum, &W->jMIN, &W->jMAX, &W->jNUM, &W->yMinSur, &W->yMaxSur,
		&W->yMin, &W->yMax, &W->yMIN, &W->yMAX, &W->ySize, &W->dy, &W->cosky, &W->sinky, &W->jtoy, &W->ytoj};
	return (axis) {&W->kMin, &W->kMax, &W->kNum, &W->kMIN, &W->kMAX, &W->kNUM, &W->zMinSur, &W->zMaxSur,
		&W->zMin, &W->zMax, &W->zMIN, &W->zMAX, &W->zSize, &W->dz, &W->coskz, &W->sinkz, &W->ktoz, &W->ztok};
}


/* ---- arrays ---- */

/* Three-level array over [iMin..iMax]x[jMin..jMax]x[kMin..kMax] on one zeroed block,
 * with shifted bases so that f[i][j][k] indexes directly.  The block layout (pointer
 * table, row table, data) is the one removeField() releases (src/world.c:58-70). */
float ***makeField(int iMin, int iMax, int jMin, int jMax, int kMin, int kMax)
{
	long ni = 1-iMin+iMax, nj = 1-jMin+jMax, nk = 1-kMin+kMax;
	float ***f;
	MALLOC(f, iMin, ni);
	MALLOC(f[iMin], jMin, ni*nj);
	MALLOC(f[iMin][jMin], kMin, ni*nj*nk);
	float **rows = f[iMin] + jMin;
	float *data = f[iMin][jMin] + kMin;
	for (long a = 0; a < ni; a++) {
		f[iMin+a] = rows + a*nj - jMin;
		for (long b = 0; b < nj; b++) f[iMin+a][jMin+b] = data + (a*nj + b)*nk - kMin;
	}
	return f;
}

float complex ***makeComplexField(int iMin, int iMax, int jMin, int jMax, int kMin, int kMax)
{
	long ni = 1-iMin+iMax, nj = 1-jMin+jMax, nk = 1-kMin+kMax;
	float complex ***f;
	MALLOC(f, iMin, ni);
	MALLOC(f[iMin], jMin, ni*nj);
	MALLOC(f[iMin][jMin], kMin, ni*nj*nk);
	float complex **rows = f[iMin] + jMin;
	float complex *data = f[iMin][jMin] + kMin;
	for (long a = 0; a < ni; a++) {
		f[iMin+a] = rows + a*nj - jMin;
		for (long b = 0; b < nj; b++) f[iMin+a][jMin+b] = data + (a*nj + b)*nk - kMin;
	}
	return f;
}

float ***makeFieldFilled(int iMin, int iMax, int jMin, int jMax, int kMin, int kMax, float v)
{
	float ***f = makeField(iMin, iMax, jMin, jMax, kMin, kMax);
	long n = (long)(1-iMin+iMax)*(1-jMin+jMax)*(1-kMin+kMax);
	float *p = &f[iMin][jMin][kMin];
	for (long a = 0; a < n; a++) p[a] = v;
	return f;
}

void resetField3(float ***f, int iMin, int iMax, int jMin, int jMax, int kMin, int kMax)
{
	long n = (long)(1-iMin+iMax)*(1-jMin+jMax)*(1-kMin+kMax);
	memset(&f[iMin][jMin][kMin], 0, n*sizeof(float));
}

/* What a visitor does to each array of a field set. */
enum { ARR_MAKE, ARR_ZERO, ARR_FREE };

static void visitArray(float ****f, const int b[6], int what)
{
	if (what == ARR_MAKE) *f = makeField(b[0], b[1], b[2], b[3], b[4], b[5]);
	else if (!*f) return;
	else if (what == ARR_ZERO) resetField3(*f, b[0], b[1], b[2], b[3], b[4], b[5]);
	else removeField((*f), b[0], b[2], b[4]);
}

/* The arrays of part m of a field set (Off: 0 for E, 1 for H): the three components over
 * the grid with one ghost layer, then per PML face (xMin yMin zMin xMax yMax zMax) its two
 * auxiliary arrays (src/world.c:164-189): on a min face from MIN+Off up to Min+Off-2 along the
 * face's axis, on a max face from Max+2 up to MAX, the full grid across.  The components are
 * skipped where `comps` is 0 (an auxiliary world's rows alias its parent's). */
static void visitFieldSet(world W, vfield *V, int Off, int what, int comps)
{
	for (int m = 0; m <= W->complexField; m++) {
		vfield *F = &V[m];
		const int g[6] = {W->iMIN-1, W->iMAX+1, W->jMIN-1, W->jMAX+1, W->kMIN-1, W->kMAX+1};
		float ****comp[3] = {&F->x, &F->y, &F->z};
		for (int a = 0; a < 3 && comps; a++) visitArray(comp[a], g, what);
		float ****slab[6][2] = {{&F->xMinPy, &F->xMinPz}, {&F->yMinPz, &F->yMinPx}, {&F->zMinPx, &F->zMinPy},
		                        {&F->xMaxPy, &F->xMaxPz}, {&F->yMaxPz, &F->yMaxPx}, {&F->zMaxPx, &F->zMaxPy}};
		for (int f = 0; f < 6; f++) {
			const axis X = axisOf(W, f % 3);
			if ((f < 3 ? *X.minSur : *X.maxSur) != PML) continue;
			int b[6] = {W->iMIN, W->iMAX, W->jMIN, W->jMAX, W->kMIN, W->kMAX};
			if (f < 3) b[2*f] = *X.MIN + Off, b[2*f+1] = *X.Min + Off - 2;
			else b[2*(f-3)] = *X.Max + 2;
			visitArray(slab[f][0], b, what);
			visitArray(slab[f][1], b, what);
		}
	}
}

/* The polarisation currents J (every pole) and K (second arrays of pole pairs) of every
 * dispersive object of C over its box, per part (src/world.c:283-315): their pointer tables
 * are made with ARR_MAKE and released with ARR_FREE. */
static void visitCurrents(world W, vfield *V, coeffs *C, int what)
{
	for (int m = 0; m <= W->complexField; m++) {
		vfield *F = &V[m];
		float ******J[3] = {&F->Jx, &F->Jy, &F->Jz}, ******K[3] = {&F->Kx, &F->Ky, &F->Kz};
		for (int a = 0; a < 3 && what == ARR_MAKE; a++) {
			MALLOC(*J[a], 0, C->N);
			MALLOC(*K[a], 0, C->N);
		}
		for (int n = 0; n < C->N; n++) {
			const int b[6] = {C->iMin[n], C->iMax[n], C->jMin[n], C->jMax[n], C->kMin[n], C->kMax[n]};
			const int NJ = C->NJ[n], NK = C->NK[n];
			for (int a = 0; a < 3; a++) {
				if (what == ARR_MAKE) {
					MALLOC((*J[a])[n], 0, NK);
					MALLOC((*K[a])[n], NJ, NK-NJ);
				}
				for (int p = 0; p < NK; p++) visitArray(&(*J[a])[n][p], b, what);
				for (int p = NJ; p < NK; p++) visitArray(&(*K[a])[n][p], b, what);
				if (what == ARR_FREE) free((*J[a])[n]), free((*K[a])[n] + NJ);
			}
		}
		for (int a = 0; a < 3 && what == ARR_FREE; a++) free(*J[a]), free(*K[a]);
	}
}

void makeAuxFieldArrays(world W, vfield *F, coeffs *C) { visitCurrents(W, F, C, ARR_MAKE); }

static vfield *newFieldSet(world W, int Off)
{
	vfield *F;
	MALLOC(F, 0, 1+W->complexField);
	visitFieldSet(W, F, Off, ARR_MAKE, 1);
	return F;
}


/* ---- step coefficients ---- */

/* CPML profile on one axis (src/world.c:98-132).  For n = 1..Num cells into a PML face
 * (half a cell further for the H grid, Offset 1): sigma = 0.8 SigmaMax (m+1) x^m and
 * kappa = 1 + (KappaMax-1) x^m with x = (n - Offset/2)/Num and m = SigmaPower, alpha =
 * AlphaMax ((Num - n + Offset/2)/Num)^AlphaPower; dt/ds of the cell is divided by kappa and
 * the recursion b = Ca, c = -+dt/ds Cb is stored at the cell. */
static void cpmlProfile(sur S, int Offset, float **b, float **c1, float **c2, float *dtds, float dt,
                        int MinSur, int MaxSur, int Min, int Max, int MIN, int NUM)
{
	if (!S.pml.SigmaMax) S.pml.SigmaMax = 1;
	if (!S.pml.KappaMax) S.pml.KappaMax = S.pml.Num < 24 ? 1 : 10;
	if (!S.pml.AlphaMax) S.pml.AlphaMax = S.pml.Num < 24 ? 0 : 0.01;
	if (!S.pml.SigmaPower) S.pml.SigmaPower = 3.5;
	if (!S.pml.AlphaPower) S.pml.AlphaPower = 1;
	if (MinSur != PML && MaxSur != PML) return;
	MALLOC(*b, MIN, NUM);
	MALLOC(*c1, MIN, NUM);
	MALLOC(*c2, MIN, NUM);
	const int at[2] = {Min - 1 + Offset, Max + 1}, dir[2] = {-1, +1}, on[2] = {MinSur == PML, MaxSur == PML};
	for (int n = 1; n <= S.pml.Num; n++) {
		float Sigma = 0.8*S.pml.SigmaMax*(S.pml.SigmaPower+1)*powf((n-0.5*Offset)/S.pml.Num, S.pml.SigmaPower);
		float Kappa = 1+(S.pml.KappaMax-1)*powf((n-0.5*Offset)/S.pml.Num, S.pml.SigmaPower);
		float Alpha = S.pml.AlphaMax*powf((S.pml.Num-(n-0.5*Offset))/S.pml.Num, S.pml.AlphaPower);
		for (int s = 0; s < 2; s++) {
			if (!on[s]) continue;
			const int q = at[s] + dir[s] * n;
			dtds[q] /= Kappa;
			float Ca = (1-0.5*(Sigma*dtds[q]+Alpha*dt))/(1+0.5*(Sigma*dtds[q]+Alpha*dt));
			float Cb = Sigma*dtds[q]/(1+0.5*(Sigma*dtds[q]+Alpha*dt));
			(*b)[q] = Ca;
			(*c1)[q] = -dtds[q]*Cb;
			(*c2)[q] = +dtds[q]*Cb;
		}
	}
}

/* Cell widths and dt/width of the E (Offset 0) or H (Offset 1) grid per axis, widths
 * mirrored across a symmetry plane, then the CPML profiles (src/world.c:135-161). */
static coeffs *stepCoeffs(world W, sur S, int Offset)
{
	coeffs *C;
	MALLOC(C, 0, 1);
	S.pml.Num = S.pml.Num ? S.pml.Num : 12;
	float **width[3] = {&C->dx, &C->dy, &C->dz}, **ratio[3] = {&C->dtdx, &C->dtdy, &C->dtdz};
	float **prof[3][3] = {{&C->xP, &C->xPy, &C->xPz}, {&C->yP, &C->yPz, &C->yPx}, {&C->zP, &C->zPx, &C->zPy}};
	for (int a = 0; a < 3; a++) {
		const axis X = axisOf(W, a);
		const int sym = *X.minSur == SYM;
		MALLOC(*width[a], sym ? -*X.MAX : *X.MIN, sym ? 2 * *X.NUM - 1 : *X.NUM);
		MALLOC(*ratio[a], *X.MIN, *X.NUM);
		float *d = *width[a], *r = *ratio[a];
		for (int i = *X.MIN; i <= *X.MAX; i++) {
			d[i] = (*X.coord)(W, i-0.5*(Offset-1)) - (*X.coord)(W, i-0.5*(Offset+1));
			r[i] = W->dt / d[i];
		}
		if (sym) for (int i = 1; i <= *X.MAX; i++) d[-i] = d[i];
	}
	for (int a = 0; a < 3; a++) {
		const axis X = axisOf(W, a);
		cpmlProfile(S, Offset, prof[a][0], prof[a][1], prof[a][2], *ratio[a], W->dt, *X.minSur, *X.maxSur, *X.Min, *X.Max, *X.MIN, *X.NUM);
	}
	C->N = 0;
	return C;
}

static void freeStepCoeffs(world W, coeffs *C)
{
	float *prof[3][3] = {{C->xP, C->xPy, C->xPz}, {C->yP, C->yPz, C->yPx}, {C->zP, C->zPx, C->zPy}};
	float *width[3] = {C->dx, C->dy, C->dz}, *ratio[3] = {C->dtdx, C->dtdy, C->dtdz};
	for (int a = 0; a < 3; a++) {
		const axis X = axisOf(W, a);
		if (prof[a][0]) for (int c = 0; c < 3; c++) free(&prof[a][c][*X.MIN]);
		free(&ratio[a][*X.MIN]);
		free(&width[a][*X.minSur != SYM ? *X.MIN : -*X.MAX]);
	}
	free(C);
}


/* ---- worlds ---- */

/* Grid, boundary codes, extents and Bloch phases of a new world (src/world.c:192-258). */
static world newWorld(res R, dom D, sur S)
{
	world W;
	MALLOC(W, 0, 1);
	float *domLo[3] = {&D.x.Min, &D.y.Min, &D.z.Min}, *domHi[3] = {&D.x.Max, &D.y.Max, &D.z.Max};
	float *surLo[3] = {&S.x.MinSur, &S.y.MinSur, &S.z.MinSur}, *surHi[3] = {&S.x.MaxSur, &S.y.MaxSur, &S.z.MaxSur};
	float *zoneLo[3] = {&R.nug.x.Min, &R.nug.y.Min, &R.nug.z.Min}, *zoneHi[3] = {&R.nug.x.Max, &R.nug.y.Max, &R.nug.z.Max};
	const int zoneM[3] = {R.nug.x.m, R.nug.y.m, R.nug.z.m};

	/* as given by the user: a graded zone by its half width; PBC with a period is Bloch */
	for (int a = 0; a < 3; a++) {
		if (*zoneLo[a] > 0) *zoneHi[a] = *zoneLo[a], *zoneLo[a] = -*zoneHi[a];
		if (*surLo[a] == PBC && *surHi[a]) *surLo[a] = BBC;
	}
	if (!S.pml.Num) S.pml.Num = 12;
	W->Dom = D, W->Res = R, W->Sur = S;      /* the domain as given: 2-D tests compare Min with Max */
	W->t = 0, W->dt = R.dt;
	W->dx = R.d.x, W->dy = R.d.y ? R.d.y : R.d.x, W->dz = R.d.z ? R.d.z : R.d.x;
	W->eV = R.spec.eV ? R.spec.eV : 1;
	W->f = INF;

	float (*uniform[3])(world, float) = {uniformX, uniformY, uniformZ}, (*graded[3])(world, float) = {gradedX, gradedY, gradedZ};
	float (*user[3])(world, float) = {R.gridFunction.itox, R.gridFunction.jtoy, R.gridFunction.ktoz};
	int (*node[3])(world, float) = {nodeX, nodeY, nodeZ};
	for (int a = 0; a < 3; a++) {
		const axis X = axisOf(W, a);
		/* a single extent is a half width; a symmetry plane mirrors the domain */
		if ((!*domHi[a] && *domLo[a] > 0) || *surLo[a] == SYM) *domLo[a] = -(*domHi[a] = *domHi[a] ? *domHi[a] : *domLo[a]);
		*X.coord = user[a] ? user[a] : zoneM[a] > 1 ? graded[a] : uniform[a];
		*X.node = node[a];
		/* unset codes inherit the x minimum side's; a max side keeps its min side's code unless
		 * it names its own (a period of PBC/BBC is not a code) */
		*X.minSur = a && *surLo[a] ? *surLo[a] : S.x.MinSur;
		*X.maxSur = *surLo[a] != BBC && *surHi[a] ? *surHi[a] : *X.minSur;
	}
	for (int a = 0; a < 3; a++) {
		const axis X = axisOf(W, a);
		const int given = *domLo[a] || *domHi[a];
		*X.Min = *X.minSur == SYM ? 0 : given ? (*X.node)(W, *domLo[a]) : -1;
		*X.Max = given ? (*X.node)(W, *domHi[a]) : 1;
		*X.Num = 1 - *X.Min + *X.Max;
		*X.MIN = *X.Min - (*X.minSur == PML ? S.pml.Num+1 : 0);
		*X.MAX = *X.Max + (*X.maxSur == PML ? S.pml.Num+1 : 0);
		*X.NUM = 1 - *X.MIN + *X.MAX;
		*X.hi = (*X.coord)(W, *X.Max);
		*X.lo = *X.minSur == SYM ? -*X.hi : (*X.coord)(W, *X.Min);
		*X.size = *X.hi - *X.lo;
		*X.HI = (*X.coord)(W, *X.MAX);
		*X.LO = *X.minSur == SYM ? -*X.HI : (*X.coord)(W, *X.MIN);
		/* phase across one period of a Bloch axis; 0 on a symmetry plane (its source sets
		 * the parity) */
		const int bloch = *surLo[a] == BBC && *surHi[a];
		*X.cos = bloch ? cosf(2*PI*(*X.hi-*X.lo)/ *surHi[a]) : *surLo[a] == SYM ? 0 : 1;
		*X.sin = bloch ? sinf(2*PI*(*X.hi-*X.lo)/ *surHi[a]) : 0;
		if (*surLo[a] == BBC) W->complexField = 1;
	}
	/* hexagonal cells: the phases of the two diagonal neighbours across y */
	const int hex = S.y.MinSur == HBC && S.y.MaxSur;
	W->coskyp = hex ? cosf(PI*(2*(W->yMax-W->yMin)/S.y.MaxSur+(W->xMax-W->xMin)/S.x.MaxSur)) : 1;
	W->coskym = hex ? cosf(PI*(2*(W->yMax-W->yMin)/S.y.MaxSur-(W->xMax-W->xMin)/S.x.MaxSur)) : 1;
	W->sinkyp = hex ? sinf(PI*(2*(W->yMax-W->yMin)/S.y.MaxSur+(W->xMax-W->xMin)/S.x.MaxSur)) : 0;
	W->sinkym = hex ? sinf(PI*(2*(W->yMax-W->yMin)/S.y.MaxSur-(W->xMax-W->xMin)/S.x.MaxSur)) : 0;
	W->xyArea = W->xSize * W->ySize, W->xzArea = W->xSize * W->zSize, W->yzArea = W->ySize * W->zSize;
	alisPrecisionDefault(&W->devPrec, &W->devDftPrec);
	return W;
}

world createWorld(dom D, res R, sur S, char *format, ...)
{
	world W = newWorld(R, D, S);
	W->E = newFieldSet(W, 0);
	W->H = newFieldSet(W, 1);
	W->CE = stepCoeffs(W, S, 0);
	W->CH = stepCoeffs(W, S, 1);
	va_list ap;
	va_start(ap, format);
	vsnprintf(W->ID, sizeof W->ID, format, ap);
	va_end(ap);
	if (W->ID[0]) {
		/* output files of a previous run with the same ID go (src/world.c:275-278) */
		printf("Deleting previous files.");
		exec("rm -fr %s %s.c?* %s[#--0-~]* %s.[#-bd-~]* 2> /dev/null", W->ID, W->ID, W->ID, W->ID);
	}
	alisRegisterWorld(W);
	return W;
}

/* A world sharing SW's geometry and coefficients with fields of its own (src/world.c:318-331). */
world cloneWorld(world SW)
{
	world W;
	MALLOC(W, 0, 1);
	*W = *SW;
	W->dev = 0, W->hostStale = 0, W->parent = 0, W->probes = 0, W->nodes = 0;
	W->E = newFieldSet(W, 0);
	W->H = newFieldSet(W, 1);
	visitCurrents(W, W->E, W->CE, ARR_MAKE);
	visitCurrents(W, W->H, W->CH, ARR_MAKE);
	alisRegisterWorld(W);
	return W;
}

/* All fields, slabs and currents to zero, no sources, the clock at 0 (src/world.c:371-380). */
void rebootWorld(world W)
{
	alisInvalidate(W);
	removeSources(W);
	visitFieldSet(W, W->E, 0, ARR_ZERO, 1);
	visitFieldSet(W, W->H, 1, ARR_ZERO, 1);
	visitCurrents(W, W->E, W->CE, ARR_ZERO);
	visitCurrents(W, W->H, W->CH, ARR_ZERO);
	W->t = W->E->t = 0;
	W->H->t = 0.5;
	W->n = 0;
	printf("\n");
}

static void releaseWorld(world W, int ownCoeffs)
{
	alisForgetWorld(W);
	removeSources(W);
	removeObjects(W);
	if (ownCoeffs) freeStepCoeffs(W, W->CE), freeStepCoeffs(W, W->CH);
	else visitCurrents(W, W->E, W->CE, ARR_FREE), visitCurrents(W, W->H, W->CH, ARR_FREE);
	const int comps = !W->isAux || !(W->expandX || W->expandY);
	visitFieldSet(W, W->E, 0, ARR_FREE, comps);
	visitFieldSet(W, W->H, 1, ARR_FREE, comps);
	free(W->E), free(W->H);
	free(W->nodes);
	free(W);
}

void deleteWorld(world W) { releaseWorld(W, 1); }
void deleteCloneWorld(world W) { releaseWorld(W, 0); }


/*
 * Array files.  The image has no HDF5 library, so h5write/h5read use a small self-describing
 * binary container under the same file name (NAME.h5): the 8 bytes "ALISARR1", int32 rank,
 * int32 dims[rank], then float32 data in C order.  Files written by h5write read back with
 * h5read; real HDF5 files are rejected with a message.
 */
void *h5read(int *dim, char *format, ...)
{
	char name[1024], tag[8];
	int rank = 0, d[3] = {1, 1, 1};
	va_list ap;
	va_start(ap, format);
	vsnprintf(name, sizeof name - 4, format, ap);
	va_end(ap);
	strcat(name, ".h5");
	FILE *fp = fopen(name, "rb");
	if (!fp || fread(tag, 1, 8, fp) != 8 || memcmp(tag, "ALISARR1", 8) || fread(&rank, 4, 1, fp) != 1 ||
	    rank < 1 || rank > 3 || fread(d, 4, rank, fp) != (size_t)rank) {
		printf("h5read: %s is missing or not an array file written by h5write.\n", name);
		exit(0);
	}
	if (dim) for (int n = 0; n < rank; n++) dim[n] = d[n];
	long total = (long)d[0]*(rank > 1 ? d[1] : 1)*(rank > 2 ? d[2] : 1);
	float *data = calloc(total, sizeof(float));
	if (fread(data, sizeof(float), total, fp) != (size_t)total) printf("h5read: %s is truncated.\n", name), exit(0);
	fclose(fp);
	if (rank == 1) return data;
	if (rank == 2) {
		float **f = calloc(d[0], sizeof(float*));
		for (int i = 0; i < d[0]; i++) f[i] = data + (long)d[1]*i;
		return f;
	}
	float ***f = calloc(d[0], sizeof(float**));
	f[0] = calloc((long)d[0]*d[1], sizeof(float*));
	for (int i = 0; i < d[0]; i++) {
		f[i] = f[0] + (long)d[1]*i;
		for (int j = 0; j < d[1]; j++) f[i][j] = data + ((long)d[1]*i + j)*d[2];
	}
	return f;
}

void h5write(int *dim, float *p, char *format, ...)
{
	char name[1024];
	int rank = 0, d[3];
	va_list ap;
	va_start(ap, format);
	vsnprintf(name, sizeof name - 4, format, ap);
	va_end(ap);
	strcat(name, ".h5");
	long total = 1;
	for (int n = 0; n < 3; n++) if (dim[n]) d[rank++] = dim[n], total *= dim[n];
	FILE *fp = fopen(name, "wb");
	if (!fp) return;
	fwrite("ALISARR1", 1, 8, fp);
	fwrite(&rank, 4, 1, fp);
	fwrite(d, 4, rank, fp);
	fwrite(p, sizeof(float), total, fp);
	fclose(fp);
}
