/*
 * World construction: grid extents, boundary codes, Bloch phases, per-axis step and CPML
 * coefficients, field storage.  Follows src/world.c:12-466 of the reference; the
 * arithmetic (including its float/double promotions) is kept identical so that every
 * coefficient array is bitwise equal to the reference's.
 */
#include <string.h>
#include "internal.h"

/* Node index nearest to coordinate x for a monotone node->coordinate map
 * (src/world.c:12-25): 16 halving steps from 0, then a +-1 nearest-neighbour check. */
static int nearestNode(world W, float (*coord)(world, float), float x)
{
	int i = 0;
	for (int b = 15; b >= 0; b--) {
		float c = coord(W, i);
		if (x > c) i += 1u << b;
		if (x < c) i -= 1u << b;
	}
	float dist = fabsf(x - coord(W, i));
	if (fabsf(x - coord(W, i + 1)) < dist) i++;
	if (fabsf(x - coord(W, i - 1)) < dist) i--;
	return i;
}

/* Graded (non-uniform) node coordinate (src/world.c:27-42): uniform spacing ds inside
 * [lo, hi], geometric growth with ratio r outside, reaching spacing m*ds after n cells.
 * r solves r^n - c*ln r - 1 = 0 with c = 10m - m/2 - 1/2 (Newton iteration). */
static float gradedCoord(int m, float lo, float hi, float ds, float i)
{
	float d = 10;
	float iLo = lo/ds-0.5, iHi = hi/ds+0.5;
	float r = d/(d-1), rPrev;
	int n = floor((d*m-0.5*m-0.5)*logf(m)/(m-1)+0.5);

	do rPrev = r, r = rPrev - (powf(rPrev,n)-(d*m-0.5*m-0.5)*logf(rPrev)-1)/(n*powf(rPrev,n-1)-(d*m-0.5*m-0.5)/rPrev);
	while (fabsf(r-rPrev) > 1E-6);

	if (i>=iHi+n) return ds * (iHi + (d*m-0.5*m-0.5) + (i-iHi-n) * m);
	if (i<=iLo-n) return ds * (iLo - (d*m-0.5*m-0.5) + (i-iLo+n) * m);
	if (i>iHi) return ds * (iHi + (powf(r,i-iHi)-1) / logf(r));
	if (i<iLo) return ds * (iLo - (powf(r,iLo-i)-1) / logf(r));
	return ds * i;
}

static int xtoi(world W, float x) { return nearestNode(W, W->itox, x); }
static int ytoj(world W, float y) { return nearestNode(W, W->jtoy, y); }
static int ztok(world W, float z) { return nearestNode(W, W->ktoz, z); }
static float itoxUniform(world W, float i) { return W->dx*i; }
static float jtoyUniform(world W, float j) { return W->dy*j; }
static float ktozUniform(world W, float k) { return W->dz*k; }
static float itoxGraded(world W, float i) { return gradedCoord(W->Res.nug.x.m, W->Res.nug.x.Min, W->Res.nug.x.Max, W->dx, i); }
static float jtoyGraded(world W, float j) { return gradedCoord(W->Res.nug.y.m, W->Res.nug.y.Min, W->Res.nug.y.Max, W->dy, j); }
static float ktozGraded(world W, float k) { return gradedCoord(W->Res.nug.z.m, W->Res.nug.z.Min, W->Res.nug.z.Max, W->dz, k); }


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


/* CPML profile on one axis (src/world.c:98-132).  For n = 1..Num cells into the layer
 * (half-cell shifted by Offset for the H grid): polynomial sigma (order SigmaPower,
 * scaled 0.8*(m+1)*SigmaMax), kappa, and linear alpha; the step ratio dtds is divided by
 * kappa and the recursion coefficients b = Ca and c = -+dtds*Cb are stored. */
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
	for (int n = 1; n <= S.pml.Num; n++) {
		float Sigma = 0.8*S.pml.SigmaMax*(S.pml.SigmaPower+1)*powf((n-0.5*Offset)/S.pml.Num, S.pml.SigmaPower);
		float Kappa = 1+(S.pml.KappaMax-1)*powf((n-0.5*Offset)/S.pml.Num, S.pml.SigmaPower);
		float Alpha = S.pml.AlphaMax*powf((S.pml.Num-(n-0.5*Offset))/S.pml.Num, S.pml.AlphaPower);
		int at[2] = { Min-n-1+Offset, Max+n+1 };
		int on[2] = { MinSur==PML, MaxSur==PML };
		for (int s = 0; s < 2; s++) {
			if (!on[s]) continue;
			int q = at[s];
			dtds[q] /= Kappa;
			float Ca = (1-0.5*(Sigma*dtds[q]+Alpha*dt))/(1+0.5*(Sigma*dtds[q]+Alpha*dt));
			float Cb = Sigma*dtds[q]/(1+0.5*(Sigma*dtds[q]+Alpha*dt));
			(*b)[q] = Ca;
			(*c1)[q] = -dtds[q]*Cb;
			(*c2)[q] = +dtds[q]*Cb;
		}
	}
}


/* Cell widths and dt/width ratios of the E (Offset 0) or H (Offset 1) grid, with the
 * CPML profile folded in (src/world.c:135-161). */
static coeffs *makeStepCoeffs(world W, sur S, int Offset)
{
	coeffs *C;
	MALLOC(C, 0, 1);
	S.pml.Num = S.pml.Num ? S.pml.Num : 12;

	MALLOC(C->dx, W->xMinSur!=SYM ? W->iMIN : -W->iMAX, W->xMinSur!=SYM ? W->iNUM : 2*W->iNUM-1);
	MALLOC(C->dy, W->yMinSur!=SYM ? W->jMIN : -W->jMAX, W->yMinSur!=SYM ? W->jNUM : 2*W->jNUM-1);
	MALLOC(C->dz, W->zMinSur!=SYM ? W->kMIN : -W->kMAX, W->zMinSur!=SYM ? W->kNUM : 2*W->kNUM-1);
	MALLOC(C->dtdx, W->iMIN, W->iNUM);
	MALLOC(C->dtdy, W->jMIN, W->jNUM);
	MALLOC(C->dtdz, W->kMIN, W->kNUM);

	for (int i = W->iMIN; i <= W->iMAX; i++) {
		C->dx[i] = W->itox(W, i-0.5*(Offset-1)) - W->itox(W, i-0.5*(Offset+1));
		C->dtdx[i] = W->dt / C->dx[i];
	}
	for (int j = W->jMIN; j <= W->jMAX; j++) {
		C->dy[j] = W->jtoy(W, j-0.5*(Offset-1)) - W->jtoy(W, j-0.5*(Offset+1));
		C->dtdy[j] = W->dt / C->dy[j];
	}
	for (int k = W->kMIN; k <= W->kMAX; k++) {
		C->dz[k] = W->ktoz(W, k-0.5*(Offset-1)) - W->ktoz(W, k-0.5*(Offset+1));
		C->dtdz[k] = W->dt / C->dz[k];
	}
	if (W->xMinSur==SYM) for (int i = 1; i <= W->iMAX; i++) C->dx[-i] = C->dx[i];
	if (W->yMinSur==SYM) for (int j = 1; j <= W->jMAX; j++) C->dy[-j] = C->dy[j];
	if (W->zMinSur==SYM) for (int k = 1; k <= W->kMAX; k++) C->dz[-k] = C->dz[k];

	cpmlProfile(S, Offset, &C->xP, &C->xPy, &C->xPz, C->dtdx, W->dt, W->xMinSur, W->xMaxSur, W->iMin, W->iMax, W->iMIN, W->iNUM);
	cpmlProfile(S, Offset, &C->yP, &C->yPz, &C->yPx, C->dtdy, W->dt, W->yMinSur, W->yMaxSur, W->jMin, W->jMax, W->jMIN, W->jNUM);
	cpmlProfile(S, Offset, &C->zP, &C->zPx, &C->zPy, C->dtdz, W->dt, W->zMinSur, W->zMaxSur, W->kMin, W->kMax, W->kMIN, W->kNUM);
	C->N = 0;
	return C;
}


/* Field components with one ghost layer on each side and the CPML auxiliary slabs of
 * every PML face; the E grid uses Offset 0 and the H grid Offset 1 (src/world.c:164-189). */
static vfield *makeFieldSet(world W, int Off)
{
	vfield *F;
	MALLOC(F, 0, 1+W->complexField);
	for (int m = 0; m <= W->complexField; m++) {
		F[m].x = makeField(W->iMIN-1, W->iMAX+1, W->jMIN-1, W->jMAX+1, W->kMIN-1, W->kMAX+1);
		F[m].y = makeField(W->iMIN-1, W->iMAX+1, W->jMIN-1, W->jMAX+1, W->kMIN-1, W->kMAX+1);
		F[m].z = makeField(W->iMIN-1, W->iMAX+1, W->jMIN-1, W->jMAX+1, W->kMIN-1, W->kMAX+1);
		if (W->xMinSur==PML) {
			F[m].xMinPy = makeField(W->iMIN+Off, W->iMin+Off-2, W->jMIN, W->jMAX, W->kMIN, W->kMAX);
			F[m].xMinPz = makeField(W->iMIN+Off, W->iMin+Off-2, W->jMIN, W->jMAX, W->kMIN, W->kMAX);
		}
		if (W->yMinSur==PML) {
			F[m].yMinPz = makeField(W->iMIN, W->iMAX, W->jMIN+Off, W->jMin+Off-2, W->kMIN, W->kMAX);
			F[m].yMinPx = makeField(W->iMIN, W->iMAX, W->jMIN+Off, W->jMin+Off-2, W->kMIN, W->kMAX);
		}
		if (W->zMinSur==PML) {
			F[m].zMinPx = makeField(W->iMIN, W->iMAX, W->jMIN, W->jMAX, W->kMIN+Off, W->kMin+Off-2);
			F[m].zMinPy = makeField(W->iMIN, W->iMAX, W->jMIN, W->jMAX, W->kMIN+Off, W->kMin+Off-2);
		}
		if (W->xMaxSur==PML) {
			F[m].xMaxPy = makeField(W->iMax+2, W->iMAX, W->jMIN, W->jMAX, W->kMIN, W->kMAX);
			F[m].xMaxPz = makeField(W->iMax+2, W->iMAX, W->jMIN, W->jMAX, W->kMIN, W->kMAX);
		}
		if (W->yMaxSur==PML) {
			F[m].yMaxPz = makeField(W->iMIN, W->iMAX, W->jMax+2, W->jMAX, W->kMIN, W->kMAX);
			F[m].yMaxPx = makeField(W->iMIN, W->iMAX, W->jMax+2, W->jMAX, W->kMIN, W->kMAX);
		}
		if (W->zMaxSur==PML) {
			F[m].zMaxPx = makeField(W->iMIN, W->iMAX, W->jMIN, W->jMAX, W->kMax+2, W->kMAX);
			F[m].zMaxPy = makeField(W->iMIN, W->iMAX, W->jMIN, W->jMAX, W->kMax+2, W->kMAX);
		}
	}
	return F;
}


/* Grid, surfaces and Bloch phases of a new world (src/world.c:192-258). */
static world makeBareWorld(res R, dom D, sur S)
{
	world W;
	MALLOC(W, 0, 1);

	if (R.nug.x.Min>0) R.nug.x.Max = R.nug.x.Min, R.nug.x.Min = -R.nug.x.Max;
	if (R.nug.y.Min>0) R.nug.y.Max = R.nug.y.Min, R.nug.y.Min = -R.nug.y.Max;
	if (R.nug.z.Min>0) R.nug.z.Max = R.nug.z.Min, R.nug.z.Min = -R.nug.z.Max;
	if (S.x.MinSur==PBC && S.x.MaxSur) S.x.MinSur = BBC;
	if (S.y.MinSur==PBC && S.y.MaxSur) S.y.MinSur = BBC;
	if (S.z.MinSur==PBC && S.z.MaxSur) S.z.MinSur = BBC;
	if (!S.pml.Num) S.pml.Num = 12;

	/* Dom keeps the domain as given (the 2-D tests compare Dom.Min with Dom.Max). */
	W->Dom = D, W->Res = R, W->Sur = S;
	W->t = 0, W->dt = R.dt;
	W->dx = R.d.x, W->dy = R.d.y ? R.d.y : R.d.x, W->dz = R.d.z ? R.d.z : R.d.x;
	W->eV = R.spec.eV ? R.spec.eV : 1;
	W->f = INF;
	if (!D.x.Max && D.x.Min>0 || S.x.MinSur==SYM) D.x.Min = -(D.x.Max = D.x.Max ? D.x.Max : D.x.Min);
	if (!D.y.Max && D.y.Min>0 || S.y.MinSur==SYM) D.y.Min = -(D.y.Max = D.y.Max ? D.y.Max : D.y.Min);
	if (!D.z.Max && D.z.Min>0 || S.z.MinSur==SYM) D.z.Min = -(D.z.Max = D.z.Max ? D.z.Max : D.z.Min);

	W->itox = R.gridFunction.itox ? R.gridFunction.itox : R.nug.x.m > 1 ? itoxGraded : itoxUniform;
	W->jtoy = R.gridFunction.jtoy ? R.gridFunction.jtoy : R.nug.y.m > 1 ? jtoyGraded : jtoyUniform;
	W->ktoz = R.gridFunction.ktoz ? R.gridFunction.ktoz : R.nug.z.m > 1 ? ktozGraded : ktozUniform;
	W->xtoi = xtoi, W->ytoj = ytoj, W->ztok = ztok;

	/* Unset surfaces inherit from the x minimum side; a max side copies its min side
	 * unless a distinct code is given (periods of PBC/BBC are not codes). */
	W->xMinSur = S.x.MinSur;
	W->yMinSur = S.y.MinSur ? S.y.MinSur : S.x.MinSur;
	W->zMinSur = S.z.MinSur ? S.z.MinSur : S.x.MinSur;
	W->xMaxSur = S.x.MinSur!=BBC && S.x.MaxSur ? S.x.MaxSur : S.x.MinSur;
	W->yMaxSur = S.y.MinSur!=BBC && S.y.MaxSur ? S.y.MaxSur : S.y.MinSur ? S.y.MinSur : S.x.MinSur;
	W->zMaxSur = S.z.MinSur!=BBC && S.z.MaxSur ? S.z.MaxSur : S.z.MinSur ? S.z.MinSur : S.x.MinSur;

	W->iMin = W->xMinSur==SYM ? 0 : D.x.Min || D.x.Max ? W->xtoi(W, D.x.Min) : -1;
	W->jMin = W->yMinSur==SYM ? 0 : D.y.Min || D.y.Max ? W->ytoj(W, D.y.Min) : -1;
	W->kMin = W->zMinSur==SYM ? 0 : D.z.Min || D.z.Max ? W->ztok(W, D.z.Min) : -1;
	W->iMax = D.x.Min || D.x.Max ? W->xtoi(W, D.x.Max) : 1, W->iNum = 1-W->iMin+W->iMax;
	W->jMax = D.y.Min || D.y.Max ? W->ytoj(W, D.y.Max) : 1, W->jNum = 1-W->jMin+W->jMax;
	W->kMax = D.z.Min || D.z.Max ? W->ztok(W, D.z.Max) : 1, W->kNum = 1-W->kMin+W->kMax;

	W->iMIN = W->iMin-(W->xMinSur==PML ? S.pml.Num+1 : 0);
	W->jMIN = W->jMin-(W->yMinSur==PML ? S.pml.Num+1 : 0);
	W->kMIN = W->kMin-(W->zMinSur==PML ? S.pml.Num+1 : 0);
	W->iMAX = W->iMax+(W->xMaxSur==PML ? S.pml.Num+1 : 0), W->iNUM = 1-W->iMIN+W->iMAX;
	W->jMAX = W->jMax+(W->yMaxSur==PML ? S.pml.Num+1 : 0), W->jNUM = 1-W->jMIN+W->jMAX;
	W->kMAX = W->kMax+(W->zMaxSur==PML ? S.pml.Num+1 : 0), W->kNUM = 1-W->kMIN+W->kMAX;

	W->xMax = W->itox(W, W->iMax), W->xMin = W->xMinSur==SYM ? -W->xMax : W->itox(W, W->iMin), W->xSize = W->xMax - W->xMin;
	W->yMax = W->jtoy(W, W->jMax), W->yMin = W->yMinSur==SYM ? -W->yMax : W->jtoy(W, W->jMin), W->ySize = W->yMax - W->yMin;
	W->zMax = W->ktoz(W, W->kMax), W->zMin = W->zMinSur==SYM ? -W->zMax : W->ktoz(W, W->kMin), W->zSize = W->zMax - W->zMin;
	W->xMAX = W->itox(W, W->iMAX), W->xMIN = W->xMinSur==SYM ? -W->xMAX : W->itox(W, W->iMIN);
	W->yMAX = W->jtoy(W, W->jMAX), W->yMIN = W->yMinSur==SYM ? -W->yMAX : W->jtoy(W, W->jMIN);
	W->zMAX = W->ktoz(W, W->kMAX), W->zMIN = W->zMinSur==SYM ? -W->zMAX : W->ktoz(W, W->kMIN);

	/* Bloch phase across one period; 0 marks a symmetry plane whose parity a source sets. */
	W->coskx = S.x.MinSur==BBC && S.x.MaxSur ? cosf(2*PI*(W->xMax-W->xMin)/S.x.MaxSur) : S.x.MinSur==SYM ? 0 : 1;
	W->cosky = S.y.MinSur==BBC && S.y.MaxSur ? cosf(2*PI*(W->yMax-W->yMin)/S.y.MaxSur) : S.y.MinSur==SYM ? 0 : 1;
	W->coskz = S.z.MinSur==BBC && S.z.MaxSur ? cosf(2*PI*(W->zMax-W->zMin)/S.z.MaxSur) : S.z.MinSur==SYM ? 0 : 1;
	W->sinkx = S.x.MinSur==BBC && S.x.MaxSur ? sinf(2*PI*(W->xMax-W->xMin)/S.x.MaxSur) : 0;
	W->sinky = S.y.MinSur==BBC && S.y.MaxSur ? sinf(2*PI*(W->yMax-W->yMin)/S.y.MaxSur) : 0;
	W->sinkz = S.z.MinSur==BBC && S.z.MaxSur ? sinf(2*PI*(W->zMax-W->zMin)/S.z.MaxSur) : 0;

	/* Hexagonal cells: phases of the two diagonal neighbours. */
	W->coskyp = S.y.MinSur==HBC && S.y.MaxSur ? cosf(PI*(2*(W->yMax-W->yMin)/S.y.MaxSur+(W->xMax-W->xMin)/S.x.MaxSur)) : 1;
	W->coskym = S.y.MinSur==HBC && S.y.MaxSur ? cosf(PI*(2*(W->yMax-W->yMin)/S.y.MaxSur-(W->xMax-W->xMin)/S.x.MaxSur)) : 1;
	W->sinkyp = S.y.MinSur==HBC && S.y.MaxSur ? sinf(PI*(2*(W->yMax-W->yMin)/S.y.MaxSur+(W->xMax-W->xMin)/S.x.MaxSur)) : 0;
	W->sinkym = S.y.MinSur==HBC && S.y.MaxSur ? sinf(PI*(2*(W->yMax-W->yMin)/S.y.MaxSur-(W->xMax-W->xMin)/S.x.MaxSur)) : 0;

	W->xyArea = W->xSize * W->ySize, W->xzArea = W->xSize * W->zSize, W->yzArea = W->ySize * W->zSize;
	if (S.x.MinSur==BBC || S.y.MinSur==BBC || S.z.MinSur==BBC) W->complexField = 1;

	alisPrecisionDefault(&W->devPrec, &W->devDftPrec);
	return W;
}


world createWorld(dom D, res R, sur S, char *format, ...)
{
	va_list ap;
	world W = makeBareWorld(R, D, S);

	W->E = makeFieldSet(W, 0);
	W->H = makeFieldSet(W, 1);
	W->CE = makeStepCoeffs(W, S, 0);
	W->CH = makeStepCoeffs(W, S, 1);

	va_start(ap, format);
	vsnprintf(W->ID, sizeof W->ID, format, ap);
	va_end(ap);
	if (W->ID[0]) {
		/* Output files of a previous run with the same ID are removed (src/world.c:275-278). */
		printf("Deleting previous files.");
		exec("rm -fr %s %s.c?* %s[#--0-~]* %s.[#-bd-~]* 2> /dev/null", W->ID, W->ID, W->ID, W->ID);
	}
	alisRegisterWorld(W);
	return W;
}


/* Polarisation-current arrays J (all poles) and K (Lorentz/critical-point poles) of every
 * dispersive object over its bounding box (src/world.c:283-315). */
void makeAuxFieldArrays(world W, vfield *F, coeffs *C)
{
	for (int m = 0; m <= W->complexField; m++) {
		MALLOC(F[m].Jx, 0, C->N); MALLOC(F[m].Jy, 0, C->N); MALLOC(F[m].Jz, 0, C->N);
		MALLOC(F[m].Kx, 0, C->N); MALLOC(F[m].Ky, 0, C->N); MALLOC(F[m].Kz, 0, C->N);
	}
	for (int n = 0; n < C->N; n++) {
		for (int m = 0; m <= W->complexField; m++) {
			MALLOC(F[m].Jx[n], 0, C->NK[n]); MALLOC(F[m].Jy[n], 0, C->NK[n]); MALLOC(F[m].Jz[n], 0, C->NK[n]);
			MALLOC(F[m].Kx[n], C->NJ[n], C->NK[n]-C->NJ[n]);
			MALLOC(F[m].Ky[n], C->NJ[n], C->NK[n]-C->NJ[n]);
			MALLOC(F[m].Kz[n], C->NJ[n], C->NK[n]-C->NJ[n]);
			for (int p = 0; p < C->NK[n]; p++) {
				F[m].Jx[n][p] = makeField(C->iMin[n], C->iMax[n], C->jMin[n], C->jMax[n], C->kMin[n], C->kMax[n]);
				F[m].Jy[n][p] = makeField(C->iMin[n], C->iMax[n], C->jMin[n], C->jMax[n], C->kMin[n], C->kMax[n]);
				F[m].Jz[n][p] = makeField(C->iMin[n], C->iMax[n], C->jMin[n], C->jMax[n], C->kMin[n], C->kMax[n]);
			}
			for (int p = C->NJ[n]; p < C->NK[n]; p++) {
				F[m].Kx[n][p] = makeField(C->iMin[n], C->iMax[n], C->jMin[n], C->jMax[n], C->kMin[n], C->kMax[n]);
				F[m].Ky[n][p] = makeField(C->iMin[n], C->iMax[n], C->jMin[n], C->jMax[n], C->kMin[n], C->kMax[n]);
				F[m].Kz[n][p] = makeField(C->iMin[n], C->iMax[n], C->jMin[n], C->jMax[n], C->kMin[n], C->kMax[n]);
			}
		}
	}
}


/* A world sharing geometry and coefficients with SW but with its own fields
 * (src/world.c:318-331). */
world cloneWorld(world SW)
{
	world W;
	MALLOC(W, 0, 1);
	*W = *SW;
	W->dev = 0, W->hostStale = 0, W->parent = 0;
	W->E = makeFieldSet(W, 0);
	W->H = makeFieldSet(W, 1);
	makeAuxFieldArrays(W, W->E, W->CE);
	makeAuxFieldArrays(W, W->H, W->CH);
	alisRegisterWorld(W);
	return W;
}


static void zeroFieldSet(world W, vfield *F, coeffs *C, int Off)
{
	for (int m = 0; m <= W->complexField; m++) {
		resetField3(F[m].x, W->iMIN-1, W->iMAX+1, W->jMIN-1, W->jMAX+1, W->kMIN-1, W->kMAX+1);
		resetField3(F[m].y, W->iMIN-1, W->iMAX+1, W->jMIN-1, W->jMAX+1, W->kMIN-1, W->kMAX+1);
		resetField3(F[m].z, W->iMIN-1, W->iMAX+1, W->jMIN-1, W->jMAX+1, W->kMIN-1, W->kMAX+1);
		if (F[m].xMinPy) resetField3(F[m].xMinPy, W->iMIN+Off, W->iMin+Off-2, W->jMIN, W->jMAX, W->kMIN, W->kMAX);
		if (F[m].xMinPz) resetField3(F[m].xMinPz, W->iMIN+Off, W->iMin+Off-2, W->jMIN, W->jMAX, W->kMIN, W->kMAX);
		if (F[m].yMinPz) resetField3(F[m].yMinPz, W->iMIN, W->iMAX, W->jMIN+Off, W->jMin+Off-2, W->kMIN, W->kMAX);
		if (F[m].yMinPx) resetField3(F[m].yMinPx, W->iMIN, W->iMAX, W->jMIN+Off, W->jMin+Off-2, W->kMIN, W->kMAX);
		if (F[m].zMinPx) resetField3(F[m].zMinPx, W->iMIN, W->iMAX, W->jMIN, W->jMAX, W->kMIN+Off, W->kMin+Off-2);
		if (F[m].zMinPy) resetField3(F[m].zMinPy, W->iMIN, W->iMAX, W->jMIN, W->jMAX, W->kMIN+Off, W->kMin+Off-2);
		if (F[m].xMaxPy) resetField3(F[m].xMaxPy, W->iMax+2, W->iMAX, W->jMIN, W->jMAX, W->kMIN, W->kMAX);
		if (F[m].xMaxPz) resetField3(F[m].xMaxPz, W->iMax+2, W->iMAX, W->jMIN, W->jMAX, W->kMIN, W->kMAX);
		if (F[m].yMaxPz) resetField3(F[m].yMaxPz, W->iMIN, W->iMAX, W->jMax+2, W->jMAX, W->kMIN, W->kMAX);
		if (F[m].yMaxPx) resetField3(F[m].yMaxPx, W->iMIN, W->iMAX, W->jMax+2, W->jMAX, W->kMIN, W->kMAX);
		if (F[m].zMaxPx) resetField3(F[m].zMaxPx, W->iMIN, W->iMAX, W->jMIN, W->jMAX, W->kMax+2, W->kMAX);
		if (F[m].zMaxPy) resetField3(F[m].zMaxPy, W->iMIN, W->iMAX, W->jMIN, W->jMAX, W->kMax+2, W->kMAX);
		for (int n = 0; n < C->N; n++) {
			for (int p = 0; p < C->NK[n]; p++) {
				resetField3(F[m].Jx[n][p], C->iMin[n], C->iMax[n], C->jMin[n], C->jMax[n], C->kMin[n], C->kMax[n]);
				resetField3(F[m].Jy[n][p], C->iMin[n], C->iMax[n], C->jMin[n], C->jMax[n], C->kMin[n], C->kMax[n]);
				resetField3(F[m].Jz[n][p], C->iMin[n], C->iMax[n], C->jMin[n], C->jMax[n], C->kMin[n], C->kMax[n]);
			}
			for (int p = C->NJ[n]; p < C->NK[n]; p++) {
				resetField3(F[m].Kx[n][p], C->iMin[n], C->iMax[n], C->jMin[n], C->jMax[n], C->kMin[n], C->kMax[n]);
				resetField3(F[m].Ky[n][p], C->iMin[n], C->iMax[n], C->jMin[n], C->jMax[n], C->kMin[n], C->kMax[n]);
				resetField3(F[m].Kz[n][p], C->iMin[n], C->iMax[n], C->jMin[n], C->jMax[n], C->kMin[n], C->kMax[n]);
			}
		}
	}
}


/* Zero all fields and sources and restart the clock (src/world.c:371-380). */
void rebootWorld(world W)
{
	alisInvalidate(W);
	removeSources(W);
	zeroFieldSet(W, W->E, W->CE, 0);
	zeroFieldSet(W, W->H, W->CH, 1);
	W->t = W->E->t = 0;
	W->H->t = 0.5;
	W->n = 0;
	printf("\n");
}


static void freeAuxFieldArrays(world W, vfield *F, coeffs *C)
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
	}
}


static void freeFieldSet(world W, vfield *F, int Off)
{
	for (int m = 0; m <= W->complexField; m++) {
		if (!W->isAux || !(W->expandX || W->expandY)) {
			removeField(F[m].x, W->iMIN-1, W->jMIN-1, W->kMIN-1);
			removeField(F[m].y, W->iMIN-1, W->jMIN-1, W->kMIN-1);
			removeField(F[m].z, W->iMIN-1, W->jMIN-1, W->kMIN-1);
		}
		if (F[m].xMinPy) removeField(F[m].xMinPy, W->iMIN+Off, W->jMIN, W->kMIN);
		if (F[m].xMinPz) removeField(F[m].xMinPz, W->iMIN+Off, W->jMIN, W->kMIN);
		if (F[m].yMinPz) removeField(F[m].yMinPz, W->iMIN, W->jMIN+Off, W->kMIN);
		if (F[m].yMinPx) removeField(F[m].yMinPx, W->iMIN, W->jMIN+Off, W->kMIN);
		if (F[m].zMinPx) removeField(F[m].zMinPx, W->iMIN, W->jMIN, W->kMIN+Off);
		if (F[m].zMinPy) removeField(F[m].zMinPy, W->iMIN, W->jMIN, W->kMIN+Off);
		if (F[m].xMaxPy) removeField(F[m].xMaxPy, W->iMax+2, W->jMIN, W->kMIN);
		if (F[m].xMaxPz) removeField(F[m].xMaxPz, W->iMax+2, W->jMIN, W->kMIN);
		if (F[m].yMaxPz) removeField(F[m].yMaxPz, W->iMIN, W->jMax+2, W->kMIN);
		if (F[m].yMaxPx) removeField(F[m].yMaxPx, W->iMIN, W->jMax+2, W->kMIN);
		if (F[m].zMaxPx) removeField(F[m].zMaxPx, W->iMIN, W->jMIN, W->kMax+2);
		if (F[m].zMaxPy) removeField(F[m].zMaxPy, W->iMIN, W->jMIN, W->kMax+2);
	}
	free(F);
}


static void freeStepCoeffs(world W, coeffs *C)
{
	if (C->xP) free(&C->xP[W->iMIN]), free(&C->xPy[W->iMIN]), free(&C->xPz[W->iMIN]);
	if (C->yP) free(&C->yP[W->jMIN]), free(&C->yPz[W->jMIN]), free(&C->yPx[W->jMIN]);
	if (C->zP) free(&C->zP[W->kMIN]), free(&C->zPx[W->kMIN]), free(&C->zPy[W->kMIN]);
	free(&C->dtdx[W->iMIN]), free(&C->dx[W->xMinSur!=SYM ? W->iMIN : -W->iMAX]);
	free(&C->dtdy[W->jMIN]), free(&C->dy[W->yMinSur!=SYM ? W->jMIN : -W->jMAX]);
	free(&C->dtdz[W->kMIN]), free(&C->dz[W->zMinSur!=SYM ? W->kMIN : -W->kMAX]);
	free(C);
}


void deleteWorld(world W)
{
	alisForgetWorld(W);
	removeSources(W);
	removeObjects(W);
	freeStepCoeffs(W, W->CE);
	freeStepCoeffs(W, W->CH);
	freeFieldSet(W, W->E, 0);
	freeFieldSet(W, W->H, 1);
	free(W);
}


void deleteCloneWorld(world W)
{
	alisForgetWorld(W);
	removeSources(W);
	removeObjects(W);
	freeAuxFieldArrays(W, W->E, W->CE);
	freeAuxFieldArrays(W, W->H, W->CH);
	freeFieldSet(W, W->E, 0);
	freeFieldSet(W, W->H, 1);
	free(W);
}


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
