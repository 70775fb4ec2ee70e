/*
 * Sources: waveforms, point dipoles, guided-mode planes, TF/SF plane waves (with their
 * auxiliary 1-D/2-D worlds) and Gaussian beams.  Follows src/source.c:11-524; waveform and
 * amplitude arithmetic is kept identical to the reference's (float with its double
 * promotions) because the injected values must agree bitwise.
 */
#include <string.h>
#include "internal.h"

/* Sine ramped up over `supplement` periods by a raised cosine (src/source.c:11-23). */
float Sine(world W, source S, float t, float phase)
{
	if (!S->N) {
		S->frequency = 2 * PI / S->wavelength;
		S->peaktime = W->dt * floorf(S->wavelength * S->supplement / W->dt);
		S->deviation = S->peaktime / PI;
		S->N = -1;
	}
	if (t < S->peaktime)
		return sinf(S->frequency * (t-S->peaktime) - phase) * (0.5 * (1 - cosf(t / S->deviation)));
	return sinf(S->frequency * (t-S->peaktime) - phase);
}

/* Gaussian-modulated sine of bandwidth `supplement` (nm) around `wavelength`
 * (src/source.c:26-35). */
float Pulse(world W, source S, float t, float phase)
{
	if (!S->N) {
		S->frequency = 2 * PI / S->wavelength;
		S->deviation = 2 * sq(S->wavelength) / (PI * S->supplement);
		S->peaktime = W->dt * floorf(PI * S->deviation / W->dt);
		S->N = 2 * floorf(PI * S->deviation / W->dt);
	}
	return sinf(S->frequency * (t-S->peaktime) - phase) * expf(-sq((t-S->peaktime) / S->deviation));
}

/* Flat band between the wavelengths `wavelength` (shortest) and `supplement`: difference
 * of two sinc-like cosines (src/source.c:38-49). */
float Band(world W, source S, float t, float phase)
{
	if (!S->N) {
		S->resolution = S->supplement > S->wavelength ? ceilf(1.25 * S->supplement / S->wavelength + 0.5) : 2;
		S->frequency = 2 * PI / S->wavelength;
		S->deviation = 2 * PI / S->wavelength / (S->resolution - 0.5);
		S->peaktime = floorf(S->wavelength * (S->resolution - 0.5) / W->dt) * W->dt;
		S->N = 2 * S->wavelength * (S->resolution - 0.5) / W->dt;
	}
	if (t == S->peaktime) return 0;
	return 2 * (cosf((t-S->peaktime) * S->deviation * S->resolution) - cosf((t-S->peaktime) * S->deviation)) / ((t-S->peaktime) * S->deviation);
}


/* Array targeted by field function F in world W, or 0 (src/source.c:59-69). */
static float ***targetArray(world W, field *F, int *isH)
{
	int h, m, a;
	if (!fieldComponent(F, &h, &m, &a)) return 0;
	if (m > W->complexField) return 0;
	vfield *V = h ? &W->H[m] : &W->E[m];
	*isH = h;
	return a == 0 ? V->x : a == 1 ? V->y : V->z;
}

/* New source appended to the E or H chain of W; initialises the waveform parameters and
 * the world's source length / period (src/source.c:52-75). */
static source createSource(world W, field F, waveform waveForm, float wavelength, float supplement, float resolution)
{
	source S;
	MALLOC(S, 0, 1);
	S->waveForm = waveForm, S->wavelength = wavelength, S->supplement = supplement, S->resolution = resolution;
	waveForm(W, S, 0, 0);

	int isH = 0;
	float ***target = targetArray(W, F, &isH);
	if (target) {
		source *chain = isH ? &W->SH : &W->SE;
		while (*chain) chain = &(*chain)->S;
		*chain = S;
		S->F = target;
	}
	if (W->N < S->N && S->N != -1) W->N = S->N;
	if (wavelength && W->T != wavelength/W->dt) W->f = 1/wavelength, W->T = wavelength/W->dt;
	if (!W->sourceType) W->sourceType = 1;
	return S;
}

static void setAmplitude(source S, float *dtdx, float *dtdy, float *dtdz, float ***C, float amplitude)
{
	for (int i = S->iMin; i <= S->iMax; i++) for (int j = S->jMin; j <= S->jMax; j++) for (int k = S->kMin; k <= S->kMax; k++)
		S->Amplitude[i][j][k] = amplitude * dtdx[i] * dtdy[j] * dtdz[k] * (C ? C[i][j][k] : 1);
}

static int isFx(field *F) { return F==Ex || F==Hy || F==Hz || F==iEx || F==iHy || F==iHz; }
static int isFy(field *F) { return F==Ey || F==Hz || F==Hx || F==iEy || F==iHz || F==iHx; }
static int isFz(field *F) { return F==Ez || F==Hx || F==Hy || F==iEz || F==iHx || F==iHy; }

/* Parity of the symmetry planes implied by a source of field F placed on them. */
static void dipoleParity(world W, field F, int i, int j, int k)
{
	if (W->xMinSur==SYM && !i) W->coskx = isFx(F) ? 1 : -1;
	if (W->yMinSur==SYM && !j) W->cosky = isFy(F) ? 1 : -1;
	if (W->zMinSur==SYM && !k) W->coskz = isFz(F) ? 1 : -1;
	if (W->xMinSur==SYM && !W->coskx) W->coskx = isFx(F) ? -1 : 1;
	if (W->yMinSur==SYM && !W->cosky) W->cosky = isFy(F) ? -1 : 1;
	if (W->zMinSur==SYM && !W->coskz) W->coskz = isFz(F) ? -1 : 1;
}

/* Point dipole on the two grid edges around (x,y,z) along the component of F, normalised
 * to unit radiated power per unit bandwidth (src/source.c:87-124). */
void (pointDipole)(world W, field F, float x, float y, float z, waveform waveForm, float wavelength, float supplement, float amplitude, float phase, ...)
{
	alisInvalidate(W);
	int i = W->xtoi(W, x), j = W->ytoj(W, y), k = W->ztok(W, z);
	dipoleParity(W, F, i, j, k);

	if (waveForm) {
		if (!W->sourceType)
			W->sourceType = (W->Dom.x.Min==W->Dom.x.Max || W->Dom.y.Min==W->Dom.y.Max || W->Dom.z.Min==W->Dom.z.Max) ? 2 : 3;
		source S = createSource(W, F, waveForm, wavelength, supplement, amplitude);
		S->iMin = S->iMax = i, S->jMin = S->jMax = j, S->kMin = S->kMax = k;
		if (isFx(F)) S->iMax++;
		if (isFy(F)) S->jMax++;
		if (isFz(F)) S->kMax++;

		if (!amplitude) amplitude = 1;
		if      (W->Dom.x.Min==W->Dom.x.Max) amplitude *= W->dx * sqrtf(2.5) * sqrtf(wavelength) / sq(W->dt);
		else if (W->Dom.y.Min==W->Dom.y.Max) amplitude *= W->dy * sqrtf(2.5) * sqrtf(wavelength) / sq(W->dt);
		else if (W->Dom.z.Min==W->Dom.z.Max) amplitude *= W->dz * sqrtf(2.5) * sqrtf(wavelength) / sq(W->dt);
		else amplitude *= sqrtf(1.5 / PI) * wavelength / sq(W->dt);

		S->phase = phase * PI / 180;
		S->Amplitude = makeField(S->iMin, S->iMax, S->jMin, S->jMax, S->kMin, S->kMax);
		if (F==Ex || F==iEx) setAmplitude(S, W->CH->dtdx, W->CE->dtdy, W->CE->dtdz, W->CE->x, amplitude*0.5);
		if (F==Ey || F==iEy) setAmplitude(S, W->CE->dtdx, W->CH->dtdy, W->CE->dtdz, W->CE->y, amplitude*0.5);
		if (F==Ez || F==iEz) setAmplitude(S, W->CE->dtdx, W->CE->dtdy, W->CH->dtdz, W->CE->z, amplitude*0.5);
		if (F==Hx || F==iHx) setAmplitude(S, W->CE->dtdx, W->CH->dtdy, W->CH->dtdz, W->CH->x, amplitude*0.25);
		if (F==Hy || F==iHy) setAmplitude(S, W->CH->dtdx, W->CE->dtdy, W->CH->dtdz, W->CH->y, amplitude*0.25);
		if (F==Hz || F==iHz) setAmplitude(S, W->CH->dtdx, W->CH->dtdy, W->CE->dtdz, W->CH->z, amplitude*0.25);
	}
	if (wavelength && W->T != wavelength/W->dt) W->f = 1/wavelength, W->T = wavelength/W->dt;
}


/* Mode-profile source on a plane normal to `axis` read from an array file
 * (src/source.c:127-220). */
static void guidedWave(world W, int axis, char *file, field F, float pos, waveform waveForm, float wavelength, float supplement, float amplitude, float phase)
{
	alisInvalidate(W);
	int c = axis == 0 ? W->xtoi(W, pos) : axis == 1 ? W->ytoj(W, pos) : W->ztok(W, pos);
	if (axis != 0 && W->xMinSur==SYM) W->coskx = isFx(F) ? 1 : -1;
	if (axis != 1 && W->yMinSur==SYM) W->cosky = isFy(F) ? 1 : -1;
	if (axis != 2 && W->zMinSur==SYM) W->coskz = isFz(F) ? 1 : -1;
	if (axis == 0) dipoleParity(W, F, c, 1, 1);
	if (axis == 1) dipoleParity(W, F, 1, c, 1);
	if (axis == 2) dipoleParity(W, F, 1, 1, c);

	if (waveForm) {
		int d[2];
		float **f = h5read(d, file);
		if (!amplitude) amplitude = 1;
		source S = createSource(W, F, waveForm, wavelength, supplement, amplitude);
		int aMin[3] = { W->iMin, W->jMin, W->kMin }, aMax[3] = { W->iMax, W->jMax, W->kMax };
		int sym[3] = { W->xMinSur==SYM, W->yMinSur==SYM, W->zMinSur==SYM };
		int u = axis == 0 ? 1 : 0, v = axis == 2 ? 1 : 2;
		if (d[0] != 1+aMax[u]+(sym[u]?aMax[u]:-aMin[u])) printf("The dimension of guided wave does not match!\n"), exit(0);
		if (d[1] != 1+aMax[v]+(sym[v]?aMax[v]:-aMin[v])) printf("The dimension of guided wave does not match!\n"), exit(0);
		S->phase = phase * PI / 180;
		S->iMin = W->iMin, S->iMax = W->iMax, S->jMin = W->jMin, S->jMax = W->jMax, S->kMin = W->kMin, S->kMax = W->kMax;
		if (axis == 0) S->iMin = S->iMax = c;
		if (axis == 1) S->jMin = S->jMax = c;
		if (axis == 2) S->kMin = S->kMax = c;
		S->Amplitude = makeField(S->iMin, S->iMax, S->jMin, S->jMax, S->kMin, S->kMax);
		for (int i = S->iMin; i <= S->iMax; i++) for (int j = S->jMin; j <= S->jMax; j++) for (int k = S->kMin; k <= S->kMax; k++) {
			int a = axis == 0 ? j : i, b = axis == 2 ? j : k;
			S->Amplitude[i][j][k] = amplitude * f[a-aMax[u]+d[0]-1][b-aMax[v]+d[1]-1];
		}
	}
	if (wavelength && W->T != wavelength/W->dt) W->f = 1/wavelength, W->T = wavelength/W->dt;
}

void (guidedWaveX)(world W, char *file, field F, float x, waveform waveForm, float wavelength, float supplement, float amplitude, float phase, ...)
{ guidedWave(W, 0, file, F, x, waveForm, wavelength, supplement, amplitude, phase); }
void (guidedWaveY)(world W, char *file, field F, float y, waveform waveForm, float wavelength, float supplement, float amplitude, float phase, ...)
{ guidedWave(W, 1, file, F, y, waveForm, wavelength, supplement, amplitude, phase); }
void (guidedWaveZ)(world W, char *file, field F, float z, waveform waveForm, float wavelength, float supplement, float amplitude, float phase, ...)
{ guidedWave(W, 2, file, F, z, waveForm, wavelength, supplement, amplitude, phase); }


/* Row aliasing of an auxiliary world narrower than its parent along x (y): every index
 * beyond its own last node reads node iMAX (alt = 0) or its ghost iMAX+1 (alt = 1), so
 * the parent can index it over its own extent (src/source.c:223-248). */
static float ***aliasRowsX(world W, world SW, float ***g, int alt)
{
	float ***f;
	MALLOC(f, W->iMIN-1, W->iNUM+2);
	for (int i = SW->iMIN-1; i <= SW->iMAX; i++) f[i] = g[i];
	for (int i = SW->iMAX+1; i <= W->iMAX+1; i++) f[i] = g[SW->iMAX+alt];
	free(&g[SW->iMIN-1]);
	return f;
}

static float ***aliasRowsY(world W, world SW, float ***g, int alt)
{
	float ***f;
	MALLOC(f, SW->iMIN-1, SW->iNUM+2);
	MALLOC(f[SW->iMIN-1], W->jMIN-1, (SW->iNUM+2)*(W->jNUM+2));
	for (int i = SW->iMIN-1; i <= SW->iMAX+1; i++) {
		f[i] = f[SW->iMIN-1] + (W->jNUM+2)*(i-SW->iMIN+1);
		for (int j = SW->jMIN-1; j <= SW->jMAX; j++) f[i][j] = g[i][j];
		for (int j = SW->jMAX+1; j <= W->jMAX+1; j++) f[i][j] = g[i][SW->jMAX+alt];
	}
	free(&g[SW->iMIN-1][SW->jMIN-1]);
	free(&g[SW->iMIN-1]);
	return f;
}

/* Background of the auxiliary world: the objects that reach outside the total-field box
 * (layers, substrates) and the background matter (src/source.c:251-266). */
static void putBackground(world W, tfsf S, matter *M, object *O)
{
	int N = 0;
	while (O[N].S) N++;
	int first = N;
	#pragma omp parallel for reduction(min:first) schedule(dynamic)
	for (int i = W->iMIN; i <= W->iMAX; i++) for (int j = W->jMIN; j <= W->jMAX; j++) for (int k = W->kMIN; k <= W->kMAX; k++) {
		if (i < S->iMin || S->iMax < i || j < S->jMin || S->jMax < j || k < S->kMin || S->kMax < k) {
			float x = W->itox(W, i), y = W->jtoy(W, j), z = W->ktoz(W, k);
			int n = 0;
			while (n < first && O[n].S && !O[n].S(O[n], x, y, z)) n++;
			if (n < first) first = n;
		}
	}
	N = first;
	if (O[N].S || M[N].e.x != 1 || M[N].e.y && M[N].e.y != 1 || M[N].e.z && M[N].e.z != 1 || M[N].P->w || M[N].P->g)
		putObjectArray(S->W, M+N, O+N);
}

/* Total-field/scattered-field box DOM with an auxiliary world that carries the incident
 * field; NIL surfaces of SUR make that axis 2 cells wide with row aliasing
 * (src/source.c:269-334). */
tfsf (createTFSF)(world W, dom DOM, sur SUR)
{
	alisInvalidate(W);
	tfsf S;
	MALLOC(S, 0, 1);
	tfsf *chain = &W->S;
	while (*chain) chain = &(*chain)->S;
	*chain = S;

	/* the auxiliary world: an axis marked NIL shrinks to two cells across (x, y: its
	 * boundary follows the main world's -- mirror, Bloch or periodic) or keeps the main
	 * domain plus a cell each side between perfectly matched layers (z) */
	dom Dom = W->Dom;
	sur Sur = SUR;
	float *domLo[3] = {&Dom.x.Min, &Dom.y.Min, &Dom.z.Min}, *domHi[3] = {&Dom.x.Max, &Dom.y.Max, &Dom.z.Max};
	float *surLo[3] = {&Sur.x.MinSur, &Sur.y.MinSur, &Sur.z.MinSur};
	const float given[3] = {SUR.x.MinSur, SUR.y.MinSur, SUR.z.MinSur};
	for (int a = 0; a < 3; a++) {
		if (given[a] != NIL) continue;
		const axis X = axisOf(W, a);
		if (a < 2) {
			const int sym = *X.minSur == SYM;
			*domLo[a] = sym ? 0 : *X.LO;
			*domHi[a] = sym ? 0 : (*X.coord)(W, 1 + (*X.node)(W, *X.LO));
			*surLo[a] = sym ? SYM : *X.minSur == BBC ? BBC : PBC;
		}
		else {
			*domLo[a] = (*X.coord)(W, (*X.node)(W, *X.lo) - 1);
			*domHi[a] = (*X.coord)(W, (*X.node)(W, *X.hi) + 1);
			Sur.z.MinSur = Sur.z.MaxSur = PML;
		}
	}

	S->W = createWorld(Dom, W->Res, Sur, "");
	strcpy(S->W->ID, W->ID);
	S->W->isAux = 1, S->W->parent = W;
	S->W->devPrec = W->devPrec, S->W->devDftPrec = W->devDftPrec;
	world A = S->W;

	/* alt per component: E.x E.y E.z H.x H.y H.z */
	static const int altY[6] = { 0, 1, 0, 1, 0, 1 }, altX[6] = { 1, 0, 0, 0, 1, 1 };
	if (SUR.y.MinSur==NIL) {
		A->expandY = 1;
		for (int n = 0; n <= A->complexField; n++) {
			A->E[n].x = aliasRowsY(W, A, A->E[n].x, altY[0]);
			A->E[n].y = aliasRowsY(W, A, A->E[n].y, altY[1]);
			A->E[n].z = aliasRowsY(W, A, A->E[n].z, altY[2]);
			A->H[n].x = aliasRowsY(W, A, A->H[n].x, altY[3]);
			A->H[n].y = aliasRowsY(W, A, A->H[n].y, altY[4]);
			A->H[n].z = aliasRowsY(W, A, A->H[n].z, altY[5]);
		}
	}
	if (SUR.x.MinSur==NIL) {
		A->expandX = 1;
		for (int n = 0; n <= A->complexField; n++) {
			A->E[n].x = aliasRowsX(W, A, A->E[n].x, altX[0]);
			A->E[n].y = aliasRowsX(W, A, A->E[n].y, altX[1]);
			A->E[n].z = aliasRowsX(W, A, A->E[n].z, altX[2]);
			A->H[n].x = aliasRowsX(W, A, A->H[n].x, altX[3]);
			A->H[n].y = aliasRowsX(W, A, A->H[n].y, altX[4]);
			A->H[n].z = aliasRowsX(W, A, A->H[n].z, altX[5]);
		}
	}

	/* the total-field box: per axis the nodes of DOM (a single extent is a half width),
	 * clipped to the grid -- a box reaching past the domain takes the whole grid -- and the
	 * first and last strip planes inside it */
	float *boxLo[3] = {&DOM.x.Min, &DOM.y.Min, &DOM.z.Min}, *boxHi[3] = {&DOM.x.Max, &DOM.y.Max, &DOM.z.Max};
	const float boxSur[3] = {SUR.x.MinSur, SUR.y.MinSur, SUR.z.MinSur};
	int *tMin[3] = {&S->iMin, &S->jMin, &S->kMin}, *tMax[3] = {&S->iMax, &S->jMax, &S->kMax};
	int *tMIN[3] = {&S->iMIN, &S->jMIN, &S->kMIN}, *tMAX[3] = {&S->iMAX, &S->jMAX, &S->kMAX};
	for (int a = 0; a < 3; a++) {
		const axis X = axisOf(W, a);
		if ((!*boxHi[a] && *boxLo[a] > 0) || boxSur[a] == SYM) *boxLo[a] = -(*boxHi[a] = *boxHi[a] ? *boxHi[a] : *boxLo[a]);
		const int flat = *boxLo[a] == *boxHi[a];
		*tMin[a] = *X.minSur == SYM ? 0 : flat ? *X.Min : *X.Min > (*X.node)(W, *boxLo[a]) ? *X.MIN : (*X.node)(W, *boxLo[a]);
		*tMax[a] = flat ? *X.Max : *X.Max < (*X.node)(W, *boxHi[a]) ? *X.MAX : (*X.node)(W, *boxHi[a]);
		*tMIN[a] = *tMin[a] == *X.MIN ? *tMin[a] : *tMin[a] + 1;
		*tMAX[a] = *tMax[a] == *X.MAX ? *tMax[a] + 1 : *tMax[a];
	}

	if (W->CE->M && W->CE->O) putBackground(W, S, W->CE->M, W->CE->O);
	return S;
}


/* One injection plane of component F at z index k: scalar amplitude and phase; on a Bloch
 * world, per-cell phases phase + kx x + ky y at the component's staggered position
 * (src/source.c:337-355). */
static void sourcePlane(world W, field F, int k, float kx, float ky, waveform waveForm, float wavelength, float supplement, float resolution, float amplitude, float phase)
{
	source S = createSource(W, F, waveForm, wavelength, supplement, resolution);
	S->amplitude = amplitude;
	S->phase = phase;
	S->iMin = W->iMIN, S->iMax = W->iMAX, S->jMin = W->jMIN, S->jMax = W->jMAX, S->kMin = k, S->kMax = k;
	if (!W->complexField) return;
	const int xStag = F==Ex || F==Hy || F==iEx || F==iHy, yStag = F==Ey || F==Hx || F==iEy || F==iHx;
	const float di = xStag ? 0.5 : 0, dj = yStag ? 0.5 : 0;
	S->Phase = makeField(S->iMin, S->iMax, S->jMin, S->jMax, S->kMin, S->kMax);
	for (int i = S->iMin; i <= S->iMax; i++) for (int j = S->jMin; j <= S->jMax; j++)
		S->Phase[i][j][k] = phase + kx * W->itox(W, i-di) + ky * W->jtoy(W, j-dj);
}

/*
 * The E and H injection planes of a plane wave of polarisation pol from direction (theta,
 * phi) (src/source.c:358-411).  The wave enters through the z face it travels away from: E
 * on plane kE, H on kH (half a cell further), in the medium of that face (index n).  The
 * in-plane wave vector sets the Bloch phases; on a symmetry plane the polarisation sets the
 * parity.  Each polarisation injects the tangential components below -- amplitude
 * base * a * b, dropped where the azimuth makes it vanish -- into the real part and, on a
 * Bloch world, into the imaginary part a quarter period later; the chain order is the one
 * listed.
 */
typedef struct {
	field *F, *iF;
	int hPlane;                 /* 0: E plane kE, 1: H plane kH */
	int needs;                  /* 1: skipped for phi = 90, 270;  2: skipped for phi = 0, 180 */
	int sign;                   /* sign of the base amplitude */
	int a, b;                   /* factors: 0 none, 1 cos(phi), 2 sin(phi), 3 cos(theta) */
} planeTerm;

static const planeTerm planeTerms[2][4] = {
	/* P: E along the meridian, H across it */
	{{Ex, iEx, 0, 1, +1, 1, 3}, {Ey, iEy, 0, 2, +1, 2, 3}, {Hx, iHx, 1, 2, +1, 2, 0}, {Hy, iHy, 1, 1, -1, 1, 0}},
	/* S: E across the meridian, H along it */
	{{Ex, iEx, 0, 2, -1, 2, 0}, {Ey, iEy, 0, 1, +1, 1, 0}, {Hx, iHx, 1, 1, +1, 1, 3}, {Hy, iHy, 1, 2, +1, 2, 3}},
};

static void sourcePlanes(world W, tfield pol(phaser,float,float), float theta, float phi, waveform waveForm, float wavelength, float supplement, float resolution, float amplitude, float phase)
{
	const int up = 90 < theta && theta < 270, down = theta < 90 || 270 < theta;
	const int kE = up ? W->kMin : down ? W->kMax : 0, kH = up ? W->kMin : down ? W->kMax+1 : 0;
	float sinth = sinf(theta*PI/180) * sqrtf(W->CE->z?1/W->CE->z[W->iMin][W->jMin][kH]:1);
	float costh = cosf(theta*PI/180) * sqrtf(W->CE->z?1/W->CE->z[W->iMin][W->jMin][kH]:1);
	float sinph = sinf(phi*PI/180);
	float cosph = cosf(phi*PI/180);
	float kx = -2 * PI * sinth * cosph / wavelength;
	float ky = -2 * PI * sinth * sinph / wavelength;

	if (W->xMinSur == BBC) W->coskx = cosf(kx*(W->xMax-W->xMin)), W->sinkx = sinf(kx*(W->xMax-W->xMin));
	if (W->yMinSur == BBC) W->cosky = cosf(ky*(W->yMax-W->yMin)), W->sinky = sinf(ky*(W->yMax-W->yMin));
	const int alongY = phi==90 || phi==270, alongX = phi==0 || phi==180, normal = theta==0 || theta==180;
	if (W->xMinSur == SYM && alongY) W->coskx = pol==Ppol ? -1 : 1;
	if (W->xMinSur == SYM && alongX && normal) W->coskx = pol==Ppol ? 1 : -1;
	if (W->yMinSur == SYM && alongX) W->cosky = pol==Ppol ? -1 : 1;
	if (W->yMinSur == SYM && alongY && normal) W->cosky = pol==Ppol ? 1 : -1;

	if (waveForm && (pol == Ppol || pol == Spol)) {
		float n = (W->CE->z ? 1/sqrtf(W->CE->z[W->iMin][W->jMin][kH]) : 1);
		float kzdz = W->CE->dz[kE] * PI * fabsf(costh) / wavelength;
		float ampE = W->CE->dtdz[kE] * (amplitude ? amplitude : 1) * 2 / ((1+n) * sqrtf(n) * n);
		float ampH = W->CH->dtdz[kH] * (amplitude ? amplitude : 1) * 2 / ((1+n) * sqrtf(n));
		/* phases of the real parts (a quarter period ahead) and of the imaginary parts; H is
		 * half a cell further along the propagation */
		const float re[2] = {(phase+90)*PI/180, (phase+90)*PI/180+kzdz}, im[2] = {phase*PI/180, phase*PI/180+kzdz};
		const float factor[4] = {1, cosph, sinph, costh};
		const planeTerm *T = planeTerms[pol == Spol];
		for (int part = 0; part <= W->complexField; part++)
			for (int q = 0; q < 4; q++) {
				if ((T[q].needs == 1 && alongY) || (T[q].needs == 2 && alongX)) continue;
				float base = T[q].hPlane ? ampH : ampE;
				if (T[q].sign < 0) base = -base;
				float amp = base * factor[T[q].a];
				if (T[q].b) amp = amp * factor[T[q].b];
				sourcePlane(W, part ? T[q].iF : T[q].F, T[q].hPlane ? kH : kE, kx, ky, waveForm, wavelength, supplement, resolution,
				            amp, part ? im[T[q].hPlane] : re[T[q].hPlane]);
			}
	}
	if (wavelength && W->T != wavelength/W->dt) W->f = 1/wavelength, W->T = wavelength/W->dt;
}

/* Plane wave incident on the total-field box Dom from direction (theta, phi)
 * (src/source.c:414-432). */
void (planewave)(world W, dom Dom, tfield pol(phaser,float,float), float theta, float phi, waveform waveForm, float wavelength, float supplement, float amplitude, float phase, ...)
{
	while (theta < 0) theta += 360;
	while (theta >= 360) theta -= 360;
	while (phi < 0) phi += 360;
	while (phi >= 360) phi -= 360;

	sur Sur = {{NIL}, {NIL}, {NIL}};
	if ((theta != 0 && theta != 180) && (phi != 90 && phi != 270)) Sur.x.MinSur = BBC;
	if ((theta != 0 && theta != 180) && (phi !=  0 && phi != 180)) Sur.y.MinSur = BBC;

	float resolution = amplitude;
	if (!amplitude) amplitude = 1;
	tfsf S = createTFSF(W, Dom, Sur);
	sourcePlanes(W, pol, theta, phi, 0, wavelength, supplement, resolution, 0, 0);
	sourcePlanes(S->W, pol, theta, phi, waveForm, wavelength, supplement, resolution, amplitude, phase);
	W->N = S->W->N;
}


/* A Gaussian beam launched from the z face it travels away from (src/source.c:435-477): in
 * the medium of index n at the launch plane, with waist w0 = lambda spot / (n sqrt(2 ln 2))
 * (spot: the full width at half maximum in wavelengths) at the focus (x, y, z), the launch
 * plane, a distance d from the focus, carries the field of the beam there -- amplitude
 * w0/w(d) e^{-r^2/w(d)^2}, phase -k r^2 / (2 R(d)) plus the Gouy phase atan(d/z0) -- with
 * the Rayleigh range z0 = pi n w0^2 / lambda, w(d) = w0 sqrt(1 + (d/z0)^2) and
 * R(d) = d (1 + (z0/d)^2); r is measured at the component's staggered position, along the
 * axes whose focus coordinate is finite.  Without a spot size the plane is uniform. */
typedef struct { float w0, z0, w, R, gouy, k; } gaussian;

static gaussian gaussianAt(float n, float wavelength, float spot, float d)
{
	gaussian g;
	g.k = 2 * PI * n / wavelength;
	g.w0 = wavelength * (spot ? spot : 1) / (n * sqrtf(-2.0 * logf(0.5)));
	g.z0 = PI * n * sq(g.w0) / wavelength;
	g.w = g.w0 * sqrtf(1 + sq(d/g.z0));
	g.R = d * (1 + sq(g.z0/d));
	g.gouy = atanf(d/g.z0);
	return g;
}

static void beam(world W, field F, float x, float y, float z, float theta, waveform waveForm, float wavelength, float supplement, float spotsize, float amplitude, float phase)
{
	while (theta < 0) theta += 360;
	while (theta >= 360) theta -= 360;
	if (W->xMinSur==SYM) W->coskx = (F==Ex || F==iEx) ? 1 : -1;
	if (W->yMinSur==SYM) W->cosky = (F==Ey || F==iEy) ? 1 : -1;
	if (waveForm) {
		const int k = (90<theta && theta<270) ? W->kMin : (theta<90 || 270<theta) ? W->kMax : 0;
		const float n = W->CE->z ? 1/sqrtf(W->CE->z[W->xtoi(W,MIN(x,W->xMax))][W->ytoj(W,MIN(y,W->yMax))][k]) : 1;
		source S = createSource(W, F, waveForm, wavelength, supplement, amplitude);
		S->iMin = W->iMIN, S->iMax = W->iMAX, S->jMin = W->jMIN, S->jMax = W->jMAX, S->kMin = k, S->kMax = k;
		const float amp = (amplitude ? amplitude : 1) / sqrtf(n*n*n);
		if (!spotsize) {
			S->amplitude = amp;
			S->phase = phase * PI / 180;
		}
		else {
			const gaussian g = gaussianAt(n, wavelength, spotsize, fabsf(z - W->ktoz(W,k)));
			S->Amplitude = makeField(S->iMin, S->iMax, S->jMin, S->jMax, S->kMin, S->kMax);
			S->Phase = makeField(S->iMin, S->iMax, S->jMin, S->jMax, S->kMin, S->kMax);
			for (int i = S->iMin; i <= S->iMax; i++)
				for (int j = S->jMin; j <= S->jMax; j++) {
					float r2 = 0;
					if (x < INF) r2 += sq(x - W->itox(W, F==Ex ? i-0.5 : i));
					if (y < INF) r2 += sq(y - W->jtoy(W, F==Ey ? j-0.5 : j));
					S->Amplitude[i][j][k] = amp * g.w0/g.w * expf(-r2/sq(g.w));
					S->Phase[i][j][k] = phase * PI / 180 - g.k * (0.5 * r2 / g.R) + g.gouy;
				}
		}
	}
	if (wavelength && W->T != wavelength/W->dt) W->f = 1/wavelength, W->T = wavelength/W->dt;
}

void (focusedBeam)(world W, dom Dom, field F, float x, float y, float z, float theta, waveform waveForm, float wavelength, float supplement, float spotsize, float amplitude, float phase, ...)
{
	alisInvalidate(W);
	if (Dom.x.Min==W->Dom.x.Min && Dom.x.Max==W->Dom.x.Max &&
	    Dom.y.Min==W->Dom.y.Min && Dom.y.Max==W->Dom.y.Max &&
	    Dom.z.Min==W->Dom.z.Min && Dom.z.Max==W->Dom.z.Max) {
		beam(W, F, x, y, z, theta, waveForm, wavelength, supplement, spotsize, amplitude, phase);
	}
	else {
		sur Sur = {{NIL}, {NIL}, {NIL}};
		if (x < INF) Sur.x.MinSur = W->Sur.x.MinSur, Sur.x.MaxSur = W->Sur.x.MaxSur;
		if (y < INF) Sur.y.MinSur = W->Sur.y.MinSur, Sur.y.MaxSur = W->Sur.y.MaxSur;
		tfsf S = createTFSF(W, Dom, Sur);
		beam(W, F, x, y, z, theta, 0, wavelength, supplement, 0, 0, 0);
		beam(S->W, F, x, y, z, theta, waveForm, wavelength, supplement, spotsize, amplitude, phase);
		W->N = S->W->N;
	}
}


static void freeSourceChain(source S)
{
	while (S) {
		source next = S->S;
		if (S->Amplitude) removeField(S->Amplitude, S->iMin, S->jMin, S->kMin);
		if (S->Phase) removeField(S->Phase, S->iMin, S->jMin, S->kMin);
		free(S);
		S = next;
	}
}

void removeSources(world W)
{
	alisInvalidate(W);
	freeSourceChain(W->SE), W->SE = 0;
	freeSourceChain(W->SH), W->SH = 0;
	for (tfsf S = W->S; S; ) {
		tfsf next = S->S;
		deleteWorld(S->W);
		free(S);
		S = next;
	}
	W->S = 0;
	W->f = 0;
	W->T = 0;
	W->N = 0;
}


/*
 * Injection of a source chain for one half step (src/update.c:161-179): each active source
 * (W->n < S->N) adds amplitude * waveform(t, phase) on its box.  Waveforms are host
 * functions, evaluated here exactly as the reference does; per-cell phases give one value
 * per cell.  The addition itself runs on the device.
 */
void sourceHalfStep(world W, source S, float t, int h)
{
	int index = 0;
	static float *cellWave;
	static long cellCap;
	for (; S; S = S->S, index++) {
		if (!(W->n < S->N)) continue;
		if (S->Phase) {
			long count = (long)(1+S->iMax-S->iMin)*(1+S->jMax-S->jMin)*(1+S->kMax-S->kMin);
			if (count > cellCap) cellWave = realloc(cellWave, count*sizeof(float)), cellCap = count;
			long c = 0;
			for (int i = S->iMin; i <= S->iMax; i++) for (int j = S->jMin; j <= S->jMax; j++) for (int k = S->kMin; k <= S->kMax; k++)
				cellWave[c++] = S->waveForm(W, S, t, S->Phase[i][j][k]);
			alisFatal("inject", alis_dev_inject(W->dev, h, index, 0, cellWave, count));
		}
		else {
			float f = S->waveForm(W, S, t, S->phase);
			alisFatal("inject", alis_dev_inject(W->dev, h, index, f, 0, 0));
		}
	}
}
