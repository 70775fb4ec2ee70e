/*
 * Time stepping: updateE / updateH drive the device half-step programs in the reference's
 * order (src/update.c:275-341), the loop timer, and the host-array forms of copyBoundary
 * and the nonlinear couplings.
 */
#include <string.h>
#include <sys/time.h>
#include "internal.h"

/* Loop predicate with a progress line (src/update.c:14-45): returns n <= N.  Step 0 (or the
 * first call) starts the clock; afterwards, at most once per wall-clock second and after the
 * last step, it prints "\r[str ][hh:mm:ss/hh:mm:ss] [n-1/N: p%]": the time since the start
 * and the estimated total, the remaining (N - n + 1) steps timed at the pace since the last
 * line (an estimate that leaves more than an hour is rounded down to whole minutes).  When
 * the loop ends, device state is copied to the host arrays, so that code after the loop may
 * read them directly. */
static void clockText(unsigned long long us, char out[32])
{
	const unsigned long long s = us / 1000000ULL;
	snprintf(out, 32, "%02llu:%02llu:%02llu", s / 3600, s / 60 % 60, s % 60);
}

int (timer)(long int n, long int N, char *str, ...)
{
	static struct { long step; unsigned long long us; } first, last;
	struct timeval tv;
	gettimeofday(&tv, 0);
	const unsigned long long now = tv.tv_sec * 1000000ULL + tv.tv_usec;
	if (!n || !first.step) {
		first.step = last.step = n, first.us = last.us = now;
	}
	else if (n > N || now / 1000000ULL > last.us / 1000000ULL) {
		const unsigned long long elapsed = now - first.us;
		unsigned long long total = elapsed + (N - n + 1) * (now - last.us) / (n - last.step);
		if (total - elapsed > 3600ULL * 1000000ULL) total -= total % (60ULL * 1000000ULL);
		last.step = n, last.us = now;
		char e[32], t[32];
		clockText(elapsed, e), clockText(total, t);
		printf("\r%s%s", str, str[0] ? " " : "");
		printf("[%s/%s] [%ld/%ld: %.1f%%]", e, t, n - 1, N, 100.0 * (n - 1) / N);
		if (n > N) printf("\n");
		fflush(stdout);
	}
	if (n > N) alisSyncAll();
	return n <= N;
}


/* E half step: dispersive currents and curl of H with CPML on the device, then the TF/SF
 * auxiliary worlds and their strips, the E sources at the H time, the ghost layers, and the
 * clock (src/update.c:275-299). */
void updateE(world W)
{
	alisCommit(W);
	alisFatal("E core", alis_dev_core(W->dev, 0));
	int s = 0;
	for (tfsf S = W->S; S; S = S->S, s++) {
		updateE(S->W);
		alisFatal("E strips", alis_dev_strips(W->dev, s, 0));
	}
	if (W->SE) sourceHalfStep(W, W->SE, W->H->t, 0);
	alisFatal("E ghosts", alis_dev_ghosts(W->dev, 0));
	W->t = W->E->t = ++W->n * W->dt;
	W->hostStale = 1;
	alisStepped();
}

/* H half step: curl of E with CPML and the ADE current update on the new E, TF/SF,
 * H sources at the E time, ghost layers (src/update.c:302-341). */
void updateH(world W)
{
	alisCommit(W);
	alisFatal("H core", alis_dev_core(W->dev, 1));
	int s = 0;
	for (tfsf S = W->S; S; S = S->S, s++) {
		updateH(S->W);
		alisFatal("H strips", alis_dev_strips(W->dev, s, 1));
	}
	if (W->SH) sourceHalfStep(W, W->SH, W->E->t, 1);
	alisFatal("H ghosts", alis_dev_ghosts(W->dev, 1));
	W->t = W->H->t = W->E->t + 0.5 * W->dt;
	W->hostStale = 1;
	alisStepped();
}


/*
 * Several worlds per call (an addition to the reference's API, which steps one world per
 * call): the half step of every world in W[0..n-1] and of their TF/SF auxiliary worlds in
 * shared launches -- one curl launch for all of them, one post launch per program round.
 * Each world's arithmetic and order are those of updateE/updateH on it alone.
 */
static void collectTree(world W, world **L, int *n, int *cap)
{
	if (*n == *cap) *cap = *cap ? 2 * *cap : 16, *L = realloc(*L, *cap * sizeof(world));
	(*L)[(*n)++] = W;
	for (tfsf S = W->S; S; S = S->S) collectTree(S->W, L, n, cap);
}

static void halfStepMany(world *Ws, int n, int h)
{
	static world *L;
	static int cap;
	int m = 0;
	for (int a = 0; a < n; a++) alisCommit(Ws[a]);
	for (int a = 0; a < n; a++) collectTree(Ws[a], &L, &m, &cap);
	alis_dev_batch *B = alisBatch(L, m);
	alisFatal(h ? "H core" : "E core", alis_dev_batch_core(B, h));
	for (int a = 0; a < m; a++) {
		world W = L[a];
		int s = 0;
		for (tfsf S = W->S; S; S = S->S, s++) alisFatal("strips", alis_dev_strips(W->dev, s, h));
		if (!h && W->SE) sourceHalfStep(W, W->SE, W->H->t, 0);
		if (h && W->SH) sourceHalfStep(W, W->SH, W->E->t, 1);
	}
	alisFatal(h ? "H ghosts" : "E ghosts", alis_dev_batch_post(B, h));
	for (int a = 0; a < m; a++) {
		world W = L[a];
		if (!h) W->t = W->E->t = ++W->n * W->dt;
		else W->t = W->H->t = W->E->t + 0.5 * W->dt;
		W->hostStale = 1;
	}
	alisStepped();
}

void updateWorldsE(world *W, int n) { halfStepMany(W, n, 0); }
void updateWorldsH(world *W, int n) { halfStepMany(W, n, 1); }


/* Host-array ghost copy (src/update.c:182-191): F = c F' + s G', G = c G' - s F'. */
void (copyBoundary)(float ***F, float ***G, float cosk, float sink, int iMin, int iMax, int jMin, int jMax, int kMin, int kMax, int iAlt, int jAlt, int kAlt)
{
	for (int i = iMin; i <= iMax; i++) for (int j = jMin; j <= jMax; j++) for (int k = kMin; k <= kMax; k++) {
		float f = F[i+iAlt][j+jAlt][k+kAlt];
		if (G) {
			float g = G[i+iAlt][j+jAlt][k+kAlt];
			F[i][j][k] = cosk * f + sink * g;
			G[i][j][k] = cosk * g - sink * f;
		}
		else F[i][j][k] = cosk * f;
	}
}


/*
 * Second-order (three-wave) coupling of the pump worlds' E into WS inside WS's dispersive
 * objects (src/update.c:344-411), on the device (alis_dev_three_wave): the worlds are
 * committed if they are not yet, and the d tensor of the n-th object is read from the n-th
 * matter entry, as the reference does.  The signal world may not be one of the pumps (the
 * reference's sequential loop would then read values it has already changed).
 */
static void threeWave(world WS, world WP1, world WP2, float Alt)
{
	if (WS == WP1 || WS == WP2) {
		printf("alis: nonLinearSHG/TWM: the signal world is one of the pump worlds\n");
		exit(1);
	}
	alisCommit(WS), alisCommit(WP1), alisCommit(WP2);
	for (int n = 0; n < WS->CE->N; n++)
		alisFatal("three-wave coupling", alis_dev_three_wave(WS->dev, WP1->dev, WP2->dev, n, &WS->CE->M[n].d[0][0], Alt));
	WS->hostStale = 1;
}

void nonLinearSHG(world WS, world WP, float Alt) { threeWave(WS, WP, WP, Alt); }
void nonLinearTWM(world WS, world WP1, world WP2, float Alt) { threeWave(WS, WP1, WP2, Alt); }
