/*
 * Device residency of worlds and phasers.  A world is committed to the GPU at its first
 * time step (after setup); setup calls made later bring the state home and drop the device
 * copy, which is rebuilt at the next step.  Host arrays of the reference layout are kept as
 * mirrors and refreshed on demand: a small box for get(), everything for the host
 * observers, and everything when timer() reports the end of a loop.
 */
#define _GNU_SOURCE             /* PTHREAD_RECURSIVE_MUTEX_INITIALIZER_NP */
#include <pthread.h>
#include <string.h>
#include "internal.h"

/* The registries of live worlds and phasers below are shared with the serving loop's setup
 * thread, which creates worlds and phasers while the stepping thread reads and releases
 * others: every access holds this (recursive) lock. */
static pthread_mutex_t registry = PTHREAD_RECURSIVE_MUTEX_INITIALIZER_NP;
#define LOCK() pthread_mutex_lock(&registry)
#define UNLOCK() pthread_mutex_unlock(&registry)

static world *liveWorlds;
/* Batches of updateWorldsE/H (and of updateE/H on worlds with TF/SF boxes), most recently
 * used first, keyed by their device worlds. */
#define NBATCH 8
static struct { alis_dev_batch *B; alis_dev_world **D; int n; } batches[NBATCH];

/* Drop every cached batch: called whenever a device world goes away, so that a later world
 * at a reused address can never meet tables of a freed one. */
static void dropBatch(void)
{
	for (int a = 0; a < NBATCH; a++) {
		if (batches[a].B) alis_dev_batch_release(batches[a].B);
		free(batches[a].D);
		batches[a].B = 0, batches[a].D = 0, batches[a].n = 0;
	}
}

/* The batch of the committed worlds L[0..n-1] (created on first use, then reused). */
alis_dev_batch *alisBatch(world *L, int n)
{
	int hit = -1;
	for (int a = 0; a < NBATCH && hit < 0; a++) {
		if (!batches[a].B || batches[a].n != n) continue;
		int same = 1;
		for (int b = 0; same && b < n; b++) same = batches[a].D[b] == L[b]->dev;
		if (same) hit = a;
	}
	if (hit < 0) {
		if (batches[NBATCH-1].B) alis_dev_batch_release(batches[NBATCH-1].B), free(batches[NBATCH-1].D);
		hit = NBATCH - 1;
		batches[hit].D = malloc(n * sizeof(alis_dev_world *));
		for (int b = 0; b < n; b++) batches[hit].D[b] = L[b]->dev;
		batches[hit].n = n;
		alisFatal("batch", alis_dev_batch_create(batches[hit].D, n, &batches[hit].B));
	}
	if (hit) {
		typeof(batches[0]) e = batches[hit];
		memmove(&batches[1], &batches[0], hit * sizeof batches[0]);
		batches[0] = e;
	}
	return batches[0].B;
}

static int nWorlds, capWorlds;
static phaser *livePhasers;
static int nPhasers, capPhasers;

void alisFatal(const char *what, int rc)
{
	if (!rc) return;
	fprintf(stderr, "alis: %s failed: %s\n", what, alis_dev_error());
	exit(1);
}

int alisPrecisionDefault(int *fieldBits, int *dftBits)
{
	const char *p = getenv("ALIS_PRECISION");
	*fieldBits = 0, *dftBits = 1;
	if (p && !strcmp(p, "fp64")) *fieldBits = 1, *dftBits = 1;
	else if (p && !strcmp(p, "ref")) *fieldBits = 0, *dftBits = 0;
	return 0;
}

static void setPrecisionTree(world W, int f64, int d64)
{
	W->devPrec = f64, W->devDftPrec = d64;
	for (tfsf S = W->S; S; S = S->S) setPrecisionTree(S->W, f64, d64);
}

void worldPrecision(world W, int fieldBits, int dftBits)
{
	alisInvalidate(W);
	setPrecisionTree(W, fieldBits == 64, dftBits == 64);
}

void alisRegisterWorld(world W)
{
	LOCK();
	if (nWorlds == capWorlds) capWorlds = capWorlds ? 2*capWorlds : 16, liveWorlds = realloc(liveWorlds, capWorlds*sizeof(world));
	liveWorlds[nWorlds++] = W;
	UNLOCK();
}

static void forgetProbes(world W);

void alisForgetWorld(world W)
{
	forgetProbes(W);
	alisForgetSpectra(W);
	alisForgetTexts(W);
	if (W->dev) {
		dropBatch();
		alis_dev_release(W->dev);
		W->dev = 0;
	}
	LOCK();
	for (int a = 0; a < nWorlds; a++) if (liveWorlds[a] == W) { liveWorlds[a] = liveWorlds[--nWorlds]; break; }
	for (int a = 0; a < nPhasers; a++) if (livePhasers[a]->W == W) livePhasers[a]->devId = -1;
	UNLOCK();
}

void alisRegisterPhaser(phaser P)
{
	LOCK();
	if (nPhasers == capPhasers) capPhasers = capPhasers ? 2*capPhasers : 16, livePhasers = realloc(livePhasers, capPhasers*sizeof(phaser));
	livePhasers[nPhasers++] = P;
	UNLOCK();
}

void alisForgetPhaser(phaser P)
{
	LOCK();
	for (int a = 0; a < nPhasers; a++) if (livePhasers[a] == P) { livePhasers[a] = livePhasers[--nPhasers]; break; }
	UNLOCK();
}

/* First element of the phaser array that contains P (arrays end with an f == 0 entry). */
phaser phaserBase(phaser P)
{
	phaser found = 0;
	LOCK();
	for (int a = 0; a < nPhasers && !found; a++) {
		phaser B = livePhasers[a];
		if (P >= B && P < B + B->count) found = B;
	}
	UNLOCK();
	return found;
}

static void commitTree(world W)
{
	if (W->dev) return;
	for (tfsf S = W->S; S; S = S->S) commitTree(S->W);
	alis_dev_world *D;
	alis_dev_world_desc desc;
	describeWorld(W, &desc);
	int rc = alis_dev_commit(&desc, W->devPrec, W->devDftPrec, &D);
	freeWorldDesc(&desc);
	alisFatal("commit", rc);
	int s = 0;
	for (tfsf S = W->S; S; S = S->S, s++) alisFatal("TF/SF link", alis_dev_link_tfsf(D, s, S->W->dev));
	W->dev = D;
	W->hostStale = 0;
}

void alisCommit(world W)
{
	if (W->dev) return;
	if (W->isAux && W->parent) {
		alisCommit(W->parent);
		return;
	}
	commitTree(W);
}

static void syncTree(world W)
{
	if (W->dev && W->hostStale) {
		alisFatal("sync", alis_dev_sync(W->dev, 0, -1, 0, -1, 0, -1));
		W->hostStale = 0;
	}
	for (tfsf S = W->S; S; S = S->S) syncTree(S->W);
}

void alisSyncPhaser(phaser P);

/* Bring fields, CPML and ADE state of W (and its TF/SF auxiliary worlds) and the
 * accumulators of phasers recording W to the host arrays. */
void syncWorld(world W)
{
	syncTree(W);
	LOCK();
	for (int a = 0; a < nPhasers; a++) if (livePhasers[a]->W == W) alisSyncPhaser(livePhasers[a]);
	UNLOCK();
}

void alisSyncPhaser(phaser P)
{
	phaser B = phaserBase(P);
	if (!B || B->devId < 0 || !B->hostStale || !B->W->dev) return;
	alisFatal("phaser sync", alis_dev_phaser_sync(B->W->dev, B->devId));
	B->hostStale = 0;
}

static void releaseTree(world W)
{
	for (tfsf S = W->S; S; S = S->S) releaseTree(S->W);
	if (W->dev) {
		dropBatch();
		alis_dev_release(W->dev);
		W->dev = 0;
	}
	W->hostStale = 0;
}

void alisInvalidate(world W)
{
	if (W->isAux && W->parent && W->parent->dev) {
		alisInvalidate(W->parent);
		return;
	}
	if (!W->dev) return;
	LOCK();
	for (int a = 0; a < nPhasers; a++) {
		phaser B = livePhasers[a];
		if (B->W != W) continue;
		alisSyncPhaser(B);
		B->devId = -1;
	}
	UNLOCK();
	syncTree(W);
	releaseTree(W);
}

/* Host arrays were modified after a sync: drop the device copy without reading it back. */
void alisDropDevice(world W)
{
	if (W->isAux && W->parent && W->parent->dev) W = W->parent;
	LOCK();
	for (int a = 0; a < nPhasers; a++) {
		phaser B = livePhasers[a];
		if (B->W == W && B->devId >= 0) alisSyncPhaser(B), B->devId = -1;
	}
	UNLOCK();
	releaseTree(W);
}

/* The host mirrors of the arrays in mask (alis_dev_sync_boxes) over a box, without
 * touching the rest of the world's mirrors. */
void alisSyncArrays(world W, int i0, int i1, int j0, int j1, int k0, int k1, int mask)
{
	if (W->isAux && W->parent) {
		syncWorld(W);
		return;
	}
	if (!W->dev || !W->hostStale) return;
	const int box[7] = {i0, i1, j0, j1, k0, k1, mask};
	alis_dev_world *D = W->dev;
	alisFatal("box sync", alis_dev_sync_boxes(&D, box, 1));
}

void alisSyncBox(world W, int i0, int i1, int j0, int j1, int k0, int k1)
{
	if (!W->dev || !W->hostStale) return;
	alisFatal("box sync", alis_dev_sync(W->dev, i0, i1, j0, j1, k0, k1));
}

/*
 * Probe boxes.  get() reads a few cells around one node per call, every step in a typical
 * program, often at several nodes of several worlds.  Each box it reads is remembered per
 * world with the half step it was last copied at; the first get() after a half step copies
 * the boxes of every world probed so far in one engine call (one gather and one copy per
 * world, one wait per stream), and later calls in the same half step read the host arrays.
 */
#define MAXPROBES 4096
typedef struct {
	world W;
	int n;
	int next;                       /* where the next lookup starts: probes are usually read
	                                   in the same order every step */
	int (*box)[7];                  /* i0 i1 j0 j1 k0 k1, arrays (alis_dev_sync_boxes) */
	unsigned long *epoch;
} probeSet;
static probeSet **probeSets;     /* every world's set (W->probes points at its own) */
static int nProbeSets, capProbeSets;
static unsigned long epochNow = 1;

void alisStepped(void) { epochNow++; }

static void forgetProbes(world W)
{
	for (int a = 0; a < nProbeSets; a++)
		if (probeSets[a]->W == W) {
			free(probeSets[a]->box), free(probeSets[a]->epoch), free(probeSets[a]);
			probeSets[a] = probeSets[--nProbeSets];
			break;
		}
	W->probes = 0;
}

void alisProbe(world W, int i0, int i1, int j0, int j1, int k0, int k1, int mask)
{
	if (!W->dev || !W->hostStale) return;
	probeSet *P = W->probes;
	if (!P) {
		if (nProbeSets == capProbeSets) capProbeSets = capProbeSets ? 2 * capProbeSets : 8, probeSets = realloc(probeSets, capProbeSets * sizeof *probeSets);
		P = probeSets[nProbeSets++] = malloc(sizeof *P);
		*P = (probeSet) {W, 0, 0, malloc(MAXPROBES * sizeof *P->box), malloc(MAXPROBES * sizeof *P->epoch)};
		W->probes = P;
	}
	const int b[7] = {i0, i1, j0, j1, k0, k1, mask};
	int at = -1;
	for (int c = 0; c < P->n && at < 0; c++) {
		const int q = (P->next + c) % P->n;
		if (!memcmp(P->box[q], b, sizeof b)) at = q;
	}
	if (at >= 0) P->next = (at + 1) % P->n;
	if (at < 0) {
		if (P->n == MAXPROBES) {              /* too many distinct probes: copy this one alone */
			alisSyncBox(W, i0, i1, j0, j1, k0, k1);
			return;
		}
		at = P->n++;
		memcpy(P->box[at], b, sizeof b);
		P->epoch[at] = 0;
	}
	if (P->epoch[at] == epochNow) return;
	/* copy every stale box of every probed world that the device holds newer */
	int m = 0;
	for (int a = 0; a < nProbeSets; a++)
		if (probeSets[a]->W->dev && probeSets[a]->W->hostStale) m += probeSets[a]->n;
	alis_dev_world **D = malloc(m * sizeof *D);
	int *boxes = malloc(7 * m * sizeof(int));
	m = 0;
	for (int a = 0; a < nProbeSets; a++) {
		probeSet *Q = probeSets[a];
		if (!Q->W->dev || !Q->W->hostStale) continue;
		for (int q = 0; q < Q->n; q++) {
			if (Q->epoch[q] == epochNow) continue;
			D[m] = Q->W->dev;
			memcpy(boxes + 7 * m, Q->box[q], sizeof b);
			Q->epoch[q] = epochNow;
			m++;
		}
	}
	alisFatal("probe sync", alis_dev_sync_boxes(D, boxes, m));
	free(D), free(boxes);
}

void alisSyncAll(void)
{
	LOCK();
	for (int a = 0; a < nWorlds; a++) if (!liveWorlds[a]->isAux) syncWorld(liveWorlds[a]);
	for (int a = 0; a < nPhasers; a++) alisSyncPhaser(livePhasers[a]);
	UNLOCK();
}
