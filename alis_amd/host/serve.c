/*
 * Serving scattering solves with continuous batching (an addition to the reference, which
 * runs one world per program).  Requests arrive over time; each is a TF/SF scattering run
 * of one Drude-Lorentz sphere with a 16-frequency phaser box.
 *
 * Two threads.  The setup thread builds each request's world once it has arrived (grid,
 * voxelised sphere, TF/SF box, phasers: host work), in arrival order, at most `slots`
 * ahead of admission.  The serving thread steps all admitted worlds together
 * (updateWorldsE/updateWorldsH/updatePhaserArrays: shared launches over worlds of
 * different sizes) and, between any two steps, admits built requests in arrival order
 * while a slot is free and their device bytes fit the budget; a request's first step
 * commits it to the device, where its fields stay until it finishes and is retired.  So
 * the GPU keeps stepping the admitted worlds while new ones are built, and a request joins
 * the batch at the first step after it is built.  Each result is the same as that request
 * served alone: the shared launches compute every cell of every world exactly as its own
 * launch does.
 */
#include <pthread.h>
#include <time.h>
#include <string.h>
#include "internal.h"
#include "alis_serve.h"

static double nowSec(void)
{
	struct timespec t;
	clock_gettime(CLOCK_MONOTONIC, &t);
	return t.tv_sec + 1e-9 * t.tv_nsec;
}

#define SERVE_DX 2.0f           /* nm */
#define SERVE_PML 8
#define SERVE_MARGIN 8          /* cells between the sphere and the domain edge, at least */

/* Largest sphere radius (cells) of a request of grid edge n: the TF/SF box (r + 3 cells)
 * and the phaser box (r + 6 cells) stay two cells inside the domain. */
int alisServeMaxRadius(int n)
{
	const int d = n - 2 * (SERVE_PML + 1);
	return (d - 2) / 2 - SERVE_MARGIN;
}

static world requestWorld(const alisRequest *q, phaser *P)
{
	const float dx = SERVE_DX;
	const int d = q->cells - 2 * (SERVE_PML + 1);            /* domain nodes per edge */
	const float lo = -dx * ((d - 2) / 2), hi = dx * (d / 2);
	dom D = {{lo, hi}, {lo, hi}, {lo, hi}};
	res R = {1, {dx}, {1240}};
	sur S = {{PML}, {PML}, {PML}, {SERVE_PML}};
	world W = createWorld(D, R, S, "");
	const float r = q->radius * dx;
	matter M = {{1.5}, {{q->plasma, 0.07}, {2.6, 0.3, 0.5}, {3.5, 0.6, 0.8}}};
	object Sphere = {Ball, {{-r, r}, {-r, r}, {-r, r}}};
	putObjects(W, M, Sphere, (matter){{1}});
	const float tf = r + 3 * dx, ph = r + 6 * dx;
	dom DTF = {{-tf, tf}, {-tf, tf}, {-tf, tf}};
	planewave(W, DTF, Ppol, 0, 0, Band, 400, 900);
	/* 16 wavelengths from 385 to 910 nm, equally spaced in frequency */
	*P = createPhasers(W, 400, 900, 35, -ph, ph, -ph, ph, -ph, ph);
	return W;
}

/* Device bytes of a request before it is committed: fields and 1/eps of the padded grid,
 * CPML slabs, ADE state of the sphere, phaser accumulators (FP64) and staging.  An upper
 * estimate; the real figure replaces it after the request's first tick. */
static double requestBytes(const alisRequest *q, int fp64)
{
	const double s = fp64 ? 8 : 4, n = q->cells, row = 32 * ceil((n + 2) / 32 + 1);
	const double grid = (n + 2) * (n + 2) * row;
	const double box = 2 * q->radius + 1, phx = 2 * (q->radius + 6) + 1;
	return grid * s * 9 + 6 * 4 * SERVE_PML * (n + 2) * row * s + box * box * box * s * 18
		+ 16.0 * 4 * 6 * phx * phx * (2 * 8 + 2 * 4) + (4 << 20);
}

static long deviceBytes(world W)
{
	long b = W->dev ? alis_dev_bytes(W->dev) : 0;
	for (tfsf S = W->S; S; S = S->S) b += deviceBytes(S->W);
	return b;
}

/* splitmix64 (Steele, Lea, Flood 2014): the state advances by the golden-ratio increment
 * and each output is the state through two xor-shift-multiply rounds. */
static unsigned long long splitmix(unsigned long long *x)
{
	unsigned long long z = (*x += 0x9e3779b97f4a7c15ULL);
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
	return z ^ (z >> 31);
}

static double uniform(unsigned long long *x)
{
	return (splitmix(x) >> 11) * 0x1p-53;
}

void alisServeRequests(unsigned long long seed, int n, double rate, alisRequest *R)
{
	static const int edges[4] = {64, 96, 128, 192};
	unsigned long long x = seed;
	double t = 0;
	for (int a = 0; a < n; a++) {
		t += -log(1 - uniform(&x)) / rate;
		R[a].arrival = t;
		R[a].cells = edges[splitmix(&x) % 4];
		const int rmax = MIN(40, alisServeMaxRadius(R[a].cells));
		R[a].radius = 10 + splitmix(&x) % (rmax - 10 + 1);
		R[a].steps = 1000 + splitmix(&x) % 5001;
		R[a].plasma = 8 + 1.5 * uniform(&x);
	}
}

typedef struct {
	int id;
	world W;
	phaser P;
	int done, steps;
	double bytes;
} active;

/* The requests built by the setup thread and not yet admitted: request id sits in slot
 * id % depth while taken <= id < built. */
typedef struct {
	pthread_mutex_t m;
	pthread_cond_t cv;          /* a request was built, taken, or the run stops */
	const alisRequest *R;
	alisResult *out;
	int n, depth, built, taken, stop;
	active *slot;
	double t0, setup;
} setupQueue;

static void waitFor(setupQueue *q, double seconds)
{
	struct timespec t;
	clock_gettime(CLOCK_MONOTONIC, &t);
	const double at = t.tv_sec + 1e-9 * t.tv_nsec + seconds;
	t.tv_sec = (time_t)at, t.tv_nsec = (long)((at - (double)t.tv_sec) * 1e9);
	pthread_cond_timedwait(&q->cv, &q->m, &t);
}

static void *setupThread(void *arg)
{
	setupQueue *q = arg;
	pthread_mutex_lock(&q->m);
	while (!q->stop && q->built < q->n) {
		if (q->built - q->taken >= q->depth) {
			pthread_cond_wait(&q->cv, &q->m);
			continue;
		}
		const double early = q->R[q->built].arrival - (nowSec() - q->t0);
		if (early > 0) {
			waitFor(q, early);
			continue;
		}
		const int id = q->built;
		pthread_mutex_unlock(&q->m);
		const double s = nowSec();
		active a = {id, 0, 0, 0, q->R[id].steps, 0};
		a.W = requestWorld(&q->R[id], &a.P);
		const double e = nowSec();
		pthread_mutex_lock(&q->m);
		q->setup += e - s;
		q->out[id].ready = e - q->t0;
		q->slot[id % q->depth] = a;
		__atomic_store_n(&q->built, id + 1, __ATOMIC_RELEASE);
		pthread_cond_broadcast(&q->cv);
	}
	pthread_mutex_unlock(&q->m);
	return 0;
}

static alisServeStats lastStats;

void alisServeLastStats(alisServeStats *s)
{
	*s = lastStats;
}

int alisServe(const alisRequest *R, int n, const alisServeConfig *C, alisResult *out)
{
	if (n < 0 || !C || C->stepsPerTick < 1 || C->slots < 1) return -1;
	for (int a = 0; a < n; a++) {
		if (a && R[a].arrival < R[a-1].arrival) return -1;                      /* arrival order */
		if (R[a].cells < 2 * (SERVE_PML + 1) + 8 || R[a].steps < 1) return -1;
		if (R[a].radius < 1 || R[a].radius > alisServeMaxRadius(R[a].cells)) return -1;
	}
	int fbits, dbits;
	alisPrecisionDefault(&fbits, &dbits);
	memset(out, 0, n * sizeof *out);
	memset(&lastStats, 0, sizeof lastStats);
	setupQueue q = {.R = R, .out = out, .n = n, .depth = C->slots};
	pthread_mutex_init(&q.m, 0);
	pthread_condattr_t ca;
	pthread_condattr_init(&ca);
	pthread_condattr_setclock(&ca, CLOCK_MONOTONIC);
	pthread_cond_init(&q.cv, &ca);
	pthread_condattr_destroy(&ca);
	q.slot = calloc(q.depth, sizeof *q.slot);
	active *A = calloc(n ? n : 1, sizeof *A);
	world *Ws = calloc(n ? n : 1, sizeof *Ws);
	phaser *Ps = calloc(n ? n : 1, sizeof *Ps);
	int finished = 0, na = 0;
	double used = 0, stepping = 0;
	long steps = 0;
	q.t0 = nowSec();
	pthread_t setup;
	if (pthread_create(&setup, 0, setupThread, &q)) return free(q.slot), free(A), free(Ws), free(Ps), -1;
	while (finished < n) {
		/* admit built requests in arrival order while a slot is free and they fit (always
		 * one on an idle device); wait for one when nothing runs */
		int joined = 0;
		pthread_mutex_lock(&q.m);
		for (;;) {
			while (q.taken < q.built && na < C->slots) {
				const double need = requestBytes(&R[q.taken], fbits);
				if (na && used + need > C->budget) break;
				active *a = &A[na++];
				*a = q.slot[q.taken % q.depth];
				a->bytes = need;
				used += need;
				out[a->id].admitted = nowSec() - q.t0;
				q.taken++;
				joined = 1;
				pthread_cond_broadcast(&q.cv);
			}
			if (na) break;
			pthread_cond_wait(&q.cv, &q.m);
		}
		pthread_mutex_unlock(&q.m);
		int k = C->stepsPerTick;
		for (int a = 0; a < na; a++) if (A[a].steps - A[a].done < k) k = A[a].steps - A[a].done;
		for (int a = 0; a < na; a++) Ws[a] = A[a].W, Ps[a] = A[a].P;
		const double s0 = nowSec();
		int s = 0;
		while (s < k) {
			updateWorldsE(Ws, na);
			updateWorldsH(Ws, na);
			updatePhaserArrays(Ws, Ps, na);
			s++;
			/* a built request and a free slot: admit it before the next step */
			if (na < C->slots && __atomic_load_n(&q.built, __ATOMIC_ACQUIRE) > q.taken) break;
		}
		stepping += nowSec() - s0;
		steps += s;
		if (joined) {
			/* the new worlds are committed now: replace estimates by the real footprint */
			used = 0;
			for (int a = 0; a < na; a++) used += A[a].bytes = deviceBytes(A[a].W);
		}
		for (int a = 0; a < na; a++) {
			A[a].done += s;
			if (out[A[a].id].batchMax < na) out[A[a].id].batchMax = na;
			out[A[a].id].bytes = A[a].bytes;
		}
		/* retire finished requests */
		for (int a = 0; a < na; ) {
			if (A[a].done < A[a].steps) { a++; continue; }
			alisResult *r = &out[A[a].id];
			for (int f = 0; f < 16 && A[a].P[f].f; f++) {
				r->spectrum[f] = phaserPoyntingOut(&A[a].P[f]);
				r->wavelength[f] = 1 / A[a].P[f].f;
			}
			r->finished = nowSec() - q.t0;
			r->latency = r->finished - R[A[a].id].arrival;
			r->cellSteps = (double)A[a].W->iNUM * A[a].W->jNUM * A[a].W->kNUM * A[a].steps;
			deletePhaser(A[a].P);
			deleteWorld(A[a].W);
			used -= A[a].bytes;
			A[a] = A[--na];
			finished++;
		}
	}
	pthread_mutex_lock(&q.m);
	q.stop = 1;
	pthread_cond_broadcast(&q.cv);
	pthread_mutex_unlock(&q.m);
	pthread_join(setup, 0);
	lastStats.makespan = nowSec() - q.t0;
	lastStats.stepping = stepping;
	lastStats.setup = q.setup;
	lastStats.steps = steps;
	pthread_cond_destroy(&q.cv);
	pthread_mutex_destroy(&q.m);
	free(q.slot);
	free(A);
	free(Ws);
	free(Ps);
	return 0;
}
