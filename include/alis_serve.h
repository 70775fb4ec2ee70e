/*
 * Serving scattering solves with continuous batching (libalis.so).  No counterpart in the
 * reference, whose programs step one world each (src/update.c:275-341); a request here is a
 * whole alis.h program -- createWorld, putObjects, planewave, createPhasers, a time loop of
 * updateE/updateH/updatePhaser, phaserPoyntingOut -- run for a caller that submits many.
 *
 * A request is a cubic grid of edge `cells` (an 8-cell CPML on every face included,
 * dx = 2 nm) holding a sphere of `radius` cells (at most alisServeMaxRadius(cells)) with a
 * Drude pole of plasma energy `plasma` eV and two Lorentz poles, lit by a 400-900 nm plane
 * wave on a TF/SF box around it, with a phaser box of 16 wavelengths (385-910 nm) outside
 * that, stepped `steps` times.  FP32 fields with FP64 DFT accumulators unless ALIS_PRECISION
 * says otherwise.
 *
 * alisServe runs a list of requests sorted by arrival time (seconds after the call).  A
 * setup thread builds each request's world on the host once it has arrived (in arrival
 * order, at most `slots` ahead of admission).  A built request is admitted, between any two
 * steps, once a slot is free and the device bytes of all admitted requests stay within the
 * budget (the first request on an idle device is always admitted); all admitted requests
 * are stepped together in shared launches (updateWorldsE/updateWorldsH), and finished ones
 * are retired at least every `stepsPerTick` steps.  Fields stay on the device from
 * admission to retirement.  Each result is bit for bit the one the request gets when
 * served alone.  alisServe itself is not reentrant: one call at a time per process.
 */
#ifndef ALIS_SERVE_H
#define ALIS_SERVE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
	double arrival;             /* seconds after the start of alisServe */
	int cells, steps;
	float radius, plasma;       /* cells, eV */
} alisRequest;

typedef struct {
	double ready;                         /* seconds after the start: its world is built */
	double admitted, finished, latency;   /* seconds after the start; finished - arrival */
	double cellSteps;                     /* grid cells x steps */
	double bytes;                         /* device bytes of the request */
	int batchMax;                         /* most requests it shared the device with, itself included */
	float wavelength[16];                 /* nm */
	float spectrum[16];                   /* power scattered out of the phaser box */
} alisResult;

typedef struct {
	double budget;              /* device bytes for all admitted requests */
	int slots;                  /* most requests admitted at once */
	int stepsPerTick;           /* most steps between admission points */
} alisServeConfig;

/* Serve R[0..n-1]; results in out[0..n-1].  0, or -1 for an invalid list or configuration. */
int alisServe(const alisRequest *R, int n, const alisServeConfig *C, alisResult *out);

typedef struct {
	double makespan;            /* seconds from the start to the last retirement */
	double stepping;            /* seconds the serving thread spent in shared steps */
	double setup;               /* seconds the setup thread spent building worlds */
	long steps;                 /* shared steps taken */
} alisServeStats;

/* Statistics of the last alisServe call. */
void alisServeLastStats(alisServeStats *s);

/* Largest sphere radius (cells) a request of grid edge `cells` can hold. */
int alisServeMaxRadius(int cells);

/* The seeded serving load: Poisson arrivals of mean `rate` per second, grid edges drawn
 * from {64, 96, 128, 192}, radii from 10 to min(40, alisServeMaxRadius), 1000-6000 steps,
 * Drude plasma energies 8-9.5 eV; every draw from one splitmix64 stream of `seed`. */
void alisServeRequests(unsigned long long seed, int n, double rate, alisRequest *R);

#ifdef __cplusplus
}
#endif
#endif
