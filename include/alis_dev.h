/*
 * alis_dev.h -- C ABI between the host library (alis.h) and the MI355X engine.
 *
 * The host library describes a world with plain C types (alis_dev_world_desc: extents,
 * boundary codes, Bloch phases, coefficient arrays and the host arrays that receive device
 * state) and commits it.  A committed world lives in HBM: fields, CPML slabs, ADE currents,
 * coefficients, sources, TF/SF strips, ghost-copy programs and phaser accumulators.  The
 * host library drives it half step by half step in the reference's order; every call is
 * asynchronous on the world's stream except the sync/read calls.  Functions return 0 on
 * success or a negative code; alis_dev_error() describes the last failure.
 *
 * Array conventions: a "host 3-D array" is a float*** over a box, k fastest, whose elements
 * form one contiguous block starting at [i0][j0][k0] (the layout of the reference's
 * makeField, src/world.c:58-70).  1-D coefficient pointers point at the element of index
 * MIN (they are indexed by grid index, like the reference's shifted arrays).
 */
#ifndef ALIS_DEV_H
#define ALIS_DEV_H

#ifdef __cplusplus
extern "C" {
#endif

#define ALIS_DEV_MAXPOLES 6

/* A dispersive object (src/object.c:178-226, 290-336): box, pole tables, fill factors on
 * the staggered E positions and the host J/K currents. */
typedef struct {
	int i0, i1, j0, j1, k0, k1;
	int nj, nk;                               /* Drude poles [0,nj), Lorentz/CP pairs [nj,nk) */
	float a[3];                               /* conductivity factors ax, ay, az */
	float JJ[3][ALIS_DEV_MAXPOLES], JK[3][ALIS_DEV_MAXPOLES], JE[3][ALIS_DEV_MAXPOLES];
	float KJ[3][ALIS_DEV_MAXPOLES], KK[3][ALIS_DEV_MAXPOLES], KE[3][ALIS_DEV_MAXPOLES];
	float ***fill[3];
	float ***J[2][3][ALIS_DEV_MAXPOLES], ***K[2][3][ALIS_DEV_MAXPOLES];   /* [m][axis][pole] */
} alis_dev_ade;

/* A source of the E (h=0) or H (h=1) chain (src/alis_c.h:305-312): target component, box,
 * scalar or per-cell amplitude, and whether its waveform has per-cell phases. */
typedef struct {
	int h, m, axis;
	int i0, i1, j0, j1, k0, k1;
	float amplitude;
	float ***Amplitude;
	int cellPhase;
} alis_dev_src;

/* A total-field/scattered-field box (src/alis_c.h:313-318). */
typedef struct {
	int iMin, iMax, jMin, jMax, kMin, kMax, iMIN, iMAX, jMIN, jMAX, kMIN, kMAX;
} alis_dev_tfsf;

typedef struct {
	int iMIN, iMAX, jMIN, jMAX, kMIN, kMAX;   /* grid including CPML */
	int iMin, iMax, jMin, jMax, kMin, kMax;   /* domain */
	int iNUM, jNUM, kNUM;
	int xMinSur, xMaxSur, yMinSur, yMaxSur, zMinSur, zMaxSur;
	int complexField;
	float coskx, sinkx, cosky, sinky, coskz, sinkz, coskyp, coskym, sinkyp, sinkym;
	const float *dt[2][3];                    /* [E,H][axis] dt/d */
	const float *cpml[2][3][3];               /* [E,H][axis][b, c_first, c_second] or NULL */
	float ***eps[3];                          /* 1/eps of the E grid (extents MIN..MAX(+1)), or NULL */
	float ***field[2][2][3];                  /* host fields [E,H][re,im][axis], ghost layer included */
	float ***slab[2][2][6][2];                /* host CPML slabs [E,H][m][face xMin yMin zMin xMax yMax zMax][a,b] */
	int nobj;
	const alis_dev_ade *obj;
	int nsrc[2];
	const alis_dev_src *src[2];
	int ntfsf;
	const alis_dev_tfsf *tfsf;
	int expandX, expandY;                     /* TF/SF auxiliary world: rows alias beyond iMAX/jMAX */
} alis_dev_world_desc;

/* A createPhasers array (src/phaser.c:21-66): nf frequencies over one box; host cfield data
 * starts per [frequency][face xMin xMax yMin yMax zMin zMax][component E1 E2 H1 H2] as
 * interleaved re/im floats, NULL for absent faces. */
typedef struct {
	int nf;
	int iMin, iMax, jMin, jMax, kMin, kMax;
	float **cf;                               /* nf*6*4 pointers */
} alis_dev_phaser_desc;

typedef struct alis_dev_world alis_dev_world;

const char *alis_dev_error(void);

/* Upload a described world (setup arrays and current host state) and return its device
 * world.  fp64 selects FP64 fields/state; dft64 selects FP64 phaser accumulators.  TF/SF
 * auxiliary worlds are committed separately and linked with alis_dev_link_tfsf.
 * Replaces the allocation side of createWorld/putObjectArray/pointDipole/createTFSF
 * (src/world.c:261-280, src/object.c:339-405, src/source.c:87-124,269-334).  The
 * descriptor's arrays must stay valid until the world is released. */
int alis_dev_commit(const alis_dev_world_desc *W, int fp64, int dft64, alis_dev_world **out);
int alis_dev_link_tfsf(alis_dev_world *D, int s, alis_dev_world *aux);
void alis_dev_release(alis_dev_world *D);

/* Half-step core: for E (h=0) conductivity scaling and -sum J of dispersive objects, the
 * curl of H and the CPML slabs (src/update.c:87-125 with Alt=+1); for H (h=1) the curl of
 * E with CPML (Alt=-1) followed by the ADE current update on the new E
 * (src/update.c:128-158).  Both real and imaginary parts of a Bloch world.
 * Fused steps (opt-in: the environment variable ALIS_FUSE=1 for worlds large enough to
 * fill the GPU, ALIS_FUSE=force for any world; never a TF/SF auxiliary world or a world in a
 * batch): the E core and its alis_dev_ghosts are deferred, and the H core runs both halves
 * in one pass over the grid (k_step) followed by the E half's end-of-step program and a
 * recomputation of H where that program changed E; the results are those of the separate
 * half steps, bit for bit.  Any other call on the world (or on a world that reads it) first
 * completes a deferred E half step the separate way. */
int alis_dev_core(alis_dev_world *D, int h);

/* Face-strip corrections of the s-th TF/SF box from its auxiliary world
 * (src/update.c:194-272).  Recorded: applied by the next alis_dev_ghosts of the same half
 * step, in the reference's order (strips of every box, then sources, then ghosts). */
int alis_dev_strips(alis_dev_world *D, int s, int h);

/* Add amplitude*wave to the box of the index-th source of the E (h=0) or H (h=1) chain
 * (src/update.c:161-179).  cellWave, when not NULL, gives one wave value per box cell
 * (k fastest) for sources with per-cell phase; otherwise `wave` applies to all cells.
 * Recorded like alis_dev_strips; sources not recorded in a half step are inactive. */
int alis_dev_inject(alis_dev_world *D, int h, int index, float wave, const float *cellWave, long ncell);

/* End of the E (h=0) or H (h=1) half step: the recorded strips and sources, then the
 * ghost-layer copies -- periodic, Bloch, PEC, symmetry and hexagonal boundaries
 * (src/update.c:182-191 with the schedules at :284-297 and :311-339). */
int alis_dev_ghosts(alis_dev_world *D, int h);

/* Copy device state to the host arrays of the descriptor: everything (fields, CPML slabs,
 * ADE currents) when i0 > i1, else E/H inside the box and the currents of objects meeting it. */
int alis_dev_sync(alis_dev_world *D, int i0, int i1, int j0, int j1, int k0, int k1);

/* Probe copies: for e = 0..n-1, the cells of box[7e..7e+5] = {i0,i1,j0,j1,k0,k1} of world
 * D[e] in the arrays that box[7e+6] selects -- bit h*6 + m*3 + axis for field array
 * (E or H, part m, axis), ALIS_DEV_PROBE_CURRENTS for the currents of its dispersive
 * objects -- to the host arrays of the descriptor: one gather launch and one copy for all
 * of them, one wait.  Serves get() (src/output.c:50-53), which reads a few cells per call. */
#define ALIS_DEV_PROBE_CURRENTS (1 << 12)
#define ALIS_DEV_PROBE_ALL ((1 << 13) - 1)
int alis_dev_sync_boxes(alis_dev_world *const *D, const int *box, int n);

/* Read / overwrite one whole field array (h: 0 E, 1 H; m: real/imag; axis 0..2) as doubles
 * in the layout [iMIN-1..iMAX+1][jMIN-1..jMAX+1][kMIN-1..kMAX+1], k fastest. */
int alis_dev_read_field(alis_dev_world *D, int h, int m, int axis, double *out);
int alis_dev_write_field(alis_dev_world *D, int h, int m, int axis, const double *in);

/* Phasers (src/phaser.c:21-110): register an array (accumulators start from the host
 * cfields), accumulate one sample into frequencies q0..q0+nq-1 (expE/expH: nq complex phase
 * factors e^{i 2 pi f t} at the E and H times as re/im pairs), copy accumulators to the
 * host cfields, read them as doubles, remove.  Read layout: [frequency][component E1 E2 H1
 * H2][cell][re,im] with the cells of the present faces concatenated in the order xMin, xMax,
 * yMin, yMax, zMin, zMax (x faces (j,k), y faces (i,k), z faces (i,j), last index fastest).
 * Float accumulators add each sample plainly, as the reference does; FP64 accumulators add
 * with Neumaier compensation, and every entry that reads them (sync, read, flux sums, far
 * field) first folds the compensations into the sums. */
int alis_dev_phaser_add(alis_dev_world *D, const alis_dev_phaser_desc *P, int *id);
int alis_dev_phaser_update(alis_dev_world *D, int id, int q0, int nq, const double *expE, const double *expH);
/* One sample into n phasers at once (entry e: world D[e], phaser id[e], frequencies
 * q0[e]..q0[e]+nq[e]-1 with factors expE[e]/expH[e]) in one launch, with the results of
 * alis_dev_phaser_update on each; the worlds join D[0]'s stream and share their precision.
 * No counterpart in the reference (one updatePhaser per phaser, src/phaser.c:83-110). */
int alis_dev_phaser_update_many(alis_dev_world *const *D, const int *id, const int *q0, const int *nq,
                                const double *const *expE, const double *const *expH, int n);
int alis_dev_phaser_sync(alis_dev_world *D, int id);
long alis_dev_phaser_size(alis_dev_world *D, int id);
int alis_dev_phaser_read(alis_dev_world *D, int id, double *out, long n);
/* Inner sums of the net flux of frequency q of phaser id (src/phaser.c:142-223 — the
 * trapezoid loops of phaserPoyntingX/Y/Z; phaserPoyntingOut, src/phaser.c:228-245, adds them
 * with the outer weights on the host).  nterm (1..12) terms of 8 ints each: kind (0: sums
 * along the rows of a face, 64 in-order partial sums per row, out[o + a*64 + l] for lane l
 * adding cells b = l, l+64, ...; 1: one sum per column, out[o + b], adding rows a = 0.. in
 * order), the face's first cell, rows na, row length nb (cell a*nb + b), the components x and
 * y (0..3, the device read layout), the offset of the weights in w (per row cell b for kind
 * 0, per row a for kind 1) and the output offset o.  Each sum adds
 * w * (Re x Re y + Im x Im y) in double.  Returns -1 for a term outside the faces or the
 * buffers. */
int alis_dev_phaser_flux_sums(alis_dev_world *D, int id, int q, int nterm, const int *terms,
                              const double *w, long nw, double *out, long nout);
/* The same sums for the nq frequencies q0 .. q0+nq-1 of the array in one launch: out holds
 * nq blocks of nout doubles, block q - q0 the outputs of frequency q. */
int alis_dev_phaser_flux_sums_range(alis_dev_world *D, int id, int q0, int nq, int nterm, const int *terms,
                                    const double *w, long nw, double *out, long nout);
/* The same sums, allocating nothing: no counterpart in the reference; for callers that
 * cannot allocate device memory (code under graph capture, tight serving loops).  Outputs,
 * weights and the term table live in a file-scope device array of 4 MiB per GPU (held until
 * the process ends), so nout + nw + (9 * nterm + 1) / 2 must not exceed 524288 doubles
 * (-1 otherwise).  Every call on one GPU shares that array: calls that overlap on one GPU
 * race.  Results are those of alis_dev_phaser_flux_sums; outputs that no term writes are
 * zero in both. */
int alis_dev_phaser_flux_sums_noalloc(alis_dev_world *D, int id, int q, int nterm, const int *terms,
                                      const double *w, long nw, double *out, long nout);
int alis_dev_phaser_remove(alis_dev_world *D, int id);

/* Near-to-far projection, the surface sums (src/farfld.c:248-459).  For each of n
 * directions ("columns"), every component of the DFT field at frequency q on every
 * PML-bounded face of the phaser box, summed against two weight vectors along the face's
 * axes: in-plane phase factors d e^{i k p} (with the mirrored term of a symmetry plane) and,
 * on x and y faces, the discrete plane-wave profile T(z) of the direction.  The host
 * library computes the profiles (a 1-D recursion per polar angle, src/farfld.c:126-245)
 * and assembles the far field from the sums (src/farfld.c:462-514).
 *
 * Axis tables run over the box (index 0 = its first cell): E and H cell widths, node and
 * half-node (i - 1/2) positions.  Profiles: nslot slots of 6 components (Ex Ey Ez Hx Hy Hz)
 * of nz complex values for absolute k = kLo .. kLo+nz-1.  A column gives (kx, ky) and, per
 * polarisation, the slot and for each T component the slot component it comes from and a
 * real factor (the azimuthal rotation; factor 0 for an absent component).
 * Output S[col][face kind 0..5][component E1 E2 H1 H2][polarisation 0..1] as re/im pairs:
 * on z faces the sums do not involve T and only polarisation 0 is written; absent faces and
 * unwritten entries are zero. */
typedef struct {
	const double *dxE, *dxH, *xN, *xH;
	const double *dyE, *dyH, *yN, *yH;
	const double *dzE, *dzH;
	int symX, symY, symZ;       /* a symmetry plane at the min face */
	double coskx, cosky, coskz;
	int nslot, kLo, nz;
	const double *T;            /* [slot][component][nz][re,im] */
} alis_dev_ff_geom;

typedef struct {
	double kx, ky;
	int slot[2];
	int src[2][6];
	double fac[2][6];
} alis_dev_ff_col;

int alis_dev_farfield(alis_dev_world *D, int id, int q, const alis_dev_ff_geom *g, int n,
                      const alis_dev_ff_col *cols, double *S);

/* Batches: several worlds -- with the TF/SF auxiliary worlds of each, which the caller
 * lists too -- stepped in shared launches.  All worlds join the first one's stream.  A half
 * step of a batch is alis_dev_batch_core (curl, CPML and ADE of every world in one launch),
 * then the usual alis_dev_strips / alis_dev_inject records per world, then
 * alis_dev_batch_post (each world's strips, sources and ghost copies; one launch per
 * program round).  Worlds keep their own state; results equal stepping them one by one.
 * A world may belong to several batches; release a batch before any of its worlds.
 * No counterpart in the reference, which steps one world per call (src/update.c:275-341). */
typedef struct alis_dev_batch alis_dev_batch;
int alis_dev_batch_create(alis_dev_world *const *D, int n, alis_dev_batch **out);
int alis_dev_batch_core(alis_dev_batch *B, int h);
int alis_dev_batch_post(alis_dev_batch *B, int h);
void alis_dev_batch_release(alis_dev_batch *B);

/* Per-row sums over a plane of the fields (real parts) for the instantaneous Poynting flux
 * poyntingX/Y/Z (src/output.c:233-332), so the host needs no copy of the world: nterm terms of
 * 11 ints -- normal axis, plane index, E component, H component, row axis, first and last row,
 * first and last inner index (the inner axis is the remaining one), weight offset, output
 * offset.  out[offset + r - first row] = sum over inner s in order of
 * w[offset + s - first] * E(cell) * (H(cell) + H(cell one step along the normal)), the H pair
 * added in the field precision, the rest in double.  Outputs no term writes are zero. */
int alis_dev_plane_sums(alis_dev_world *D, int nterm, const int *terms, const double *w, long nw, double *out, long nout);

/* Column sums of the power absorbed by the n-th dispersive object (objectAbsorption,
 * src/output.c:375-435; the current of src/output.c:25-47): for each column (i, j) of the box
 * (inside the object's box), over k = k0..k1 in order, the terms dz E_q current_q of q = x, y,
 * z with dz from dzE (x, y) or dzH (z), indexed k - k0, skipped where 1/eps is 0, the terms in
 * the field precision and the sums in double.  out[9 ((i-i0)(j1-j0+1) + j-j0) + ...]: the three
 * sums, the three terms at k0, the three terms at k1. */
int alis_dev_absorption(alis_dev_world *D, int n, int i0, int i1, int j0, int j1, int k0, int k1,
                        const double *dzE, const double *dzH, double *out);

/* Probe spectra: the running DFT of writeSpectrum (src/output.c:177-230, whose loop at
 * :212-216 adds one sample of every probe to every frequency bin per call) in FP64 with
 * compensated summation.  A spectrum has nprobe series and nbin bins; the sample of scaled
 * time tau adds v[p] e^{i 2 pi tau (nTop - b)} to bin b of probe p.  Samples are staged on
 * the host and added on the device in call order, 256 at a time, on a stream of their own
 * (no world is involved); read flushes the staged samples and returns the sums as
 * out[(p nbin + b) 2 + re/im].  Replaces the float complex accumulation of the reference. */
typedef struct alis_dev_spectrum alis_dev_spectrum;
int alis_dev_spectrum_create(int nprobe, int nbin, int nTop, alis_dev_spectrum **out);
int alis_dev_spectrum_add(alis_dev_spectrum *S, double tau, const double *v);
int alis_dev_spectrum_read(alis_dev_spectrum *S, double *out);
void alis_dev_spectrum_release(alis_dev_spectrum *S);

/* Second-order (three-wave) coupling, nonLinearSHG / nonLinearTWM (src/update.c:344-411):
 * over the cells of the n-th dispersive object of S (its box without the upper plane), for
 * each part, the E field of S gains 2 alt d[a][b] (1/eps)(fill) times products of the pump
 * worlds' E components (P1, P2; averaged to the component's position), d being 18 floats
 * [3][6], with the reference's float and double promotions.  S, P1 and P2 share one grid
 * and precision, and S is neither pump; the pumps join S's stream. */
int alis_dev_three_wave(alis_dev_world *S, alis_dev_world *P1, alis_dev_world *P2, int n, const float *d, float alt);

/* Wait for all work of D. */
int alis_dev_synchronize(alis_dev_world *D);

/* Timing of the curl kernels with device events on the world's stream: enable (resets the
 * totals) or disable, then read total milliseconds and launch count per half (h = 0 E,
 * 1 H) or of the fused step kernel (h = 2, see alis_dev_core). */
int alis_dev_timing(alis_dev_world *D, int enable);
int alis_dev_timing_read(alis_dev_world *D, int h, double *ms, long *launches);

/* Average bytes per grid cell that the single-world E curl reads for 1/eps: 3 field-sized
 * values in row segments (64 cells along z) that hold more than one (x, y, z) triple, one
 * table entry per segment, 0 without objects.  Measurement helper (the algorithmic bytes of
 * the E launch); no counterpart in the reference. */
double alis_dev_eps_bytes(const alis_dev_world *D);

/* Device-resident bytes of D: fields, coefficients, CPML and ADE state, source tables,
 * phaser accumulators and staging (not the TF/SF auxiliary worlds, which have their own). */
long alis_dev_bytes(alis_dev_world *D);

#ifdef __cplusplus
}
#endif
#endif
