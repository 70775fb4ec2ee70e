// MI355X FDTD engine: device-resident worlds and the half-step programs that the host
// library (alis.h) drives.  One world = one HIP stream; every launch is asynchronous.
//
// Layout in HBM.  Every field component is its own array over
// [iMIN-1..iMAX+1] x [jMIN-1..jMAX+1] x [kMIN-1..kMAX+1] (ghost layer included), z
// contiguous, each (i,j) row padded to a multiple of 128 bytes and shifted so that k = kMIN
// starts on a 128-byte boundary: a 64-lane wavefront reading 64 consecutive k touches
// exactly two cache lines.  CPML slabs normal to x and y use the same padded rows; slabs
// normal to z, ADE currents and fill factors are compact over their boxes.
//
// Arithmetic.  Kernels keep the reference's operation order per cell (curl, then the CPML
// slabs in face order xMin,yMin,zMin,xMax,yMax,zMax; ADE scaling, then -J per pole; strips
// and sources in chain order) and are compiled without contraction, so FP32 results are
// bitwise those of the reference's float code.  FP64 worlds run the same programs on
// doubles (coefficients widened from the float setup).
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <vector>
#include <string>
#include <cstring>
#include <cstdio>
#include <cstdint>
#include <cmath>
#include <algorithm>
#include <map>
#include <mutex>

#include "alis_dev.h"

namespace {

// Boundary codes of the reference (src/alis_c.h:15-21).
enum { SUR_SYM = 1, SUR_PEC = 2, SUR_PML = 3, SUR_PBC = 4, SUR_HBC = 5, SUR_BBC = 6 };

std::string g_error;

int fail(const char *what, hipError_t e)
{
	g_error = std::string(what) + ": " + hipGetErrorString(e);
	return -1;
}
int fail(const std::string &what)
{
	g_error = what;
	return -1;
}

#define HIPCHECK(call) do { hipError_t e_ = (call); if (e_ != hipSuccess) return fail(#call, e_); } while (0)

#ifndef TPB_TILE
#define TPB_TILE 1
#endif
#ifndef K_SHUFFLE
#define K_SHUFFLE true       // k_curl: k neighbours passed between lanes (see curl_cell.inc, KSH)
#endif
constexpr int ROW_ALIGN_BYTES = 128;
constexpr int MAXPOLE = ALIS_DEV_MAXPOLES;

// Index map of one device array: (i - i0)*si + (j - j0)*sj + (k - k0) + koff.
struct Lay {
	int i0, j0, k0, koff;
	long si, sj;
	__host__ __device__ long at(int i, int j, int k) const { return (long)(i - i0) * si + (long)(j - j0) * sj + (k - k0) + koff; }
};

// Lay::at for a layout read in place from a table in global memory.
template<class L> __device__ inline long layAt(L &lay, int i, int j, int k)
{
	return (long)(i - lay.i0) * lay.si + (long)(j - lay.j0) * lay.sj + (k - lay.k0) + lay.koff;
}

struct Box {
	int i0, i1, j0, j1, k0, k1;
	long count() const { return (i1 < i0 || j1 < j0 || k1 < k0) ? 0 : (long)(i1-i0+1)*(j1-j0+1)*(k1-k0+1); }
};

// ---------------------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------------------

// Four values (CPML coefficients) as one vector for a single load.
template<class T> struct Quad;
template<> struct Quad<float> { typedef float V __attribute__((vector_size(16))); };
template<> struct Quad<double> { typedef double V __attribute__((vector_size(32))); };

template<class T> struct CurlFace {
	T *pa, *pb;                 // the two auxiliary components of this face
	T *qa, *qb;                 // where their new values go (the same arrays but in k_step
	                            // and k_fix, which write the world's other set of arrays)
	const T *bc;                // recursion coefficients (b, ca, cb, 0) along the face axis
	                            // (index - MIN), four to an index for one vector load
	Lay L;
	int lo, hi, on;
};

// A dispersive object as the kernels see it (device memory, one table per world):
// its box clipped to the grid, the compact layout of its arrays, pole coefficients,
// fill factors on the staggered E positions and the J/K currents per real/imaginary part.
template<class T> struct ObjDev {
	int i0, i1, j0, j1, k0, k1;
	Lay LO;
	int NJ, NK, scale;          // scale: any conductivity factor non-zero
	T a[3];
	T JJ[3][MAXPOLE], JK[3][MAXPOLE], JE[3][MAXPOLE], KJ[3][MAXPOLE], KK[3][MAXPOLE], KE[3][MAXPOLE];
	const T *fill[3];
	T *J[2][3][MAXPOLE], *K[2][3][MAXPOLE];
};

// A z face of a phaser whose samples the H curl stores (k_curl, ROWS): of the face (i0..i1,
// j0..j1) at k, the cells (si0..si1, sj0..sj1), at their places from `base` in the ring slot
// [4][cells] (E x, E y, H x(k) + H x(k+1), H y(k) + H y(k+1)), as k_phs_sample stores them.
template<class T> struct PhsZ {
	int k, i0, i1, j0, j1;      // the face
	int si0, si1, sj0, sj1;     // the cells sampled here
	long cells, base;
	T *slot;
};
constexpr int MAXZS = 4;

template<class T> struct CurlArgs {
	T *F0, *F1, *F2;
	T *O0, *O1, *O2;            // where the new values go (F* but in k_step and k_fix)
	const T *G0, *G1, *G2;
	const T *C0, *C1, *C2;
	const typename Quad<T>::V *seg;  // k_curl E: per 64-cell row segment (tile), 1/eps
	                            // (x, y, z, 1) when the segment holds one triple, else 0s
	const T *dtx, *dty, *dtz;   // dt/d along each axis, index - MIN
	Lay L;
	int iMIN, jMIN, kMIN, ni, nj, nk;
	int alt;                    // +1: E from H, -1: H from E
	CurlFace<T> face[6];        // xMin yMin zMin xMax yMax zMax
	const ObjDev<T> *obj;       // dispersive objects (E: scaling and -sum J before the curl;
	int nobj, m;                //  H: current update on the new E), part m
	const ObjDev<T> *objOut;    // the table whose currents receive the new values (obj but
	                            //  in k_step and k_fix)
	const int *otStart;         // per 64 x 4 tile (k_curl's tiles): its objects are
	const int *otList;          //  otList[otStart[t] .. otStart[t+1]) ascending; null: all
	int flatPitch;              // batched launches of packed short rows: row length with ghosts
	int nzs;                    // H launches of k_curl: phaser z faces to sample (see PhsZ)
	PhsZ<T> zs[MAXZS];
};

// Fields and coefficients live in global memory, but to the compiler a pointer that a kernel
// loads from memory (the batched launches' argument tables) is generic, and generic (flat)
// accesses cost more than global ones (about 20 % on a batched 256^3 curl): the cell update
// reaches every array through a global-space pointer.
template<class T> __device__ inline const __attribute__((address_space(1))) T *glb(const T *p)
{
	return (const __attribute__((address_space(1))) T *)p;
}
template<class T> __device__ inline __attribute__((address_space(1))) T *glb(T *p)
{
	return (__attribute__((address_space(1))) T *)p;
}

// Streaming access: arrays touched once per launch (the updated field, 1/eps, CPML psi)
// are read and written non-temporally so they do not displace the other field's rows and
// planes, which neighbouring workgroups read again from L2.
template<bool NT, class T> __device__ inline T ldS(const T *p)
{
	if (NT) return __builtin_nontemporal_load(glb(p));
	return *glb(p);
}
template<bool NT, class T> __device__ inline void stS(T *p, T v)
{
	if (NT) __builtin_nontemporal_store(v, glb(p));
	else *glb(p) = v;
}

// Curl update of one half step plus the CPML slabs (src/update.c:48-125).  One thread per
// cell; 64 lanes along z, 4 rows along y per workgroup, one x plane per grid z index.
// NT: non-temporal access to the single-use arrays; NTG also to the centre values of the
// other field.
// BY: rows along y per workgroup (64 x BY threads).  ROWS: a one-dimensional grid that
// deals rows (the tiles along k of one block of BY rows in one x plane) to the 8 XCDs in turn,
// every tile of a row on one XCD in consecutive slots, so that the line a tile reads beyond
// its last k is fetched into the L2 that the next tile reads it from (the dispatcher deals
// consecutive workgroup ids to XCDs in turn; with the plain grid the tiles of a row land on
// different XCDs).  Measured on the 256^3 nanoblock: H 139.8 -> 137.1 us, E unchanged
// (149 us either way, so E keeps the plain grid); groups of 2 or 4 rows per XCD, which would
// also keep the j+1 row local, lose 5-11 %.

template<class T, int CM, bool NT = true, bool NTG = false, bool MZ = false, bool PC = false>
__device__ __forceinline__ void curlAt(const CurlArgs<T> &a, int i, int j, int k, T (&coef)[6]);
template<class T, int CM, bool NT, bool MZ>
__device__ __forceinline__ void curlAtGiven(const CurlArgs<T> &a, int i, int j, int k, const T (&coef)[6]);

// Cell (bx, by, bz) of a 64 x 4 tile, of part a (PAIR: then of part *b, the imaginary part of
// the same world).
template<class T, int CM, bool PAIR>
__device__ __forceinline__ void curlCell(const CurlArgs<T> &a, const CurlArgs<T> *b, int bx, int by, int bz)
{
	const int k = a.kMIN + bx * 64 + (int)threadIdx.x;
	const int j = a.jMIN + by * 4 + (int)threadIdx.y;
	const int i = a.iMIN + bz;
	if (k >= a.kMIN + a.nk || j >= a.jMIN + a.nj) return;
	T coef[6];
	curlAt<T, CM>(a, i, j, k, coef);
	if (PAIR) curlAtGiven<T, CM, true, false>(*b, i, j, k, coef);
}

// The update of cell (i, j, k) in batched launches: the CPML faces reuse the curl's
// neighbour values, and the psi slabs go through the cache (the short rows of small worlds
// cover a slab with partial lines that neighbouring tiles complete).  Measured on 64 k-point
// worlds: 5 % (E) and 14 % (H) faster than reloading and streaming; slower in k_curl.
template<class T, int CM, bool NT, bool NTG, bool MZ, bool PC>
__device__ __forceinline__ void curlAt(const CurlArgs<T> &a, int i, int j, int k, T (&coef)[6])
{
	constexpr bool SHARE = true, NTP = false, MERGEZ = MZ, UNI = !MZ, SEP = false, wr = true, NTC = NT && !PC;
	constexpr int KSH = 0;
#include "curl_cell.inc"
	coef[0] = dx, coef[1] = dy, coef[2] = dz, coef[3] = cx, coef[4] = cy, coef[5] = cz;
}

// The same for the imaginary part of a Bloch world in the thread that updated the real part:
// dt/d and 1/eps are the real part's (coef), not loaded again.
template<class T, int CM, bool NT, bool MZ>
__device__ __forceinline__ void curlAtGiven(const CurlArgs<T> &a, int i, int j, int k, const T (&coef)[6])
{
	constexpr bool NTG = false, SHARE = true, NTP = false, MERGEZ = MZ, UNI = !MZ, SEP = false, wr = true, NTC = NT;
	constexpr int KSH = 0;
#define CURL_COEF
#include "curl_cell.inc"
#undef CURL_COEF
}

template<class T, int CM, bool ROWS, bool NT = true, bool NTG = false, int BY = 4, bool ZS = false, bool KSHUF = K_SHUFFLE>
__global__ __launch_bounds__(64 * BY) void k_curl(CurlArgs<T> a)
{
	int bx = (int)blockIdx.x, by = (int)blockIdx.y, bz = (int)blockIdx.z;
	if (ROWS) {
		// workgroup w: XCD w % 8, its slot s = w / 8; row r = (s / nkb) * 8 + w % 8
		const int nkb = (a.nk + 63) / 64, njb = (a.nj + BY - 1) / BY;
		const int w = (int)blockIdx.x, s = w >> 3, r = (s / nkb) * 8 + (w & 7);
		if (r >= njb * a.ni) return;
		bx = s % nkb, by = r % njb, bz = r / njb;
	}
	const int k = a.kMIN + bx * 64 + (int)threadIdx.x;
	const int j = a.jMIN + by * BY + (int)threadIdx.y;
	const int i = a.iMIN + bz;
	if (k >= a.kMIN + a.nk || j >= a.jMIN + a.nj) return;
	constexpr bool SHARE = false, NTP = NT, MERGEZ = false, UNI = true, SEP = false, wr = true, NTC = NT;
	// E only: the H launch (previous lane, lane 0 loading) measured 137.8 -> 145.7 us with it,
	// the E launch 147.8 -> 146.6 us (256^3 nanoblock)
	constexpr int KSH = KSHUF && !ROWS ? 1 : 0;
#include "curl_cell.inc"
	// H: the samples of the phasers' z faces from the values at hand -- the new H here and at
	// k + 1 (the next lane: the host takes faces whose k + 1 lies in the same wavefront) and E,
	// final after its end-of-half-step program -- in place of k_phs_sample reading a strided
	// plane (the host takes faces that the H program does not change)
	if (ZS && ROWS) {
		for (int z = 0; z < a.nzs; z++) {
			const PhsZ<T> &Z = a.zs[z];
			// the wavefront's row and 64 cells (bx) hold a sampled cell of the face: the same
			// for all its lanes, so only these waves exchange values
			if (i < Z.si0 || i > Z.si1 || j < Z.sj0 || j > Z.sj1 || (Z.k - a.kMIN) >> 6 != bx) continue;
			const T hx1 = __shfl_down(fx, 1), hy1 = __shfl_down(fy, 1);
			if (k != Z.k) continue;
			const long t = Z.base + (long)(i - Z.i0) * (Z.j1 - Z.j0 + 1) + (j - Z.j0);
			T *const q = Z.slot;
			q[t] = gx, q[Z.cells + t] = gy, q[2 * Z.cells + t] = fx + hx1, q[3 * Z.cells + t] = fy + hy1;
		}
	}
}

// The entry owning workgroup b of a batch launch: start[] is increasing; every wave finds it
// with one load per lane and a ballot (no chain of dependent loads).
__device__ __forceinline__ int batchEntry(const int *start, int n, int b)
{
	int e = 0;
	for (int c = 0; c < n; c += 64) {
		const int q = c + (int)__lane_id();
		const unsigned long long m = __ballot(q < n && start[q] <= b);
		if (!m) break;
		e = c + __popcll(m) - 1;
		if (__popcll(m) < 64) break;
	}
	return __builtin_amdgcn_readfirstlane(e);
}

// Several worlds (and real/imaginary parts) in one launch: entry e owns workgroups
// [start[e], start[e+1]) of a one-dimensional grid; each workgroup updates TPB consecutive
// tiles of its entry (tiles as in k_curl).
// FLAT: 1 or 0 when every entry of the launch has packed rows or none has; CM: 2 when every
// entry takes 1/eps from its per-segment table (and arrays where a segment is mixed), 1 when
// every entry reads 1/eps arrays, 0 when none does; -1 to decide per entry.  Specialised
// launches compile to leaner code (64 k-point worlds: E curl 380 -> 332 us); the host runs
// one launch per class of entries.  H launches (alt < 0), which read the next plane of E, run each entry's
// tiles in descending order, so that plane was read as a centre just before (64 k-point
// worlds: H curl 331 -> 306 us; the 256^3 k_curl shows no difference).
// PAIR: every entry is both parts of a Bloch world (arguments 2e and 2e + 1, one geometry):
// each thread updates its cell of the real part and then of the imaginary part, which reads
// the same 1/eps values again, from the cache (the real part loads them cached).
template<class T, int TPB, int FLAT = -1, int CM = -1, bool PAIR = false>
__global__ __launch_bounds__(256) void k_curl_batch(const CurlArgs<T> *__restrict__ args, const int *__restrict__ start, int n)
{
	const int b = (int)blockIdx.x;
	const int e = batchEntry(start, n, b);
	const CurlArgs<T> &a = args[PAIR ? 2 * e : e];
	if (FLAT == 1 || (FLAT < 0 && a.flatPitch)) {
		// packed rows shorter than a wavefront: a tile is 256 consecutive cells of one x
		// plane from the plane's first (ghost) row on (ghost rows, row ghosts and plane
		// padding included, their lanes idle), so the lanes of a wavefront stay busy across
		// row ends; planes are padded to 256 bytes, so every wavefront's access is aligned
		const int pitch = a.flatPitch, per = (int)((a.L.si + 255) / 256), tiles = per * a.ni;
		const int lid = (int)(threadIdx.y * 64 + threadIdx.x);
		for (int t0 = (b - start[e]) * TPB, t1 = min(t0 + TPB, tiles); t0 < t1; t0++) {
			const int t = a.alt < 0 ? tiles - 1 - t0 : t0;
			const int off = (t % per) * 256 + lid, jj = off / pitch - 1, kk = off - (jj + 1) * pitch;
			if (jj < 0 || jj >= a.nj || kk < 1 || kk > a.nk) continue;
			const int i = a.iMIN + t / per, j = a.jMIN + jj, k = a.kMIN - 1 + kk;
			T coef[6];
			const int cm = CM >= 0 ? CM : a.seg ? 2 : a.C0 ? 1 : 0;
			if (cm == 2) curlAt<T, 2, true, false, true, PAIR>(a, i, j, k, coef);
			else if (cm == 1) curlAt<T, 1, true, false, true, PAIR>(a, i, j, k, coef);
			else curlAt<T, 0, true, false, true, PAIR>(a, i, j, k, coef);
			if (PAIR) {
				// the imaginary part: the same world's dt/d and 1/eps
				const CurlArgs<T> &a1 = args[2 * e + 1];
				if (cm == 2) curlAtGiven<T, 2, true, true>(a1, i, j, k, coef);
				else if (cm == 1) curlAtGiven<T, 1, true, true>(a1, i, j, k, coef);
				else curlAtGiven<T, 0, true, true>(a1, i, j, k, coef);
			}
		}
		return;
	}
	const int nkb = (a.nk + 63) / 64, njb = (a.nj + 3) / 4, tiles = nkb * njb * a.ni;
	for (int t0 = (b - start[e]) * TPB, t1 = min(t0 + TPB, tiles); t0 < t1; t0++) {
		const int t = a.alt < 0 ? tiles - 1 - t0 : t0;
		const int bx = t % nkb, by = (t / nkb) % njb, bz = t / (nkb * njb);
		if (CM == 2 || (CM < 0 && a.seg)) curlCell<T, 2, PAIR>(a, args + 2 * e + 1, bx, by, bz);
		else if (CM == 1 || (CM < 0 && a.C0)) curlCell<T, 1, PAIR>(a, args + 2 * e + 1, bx, by, bz);
		else curlCell<T, 0, PAIR>(a, args + 2 * e + 1, bx, by, bz);
	}
}

__device__ inline void boxCell(long t, int ni, int nj, int nk, int &di, int &dj, int &dk)
{
	dk = (int)(t % nk);
	long r = t / nk;
	dj = (int)(r % nj);
	di = (int)(r / nj);
	(void)ni;
}

// One time step of a world in one pass (the curl halves of src/update.c:275-341): the E
// update and then the H update on the new E, read from the world's current set of arrays
// (fields, CPML psi, dispersive currents) and written to its other set, so that no
// workgroup overwrites a value another one still reads.  A workgroup sweeps a chunk of x
// planes for R rows (each wave one row, all of its k cells, NC per lane): per plane it
// updates E, passes the new values through LDS, then updates H from them, so each field
// element comes from HBM once and goes back once per step (the separate half steps read
// the other field a second time).  The E values H needs beyond the workgroup's cells -- row
// j0 - 1 (wave 0) and plane i0 - 1 (the sweep's first iteration) -- are computed again and
// not stored; outside the update range the arrays' values are read (cells no step writes).
// The end-of-E-half program (strips, sources, ghosts) runs afterwards on the new set, and
// k_fix recomputes H on the cells whose E inputs it changed.
// Workgroups are dealt to the 8 XCDs by chunk: all row groups of a chunk run on one XCD,
// side by side, so the rows one group reads of its neighbour's come from that XCD's L2.
template<class T> struct StepArgs {
	CurlArgs<T> e, h;           // the two halves (alt +1 and -1): F, G, psi and currents of
	                            // the current set; O*, qa/qb and objOut of the other
};

// sa: the halves' arguments (device memory, read as needed rather than held in registers
// across the sweep); R rows per group; wave (c, w) of the workgroup (threadIdx.y, .z) owns
// the 64 cells from kMIN + 64 c of row j0 - 1 + w, so every wave has one cell per lane per
// half step and plane, as in k_curl; P = 64 * ceil(nk / 64) + 1 values per LDS row and
// component, whose slot 0 holds k = kMIN - 1.  LDS: two planes (this one and the one
// before, alternating) of R + 1 rows of the three E components, so one barrier per plane
// separates the E writes from the H reads.  The values a plane's E update streams from the
// current set -- E and H at the cell, H one cell up along each axis -- are loaded into
// registers a plane ahead (H at the cell two planes ahead: it is the next plane's i + 1
// neighbour first), so their latency passes while the plane before is updated.
template<class T, int CME>
__global__ __launch_bounds__(1024, sizeof(T) == 4 ? 8 : 4) void k_step(const StepArgs<T> *__restrict__ sa, int R, int P, int chunk, int nchunk, int ngroup)
{
	extern __shared__ __align__(16) unsigned char stepLds[];
	const int lane = (int)threadIdx.x, c = (int)threadIdx.y, w = (int)threadIdx.z;
	const int b = (int)blockIdx.x, slot = b >> 3;
	const int g = slot % ngroup, q = (b & 7) + 8 * (slot / ngroup);
	if (q >= nchunk) return;
	const int iMIN = sa->e.iMIN, jMIN = sa->e.jMIN, kMIN = sa->e.kMIN, nj = sa->e.nj, nk = sa->e.nk;
	const int j = jMIN + g * R - 1 + w, k = kMIN + c * 64 + lane, x = 1 + c * 64 + lane;
	const int ib = iMIN + q * chunk, ie = min(ib + chunk, iMIN + sa->e.ni) - 1;
	const bool rowIn = j >= jMIN && j < jMIN + nj, live = j < jMIN + nj && k < kMIN + nk;
	const int plane = (R + 1) * 3 * P;
	const Lay L = sa->e.L;
	const T *const E0 = sa->e.F0, *const E1 = sa->e.F1, *const E2 = sa->e.F2;
	const T *const H0 = sa->e.G0, *const H1 = sa->e.G1, *const H2 = sa->e.G2;
	// pipeline: e (E at plane i), h (H at i), u (H at i + 1), n (H x, z at j + 1 and x, y at
	// k + 1, plane i), s (E at k = kMIN - 1, plane i; lane x == 1); the next plane's in p*
	T e0 = 0, e1 = 0, e2 = 0, h0 = 0, h1 = 0, h2 = 0, u0 = 0, u1 = 0, u2 = 0;
	T n0 = 0, n1 = 0, n2 = 0, n3 = 0, s0 = 0, s1 = 0, s2 = 0;
	auto loadE = [&](int ii, T &a0, T &a1, T &a2) {
		const long id = L.at(ii, j, k);
		a0 = ldS<true>(E0 + id), a1 = ldS<true>(E1 + id), a2 = ldS<true>(E2 + id);
	};
	auto loadH = [&](int ii, T &a0, T &a1, T &a2) {
		const long id = L.at(ii, j, k);
		a0 = *glb(H0 + id), a1 = *glb(H1 + id), a2 = *glb(H2 + id);
	};
	auto loadN = [&](int ii, T &a0, T &a1, T &a2, T &a3) {
		const long id = L.at(ii, j, k);
		a0 = *glb(H0 + id + L.sj), a1 = *glb(H2 + id + L.sj), a2 = *glb(H0 + id + 1), a3 = *glb(H1 + id + 1);
	};
	auto loadS = [&](int ii, T &a0, T &a1, T &a2) {
		const long id = L.at(ii, j, kMIN - 1);
		a0 = *glb(E0 + id), a1 = *glb(E1 + id), a2 = *glb(E2 + id);
	};
	if (live) {
		loadE(ib - 1, e0, e1, e2);
		if (x == 1) loadS(ib - 1, s0, s1, s2);
		if (rowIn) {
			loadH(ib - 1, h0, h1, h2), loadH(ib, u0, u1, u2), loadN(ib - 1, n0, n1, n2, n3);
		}
	}
	for (int i = ib - 1; i <= ie; i++) {
		T pe0 = 0, pe1 = 0, pe2 = 0, pu0 = 0, pu1 = 0, pu2 = 0, pn0 = 0, pn1 = 0, pn2 = 0, pn3 = 0, ps0 = 0, ps1 = 0, ps2 = 0;
		if (live && i < ie) {
			loadE(i + 1, pe0, pe1, pe2);
			if (x == 1) loadS(i + 1, ps0, ps1, ps2);
			if (rowIn) {
				loadN(i + 1, pn0, pn1, pn2, pn3);
				if (i + 2 <= ie + 1) loadH(i + 2, pu0, pu1, pu2);
			}
		}
		T *const mine = (T *)stepLds + (i & 1) * plane + w * 3 * P;
		const bool own = w > 0 && rowIn && i >= ib;
		if (live) {
			T ex = e0, ey = e1, ez = e2;
			if (rowIn && i >= iMIN) {
				const T given0 = e0, given1 = e1, given2 = e2;
				const T gx = h0, gy = h1, gz = h2;
				const T g0j = n0, g2j = n1, g0k = n2, g1k = n3, g1i = u1, g2i = u2;
				const CurlArgs<T> &a = sa->e;
				const bool wr = own;
				constexpr int CM = CME;
				constexpr bool NT = true, NTG = false, SHARE = true, NTP = true, MERGEZ = false, UNI = true, SEP = true, NTC = true;
				constexpr int KSH = 0;
#define CURL_GIVEN
#define CURL_GIVEN_F
#include "curl_cell.inc"
#undef CURL_GIVEN_F
#undef CURL_GIVEN
				ex = fx, ey = fy, ez = fz;
			}
			mine[x] = ex, mine[P + x] = ey, mine[2 * P + x] = ez;
			if (x == 1) mine[0] = s0, mine[P] = s1, mine[2 * P] = s2;
		}
		__syncthreads();
		if (own && live) {
			const T *const below = mine - 3 * P;
			const T *const before = (T *)stepLds + ((i + 1) & 1) * plane + w * 3 * P;   // plane i - 1
			const T given0 = h0, given1 = h1, given2 = h2;
			const T gx = mine[x], gy = mine[P + x], gz = mine[2 * P + x];
			const T g0k = mine[x - 1], g1k = mine[P + x - 1];
			const T g0j = below[x], g2j = below[2 * P + x];
			const T g1i = before[P + x], g2i = before[2 * P + x];
			const CurlArgs<T> &a = sa->h;
			constexpr int CM = 0;
			constexpr bool NT = true, NTG = false, SHARE = true, NTP = true, MERGEZ = false, UNI = true, SEP = true, wr = true, NTC = true;
			constexpr int KSH = 0;
#define CURL_GIVEN
#define CURL_GIVEN_F
#include "curl_cell.inc"
#undef CURL_GIVEN_F
#undef CURL_GIVEN
		}
		e0 = pe0, e1 = pe1, e2 = pe2, h0 = u0, h1 = u1, h2 = u2, u0 = pu0, u1 = pu1, u2 = pu2;
		n0 = pn0, n1 = pn1, n2 = pn2, n3 = pn3, s0 = ps0, s1 = ps1, s2 = ps2;
	}
}

// Cells copied from one set of a world's arrays to the other (see the carry lists of
// buildPost): one thread per cell of a list of boxes.
constexpr int MAXCARRY = 16;
template<class T> struct CarryArgs {
	const T *src[MAXCARRY];
	T *dst[MAXCARRY];
	Box b[MAXCARRY];
	long off[MAXCARRY + 1];
	Lay L;
	int n;
};

template<class T>
__global__ __launch_bounds__(256) void k_carry(CarryArgs<T> a)
{
	const long t = (long)blockIdx.x * 256 + threadIdx.x;
	if (t >= a.off[a.n]) return;
	int q = 0;
	while (q + 1 < a.n && a.off[q + 1] <= t) q++;
	const Box &B = a.b[q];
	int di, dj, dk;
	boxCell(t - a.off[q], B.i1 - B.i0 + 1, B.j1 - B.j0 + 1, B.k1 - B.k0 + 1, di, dj, dk);
	const long id = a.L.at(B.i0 + di, B.j0 + dj, B.k0 + dk);
	a.dst[q][id] = a.src[q][id];
}

// H again on the cells whose E inputs the end-of-E-half program changed after k_step (the
// boxes of its operations, one cell further up each axis): from the old set's H, psi and
// currents and the new set's final E, into the new set -- the values the separate half
// steps give.  One thread per cell of a list of boxes.
constexpr int MAXFIX = 16;
struct FixBoxes {
	Box b[MAXFIX];
	long off[MAXFIX + 1];       // first thread of each box; off[n] = all
	int n;
};

template<class T>
__global__ __launch_bounds__(256) void k_fix(CurlArgs<T> a, FixBoxes f)
{
	const long t0 = (long)blockIdx.x * 256 + threadIdx.x;
	if (t0 >= f.off[f.n]) return;
	int q = 0;
	while (q + 1 < f.n && f.off[q + 1] <= t0) q++;
	const Box &B = f.b[q];
	int di, dj, dk;
	boxCell(t0 - f.off[q], B.i1 - B.i0 + 1, B.j1 - B.j0 + 1, B.k1 - B.k0 + 1, di, dj, dk);
	const int i = B.i0 + di, j = B.j0 + dj, k = B.k0 + dk;
	constexpr int CM = 0;
	constexpr bool NT = false, NTG = false, SHARE = true, NTP = false, MERGEZ = false, UNI = false, SEP = true, wr = true, NTC = false;
	constexpr int KSH = 0;
#include "curl_cell.inc"
}

// The end of a half step in one launch: TF/SF strips and sources (additive operations,
// src/update.c:161-179,194-272) followed by the ghost copies (src/update.c:182-191 with the
// schedules at :284-297 and :311-339).  The operations form a program in the reference's
// call order.  The host compiles each launch's program into pieces, boxes of cells whose
// sequential result is computed the same way (buildPieces), and the launch runs one thread
// per piece cell:
//  * an additive piece holds cells whose first additive operation on the array is the
//    piece's; its list is every additive operation covering them, in program order, and
//    the thread applies those switched on this half step;
//  * a copy piece holds cells whose last ghost copy is the piece's; its list is the copy
//    chain from that copy back to a cell that no earlier copy writes (the base cell, at a
//    fixed offset), replayed forward from the base cell's value.
// Cells that a later copy overwrites belong to no earlier piece.  The host splits a
// program into several launches where a chain's base cell is written by the same launch
// (see buildPost), so every launch only reads cells it does not write.
enum { OP_COPY = 0, OP_STRIP = 1, OP_SOURCE = 2 };

template<class T> struct PostOp {
	int kind, slot;             // slot: TF/SF box (STRIP) or source index (SOURCE)
	T *F, *G;                   // target; COPY: imaginary partner or null
	const T *A;                 // STRIP: auxiliary field
	const T *C;                 // STRIP: 1/eps or null; SOURCE: amplitude array or null
	const T *W;                 // SOURCE: per-cell waves or null
	Lay LF, LA, LC;
	Box b;
	int di, dj, dk;
	T c, s;                     // COPY: cos, sin; STRIP: dt/ds; SOURCE: scalar amplitude
	int exX, exY, auxIMAX, auxJMAX, altX, altY;
};

constexpr int MAXSRC = 128;     // sources per chain (E or H) of one world
constexpr int MAXCHAIN = 24;    // copies per launch
constexpr int MAXPOST = 48;     // operations per launch
constexpr int CURL_TPB = 2;     // tiles per workgroup of a batched curl launch (measured best of 1..16)
constexpr int CURL_TPB_PAIR = 1;  // the same for entries of both parts of a Bloch world
constexpr int CURL_TPB_TILE = TPB_TILE;  // the same for entries of 64 x 4 tiles (rows of 64 cells or more)

constexpr int POST_WG = 64;     // threads (cells) per workgroup of k_post: one wavefront (64 k-point
                                // worlds: H program 84 -> 79 us, E 35 -> 31 us against 256)

struct PostPiece {
	int kind;                   // OP_COPY, or OP_STRIP for an additive piece
	int n;                      // operations in list
	int si, sj, sk;             // copy piece: base cell - cell
	Box b;                      // the cells
	long off;                   // first thread in the launch (a multiple of POST_WG)
	unsigned char list[MAXPOST];// operation indices (copy piece: its chain, last copy first)
};

template<class T> struct PostArgs {
	const PostOp<T> *ops;
	const PostPiece *pieces;    // per workgroup of the launch: the piece it runs
	long total;
	unsigned strips;            // TF/SF boxes whose strips apply this half step (bit s)
	unsigned src[MAXSRC / 32];  // sources active this half step
	T wave[MAXSRC];             // their scalar wave values
};

template<class T> __device__ inline bool opOn(const PostArgs<T> &a, const PostOp<T> &o)
{
	if (o.kind == OP_STRIP) return (a.strips >> o.slot) & 1u;
	if (o.kind == OP_SOURCE) return (a.src[o.slot >> 5] >> (o.slot & 31)) & 1u;
	return true;
}

// F(i,j,k) + dt/ds G_aux (1/eps), or F + amplitude wave, as the reference evaluates them.
template<class T> __device__ inline T addOne(const PostArgs<T> &a, const PostOp<T> &o, T f, int i, int j, int k)
{
	if (o.kind == OP_STRIP) {
		int ia = i + o.di, ja = j + o.dj;
		if (o.exX && ia > o.auxIMAX) ia = o.auxIMAX + o.altX;
		if (o.exY && ja > o.auxJMAX) ja = o.auxJMAX + o.altY;
		const T g = o.A[o.LA.at(ia, ja, k + o.dk)];
		return o.C ? f + o.c * g * o.C[o.LC.at(i, j, k)] : f + o.c * g;
	}
	const long c = o.LC.at(i, j, k);
	const T amp = o.C ? o.C[c] : o.c;
	const T w = o.W ? o.W[c] : a.wave[o.slot];
	return f + amp * w;
}

// One workgroup (number blk within its launch) of k_post.  Pieces start on workgroup
// boundaries, so the piece and its operations are the same for the whole workgroup; the
// launch's table holds each workgroup's piece, so a workgroup reaches its cell through three
// dependent loads (piece, operation, cell) where a search of the pieces added about seven
// (measured with 256-thread workgroups: 64 k-point worlds' H program 95 -> 84 us; 256^3
// nanoblock H program 17.5 -> 16.1 us).
template<class T>
__device__ __forceinline__ void postBlock(const PostArgs<T> &a, int blk)
{
	const PostPiece &P = a.pieces[blk];
	const int ni = P.b.i1 - P.b.i0 + 1, nj = P.b.j1 - P.b.j0 + 1, nk = P.b.k1 - P.b.k0 + 1;
	const int t = blk * POST_WG + (int)threadIdx.x - (int)P.off;
	if (t >= ni * nj * nk) return;
	const int dk = t % nk, r = t / nk, dj = r % nj, di = r / nj;
	const int i = P.b.i0 + di, j = P.b.j0 + dj, k = P.b.k0 + dk;
	const PostOp<T> &o = a.ops[P.list[0]];
	const long d = o.LF.at(i, j, k);
	if (P.kind == OP_COPY) {
		const long s = o.LF.at(i + P.si, j + P.sj, k + P.sk);
		T f = o.F[s], g = o.G ? o.G[s] : T(0);
		for (int u = P.n - 1; u >= 0; u--) {
			const PostOp<T> &r = a.ops[P.list[u]];
			if (o.G) {
				const T fs = f, gs = g;
				f = r.c * fs + r.s * gs;
				g = r.c * gs - r.s * fs;
			}
			else f = r.c * f;
		}
		o.F[d] = f;
		if (o.G) o.G[d] = g;
		return;
	}
	bool on = false;
	T f = T(0);
	for (int u = 0; u < P.n; u++) {
		const PostOp<T> &r = a.ops[P.list[u]];
		if (!opOn(a, r)) continue;
		if (!on) f = o.F[d], on = true;
		f = addOne(a, r, f, i, j, k);
	}
	if (on) o.F[d] = f;
}

template<class T>
__global__ __launch_bounds__(POST_WG) void k_post(PostArgs<T> a)
{
	postBlock(a, (int)blockIdx.x);
}

// Several worlds' programs in one launch: entry e owns workgroups [start[e], start[e+1]).
template<class T>
__global__ __launch_bounds__(POST_WG) void k_post_batch(const PostArgs<T> *__restrict__ args, const int *__restrict__ start, int n)
{
	const int b = (int)blockIdx.x;
	const int e = batchEntry(start, n, b);
	postBlock(args[e], b - start[e]);
}

// Phaser running DFT (src/phaser.c:83-110): one thread per face cell and block of PHS_QPT
// frequencies (the frequency blocks multiply the threads, so even a small phaser box fills
// the chip).  The face components are read once per thread; accumulators stream coalesced
// per [frequency][component] row.  A single world's updates are deferred (k_phs_sample,
// k_phs_flush, and k_phs_sample_batch / k_phs_flush_batch for several phasers at once).
struct PhsFaceGeom {
	int kind;                   // 0 xMin 1 xMax 2 yMin 3 yMax 4 zMin 5 zMax
	int fixed;                  // the face's fixed index
	int n2;                     // cells along the fast face axis
	long off;                   // first cell of this face in the concatenated cell list
	int si0, si1, sj0, sj1;     // z faces: cells (i, j) the H curl sampled (k_phs_sample skips
	                            // them), none when si0 > si1
};

constexpr int PHS_QPT = 4;      // frequencies per thread (measured: 1 < 2 < 4)

// A complex accumulator (re, im) as one vector for 8- or 16-byte accesses.
template<class TA> struct Pair;
template<> struct Pair<float> { typedef float V __attribute__((vector_size(8))); };
template<> struct Pair<double> { typedef double V __attribute__((vector_size(16))); };

template<class T> struct PhsArgs {
	const T *E[3], *H[3];
	Lay L;
	int iMin, jMin, kMin, nface;
	PhsFaceGeom face[6];
	long ncell;
	void *acc;                  // [freq][comp 4][cell] complex (TA pairs)
	void *cmp;                  // FP64 accumulators: their compensations, laid out as acc
	int q0, nq;                 // the launch's frequencies (of the accumulators)
};

// The sampled fields of face cell t of a phaser: the two tangential E components and the two
// sums of the tangential H components across the face (src/phaser.c:83-110).
template<class T>
__device__ __forceinline__ void phaserFields(const PhsArgs<T> &a, long t, T &e1, T &e2, T &h1, T &h2)
{
	int f = 0;
	while (f + 1 < a.nface && t >= a.face[f + 1].off) f++;
	const PhsFaceGeom &g = a.face[f];
	long local = t - g.off;
	int u = (int)(local / g.n2), v = (int)(local % g.n2);
	int i, j, k;
	if (g.kind < 2) {
		i = g.fixed, j = a.jMin + u, k = a.kMin + v;
		long id = a.L.at(i, j, k), s = a.L.si;
		e1 = a.E[1][id], e2 = a.E[2][id];
		h1 = a.H[1][id] + a.H[1][id + s], h2 = a.H[2][id] + a.H[2][id + s];
	}
	else if (g.kind < 4) {
		i = a.iMin + u, j = g.fixed, k = a.kMin + v;
		long id = a.L.at(i, j, k), s = a.L.sj;
		e1 = a.E[2][id], e2 = a.E[0][id];
		h1 = a.H[2][id] + a.H[2][id + s], h2 = a.H[0][id] + a.H[0][id + s];
	}
	else {
		i = a.iMin + u, j = a.jMin + v, k = g.fixed;
		long id = a.L.at(i, j, k);
		e1 = a.E[0][id], e2 = a.E[1][id];
		h1 = a.H[0][id] + a.H[0][id + 1], h2 = a.H[1][id] + a.H[1][id + 1];
	}
}

// Workgroup w of a one-dimensional phaser grid of ny frequency blocks per cell block: the
// frequency blocks of one cell block run on one XCD (the dispatcher deals workgroup ids to
// the 8 XCDs in turn), so that XCD's L2 fetches the cells' field lines once instead of every
// XCD fetching them (the cell blocks are padded to a multiple of 8).
__device__ inline void phaserBlock(int w, int ny, int &xb, int &yb)
{
	xb = w % 8 + 8 * (w / (8 * ny));
	yb = (w / 8) % ny;
}

// Deferred phaser updates: a sample stores the face fields of one step (ring slot, [4][cell]);
// a flush adds the pending samples to the accumulators in sample order, the same additions as
// one immediate update per sample, reading and writing the accumulators once for all of them.
template<class T>
__global__ __launch_bounds__(256) void k_phs_sample(PhsArgs<T> a, T *__restrict__ slot, long n)
{
	const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= n) return;
	{
		int f = 0;
		while (f + 1 < a.nface && t >= a.face[f + 1].off) f++;
		const PhsFaceGeom &g = a.face[f];
		if (g.kind >= 4 && g.si0 <= g.si1) {        // cells the H curl sampled
			const long local = t - g.off;
			const int i = a.iMin + (int)(local / g.n2), j = a.jMin + (int)(local % g.n2);
			if (i >= g.si0 && i <= g.si1 && j >= g.sj0 && j <= g.sj1) return;
		}
	}
	T e1, e2, h1, h2;
	phaserFields(a, t, e1, e2, h1, h2);
	slot[t] = e1, slot[a.ncell + t] = e2, slot[2 * a.ncell + t] = h1, slot[3 * a.ncell + t] = h2;
}

// s += y with Neumaier's compensation (Neumaier 1974, the improved Kahan-Babuska sum): c
// collects the low-order part each addition drops; the sum's value is s + c.
__device__ __forceinline__ void neumaierAdd(double &s, double &c, double y)
{
	const double t = s + y;
	c += fabs(s) >= fabs(y) ? (s - t) + y : (y - t) + s;
	s = t;
}

// ring: ns samples [s][4][cell]; ex: per sample `stride` doubles, frequency q of the launch at
// ex[s * stride + 4 q] = (eE re, eE im, eH re, eH im).  Face cell t, frequencies qa ...
// Float accumulators add plainly (the reference's sums); FP64 ones with compensation.
template<class T, class TA>
__device__ __forceinline__ void phaserFlushCell(const PhsArgs<T> &a, long t, int qa, const T *__restrict__ ring, int ns,
                                                const double *__restrict__ ex, int stride)
{
	typedef typename Pair<TA>::V V;
	constexpr bool compensated = sizeof(TA) == 8;
	V *acc = (V *)a.acc, *cmp = (V *)a.cmp;
	const TA half = TA(0.5);
	const int nq = min(PHS_QPT, a.nq - qa);
	V x[PHS_QPT][4], c[PHS_QPT][4];
#pragma unroll
	for (int u = 0; u < PHS_QPT; u++)
		if (u < nq) {
			const long o = (long)(a.q0 + qa + u) * 4 * a.ncell + t;
			for (int m = 0; m < 4; m++) {
				x[u][m] = __builtin_nontemporal_load(acc + o + m * a.ncell);
				if constexpr (compensated) c[u][m] = __builtin_nontemporal_load(cmp + o + m * a.ncell);
			}
		}
	for (int s = 0; s < ns; s++) {
		const T *r = ring + (long)s * 4 * a.ncell + t;
		const T e1 = r[0], e2 = r[a.ncell], h1 = r[2 * a.ncell], h2 = r[3 * a.ncell];
#pragma unroll
		for (int u = 0; u < PHS_QPT; u++)
			if (u < nq) {
				const double *f = ex + (long)s * stride + 4 * (qa + u);
				const TA er = TA(f[0]), ei = TA(f[1]), hr = TA(f[2]), hi = TA(f[3]);
				const TA y[4][2] = { { er * TA(e1), ei * TA(e1) }, { er * TA(e2), ei * TA(e2) },
				                     { hr * TA(h1) * half, hi * TA(h1) * half }, { hr * TA(h2) * half, hi * TA(h2) * half } };
				for (int m = 0; m < 4; m++)
					for (int p = 0; p < 2; p++) {
						if constexpr (compensated) {
							double sm = x[u][m][p], cm = c[u][m][p];
							neumaierAdd(sm, cm, y[m][p]);
							x[u][m][p] = sm, c[u][m][p] = cm;
						}
						else x[u][m][p] = x[u][m][p] + y[m][p];
					}
			}
	}
#pragma unroll
	for (int u = 0; u < PHS_QPT; u++)
		if (u < nq) {
			const long o = (long)(a.q0 + qa + u) * 4 * a.ncell + t;
			for (int m = 0; m < 4; m++) {
				__builtin_nontemporal_store(x[u][m], acc + o + m * a.ncell);
				if constexpr (compensated) __builtin_nontemporal_store(c[u][m], cmp + o + m * a.ncell);
			}
		}
}

// Fold the compensations of FP64 accumulators into their sums before anything reads them:
// acc += cmp and cmp = 0 where cmp is not zero (a zero leaves the sum's bits alone).
__global__ __launch_bounds__(256) void k_phs_fold(double *__restrict__ acc, double *__restrict__ cmp, long n)
{
	const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= n) return;
	const double c = cmp[t];
	if (c != 0) acc[t] += c, cmp[t] = 0;
}

template<class T, class TA>
__global__ __launch_bounds__(256) void k_phs_flush(PhsArgs<T> a, const T *__restrict__ ring, int ns,
                                                   const double *__restrict__ ex, int stride, int ny)
{
	int xb, yb;
	phaserBlock((int)blockIdx.x, ny, xb, yb);
	const long t = (long)xb * 256 + threadIdx.x;
	if (t >= a.ncell) return;
	phaserFlushCell<T, TA>(a, t, yb * PHS_QPT, ring, ns, ex, stride);
}

// The same for several phasers (of several worlds) in one launch each: the samples of every
// entry (acc: its ring slot), and the flushes of the entries whose ring is full.
template<class T>
__global__ __launch_bounds__(256) void k_phs_sample_batch(const PhsArgs<T> *__restrict__ args, const int *__restrict__ start, int n)
{
	const int b = (int)blockIdx.x;
	const int e = batchEntry(start, n, b);
	const PhsArgs<T> &a = args[e];
	const long t = (long)(b - start[e]) * 256 + threadIdx.x;
	if (t >= a.ncell) return;
	T e1, e2, h1, h2;
	phaserFields(a, t, e1, e2, h1, h2);
	T *slot = (T *)a.acc;
	slot[t] = e1, slot[a.ncell + t] = e2, slot[2 * a.ncell + t] = h1, slot[3 * a.ncell + t] = h2;
}

template<class T> struct PhsFlush {
	PhsArgs<T> a;
	const T *ring;
	const double *ex;
	int ns, stride;
};

template<class T, class TA>
__global__ __launch_bounds__(256) void k_phs_flush_batch(const PhsFlush<T> *__restrict__ args, const int *__restrict__ start, int n, int nb, int ny)
{
	int b, yb;
	phaserBlock((int)blockIdx.x, ny, b, yb);
	if (b >= nb) return;
	const int e = batchEntry(start, n, b);
	const PhsFlush<T> &F = args[e];
	const long t = (long)(b - start[e]) * 256 + threadIdx.x;
	if (t >= F.a.ncell || yb * PHS_QPT >= F.a.nq) return;
	phaserFlushCell<T, TA>(F.a, t, yb * PHS_QPT, F.ring, F.ns, F.ex, F.stride);
}

// Second-order (three-wave) coupling (src/update.c:344-411): one thread per cell of the box.
// A component of a pump field at another component's position is the mean of four values
// (their sum in the field precision, then 0.25 times it in double); each term adds
// 2 alt d (1/eps) fill (product) to the signal field in the order of the reference, where a
// product of two values at the component's own position stays in the field precision and
// any other product is double, as the reference's C expressions promote them.
template<class T> struct ThreeWaveArgs {
	T *F[3];                    // signal E (one part)
	const T *A[3], *B[3];       // pump E of the two pump worlds (the same part)
	const T *C[3];              // the signal world's 1/eps
	const T *fill[3];           // the object's fill factors, layout LO
	Lay L, LO;
	Box b;                      // the cells
	float d[3][6];
	float alt;
};

template<class T> __device__ __forceinline__ void twAdd(T &f, T c, T g) { f = f + c * g; }
template<class T> __device__ __forceinline__ void twAdd(T &f, T c, double g) { f = (T)((double)f + (double)c * g); }
__device__ __forceinline__ void twAdd(double &f, double c, double g) { f = f + c * g; }

template<class T>
__global__ __launch_bounds__(256) void k_three_wave(ThreeWaveArgs<T> a)
{
	const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	const int nk = a.b.k1 - a.b.k0 + 1, nj = a.b.j1 - a.b.j0 + 1, ni = a.b.i1 - a.b.i0 + 1;
	if (t >= (long)ni * nj * nk) return;
	const int k = a.b.k0 + (int)(t % nk), j = a.b.j0 + (int)((t / nk) % nj), i = a.b.i0 + (int)(t / ((long)nk * nj));
	const Lay &L = a.L;
	const long id = L.at(i, j, k), o = a.LO.at(i, j, k), si = L.si, sj = L.sj;
	auto at = [&](const T *const *G, int c, long off) { return *glb(G[c] + id + off); };
	// a pump component at the x, y or z component's position (src/update.c:344-352)
	auto yx = [&](const T *const *G) { return 0.25 * (double)(at(G, 1, 0) + at(G, 1, -si) + at(G, 1, sj) + at(G, 1, sj - si)); };
	auto zx = [&](const T *const *G) { return 0.25 * (double)(at(G, 2, 0) + at(G, 2, -si) + at(G, 2, 1) + at(G, 2, 1 - si)); };
	auto zy = [&](const T *const *G) { return 0.25 * (double)(at(G, 2, 0) + at(G, 2, -sj) + at(G, 2, 1) + at(G, 2, 1 - sj)); };
	auto xy = [&](const T *const *G) { return 0.25 * (double)(at(G, 0, 0) + at(G, 0, -sj) + at(G, 0, si) + at(G, 0, si - sj)); };
	auto xz = [&](const T *const *G) { return 0.25 * (double)(at(G, 0, 0) + at(G, 0, -1) + at(G, 0, si) + at(G, 0, si - 1)); };
	auto yz = [&](const T *const *G) { return 0.25 * (double)(at(G, 1, 0) + at(G, 1, -1) + at(G, 1, sj) + at(G, 1, sj - 1)); };
	const T ax = at(a.A, 0, 0), ay = at(a.A, 1, 0), az = at(a.A, 2, 0), bx = at(a.B, 0, 0), by = at(a.B, 1, 0), bz = at(a.B, 2, 0);
	for (int c = 0; c < 3; c++) {
		T f = *glb(a.F[c] + id);
		const T w = *glb(a.C[c] + id), fl = *glb(a.fill[c] + o);
		auto coef = [&](int q) { return (T)((T)((T)((T)(2 * a.alt) * (T)a.d[c][q]) * w) * fl); };
		const float *d = a.d[c];
		if (c == 0) {
			if (d[0]) twAdd(f, coef(0), ax * bx);
			if (d[1]) twAdd(f, coef(1), yx(a.A) * yx(a.B));
			if (d[2]) twAdd(f, coef(2), zx(a.A) * zx(a.B));
			if (d[3]) twAdd(f, coef(3), yx(a.A) * zx(a.B)), twAdd(f, coef(3), zx(a.A) * yx(a.B));
			if (d[4]) twAdd(f, coef(4), zx(a.A) * (double)bx), twAdd(f, coef(4), (double)ax * zx(a.B));
			if (d[5]) twAdd(f, coef(5), (double)ax * yx(a.B)), twAdd(f, coef(5), yx(a.A) * (double)bx);
		}
		else if (c == 1) {
			if (d[0]) twAdd(f, coef(0), xy(a.A) * xy(a.B));
			if (d[1]) twAdd(f, coef(1), ay * by);
			if (d[2]) twAdd(f, coef(2), zy(a.A) * zy(a.B));
			if (d[3]) twAdd(f, coef(3), (double)ay * zy(a.B)), twAdd(f, coef(3), zy(a.A) * (double)by);
			if (d[4]) twAdd(f, coef(4), zy(a.A) * xy(a.B)), twAdd(f, coef(4), xy(a.A) * zy(a.B));
			if (d[5]) twAdd(f, coef(5), xy(a.A) * (double)by), twAdd(f, coef(5), (double)ay * xy(a.B));
		}
		else {
			if (d[0]) twAdd(f, coef(0), xz(a.A) * xz(a.B));
			if (d[1]) twAdd(f, coef(1), yz(a.A) * yz(a.B));
			if (d[2]) twAdd(f, coef(2), az * bz);
			if (d[3]) twAdd(f, coef(3), yz(a.A) * (double)bz), twAdd(f, coef(3), (double)az * yz(a.B));
			if (d[4]) twAdd(f, coef(4), (double)az * xz(a.B)), twAdd(f, coef(4), xz(a.A) * (double)bz);
			if (d[5]) twAdd(f, coef(5), xz(a.A) * yz(a.B)), twAdd(f, coef(5), yz(a.A) * xz(a.B));
		}
		*glb(a.F[c] + id) = f;
	}
}

// Per-call launch arguments of batched launches from pinned host memory into device memory,
// in stream order: a copy-engine transfer there would add a round trip between the
// launches around it (about 25 us measured between the batched curl and post launches).
__global__ __launch_bounds__(256) void k_fetch(unsigned *__restrict__ dst, const unsigned *__restrict__ src, int words)
{
	for (int w = (int)(blockIdx.x * blockDim.x + threadIdx.x); w < words; w += (int)(gridDim.x * blockDim.x)) dst[w] = src[w];
}

// Several boxes of several arrays (of one or several worlds, float or double) into one float
// buffer: segment s covers dst[off[s] .. off[s+1]) with the cells of its box, k fastest.
struct GatherSeg {
	const void *src;
	Lay L;
	Box b;
	long off;
	int dbl;                    // the source array holds doubles
	int pad;                    // zero: tables are compared byte for byte
};

__global__ __launch_bounds__(256) void k_gather(const GatherSeg *__restrict__ seg, int nseg, long total, float *dst)
{
	const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= total) return;
	int lo = 0, hi = nseg - 1;
	while (lo < hi) {
		const int mid = (lo + hi + 1) / 2;
		if (seg[mid].off <= t) lo = mid;
		else hi = mid - 1;
	}
	const GatherSeg &g = seg[lo];
	const long l = t - g.off;
	const int nk = g.b.k1 - g.b.k0 + 1, nj = g.b.j1 - g.b.j0 + 1;
	const int dk = (int)(l % nk), dj = (int)(l / nk % nj), di = (int)(l / nk / nj);
	const long at = g.L.at(g.b.i0 + di, g.b.j0 + dj, g.b.k0 + dk);
	dst[t] = g.dbl ? (float)((const double *)g.src)[at] : ((const float *)g.src)[at];
}

// Pack a box of a device array into a compact buffer (float or double), and back.
template<class T, class U>
__global__ __launch_bounds__(256) void k_pack(const T *src, Lay L, Box b, U *dst)
{
	const int ni = b.i1 - b.i0 + 1, nj = b.j1 - b.j0 + 1, nk = b.k1 - b.k0 + 1;
	long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= (long)ni * nj * nk) return;
	int di, dj, dk;
	boxCell(t, ni, nj, nk, di, dj, dk);
	dst[t] = (U)src[L.at(b.i0 + di, b.j0 + dj, b.k0 + dk)];
}

template<class T, class U>
__global__ __launch_bounds__(256) void k_unpack(const U *src, Lay L, Box b, T *dst)
{
	const int ni = b.i1 - b.i0 + 1, nj = b.j1 - b.j0 + 1, nk = b.k1 - b.k0 + 1;
	long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= (long)ni * nj * nk) return;
	int di, dj, dk;
	boxCell(t, ni, nj, nk, di, dj, dk);
	dst[L.at(b.i0 + di, b.j0 + dj, b.k0 + dk)] = (T)src[t];
}

template<class TA>
__global__ __launch_bounds__(256) void k_acc_to_float(const TA *src, float *dst, long n)
{
	long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t < n) dst[t] = (float)src[t];
}

template<class TA>
__global__ __launch_bounds__(256) void k_acc_to_double(const TA *src, double *dst, long n)
{
	long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t < n) dst[t] = (double)src[t];
}

inline unsigned blocksFor(long n, int per = 256) { return (unsigned)((n + per - 1) / per); }

// A fill kernel for buffers that must start from a value (used instead of a memset node,
// which callers capturing the work into a graph should not meet).
template<class T>
__global__ __launch_bounds__(256) void k_fill(T *p, long n, T v)
{
	const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t < n) p[t] = v;
}

// Probe spectra (the running DFT of writeSpectrum, src/output.c:177-230) in FP64 with
// compensated (Neumaier) summation.  acc holds, per probe p and frequency bin b, the running
// sum (re, im) at acc[2 (p nb + b)] and its compensation at acc[2 (p nb + b) + 1]; a thread
// takes one bin and up to SPEC_PPT probes and adds ns staged samples (tau, v[0..np-1]) in
// sample order, each with the factor e^{i 2 pi tau (nTop - b)} (one sincos per sample and
// bin, shared by the thread's probes).
constexpr int SPEC_PPT = 8;

__device__ inline void neumaier(double &s, double &c, double x)
{
	const double t = s + x;
	c += fabs(s) >= fabs(x) ? (s - t) + x : (x - t) + s;
	s = t;
}

__global__ __launch_bounds__(256) void k_spectrum_add(double2 *__restrict__ acc, int np, int nb, int nTop,
                                                      const double *__restrict__ samp, int ns)
{
	const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
	const int b = t % nb, p0 = t / nb * SPEC_PPT;
	if (p0 >= np) return;
	const int pn = min(SPEC_PPT, np - p0);
	double2 sum[SPEC_PPT], cmp[SPEC_PPT];
#pragma unroll
	for (int u = 0; u < SPEC_PPT; u++)
		if (u < pn) sum[u] = acc[2 * ((long)(p0 + u) * nb + b)], cmp[u] = acc[2 * ((long)(p0 + u) * nb + b) + 1];
	const double w = 2 * M_PI;
	const int harmonic = nTop - b;
	for (int q = 0; q < ns; q++) {
		const double *r = samp + (long)q * (1 + np);
		double sn, cs;
		sincos(w * r[0] * harmonic, &sn, &cs);
#pragma unroll
		for (int u = 0; u < SPEC_PPT; u++)
			if (u < pn) {
				const double v = r[1 + p0 + u];
				neumaier(sum[u].x, cmp[u].x, v * cs);
				neumaier(sum[u].y, cmp[u].y, v * sn);
			}
	}
#pragma unroll
	for (int u = 0; u < SPEC_PPT; u++)
		if (u < pn) acc[2 * ((long)(p0 + u) * nb + b)] = sum[u], acc[2 * ((long)(p0 + u) * nb + b) + 1] = cmp[u];
}

// Per-row sums over a plane of a world's fields for the instantaneous Poynting flux
// (src/output.c:233-332).  Term t (PLANE_TERM ints: normal axis, plane index, E component,
// H component, row axis, first row, last row, first and last inner index, weight offset,
// output offset, first row of the launch) gives one sum per row r:
//     sum over inner s, in order, of w[s] * E_e(cell) * (H_h(cell) + H_h(cell + 1 along the normal))
// with the cell at (normal = plane index, row axis = r, inner axis = s), the real parts, the
// H pair added in the field precision and the rest in double.  One thread per row.
constexpr int PLANE_TERM = 12;

template<class T> struct Arrays6 { const T *p[6]; };   // E x y z, H x y z (or E x y z, 1/eps x y z)

template<class T>
__global__ __launch_bounds__(256) void k_plane_sums(Arrays6<T> F, Lay L, const int *__restrict__ terms, int nterm,
                                                    long rows, const double *__restrict__ w, double *__restrict__ out)
{
	const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (g >= rows) return;
	int t = 0;
	while (t + 1 < nterm && g >= terms[PLANE_TERM * (t + 1) + 11]) t++;
	const int *u = terms + PLANE_TERM * t;
	const int nrm = u[0], row = u[4], inner = 3 - u[0] - u[4];
	const int r = u[5] + (int)(g - u[11]);
	const T *E = F.p[u[2]], *H = F.p[3 + u[3]];
	int c[3];
	c[nrm] = u[1], c[row] = r;
	const long step = inner == 0 ? L.si : inner == 1 ? L.sj : 1, shift = nrm == 0 ? L.si : nrm == 1 ? L.sj : 1;
	c[inner] = u[7];
	long id = L.at(c[0], c[1], c[2]);
	double sum = 0;
	for (int q = 0; q <= u[8] - u[7]; q++, id += step)
		sum += w[u[9] + q] * (double)E[id] * (double)(H[id] + H[id + shift]);
	out[u[10] + (r - u[5])] = sum;
}

// Column sums of the power absorbed by a dispersive object (src/output.c:375-398): for each
// column (i, j) of the box, over k in order, the three terms dz E_q current_q (q = x, y, z;
// dz = E widths for x and y, H widths for z, passed as weights), skipped where 1/eps is 0;
// current_q = (a_q ? 2 fill_q E_q / (1 + 1/a_q) : 0) + sum_p J_p,q, over 1/eps_q, in the field
// precision (src/output.c:25-47).  out per column: the three sums, then the three terms of the
// first cell and of the last cell (the host's boundary corrections).
template<class T>
__global__ __launch_bounds__(256) void k_absorption(Arrays6<T> EC, Lay L, const ObjDev<T> *__restrict__ obj, int n, Box b,
                                                    const double *__restrict__ dzE, const double *__restrict__ dzH, double *__restrict__ out)
{
	const int nj = b.j1 - b.j0 + 1;
	const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (col >= (long)(b.i1 - b.i0 + 1) * nj) return;
	const int i = b.i0 + (int)(col / nj), j = b.j0 + (int)(col % nj);
	const ObjDev<T> &O = obj[n];
	double sum[3] = {0, 0, 0}, first[3] = {0, 0, 0}, last[3] = {0, 0, 0};
	for (int k = b.k0; k <= b.k1; k++) {
		const long id = L.at(i, j, k), o = O.LO.at(i, j, k);
#pragma unroll
		for (int q = 0; q < 3; q++) {
			const T c = EC.p[3 + q][id];
			if (c == T(0)) continue;
			const T e = EC.p[q][id];
			T f = 0;
			if (O.a[q] != T(0)) f += T(2) * O.fill[q][o] * e / (T(1) + T(1) / O.a[q]);
			for (int p = 0; p < O.NK; p++) f += O.J[0][q][p][o];
			const T d = (T)(q < 2 ? dzE[k - b.k0] : dzH[k - b.k0]);
			const double term = (double)(d * e * (f / c));
			sum[q] += term;
			if (k == b.k0) first[q] = term;
			if (k == b.k1) last[q] = term;
		}
	}
	double *r = out + col * 9;
	for (int q = 0; q < 3; q++) r[q] = sum[q], r[3 + q] = first[q], r[6 + q] = last[q];
}

// Inner sums of the net flux Re(E x H*) of one frequency over the phaser faces
// (src/phaser.c:142-223).  Term t of the table (FLUX_TERM ints: kind, first cell, rows na, row
// length nb, components x and y, weight offset, output offset, first workgroup) sums
// w * Re(x conj(y)) over a face stored row by row (cell a * nb + b):
//  * kind 0 along each row: a wavefront per row, lane l adds b = l, l + 64, ... in order and
//    writes its partial sum to out[a * 64 + l] (the host adds the 64 in order);
//  * kind 1 down each column: a thread per column b adds a = 0 .. na - 1 in order.
// Both read every row with consecutive lanes on consecutive cells and use no cross-lane
// operation.  The accumulators are read once: about 2 x 16 bytes per face cell and term.
constexpr int FLUX_TERM = 9;

// Doubles of the file-scope `scratch` array that alis_dev_phaser_flux_sums_noalloc uses for
// its outputs, weights and term table (4 MiB per GPU, held until the process ends).
constexpr long FLUX_SCRATCH_DOUBLES = 1L << 19;

template<class TA>
__global__ __launch_bounds__(256) void k_flux_sums(const typename Pair<TA>::V *__restrict__ acc, long ncell,
                                                   const int *__restrict__ terms, int nterm,
                                                   const double *__restrict__ w, double *__restrict__ out, long nout)
{
	typedef typename Pair<TA>::V V;
	const int b = (int)blockIdx.x;
	// blockIdx.y: the frequency of a range launch (its accumulators and its outputs)
	acc += (long)blockIdx.y * 4 * ncell, out += (long)blockIdx.y * nout;
	int t = 0;
	while (t + 1 < nterm && b >= terms[FLUX_TERM * (t + 1) + 8]) t++;
	const int *T = terms + FLUX_TERM * t;
	const int na = T[2], nb = T[3], lb = b - T[8];
	const V *X = acc + (long)T[4] * ncell + T[1], *Y = acc + (long)T[5] * ncell + T[1];
	const double *wt = w + T[6];
	double *o = out + T[7];
	if (T[0] == 0) {
		const int a = lb * 4 + (int)threadIdx.x / 64, l = (int)threadIdx.x % 64;
		if (a >= na) return;
		double r = 0;
		for (int c = l; c < nb; c += 64) {
			const V x = X[(long)a * nb + c], y = Y[(long)a * nb + c];
			r += wt[c] * ((double)x[0] * (double)y[0] + (double)x[1] * (double)y[1]);
		}
		o[(long)a * 64 + l] = r;
	}
	else {
		const int c = lb * 256 + (int)threadIdx.x;
		if (c >= nb) return;
		double r = 0;
		for (int a = 0; a < na; a++) {
			const V x = X[(long)a * nb + c], y = Y[(long)a * nb + c];
			r += wt[a] * ((double)x[0] * (double)y[0] + (double)x[1] * (double)y[1]);
		}
		o[c] = r;
	}
}


}  // namespace

// Scratch memory of alis_dev_phaser_flux_sums_noalloc, for callers that cannot allocate:
// one copy per GPU, shared by every such call on that GPU (calls that overlap on one GPU
// race).  Host code reaches it only through the address that hipGetSymbolAddress returns.
static __device__ __align__(256) double scratch[FLUX_SCRATCH_DOUBLES];

// ---------------------------------------------------------------------------------------
// Device world
// ---------------------------------------------------------------------------------------

// Launch kernel k with its start and end bound to events e0 and e1 (timing, or an event
// another stream waits for); a plain launch, which costs the host less, when there are none
// (64^3 FP64 dipole, five launches a step: 0.080 -> 0.073-0.079 ms a step).
template<class... P, class... A>
inline void launchEv(void (*k)(P...), dim3 grid, dim3 block, size_t lds, hipStream_t st, hipEvent_t e0, hipEvent_t e1, A... args)
{
	if (e0 || e1) hipExtLaunchKernelGGL(k, grid, block, lds, st, e0, e1, 0, args...);
	else hipLaunchKernelGGL(k, grid, block, lds, st, args...);
}

struct DevSlab {
	bool on = false;
	int lo = 0, hi = -1;
	Lay L{};
	Box b{};                    // host extents of the slab array
	long n = 0;                 // elements of each array
	void *p[2][2] = {};         // [m][a/b]
	void *alt[2][2] = {};       // the other set (k_step), allocated at the first fused step
};

struct DevObj {
	Box box{};                  // object box (C->iMin[n]..C->iMax[n] ...)
	Box clip{};                 // box clipped to the grid [MIN..MAX]
	Lay LO{};
	int NJ = 0, NK = 0;
	float a[3] = {};
	float JJ[3][MAXPOLE] = {}, JK[3][MAXPOLE] = {}, JE[3][MAXPOLE] = {}, KJ[3][MAXPOLE] = {}, KK[3][MAXPOLE] = {}, KE[3][MAXPOLE] = {};
	void *fill[3] = {};
	void *J[2][3][MAXPOLE] = {}, *K[2][3][MAXPOLE] = {};
	void *Jalt[2][3][MAXPOLE] = {}, *Kalt[2][3][MAXPOLE] = {};   // the other set (k_step)
};

struct DevSrc {
	int h = 0, m = 0, axis = 0;
	Box b{};
	Lay LB{};                   // compact box layout
	float amp = 0;
	void *ampArr = nullptr;     // per-cell amplitude (T) or null
	bool cellWave = false;
	void *waveDev = nullptr;    // per-cell wave buffer (T)
	float *wavePinned = nullptr;
	hipEvent_t waveDone = nullptr;
	long ncell = 0;
};

struct PostSeg {
	void *ops = nullptr;        // PostOp<T>[] in device memory
	void *pieces = nullptr;     // PostPiece[total / POST_WG] in device memory: each workgroup's
	                            //  piece
	long total = 0;             // threads (a multiple of POST_WG)
};

struct DevTfsf {
	alis_dev_tfsf S{};
	alis_dev_world *aux = nullptr;
};

// Launch arguments staged in pinned host memory and their device copy (see fetchArgs).
struct Staging {
	void *pinned = nullptr, *dev = nullptr;
	size_t bytes = 0;
	hipEvent_t done = nullptr;       // the last copy from `pinned` has been made
};

struct DevPhaser {
	bool live = false;
	std::vector<float *> cf;    // host cfield data starts [q][face][comp]
	int nf = 0, nface = 0;
	int iMin = 0, iMax = 0, jMin = 0, jMax = 0, kMin = 0, kMax = 0;
	PhsFaceGeom face[6]{};
	long ncell = 0;
	void *acc = nullptr;        // TA
	void *cmp = nullptr;        // FP64 accumulators: compensations (same allocation, after acc)
	bool dirty = false;         // cmp may hold nonzero terms (folded before reads)
	float *stage = nullptr;     // device float buffer for host syncs
	size_t bytes = 0;           // device bytes of acc and stage (and the ring)
	// deferred updates: samples of the face fields not yet added to acc
	void *ring = nullptr;       // [PHS_RING][4][ncell] T
	double *exPinned = nullptr; // [PHS_RING][nf][4] their DFT factors, staged
	double *exDev = nullptr;
	hipEvent_t exDone = nullptr;  // the staging has been copied
	int pend = 0, pq0 = 0, pnq = 0;  // samples pending and their frequency range
	bool pre = false;           // the H curl stored the z faces' sample of slot preSlot
	int preSlot = -1;           //  since the fields last changed
};

constexpr int PHS_RING = 32;    // samples per deferred flush

struct alis_dev_world {
	alis_dev_world_desc W{};    // copy of the descriptor (arrays below keep its tables alive)
	std::vector<alis_dev_ade> objDesc;
	std::vector<alis_dev_src> srcDesc[2];
	std::vector<alis_dev_tfsf> tfDesc;
	int fp64 = 0, dft64 = 0, cplx = 0;
	int device = 0;             // the GPU that holds the world (current at commit)
	size_t esz = 4;
	int iMIN, iMAX, jMIN, jMAX, kMIN, kMAX, iMin, iMax, jMin, jMax, kMin, kMax, iNUM, jNUM, kNUM;
	Lay L{};
	long fsize = 0;             // elements per field array
	void *F[2][2][3] = {};      // [h][m][axis]
	void *C[3] = {};            // inverse permittivity of the E grid, or null
	void *dt[2][3] = {};        // [h][axis] dt/d arrays (T), index - MIN
	void *cp[2][3][3] = {};     // [h][axis][b, ca, cb] CPML arrays or null
	void *cpk[2][3] = {};       // [h][axis] the same packed (b, ca, cb, 0) per index
	void *seg = nullptr;        // 1/eps per 64-cell row segment (see rowSegments) or null
	double epsBytes = 0;        // bytes per cell the E curl reads for 1/eps
	DevSlab slab[2][6];
	std::vector<DevObj> obj;
	void *objTab = nullptr;     // ObjDev<T>[obj.size()] in device memory
	void *objTile = nullptr;    // int[]: per-tile object lists (see objectTiles) or null
	std::vector<DevSrc> src[2];
	std::vector<DevTfsf> tf;
	// end-of-half-step programs (k_post launches), built at the first ghosts call after the
	// commit or a TF/SF link; strips and sources are recorded during the half step
	bool postBuilt = false;
	std::vector<PostSeg> post[2];
	unsigned pendStrips[2] = {0, 0};
	std::vector<unsigned> pendSrc[2];
	std::vector<double> pendWave[2];
	std::vector<DevPhaser> phs;
	hipStream_t st = nullptr;   // shared by a world and its TF/SF auxiliary worlds
	int *streamRefs = nullptr;  // worlds using st; the last one destroys it
	Staging phsStage, phfStage;      // staging of batched phaser samples and flushes led by this world
	void *ffWork = nullptr;          // far-field workspace
	size_t ffWorkBytes = 0;
	void *scratch = nullptr;    // pack/unpack buffer (double-sized)
	long scratchElems = 0;
	std::vector<void *> allocs;
	long bytes = 0;
	bool timing = false;
	hipEvent_t halfDone = nullptr;  // the end of the last postHalf (bound to its last launch)
	std::vector<hipEvent_t> evPairs[3];  // E curl, H curl, fused step (k_step)
	std::vector<hipEvent_t> evFree;  // timing events already read, for reuse
	double evMs[3] = {0, 0, 0};
	long evCount[3] = {0, 0, 0};
	// Fused steps (k_step): the world holds a second set of the arrays a step updates and
	// alternates between the sets; an E half step is deferred until its H half arrives.
	int fuseState = 0;          // 0 not decided yet, 1 fused steps allowed, -1 not
	bool isAux = false;         // the TF/SF auxiliary world of another world
	int inBatch = 0;            // live batches whose launch arguments name this world's arrays
	int pendE = 0;              // 1: E curl deferred; 2: and its end-of-half-step program
	bool altReady = false;      // the other set exists
	bool altStale = false;      // cells no step writes may differ between the sets
	int par = 0;                // which set is current (0: the one allocated at commit)
	void *Falt[2][2][3] = {};
	void *objTabAlt = nullptr;  // the object table naming the other set's currents
	std::vector<PostSeg> postAlt[2];  // end-of-half-step programs of the other set
	bool postBuiltAlt = false;
	void *stepArgs = nullptr;   // StepArgs<T>[2 sets][2 parts]: k_step's arguments per current set
	struct Carry { int axis; Box b; };
	std::vector<Carry> carry[2];  // cells a fused step's end-of-half programs take from the other set
	bool carryDue[2] = {false, false};
	std::vector<Box> hWrites;   // the boxes the end-of-H-half program writes
	std::vector<Box> fix;       // boxes k_fix recomputes after the E program
	bool fixOk = false;         // every additive operation lies in the update range and the
	                            //  boxes fit one k_fix launch
};

namespace {

int devAlloc(alis_dev_world *D, void **p, size_t bytes)
{
	if (!bytes) bytes = 16;
	HIPCHECK(hipMalloc(p, bytes));
	HIPCHECK(hipMemsetAsync(*p, 0, bytes, D->st));
	D->allocs.push_back(*p);
	D->bytes += bytes;
	return 0;
}

// Upload a 1-D float host array of n entries (widened to T).
template<class T>
int upload1(alis_dev_world *D, void **dst, const float *src, int n)
{
	if (devAlloc(D, dst, n * sizeof(T))) return -1;
	std::vector<T> tmp(n);
	for (int a = 0; a < n; a++) tmp[a] = (T)src[a];
	HIPCHECK(hipMemcpyAsync(*dst, tmp.data(), n * sizeof(T), hipMemcpyHostToDevice, D->st));
	HIPCHECK(hipStreamSynchronize(D->st));
	return 0;
}

int ensureScratch(alis_dev_world *D, long elems)
{
	if (elems <= D->scratchElems) return 0;
	if (D->scratch) {
		HIPCHECK(hipStreamSynchronize(D->st));
		HIPCHECK(hipFree(D->scratch));
		D->bytes -= D->scratchElems * sizeof(double);
	}
	HIPCHECK(hipMalloc(&D->scratch, elems * sizeof(double)));
	D->scratchElems = elems;
	D->bytes += elems * sizeof(double);
	return 0;
}

// Copy the host block of a box (float, k fastest) into a device array of layout L.
template<class T>
int uploadBox(alis_dev_world *D, void *dst, const Lay &L, const Box &b, const float *host)
{
	long n = b.count();
	if (!n) return 0;
	if (ensureScratch(D, n)) return -1;
	HIPCHECK(hipMemcpyAsync(D->scratch, host, n * sizeof(float), hipMemcpyHostToDevice, D->st));
	k_unpack<T, float><<<blocksFor(n), 256, 0, D->st>>>((const float *)D->scratch, L, b, (T *)dst);
	HIPCHECK(hipGetLastError());
	HIPCHECK(hipStreamSynchronize(D->st));
	return 0;
}

// Copy a box of a device array into a host float block (k fastest).
template<class T>
int downloadBox(alis_dev_world *D, const void *src, const Lay &L, const Box &b, float *host)
{
	long n = b.count();
	if (!n) return 0;
	if (ensureScratch(D, n)) return -1;
	k_pack<T, float><<<blocksFor(n), 256, 0, D->st>>>((const T *)src, L, b, (float *)D->scratch);
	HIPCHECK(hipGetLastError());
	HIPCHECK(hipMemcpyAsync(host, D->scratch, n * sizeof(float), hipMemcpyDeviceToHost, D->st));
	HIPCHECK(hipStreamSynchronize(D->st));
	return 0;
}

// Box sync into a host float*** array with arbitrary extents: pack, download, scatter rows.
template<class T>
int downloadSubBox(alis_dev_world *D, const void *src, const Lay &L, const Box &b, float ***host)
{
	long n = b.count();
	if (!n) return 0;
	std::vector<float> buf(n);
	if (downloadBox<T>(D, src, L, b, buf.data())) return -1;
	long c = 0;
	int nk = b.k1 - b.k0 + 1;
	for (int i = b.i0; i <= b.i1; i++)
		for (int j = b.j0; j <= b.j1; j++, c += nk)
			memcpy(&host[i][j][b.k0], &buf[c], nk * sizeof(float));
	return 0;
}

// Field arrays: rows along k.  Rows are padded and shifted so that k = kMIN starts on a
// 128-byte boundary, except rows (ghosts included) no longer than a wavefront: those are
// packed end to end, where the padding and the ghost cell would cost a partly used sector
// per row and array, and the batched curl runs them in flat tiles across row ends, its
// planes padded to 256 bytes so that the tiles, which start at plane origins, keep every
// wavefront's access aligned (64 k-point worlds: E curl 17 % faster than unpadded planes
// tiled from the first active row).  (Longer rows lose when packed: 64-194-cell rows of the
// serving worlds ran 14 % slower.)
Lay fieldLayout(const alis_dev_world *D)
{
	const int align = ROW_ALIGN_BYTES / (int)D->esz;
	int koff = align;
	const int nkTot = koff + D->kNUM + 1;                      // offsets 0..koff+kNUM used
	long SK = ((nkTot + align - 1) / align) * align;
	if (D->kNUM + 2 <= 64) koff = 1, SK = D->kNUM + 2;
	Lay L;
	L.i0 = D->iMIN - 1, L.j0 = D->jMIN - 1, L.k0 = D->kMIN, L.koff = koff;
	L.sj = SK;
	L.si = (long)(D->jNUM + 2) * SK;
	if (koff == 1) {
		const long plane = 256 / (long)D->esz;
		L.si = (L.si + plane - 1) / plane * plane;
	}
	return L;
}

Box fieldBox(const alis_dev_world *D)
{
	return Box{D->iMIN - 1, D->iMAX + 1, D->jMIN - 1, D->jMAX + 1, D->kMIN - 1, D->kMAX + 1};
}

Lay compactLayout(const Box &b)
{
	Lay L;
	L.i0 = b.i0, L.j0 = b.j0, L.k0 = b.k0, L.koff = 0;
	L.sj = b.k1 - b.k0 + 1;
	L.si = (long)(b.j1 - b.j0 + 1) * L.sj;
	return L;
}

float *hostBlock(float ***f, const Box &b) { return &f[b.i0][b.j0][b.k0]; }

// Ghost-copy schedule of the reference (src/update.c:284-297 for E, :311-339 for H) as a
// list of copy operations in call order.
struct CopySpec { int axis; float c, s; Box b; int di, dj, dk; };

std::vector<CopySpec> ghostSchedule(const alis_dev_world_desc *W, int h)
{
	std::vector<CopySpec> v;
	const int iAlt = (W->xMinSur==SUR_PBC || W->xMinSur==SUR_BBC) ? W->iNUM-1 : 1;
	const int jAlt = (W->yMinSur==SUR_PBC || W->yMinSur==SUR_BBC) ? W->jNUM-1 : 1;
	const int kAlt = (W->zMinSur==SUR_PBC || W->zMinSur==SUR_BBC) ? W->kNUM-1 : 1;
	auto add = [&](int axis, float c, float s, int i0, int i1, int j0, int j1, int k0, int k1, int di, int dj, int dk) {
		v.push_back(CopySpec{axis, c, s, Box{i0, i1, j0, j1, k0, k1}, di, dj, dk});
	};
	if (h == 0) {
		add(0, W->coskx, -W->sinkx, W->iMAX+1, W->iMAX+1, W->jMIN, W->jMAX, W->kMIN, W->kMAX, -iAlt, 0, 0);
		add(2, W->coskz, -W->sinkz, W->iMIN, W->iMAX, W->jMIN, W->jMAX, W->kMAX+1, W->kMAX+1, 0, 0, -kAlt);
		if (W->yMinSur==SUR_HBC) {
			add(1, W->coskym, -W->sinkym, W->iMIN, 0, W->jMAX+1, W->jMAX+1, W->kMIN, W->kMAX, W->iMAX, 1-W->jNUM, 0);
			add(1, W->coskyp, -W->sinkyp, 1, W->iMAX+1, W->jMAX+1, W->jMAX+1, W->kMIN, W->kMAX, W->iMIN, 1-W->jNUM, 0);
		}
		else add(1, W->cosky, -W->sinky, W->iMIN, W->iMAX, W->jMAX+1, W->jMAX+1, W->kMIN, W->kMAX, 0, -jAlt, 0);
		return v;
	}
	const float cxm = W->xMaxSur==SUR_PEC ? 1 : W->coskx, czm = W->zMaxSur==SUR_PEC ? 1 : W->coskz, cym = W->yMaxSur==SUR_PEC ? 1 : W->cosky;
	add(1, cxm, -W->sinkx, W->iMAX+1, W->iMAX+1, W->jMIN, W->jMAX+1, W->kMIN, W->kMAX+1, -iAlt, 0, 0);
	add(2, cxm, -W->sinkx, W->iMAX+1, W->iMAX+1, W->jMIN, W->jMAX+1, W->kMIN, W->kMAX+1, -iAlt, 0, 0);
	add(0, czm, -W->sinkz, W->iMIN, W->iMAX+1, W->jMIN, W->jMAX+1, W->kMAX+1, W->kMAX+1, 0, 0, -kAlt);
	add(1, czm, -W->sinkz, W->iMIN, W->iMAX+1, W->jMIN, W->jMAX+1, W->kMAX+1, W->kMAX+1, 0, 0, -kAlt);
	add(1, W->coskx, W->sinkx, W->iMIN, W->iMIN, W->jMIN, W->jMAX+1, W->kMIN, W->kMAX+1, iAlt, 0, 0);
	add(2, W->coskx, W->sinkx, W->iMIN, W->iMIN, W->jMIN, W->jMAX+1, W->kMIN, W->kMAX+1, iAlt, 0, 0);
	add(0, W->coskz, W->sinkz, W->iMIN, W->iMAX+1, W->jMIN, W->jMAX+1, W->kMIN, W->kMIN, 0, 0, kAlt);
	add(1, W->coskz, W->sinkz, W->iMIN, W->iMAX+1, W->jMIN, W->jMAX+1, W->kMIN, W->kMIN, 0, 0, kAlt);
	if (W->yMinSur==SUR_HBC) {
		add(0, W->coskyp, W->sinkyp, W->iMIN, 0, W->jMIN, W->jMIN, W->kMIN, W->kMAX, W->iMAX, W->jNUM-1, 0);
		add(2, W->coskyp, W->sinkyp, W->iMIN, 0, W->jMIN, W->jMIN, W->kMIN, W->kMAX, W->iMAX, W->jNUM-1, 0);
		add(0, W->coskym, -W->sinkym, W->iMIN, 0, W->jMAX+1, W->jMAX+1, W->kMIN, W->kMAX, W->iMAX, 1-W->jNUM, 0);
		add(2, W->coskym, -W->sinkym, W->iMIN, 0, W->jMAX+1, W->jMAX+1, W->kMIN, W->kMAX, W->iMAX, 1-W->jNUM, 0);
		add(0, W->coskym, W->sinkym, 1, W->iMAX+1, W->jMIN, W->jMIN, W->kMIN, W->kMAX, W->iMIN, W->jNUM-1, 0);
		add(2, W->coskym, W->sinkym, 1, W->iMAX+1, W->jMIN, W->jMIN, W->kMIN, W->kMAX, W->iMIN, W->jNUM-1, 0);
		add(0, W->coskyp, -W->sinkyp, 1, W->iMAX+1, W->jMAX+1, W->jMAX+1, W->kMIN, W->kMAX, W->iMIN, 1-W->jNUM, 0);
		add(2, W->coskyp, -W->sinkyp, 1, W->iMAX+1, W->jMAX+1, W->jMAX+1, W->kMIN, W->kMAX, W->iMIN, 1-W->jNUM, 0);
	}
	else {
		add(2, cym, -W->sinky, W->iMIN, W->iMAX+1, W->jMAX+1, W->jMAX+1, W->kMIN, W->kMAX+1, 0, -jAlt, 0);
		add(0, cym, -W->sinky, W->iMIN, W->iMAX+1, W->jMAX+1, W->jMAX+1, W->kMIN, W->kMAX+1, 0, -jAlt, 0);
		add(2, W->cosky, W->sinky, W->iMIN, W->iMAX+1, W->jMIN, W->jMIN, W->kMIN, W->kMAX+1, 0, jAlt, 0);
		add(0, W->cosky, W->sinky, W->iMIN, W->iMAX+1, W->jMIN, W->jMIN, W->kMIN, W->kMAX+1, 0, jAlt, 0);
	}
	return v;
}

bool boxesMeet(const Box &a, const Box &b)
{
	return !(a.i1 < b.i0 || b.i1 < a.i0 || a.j1 < b.j0 || b.j1 < a.j0 || a.k1 < b.k0 || b.k1 < a.k0);
}

inline bool inBoxH(const Box &b, int i, int j, int k)
{
	return i >= b.i0 && i <= b.i1 && j >= b.j0 && j <= b.j1 && k >= b.k0 && k <= b.k1;
}

// Would a launch of the additive operations A and the copies C read a cell it writes?  The
// only reads of target arrays are the base cells of copy chains (see k_post): a base cell
// must be neither a copy destination nor an additive target of the same launch.
template<class T>
bool postConflict(const std::vector<PostOp<T>> &A, const std::vector<PostOp<T>> &C)
{
	const int n = (int)C.size();
	for (int p = 0; p < n; p++) {
		const Box &b = C[p].b;
		for (int i = b.i0; i <= b.i1; i++)
			for (int j = b.j0; j <= b.j1; j++)
				for (int k = b.k0; k <= b.k1; k++) {
					bool last = true;
					for (int q = p + 1; q < n && last; q++) if (C[q].F == C[p].F && inBoxH(C[q].b, i, j, k)) last = false;
					if (!last) continue;
					int ci = i, cj = j, ck = k, t = p;
					for (;;) {
						ci += C[t].di, cj += C[t].dj, ck += C[t].dk;
						int prev = -1;
						for (int u = t - 1; u >= 0; u--) if (C[u].F == C[p].F && inBoxH(C[u].b, ci, cj, ck)) { prev = u; break; }
						if (prev < 0) break;
						t = prev;
					}
					for (int q = 0; q < n; q++) if (C[q].F == C[p].F && inBoxH(C[q].b, ci, cj, ck)) return true;
					for (const PostOp<T> &o : A)
						if ((o.F == C[p].F || (C[p].G && o.F == C[p].G)) && inBoxH(o.b, ci, cj, ck)) return true;
				}
	}
	return false;
}

// The work of cell (i,j,k) of operation p in a launch's program v (additive operations
// first, nadd of them), as the key (kind, base offset, operation list); false when another
// piece computes the cell.  Follows the sequential semantics described at k_post.
template<class T>
bool cellWork(const std::vector<PostOp<T>> &v, int nadd, int p, int i, int j, int k, std::vector<int> &key)
{
	const PostOp<T> &o = v[p];
	const int n = (int)v.size();
	key.assign(4, 0);
	if (p < nadd) {
		for (int r = 0; r < p; r++) if (v[r].F == o.F && inBoxH(v[r].b, i, j, k)) return false;
		for (int q = nadd; q < n; q++) if ((v[q].F == o.F || v[q].G == o.F) && inBoxH(v[q].b, i, j, k)) return false;
		key[0] = OP_STRIP;
		for (int q = p; q < nadd; q++) if (v[q].F == o.F && inBoxH(v[q].b, i, j, k)) key.push_back(q);
		return true;
	}
	for (int q = p + 1; q < n; q++) if (v[q].F == o.F && inBoxH(v[q].b, i, j, k)) return false;
	int ci = i, cj = j, ck = k, q = p;
	for (;;) {
		key.push_back(q);
		ci += v[q].di, cj += v[q].dj, ck += v[q].dk;
		int prev = -1;
		for (int u = q - 1; u >= nadd; u--) if (v[u].F == o.F && inBoxH(v[u].b, ci, cj, ck)) { prev = u; break; }
		if (prev < 0) break;
		q = prev;
	}
	key[0] = OP_COPY, key[1] = ci - i, key[2] = cj - j, key[3] = ck - k;
	return true;
}

// Compile a launch's program into pieces: per operation, the cells of its box labelled by
// their work, then grown greedily into boxes of equal work (along k, then j, then i).
template<class T>
void buildPieces(const std::vector<PostOp<T>> &v, int nadd, std::vector<PostPiece> &out, long &total)
{
	std::map<std::vector<int>, int> ids;
	std::vector<std::vector<int>> keys;
	std::vector<int> key, last;
	int lastId = -1;
	total = 0;
	for (int p = 0; p < (int)v.size(); p++) {
		const Box &b = v[p].b;
		const int ni = b.i1 - b.i0 + 1, nj = b.j1 - b.j0 + 1, nk = b.k1 - b.k0 + 1;
		if (ni <= 0 || nj <= 0 || nk <= 0) continue;
		std::vector<int> sig((size_t)ni * nj * nk);
		size_t c = 0;
		for (int di = 0; di < ni; di++)
			for (int dj = 0; dj < nj; dj++)
				for (int dk = 0; dk < nk; dk++, c++) {
					if (!cellWork(v, nadd, p, b.i0 + di, b.j0 + dj, b.k0 + dk, key)) { sig[c] = -1; continue; }
					if (lastId < 0 || key != last) {
						auto it = ids.find(key);
						if (it == ids.end()) it = ids.emplace(key, (int)keys.size()).first, keys.push_back(key);
						lastId = it->second, last = key;
					}
					sig[c] = lastId;
				}
		auto at = [&](int di, int dj, int dk) -> int & { return sig[((size_t)di * nj + dj) * nk + dk]; };
		for (int di = 0; di < ni; di++)
			for (int dj = 0; dj < nj; dj++)
				for (int dk = 0; dk < nk; dk++) {
					const int id = at(di, dj, dk);
					if (id < 0) continue;
					int dk1 = dk, dj1 = dj, di1 = di;
					while (dk1 + 1 < nk && at(di, dj, dk1 + 1) == id) dk1++;
					auto rowSame = [&](int ii, int jj) {
						for (int kk = dk; kk <= dk1; kk++) if (at(ii, jj, kk) != id) return false;
						return true;
					};
					while (dj1 + 1 < nj && rowSame(di, dj1 + 1)) dj1++;
					for (;;) {
						if (di1 + 1 >= ni) break;
						bool same = true;
						for (int jj = dj; jj <= dj1 && same; jj++) same = rowSame(di1 + 1, jj);
						if (!same) break;
						di1++;
					}
					for (int ii = di; ii <= di1; ii++)
						for (int jj = dj; jj <= dj1; jj++)
							for (int kk = dk; kk <= dk1; kk++) at(ii, jj, kk) = -1;
					const std::vector<int> &K = keys[id];
					PostPiece P{};
					P.kind = K[0];
					P.si = K[1], P.sj = K[2], P.sk = K[3];
					P.n = (int)K.size() - 4;
					for (int u = 0; u < P.n; u++) P.list[u] = (unsigned char)K[4 + u];
					P.b = Box{b.i0 + di, b.i0 + di1, b.j0 + dj, b.j0 + dj1, b.k0 + dk, b.k0 + dk1};
					P.off = total;
					total += (P.b.count() + POST_WG - 1) / POST_WG * POST_WG;
					out.push_back(P);
				}
	}
}

// 1/eps per row segment of k_curl's tiles (64 cells along k): (x, y, z, 1) when every cell
// of the segment holds the same triple (bit patterns), else zeros.  A wavefront of the E
// curl covers one segment: on a uniform one it takes 1/eps from this table instead of
// reading three arrays (most segments of a world lie inside one material; the values, and
// so the results, are the same).
template<class T>
int rowSegments(alis_dev_world *D)
{
	const alis_dev_world_desc *W = &D->W;
	const int nkb = (D->kNUM + 63) / 64;
	const long nseg = (long)D->iNUM * D->jNUM * nkb;
	std::vector<T> tab(4 * nseg, T(0));
	long uniform = 0, mixedCells = 0;
	for (int i = D->iMIN; i <= D->iMAX; i++)
		for (int j = D->jMIN; j <= D->jMAX; j++)
			for (int b = 0; b < nkb; b++) {
				const int k0 = D->kMIN + 64 * b, k1 = std::min(k0 + 63, D->kMAX);
				bool same = true;
				for (int a = 0; a < 3 && same; a++) {
					const float *row = W->eps[a][i][j];
					for (int k = k0 + 1; k <= k1 && same; k++) same = !memcmp(&row[k], &row[k0], sizeof(float));
				}
				T *e = &tab[4 * (((long)(i - D->iMIN) * D->jNUM + (j - D->jMIN)) * nkb + b)];
				if (same) {
					for (int a = 0; a < 3; a++) e[a] = (T)W->eps[a][i][j][k0];
					e[3] = T(1);
					uniform++;
				}
				else mixedCells += k1 - k0 + 1;
			}
	// a table that spares the arrays in under a quarter of the cells (rows that all cross
	// material boundaries) only adds its own reads: the curl then reads the arrays alone
	const double cells = (double)D->iNUM * D->jNUM * D->kNUM;
	(void)uniform;
	if (mixedCells > 0.75 * cells) return 0;
	if (devAlloc(D, &D->seg, tab.size() * sizeof(T))) return -1;
	HIPCHECK(hipMemcpyAsync(D->seg, tab.data(), tab.size() * sizeof(T), hipMemcpyHostToDevice, D->st));
	HIPCHECK(hipStreamSynchronize(D->st));
	D->epsBytes = (mixedCells * 3.0 * D->esz + nseg * 4.0 * D->esz) / cells;
	return 0;
}

// Per-tile object lists for worlds with more dispersive objects than OBJ_SCAN: for every
// 64 x 4 tile of the curl (an x plane, 4 rows, 64 cells along k) the objects whose clipped
// box meets it, in ascending order, so that a cell visits only those (and adds their
// currents in the same order as a scan of all objects would).  Layout: ntile + 1 starts,
// then the lists.
constexpr int OBJ_SCAN = 2;

int objectTiles(alis_dev_world *D)
{
	if ((int)D->obj.size() <= OBJ_SCAN) return 0;
	const int njb = (D->jNUM + 3) / 4, nkb = (D->kNUM + 63) / 64;
	const long ntile = (long)D->iNUM * njb * nkb;
	std::vector<int> start(ntile + 1, 0);
	auto each = [&](const Box &b, auto &&f) {
		if (b.i1 < b.i0 || b.j1 < b.j0 || b.k1 < b.k0) return;
		for (int i = b.i0; i <= b.i1; i++)
			for (int jb = (b.j0 - D->jMIN) / 4; jb <= (b.j1 - D->jMIN) / 4; jb++)
				for (int kb = (b.k0 - D->kMIN) / 64; kb <= (b.k1 - D->kMIN) / 64; kb++)
					f(((long)(i - D->iMIN) * njb + jb) * nkb + kb);
	};
	for (const DevObj &O : D->obj) each(O.clip, [&](long t) { start[t + 1]++; });
	for (long t = 0; t < ntile; t++) start[t + 1] += start[t];
	std::vector<int> table(start.begin(), start.end());
	table.resize(ntile + 1 + start[ntile]);
	std::vector<int> fill(start.begin(), start.end() - 1);
	for (int n = 0; n < (int)D->obj.size(); n++)
		each(D->obj[n].clip, [&](long t) { table[ntile + 1 + fill[t]++] = n; });
	if (devAlloc(D, &D->objTile, table.size() * sizeof(int))) return -1;
	HIPCHECK(hipMemcpyAsync(D->objTile, table.data(), table.size() * sizeof(int), hipMemcpyHostToDevice, D->st));
	HIPCHECK(hipStreamSynchronize(D->st));
	return 0;
}

// Commit: allocate and upload everything of the descriptor.
// The device table of D's dispersive objects (ObjDev<T>), naming the currents of its
// commit-time set (alt false) or of its other set (alt true, see k_step).
template<class T>
int uploadObjTable(alis_dev_world *D, bool alt, void **dst)
{
	std::vector<ObjDev<T>> tab(D->obj.size());
	for (size_t n = 0; n < D->obj.size(); n++) {
		const DevObj &O = D->obj[n];
		ObjDev<T> &t = tab[n];
		memset(&t, 0, sizeof t);
		t.i0 = O.clip.i0, t.i1 = O.clip.i1, t.j0 = O.clip.j0, t.j1 = O.clip.j1, t.k0 = O.clip.k0, t.k1 = O.clip.k1;
		t.LO = O.LO;
		t.NJ = O.NJ, t.NK = O.NK;
		t.scale = O.a[0] || O.a[1] || O.a[2];
		for (int c = 0; c < 3; c++) {
			t.a[c] = (T)O.a[c];
			t.fill[c] = (const T *)O.fill[c];
			for (int p = 0; p < MAXPOLE; p++) {
				t.JJ[c][p] = (T)O.JJ[c][p], t.JK[c][p] = (T)O.JK[c][p], t.JE[c][p] = (T)O.JE[c][p];
				t.KJ[c][p] = (T)O.KJ[c][p], t.KK[c][p] = (T)O.KK[c][p], t.KE[c][p] = (T)O.KE[c][p];
				for (int m = 0; m <= D->cplx; m++) {
					t.J[m][c][p] = (T *)(alt ? O.Jalt : O.J)[m][c][p];
					t.K[m][c][p] = (T *)(alt ? O.Kalt : O.K)[m][c][p];
				}
			}
		}
	}
	if (devAlloc(D, dst, tab.size() * sizeof(ObjDev<T>))) return -1;
	HIPCHECK(hipMemcpyAsync(*dst, tab.data(), tab.size() * sizeof(ObjDev<T>), hipMemcpyHostToDevice, D->st));
	HIPCHECK(hipStreamSynchronize(D->st));
	return 0;
}

template<class T>
int commitWorld(alis_dev_world *D)
{
	const alis_dev_world_desc *W = &D->W;
	D->L = fieldLayout(D);
	D->fsize = (long)(D->iNUM + 2) * D->L.si;
	const Box fb = fieldBox(D);

	for (int h = 0; h < 2; h++)
		for (int m = 0; m <= D->cplx; m++)
			for (int a = 0; a < 3; a++) {
				if (!W->field[h][m][a]) return fail("field array missing in the descriptor");
				if (devAlloc(D, &D->F[h][m][a], D->fsize * sizeof(T))) return -1;
				if (uploadBox<T>(D, D->F[h][m][a], D->L, fb, hostBlock(W->field[h][m][a], fb))) return -1;
			}

	// inverse permittivity (reference extents: one extra node along the component axis)
	if (W->eps[0] && W->eps[1] && W->eps[2]) {
		for (int a = 0; a < 3; a++) {
			Box b{D->iMIN, D->iMAX + (a==0), D->jMIN, D->jMAX + (a==1), D->kMIN, D->kMAX + (a==2)};
			if (devAlloc(D, &D->C[a], D->fsize * sizeof(T))) return -1;
			if (uploadBox<T>(D, D->C[a], D->L, b, hostBlock(W->eps[a], b))) return -1;
		}
		if (rowSegments<T>(D)) return -1;
	}

	// per-axis step ratios and CPML coefficients
	const int NUMs[3] = { D->iNUM, D->jNUM, D->kNUM };
	for (int h = 0; h < 2; h++)
		for (int a = 0; a < 3; a++) {
			if (upload1<T>(D, &D->dt[h][a], W->dt[h][a], NUMs[a])) return -1;
			if (W->cpml[h][a][0])
			{
				for (int c = 0; c < 3; c++) if (upload1<T>(D, &D->cp[h][a][c], W->cpml[h][a][c], NUMs[a])) return -1;
				std::vector<float> packed(4 * (size_t)NUMs[a], 0.f);
				for (int q = 0; q < NUMs[a]; q++)
					for (int c = 0; c < 3; c++) packed[4 * q + c] = W->cpml[h][a][c][q];
				if (upload1<T>(D, &D->cpk[h][a], packed.data(), 4 * NUMs[a])) return -1;
			}
		}

	// CPML slabs: ranges per face for E (Alt=+1, Offset 0) and H (Alt=-1, Offset 1)
	const int MINa[3] = { D->iMIN, D->jMIN, D->kMIN }, MAXa[3] = { D->iMAX, D->jMAX, D->kMAX };
	const int Mina[3] = { D->iMin, D->jMin, D->kMin }, Maxa[3] = { D->iMax, D->jMax, D->kMax };
	for (int h = 0; h < 2; h++) {
		const int alt = h ? -1 : 1;
		for (int f = 0; f < 6; f++) {
			const int ax = f % 3;
			DevSlab &S = D->slab[h][f];
			const bool minFace = f < 3;
			S.on = minFace ? MINa[ax] < Mina[ax] : MAXa[ax] > Maxa[ax];
			if (!S.on) continue;
			if (!D->cp[h][ax][0]) return fail("CPML face without a CPML profile");
			S.lo = minFace ? MINa[ax] - (alt - 1) / 2 : Maxa[ax] + 2;
			S.hi = minFace ? Mina[ax] - (alt + 3) / 2 : MAXa[ax];
			Box b{D->iMIN, D->iMAX, D->jMIN, D->jMAX, D->kMIN, D->kMAX};
			if (ax == 0) b.i0 = S.lo, b.i1 = S.hi;
			if (ax == 1) b.j0 = S.lo, b.j1 = S.hi;
			if (ax == 2) b.k0 = S.lo, b.k1 = S.hi;
			S.b = b;
			if (ax == 2) S.L = compactLayout(b);
			else {
				S.L = D->L;
				S.L.i0 = b.i0, S.L.j0 = b.j0;
				S.L.si = (long)(b.j1 - b.j0 + 1) * D->L.sj;
			}
			const long n = (long)(b.i1 - b.i0 + 1) * S.L.si;
			S.n = n;
			for (int m = 0; m <= D->cplx; m++)
				for (int s2 = 0; s2 < 2; s2++) {
					float ***hostp = W->slab[h][m][f][s2];
					if (!hostp) return fail("CPML slab missing on the host");
					if (devAlloc(D, &S.p[m][s2], n * sizeof(T))) return -1;
					if (uploadBox<T>(D, S.p[m][s2], S.L, b, hostBlock(hostp, b))) return -1;
				}
		}
	}

	// dispersive objects
	for (int n = 0; n < W->nobj; n++) {
		const alis_dev_ade &A = W->obj[n];
		DevObj O;
		O.box = Box{A.i0, A.i1, A.j0, A.j1, A.k0, A.k1};
		O.clip = Box{std::max(D->iMIN, A.i0), std::min(D->iMAX, A.i1),
		             std::max(D->jMIN, A.j0), std::min(D->jMAX, A.j1),
		             std::max(D->kMIN, A.k0), std::min(D->kMAX, A.k1)};
		O.LO = compactLayout(O.box);
		O.NJ = A.nj, O.NK = A.nk;
		if (O.NK > MAXPOLE || O.NJ > O.NK) return fail("bad pole counts");
		for (int c = 0; c < 3; c++) {
			O.a[c] = A.a[c];
			for (int p = 0; p < O.NK; p++) {
				O.JJ[c][p] = A.JJ[c][p], O.JE[c][p] = A.JE[c][p];
				if (p >= O.NJ) O.JK[c][p] = A.JK[c][p], O.KJ[c][p] = A.KJ[c][p], O.KK[c][p] = A.KK[c][p], O.KE[c][p] = A.KE[c][p];
			}
		}
		const long n3 = O.box.count();
		for (int c = 0; c < 3; c++) {
			if (devAlloc(D, &O.fill[c], n3 * sizeof(T))) return -1;
			if (uploadBox<T>(D, O.fill[c], O.LO, O.box, hostBlock(A.fill[c], O.box))) return -1;
		}
		for (int m = 0; m <= D->cplx; m++)
			for (int c = 0; c < 3; c++)
				for (int p = 0; p < O.NK; p++) {
					if (devAlloc(D, &O.J[m][c][p], n3 * sizeof(T))) return -1;
					if (uploadBox<T>(D, O.J[m][c][p], O.LO, O.box, hostBlock(A.J[m][c][p], O.box))) return -1;
					if (p >= O.NJ) {
						if (devAlloc(D, &O.K[m][c][p], n3 * sizeof(T))) return -1;
						if (uploadBox<T>(D, O.K[m][c][p], O.LO, O.box, hostBlock(A.K[m][c][p], O.box))) return -1;
					}
				}
		D->obj.push_back(O);
	}
	if (!D->obj.empty()) {
		if (uploadObjTable<T>(D, false, &D->objTab)) return -1;
		if (objectTiles(D)) return -1;
	}

	// sources
	for (int h = 0; h < 2; h++)
		for (int q = 0; q < W->nsrc[h]; q++) {
			const alis_dev_src &S = W->src[h][q];
			DevSrc s;
			s.h = h, s.m = S.m, s.axis = S.axis;
			if (S.h != h || S.m < 0 || S.m > D->cplx || S.axis < 0 || S.axis > 2) return fail("source targets an array outside this world");
			s.b = Box{S.i0, S.i1, S.j0, S.j1, S.k0, S.k1};
			const Box fbx = fieldBox(D);
			if (s.b.i0 < fbx.i0 || s.b.i1 > fbx.i1 || s.b.j0 < fbx.j0 || s.b.j1 > fbx.j1 || s.b.k0 < fbx.k0 || s.b.k1 > fbx.k1)
				return fail("source box outside the grid");
			s.LB = compactLayout(s.b);
			s.ncell = s.b.count();
			s.amp = S.amplitude;
			if (S.Amplitude) {
				if (devAlloc(D, &s.ampArr, s.ncell * sizeof(T))) return -1;
				if (uploadBox<T>(D, s.ampArr, s.LB, s.b, hostBlock(S.Amplitude, s.b))) return -1;
			}
			if (S.cellPhase) {
				s.cellWave = true;
				if (devAlloc(D, &s.waveDev, s.ncell * sizeof(T))) return -1;
				HIPCHECK(hipHostMalloc((void **)&s.wavePinned, s.ncell * sizeof(T), hipHostMallocDefault));
				HIPCHECK(hipEventCreateWithFlags(&s.waveDone, hipEventDisableTiming));
				HIPCHECK(hipEventRecord(s.waveDone, D->st));
			}
			D->src[h].push_back(s);
		}

	for (int h = 0; h < 2; h++) {
		D->pendSrc[h].assign((D->src[h].size() + 31) / 32, 0u);
		D->pendWave[h].assign(D->src[h].size(), 0.0);
	}
	HIPCHECK(hipStreamSynchronize(D->st));
	return 0;
}

template<class T>
CurlArgs<T> curlArgs(alis_dev_world *D, int h, int m)
{
	CurlArgs<T> a{};
	a.F0 = (T *)D->F[h][m][0], a.F1 = (T *)D->F[h][m][1], a.F2 = (T *)D->F[h][m][2];
	a.O0 = a.F0, a.O1 = a.F1, a.O2 = a.F2;
	a.G0 = (const T *)D->F[1-h][m][0], a.G1 = (const T *)D->F[1-h][m][1], a.G2 = (const T *)D->F[1-h][m][2];
	if (h == 0) a.C0 = (const T *)D->C[0], a.C1 = (const T *)D->C[1], a.C2 = (const T *)D->C[2];
	if (h == 0) a.seg = (const typename Quad<T>::V *)D->seg;
	a.dtx = (const T *)D->dt[h][0], a.dty = (const T *)D->dt[h][1], a.dtz = (const T *)D->dt[h][2];
	a.L = D->L;
	a.iMIN = D->iMIN, a.jMIN = D->jMIN, a.kMIN = D->kMIN;
	a.ni = D->iNUM, a.nj = D->jNUM, a.nk = D->kNUM;
	a.alt = h ? -1 : 1;
	for (int f = 0; f < 6; f++) {
		const DevSlab &S = D->slab[h][f];
		CurlFace<T> &P = a.face[f];
		P.on = S.on;
		if (!S.on) continue;
		const int ax = f % 3;
		P.pa = (T *)S.p[m][0], P.pb = (T *)S.p[m][1];
		P.qa = P.pa, P.qb = P.pb;
		P.bc = (const T *)D->cpk[h][ax];
		P.L = S.L, P.lo = S.lo, P.hi = S.hi;
	}
	a.obj = a.objOut = (const ObjDev<T> *)D->objTab;
	a.nobj = (int)D->obj.size();
	if (D->objTile) {
		const long ntile = (long)D->iNUM * ((D->jNUM + 3) / 4) * ((D->kNUM + 63) / 64);
		a.otStart = (const int *)D->objTile, a.otList = (const int *)D->objTile + ntile + 1;
	}
	a.m = m;
	return a;
}

// A pair of timing events of D: reused ones where alis_dev_timing_read has read them.
int timingEvents(alis_dev_world *D, hipEvent_t &e0, hipEvent_t &e1)
{
	for (hipEvent_t *e : {&e0, &e1}) {
		if (!D->evFree.empty()) *e = D->evFree.back(), D->evFree.pop_back();
		else HIPCHECK(hipEventCreate(e));
	}
	return 0;
}

// The phasers' samples the last H curl stored no longer hold once the fields change.
void dropPresamples(alis_dev_world *D)
{
	for (DevPhaser &P : D->phs) P.pre = false;
}

// The z faces of D's phasers that the H curl can sample (PhsZ): a face whose k + 1 lies in
// the same wavefront as k (k_curl's waves cover 64 cells from kMIN on), over the rectangle of
// its cells that the end-of-H-half program leaves alone at k and k + 1 (shrunk from the edges
// its boxes reach: the ghost planes of periodic boundaries), so the values the curl has are
// the ones the phaser update would read; k_phs_sample samples the other cells.
template<class T>
int zSamples(alis_dev_world *D, PhsZ<T> *zs, std::vector<DevPhaser *> &owners)
{
	int n = 0;
	if (!D->postBuilt) return 0;
	for (DevPhaser &P : D->phs) {
		if (!P.live || !P.ncell) continue;
		const int before = n;
		bool ok = true;
		for (int f = 0; f < P.nface && ok; f++) {
			const PhsFaceGeom &g = P.face[f];
			if (g.kind < 4) continue;
			const int k = g.fixed;
			if ((k - D->kMIN) % 64 == 63 || k + 1 > D->kMAX || n == MAXZS) { ok = false; break; }
			int i0 = P.iMin, i1 = P.iMax, j0 = P.jMin, j1 = P.jMax;
			for (const Box &b : D->hWrites) {
				if (b.k1 < k || b.k0 > k + 1 || b.i1 < i0 || b.i0 > i1 || b.j1 < j0 || b.j0 > j1) continue;
				if (b.i1 - b.i0 < b.j1 - b.j0 && b.i1 < (i0 + i1) / 2) i0 = b.i1 + 1;          // a plane at low i
				else if (b.i1 - b.i0 < b.j1 - b.j0 && b.i0 > (i0 + i1) / 2) i1 = b.i0 - 1;
				else if (b.j1 < (j0 + j1) / 2) j0 = b.j1 + 1;
				else if (b.j0 > (j0 + j1) / 2) j1 = b.j0 - 1;
				else ok = false;
			}
			if (i0 > i1 || j0 > j1) ok = false;
			zs[n++] = PhsZ<T>{k, P.iMin, P.iMax, P.jMin, P.jMax, i0, i1, j0, j1, P.ncell, g.off,
			                  (T *)P.ring + (size_t)P.pend * 4 * P.ncell};
		}
		if (!ok || n == before) {
			n = before;
			continue;
		}
		// the rectangles, for k_phs_sample to skip
		for (int f = 0, z = before; f < P.nface; f++) {
			PhsFaceGeom &g = P.face[f];
			if (g.kind < 4) continue;
			g.si0 = zs[z].si0, g.si1 = zs[z].si1, g.sj0 = zs[z].sj0, g.sj1 = zs[z].sj1;
			z++;
		}
		owners.push_back(&P);
	}
	// worth it where the sampled cells are many: the sampling code costs every H wavefront
	// a little (256^3 nanoblock, 9 800 face cells: H curl +3 us against 0.7 us saved), the
	// strided reads it replaces cost a cache line a value (641x641x320 metasurface, two faces
	// of 408 000 cells: 70 us a step saved)
	// (ALIS_PRESAMPLE=force: every eligible face, 0: none)
	long cells = 0;
	for (int z = 0; z < n; z++) cells += (long)(zs[z].si1 - zs[z].si0 + 1) * (zs[z].sj1 - zs[z].sj0 + 1);
	const char *env = getenv("ALIS_PRESAMPLE");
	const bool force = env && !strcmp(env, "force"), off = env && !strcmp(env, "0");
	if (off || (!force && 2 * cells < (long)D->iNUM * D->jNUM)) {
		owners.clear();
		return 0;
	}
	return n;
}

template<class T>
int coreHalf(alis_dev_world *D, int h)
{
	dropPresamples(D);
	hipEvent_t e0 = nullptr, e1 = nullptr;
	if (D->timing) {
		if (timingEvents(D, e0, e1)) return -1;
	}
	// E: grid (k tiles, y tiles, x planes); H: rows dealt to XCDs (k_curl's ROWS), rows
	// padded to a multiple of 8
	const long rows = (long)((D->jNUM + 3) / 4) * D->iNUM;
	const dim3 grid = h == 0 ? dim3((D->kNUM + 63) / 64, (D->jNUM + 3) / 4, D->iNUM)
	                         : dim3((unsigned)((rows + 7) / 8 * 8 * ((D->kNUM + 63) / 64))), block(64, 4);
	// timing events are bound to the launches themselves (their start and end), so timing
	// adds no marker between launches
	std::vector<DevPhaser *> sampled;
	for (int m = 0; m <= D->cplx; m++) {
		CurlArgs<T> a = curlArgs<T>(D, h, m);
		if (h == 1 && m == 0) a.nzs = zSamples<T>(D, a.zs, sampled);   // phasers sample part 0
		hipEvent_t s0 = m == 0 ? e0 : nullptr, s1 = m == D->cplx ? e1 : nullptr;
		if (h == 0 && D->seg) launchEv(k_curl<T, 2, false>, grid, block, 0, D->st, s0, s1, a);
		else if (h == 0 && D->C[0]) launchEv(k_curl<T, 1, false>, grid, block, 0, D->st, s0, s1, a);
		else if (h == 0) launchEv(k_curl<T, 0, false>, grid, block, 0, D->st, s0, s1, a);
		else if (a.nzs) launchEv(k_curl<T, 0, true, true, false, 4, true>, grid, block, 0, D->st, s0, s1, a);
		else launchEv(k_curl<T, 0, true>, grid, block, 0, D->st, s0, s1, a);
	}
	if (e1) {
		D->evPairs[h].push_back(e0);
		D->evPairs[h].push_back(e1);
	}
	HIPCHECK(hipGetLastError());
	for (DevPhaser *P : sampled) P->pre = true, P->preSlot = P->pend;
	return 0;
}

// TF/SF strips of box s (src/update.c:205-272): F += dt/ds * G_aux * (1/eps) on the 12 face
// strips, with auxiliary indices beyond its last row aliased as in the reference.
template<class T>
int stripOps(alis_dev_world *D, int s, int h, std::vector<PostOp<T>> &ops)
{
	if (s >= (int)D->tf.size() || !D->tf[s].aux) return fail("TF/SF box not linked");
	const alis_dev_world_desc *W = &D->W;
	const alis_dev_tfsf *S = &D->tf[s].S;
	alis_dev_world *A = D->tf[s].aux;
	const alis_dev_world_desc *AW = &A->W;
	static const int altY[6] = { 0, 1, 0, 1, 0, 1 }, altX[6] = { 1, 0, 0, 0, 1, 1 };
	const size_t first = ops.size();
	const float *dtx = W->dt[h][0] - W->iMIN, *dty = W->dt[h][1] - W->jMIN, *dtz = W->dt[h][2] - W->kMIN;
	auto strip = [&](int m, int fa, int ga, float dtds, int i0, int i1, int j0, int j1, int k0, int k1, int di, int dj, int dk) {
		PostOp<T> o{};
		o.kind = OP_STRIP;
		o.slot = s;
		o.F = (T *)D->F[h][m][fa];
		o.A = (const T *)A->F[1-h][m][ga];
		o.C = h == 0 ? (const T *)D->C[fa] : nullptr;
		o.LF = D->L, o.LA = A->L, o.LC = D->L;
		o.b = Box{i0, i1, j0, j1, k0, k1};
		o.di = di, o.dj = dj, o.dk = dk;
		o.c = (T)dtds;
		const int comp = (1-h) * 3 + ga;   // aux component index E.x..H.z
		o.exX = AW->expandX, o.exY = AW->expandY;
		o.auxIMAX = AW->iMAX, o.auxJMAX = AW->jMAX;
		o.altX = altX[comp], o.altY = altY[comp];
		ops.push_back(o);
	};
	for (int m = 0; m <= W->complexField; m++) {
		if (m > AW->complexField) return fail("complex main world with a real auxiliary world");
		if (h == 0) {
			if (S->iMin > W->iMIN) {
				strip(m, 2, 1, -dtx[S->iMin], S->iMin, S->iMin, S->jMin, S->jMax, S->kMIN, S->kMAX, 0, 0, 0);
				strip(m, 1, 2, +dtx[S->iMin], S->iMin, S->iMin, S->jMIN, S->jMAX, S->kMin, S->kMax, 0, 0, 0);
			}
			if (S->jMin > W->jMIN) {
				strip(m, 0, 2, -dty[S->jMin], S->iMIN, S->iMAX, S->jMin, S->jMin, S->kMin, S->kMax, 0, 0, 0);
				strip(m, 2, 0, +dty[S->jMin], S->iMin, S->iMax, S->jMin, S->jMin, S->kMIN, S->kMAX, 0, 0, 0);
			}
			if (S->kMin > W->kMIN) {
				strip(m, 1, 0, -dtz[S->kMin], S->iMin, S->iMax, S->jMIN, S->jMAX, S->kMin, S->kMin, 0, 0, 0);
				strip(m, 0, 1, +dtz[S->kMin], S->iMIN, S->iMAX, S->jMin, S->jMax, S->kMin, S->kMin, 0, 0, 0);
			}
			if (S->iMax < W->iMAX) {
				strip(m, 2, 1, +dtx[S->iMax], S->iMax, S->iMax, S->jMin, S->jMax, S->kMIN, S->kMAX, 1, 0, 0);
				strip(m, 1, 2, -dtx[S->iMax], S->iMax, S->iMax, S->jMIN, S->jMAX, S->kMin, S->kMax, 1, 0, 0);
			}
			if (S->jMax < W->jMAX) {
				strip(m, 0, 2, +dty[S->jMax], S->iMIN, S->iMAX, S->jMax, S->jMax, S->kMin, S->kMax, 0, 1, 0);
				strip(m, 2, 0, -dty[S->jMax], S->iMin, S->iMax, S->jMax, S->jMax, S->kMIN, S->kMAX, 0, 1, 0);
			}
			if (S->kMax < W->kMAX) {
				strip(m, 1, 0, +dtz[S->kMax], S->iMin, S->iMax, S->jMIN, S->jMAX, S->kMax, S->kMax, 0, 0, 1);
				strip(m, 0, 1, -dtz[S->kMax], S->iMIN, S->iMAX, S->jMin, S->jMax, S->kMax, S->kMax, 0, 0, 1);
			}
		}
		else {
			if (S->iMin > W->iMIN) {
				strip(m, 2, 1, +dtx[S->iMin], S->iMin, S->iMin, S->jMIN, S->jMAX, S->kMin, S->kMax, 0, 0, 0);
				strip(m, 1, 2, -dtx[S->iMin], S->iMin, S->iMin, S->jMin, S->jMax, S->kMIN, S->kMAX, 0, 0, 0);
			}
			if (S->jMin > W->jMIN) {
				strip(m, 0, 2, +dty[S->jMin], S->iMin, S->iMax, S->jMin, S->jMin, S->kMIN, S->kMAX, 0, 0, 0);
				strip(m, 2, 0, -dty[S->jMin], S->iMIN, S->iMAX, S->jMin, S->jMin, S->kMin, S->kMax, 0, 0, 0);
			}
			if (S->kMin > W->kMIN) {
				strip(m, 1, 0, +dtz[S->kMin], S->iMIN, S->iMAX, S->jMin, S->jMax, S->kMin, S->kMin, 0, 0, 0);
				strip(m, 0, 1, -dtz[S->kMin], S->iMin, S->iMax, S->jMIN, S->jMAX, S->kMin, S->kMin, 0, 0, 0);
			}
			if (S->iMax < W->iMAX) {
				strip(m, 2, 1, -dtx[S->iMax+1], S->iMax+1, S->iMax+1, S->jMIN, S->jMAX, S->kMin, S->kMax, -1, 0, 0);
				strip(m, 1, 2, +dtx[S->iMax+1], S->iMax+1, S->iMax+1, S->jMin, S->jMax, S->kMIN, S->kMAX, -1, 0, 0);
			}
			if (S->jMax < W->jMAX) {
				strip(m, 0, 2, -dty[S->jMax+1], S->iMin, S->iMax, S->jMax+1, S->jMax+1, S->kMIN, S->kMAX, 0, -1, 0);
				strip(m, 2, 0, +dty[S->jMax+1], S->iMIN, S->iMAX, S->jMax+1, S->jMax+1, S->kMin, S->kMax, 0, -1, 0);
			}
			if (S->kMax < W->kMAX) {
				strip(m, 1, 0, -dtz[S->kMax+1], S->iMIN, S->iMAX, S->jMin, S->jMax, S->kMax+1, S->kMax+1, 0, 0, -1);
				strip(m, 0, 1, +dtz[S->kMax+1], S->iMin, S->iMax, S->jMIN, S->jMAX, S->kMax+1, S->kMax+1, 0, 0, -1);
			}
		}
	}
	// every strip must stay inside both grids
	const Box fb = fieldBox(D);
	for (size_t n = first; n < ops.size(); n++) {
		const PostOp<T> &o = ops[n];
		if (o.b.count() == 0) continue;
		if (o.b.i0 < fb.i0 || o.b.i1 > fb.i1 || o.b.j0 < fb.j0 || o.b.j1 > fb.j1 || o.b.k0 < fb.k0 || o.b.k1 > fb.k1)
			return fail("TF/SF strip outside the grid");
		const int ia0 = o.b.i0 + o.di, ja0 = o.b.j0 + o.dj, ka0 = o.b.k0 + o.dk, ka1 = o.b.k1 + o.dk;
		const int ia1 = AW->expandX ? std::min(o.b.i1 + o.di, AW->iMAX + 1) : o.b.i1 + o.di;
		const int ja1 = AW->expandY ? std::min(o.b.j1 + o.dj, AW->jMAX + 1) : o.b.j1 + o.dj;
		if (ia0 < AW->iMIN - 1 || ia1 > AW->iMAX + 1 || ja0 < AW->jMIN - 1 || ja1 > AW->jMAX + 1 || ka0 < AW->kMIN - 1 || ka1 > AW->kMAX + 1)
			return fail("TF/SF box reaches outside its auxiliary grid");
	}
	return 0;
}

// The parts of box b outside box r, as disjoint boxes.
std::vector<Box> outsideParts(Box b, const Box &r)
{
	std::vector<Box> out;
	auto take = [&](const Box &x) { if (x.count()) out.push_back(x); };
	take(Box{b.i0, std::min(b.i1, r.i0 - 1), b.j0, b.j1, b.k0, b.k1});
	take(Box{std::max(b.i0, r.i1 + 1), b.i1, b.j0, b.j1, b.k0, b.k1});
	b.i0 = std::max(b.i0, r.i0), b.i1 = std::min(b.i1, r.i1);
	if (b.i0 > b.i1) return out;
	take(Box{b.i0, b.i1, b.j0, std::min(b.j1, r.j0 - 1), b.k0, b.k1});
	take(Box{b.i0, b.i1, std::max(b.j0, r.j1 + 1), b.j1, b.k0, b.k1});
	b.j0 = std::max(b.j0, r.j0), b.j1 = std::min(b.j1, r.j1);
	if (b.j0 > b.j1) return out;
	take(Box{b.i0, b.i1, b.j0, b.j1, b.k0, std::min(b.k1, r.k0 - 1)});
	take(Box{b.i0, b.i1, b.j0, b.j1, std::max(b.k0, r.k1 + 1), b.k1});
	return out;
}

// The H cells a fused step computes again after the E program (k_fix): H at a cell reads E
// there and one cell below along each axis, so each box the program writes, grown by one
// cell upwards and clipped to the update range.  Boxes that overlap much (the two strips of
// one TF/SF face) are merged, and then the pairs whose union adds the fewest cells until at
// most MAXFIX remain; computing H again on a cell the program did not reach reproduces the
// value k_step wrote.
std::vector<Box> fixBoxes(const std::vector<Box> &ops, const Box &range)
{
	std::vector<Box> v;
	for (Box b : ops) {
		b = Box{std::max(b.i0, range.i0), std::min(b.i1 + 1, range.i1), std::max(b.j0, range.j0), std::min(b.j1 + 1, range.j1),
		        std::max(b.k0, range.k0), std::min(b.k1 + 1, range.k1)};
		if (b.count()) v.push_back(b);
	}
	auto unite = [](const Box &a, const Box &b) {
		return Box{std::min(a.i0, b.i0), std::max(a.i1, b.i1), std::min(a.j0, b.j0), std::max(a.j1, b.j1),
		           std::min(a.k0, b.k0), std::max(a.k1, b.k1)};
	};
	for (bool merged = true; merged;) {
		merged = false;
		for (size_t x = 0; x < v.size() && !merged; x++)
			for (size_t y = x + 1; y < v.size() && !merged; y++)
				if (unite(v[x], v[y]).count() * 10 <= (v[x].count() + v[y].count()) * 11) {
					v[x] = unite(v[x], v[y]);
					v.erase(v.begin() + y);
					merged = true;
				}
	}
	while (v.size() > (size_t)MAXFIX) {
		size_t bx = 0, by = 1;
		long best = -1;
		for (size_t x = 0; x < v.size(); x++)
			for (size_t y = x + 1; y < v.size(); y++) {
				const long grow = unite(v[x], v[y]).count() - v[x].count() - v[y].count();
				if (best < 0 || grow < best) best = grow, bx = x, by = y;
			}
		v[bx] = unite(v[bx], v[by]);
		v.erase(v.begin() + by);
	}
	return v;
}

// Build the end-of-half-step programs of D: strips of every TF/SF box, sources of the
// chain, ghost copies; split into launches where a copy chain would read a cell that its
// launch writes.
template<class T>
int buildPost(alis_dev_world *D)
{
	if ((int)D->tf.size() > 32) return fail("more than 32 TF/SF boxes in one world");
	std::vector<std::vector<PostOp<T>>> segOps[2];
	std::vector<int> segAdd[2];
	const Box range{D->iMIN, D->iMAX, D->jMIN, D->jMAX, D->kMIN, D->kMAX};
	std::vector<Box> eOps;      // the cells the E program writes
	std::vector<PostOp<T>> copyOps[2];
	D->fixOk = true;
	for (int h = 0; h < 2; h++) {
		if ((int)D->src[h].size() > MAXSRC) return fail("more than 128 sources in one chain");
		std::vector<PostOp<T>> adds, copies;
		for (int s = 0; s < (int)D->tf.size(); s++)
			if (stripOps<T>(D, s, h, adds)) return -1;
		for (int q = 0; q < (int)D->src[h].size(); q++) {
			const DevSrc &S = D->src[h][q];
			PostOp<T> o{};
			o.kind = OP_SOURCE;
			o.slot = q;
			o.F = (T *)D->F[h][S.m][S.axis];
			o.LF = D->L;
			o.C = (const T *)S.ampArr;
			o.W = S.cellWave ? (const T *)S.waveDev : nullptr;
			o.LC = S.LB;
			o.b = S.b;
			o.c = (T)S.amp;
			adds.push_back(o);
		}
		for (const CopySpec &c : ghostSchedule(&D->W, h)) {
			PostOp<T> o{};
			o.kind = OP_COPY;
			o.F = (T *)D->F[h][0][c.axis];
			o.G = D->cplx ? (T *)D->F[h][1][c.axis] : nullptr;
			o.LF = D->L;
			o.b = c.b;
			o.di = c.di, o.dj = c.dj, o.dk = c.dk;
			o.c = (T)c.c, o.s = (T)c.s;
			copies.push_back(o);
		}
		auto nonEmpty = [](std::vector<PostOp<T>> &v) {
			v.erase(std::remove_if(v.begin(), v.end(), [](const PostOp<T> &o) { return o.b.count() == 0; }), v.end());
		};
		nonEmpty(adds), nonEmpty(copies);
		// fused steps keep the two sets equal on the cells no step writes, which an additive
		// operation outside the update range (applied on some steps only) would break
		for (const PostOp<T> &o : adds)
			if (o.b.i0 < range.i0 || o.b.i1 > range.i1 || o.b.j0 < range.j0 || o.b.j1 > range.j1 || o.b.k0 < range.k0 || o.b.k1 > range.k1)
				D->fixOk = false;
		if (h == 0) {
			for (const PostOp<T> &o : adds) eOps.push_back(o.b);
			for (const PostOp<T> &o : copies) eOps.push_back(o.b);
		}
		else {
			D->hWrites.clear();
			for (const PostOp<T> &o : adds) D->hWrites.push_back(o.b);
			for (const PostOp<T> &o : copies) D->hWrites.push_back(o.b);
		}
		copyOps[h] = copies;
		if (copies.size() <= (size_t)MAXCHAIN && adds.size() + copies.size() <= (size_t)MAXPOST && !postConflict(adds, copies)) {
			std::vector<PostOp<T>> all = adds;
			all.insert(all.end(), copies.begin(), copies.end());
			if (!all.empty()) segOps[h].push_back(all), segAdd[h].push_back((int)adds.size());
			continue;
		}
		for (size_t a0 = 0; a0 < adds.size(); a0 += MAXPOST) {
			std::vector<PostOp<T>> chunk(adds.begin() + a0, adds.begin() + std::min(adds.size(), a0 + MAXPOST));
			segOps[h].push_back(chunk), segAdd[h].push_back((int)chunk.size());
		}
		std::vector<PostOp<T>> group;
		const std::vector<PostOp<T>> none;
		for (const PostOp<T> &o : copies) {
			std::vector<PostOp<T>> trial = group;
			trial.push_back(o);
			if (!group.empty() && (trial.size() > (size_t)MAXCHAIN || postConflict(none, trial))) {
				segOps[h].push_back(group), segAdd[h].push_back(0);
				group.assign(1, o);
			}
			else group.swap(trial);
		}
		if (!group.empty()) segOps[h].push_back(group), segAdd[h].push_back(0);
	}
	for (int h = 0; h < 2; h++) {
		D->post[h].clear();
		D->carry[h].clear();
		for (size_t g = 0; g < segOps[h].size(); g++) {
			const std::vector<PostOp<T>> &v = segOps[h][g];
			std::vector<PostPiece> pieces;
			PostSeg seg;
			buildPieces(v, segAdd[h][g], pieces, seg.total);
			if (pieces.empty()) continue;
			// A copy chain starts from a cell no earlier copy of its launch writes.  Where that
			// cell lies outside the update range and a copy of the program writes it (later:
			// the program reads the value the previous step left there), a fused step's new
			// set holds the value of two steps before: those cells are carried over first.
			for (const PostPiece &P : pieces) {
				if (P.kind != OP_COPY) continue;
				const PostOp<T> &o = v[P.list[0]];
				int axis = -1;
				for (int a = 0; a < 3; a++) if (o.F == (T *)D->F[h][0][a]) axis = a;
				const Box base{P.b.i0 + P.si, P.b.i1 + P.si, P.b.j0 + P.sj, P.b.j1 + P.sj, P.b.k0 + P.sk, P.b.k1 + P.sk};
				for (const PostOp<T> &t : copyOps[h]) {
					if (t.F != o.F) continue;
					for (const Box &part : outsideParts(t.b, range)) {
						const Box x{std::max(base.i0, part.i0), std::min(base.i1, part.i1), std::max(base.j0, part.j0),
						            std::min(base.j1, part.j1), std::max(base.k0, part.k0), std::min(base.k1, part.k1)};
						if (x.count() && axis >= 0) D->carry[h].push_back({axis, x});
					}
				}
			}
			if (seg.total >= (1L << 31)) return fail("end-of-half-step launch over 2^31 threads");
			std::vector<PostPiece> perBlock(seg.total / POST_WG);
			for (const PostPiece &P : pieces)
				for (long b = P.off / POST_WG; b < (P.off + P.b.count() + POST_WG - 1) / POST_WG; b++) perBlock[b] = P;
			if (devAlloc(D, &seg.ops, v.size() * sizeof(PostOp<T>))) return -1;
			if (devAlloc(D, &seg.pieces, perBlock.size() * sizeof(PostPiece))) return -1;
			HIPCHECK(hipMemcpyAsync(seg.ops, v.data(), v.size() * sizeof(PostOp<T>), hipMemcpyHostToDevice, D->st));
			HIPCHECK(hipMemcpyAsync(seg.pieces, perBlock.data(), perBlock.size() * sizeof(PostPiece), hipMemcpyHostToDevice, D->st));
			HIPCHECK(hipStreamSynchronize(D->st));   // the host vectors end with this iteration
			D->post[h].push_back(seg);
		}
	}
	HIPCHECK(hipStreamSynchronize(D->st));
	D->postBuilt = true;
	D->fix = fixBoxes(eOps, range);
	long fixCells = 0;
	for (const Box &b : D->fix) fixCells += b.count();
	if (D->fix.size() > (size_t)MAXFIX || fixCells * 4 > range.count()) D->fixOk = false;   // not worth it
	for (int h = 0; h < 2; h++)
		if (D->carry[h].size() * (1 + D->cplx) > (size_t)MAXCARRY) D->fixOk = false;
	return 0;
}

// Run the end of half step h: the recorded strips and sources, then the ghost copies.
template<class T>
int postHalf(alis_dev_world *D, int h)
{
	if (!D->postBuilt && buildPost<T>(D)) return -1;
	if (D->carryDue[h]) {
		D->carryDue[h] = false;
		CarryArgs<T> c{};
		for (const alis_dev_world::Carry &x : D->carry[h])
			for (int m = 0; m <= D->cplx; m++) {
				c.src[c.n] = (const T *)D->Falt[h][m][x.axis], c.dst[c.n] = (T *)D->F[h][m][x.axis];
				c.b[c.n] = x.b, c.off[c.n + 1] = c.off[c.n] + x.b.count();
				c.n++;
			}
		c.L = D->L;
		if (c.n) {
			k_carry<T><<<blocksFor(c.off[c.n]), 256, 0, D->st>>>(c);
			HIPCHECK(hipGetLastError());
		}
	}
	PostArgs<T> a{};
	a.strips = D->pendStrips[h];
	for (size_t w = 0; w < D->pendSrc[h].size(); w++) a.src[w] = D->pendSrc[h][w];
	for (size_t q = 0; q < D->pendWave[h].size(); q++) a.wave[q] = (T)D->pendWave[h][q];
	// Strips read the auxiliary fields of the other kind, which the auxiliary world's half
	// steps write on its own stream: they wait for its last half step, and its next one
	// waits for them.  Both events are bound to the worlds' last post launches, so the
	// ordering adds no marker to either stream.
	auto sideStream = [&](size_t s) { return ((a.strips >> s) & 1u) && D->tf[s].aux->st != D->st; };
	for (size_t s = 0; s < D->tf.size(); s++) {
		if (!sideStream(s)) continue;
		alis_dev_world *X = D->tf[s].aux;
		if (!X->halfDone) {
			HIPCHECK(hipEventCreateWithFlags(&X->halfDone, hipEventDisableTiming));
			HIPCHECK(hipEventRecord(X->halfDone, X->st));
		}
		HIPCHECK(hipStreamWaitEvent(D->st, X->halfDone, 0));
	}
	// the end of this program: an event the auxiliary world's stream waits for (a world with
	// strips on another stream), or the main world's (an auxiliary world)
	bool bind = D->isAux;
	for (size_t s = 0; s < D->tf.size(); s++) bind = bind || sideStream(s);
	if (bind && !D->halfDone) HIPCHECK(hipEventCreateWithFlags(&D->halfDone, hipEventDisableTiming));
	int last = -1;
	for (int g = 0; g < (int)D->post[h].size(); g++) if (D->post[h][g].total) last = g;
	for (int g = 0; g <= last; g++) {
		const PostSeg &G = D->post[h][g];
		if (!G.total) continue;
		a.ops = (const PostOp<T> *)G.ops;
		a.pieces = (const PostPiece *)G.pieces;
		a.total = G.total;
		if (g == last) launchEv(k_post<T>, dim3(blocksFor(G.total, POST_WG)), dim3(POST_WG), 0, D->st, nullptr, bind ? D->halfDone : nullptr, a);
		else k_post<T><<<blocksFor(G.total, POST_WG), POST_WG, 0, D->st>>>(a);
		HIPCHECK(hipGetLastError());
	}
	if (last < 0 && bind) HIPCHECK(hipEventRecord(D->halfDone, D->st));
	for (size_t s = 0; s < D->tf.size(); s++)
		if (sideStream(s)) HIPCHECK(hipStreamWaitEvent(D->tf[s].aux->st, D->halfDone, 0));
	D->pendStrips[h] = 0;
	std::fill(D->pendSrc[h].begin(), D->pendSrc[h].end(), 0u);
	return 0;
}

// Record source `index` of chain h for the end of this half step (src/update.c:161-179);
// per-cell waves are uploaded now, in stream order before the launch that uses them.
template<class T>
int injectRecord(alis_dev_world *D, int h, int index, float wave, const float *cellWave, long ncell)
{
	if (index < 0 || index >= (int)D->src[h].size()) return fail("source index out of range");
	DevSrc &s = D->src[h][index];
	if (cellWave) {
		if (!s.cellWave || ncell != s.ncell) return fail("per-cell wave does not match the source box");
		HIPCHECK(hipEventSynchronize(s.waveDone));
		T *pin = (T *)s.wavePinned;
		for (long c = 0; c < ncell; c++) pin[c] = (T)cellWave[c];
		HIPCHECK(hipMemcpyAsync(s.waveDev, pin, ncell * sizeof(T), hipMemcpyHostToDevice, D->st));
		HIPCHECK(hipEventRecord(s.waveDone, D->st));
	}
	else if (s.cellWave) return fail("source with per-cell phases needs per-cell waves");
	D->pendSrc[h][index >> 5] |= 1u << (index & 31);
	D->pendWave[h][index] = wave;
	return 0;
}

// ---- fused steps (k_step) ----

// Exchange the roles of D's two sets of arrays: what k_step wrote becomes current.
void swapSets(alis_dev_world *D)
{
	for (int h = 0; h < 2; h++)
		for (int m = 0; m < 2; m++)
			for (int a = 0; a < 3; a++) std::swap(D->F[h][m][a], D->Falt[h][m][a]);
	for (int h = 0; h < 2; h++)
		for (DevSlab &S : D->slab[h])
			for (int m = 0; m < 2; m++) std::swap(S.p[m][0], S.alt[m][0]), std::swap(S.p[m][1], S.alt[m][1]);
	for (DevObj &O : D->obj)
		for (int m = 0; m < 2; m++)
			for (int c = 0; c < 3; c++)
				for (int p = 0; p < MAXPOLE; p++) std::swap(O.J[m][c][p], O.Jalt[m][c][p]), std::swap(O.K[m][c][p], O.Kalt[m][c][p]);
	std::swap(D->objTab, D->objTabAlt);
	for (int h = 0; h < 2; h++) std::swap(D->post[h], D->postAlt[h]);
	std::swap(D->postBuilt, D->postBuiltAlt);
	D->par ^= 1;
}

// The other set: allocated at the first fused step, and made equal to the current one then
// and whenever cells outside a step's reach may have changed (alis_dev_write_field).
template<class T>
int ensureAlt(alis_dev_world *D)
{
	if (!D->altReady) {
		for (int h = 0; h < 2; h++)
			for (int m = 0; m <= D->cplx; m++)
				for (int a = 0; a < 3; a++)
					if (devAlloc(D, &D->Falt[h][m][a], D->fsize * sizeof(T))) return -1;
		for (int h = 0; h < 2; h++)
			for (DevSlab &S : D->slab[h])
				for (int m = 0; S.on && m <= D->cplx; m++)
					for (int x = 0; x < 2; x++)
						if (devAlloc(D, &S.alt[m][x], S.n * sizeof(T))) return -1;
		for (DevObj &O : D->obj)
			for (int m = 0; m <= D->cplx; m++)
				for (int c = 0; c < 3; c++)
					for (int p = 0; p < O.NK; p++) {
						if (devAlloc(D, &O.Jalt[m][c][p], O.box.count() * sizeof(T))) return -1;
						if (p >= O.NJ && devAlloc(D, &O.Kalt[m][c][p], O.box.count() * sizeof(T))) return -1;
					}
		if (!D->obj.empty() && uploadObjTable<T>(D, true, &D->objTabAlt)) return -1;
		D->altReady = true, D->altStale = true;
	}
	if (D->altStale) {
		auto copy = [&](void *dst, const void *src, long n) {
			return hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToDevice, D->st);
		};
		for (int h = 0; h < 2; h++)
			for (int m = 0; m <= D->cplx; m++)
				for (int a = 0; a < 3; a++) HIPCHECK(copy(D->Falt[h][m][a], D->F[h][m][a], D->fsize));
		for (int h = 0; h < 2; h++)
			for (DevSlab &S : D->slab[h])
				for (int m = 0; S.on && m <= D->cplx; m++)
					for (int x = 0; x < 2; x++) HIPCHECK(copy(S.alt[m][x], S.p[m][x], S.n));
		for (DevObj &O : D->obj)
			for (int m = 0; m <= D->cplx; m++)
				for (int c = 0; c < 3; c++)
					for (int p = 0; p < O.NK; p++) {
						HIPCHECK(copy(O.Jalt[m][c][p], O.J[m][c][p], O.box.count()));
						if (p >= O.NJ) HIPCHECK(copy(O.Kalt[m][c][p], O.K[m][c][p], O.box.count()));
					}
		D->altStale = false;
	}
	return 0;
}

// Launch shape of k_step for D: NC waves per row (64 cells each), R rows per group and the
// row below them (R + 1 = 16 / NC rows, at most 9, fewer where the two planes of LDS rows
// would pass 64 KiB), P values per LDS row and component, x planes per chunk (8;
// ALIS_STEP_CHUNK and ALIS_STEP_ROWS override).
struct StepShape { int NC, R, P, chunk, ngroup, nchunk; size_t lds; };

StepShape stepShape(const alis_dev_world *D)
{
	StepShape s{};
	s.NC = (D->kNUM + 63) / 64;
	s.P = 64 * s.NC + 1;
	const size_t row = 2 * 3 * (size_t)s.P * D->esz;
	s.R = std::min({8, 16 / std::max(s.NC, 1) - 1, (int)(65536 / row) - 1});
	if (const char *e = getenv("ALIS_STEP_ROWS")) s.R = std::min(s.R, atoi(e));
	if (s.R < 1) s.R = 1;
	s.lds = (s.R + 1) * row;
	s.ngroup = (D->jNUM + s.R - 1) / s.R;
	s.chunk = 8;
	if (const char *e = getenv("ALIS_STEP_CHUNK")) s.chunk = std::max(1, atoi(e));
	s.nchunk = (D->iNUM + s.chunk - 1) / s.chunk;
	return s;
}

// Whether D's steps run fused: decided once per world (again after a TF/SF link).  Off
// unless ALIS_FUSE is 1 (worlds that give k_step two workgroups per CU and at least 3 rows
// per group) or force (every world whose results it can reproduce): on the 256^3 nanoblock
// k_step takes 560 us against 286 us for the two separate curl launches.
template<class T>
int fusedSteps(alis_dev_world *D, bool &on)
{
	on = false;
	if (D->isAux || D->inBatch) return 0;
	if (!D->fuseState) {
		D->fuseState = -1;
		const char *env = getenv("ALIS_FUSE");
		for (const DevTfsf &t : D->tf)
			if (!t.aux) return 0;           // its programs cannot be built yet
		if (!D->postBuilt && buildPost<T>(D)) return -1;
		const StepShape s = stepShape(D);
		const bool force = env && !strcmp(env, "force");   // every world where the results allow it
		const bool on = force || (env && !strcmp(env, "1"));
		if (on && D->fixOk && s.NC <= 8 && (force || (s.R >= 3 && (long)s.ngroup * s.nchunk >= 512)))
			D->fuseState = 1;
	}
	on = D->fuseState > 0;
	return 0;
}

// Point the results of a half's arguments (built on the current set) at the other set.
template<class T>
void toOtherSet(alis_dev_world *D, int h, int m, CurlArgs<T> &a)
{
	a.O0 = (T *)D->Falt[h][m][0], a.O1 = (T *)D->Falt[h][m][1], a.O2 = (T *)D->Falt[h][m][2];
	for (int f = 0; f < 6; f++)
		if (a.face[f].on) a.face[f].qa = (T *)D->slab[h][f].alt[m][0], a.face[f].qb = (T *)D->slab[h][f].alt[m][1];
	a.objOut = (const ObjDev<T> *)D->objTabAlt;
}

// k_step's argument table: for each set as the current one, each part.
template<class T>
int stepTable(alis_dev_world *D)
{
	StepArgs<T> t[2][2] = {};
	for (int pass = 0; pass < 2; pass++) {
		for (int m = 0; m <= D->cplx; m++) {
			StepArgs<T> &s = t[D->par][m];
			s.e = curlArgs<T>(D, 0, m), s.h = curlArgs<T>(D, 1, m);
			toOtherSet<T>(D, 0, m, s.e), toOtherSet<T>(D, 1, m, s.h);
		}
		swapSets(D);
	}
	if (devAlloc(D, &D->stepArgs, sizeof t)) return -1;
	HIPCHECK(hipMemcpyAsync(D->stepArgs, t, sizeof t, hipMemcpyHostToDevice, D->st));
	HIPCHECK(hipStreamSynchronize(D->st));
	return 0;
}

// A whole step of D after its deferred E half: k_step from the current set into the other,
// the sets exchanged, the E half's end-of-step program on the new set, H again where that
// program changed E (k_fix: old H, psi and currents, final E).
template<class T>
int fusedStep(alis_dev_world *D)
{
	dropPresamples(D);
	if (ensureAlt<T>(D)) return -1;
	if (!D->stepArgs && stepTable<T>(D)) return -1;
	hipEvent_t e0 = nullptr, e1 = nullptr;
	if (D->timing && timingEvents(D, e0, e1)) return -1;
	const StepShape z = stepShape(D);
	const dim3 grid((unsigned)(8 * z.ngroup * ((z.nchunk + 7) / 8))), block(64, z.NC, z.R + 1);
	for (int m = 0; m <= D->cplx; m++) {
		const StepArgs<T> *sa = (const StepArgs<T> *)D->stepArgs + 2 * D->par + m;
		hipEvent_t s0 = m == 0 ? e0 : nullptr, s1 = m == D->cplx ? e1 : nullptr;
		if (D->seg) hipExtLaunchKernelGGL(k_step<T, 2>, grid, block, z.lds, D->st, s0, s1, 0, sa, z.R, z.P, z.chunk, z.nchunk, z.ngroup);
		else if (D->C[0]) hipExtLaunchKernelGGL(k_step<T, 1>, grid, block, z.lds, D->st, s0, s1, 0, sa, z.R, z.P, z.chunk, z.nchunk, z.ngroup);
		else hipExtLaunchKernelGGL(k_step<T, 0>, grid, block, z.lds, D->st, s0, s1, 0, sa, z.R, z.P, z.chunk, z.nchunk, z.ngroup);
		HIPCHECK(hipGetLastError());
	}
	if (e1) {
		D->evPairs[2].push_back(e0);
		D->evPairs[2].push_back(e1);
	}
	swapSets(D);
	D->pendE = 0;
	D->carryDue[0] = D->carryDue[1] = true;   // both programs of this step run on the new set
	if (postHalf<T>(D, 0)) return -1;
	if (D->fix.empty()) return 0;
	FixBoxes f{};
	f.n = (int)D->fix.size();
	for (int q = 0; q < f.n; q++) f.b[q] = D->fix[q], f.off[q + 1] = f.off[q] + D->fix[q].count();
	for (int m = 0; m <= D->cplx; m++) {
		CurlArgs<T> a = curlArgs<T>(D, 1, m);       // the new set: results, and the final E
		a.O0 = a.F0, a.O1 = a.F1, a.O2 = a.F2;
		a.F0 = (T *)D->Falt[1][m][0], a.F1 = (T *)D->Falt[1][m][1], a.F2 = (T *)D->Falt[1][m][2];
		for (int g = 0; g < 6; g++)
			if (a.face[g].on) {
				CurlFace<T> &P = a.face[g];
				P.qa = P.pa, P.qb = P.pb;
				P.pa = (T *)D->slab[1][g].alt[m][0], P.pb = (T *)D->slab[1][g].alt[m][1];
			}
		a.objOut = a.obj;
		a.obj = (const ObjDev<T> *)D->objTabAlt;
		k_fix<T><<<blocksFor(f.off[f.n]), 256, 0, D->st>>>(a, f);
		HIPCHECK(hipGetLastError());
	}
	return 0;
}

// Complete a deferred E half step the separate way (a call that reads or changes the world
// between its E and H halves).
template<class T>
int flushE(alis_dev_world *D)
{
	const int p = D->pendE;
	if (!p) return 0;
	D->pendE = 0;
	if (coreHalf<T>(D, 0)) return -1;
	return p == 2 ? postHalf<T>(D, 0) : 0;
}

int flushE(alis_dev_world *D)
{
	return D->fp64 ? flushE<double>(D) : flushE<float>(D);
}

template<class T>
int coreEntry(alis_dev_world *D, int h)
{
	if (h == 1 && D->pendE == 2) return fusedStep<T>(D);
	if (flushE<T>(D)) return -1;
	if (h == 0) {
		D->carryDue[1] = false;
		bool on;
		if (fusedSteps<T>(D, on)) return -1;
		if (on) {
			D->pendE = 1;
			return 0;
		}
	}
	return coreHalf<T>(D, h);
}

template<class T>
int syncHost(alis_dev_world *D, Box box)
{
	const alis_dev_world_desc *W = &D->W;
	const bool whole = box.i0 > box.i1;
	const Box fb = fieldBox(D);
	for (int h = 0; h < 2; h++)
		for (int m = 0; m <= D->cplx; m++)
			for (int a = 0; a < 3; a++) {
				float ***hf = W->field[h][m][a];
				Box b{std::max(box.i0, fb.i0), std::min(box.i1, fb.i1), std::max(box.j0, fb.j0), std::min(box.j1, fb.j1),
				      std::max(box.k0, fb.k0), std::min(box.k1, fb.k1)};
				// whole arrays go straight into the host block (also right for aliased rows)
				if (whole || W->expandX || W->expandY) {
					if (downloadBox<T>(D, D->F[h][m][a], D->L, fb, hostBlock(hf, fb))) return -1;
				}
				else if (b.count() && downloadSubBox<T>(D, D->F[h][m][a], D->L, b, hf)) return -1;
			}
	for (size_t n = 0; n < D->obj.size(); n++) {
		const DevObj &O = D->obj[n];
		const alis_dev_ade &A = W->obj[n];
		if (!whole && !boxesMeet(O.box, box)) continue;
		for (int m = 0; m <= D->cplx; m++)
			for (int c = 0; c < 3; c++)
				for (int p = 0; p < O.NK; p++) {
					if (downloadBox<T>(D, O.J[m][c][p], O.LO, O.box, hostBlock(A.J[m][c][p], O.box))) return -1;
					if (whole && p >= O.NJ && downloadBox<T>(D, O.K[m][c][p], O.LO, O.box, hostBlock(A.K[m][c][p], O.box))) return -1;
				}
	}
	if (!whole) return 0;
	for (int h = 0; h < 2; h++)
		for (int f = 0; f < 6; f++) {
			const DevSlab &S = D->slab[h][f];
			if (!S.on) continue;
			for (int m = 0; m <= D->cplx; m++)
				for (int s2 = 0; s2 < 2; s2++)
					if (downloadBox<T>(D, S.p[m][s2], S.L, S.b, hostBlock(W->slab[h][m][f][s2], S.b))) return -1;
		}
	return 0;
}

template<class T>
int readField(alis_dev_world *D, int h, int m, int axis, double *out)
{
	const Box fb = fieldBox(D);
	long n = fb.count();
	if (ensureScratch(D, n)) return -1;
	k_pack<T, double><<<blocksFor(n), 256, 0, D->st>>>((const T *)D->F[h][m][axis], D->L, fb, (double *)D->scratch);
	HIPCHECK(hipGetLastError());
	HIPCHECK(hipMemcpyAsync(out, D->scratch, n * sizeof(double), hipMemcpyDeviceToHost, D->st));
	HIPCHECK(hipStreamSynchronize(D->st));
	return 0;
}

template<class T>
int writeField(alis_dev_world *D, int h, int m, int axis, const double *in)
{
	const Box fb = fieldBox(D);
	long n = fb.count();
	if (ensureScratch(D, n)) return -1;
	HIPCHECK(hipMemcpyAsync(D->scratch, in, n * sizeof(double), hipMemcpyHostToDevice, D->st));
	k_unpack<T, double><<<blocksFor(n), 256, 0, D->st>>>((const double *)D->scratch, D->L, fb, (T *)D->F[h][m][axis]);
	HIPCHECK(hipGetLastError());
	HIPCHECK(hipStreamSynchronize(D->st));
	return 0;
}

// ---- phasers ----

template<class T>
PhsArgs<T> phaserArgs(alis_dev_world *D, DevPhaser &P, int qa, int qn);
void fetchArgs(void *dev, const void *pinned, size_t bytes, hipStream_t st, hipEvent_t done);

// Add the pending samples of P to its accumulators (one k_phs_flush launch).
template<class T, class TA>
int phaserFlush(alis_dev_world *D, DevPhaser &P)
{
	if (!P.pend) return 0;
	fetchArgs(P.exDev, P.exPinned, (size_t)P.pend * 4 * P.pnq * sizeof(double), D->st, P.exDone);
	PhsArgs<T> a = phaserArgs<T>(D, P, P.pq0, P.pnq);
	a.nq = P.pnq;
	const int ny = (a.nq + PHS_QPT - 1) / PHS_QPT;
	const long nb = (blocksFor(P.ncell) + 7) / 8 * 8;
	k_phs_flush<T, TA><<<nb * ny, 256, 0, D->st>>>(a, (const T *)P.ring, P.pend, P.exDev, 4 * P.pnq, ny);
	HIPCHECK(hipGetLastError());
	P.pend = 0;
	P.dirty = sizeof(TA) == 8;
	return 0;
}

// Make P's accumulators readable: add the pending samples, then fold the compensations of
// FP64 accumulators into their sums.
int phaserFlushAny(alis_dev_world *D, DevPhaser &P)
{
	int rc = 0;
	if (P.pend) {
		if (D->fp64) rc = D->dft64 ? phaserFlush<double, double>(D, P) : phaserFlush<double, float>(D, P);
		else rc = D->dft64 ? phaserFlush<float, double>(D, P) : phaserFlush<float, float>(D, P);
	}
	if (rc || !P.dirty) return rc;
	const long n = (long)P.nf * 4 * P.ncell * 2;
	k_phs_fold<<<blocksFor(n), 256, 0, D->st>>>((double *)P.acc, (double *)P.cmp, n);
	HIPCHECK(hipGetLastError());
	P.dirty = false;
	return 0;
}

// One sample of P (src/phaser.c:83-110): the face fields go to the ring and the factors to
// the staging; the accumulators take the samples PHS_RING at a time, or before anything reads
// them, in sample order (bitwise the running sums of one update per sample).
template<class T, class TA>
int phaserUpdate(alis_dev_world *D, DevPhaser &P, int qa, int qn, const double *expE, const double *expH)
{
	if (P.pend && (qa != P.pq0 || qn != P.pnq) && phaserFlush<T, TA>(D, P)) return -1;
	if (!P.pend) {
		HIPCHECK(hipEventSynchronize(P.exDone));   // the last flush has copied the staging
		P.pq0 = qa, P.pnq = qn;
	}
	PhsArgs<T> a = phaserArgs<T>(D, P, qa, qn);
	const bool pre = P.pre && P.preSlot == P.pend;   // the H curl stored the z faces' inner cells
	P.pre = false;
	for (int f = 0; !pre && f < a.nface; f++) a.face[f].si0 = 1, a.face[f].si1 = 0;
	k_phs_sample<T><<<blocksFor(P.ncell), 256, 0, D->st>>>(a, (T *)P.ring + (size_t)P.pend * 4 * P.ncell, P.ncell);
	HIPCHECK(hipGetLastError());
	double *x = P.exPinned + (size_t)P.pend * 4 * qn;
	for (int q = 0; q < qn; q++)
		x[4 * q] = expE[2 * q], x[4 * q + 1] = expE[2 * q + 1], x[4 * q + 2] = expH[2 * q], x[4 * q + 3] = expH[2 * q + 1];
	if (++P.pend == PHS_RING) return phaserFlush<T, TA>(D, P);
	return 0;
}

// Copy launch arguments staged in pinned host memory to the device in stream order (k_fetch);
// `done` (bound to the copy launch, no marker) tells when the staging can be rewritten.
void fetchArgs(void *dev, const void *pinned, size_t bytes, hipStream_t st, hipEvent_t done)
{
	const int words = (int)(bytes / 4), blocks = std::max(1, std::min(64, (words + 255) / 256));
	hipExtLaunchKernelGGL(k_fetch, dim3(blocks), dim3(256), 0, st, nullptr, done, 0, (unsigned *)dev, (const unsigned *)pinned, words);
}

// Room for `bytes` of launch arguments in S; waits until the last copy from it was made.
int stagingFor(Staging &S, size_t bytes, hipStream_t st)
{
	if (bytes > S.bytes) {
		if (S.pinned) {
			HIPCHECK(hipStreamSynchronize(st));
			HIPCHECK(hipHostFree(S.pinned));
			HIPCHECK(hipFree(S.dev));
		}
		S.pinned = nullptr, S.dev = nullptr, S.bytes = 0;
		HIPCHECK(hipHostMalloc(&S.pinned, bytes, hipHostMallocDefault));
		HIPCHECK(hipMalloc(&S.dev, bytes));
		S.bytes = bytes;
		if (!S.done) HIPCHECK(hipEventCreateWithFlags(&S.done, hipEventDisableTiming));
	}
	else HIPCHECK(hipEventSynchronize(S.done));
	return 0;
}

// One sample into each of several phasers (entries of one field and accumulator type; the
// worlds share L's stream), deferred as phaserUpdate does it, in shared launches: one stores
// every entry's sample, one flushes the entries whose ring is full.  Arguments, workgroup
// starts and the flushed samples' DFT factors go through L's pinned staging.
template<class T, class TA>
int phaserBatchUpdate(alis_dev_world *L, alis_dev_world *const *D, const int *id, const int *q0, const int *nq,
                      const double *const *expE, const double *const *expH, int n)
{
	std::vector<PhsArgs<T>> sa;
	std::vector<int> sstart, full;
	int sblocks = 0;
	for (int e = 0; e < n; e++) {
		DevPhaser &P = D[e]->phs[id[e]];
		if (!P.ncell || !nq[e]) continue;
		if (P.pend && (q0[e] != P.pq0 || nq[e] != P.pnq) && phaserFlush<T, TA>(D[e], P)) return -1;
		if (!P.pend) {
			HIPCHECK(hipEventSynchronize(P.exDone));
			P.pq0 = q0[e], P.pnq = nq[e];
		}
		P.pre = false;              // this path samples every face itself
		PhsArgs<T> a = phaserArgs<T>(D[e], P, q0[e], nq[e]);
		a.acc = (T *)P.ring + (size_t)P.pend * 4 * P.ncell;
		sa.push_back(a);
		sstart.push_back(sblocks);
		sblocks += (int)blocksFor(P.ncell);
		double *x = P.exPinned + (size_t)P.pend * 4 * P.pnq;
		for (int q = 0; q < nq[e]; q++)
			x[4 * q] = expE[e][2 * q], x[4 * q + 1] = expE[e][2 * q + 1], x[4 * q + 2] = expH[e][2 * q], x[4 * q + 3] = expH[e][2 * q + 1];
		if (++P.pend == PHS_RING) full.push_back(e);
	}
	if (sa.empty()) return 0;
	{
		const size_t ab = (sa.size() * sizeof(PhsArgs<T>) + 255) / 256 * 256, bytes = ab + sstart.size() * sizeof(int);
		if (stagingFor(L->phsStage, bytes, L->st)) return -1;
		memcpy(L->phsStage.pinned, sa.data(), sa.size() * sizeof(PhsArgs<T>));
		memcpy((char *)L->phsStage.pinned + ab, sstart.data(), sstart.size() * sizeof(int));
		fetchArgs(L->phsStage.dev, L->phsStage.pinned, bytes, L->st, L->phsStage.done);
		k_phs_sample_batch<T><<<sblocks, 256, 0, L->st>>>((const PhsArgs<T> *)L->phsStage.dev, (const int *)((char *)L->phsStage.dev + ab), (int)sa.size());
		HIPCHECK(hipGetLastError());
	}
	if (full.empty()) return 0;
	std::vector<PhsFlush<T>> fa;
	std::vector<int> fstart;
	int fblocks = 0, ny = 1;
	size_t exBytes = 0;
	for (int e : full) {
		const DevPhaser &P = D[e]->phs[id[e]];
		exBytes += (size_t)P.pend * 4 * P.pnq * sizeof(double);
	}
	const size_t ab = (full.size() * sizeof(PhsFlush<T>) + 255) / 256 * 256;
	const size_t sb = (full.size() * sizeof(int) + 255) / 256 * 256, bytes = ab + sb + exBytes;
	if (stagingFor(L->phfStage, bytes, L->st)) return -1;
	size_t exOff = ab + sb;
	for (int e : full) {
		DevPhaser &P = D[e]->phs[id[e]];
		PhsFlush<T> F;
		F.a = phaserArgs<T>(D[e], P, P.pq0, P.pnq);
		F.a.nq = P.pnq;
		F.ring = (const T *)P.ring;
		F.ex = (const double *)((char *)L->phfStage.dev + exOff);
		F.ns = P.pend, F.stride = 4 * P.pnq;
		const size_t xb = (size_t)P.pend * 4 * P.pnq * sizeof(double);
		memcpy((char *)L->phfStage.pinned + exOff, P.exPinned, xb);
		exOff += xb;
		fa.push_back(F);
		fstart.push_back(fblocks);
		fblocks += (int)blocksFor(P.ncell);
		ny = std::max(ny, (P.pnq + PHS_QPT - 1) / PHS_QPT);
		P.pend = 0;
		P.dirty = sizeof(TA) == 8;
	}
	memcpy(L->phfStage.pinned, fa.data(), fa.size() * sizeof(PhsFlush<T>));
	memcpy((char *)L->phfStage.pinned + ab, fstart.data(), fstart.size() * sizeof(int));
	fetchArgs(L->phfStage.dev, L->phfStage.pinned, bytes, L->st, L->phfStage.done);
	const long nb = (fblocks + 7) / 8 * 8;
	const PhsFlush<T> *fargs = (const PhsFlush<T> *)L->phfStage.dev;
	const int *fst = (const int *)((char *)L->phfStage.dev + ab);
	k_phs_flush_batch<T, TA><<<nb * ny, 256, 0, L->st>>>(fargs, fst, (int)fa.size(), fblocks, ny);
	HIPCHECK(hipGetLastError());
	return 0;
}

// The launch arguments of phaser P for frequencies qa .. qa+qn-1.
template<class T>
PhsArgs<T> phaserArgs(alis_dev_world *D, DevPhaser &P, int qa, int qn)
{
	PhsArgs<T> a{};
	for (int c = 0; c < 3; c++) a.E[c] = (const T *)D->F[0][0][c], a.H[c] = (const T *)D->F[1][0][c];
	a.L = D->L;
	a.iMin = P.iMin, a.jMin = P.jMin, a.kMin = P.kMin, a.nface = P.nface;
	for (int f = 0; f < P.nface; f++) a.face[f] = P.face[f];
	a.ncell = P.ncell;
	a.acc = P.acc;
	a.cmp = P.cmp;
	a.q0 = qa, a.nq = qn;
	return a;
}

// Make D use the stream of `to` (reference counted; the last user destroys a stream).
int joinStream(alis_dev_world *D, alis_dev_world *to)
{
	if (D->st == to->st) return 0;
	HIPCHECK(hipStreamSynchronize(D->st));
	if (--*D->streamRefs == 0) {
		HIPCHECK(hipStreamDestroy(D->st));
		delete D->streamRefs;
	}
	D->st = to->st;
	D->streamRefs = to->streamRefs;
	++*to->streamRefs;
	return 0;
}

// ---- batches: several worlds stepped in shared launches ----

}  // namespace

struct alis_dev_batch {
	std::vector<alis_dev_world *> W;
	bool fp64 = false;
	hipStream_t st = nullptr;
	void *curl[2] = {};         // CurlArgs<T>[] per half step
	void *curlStart[2] = {};    // int[] workgroup starts (from 0 within each class)
	// the entries in classes of one tile kind and one 1/eps source, each class contiguous in
	// the tables and run by its own specialised launch
	struct CurlClass { int first, arg, n, blocks, flat, cm, pair; };   // first: start index; arg: argument index
	std::vector<CurlClass> curlClass[2];
	bool postReady = false;
	struct Round {
		std::vector<int> world;     // index into W
		std::vector<int> seg;       // segment of that world
		void *args = nullptr;       // PostArgs<T>[] (device, in postArgs)
		void *pinned = nullptr;     // its staging (in postPinned)
		void *start = nullptr;      // int[] workgroup starts (device)
		long blocks = 0;
	};
	std::vector<Round> rounds[2];
	// every round's arguments of a half step: staged together, fetched in one copy
	void *postArgs[2] = {}, *postPinned[2] = {};
	size_t postBytes[2] = {};
	hipEvent_t postDone[2] = {};    // the last copy from postPinned has been made
	std::vector<void *> allocs;
	bool held = false;              // counted in its worlds' inBatch
};

namespace {

int batchAlloc(alis_dev_batch *B, void **p, size_t bytes)
{
	HIPCHECK(hipMalloc(p, bytes ? bytes : 16));
	B->allocs.push_back(*p);
	return 0;
}

template<class T>
int batchBuildCurl(alis_dev_batch *B)
{
	for (int h = 0; h < 2; h++) {
		// an entry: one part of a world, or both parts of a Bloch world (pair)
		struct Entry { CurlArgs<T> a[2]; long tiles; int cls; };
		const char *pe = getenv("ALIS_BATCH_PAIRS");
		const bool pairs = !(pe && !strcmp(pe, "0"));
		std::vector<Entry> all;
		for (alis_dev_world *D : B->W) {
			const bool pair = pairs && D->cplx;
			for (int m = 0; m <= D->cplx; m++) {
				CurlArgs<T> a = curlArgs<T>(D, h, m);
				// tiles as k_curl_batch counts them: flat tiles for packed rows shorter than a
				// wavefront (koff 1: rows end to end), 64 x 4 cells otherwise
				const bool flat = D->L.koff == 1 && D->kNUM + 2 <= 64;
				a.flatPitch = flat ? D->kNUM + 2 : 0;
				const long tiles = flat ? (D->L.si + 255) / 256 * D->iNUM
				                        : (long)((D->kNUM + 63) / 64) * ((D->jNUM + 3) / 4) * D->iNUM;
				// 1/eps: 2 per-segment table (and arrays where a segment is mixed), 1 arrays, 0 none
				const int cm = a.seg ? 2 : a.C0 ? 1 : 0;
				if (pair && m == 1) all.back().a[1] = a;
				else all.push_back(Entry{{a, a}, tiles, (pair ? 6 : 0) + (flat ? 3 : 0) + cm});
			}
		}
		std::stable_sort(all.begin(), all.end(), [](const Entry &x, const Entry &y) { return x.cls < y.cls; });
		std::vector<CurlArgs<T>> args;
		std::vector<int> start;
		B->curlClass[h].clear();
		for (const Entry &e : all) {
			const int pair = e.cls / 6, flat = e.cls / 3 % 2, cm = e.cls % 3;
			if (B->curlClass[h].empty() || B->curlClass[h].back().pair * 6 + B->curlClass[h].back().flat * 3 + B->curlClass[h].back().cm != e.cls)
				B->curlClass[h].push_back({(int)start.size(), (int)args.size(), 0, 0, flat, cm, pair});
			alis_dev_batch::CurlClass &c = B->curlClass[h].back();
			args.push_back(e.a[0]);
			if (pair) args.push_back(e.a[1]);
			start.push_back(c.blocks);
			c.n++;
			const int tpb = pair ? CURL_TPB_PAIR : flat ? CURL_TPB : CURL_TPB_TILE;
			c.blocks += (int)((e.tiles + tpb - 1) / tpb);
		}
		if (batchAlloc(B, &B->curl[h], args.size() * sizeof(CurlArgs<T>))) return -1;
		if (batchAlloc(B, &B->curlStart[h], start.size() * sizeof(int))) return -1;
		HIPCHECK(hipMemcpyAsync(B->curl[h], args.data(), args.size() * sizeof(CurlArgs<T>), hipMemcpyHostToDevice, B->st));
		HIPCHECK(hipMemcpyAsync(B->curlStart[h], start.data(), start.size() * sizeof(int), hipMemcpyHostToDevice, B->st));
	}
	HIPCHECK(hipStreamSynchronize(B->st));
	return 0;
}

template<class T>
int batchBuildPost(alis_dev_batch *B)
{
	for (alis_dev_world *D : B->W)
		if (!D->postBuilt && buildPost<T>(D)) return -1;
	for (int h = 0; h < 2; h++) {
		size_t rounds = 0;
		for (alis_dev_world *D : B->W) rounds = std::max(rounds, D->post[h].size());
		B->rounds[h].assign(rounds, alis_dev_batch::Round{});
		for (size_t r = 0; r < rounds; r++) {
			alis_dev_batch::Round &R = B->rounds[h][r];
			std::vector<int> start;
			for (int w = 0; w < (int)B->W.size(); w++) {
				alis_dev_world *D = B->W[w];
				if (r >= D->post[h].size() || !D->post[h][r].total) continue;
				R.world.push_back(w), R.seg.push_back((int)r);
				start.push_back((int)R.blocks);
				R.blocks += D->post[h][r].total / POST_WG;
			}
			if (R.world.empty()) continue;
			if (batchAlloc(B, &R.start, start.size() * sizeof(int))) return -1;
			HIPCHECK(hipMemcpyAsync(R.start, start.data(), start.size() * sizeof(int), hipMemcpyHostToDevice, B->st));
		}
		size_t bytes = 0;
		for (const alis_dev_batch::Round &R : B->rounds[h]) bytes += (R.world.size() * sizeof(PostArgs<T>) + 255) / 256 * 256;
		if (!bytes) continue;
		if (batchAlloc(B, &B->postArgs[h], bytes)) return -1;
		HIPCHECK(hipHostMalloc(&B->postPinned[h], bytes, hipHostMallocDefault));
		HIPCHECK(hipEventCreateWithFlags(&B->postDone[h], hipEventDisableTiming));
		HIPCHECK(hipEventRecord(B->postDone[h], B->st));
		B->postBytes[h] = bytes;
		size_t off = 0;
		for (alis_dev_batch::Round &R : B->rounds[h]) {
			if (R.world.empty()) continue;
			R.args = (char *)B->postArgs[h] + off, R.pinned = (char *)B->postPinned[h] + off;
			off += (R.world.size() * sizeof(PostArgs<T>) + 255) / 256 * 256;
		}
	}
	HIPCHECK(hipStreamSynchronize(B->st));
	B->postReady = true;
	return 0;
}

template<class T>
int batchCore(alis_dev_batch *B, int h)
{
	alis_dev_world *D0 = B->W[0];
	for (alis_dev_world *W : B->W) dropPresamples(W);
	hipEvent_t e0 = nullptr, e1 = nullptr;
	if (D0->timing) {
		if (timingEvents(D0, e0, e1)) return -1;
	}
	const std::vector<alis_dev_batch::CurlClass> &cls = B->curlClass[h];
	for (size_t q = 0; q < cls.size(); q++) {
		const alis_dev_batch::CurlClass &c = cls[q];
		const CurlArgs<T> *args = (const CurlArgs<T> *)B->curl[h] + c.arg;
		const int *start = (const int *)B->curlStart[h] + c.first;
		const dim3 grid(c.blocks), block(64, 4);
		hipEvent_t s0 = q == 0 ? e0 : nullptr, s1 = q + 1 == cls.size() ? e1 : nullptr;
#define CURL_BATCH(F, C) launchEv(k_curl_batch<T, F ? CURL_TPB : CURL_TPB_TILE, F, C>, grid, block, 0, B->st, s0, s1, args, start, c.n)
#define CURL_PAIR(F, C) launchEv(k_curl_batch<T, CURL_TPB_PAIR, F, C, true>, grid, block, 0, B->st, s0, s1, args, start, c.n)
		if (c.pair) {
			if (c.flat) {
				if (c.cm == 2) CURL_PAIR(1, 2);
				else if (c.cm == 1) CURL_PAIR(1, 1);
				else CURL_PAIR(1, 0);
			}
			else {
				if (c.cm == 2) CURL_PAIR(0, 2);
				else if (c.cm == 1) CURL_PAIR(0, 1);
				else CURL_PAIR(0, 0);
			}
		}
		else if (c.flat) {
			if (c.cm == 2) CURL_BATCH(1, 2);
			else if (c.cm == 1) CURL_BATCH(1, 1);
			else CURL_BATCH(1, 0);
		}
		else {
			if (c.cm == 2) CURL_BATCH(0, 2);
			else if (c.cm == 1) CURL_BATCH(0, 1);
			else CURL_BATCH(0, 0);
		}
#undef CURL_PAIR
#undef CURL_BATCH
		HIPCHECK(hipGetLastError());
	}
	if (e1) {
		D0->evPairs[h].push_back(e0);
		D0->evPairs[h].push_back(e1);
	}
	return 0;
}

template<class T>
int batchPost(alis_dev_batch *B, int h)
{
	if (!B->postReady && batchBuildPost<T>(B)) return -1;
	if (!B->postBytes[h]) return 0;
	HIPCHECK(hipEventSynchronize(B->postDone[h]));
	for (alis_dev_batch::Round &R : B->rounds[h]) {
		if (R.world.empty()) continue;
		PostArgs<T> *a = (PostArgs<T> *)R.pinned;
		for (size_t e = 0; e < R.world.size(); e++) {
			alis_dev_world *D = B->W[R.world[e]];
			const PostSeg &g = D->post[h][R.seg[e]];
			memset(&a[e], 0, sizeof(PostArgs<T>));
			a[e].ops = (const PostOp<T> *)g.ops;
			a[e].pieces = (const PostPiece *)g.pieces;
			a[e].total = g.total;
			a[e].strips = D->pendStrips[h];
			for (size_t w = 0; w < D->pendSrc[h].size(); w++) a[e].src[w] = D->pendSrc[h][w];
			for (size_t q = 0; q < D->pendWave[h].size(); q++) a[e].wave[q] = (T)D->pendWave[h][q];
		}
	}
	fetchArgs(B->postArgs[h], B->postPinned[h], B->postBytes[h], B->st, B->postDone[h]);
	for (alis_dev_batch::Round &R : B->rounds[h]) {
		if (R.world.empty()) continue;
		k_post_batch<T><<<R.blocks, POST_WG, 0, B->st>>>((const PostArgs<T> *)R.args, (const int *)R.start, (int)R.world.size());
		HIPCHECK(hipGetLastError());
	}
	for (alis_dev_world *D : B->W) {
		D->pendStrips[h] = 0;
		std::fill(D->pendSrc[h].begin(), D->pendSrc[h].end(), 0u);
	}
	return 0;
}

}  // namespace

// ---------------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------------

extern "C" {

const char *alis_dev_error(void) { return g_error.c_str(); }

int alis_dev_commit(const alis_dev_world_desc *W, int fp64, int dft64, alis_dev_world **out)
{
	*out = nullptr;
	int ndev = 0;
	if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail("no GPU is visible to the engine");
	alis_dev_world *D = new alis_dev_world;
	D->W = *W;
	D->objDesc.assign(W->obj, W->obj + W->nobj);
	D->W.obj = D->objDesc.data();
	for (int h = 0; h < 2; h++) {
		D->srcDesc[h].assign(W->src[h], W->src[h] + W->nsrc[h]);
		D->W.src[h] = D->srcDesc[h].data();
	}
	D->tfDesc.assign(W->tfsf, W->tfsf + W->ntfsf);
	D->W.tfsf = D->tfDesc.data();
	if (hipGetDevice(&D->device) != hipSuccess) {
		delete D;
		return fail("hipGetDevice failed");
	}
	D->fp64 = fp64 ? 1 : 0, D->dft64 = dft64 ? 1 : 0;
	D->cplx = W->complexField ? 1 : 0;
	D->esz = D->fp64 ? 8 : 4;
	D->iMIN = W->iMIN, D->iMAX = W->iMAX, D->jMIN = W->jMIN, D->jMAX = W->jMAX, D->kMIN = W->kMIN, D->kMAX = W->kMAX;
	D->iMin = W->iMin, D->iMax = W->iMax, D->jMin = W->jMin, D->jMax = W->jMax, D->kMin = W->kMin, D->kMax = W->kMax;
	D->iNUM = W->iNUM, D->jNUM = W->jNUM, D->kNUM = W->kNUM;
	if (hipStreamCreateWithFlags(&D->st, hipStreamNonBlocking) != hipSuccess) {
		delete D;
		return fail("hipStreamCreate failed");
	}
	D->streamRefs = new int(1);
	int rc = D->fp64 ? commitWorld<double>(D) : commitWorld<float>(D);
	if (rc) {
		std::string keep = g_error;
		alis_dev_release(D);
		g_error = keep;
		return rc;
	}
	for (int s = 0; s < W->ntfsf; s++) {
		DevTfsf t;
		t.S = W->tfsf[s];
		D->tf.push_back(t);
	}
	*out = D;
	return 0;
}

int alis_dev_link_tfsf(alis_dev_world *D, int s, alis_dev_world *aux)
{
	if (s < 0 || s >= (int)D->tf.size()) return fail("TF/SF index out of range");
	if (aux->fp64 != D->fp64) return fail("TF/SF auxiliary world precision differs");
	if (flushE(D) || flushE(aux)) return -1;
	D->tf[s].aux = aux;
	D->postBuilt = D->postBuiltAlt = false;
	D->fuseState = 0;
	aux->isAux = true;
	// The auxiliary world keeps its own stream, so its half steps run beside the main
	// world's curl launches; postHalf orders its strips against them with events.
	return 0;
}

void alis_dev_release(alis_dev_world *D)
{
	if (!D) return;
	if (D->st) hipStreamSynchronize(D->st);
	for (void *p : D->allocs) hipFree(p);
	for (int h = 0; h < 2; h++)
		for (DevSrc &s : D->src[h]) {
			if (s.wavePinned) hipHostFree(s.wavePinned);
			if (s.waveDone) hipEventDestroy(s.waveDone);
		}
	for (DevPhaser &P : D->phs)
		if (P.live) {
			hipFree(P.acc), hipFree(P.stage), hipFree(P.ring), hipFree(P.exDev);
			hipHostFree(P.exPinned), hipEventDestroy(P.exDone);
		}
	if (D->halfDone) hipEventDestroy(D->halfDone);
	for (int h = 0; h < 3; h++) for (hipEvent_t e : D->evPairs[h]) (void)hipEventDestroy(e);
	for (hipEvent_t e : D->evFree) (void)hipEventDestroy(e);
	if (D->scratch) hipFree(D->scratch);
	if (D->ffWork) hipFree(D->ffWork);
	for (Staging *S : {&D->phsStage, &D->phfStage}) {
		if (S->pinned) hipHostFree(S->pinned);
		if (S->dev) hipFree(S->dev);
		if (S->done) hipEventDestroy(S->done);
	}
	if (D->st && --*D->streamRefs == 0) {
		hipStreamDestroy(D->st);
		delete D->streamRefs;
	}
	delete D;
}

int alis_dev_core(alis_dev_world *D, int h)
{
	if (h < 0 || h > 1) return fail("no such half step");
	return D->fp64 ? coreEntry<double>(D, h) : coreEntry<float>(D, h);
}

int alis_dev_strips(alis_dev_world *D, int s, int h)
{
	if (h < 0 || h > 1 || s < 0 || s >= (int)D->tf.size() || !D->tf[s].aux) return fail("TF/SF box not linked");
	D->pendStrips[h] |= 1u << s;
	return 0;
}

int alis_dev_inject(alis_dev_world *D, int h, int index, float wave, const float *cellWave, long ncell)
{
	if (h < 0 || h > 1) return fail("no such source chain");
	return D->fp64 ? injectRecord<double>(D, h, index, wave, cellWave, ncell) : injectRecord<float>(D, h, index, wave, cellWave, ncell);
}

int alis_dev_ghosts(alis_dev_world *D, int h)
{
	if (h < 0 || h > 1) return fail("no such half step");
	if (h == 0 && D->pendE == 1) {      // deferred with its E core (see alis_dev_core)
		D->pendE = 2;
		return 0;
	}
	if (flushE(D)) return -1;
	return D->fp64 ? postHalf<double>(D, h) : postHalf<float>(D, h);
}

int alis_dev_sync(alis_dev_world *D, int i0, int i1, int j0, int j1, int k0, int k1)
{
	if (flushE(D)) return -1;
	Box b{i0, i1, j0, j1, k0, k1};
	return D->fp64 ? syncHost<double>(D, b) : syncHost<float>(D, b);
}

// Probe copies (alis_dev_sync_boxes): the segments of the selected field arrays and of the
// currents of dispersive objects inside each box, for every world, gathered by one launch into one
// buffer and copied by one transfer into pinned memory; the host then scatters the values
// into the descriptors' arrays.  The segment table is uploaded only when it changes (a
// program reads the same probes every step).
extern "C++" {
struct GatherHost { float ***host; Box b; long off; };

struct GatherState {
	void *dev = nullptr;        // segment table, then the gathered values
	size_t devBytes = 0, tabBytes = 0;
	float *pinned = nullptr;
	size_t pinnedBytes = 0;
	std::vector<GatherSeg> table;   // the table on the device (valid when tabBytes > 0)
	hipEvent_t ev = nullptr, done = nullptr;
};

// Gathers of at most this many bytes are written by the kernel straight into pinned host
// memory (no copy launch); larger ones are gathered on the device and copied.
constexpr size_t GATHER_DIRECT_BYTES = 1 << 20;
GatherState g_gather;

void gatherWorld(alis_dev_world *D, const Box &box, int mask, std::vector<GatherSeg> &seg, std::vector<GatherHost> &out, long &total)
{
	const alis_dev_world_desc *W = &D->W;
	const Box fb = fieldBox(D);
	auto add = [&](const void *src, const Lay &L, const Box &b, float ***host) {
		if (!b.count()) return;
		seg.push_back(GatherSeg{src, L, b, total, D->fp64, 0});
		out.push_back(GatherHost{host, b, total});
		total += b.count();
	};
	const Box b{std::max(box.i0, fb.i0), std::min(box.i1, fb.i1), std::max(box.j0, fb.j0), std::min(box.j1, fb.j1),
	            std::max(box.k0, fb.k0), std::min(box.k1, fb.k1)};
	for (int h = 0; h < 2; h++)
		for (int m = 0; m <= D->cplx; m++)
			for (int a = 0; a < 3; a++)
				if (mask >> (h * 6 + m * 3 + a) & 1) add(D->F[h][m][a], D->L, b, W->field[h][m][a]);
	for (size_t n = 0; (mask & ALIS_DEV_PROBE_CURRENTS) && n < D->obj.size(); n++) {
		const DevObj &O = D->obj[n];
		const Box c{std::max(box.i0, O.box.i0), std::min(box.i1, O.box.i1), std::max(box.j0, O.box.j0), std::min(box.j1, O.box.j1),
		            std::max(box.k0, O.box.k0), std::min(box.k1, O.box.k1)};
		if (c.i0 > c.i1 || c.j0 > c.j1 || c.k0 > c.k1) continue;
		for (int m = 0; m <= D->cplx; m++)
			for (int q = 0; q < 3; q++)
				for (int p = 0; p < O.NK; p++) add(O.J[m][q][p], O.LO, c, W->obj[n].J[m][q][p]);
	}
}
}  // extern "C++"

int alis_dev_sync_boxes(alis_dev_world *const *D, const int *box, int n)
{
	for (int e = 0; e < n; e++) if (flushE(D[e])) return -1;
	if (n < 1) return 0;
	std::vector<GatherSeg> seg;
	std::vector<GatherHost> out;
	long total = 0;
	for (int e = 0; e < n; e++) {
		const int *b = box + 7 * e;
		gatherWorld(D[e], Box{b[0], b[1], b[2], b[3], b[4], b[5]}, b[6], seg, out, total);
	}
	if (!total) return 0;
	GatherState &G = g_gather;
	// the gather runs on the first world's stream after the work of every other world's stream
	const hipStream_t st = D[0]->st;
	if (!G.ev) HIPCHECK(hipEventCreateWithFlags(&G.ev, hipEventDisableTiming));
	std::vector<hipStream_t> joined{st};
	for (int e = 1; e < n; e++) {
		if (std::find(joined.begin(), joined.end(), D[e]->st) != joined.end()) continue;
		joined.push_back(D[e]->st);
		HIPCHECK(hipEventRecord(G.ev, D[e]->st));
		HIPCHECK(hipStreamWaitEvent(st, G.ev, 0));
	}
	const size_t segB = (seg.size() * sizeof(GatherSeg) + 255) / 256 * 256, need = segB + total * sizeof(float);
	if (need > G.devBytes) {
		HIPCHECK(hipDeviceSynchronize());
		if (G.dev) HIPCHECK(hipFree(G.dev));
		G.dev = nullptr, G.devBytes = 0, G.tabBytes = 0;
		HIPCHECK(hipMalloc(&G.dev, need));
		G.devBytes = need;
	}
	if ((size_t)total * sizeof(float) > G.pinnedBytes) {
		if (G.pinned) HIPCHECK(hipHostFree(G.pinned));
		G.pinned = nullptr, G.pinnedBytes = 0;
		HIPCHECK(hipHostMalloc((void **)&G.pinned, total * sizeof(float), hipHostMallocDefault));
		G.pinnedBytes = total * sizeof(float);
	}
	const bool same = G.tabBytes == segB && G.table.size() == seg.size() && !memcmp(G.table.data(), seg.data(), seg.size() * sizeof(GatherSeg));
	if (!same) {
		G.table = seg;
		G.tabBytes = segB;
		HIPCHECK(hipMemcpyAsync(G.dev, G.table.data(), G.table.size() * sizeof(GatherSeg), hipMemcpyHostToDevice, st));
	}
	const bool direct = total * sizeof(float) <= GATHER_DIRECT_BYTES;
	float *vals = direct ? G.pinned : (float *)((char *)G.dev + segB);
	k_gather<<<blocksFor(total), 256, 0, st>>>((const GatherSeg *)G.dev, (int)seg.size(), total, vals);
	HIPCHECK(hipGetLastError());
	if (!direct) HIPCHECK(hipMemcpyAsync(G.pinned, vals, total * sizeof(float), hipMemcpyDeviceToHost, st));
	// the caller reads the values at once (get() returns one): poll the completion instead of
	// a blocking wait, which sleeps and wakes the thread some time after the work is done
	if (!G.done) HIPCHECK(hipEventCreateWithFlags(&G.done, hipEventDisableTiming));
	HIPCHECK(hipEventRecord(G.done, st));
	hipError_t q;
	while ((q = hipEventQuery(G.done)) == hipErrorNotReady) {}
	if (q != hipSuccess) return fail("hipEventQuery", q);
	for (const GatherHost &g : out) {
		const float *v = G.pinned + g.off;
		const int nk = g.b.k1 - g.b.k0 + 1;
		for (int i = g.b.i0; i <= g.b.i1; i++)
			for (int j = g.b.j0; j <= g.b.j1; j++, v += nk) memcpy(&g.host[i][j][g.b.k0], v, nk * sizeof(float));
	}
	return 0;
}

int alis_dev_read_field(alis_dev_world *D, int h, int m, int axis, double *out)
{
	if (flushE(D)) return -1;
	if (m > D->cplx || h < 0 || h > 1 || axis < 0 || axis > 2) return fail("no such field array");
	return D->fp64 ? readField<double>(D, h, m, axis, out) : readField<float>(D, h, m, axis, out);
}

int alis_dev_write_field(alis_dev_world *D, int h, int m, int axis, const double *in)
{
	if (flushE(D)) return -1;
	if (m > D->cplx || h < 0 || h > 1 || axis < 0 || axis > 2) return fail("no such field array");
	D->altStale = true;         // it may write cells no step writes (see ensureAlt)
	dropPresamples(D);
	return D->fp64 ? writeField<double>(D, h, m, axis, in) : writeField<float>(D, h, m, axis, in);
}

int alis_dev_phaser_add(alis_dev_world *D, const alis_dev_phaser_desc *P, int *id)
{
	if (flushE(D)) return -1;
	DevPhaser Q;
	Q.nf = P->nf;
	Q.iMin = P->iMin, Q.iMax = P->iMax, Q.jMin = P->jMin, Q.jMax = P->jMax, Q.kMin = P->kMin, Q.kMax = P->kMax;
	Q.cf.assign(P->cf, P->cf + (size_t)Q.nf * 24);
	const int ni = Q.iMax - Q.iMin + 1, nj = Q.jMax - Q.jMin + 1, nk = Q.kMax - Q.kMin + 1;
	const int fixed[6] = { Q.iMin, Q.iMax, Q.jMin, Q.jMax, Q.kMin, Q.kMax };
	const int lo[3] = { D->iMIN, D->jMIN, D->kMIN }, hi[3] = { D->iMAX, D->jMAX, D->kMAX };
	if (Q.iMin < lo[0] || Q.iMax > hi[0] || Q.jMin < lo[1] || Q.jMax > hi[1] || Q.kMin < lo[2] || Q.kMax > hi[2] || ni < 1 || nj < 1 || nk < 1)
		return fail("phaser box outside the grid");
	long off = 0;
	for (int kind = 0; kind < 6; kind++) {
		if (!Q.cf[kind * 4]) continue;
		PhsFaceGeom g;
		g.kind = kind;
		g.fixed = fixed[kind];
		long cells = kind < 2 ? (long)nj * nk : kind < 4 ? (long)ni * nk : (long)ni * nj;
		g.n2 = kind < 4 ? nk : nj;
		g.off = off;
		g.si0 = 1, g.si1 = 0, g.sj0 = 1, g.sj1 = 0;
		off += cells;
		Q.face[Q.nface++] = g;
	}
	Q.ncell = off;
	const long n = (long)Q.nf * 4 * Q.ncell * 2;
	const size_t accBytes = (size_t)std::max(n, 1L) * (D->dft64 ? 8 : 4);
	HIPCHECK(hipMalloc(&Q.acc, D->dft64 ? 2 * accBytes : accBytes));
	if (D->dft64) {
		Q.cmp = (char *)Q.acc + accBytes;
		k_fill<double><<<blocksFor(std::max(n, 1L)), 256, 0, D->st>>>((double *)Q.cmp, std::max(n, 1L), 0.0);
		HIPCHECK(hipGetLastError());
	}
	HIPCHECK(hipMalloc((void **)&Q.stage, (size_t)std::max(n, 1L) * sizeof(float)));
	const size_t ringBytes = (size_t)PHS_RING * 4 * std::max(Q.ncell, 1L) * D->esz;
	const size_t exBytes = (size_t)PHS_RING * 4 * std::max(Q.nf, 1) * sizeof(double);
	HIPCHECK(hipMalloc(&Q.ring, ringBytes));
	HIPCHECK(hipMalloc((void **)&Q.exDev, exBytes));
	HIPCHECK(hipHostMalloc((void **)&Q.exPinned, exBytes, hipHostMallocDefault));
	HIPCHECK(hipEventCreateWithFlags(&Q.exDone, hipEventDisableTiming));
	Q.bytes = (D->dft64 ? 2 : 1) * accBytes + (size_t)std::max(n, 1L) * sizeof(float) + ringBytes + exBytes;
	D->bytes += Q.bytes;
	// start from the host accumulators (normally zero)
	std::vector<float> host(std::max(n, 1L));
	for (int q = 0; q < Q.nf; q++)
		for (int f = 0; f < Q.nface; f++) {
			const long cells = (f + 1 < Q.nface ? Q.face[f + 1].off : Q.ncell) - Q.face[f].off;
			for (int comp = 0; comp < 4; comp++)
				memcpy(&host[2 * (((size_t)q * 4 + comp) * Q.ncell + Q.face[f].off)], Q.cf[((size_t)q * 6 + Q.face[f].kind) * 4 + comp], cells * 2 * sizeof(float));
		}
	if (D->dft64) {
		std::vector<double> hd(host.begin(), host.end());
		HIPCHECK(hipMemcpyAsync(Q.acc, hd.data(), hd.size() * sizeof(double), hipMemcpyHostToDevice, D->st));
		HIPCHECK(hipStreamSynchronize(D->st));
	}
	else {
		HIPCHECK(hipMemcpyAsync(Q.acc, host.data(), host.size() * sizeof(float), hipMemcpyHostToDevice, D->st));
		HIPCHECK(hipStreamSynchronize(D->st));
	}
	Q.live = true;
	D->phs.push_back(Q);
	*id = (int)D->phs.size() - 1;
	return 0;
}

int alis_dev_phaser_update(alis_dev_world *D, int id, int q0, int nq, const double *expE, const double *expH)
{
	if (flushE(D)) return -1;
	if (id < 0 || id >= (int)D->phs.size() || !D->phs[id].live) return fail("no such phaser");
	DevPhaser &P = D->phs[id];
	if (q0 < 0 || nq < 0 || q0 + nq > P.nf) return fail("phaser frequency range out of bounds");
	if (!P.ncell || !nq) return 0;
	if (D->fp64) return D->dft64 ? phaserUpdate<double, double>(D, P, q0, nq, expE, expH) : phaserUpdate<double, float>(D, P, q0, nq, expE, expH);
	return D->dft64 ? phaserUpdate<float, double>(D, P, q0, nq, expE, expH) : phaserUpdate<float, float>(D, P, q0, nq, expE, expH);
}

int alis_dev_phaser_update_many(alis_dev_world *const *D, const int *id, const int *q0, const int *nq,
                                const double *const *expE, const double *const *expH, int n)
{
	for (int e = 0; e < n; e++) if (flushE(D[e])) return -1;
	if (n < 1) return 0;
	for (int e = 0; e < n; e++) {
		if (id[e] < 0 || id[e] >= (int)D[e]->phs.size() || !D[e]->phs[id[e]].live) return fail("no such phaser");
		if (q0[e] < 0 || nq[e] < 0 || q0[e] + nq[e] > D[e]->phs[id[e]].nf) return fail("phaser frequency range out of bounds");
		if (D[e]->fp64 != D[0]->fp64 || D[e]->dft64 != D[0]->dft64)
			return fail("phasers updated together must share their field and accumulator precision");
	}
	for (int e = 1; e < n; e++) if (joinStream(D[e], D[0])) return -1;
	if (D[0]->fp64) return D[0]->dft64 ? phaserBatchUpdate<double, double>(D[0], D, id, q0, nq, expE, expH, n)
	                                   : phaserBatchUpdate<double, float>(D[0], D, id, q0, nq, expE, expH, n);
	return D[0]->dft64 ? phaserBatchUpdate<float, double>(D[0], D, id, q0, nq, expE, expH, n)
	                   : phaserBatchUpdate<float, float>(D[0], D, id, q0, nq, expE, expH, n);
}

long alis_dev_phaser_size(alis_dev_world *D, int id)
{
	if (id < 0 || id >= (int)D->phs.size() || !D->phs[id].live) return -1;
	return (long)D->phs[id].nf * 4 * D->phs[id].ncell * 2;
}

int alis_dev_phaser_read(alis_dev_world *D, int id, double *out, long n)
{
	if (flushE(D)) return -1;
	long need = alis_dev_phaser_size(D, id);
	if (need < 0 || n < need) return fail("phaser read buffer too small");
	DevPhaser &P = D->phs[id];
	if (!need) return 0;
	if (phaserFlushAny(D, P)) return -1;
	if (ensureScratch(D, need)) return -1;
	if (D->dft64) k_acc_to_double<double><<<blocksFor(need), 256, 0, D->st>>>((const double *)P.acc, (double *)D->scratch, need);
	else k_acc_to_double<float><<<blocksFor(need), 256, 0, D->st>>>((const float *)P.acc, (double *)D->scratch, need);
	HIPCHECK(hipGetLastError());
	HIPCHECK(hipMemcpyAsync(out, D->scratch, need * sizeof(double), hipMemcpyDeviceToHost, D->st));
	HIPCHECK(hipStreamSynchronize(D->st));
	return 0;
}

extern "C++" {
namespace {

// A checked flux-sums request: the kernel's term table and the doubles of device memory it
// needs (outputs, weights, then the table).
struct FluxPlan {
	int id = 0, q = 0, nq = 1, nterm = 0;
	long nw = 0, nout = 0, blocks = 0, doubles = 0;
	std::vector<int> t;
};

// Frequencies q .. q+nq-1, each with nout outputs.
int fluxPlan(alis_dev_world *D, int id, int q, int nq, int nterm, const int *terms, long nw, long nout, FluxPlan &F)
{
	if (alis_dev_phaser_size(D, id) < 0) return fail("no such phaser");
	const DevPhaser &P = D->phs[id];
	if (q < 0 || nq < 1 || q + nq > P.nf || nterm < 1 || nterm > 12 || nw < 1 || nout < 1)
		return fail("flux sums: bad frequency range, term count or sizes");
	F.id = id, F.q = q, F.nq = nq, F.nterm = nterm, F.nw = nw, F.nout = nout;
	F.t.assign((size_t)FLUX_TERM * nterm, 0);
	F.blocks = 0;
	for (int n = 0; n < nterm; n++) {
		const int *u = terms + 8 * n;
		const long na = u[2], nb = u[3], len = u[0] == 0 ? na * 64 : nb, wl = u[0] == 0 ? nb : na;
		if ((u[0] != 0 && u[0] != 1) || na < 1 || nb < 1 || u[1] < 0 || u[1] + na * nb > P.ncell || u[4] < 0 || u[4] > 3
		    || u[5] < 0 || u[5] > 3 || u[6] < 0 || u[6] + wl > nw || u[7] < 0 || u[7] + len > nout)
			return fail("flux sums: term outside the phaser faces or the buffers");
		std::copy(u, u + 8, F.t.begin() + (size_t)FLUX_TERM * n);
		F.t[(size_t)FLUX_TERM * n + 8] = (int)F.blocks;
		F.blocks += u[0] == 0 ? (na + 3) / 4 : (nb + 255) / 256;
	}
	F.doubles = nq * nout + nw + ((long)F.t.size() + 1) / 2;
	return 0;
}

// Run a planned request with `dev` (F.doubles doubles of device memory) holding the outputs,
// the weights and the term table.  Outputs no term writes read back as zero.
int fluxRun(alis_dev_world *D, const FluxPlan &F, const double *w, double *out, double *dev)
{
	DevPhaser &P = D->phs[F.id];
	if (phaserFlushAny(D, P)) return -1;
	const long nall = (long)F.nq * F.nout;
	double *dOut = dev, *dW = dOut + nall;
	int *dT = (int *)(dW + F.nw);
	k_fill<double><<<blocksFor(nall), 256, 0, D->st>>>(dOut, nall, 0.0);
	HIPCHECK(hipGetLastError());
	HIPCHECK(hipMemcpyAsync(dW, w, F.nw * sizeof(double), hipMemcpyHostToDevice, D->st));
	HIPCHECK(hipMemcpyAsync(dT, F.t.data(), F.t.size() * sizeof(int), hipMemcpyHostToDevice, D->st));
	const long base = (long)F.q * 4 * P.ncell;
	const dim3 grid((unsigned)F.blocks, (unsigned)F.nq);
	if (D->dft64) k_flux_sums<double><<<grid, 256, 0, D->st>>>((const Pair<double>::V *)P.acc + base, P.ncell, dT, F.nterm, dW, dOut, F.nout);
	else k_flux_sums<float><<<grid, 256, 0, D->st>>>((const Pair<float>::V *)P.acc + base, P.ncell, dT, F.nterm, dW, dOut, F.nout);
	HIPCHECK(hipGetLastError());
	HIPCHECK(hipMemcpyAsync(out, dOut, nall * sizeof(double), hipMemcpyDeviceToHost, D->st));
	HIPCHECK(hipStreamSynchronize(D->st));
	return 0;
}

// Device address of this GPU's copy of `scratch`: looked up with hipGetSymbolAddress on the
// first call per device (with that device current) and kept.  In host code the array's name
// only stands for the symbol; kernels must be given this address.
int fluxScratch(alis_dev_world *D, double **out)
{
	static std::mutex lock;
	static std::map<int, double *> addr;
	std::lock_guard<std::mutex> g(lock);
	auto it = addr.find(D->device);
	if (it == addr.end()) {
		int cur = 0;
		HIPCHECK(hipGetDevice(&cur));
		if (cur != D->device) HIPCHECK(hipSetDevice(D->device));
		void *p = nullptr;
		const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (cur != D->device) HIPCHECK(hipSetDevice(cur));
		if (e != hipSuccess) return fail("hipGetSymbolAddress(scratch)", e);
		if (!p) return fail("hipGetSymbolAddress(scratch) returned no address");
		it = addr.emplace(D->device, (double *)p).first;
	}
	*out = it->second;
	return 0;
}

}  // namespace
}  // extern "C++"

int alis_dev_phaser_flux_sums(alis_dev_world *D, int id, int q, int nterm, const int *terms,
                              const double *w, long nw, double *out, long nout)
{
	if (flushE(D)) return -1;
	FluxPlan F;
	if (fluxPlan(D, id, q, 1, nterm, terms, nw, nout, F)) return -1;
	if (ensureScratch(D, F.doubles)) return -1;
	return fluxRun(D, F, w, out, (double *)D->scratch);
}

int alis_dev_phaser_flux_sums_range(alis_dev_world *D, int id, int q0, int nq, int nterm, const int *terms,
                                    const double *w, long nw, double *out, long nout)
{
	if (flushE(D)) return -1;
	FluxPlan F;
	if (fluxPlan(D, id, q0, nq, nterm, terms, nw, nout, F)) return -1;
	if (ensureScratch(D, F.doubles)) return -1;
	return fluxRun(D, F, w, out, (double *)D->scratch);
}

int alis_dev_phaser_flux_sums_noalloc(alis_dev_world *D, int id, int q, int nterm, const int *terms,
                                      const double *w, long nw, double *out, long nout)
{
	if (flushE(D)) return -1;
	FluxPlan F;
	if (fluxPlan(D, id, q, 1, nterm, terms, nw, nout, F)) return -1;
	if (F.doubles > FLUX_SCRATCH_DOUBLES)
		return fail("flux sums without allocation: " + std::to_string(F.doubles * 8) + " bytes of outputs, weights and terms exceed the " +
		            std::to_string(FLUX_SCRATCH_DOUBLES * 8) + "-byte scratch array");
	double *s = nullptr;
	if (fluxScratch(D, &s)) return -1;
	return fluxRun(D, F, w, out, s);
}

int alis_dev_phaser_sync(alis_dev_world *D, int id)
{
	if (flushE(D)) return -1;
	long need = alis_dev_phaser_size(D, id);
	if (need < 0) return fail("no such phaser");
	DevPhaser &Q = D->phs[id];
	if (!need) return 0;
	if (phaserFlushAny(D, Q)) return -1;
	if (D->dft64) k_acc_to_float<double><<<blocksFor(need), 256, 0, D->st>>>((const double *)Q.acc, Q.stage, need);
	else HIPCHECK(hipMemcpyAsync(Q.stage, Q.acc, need * sizeof(float), hipMemcpyDeviceToDevice, D->st));
	HIPCHECK(hipGetLastError());
	std::vector<float> host(need);
	HIPCHECK(hipMemcpyAsync(host.data(), Q.stage, need * sizeof(float), hipMemcpyDeviceToHost, D->st));
	HIPCHECK(hipStreamSynchronize(D->st));
	for (int q = 0; q < Q.nf; q++)
		for (int f = 0; f < Q.nface; f++) {
			const long cells = (f + 1 < Q.nface ? Q.face[f + 1].off : Q.ncell) - Q.face[f].off;
			for (int comp = 0; comp < 4; comp++)
				memcpy(Q.cf[((size_t)q * 6 + Q.face[f].kind) * 4 + comp], &host[2 * (((size_t)q * 4 + comp) * Q.ncell + Q.face[f].off)], cells * 2 * sizeof(float));
		}
	return 0;
}

int alis_dev_phaser_remove(alis_dev_world *D, int id)
{
	if (id < 0 || id >= (int)D->phs.size() || !D->phs[id].live) return fail("no such phaser");
	DevPhaser &P = D->phs[id];
	HIPCHECK(hipStreamSynchronize(D->st));
	HIPCHECK(hipFree(P.acc));
	HIPCHECK(hipFree(P.stage));
	HIPCHECK(hipFree(P.ring));
	HIPCHECK(hipFree(P.exDev));
	HIPCHECK(hipHostFree(P.exPinned));
	HIPCHECK(hipEventDestroy(P.exDone));
	P.pend = 0;
	D->bytes -= P.bytes;
	P.live = false;
	return 0;
}

// ---- near-to-far surface sums (src/farfld.c:248-459) ----
//
// For a face F[u][w] (u its first axis, w its last, k-fastest like the accumulators) and a
// batch of columns, S(col) = sum_v alpha_v(col) sum_p F[.][.] beta_p(col): beta runs along
// the face's in-plane phase axis p and is contracted with F by a complex GEMM (k_ff_gemm),
// alpha runs along the other axis v and is applied by a reduction kernel.  z faces: p = y,
// v = x, both phase vectors.  x faces: p = y, v = z; y faces: p = x, v = z, where alpha is
// the plane-wave profile T(z) and depends on the polarisation.  A beta matrix is shared by
// the two components of a block and the two faces of an orientation.

namespace {

struct FFPhase {               // one in-plane phase vector along an axis
	const double *d, *pos;     // cell widths and positions (device)
	int n, half, sym, useKy;
	double cosk;
};

// Weight of cell p: E widths at nodes, trapezoid ends, mirror image subtracted (half = 0),
// or H widths at half nodes, first cell left out, mirror image added (half = 1).
__device__ inline double2 ffPhaseAt(const FFPhase &a, int p, double k)
{
	double w = a.half ? (p ? 1.0 : 0.0) : (p == 0 ? 0.5 : 0.0) + (p == a.n - 1 ? 0.5 : 0.0) + (p > 0 && p < a.n - 1 ? 1.0 : 0.0);
	double sn, cs;
	sincos(k * a.pos[p], &sn, &cs);
	double re = a.d[p] * cs, im = a.d[p] * sn;
	if (a.sym) {
		const double sg = a.half ? 1.0 : -1.0;
		const double mr = re * a.cosk, mi = -im * a.cosk;      // conj(v) cosk
		re += sg * mr, im += sg * mi;
	}
	return make_double2(w * re, w * im);
}

// beta[p][col] for columns c0..c0+nc-1 (col fastest).
__global__ void k_ff_beta(double2 *beta, FFPhase a, const alis_dev_ff_col *cols, int nc)
{
	const int col = blockIdx.x * blockDim.x + threadIdx.x, p = blockIdx.y;
	if (col >= nc) return;
	const double k = a.useKy ? cols[col].ky : cols[col].kx;
	beta[(size_t)p * nc + col] = ffPhaseAt(a, p, k);
}

struct FFReduce {
	const double2 *C;          // nm matrices [v][col]
	int nm, nv, nc;
	int out[4];                // output slot kind*4+comp of each matrix
	int tcomp[4];              // T component of each matrix (profile weights)
	int profile, half;         // weights: phase vector (z faces) or profile (x, y faces)
	FFPhase ph;                // phase weights
	const double *dz;          // profile weights: cell widths along z
	const double2 *T;          // [slot][6][nz]
	int nz, kLo, k0, symZ;
	double coskz;
};

__device__ inline double2 cmad(double2 acc, double2 a, double2 b)
{
	acc.x += a.x * b.x - a.y * b.y;
	acc.y += a.x * b.y + a.y * b.x;
	return acc;
}

// Partial sums over the split-th range of v: part[((split * 4 + m) * 2 + pol) * nc + col]
// = sum_v alpha_v(col, pol) C_m[v][col].  Each thread owns one column of one range; the
// ranges run in parallel (blockIdx.y) and k_ff_finish adds them in order.
constexpr int FF_SPLIT = 16;

__global__ void k_ff_reduce(double2 *part, FFReduce r, const alis_dev_ff_col *cols)
{
	const int col = blockIdx.x * blockDim.x + threadIdx.x, split = blockIdx.y;
	if (col >= r.nc) return;
	const int per = (r.nv + FF_SPLIT - 1) / FF_SPLIT, v0 = split * per, v1 = min(r.nv, v0 + per);
	const alis_dev_ff_col &c = cols[col];
	double2 acc[4][2] = {};
	if (!r.profile) {
		const double k = r.ph.useKy ? c.ky : c.kx;
		for (int v = v0; v < v1; v++) {
			const double2 a = ffPhaseAt(r.ph, v, k);
#pragma unroll
			for (int m = 0; m < 4; m++)
				if (m < r.nm) acc[m][0] = cmad(acc[m][0], a, r.C[((size_t)m * r.nv + v) * r.nc + col]);
		}
	}
	else {
#pragma unroll
		for (int pol = 0; pol < 2; pol++)
#pragma unroll
			for (int m = 0; m < 4; m++) {
				if (m >= r.nm) continue;
				const int src = c.src[pol][r.tcomp[m]];
				const double fac = c.fac[pol][r.tcomp[m]];
				if (src < 0 || fac == 0.0) continue;
				const double2 *Tb = r.T + ((size_t)c.slot[pol] * 6 + src) * r.nz - r.kLo;   // index by absolute k
				for (int v = v0; v < v1; v++) {
					if (r.half && !v) continue;
					const int k = r.k0 + v;
					double2 t = Tb[k];
					if (r.symZ) {
						const double2 m2 = Tb[r.half ? 1 - k : -k];
						const double sg = r.half ? r.coskz : -r.coskz;
						t.x += sg * m2.x, t.y += sg * m2.y;
					}
					const double w = r.half ? 1.0 : (v == 0 ? 0.5 : 0.0) + (v == r.nv - 1 ? 0.5 : 0.0) + (v > 0 && v < r.nv - 1 ? 1.0 : 0.0);
					const double s = fac * w * r.dz[v];
					acc[m][pol] = cmad(acc[m][pol], make_double2(s * t.x, s * t.y), r.C[((size_t)m * r.nv + v) * r.nc + col]);
				}
			}
	}
#pragma unroll
	for (int m = 0; m < 4; m++)
#pragma unroll
		for (int pol = 0; pol < 2; pol++)
			part[((size_t)(split * 4 + m) * 2 + pol) * r.nc + col] = acc[m][pol];
}

// S[col][kind][comp][pol] = the partial sums of k_ff_reduce, added in range order.
__global__ void k_ff_finish(double2 *S, const double2 *part, FFReduce r)
{
	const int col = blockIdx.x * blockDim.x + threadIdx.x;
	if (col >= r.nc) return;
	for (int m = 0; m < r.nm; m++)
		for (int pol = 0; pol < (r.profile ? 2 : 1); pol++) {
			double2 s = make_double2(0.0, 0.0);
			for (int split = 0; split < FF_SPLIT; split++) {
				const double2 p = part[((size_t)(split * 4 + m) * 2 + pol) * r.nc + col];
				s.x += p.x, s.y += p.y;
			}
			S[((size_t)col * 24 + r.out[m]) * 2 + pol] = s;
		}
}

// The face contraction C_m[v][col] = sum_p A_m[v][p] B[p][col] of complex doubles on FP64
// MFMA (v_mfma_f64_16x16x4f64: lane l supplies A[l%16][l/16] and B[l/16][l%16], and slot r of
// its result holds D[l/16 + 4r][l%16]).  A_m is a face component F stored with p fastest
// (trans = 0: A[v][p] = F[v np + p]) or v fastest (trans = 1: A[v][p] = F[p nv + v]); B is
// beta[p][col].  A workgroup of four wavefronts computes a 64 x 64 tile of one matrix m
// (blockIdx.z); each wavefront a 32 x 32 quarter as 2 x 2 tiles of 16 x 16, real and
// imaginary parts in their own accumulators: Cr += Ar Br - Ai Bi, Ci += Ar Bi + Ai Br.  The
// k loop stages 16 values of p of A and B in LDS per round (zero beyond the matrix edges).
typedef double v4d __attribute__((ext_vector_type(4)));
constexpr int FFG_T = 64, FFG_K = 16;

struct FFGemm {
	const double2 *A[4];       // the face components of the launch's matrices
	double2 *C[4];             // their products [v][col]
	const double2 *B;          // beta [p][col]
	int nv, np, nc, trans;
};

__global__ __launch_bounds__(256) void k_ff_gemm(FFGemm g)
{
	__shared__ double Ar[FFG_T][FFG_K + 1], Ai[FFG_T][FFG_K + 1];
	__shared__ double Br[FFG_K][FFG_T + 1], Bi[FFG_K][FFG_T + 1];
	const int m = blockIdx.z, t = (int)threadIdx.x, lane = t & 63, wave = t >> 6;
	const int v0 = (int)blockIdx.y * FFG_T, c0 = (int)blockIdx.x * FFG_T;
	const double2 *A = g.A[m];
	const int wv = (wave >> 1) * 32, wc = (wave & 1) * 32;
	v4d cr[2][2], ci[2][2];
#pragma unroll
	for (int a = 0; a < 2; a++)
#pragma unroll
		for (int b = 0; b < 2; b++) cr[a][b] = v4d{0, 0, 0, 0}, ci[a][b] = v4d{0, 0, 0, 0};
	// A (64 rows v x 16 p) and B (16 rows p x 64 cols) of the round at p0, one complex of each
	// per thread and pass, fetched into registers one round ahead of their use
	double2 xa[4], xb[4];
	auto fetch = [&](int p0) {
#pragma unroll
		for (int pass = 0; pass < 4; pass++) {
			int v, p;
			if (!g.trans) v = pass * 16 + (t >> 4), p = t & 15;          // p fastest in memory
			else v = t & 63, p = pass * 4 + (t >> 6);                     // v fastest in memory
			const int gv = v0 + v, gp = p0 + p;
			xa[pass] = make_double2(0.0, 0.0);
			if (gv < g.nv && gp < g.np) xa[pass] = g.trans ? A[(long)gp * g.nv + gv] : A[(long)gv * g.np + gp];
			const int bp = pass * 4 + (t >> 6), bc = t & 63;
			xb[pass] = make_double2(0.0, 0.0);
			if (p0 + bp < g.np && c0 + bc < g.nc) xb[pass] = g.B[(long)(p0 + bp) * g.nc + c0 + bc];
		}
	};
	fetch(0);
	for (int p0 = 0; p0 < g.np; p0 += FFG_K) {
#pragma unroll
		for (int pass = 0; pass < 4; pass++) {
			int v, p;
			if (!g.trans) v = pass * 16 + (t >> 4), p = t & 15;
			else v = t & 63, p = pass * 4 + (t >> 6);
			Ar[v][p] = xa[pass].x, Ai[v][p] = xa[pass].y;
			const int bp = pass * 4 + (t >> 6), bc = t & 63;
			Br[bp][bc] = xb[pass].x, Bi[bp][bc] = xb[pass].y;
		}
		__syncthreads();
		if (p0 + FFG_K < g.np) fetch(p0 + FFG_K);
#pragma unroll
		for (int ks = 0; ks < FFG_K; ks += 4) {
			const int k = ks + (lane >> 4);
			double ar[2], ai[2], br[2], bi[2];
#pragma unroll
			for (int a = 0; a < 2; a++) {
				ar[a] = Ar[wv + 16 * a + (lane & 15)][k], ai[a] = Ai[wv + 16 * a + (lane & 15)][k];
				br[a] = Br[k][wc + 16 * a + (lane & 15)], bi[a] = Bi[k][wc + 16 * a + (lane & 15)];
			}
#pragma unroll
			for (int a = 0; a < 2; a++)
#pragma unroll
				for (int b = 0; b < 2; b++) {
					cr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], br[b], cr[a][b], 0, 0, 0);
					cr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai[a], bi[b], cr[a][b], 0, 0, 0);
					ci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], bi[b], ci[a][b], 0, 0, 0);
					ci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[a], br[b], ci[a][b], 0, 0, 0);
				}
		}
		__syncthreads();
	}
	double2 *C = g.C[m];
#pragma unroll
	for (int a = 0; a < 2; a++)
#pragma unroll
		for (int b = 0; b < 2; b++)
#pragma unroll
			for (int r = 0; r < 4; r++) {
				const int v = v0 + wv + 16 * a + (lane >> 4) + 4 * r, c = c0 + wc + 16 * b + (lane & 15);
				if (v < g.nv && c < g.nc) C[(long)v * g.nc + c] = make_double2(cr[a][b][r], ci[a][b][r]);
			}
}

__global__ void k_ff_widen(const float *src, double *dst, long n)
{
	const long a = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (a < n) dst[a] = src[a];
}

int ffWorkspace(alis_dev_world *D, size_t bytes)
{
	if (bytes <= D->ffWorkBytes) return 0;
	if (D->ffWork) {
		HIPCHECK(hipStreamSynchronize(D->st));
		HIPCHECK(hipFree(D->ffWork));
		D->bytes -= D->ffWorkBytes;
	}
	D->ffWork = nullptr, D->ffWorkBytes = 0;
	HIPCHECK(hipMalloc(&D->ffWork, bytes));
	D->ffWorkBytes = bytes;
	D->bytes += bytes;
	return 0;
}


}  // namespace

int alis_dev_farfield(alis_dev_world *D, int id, int q, const alis_dev_ff_geom *g, int n,
                      const alis_dev_ff_col *cols, double *S)
{
	if (id < 0 || id >= (int)D->phs.size() || !D->phs[id].live) return fail("no such phaser");
	DevPhaser &Q = D->phs[id];
	if (q < 0 || q >= Q.nf) return fail("phaser frequency out of range");
	if (n < 0) return fail("negative column count");
	if (phaserFlushAny(D, Q)) return -1;
	memset(S, 0, (size_t)n * 96 * sizeof(double));
	if (!n) return 0;
	const int ni = Q.iMax - Q.iMin + 1, nj = Q.jMax - Q.jMin + 1, nk = Q.kMax - Q.kMin + 1;
	bool side = false;
	for (int f = 0; f < Q.nface; f++) side |= Q.face[f].kind < 4;
	if (side && (!g->T || g->nz < 1)) return fail("far field: x and y faces need the plane-wave profiles");
	if (side) {
		// the reduction reads T at k0..k1 and, with a z symmetry plane, at -k and 1-k
		const int lo = g->symZ ? std::min(Q.kMin, -Q.kMax) : Q.kMin, hi = g->symZ ? std::max(Q.kMax, 1 - Q.kMin) : Q.kMax;
		if (g->kLo > lo || g->kLo + g->nz - 1 < hi) return fail("far field: profiles do not cover the phaser box");
		if (g->nslot < 1) return fail("far field: no profile slots");
		for (int c = 0; c < n; c++)
			for (int p = 0; p < 2; p++) {
				if (cols[c].slot[p] < 0 || cols[c].slot[p] >= g->nslot) return fail("far field: profile slot out of range");
				for (int t = 0; t < 6; t++) if (cols[c].src[p][t] >= 6) return fail("far field: no such profile component");
			}
	}

	// columns per chunk: the C matrices (4 per block) stay within 256 MiB
	const int nvMax = std::max(std::max(ni, nj), nk), npMax = std::max(ni, nj);
	int NC = (int)std::min<long>(8192, (256L << 20) / (4L * nvMax * 16));
	NC = std::max(64, NC / 64 * 64);
	NC = std::min(NC, n);
	// workspace (bytes, 256-aligned pieces)
	auto al = [](size_t b) { return (b + 255) / 256 * 256; };
	const size_t tabB = al((size_t)(4 * ni + 4 * nj + 2 * nk) * 8);
	const size_t tB = side ? al((size_t)g->nslot * 6 * g->nz * 16) : 0;
	const size_t colB = al((size_t)n * sizeof(alis_dev_ff_col));
	const size_t fB = D->dft64 ? 0 : al((size_t)4 * Q.ncell * 16);
	const size_t sB = al((size_t)n * 96 * 8);
	const size_t betaB = al((size_t)npMax * NC * 16), cB = al((size_t)4 * nvMax * NC * 16);
	const size_t partB = al((size_t)FF_SPLIT * 4 * 2 * NC * 16);
	if (ffWorkspace(D, tabB + tB + colB + fB + sB + betaB + cB + partB)) return -1;
	char *w = (char *)D->ffWork;
	double *tab = (double *)w;                          w += tabB;
	double2 *Tdev = (double2 *)w;                       w += tB;
	alis_dev_ff_col *cdev = (alis_dev_ff_col *)w;       w += colB;
	double *Fw = (double *)w;                           w += fB;
	double2 *Sdev = (double2 *)w;                       w += sB;
	double2 *beta = (double2 *)w;                       w += betaB;
	double2 *Cm = (double2 *)w;                         w += cB;
	double2 *part = (double2 *)w;

	// tables: dxE dxH xN xH (ni) dyE dyH yN yH (nj) dzE dzH (nk)
	std::vector<double> t;
	t.reserve(4 * ni + 4 * nj + 2 * nk);
	for (const double *a : {g->dxE, g->dxH, g->xN, g->xH}) t.insert(t.end(), a, a + ni);
	for (const double *a : {g->dyE, g->dyH, g->yN, g->yH}) t.insert(t.end(), a, a + nj);
	for (const double *a : {g->dzE, g->dzH}) t.insert(t.end(), a, a + nk);
	HIPCHECK(hipMemcpyAsync(tab, t.data(), t.size() * 8, hipMemcpyHostToDevice, D->st));
	if (side) HIPCHECK(hipMemcpyAsync(Tdev, g->T, (size_t)g->nslot * 6 * g->nz * 16, hipMemcpyHostToDevice, D->st));
	HIPCHECK(hipMemcpyAsync(cdev, cols, (size_t)n * sizeof(alis_dev_ff_col), hipMemcpyHostToDevice, D->st));
	HIPCHECK(hipMemsetAsync(Sdev, 0, (size_t)n * 96 * 8, D->st));
	const double *dxE = tab, *dxH = dxE + ni, *xN = dxH + ni, *xH = xN + ni;
	const double *dyE = xH + ni, *dyH = dyE + nj, *yN = dyH + nj, *yH = yN + nj;
	const double *dzE = yH + nj, *dzH = dzE + nk;

	// the accumulators of frequency q as doubles: [comp][cell][re,im]
	const double2 *F;
	if (D->dft64) F = (const double2 *)Q.acc + (size_t)q * 4 * Q.ncell;
	else {
		const long m = 4 * Q.ncell * 2;
		k_ff_widen<<<blocksFor(m), 256, 0, D->st>>>((const float *)Q.acc + (size_t)q * 4 * Q.ncell * 2, Fw, m);
		F = (const double2 *)Fw;
	}

	const FFPhase XN = {dxE, xN, ni, 0, g->symX, 0, g->coskx}, XH = {dxH, xH, ni, 1, g->symX, 0, g->coskx};
	const FFPhase YN = {dyE, yN, nj, 0, g->symY, 1, g->cosky}, YH = {dyH, yH, nj, 1, g->symY, 1, g->cosky};
	struct Block { int kinds[2]; int comps[2]; FFPhase beta; int profile, half; FFPhase alpha; int tcomp[2]; bool trans; int np, nv; };
	const Block blocks[6] = {
		// z faces F[i][j]: contract over j
		{{4, 5}, {0, 3}, YN, 0, 0, XH, {0, 0}, false, nj, ni},
		{{4, 5}, {1, 2}, YH, 0, 0, XN, {0, 0}, false, nj, ni},
		// x faces F[j][k]: contract over j; T: comp 0 Hz, 3 Ey (Z_N), 1 Hy, 2 Ez (Z_H)
		{{0, 1}, {0, 3}, YH, 1, 0, XN, {5, 1}, true, nj, nk},
		{{0, 1}, {1, 2}, YN, 1, 1, XN, {4, 2}, true, nj, nk},
		// y faces F[i][k]: contract over i; T: comp 0 Hx, 3 Ez (Z_H), 1 Hz, 2 Ex (Z_N)
		{{2, 3}, {0, 3}, XN, 1, 1, XN, {3, 2}, true, ni, nk},
		{{2, 3}, {1, 2}, XH, 1, 0, XN, {5, 0}, true, ni, nk},
	};
	for (int c0 = 0; c0 < n; c0 += NC) {
		const int nc = std::min(NC, n - c0);
		for (const Block &b : blocks) {
			FFReduce r{};
			FFGemm gm{};
			r.C = Cm, r.nv = b.nv, r.nc = nc;
			for (int f = 0; f < Q.nface; f++) {
				const PhsFaceGeom &fg = Q.face[f];
				if (fg.kind != b.kinds[0] && fg.kind != b.kinds[1]) continue;
				for (int cc = 0; cc < 2; cc++) {
					const int m = r.nm++;
					r.out[m] = fg.kind * 4 + b.comps[cc];
					r.tcomp[m] = b.tcomp[cc];
					gm.A[m] = F + (size_t)b.comps[cc] * Q.ncell + fg.off;
					gm.C[m] = Cm + (size_t)m * b.nv * nc;
				}
			}
			if (!r.nm) continue;
			// beta[p][col], then C_m[v][col] = sum_p F_m[.][.] beta[p][col] for the block's matrices
			k_ff_beta<<<dim3(blocksFor(nc), b.np), 256, 0, D->st>>>(beta, b.beta, cdev + c0, nc);
			HIPCHECK(hipGetLastError());
			gm.B = beta, gm.nv = b.nv, gm.np = b.np, gm.nc = nc, gm.trans = b.trans;
			k_ff_gemm<<<dim3((nc + FFG_T - 1) / FFG_T, (b.nv + FFG_T - 1) / FFG_T, r.nm), 256, 0, D->st>>>(gm);
			HIPCHECK(hipGetLastError());
			r.profile = b.profile, r.half = b.half, r.ph = b.alpha;
			r.dz = b.half ? dzH : dzE;
			r.T = Tdev, r.nz = g->nz, r.kLo = g->kLo, r.k0 = Q.kMin, r.symZ = g->symZ, r.coskz = g->coskz;
			k_ff_reduce<<<dim3(blocksFor(nc), FF_SPLIT), 256, 0, D->st>>>(part, r, cdev + c0);
			HIPCHECK(hipGetLastError());
			k_ff_finish<<<blocksFor(nc), 256, 0, D->st>>>(Sdev + (size_t)c0 * 48, part, r);
			HIPCHECK(hipGetLastError());
		}
	}
	HIPCHECK(hipMemcpyAsync(S, Sdev, (size_t)n * 96 * 8, hipMemcpyDeviceToHost, D->st));
	HIPCHECK(hipStreamSynchronize(D->st));
	return 0;
}

int alis_dev_batch_create(alis_dev_world *const *D, int n, alis_dev_batch **out)
{
	for (int w = 0; w < n; w++) if (flushE(D[w])) return -1;
	*out = nullptr;
	if (n < 1) return fail("empty batch");
	alis_dev_batch *B = new alis_dev_batch;
	B->W.assign(D, D + n);
	B->fp64 = D[0]->fp64;
	for (int w = 0; w < n; w++) {
		if (D[w]->fp64 != B->fp64) { delete B; return fail("worlds of one batch must share their precision"); }
		for (int v = 0; v < w; v++) if (D[v] == D[w]) { delete B; return fail("a world appears twice in a batch"); }
		if (joinStream(D[w], D[0])) { delete B; return -1; }
	}
	B->st = D[0]->st;
	int rc = B->fp64 ? batchBuildCurl<double>(B) : batchBuildCurl<float>(B);
	if (rc) { alis_dev_batch_release(B); return rc; }
	for (alis_dev_world *W : B->W) W->inBatch++;   // their arrays stay where the arguments name them
	B->held = true;
	*out = B;
	return 0;
}

// The worlds of a batch may have joined another stream since the batch was built (a later
// batch led by another world): bring them back onto the first world's stream, which the
// batch launches on.  joinStream waits for the work on the stream a world leaves.
static int batchStream(alis_dev_batch *B)
{
	for (alis_dev_world *W : B->W)
		if (W->st != B->W[0]->st && joinStream(W, B->W[0])) return -1;
	B->st = B->W[0]->st;
	return 0;
}

int alis_dev_batch_core(alis_dev_batch *B, int h)
{
	if (h < 0 || h > 1) return fail("no such half step");
	if (batchStream(B)) return -1;
	return B->fp64 ? batchCore<double>(B, h) : batchCore<float>(B, h);
}

int alis_dev_batch_post(alis_dev_batch *B, int h)
{
	if (h < 0 || h > 1) return fail("no such half step");
	if (batchStream(B)) return -1;
	return B->fp64 ? batchPost<double>(B, h) : batchPost<float>(B, h);
}

void alis_dev_batch_release(alis_dev_batch *B)
{
	if (!B) return;
	for (alis_dev_world *W : B->W) hipStreamSynchronize(W->st);   // the batch's launches ran on one of these
	if (B->held) for (alis_dev_world *W : B->W) W->inBatch--;
	for (void *p : B->allocs) hipFree(p);
	for (int h = 0; h < 2; h++) {
		if (B->postPinned[h]) hipHostFree(B->postPinned[h]);
		if (B->postDone[h]) hipEventDestroy(B->postDone[h]);
	}
	delete B;
}

// ---- observers: Poynting flux through planes, absorbed power ----

extern "C++" {
namespace {

template<class T>
int planeSums(alis_dev_world *D, int nterm, const int *terms, const double *w, long nw, double *out, long nout)
{
	const Box fb = fieldBox(D);
	const int lo[3] = {fb.i0, fb.j0, fb.k0}, hi[3] = {fb.i1, fb.j1, fb.k1};
	std::vector<int> t((size_t)PLANE_TERM * nterm);
	long rows = 0;
	for (int n = 0; n < nterm; n++) {
		const int *u = terms + 11 * n;
		const int nrm = u[0], row = u[4], inner = 3 - u[0] - u[4];
		const long nr = (long)u[6] - u[5] + 1, ns = (long)u[8] - u[7] + 1;
		if (nrm < 0 || nrm > 2 || row < 0 || row > 2 || row == nrm || u[2] < 0 || u[2] > 2 || u[3] < 0 || u[3] > 2 || nr < 1 || ns < 1
		    || u[1] < lo[nrm] || u[1] + 1 > hi[nrm] || u[5] < lo[row] || u[6] > hi[row] || u[7] < lo[inner] || u[8] > hi[inner]
		    || u[9] < 0 || u[9] + ns > nw || u[10] < 0 || u[10] + nr > nout)
			return fail("plane sums: term outside the grid or the buffers");
		std::copy(u, u + 11, t.begin() + (size_t)PLANE_TERM * n);
		t[(size_t)PLANE_TERM * n + 11] = (int)rows;
		rows += nr;
	}
	const long tab = ((long)t.size() + 1) / 2;
	if (ensureScratch(D, nout + nw + tab)) return -1;
	double *dOut = (double *)D->scratch, *dW = dOut + nout;
	int *dT = (int *)(dW + nw);
	k_fill<double><<<blocksFor(nout), 256, 0, D->st>>>(dOut, nout, 0.0);
	HIPCHECK(hipMemcpyAsync(dW, w, nw * sizeof(double), hipMemcpyHostToDevice, D->st));
	HIPCHECK(hipMemcpyAsync(dT, t.data(), t.size() * sizeof(int), hipMemcpyHostToDevice, D->st));
	Arrays6<T> F;
	for (int a = 0; a < 3; a++) F.p[a] = (const T *)D->F[0][0][a], F.p[3 + a] = (const T *)D->F[1][0][a];
	k_plane_sums<T><<<blocksFor(rows), 256, 0, D->st>>>(F, D->L, dT, nterm, rows, dW, dOut);
	HIPCHECK(hipGetLastError());
	HIPCHECK(hipMemcpyAsync(out, dOut, nout * sizeof(double), hipMemcpyDeviceToHost, D->st));
	HIPCHECK(hipStreamSynchronize(D->st));
	return 0;
}

template<class T>
int absorption(alis_dev_world *D, int n, const Box &b, const double *dzE, const double *dzH, double *out)
{
	const long cols = (long)(b.i1 - b.i0 + 1) * (b.j1 - b.j0 + 1), nk = b.k1 - b.k0 + 1;
	if (ensureScratch(D, 9 * cols + 2 * nk)) return -1;
	double *dOut = (double *)D->scratch, *dz = dOut + 9 * cols;
	HIPCHECK(hipMemcpyAsync(dz, dzE, nk * sizeof(double), hipMemcpyHostToDevice, D->st));
	HIPCHECK(hipMemcpyAsync(dz + nk, dzH, nk * sizeof(double), hipMemcpyHostToDevice, D->st));
	Arrays6<T> EC;
	for (int a = 0; a < 3; a++) EC.p[a] = (const T *)D->F[0][0][a], EC.p[3 + a] = (const T *)D->C[a];
	k_absorption<T><<<blocksFor(cols), 256, 0, D->st>>>(EC, D->L, (const ObjDev<T> *)D->objTab, n, b, dz, dz + nk, dOut);
	HIPCHECK(hipGetLastError());
	HIPCHECK(hipMemcpyAsync(out, dOut, 9 * cols * sizeof(double), hipMemcpyDeviceToHost, D->st));
	HIPCHECK(hipStreamSynchronize(D->st));
	return 0;
}

}  // namespace
}  // extern "C++"

int alis_dev_plane_sums(alis_dev_world *D, int nterm, const int *terms, const double *w, long nw, double *out, long nout)
{
	if (flushE(D)) return -1;
	if (nterm < 1 || nw < 1 || nout < 1) return fail("plane sums: no terms or no buffers");
	return D->fp64 ? planeSums<double>(D, nterm, terms, w, nw, out, nout) : planeSums<float>(D, nterm, terms, w, nw, out, nout);
}

int alis_dev_absorption(alis_dev_world *D, int n, int i0, int i1, int j0, int j1, int k0, int k1,
                        const double *dzE, const double *dzH, double *out)
{
	if (flushE(D)) return -1;
	if (n < 0 || n >= (int)D->obj.size()) return fail("absorption: no such dispersive object");
	if (!D->C[0]) return fail("absorption: the world has no permittivity arrays");
	const Box &o = D->obj[n].clip;
	if (i0 > i1 || j0 > j1 || k0 > k1 || i0 < o.i0 || i1 > o.i1 || j0 < o.j0 || j1 > o.j1 || k0 < o.k0 || k1 > o.k1)
		return fail("absorption: columns outside the object's box");
	const Box b{i0, i1, j0, j1, k0, k1};
	return D->fp64 ? absorption<double>(D, n, b, dzE, dzH, out) : absorption<float>(D, n, b, dzE, dzH, out);
}

// ---- probe spectra ----

constexpr int SPEC_RING = 256;  // samples staged per k_spectrum_add launch

struct alis_dev_spectrum {
	int np = 0, nb = 0, nTop = 0;
	double2 *acc = nullptr;     // [probe][bin] (sum, compensation)
	double *stage[2] = {};      // device copies of the staged samples
	double *pinned[2] = {};     // samples (tau, v[np]) staged in pinned host memory
	hipEvent_t done[2] = {};    // the last launch that read stage/pinned[s] has ended
	int cur = 0, pend = 0;
};

namespace {

// Spectra share one stream of their own, beside the worlds' streams.
hipStream_t g_specStream = nullptr;
int g_specRefs = 0;

int spectrumFlush(alis_dev_spectrum *S)
{
	if (!S->pend) return 0;
	const int c = S->cur;
	HIPCHECK(hipMemcpyAsync(S->stage[c], S->pinned[c], (size_t)S->pend * (1 + S->np) * sizeof(double), hipMemcpyHostToDevice, g_specStream));
	const long threads = (long)S->nb * ((S->np + SPEC_PPT - 1) / SPEC_PPT);
	k_spectrum_add<<<blocksFor(threads), 256, 0, g_specStream>>>(S->acc, S->np, S->nb, S->nTop, S->stage[c], S->pend);
	HIPCHECK(hipGetLastError());
	HIPCHECK(hipEventRecord(S->done[c], g_specStream));
	S->cur ^= 1, S->pend = 0;
	return 0;
}

}  // namespace

int alis_dev_spectrum_create(int nprobe, int nbin, int nTop, alis_dev_spectrum **out)
{
	*out = nullptr;
	int ndev = 0;
	if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail("no GPU is visible to the engine");
	if (nprobe < 1 || nbin < 1) return fail("spectrum: no probes or no frequency bins");
	if (!g_specStream) HIPCHECK(hipStreamCreateWithFlags(&g_specStream, hipStreamNonBlocking));
	alis_dev_spectrum *S = new alis_dev_spectrum;
	g_specRefs++;
	S->np = nprobe, S->nb = nbin, S->nTop = nTop;
	const long n = 2L * nprobe * nbin;
	auto bad = [&](const char *what, hipError_t e) { alis_dev_spectrum_release(S); return fail(what, e); };
	hipError_t e;
	if ((e = hipMalloc((void **)&S->acc, n * sizeof(double2))) != hipSuccess) return bad("hipMalloc", e);
	k_fill<double><<<blocksFor(2 * n), 256, 0, g_specStream>>>((double *)S->acc, 2 * n, 0.0);
	for (int c = 0; c < 2; c++) {
		const size_t bytes = (size_t)SPEC_RING * (1 + nprobe) * sizeof(double);
		if ((e = hipMalloc((void **)&S->stage[c], bytes)) != hipSuccess) return bad("hipMalloc", e);
		if ((e = hipHostMalloc((void **)&S->pinned[c], bytes, hipHostMallocDefault)) != hipSuccess) return bad("hipHostMalloc", e);
		if ((e = hipEventCreateWithFlags(&S->done[c], hipEventDisableTiming)) != hipSuccess) return bad("hipEventCreate", e);
		if ((e = hipEventRecord(S->done[c], g_specStream)) != hipSuccess) return bad("hipEventRecord", e);
	}
	*out = S;
	return 0;
}

int alis_dev_spectrum_add(alis_dev_spectrum *S, double tau, const double *v)
{
	if (!S->pend) HIPCHECK(hipEventSynchronize(S->done[S->cur]));   // the last launch from this staging ended
	double *r = S->pinned[S->cur] + (size_t)S->pend * (1 + S->np);
	r[0] = tau;
	memcpy(r + 1, v, S->np * sizeof(double));
	if (++S->pend == SPEC_RING) return spectrumFlush(S);
	return 0;
}

int alis_dev_spectrum_read(alis_dev_spectrum *S, double *out)
{
	if (spectrumFlush(S)) return -1;
	const long n = 2L * S->np * S->nb;
	std::vector<double2> a(n);
	HIPCHECK(hipMemcpyAsync(a.data(), S->acc, n * sizeof(double2), hipMemcpyDeviceToHost, g_specStream));
	HIPCHECK(hipStreamSynchronize(g_specStream));
	for (long e = 0; e < n / 2; e++) out[2 * e] = a[2 * e].x + a[2 * e + 1].x, out[2 * e + 1] = a[2 * e].y + a[2 * e + 1].y;
	return 0;
}

void alis_dev_spectrum_release(alis_dev_spectrum *S)
{
	if (!S) return;
	if (g_specStream) (void)hipStreamSynchronize(g_specStream);
	if (S->acc) (void)hipFree(S->acc);
	for (int c = 0; c < 2; c++) {
		if (S->stage[c]) (void)hipFree(S->stage[c]);
		if (S->pinned[c]) (void)hipHostFree(S->pinned[c]);
		if (S->done[c]) (void)hipEventDestroy(S->done[c]);
	}
	delete S;
	if (--g_specRefs == 0 && g_specStream) {
		(void)hipStreamDestroy(g_specStream);
		g_specStream = nullptr;
	}
}

int alis_dev_three_wave(alis_dev_world *S, alis_dev_world *P1, alis_dev_world *P2, int n, const float *d, float alt)
{
	if (flushE(S) || flushE(P1) || flushE(P2)) return -1;
	dropPresamples(S);
	if (S == P1 || S == P2) return fail("three-wave coupling: the signal world is one of the pumps");
	for (alis_dev_world *P : {P1, P2})
		if (P->fp64 != S->fp64 || P->cplx != S->cplx || P->iMIN != S->iMIN || P->iMAX != S->iMAX || P->jMIN != S->jMIN
		    || P->jMAX != S->jMAX || P->kMIN != S->kMIN || P->kMAX != S->kMAX || P->L.si != S->L.si || P->L.sj != S->L.sj
		    || P->L.koff != S->L.koff)
			return fail("three-wave coupling: the pump and signal worlds differ in grid or precision");
	if (n < 0 || n >= (int)S->obj.size() || !S->C[0]) return fail("three-wave coupling: no such dispersive object");
	const DevObj &O = S->obj[n];
	// the cells of src/update.c:357-359: the object's box without its upper plane
	const Box b{O.box.i0, O.box.i1 - 1, O.box.j0, O.box.j1 - 1, O.box.k0, O.box.k1 - 1};
	if (b.i0 < S->iMIN || b.i1 > S->iMAX || b.j0 < S->jMIN || b.j1 > S->jMAX || b.k0 < S->kMIN || b.k1 > S->kMAX)
		return fail("three-wave coupling: object box outside the grid");
	if (!b.count()) return 0;
	// the pumps' state must be current, and their next updates must wait for the reads
	if (joinStream(P1, S) || joinStream(P2, S)) return -1;
	auto run = [&](auto zero) -> int {
		typedef decltype(zero) T;
		for (int m = 0; m <= S->cplx; m++) {
			ThreeWaveArgs<T> a{};
			for (int c = 0; c < 3; c++) {
				a.F[c] = (T *)S->F[0][m][c], a.A[c] = (const T *)P1->F[0][m][c], a.B[c] = (const T *)P2->F[0][m][c];
				a.C[c] = (const T *)S->C[c], a.fill[c] = (const T *)O.fill[c];
			}
			a.L = S->L, a.LO = O.LO, a.b = b;
			memcpy(a.d, d, sizeof a.d);
			a.alt = alt;
			k_three_wave<T><<<blocksFor(b.count()), 256, 0, S->st>>>(a);
			HIPCHECK(hipGetLastError());
		}
		return 0;
	};
	return S->fp64 ? run(0.0) : run(0.0f);
}

int alis_dev_synchronize(alis_dev_world *D)
{
	if (flushE(D)) return -1;
	HIPCHECK(hipStreamSynchronize(D->st));
	return 0;
}

double alis_dev_eps_bytes(const alis_dev_world *D)
{
	return D->seg ? D->epsBytes : D->C[0] ? 3.0 * D->esz : 0.0;
}

int alis_dev_timing(alis_dev_world *D, int enable)
{
	HIPCHECK(hipStreamSynchronize(D->st));
	for (int h = 0; h < 3; h++) {
		D->evFree.insert(D->evFree.end(), D->evPairs[h].begin(), D->evPairs[h].end());
		D->evPairs[h].clear();
		D->evMs[h] = 0, D->evCount[h] = 0;
	}
	D->timing = enable != 0;
	// events for the launches to come, made before them (creating one costs a few
	// microseconds of host time, which would otherwise fall between a small world's launches)
	while (D->timing && D->evFree.size() < 4096) {
		hipEvent_t e;
		HIPCHECK(hipEventCreate(&e));
		D->evFree.push_back(e);
	}
	return 0;
}

int alis_dev_timing_read(alis_dev_world *D, int h, double *ms, long *launches)
{
	if (h < 0 || h > 2) return fail("no such timed launch");
	HIPCHECK(hipStreamSynchronize(D->st));
	auto &v = D->evPairs[h];
	for (size_t a = 0; a + 1 < v.size(); a += 2) {
		float t = 0;
		HIPCHECK(hipEventElapsedTime(&t, v[a], v[a + 1]));
		D->evMs[h] += t;
		D->evCount[h]++;
		D->evFree.push_back(v[a]);
		D->evFree.push_back(v[a + 1]);
	}
	v.clear();
	*ms = D->evMs[h];
	*launches = D->evCount[h];
	return 0;
}

long alis_dev_bytes(alis_dev_world *D) { return D->bytes; }

}  // extern "C"
