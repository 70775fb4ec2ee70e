/*
 * Shape predicates: int S(object O, x, y, z) is 1 inside.  Same inside tests, extents
 * convention and float arithmetic as src/shapes.c:9-235 (fill factors depend on them
 * bitwise).
 */
#include "internal.h"

typedef struct { float x0, x1, y0, y1, z0, z1; } extent;

/* Extents of a basic shape (src/shapes.c:9-12).  f = {{x0,x1},{y0,y1},{z0,z1}} when the
 * z range is non-empty; otherwise f = {{cx,cy,cz},{hx,hy,hz}} (centre and half widths),
 * decoded with the reference's operator precedence for the y and z bounds. */
static extent extentOf(const object *O)
{
	extent e;
	if (O->f[2][1]-O->f[2][0]) {
		e.x0 = O->f[0][0], e.x1 = O->f[0][1];
		e.y0 = O->f[1][0], e.y1 = O->f[1][1];
		e.z0 = O->f[2][0], e.z1 = O->f[2][1];
	}
	else {
		e.x0 = O->f[0][0]-O->f[1][0];
		e.y0 = O->f[0][1]-O->f[1][1] ? O->f[1][1] : O->f[1][0];
		e.z0 = O->f[0][2]-O->f[1][2] ? O->f[1][2] : O->f[1][0];
		e.x1 = O->f[0][0]+O->f[1][0];
		e.y1 = O->f[0][1]+O->f[1][1] ? O->f[1][1] : O->f[1][0];
		e.z1 = O->f[0][2]+O->f[1][2] ? O->f[1][2] : O->f[1][0];
	}
	return e;
}

/* Normalised offset from the centre along one axis: -1 and 1 at the two faces. */
#define U(v,a,b) ((2*(v)-(a)-(b))/((b)-(a)))

int Box(object O, float x, float y, float z)
{
	extent e = extentOf(&O);
	return e.x0 <= x && x <= e.x1 && e.y0 <= y && y <= e.y1 && e.z0 <= z && z <= e.z1;
}

int Ball(object O, float x, float y, float z)
{
	extent e = extentOf(&O);
	return sq(U(x,e.x0,e.x1))+sq(U(y,e.y0,e.y1))+sq(U(z,e.z0,e.z1)) < 1;
}

int Diamond(object O, float x, float y, float z)
{
	extent e = extentOf(&O);
	return fabsf(U(x,e.x0,e.x1))+fabsf(U(y,e.y0,e.y1))+fabsf(U(z,e.z0,e.z1)) <= 1;
}

int BiconeX(object O, float x, float y, float z)
{
	extent e = extentOf(&O);
	return fabsf(U(x,e.x0,e.x1))+sqrt(sq(U(y,e.y0,e.y1))+sq(U(z,e.z0,e.z1))) < 1;
}

int BiconeY(object O, float x, float y, float z)
{
	extent e = extentOf(&O);
	return fabsf(U(y,e.y0,e.y1))+sqrt(sq(U(z,e.z0,e.z1))+sq(U(x,e.x0,e.x1))) < 1;
}

int BiconeZ(object O, float x, float y, float z)
{
	extent e = extentOf(&O);
	return fabsf(U(z,e.z0,e.z1))+sqrt(sq(U(x,e.x0,e.x1))+sq(U(y,e.y0,e.y1))) < 1;
}

int RodX(object O, float x, float y, float z)
{
	extent e = extentOf(&O);
	return e.x0 <= x && x <= e.x1 && sq(U(y,e.y0,e.y1))+sq(U(z,e.z0,e.z1)) < 1;
}

int RodY(object O, float x, float y, float z)
{
	extent e = extentOf(&O);
	return e.y0 <= y && y <= e.y1 && sq(U(x,e.x0,e.x1))+sq(U(z,e.z0,e.z1)) < 1;
}

int RodZ(object O, float x, float y, float z)
{
	extent e = extentOf(&O);
	return e.z0 <= z && z <= e.z1 && sq(U(x,e.x0,e.x1))+sq(U(y,e.y0,e.y1)) < 1;
}

int DRodX(object O, float x, float y, float z)
{
	extent e = extentOf(&O);
	return e.x0 <= x && x <= e.x1 && fabsf(U(y,e.y0,e.y1))+fabsf(U(z,e.z0,e.z1)) <= 1;
}

int DRodY(object O, float x, float y, float z)
{
	extent e = extentOf(&O);
	return e.y0 <= y && y <= e.y1 && fabsf(U(z,e.z0,e.z1))+fabsf(U(x,e.x0,e.x1)) <= 1;
}

int DRodZ(object O, float x, float y, float z)
{
	extent e = extentOf(&O);
	return e.z0 <= z && z <= e.z1 && fabsf(U(x,e.x0,e.x1))+fabsf(U(y,e.y0,e.y1)) <= 1;
}

/* Hexagonal rods: axis first letter, flat-to-flat direction second (src/shapes.c:89-123). */
int HRodXY(object O, float x, float y, float z)
{
	extent e = extentOf(&O);
	return e.x0 <= x && x <= e.x1 && e.z0 < z && z < e.z1 && fabsf(U(z,e.z0,e.z1))+fabsf(U(y,e.y0,e.y1))*2 < 2;
}

int HRodYZ(object O, float x, float y, float z)
{
	extent e = extentOf(&O);
	return e.y0 <= y && y <= e.y1 && e.x0 < x && x < e.x1 && fabsf(U(x,e.x0,e.x1))+fabsf(U(z,e.z0,e.z1))*2 < 2;
}

int HRodZX(object O, float x, float y, float z)
{
	extent e = extentOf(&O);
	return e.z0 <= z && z <= e.z1 && e.y0 < y && y < e.y1 && fabsf(U(y,e.y0,e.y1))+fabsf(U(x,e.x0,e.x1))*2 < 2;
}

int HRodXZ(object O, float x, float y, float z)
{
	extent e = extentOf(&O);
	return e.x0 <= x && x <= e.x1 && e.y0 < y && y < e.y1 && fabsf(U(y,e.y0,e.y1))+fabsf(U(z,e.z0,e.z1))*2 < 2;
}

int HRodYX(object O, float x, float y, float z)
{
	extent e = extentOf(&O);
	return e.y0 <= y && y <= e.y1 && e.z0 < z && z < e.z1 && fabsf(U(z,e.z0,e.z1))+fabsf(U(x,e.x0,e.x1))*2 < 2;
}

int HRodZY(object O, float x, float y, float z)
{
	extent e = extentOf(&O);
	return e.z0 <= z && z <= e.z1 && e.x0 < x && x < e.x1 && fabsf(U(x,e.x0,e.x1))+fabsf(U(y,e.y0,e.y1))*2 < 2;
}

/* Bounds of a basic shape, widened by a margin that covers the rounding of its inside test:
 * every point the shape contains lies inside b = {x0, x1, y0, y1, z0, z1}.  0 for compound
 * and imported shapes, whose points can lie anywhere. */
int shapeBounds(const object *O, float b[6])
{
	static shape *const basic[] = { Box, Ball, Diamond, BiconeX, BiconeY, BiconeZ, RodX, RodY, RodZ,
		DRodX, DRodY, DRodZ, HRodXY, HRodYZ, HRodZX, HRodXZ, HRodYX, HRodZY };
	int known = 0;
	for (unsigned a = 0; a < sizeof basic / sizeof basic[0]; a++) known |= O->S == basic[a];
	if (!known) return 0;
	extent e = extentOf(O);
	const float lo[3] = { fminf(e.x0, e.x1), fminf(e.y0, e.y1), fminf(e.z0, e.z1) };
	const float hi[3] = { fmaxf(e.x0, e.x1), fmaxf(e.y0, e.y1), fmaxf(e.z0, e.z1) };
	for (int a = 0; a < 3; a++) {
		const float m = 1e-3f * (hi[a] - lo[a]) + 1e-3f * fmaxf(fabsf(lo[a]), fabsf(hi[a])) + 1e-6f;
		b[2*a] = lo[a] - m, b[2*a+1] = hi[a] + m;
		if (!(b[2*a] <= b[2*a+1])) b[2*a] = -INF, b[2*a+1] = INF;   /* NaN-safe */
	}
	return 1;
}

/* Compound shapes; O.p points at the operand object(s). */
int Translation(object O, float x, float y, float z)
{
	object *o = O.p;
	return o->S(*o, x-O.f[0][0], y-O.f[0][1], z-O.f[0][2]);
}

/* Rotation by angles f[0] (degrees, about x, y, z) around the centre f[1]. */
int Rotation(object O, float x, float y, float z)
{
	object *o = O.p;
	x -= O.f[1][0], y -= O.f[1][1], z -= O.f[1][2];
	float cx = cosf(O.f[0][0]*PI/180.), sx = sinf(O.f[0][0]*PI/180.);
	float cy = cosf(O.f[0][1]*PI/180.), sy = sinf(O.f[0][1]*PI/180.);
	float cz = cosf(O.f[0][2]*PI/180.), sz = sinf(O.f[0][2]*PI/180.);
	float u = cz*x+sz*y, v = cz*y-sz*x;
	return o->S(*o, cy*u-sy*z, cx*v+sx*(cy*z+sy*u), cx*(cy*z+sy*u)-sx*v);
}

int Combination(object O, float x, float y, float z)
{
	object *o = O.p;
	int count = O.f[0][0];
	for (int n = 0; n < count; n++) if (o[n].S(o[n], x, y, z)) return 1;
	return 0;
}

int Intersection(object O, float x, float y, float z)
{
	object *o = O.p;
	int count = O.f[0][0];
	for (int n = 0; n < count; n++) if (!o[n].S(o[n], x, y, z)) return 0;
	return 1;
}

int Difference(object O, float x, float y, float z)
{
	object *o = O.p;
	int count = O.f[0][0];
	if (!o[0].S(o[0], x, y, z)) return 0;
	for (int n = 1; n < count; n++) if (o[n].S(o[n], x, y, z)) return 0;
	return 1;
}

/* Fold one coordinate into the unit cell of a lattice with period a and count c
 * (c > 0: finite array, centred) around the operand's extent [lo, hi]. */
static float foldCoord(float v, float a, float c, float lo, float hi)
{
	if (c) v += 0.5 * (c-1) * a;
	float cell = floor((v + 0.5 * (a - lo - hi)) / a);
	v -= a * (c > 0 ? MAX(0, MIN(cell, c-1)) : cell);
	return v;
}

int Lattice(object O, float x, float y, float z)
{
	object *o = O.p;
	if (O.f[0][0]) x = foldCoord(x, O.f[0][0], O.f[0][1], o->f[0][0], o->f[0][1]);
	if (O.f[1][0]) y = foldCoord(y, O.f[1][0], O.f[1][1], o->f[1][0], o->f[1][1]);
	if (O.f[2][0]) z = foldCoord(z, O.f[2][0], O.f[2][1], o->f[2][0], o->f[2][1]);
	return o->S(*o, x, y, z) ? 1 : 0;
}

/* Two interleaved lattices: the second has one fewer element (or is infinite). */
int DLattice(object O, float x, float y, float z)
{
	if (Lattice(O, x, y, z)) return 1;
	for (int a = 0; a < 3; a++) if (O.f[a][0]) O.f[a][1] -= O.f[a][1] ? 1 : 2;
	return Lattice(O, x, y, z) ? 1 : 0;
}

/* Circular lattices: f[0][0] copies rotated about an axis through f[1][0], f[1][1]. */
int CLatticeX(object O, float x, float y, float z)
{
	object *o = O.p;
	float y0 = O.f[1][0], z0 = O.f[1][1];
	int count = O.f[0][0];
	for (int n = 0; n < count; n++) {
		float th = 2 * PI * n / count;
		if (o->S(*o, x, (y-y0)*cosf(th) + (z-z0)*sinf(th) + y0, (z-z0)*cosf(th) - (y-y0)*sinf(th) + z0)) return 1;
	}
	return 0;
}

int CLatticeY(object O, float x, float y, float z)
{
	object *o = O.p;
	float x0 = O.f[1][0], z0 = O.f[1][1];
	int count = O.f[0][0];
	for (int n = 0; n < count; n++) {
		float th = 2 * PI * n / count;
		if (o->S(*o, (x-x0)*cosf(th) + (z-z0)*sinf(th) + x0, y, (z-z0)*cosf(th) - (x-x0)*sinf(th) + z0)) return 1;
	}
	return 0;
}

int CLatticeZ(object O, float x, float y, float z)
{
	object *o = O.p;
	float x0 = O.f[1][0], y0 = O.f[1][1];
	int count = O.f[0][0];
	for (int n = 0; n < count; n++) {
		float th = 2 * PI * n / count;
		if (o->S(*o, (x-x0)*cosf(th) + (y-y0)*sinf(th) + x0, (y-y0)*cosf(th) - (x-x0)*sinf(th) + y0, z)) return 1;
	}
	return 0;
}

/* Height-map solid: O.p is a float** map over [f[0][0],f[0][1]] x [f[1][0],f[1][1]] with
 * f[0][2] x f[1][2] samples; inside where the sampled height lies in [f[2][0], f[2][1]]. */
int ImportContour(object O, float x, float y, float z)
{
	float **map = O.p;
	if (O.f[0][0] < x && x < O.f[0][1] && O.f[1][0] < y && y < O.f[1][1]) {
		int i = O.f[0][2] * (x - O.f[0][0]) / (O.f[0][1] - O.f[0][0]);
		int j = O.f[1][2] * (y - O.f[1][0]) / (O.f[1][1] - O.f[1][0]);
		if (O.f[2][0] <= map[i][j] && map[i][j] <= O.f[2][1]) return 1;
	}
	return 0;
}
