/*
 * A whole alis.h program, end to end: a y-oriented point dipole at the centre of a vacuum
 * box, an eighth of it simulated behind three symmetry planes, perfectly matched
 * layers on the far faces, a phaser array over 500-900 nm, stepped under timer() with
 * updateE/updateH/updatePhaser.  It prints, at 700 nm: the radiated power from the phaser's
 * surface (phaserPoyntingOut), the far-field intensity along z times 8 pi / 3 (farFieldI) and
 * twice the power into the upper hemisphere (farFieldFlux).  The dipole source is normalised
 * to unit radiated power and its pattern is sin^2 of the angle to y, so all three are 1 up to
 * the discretisation error.
 *
 * usage: dipole_power   (prints the three numbers on the last line)
 */
#include "alis.h"

int main(void)
{
	dom D = {{360}, {360}, {360}};
	res R = {5, {10}, {1240}};
	sur S = {{SYM, PML}, {SYM, PML}, {SYM, PML}};
	world W = createWorld(D, R, S, "");
	pointDipole(W, Ey, 0, 0, 0, Band, 500, 900);
	phaser P = createPhasers(W, 500, 900, 20);
	for (int n = 1; timer(n, P->N); n++) {
		updateE(W);
		updateH(W);
		updatePhaser(W, P);
	}
	phaser Q = choosePhaser(P, 700);
	printf("%.9g %.9g %.9g\n", phaserPoyntingOut(Q), farFieldI(Q, Unpol, 0, 0) * 8 * PI / 3, farFieldFlux(Q, Unpol, 0, 0, 90) * 2);
	return 0;
}
