/* Headless driver of the reference 3D example mbusu (OpenCAL-OMP/examples/
 * calomp_mbusu-glut): the step loop of its window program (mbusu_gui.c:69-96:
 * step++, calRunCAStep3D until the stop condition) without the window.
 * Built twice by this directory's Makefile: on the reference OpenCAL-OMP
 * engine (calomp_mbusu) and on this engine with the example's elementary
 * process run by the generic device program (dropin_mbusu, -DCALHIP_DROPIN:
 * one added call, calhip_bind_mbusu(), after the model is defined).
 *
 *   <binary> [max_steps [dump_file]]
 *
 * Runs until the example's stop condition (it writes ris10g.txt and, every
 * step, deltat.txt in the working directory) or max_steps > 0, prints the
 * step count, and optionally writes the current layer of the example's eight
 * substates (teta, moist_cont, psi, k, h, dqdh, convergence, moist_diff) as
 * raw doubles in (k, i, j) order.  Test infrastructure: the product does not
 * build or run it. */
#include <stdio.h>
#include <stdlib.h>

#include "mbusu.h"

#ifdef CALHIP_DROPIN
void calhip_bind_mbusu(void);
#endif

static void dump(const char* path)
{
	struct CALSubstate3Dr* subs[8] = {Q.teta, Q.moist_cont, Q.psi, Q.k, Q.h, Q.dqdh, Q.convergence, Q.moist_diff};
	FILE* f = fopen(path, "wb");
	int s, i, j, k;
	if (!f) return;
	for (s = 0; s < 8; s++)
		for (k = 0; k < mbusu->slices; k++)
			for (i = 0; i < mbusu->rows; i++)
				for (j = 0; j < mbusu->columns; j++) {
					double v = calGet3Dr(mbusu, subs[s], i, j, k);
					fwrite(&v, sizeof v, 1, f);
				}
	fclose(f);
}

int main(int argc, char** argv)
{
	const int max_steps = argc > 1 ? atoi(argv[1]) : 0;
	mbusuCADef();
#ifdef CALHIP_DROPIN
	calhip_bind_mbusu();
#endif
	for (;;) {
		mbusuSimulation->step++;
		if (!calRunCAStep3D(mbusuSimulation)) break;
		if (max_steps > 0 && mbusuSimulation->step >= max_steps) break;
	}
	printf("steps %d\n", mbusuSimulation->step);
	if (argc > 2) dump(argv[2]);
	mbusuExit();
	return 0;
}
