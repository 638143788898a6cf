/* Headless driver of the reference example cal_mod2CA3D-glut
 * (OpenCAL/examples/cal_mod2CA3D-glut/source): the model file mod2CA3D.c,
 * unchanged (65^3 toroidal cells, Moore 3D neighbourhood, one byte
 * substate, each cell becomes the parity of its 27-cell sum, one seed cell),
 * and the step loop of its window program (gui.c:68-92: step++,
 * calRunCAStep3D until the stop condition) without the window.  Built twice
 * by this directory's Makefile: on the reference OpenCAL-OMP 3D engine
 * (calomp_mod2CA3D) and on this engine with the process run by the generic
 * device program (dropin_mod2CA3D, -DCALHIP_DROPIN: the bind calls below,
 * after the model is defined).
 *
 *   <binary> [dump_file]
 *
 * Runs the example's 1000 steps, prints the step count and the number of
 * live cells, and with a path writes the final byte of every cell in the
 * reference's (slice, row, column) buffer order.  Test infrastructure: the
 * product does not build or run it. */
#include <stdio.h>

#include "mod2CA3D.h"

#ifdef CALHIP_DROPIN
#include <calhip2D.h>
#include <calhip3D.h>

extern const char calhip_process_source[];
void life3DTransitionFunction(struct CALModel3D* ca, int i, int j, int k);

static void bind_device_program(void)
{
	calhipAddProcessSource3D(life3D, calhip_process_source);
	calhipBindProcess3D(life3D, life3DTransitionFunction, "life3DTransitionFunction");
	calhipBindGlobal3D(life3D, "Q", &Q, sizeof Q);
}
#endif

int main(int argc, char** argv)
{
	int i, j, k, live = 0;
	life3DCADef();
#ifdef CALHIP_DROPIN
	bind_device_program();
#endif
	for (;;) {
		life3Dsimulation->step++;
		if (!calRunCAStep3D(life3Dsimulation)) break;
	}
#ifdef CALHIP_DROPIN
	if (calhipGetLastError()) {
		printf("engine error: %s\n", calhipGetErrorString());
		return 1;
	}
#endif
	for (k = 0; k < life3D->slices; k++)
		for (i = 0; i < life3D->rows; i++)
			for (j = 0; j < life3D->columns; j++) live += calGet3Db(life3D, Q.life, i, j, k) != 0;
	printf("steps %d live %d\n", life3Dsimulation->step, live);
	if (argc > 1) {
		FILE* f = fopen(argv[1], "wb");
		if (!f) return 1;
		for (k = 0; k < life3D->slices; k++)
			for (i = 0; i < life3D->rows; i++)
				for (j = 0; j < life3D->columns; j++) {
					const unsigned char v = (unsigned char)calGet3Db(life3D, Q.life, i, j, k);
					fwrite(&v, 1, 1, f);
				}
		fclose(f);
	}
	life3DExit();
	return 0;
}
