/* Headless driver of the reference example cal_sciddicaT-unsafe-glut
 * (OpenCAL/examples/cal_sciddicaT-unsafe-glut/source): the model file
 * sciddicaTunsafe.c, unchanged, and the step loop of its window program
 * (gui.c:84-114: step++, calRunCAStep2D until the stop condition, then the
 * model's save) without the window.  Built twice by this directory's
 * Makefile: on the reference OpenCAL-OMP engine (calomp_sciddicaT_unsafe;
 * with OMP_NUM_THREADS=1 its sweep visits the cells in the serial engine's
 * row-major order) and on this engine with the model's two elementary
 * processes run by the generic device program (dropin_sciddicaT_unsafe,
 * -DCALHIP_DROPIN: the bind calls below, after the model is defined, and
 * calhipSyncHost2D before the model's fields are read directly).
 *
 *   <binary> [dump_file]
 *
 * Reads ./data/dem.txt and ./data/source.txt (610 x 496), runs the
 * example's 4000 steps, writes ./data/width_final.txt as the example does,
 * prints the step count and the final number of active cells (or, on this
 * engine, the engine's error and exit status 1), and with a
 * path writes the final current layer of h as raw doubles, row-major.
 * Test infrastructure: the product does not build or run it. */
#include <stdio.h>

#include "sciddicaTunsafe.h"

#ifdef CALHIP_DROPIN
#include <calhip2D.h>

extern const char calhip_process_source[];
void sciddicaT_transition_function(struct CALModel2D* sciddicaT, int i, int j);
void sciddicaT_remove_inactive_cells(struct CALModel2D* sciddicaT, int i, int j);

static void bind_device_program(void)
{
	calhipAddProcessSource2D(sciddicaT, calhip_process_source);
	calhipBindProcess2D(sciddicaT, sciddicaT_transition_function, "sciddicaT_transition_function");
	calhipBindProcess2D(sciddicaT, sciddicaT_remove_inactive_cells, "sciddicaT_remove_inactive_cells");
	calhipBindGlobal2D(sciddicaT, "Q", &Q, sizeof Q);
	calhipBindGlobal2D(sciddicaT, "P", &P, sizeof P);
}
#endif

int main(int argc, char** argv)
{
	sciddicaTCADef();
#ifdef CALHIP_DROPIN
	bind_device_program();
#endif
	for (;;) {
		sciddicaTsimulation->step++;
		if (!calRunCAStep2D(sciddicaTsimulation)) break;
	}
#ifdef CALHIP_DROPIN
	if (calhipGetLastError()) {
		printf("engine error: %s\n", calhipGetErrorString());
		return 1;
	}
	/* the active-set fields printed below are host mirrors of the device's */
	calhipSyncHost2D(sciddicaT);
#endif
	sciddicaTSaveConfig();
	printf("steps %d active %d\n", sciddicaTsimulation->step, sciddicaT->A.size_current);
	if (argc > 1) {
		FILE* f = fopen(argv[1], "wb");
		int i, j;
		if (!f) return 1;
		for (i = 0; i < sciddicaT->rows; i++)
			for (j = 0; j < sciddicaT->columns; j++) {
				const double v = calGet2Dr(sciddicaT, Q.h, i, j);
				fwrite(&v, sizeof v, 1, f);
			}
		fclose(f);
	}
	printf("Simulation terminated\n");
	sciddicaTExit();
	return 0;
}
