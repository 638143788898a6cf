/* Harness around the reference Life transition function
 * (OpenCAL-OMP/examples/calomp_life-glut/source/life.c:16-27, linked
 * unmodified).  The reference rule reads an int substate through the global
 * `life`; the harness builds a Moore/toroidal model with one int substate,
 * loads the caller's 0/1 grid and runs calRun2D for `steps` steps.
 * Test infrastructure only. */
#include "life.h"
#include <string.h>

void lifeTransitionFunction(struct CALModel2D*, int, int);

int ref_life_run(int rows, int cols, const unsigned char* alive0, int steps,
                 unsigned char* out)
{
	size_t n = (size_t)rows * cols, c;
	struct CALSubstate2Di* Q;
	struct CALModel2D* m = calCADef2D(rows, cols, CAL_MOORE_NEIGHBORHOOD_2D,
	                                  CAL_SPACE_TOROIDAL, CAL_NO_OPT);
	struct CALRun2D* run = calRunDef2D(m, 1, steps, CAL_UPDATE_IMPLICIT);
	if (!m || !run) return -1;
	Q = calAddSubstate2Di(m);
	life.model = m;
	life.Q = (void*)Q;
	life.run = run;
	calAddElementaryProcess2D(m, lifeTransitionFunction);
	for (c = 0; c < n; c++) {
		Q->current[c] = alive0[c];
		Q->next[c] = alive0[c];
	}
	calRun2D(run);
	for (c = 0; c < n; c++) out[c] = (unsigned char)Q->current[c];
	calRunFinalize2D(run);
	calFinalize2D(m);
	return 0;
}
