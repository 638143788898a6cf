/* Harness around the reference active-cells SciddicaT transition functions
 * (OpenCAL/examples/cal_sciddicaT-activecells-glut/source/sciddicaT.c, linked
 * unmodified against the OpenCAL-OMP engine).  It mirrors sciddicaTCADef()
 * (:160-186) and the GUI step loop (gui.c:83-92: step++, calRunCAStep2D until
 * it returns false) with caller-given grids, parameters and step count.
 * Test infrastructure only. */
#include "sciddicaT.h"
#include <string.h>

void sciddicaT_flows_computation(struct CALModel2D*, int, int);
void sciddicaT_width_update(struct CALModel2D*, int, int);
void sciddicaT_remove_inactive_cells(struct CALModel2D*, int, int);
void sciddicaTSimulationInit(struct CALModel2D*);
void sciddicaTSteering(struct CALModel2D*);

static double g_r, g_eps;
static int g_steps;

static void harness_init(struct CALModel2D* ca)
{
	sciddicaTSimulationInit(ca);
	P.r = g_r;
	P.epsilon = g_eps;
}

static CALbyte harness_stop(struct CALModel2D* ca)
{
	(void)ca;
	return sciddicaTsimulation->step >= g_steps ? CAL_TRUE : CAL_FALSE;
}

static void load(struct CALSubstate2Dr* Q_, const double* v, size_t n)
{
	memcpy(Q_->current, v, n * sizeof(double));
	memcpy(Q_->next, v, n * sizeof(double));
}

/* active_counts (optional, length steps): A.size_current after each step. */
int ref_sciddicat_active_run(int rows, int cols, const double* z, const double* h,
                             double p_r, double p_eps, int steps,
                             double* out_z, double* out_h, int* active_counts)
{
	size_t n = (size_t)rows * cols;
	int k;
	CALbyte again;
	sciddicaT = calCADef2D(rows, cols, CAL_VON_NEUMANN_NEIGHBORHOOD_2D,
	                       CAL_SPACE_TOROIDAL, CAL_OPT_ACTIVE_CELLS);
	sciddicaTsimulation = calRunDef2D(sciddicaT, 1, CAL_RUN_LOOP, CAL_UPDATE_IMPLICIT);
	calAddElementaryProcess2D(sciddicaT, sciddicaT_flows_computation);
	calAddElementaryProcess2D(sciddicaT, sciddicaT_width_update);
	calAddElementaryProcess2D(sciddicaT, sciddicaT_remove_inactive_cells);
	Q.z = calAddSubstate2Dr(sciddicaT);
	Q.h = calAddSubstate2Dr(sciddicaT);
	for (k = 0; k < NUMBER_OF_OUTFLOWS; k++) Q.f[k] = calAddSubstate2Dr(sciddicaT);
	load(Q.z, z, n);
	load(Q.h, h, n);
	g_r = p_r;
	g_eps = p_eps;
	g_steps = steps;
	calRunAddInitFunc2D(sciddicaTsimulation, harness_init);
	calRunInitSimulation2D(sciddicaTsimulation);
	calRunAddSteeringFunc2D(sciddicaTsimulation, sciddicaTSteering);
	calRunAddStopConditionFunc2D(sciddicaTsimulation, harness_stop);
	do {
		sciddicaTsimulation->step++;
		again = calRunCAStep2D(sciddicaTsimulation);
		if (active_counts)
			active_counts[sciddicaTsimulation->step - 1] = sciddicaT->A.size_current;
	} while (again);
	memcpy(out_z, Q.z->current, n * sizeof(double));
	memcpy(out_h, Q.h->current, n * sizeof(double));
	calRunFinalize2D(sciddicaTsimulation);
	calFinalize2D(sciddicaT);
	return sciddicaTsimulation ? 0 : -1;
}
