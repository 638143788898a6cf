/* SciddicaT written against the OpenCAL C API, as a user of OpenCAL writes it
 * (OpenCAL/examples/cal_sciddicaT/source/sciddicaT.c and the active-cells
 * variant cal_sciddicaT-activecells-glut/source/sciddicaT.c), plus the one
 * call that moves it to the MI355X: calhipBindSciddicaT2D.
 *
 * The elementary processes below are the model's definition; on this engine
 * they are bound to the SciddicaT device program and the GPU executes them.
 * The sciddicat_* functions at the end expose runs to tests and benchmarks. */
#include <cal2D.h>
#include <cal2DIO.h>
#include <cal2DRun.h>
#include <calhip2D.h>
#include <stdlib.h>
#include <string.h>

#define NUMBER_OF_OUTFLOWS 4

struct sciddicaTSubstates {
	struct CALSubstate2Dr* z;
	struct CALSubstate2Dr* h;
	struct CALSubstate2Dr* f[NUMBER_OF_OUTFLOWS];
};

struct sciddicaTParameters {
	CALParameterr epsilon;
	CALParameterr r;
};

typedef struct {
	struct CALModel2D* model;
	struct CALRun2D* run;
	struct sciddicaTSubstates Q;
	struct sciddicaTParameters P;
	int active;
	int steps;
} sciddicat_t;

/* the model whose callbacks are running (OpenCAL models keep Q and P global) */
static sciddicat_t* S;

static void sciddicaT_flows_computation(struct CALModel2D* ca, int i, int j)
{
	CALbyte eliminated[5] = {CAL_FALSE, CAL_FALSE, CAL_FALSE, CAL_FALSE, CAL_FALSE};
	CALbyte again;
	CALint cnt, n;
	CALreal average, m, u[5];
	if (calGet2Dr(ca, S->Q.h, i, j) <= S->P.epsilon) return;
	m = calGet2Dr(ca, S->Q.h, i, j) - S->P.epsilon;
	u[0] = calGet2Dr(ca, S->Q.z, i, j) + S->P.epsilon;
	for (n = 1; n < ca->sizeof_X; n++) u[n] = calGetX2Dr(ca, S->Q.z, i, j, n) + calGetX2Dr(ca, S->Q.h, i, j, n);
	do {
		again = CAL_FALSE;
		average = m;
		cnt = 0;
		for (n = 0; n < ca->sizeof_X; n++)
			if (!eliminated[n]) {
				average += u[n];
				cnt++;
			}
		if (cnt != 0) average /= cnt;
		for (n = 0; n < ca->sizeof_X; n++)
			if (average <= u[n] && !eliminated[n]) {
				eliminated[n] = CAL_TRUE;
				again = CAL_TRUE;
			}
	} while (again);
	for (n = 1; n < ca->sizeof_X; n++) {
		if (eliminated[n]) {
			calSet2Dr(ca, S->Q.f[n - 1], i, j, 0.0);
		} else {
			calSet2Dr(ca, S->Q.f[n - 1], i, j, (average - u[n]) * S->P.r);
			if (S->active) calAddActiveCellX2D(ca, i, j, n);
		}
	}
}

static void sciddicaT_width_update(struct CALModel2D* ca, int i, int j)
{
	CALreal h_next = calGet2Dr(ca, S->Q.h, i, j);
	CALint n;
	for (n = 1; n < ca->sizeof_X; n++)
		h_next += calGetX2Dr(ca, S->Q.f[NUMBER_OF_OUTFLOWS - n], i, j, n) - calGet2Dr(ca, S->Q.f[n - 1], i, j);
	calSet2Dr(ca, S->Q.h, i, j, h_next);
}

static void sciddicaT_remove_inactive_cells(struct CALModel2D* ca, int i, int j)
{
	if (calGet2Dr(ca, S->Q.h, i, j) <= S->P.epsilon) calRemoveActiveCell2D(ca, i, j);
}

static void sciddicaT_simulation_init(struct CALModel2D* ca)
{
	CALint i, j, k;
	for (k = 0; k < NUMBER_OF_OUTFLOWS; k++) calInitSubstate2Dr(ca, S->Q.f[k], 0);
	for (i = 0; i < ca->rows; i++)
		for (j = 0; j < ca->columns; j++) {
			CALreal h = calGet2Dr(ca, S->Q.h, i, j);
			if (h > 0.0) {
				calSet2Dr(ca, S->Q.z, i, j, calGet2Dr(ca, S->Q.z, i, j) - h);
				if (S->active) calAddActiveCell2D(ca, i, j);
			}
		}
}

static void sciddicaT_steering(struct CALModel2D* ca)
{
	CALint k;
	for (k = 0; k < NUMBER_OF_OUTFLOWS; k++) {
		if (S->active)
			calInitSubstateNext2Dr(ca, S->Q.f[k], 0);
		else
			calInitSubstate2Dr(ca, S->Q.f[k], 0);
	}
}

static CALbyte sciddicaT_stop(struct CALModel2D* ca)
{
	(void)ca;
	return S->run->step >= S->steps ? CAL_TRUE : CAL_FALSE;
}

/* A model's own global transition that runs the processes one at a time,
 * as models with a custom transition do (e.g.
 * OpenCAL/examples/cal_sciddicaS3hex-unsafe-explicit-glut/source/
 * sciddicaS3Hex.c:218-229): each calApplyElementaryProcess2D followed by
 * calUpdate2D, or (full grid) by calUpdateSubstate2Dr of the substates the
 * process wrote. */
static void sciddicaT_explicit_update2D(struct CALModel2D* ca)
{
	calApplyElementaryProcess2D(ca, sciddicaT_flows_computation);
	calUpdate2D(ca);
	calApplyElementaryProcess2D(ca, sciddicaT_width_update);
	calUpdate2D(ca);
	if (S->active) {
		calApplyElementaryProcess2D(ca, sciddicaT_remove_inactive_cells);
		calUpdate2D(ca);
	}
}

static void sciddicaT_explicit_substates(struct CALModel2D* ca)
{
	CALint k;
	calApplyElementaryProcess2D(ca, sciddicaT_flows_computation);
	for (k = 0; k < NUMBER_OF_OUTFLOWS; k++) calUpdateSubstate2Dr(ca, S->Q.f[k]);
	calApplyElementaryProcess2D(ca, sciddicaT_width_update);
	calUpdateSubstate2Dr(ca, S->Q.h);
}

/* ---------------------------------------------------------------- entry points */

/* Build the model exactly as the reference examples do, load z and h (the
 * state before the init process), bind the device program.  active selects
 * the CAL_OPT_ACTIVE_CELLS variant (3 processes, CAL_RUN_LOOP + stop). */
void* sciddicat_create(int rows, int cols, const double* z, const double* h, double p_r, double p_eps, int steps,
                       int active)
{
	size_t n = (size_t)rows * cols;
	int k;
	sciddicat_t* s = (sciddicat_t*)calloc(1, sizeof(sciddicat_t));
	if (!s) return NULL;
	S = s;
	s->active = active;
	s->steps = steps;
	s->model = calCADef2D(rows, cols, CAL_VON_NEUMANN_NEIGHBORHOOD_2D, CAL_SPACE_TOROIDAL,
	                      active ? CAL_OPT_ACTIVE_CELLS : CAL_NO_OPT);
	if (!s->model) return NULL;
	s->run = calRunDef2D(s->model, 1, active ? CAL_RUN_LOOP : steps, CAL_UPDATE_IMPLICIT);
	calAddElementaryProcess2D(s->model, sciddicaT_flows_computation);
	calAddElementaryProcess2D(s->model, sciddicaT_width_update);
	if (active) calAddElementaryProcess2D(s->model, sciddicaT_remove_inactive_cells);
	s->Q.z = calAddSubstate2Dr(s->model);
	s->Q.h = calAddSubstate2Dr(s->model);
	for (k = 0; k < NUMBER_OF_OUTFLOWS; k++) s->Q.f[k] = calAddSubstate2Dr(s->model);
	memcpy(s->Q.z->current, z, n * sizeof(double));
	memcpy(s->Q.z->next, z, n * sizeof(double));
	memcpy(s->Q.h->current, h, n * sizeof(double));
	memcpy(s->Q.h->next, h, n * sizeof(double));
	s->P.r = p_r;
	s->P.epsilon = p_eps;
	calRunAddInitFunc2D(s->run, sciddicaT_simulation_init);
	calRunAddSteeringFunc2D(s->run, sciddicaT_steering);
	if (active) calRunAddStopConditionFunc2D(s->run, sciddicaT_stop);
	if (calhipBindSciddicaT2D(s->model, sciddicaT_flows_computation, sciddicaT_width_update,
	                          active ? sciddicaT_remove_inactive_cells : NULL, s->Q.z, s->Q.h, s->Q.f, &s->P.r,
	                          &s->P.epsilon) ||
	    calhipBindSciddicaTSteering2D(s->model, sciddicaT_steering))
		return NULL;
	return s;
}

/* Install a global transition of the model's own (mode 1: each process
 * followed by calUpdate2D; mode 2, full grid: followed by calUpdateSubstate2Dr
 * of what it wrote); the run loop then calls it with the implicit update. */
int sciddicat_set_explicit(void* handle, int mode)
{
	sciddicat_t* s = (sciddicat_t*)handle;
	if (mode == 1)
		calRunAddGlobalTransitionFunc2D(s->run, sciddicaT_explicit_update2D);
	else if (mode == 2 && !s->active)
		calRunAddGlobalTransitionFunc2D(s->run, sciddicaT_explicit_substates);
	else
		return -1;
	return 0;
}

/* Drop the stop condition (the active variant's step limit), so that
 * calhipRunSteps2D may hand its steps to the device program together. */
int sciddicat_clear_stop(void* handle)
{
	sciddicat_t* s = (sciddicat_t*)handle;
	calRunAddStopConditionFunc2D(s->run, NULL);
	return 0;
}

/* calRun2D: init, all steps, finalize (state back on the host afterwards). */
int sciddicat_run(void* handle)
{
	sciddicat_t* s = (sciddicat_t*)handle;
	S = s;
	if (s->active) {
		/* the active example runs init once and then steps until the stop
		 * condition (cal_sciddicaT-activecells-glut/source/gui.c:83-92) */
		calRunInitSimulation2D(s->run);
		while (1) {
			s->run->step++;
			if (!calRunCAStep2D(s->run)) break;
			if (calhipGetLastError()) break;
		}
		calhipSyncHost2D(s->model);
	} else {
		calRun2D(s->run);
	}
	return calhipGetLastError();
}

/* Step-wise driving for timing: init, then `n` calls of calRunCAStep2D. */
int sciddicat_init(void* handle)
{
	sciddicat_t* s = (sciddicat_t*)handle;
	S = s;
	calRunInitSimulation2D(s->run);
	return calhipGetLastError();
}

int sciddicat_steps(void* handle, int n)
{
	sciddicat_t* s = (sciddicat_t*)handle;
	int k;
	S = s;
	for (k = 0; k < n; k++) {
		s->run->step++;
		calRunCAStep2D(s->run);
	}
	return calhipGetLastError();
}

/* n steps of calRun2D's loop (calhipRunSteps2D): steps may be batched. */
int sciddicat_run_steps(void* handle, int n)
{
	sciddicat_t* s = (sciddicat_t*)handle;
	S = s;
	s->run->step++;
	if (calhipRunSteps2D(s->run, n) < 0) return -1;
	s->run->step--;   /* sciddicat_steps numbers a step before running it */
	return calhipGetLastError();
}

int sciddicat_sync(void* handle) { return calhipSynchronize2D(((sciddicat_t*)handle)->model); }

/* Copy z, h (current layers) and the active count out (after a host sync). */
int sciddicat_read(void* handle, double* out_z, double* out_h, int* active_count)
{
	sciddicat_t* s = (sciddicat_t*)handle;
	size_t n = (size_t)s->model->rows * s->model->columns;
	if (calhipSyncHost2D(s->model)) return -1;
	if (out_z) memcpy(out_z, s->Q.z->current, n * sizeof(double));
	if (out_h) memcpy(out_h, s->Q.h->current, n * sizeof(double));
	if (active_count) *active_count = s->model->A.size_current;
	return 0;
}

/* Flow substate k (current layer) after a host sync, for checks of the
 * transient-flow bookkeeping. */
int sciddicat_read_flow(void* handle, int k, double* out)
{
	sciddicat_t* s = (sciddicat_t*)handle;
	size_t n = (size_t)s->model->rows * s->model->columns;
	if (calhipSyncHost2D(s->model)) return -1;
	memcpy(out, s->Q.f[k]->current, n * sizeof(double));
	return 0;
}

struct CALModel2D* sciddicat_model(void* handle) { return ((sciddicat_t*)handle)->model; }

/* A host edit between steps: h = v at (i, j) in both layers (calInit2Dr). */
int sciddicat_set_h(void* handle, int i, int j, double v)
{
	sciddicat_t* s = (sciddicat_t*)handle;
	calInit2Dr(s->model, s->Q.h, i, j, v);
	return calhipGetLastError() ? -1 : 0;
}

void sciddicat_destroy(void* handle)
{
	sciddicat_t* s = (sciddicat_t*)handle;
	if (!s) return;
	calRunFinalize2D(s->run);
	calFinalize2D(s->model);
	if (S == s) S = NULL;
	free(s);
}
