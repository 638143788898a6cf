/* Harness around the reference SciddicaT transition functions
 * (OpenCAL-OMP/examples/calomp_sciddicaT/source/sciddicaT.c, linked unmodified)
 * and the reference OpenCAL-OMP engine.  It rebuilds the model exactly as the
 * reference main() does (sciddicaT.c:143-184) but with caller-given grids,
 * parameters and step count, so the oracle restatement can be checked on
 * seeded synthetic inputs.  Test infrastructure only. */
#include <cal2D.h>
#include <cal2DRun.h>
#include <string.h>

#define NUMBER_OF_OUTFLOWS 4
struct sciddicaTSubstates {
	struct CALSubstate2Dr *z;
	struct CALSubstate2Dr *h;
	struct CALSubstate2Dr *f[NUMBER_OF_OUTFLOWS];
};
struct sciddicaTParameters {
	CALParameterr epsilon;
	CALParameterr r;
};
extern struct sciddicaTSubstates Q;
extern struct sciddicaTParameters P;
void sciddicaT_flows_computation(struct CALModel2D*, int, int);
void sciddicaT_width_update(struct CALModel2D*, int, int);
void sciddicaT_simulation_init(struct CALModel2D*);
void sciddicaTSteering(struct CALModel2D*);

static double g_r, g_eps;

static void harness_init(struct CALModel2D* ca)
{
	sciddicaT_simulation_init(ca);   /* sets P to the example's constants */
	P.r = g_r;
	P.epsilon = g_eps;
}

static void load(struct CALSubstate2Dr* Q_, const double* v, size_t n)
{
	/* calLoadSubstate2Dr semantics (OpenCAL-OMP/source/cal2DIO.c:46-51) */
	memcpy(Q_->current, v, n * sizeof(double));
	memcpy(Q_->next, v, n * sizeof(double));
}

int ref_sciddicat_run(int rows, int cols, const double* z, const double* h,
                      double p_r, double p_eps, int steps,
                      double* out_z, double* out_h)
{
	size_t n = (size_t)rows * cols;
	int k;
	struct CALModel2D* m = calCADef2D(rows, cols, CAL_VON_NEUMANN_NEIGHBORHOOD_2D,
	                                  CAL_SPACE_TOROIDAL, CAL_NO_OPT);
	struct CALRun2D* run = calRunDef2D(m, 1, steps, CAL_UPDATE_IMPLICIT);
	if (!m || !run) return -1;
	calAddElementaryProcess2D(m, sciddicaT_flows_computation);
	calAddElementaryProcess2D(m, sciddicaT_width_update);
	Q.z = calAddSubstate2Dr(m);
	Q.h = calAddSubstate2Dr(m);
	for (k = 0; k < NUMBER_OF_OUTFLOWS; k++) Q.f[k] = calAddSubstate2Dr(m);
	load(Q.z, z, n);
	load(Q.h, h, n);
	g_r = p_r;
	g_eps = p_eps;
	calRunAddInitFunc2D(run, harness_init);
	calRunAddSteeringFunc2D(run, sciddicaTSteering);
	calRun2D(run);
	memcpy(out_z, Q.z->current, n * sizeof(double));
	memcpy(out_h, Q.h->current, n * sizeof(double));
	calRunFinalize2D(run);
	calFinalize2D(m);
	return 0;
}
