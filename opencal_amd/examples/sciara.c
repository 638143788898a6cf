/* SCIARA-fv2 written against the OpenCAL C API as the reference example
 * builds it (OpenCAL-OMP/examples/calomp_sciara-fv2/source/Sciara.cpp:185-356
 * and main.cpp), with caller-given arrays in place of the GIS files, plus the
 * one call that moves it to the MI355X: calhipBindSciara2D.
 *
 * The init callback (border mask, power-law parameters, vents in the active
 * set), the steering and the stop condition are the model's own host code.
 * The elementary processes are registered by identity; on this engine the
 * SCIARA device program executes them (there is no CPU execution path), and
 * the steering runs as its device twin unless the binding leaves it out. */
#include <cal2D.h>
#include <cal2DRun.h>
#include <calhip2D.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NUMBER_OF_OUTFLOWS 8

typedef struct {
	struct CALModel2D* model;
	struct CALRun2D* run;
	struct CALSubstate2Dr *Sz, *Slt, *St, *f[NUMBER_OF_OUTFLOWS];
	struct CALSubstate2Db* Mb;
	struct CALSubstate2Di* Mv;
	struct CALSubstate2Dr *Msl, *Sz_t0;
	CALreal Pclock, PTsol, PTvent, Pr_Tsol, Pr_Tvent, Phc_Tsol, Phc_Tvent;
	CALreal Pcool, Prho, Pepsilon, Psigma, Pcv, Pc, Pac, rad2;
	CALreal a, b, c, d;
	CALreal elapsed_time, effusion_duration;
	unsigned int emission_time;
	int n_vents, n_rates;
	int* vent_rc;
	CALreal* rates;
	int active;
} sciara_t;

static sciara_t* SC;

/* Processes of Sciara.cpp:3-175, executed by the device program. */
static void sciara_vents(struct CALModel2D* ca, int i, int j) { (void)ca, (void)i, (void)j; }
static void sciara_flows(struct CALModel2D* ca, int i, int j) { (void)ca, (void)i, (void)j; }
static void sciara_width(struct CALModel2D* ca, int i, int j) { (void)ca, (void)i, (void)j; }
static void sciara_temperature(struct CALModel2D* ca, int i, int j) { (void)ca, (void)i, (void)j; }
static void sciara_remove(struct CALModel2D* ca, int i, int j) { (void)ca, (void)i, (void)j; }

/* border mask: grid edge cells and cells next to a negative altitude
 * (Sciara.cpp:204-252) */
static void mark_border(struct CALModel2D* ca, int i, int j)
{
	calSetCurrent2Db(ca, SC->Mb, i, j, CAL_TRUE);
	if (SC->active) calAddActiveCell2D(ca, i, j);
}

static void make_border(struct CALModel2D* ca)
{
	int i, j, k;
	for (j = 0; j < ca->columns; j++)
		if (calGet2Dr(ca, SC->Sz, 0, j) >= 0) mark_border(ca, 0, j);
	for (j = 0; j < ca->columns; j++)
		if (calGet2Dr(ca, SC->Sz, ca->rows - 1, j) >= 0) mark_border(ca, ca->rows - 1, j);
	for (i = 0; i < ca->rows; i++)
		if (calGet2Dr(ca, SC->Sz, i, 0) >= 0) mark_border(ca, i, 0);
	for (i = 0; i < ca->rows; i++)
		if (calGet2Dr(ca, SC->Sz, i, ca->columns - 1) >= 0) mark_border(ca, i, ca->columns - 1);
	for (i = 1; i < ca->rows - 1; i++)
		for (j = 1; j < ca->columns - 1; j++)
			if (calGet2Dr(ca, SC->Sz, i, j) >= 0)
				for (k = 1; k < ca->sizeof_X; k++)
					if (calGetX2Dr(ca, SC->Sz, i, j, k) < 0) {
						mark_border(ca, i, j);
						break;
					}
}

/* k1, k2 of 10^(k1 + k2 T) through (PTsol, sol) and (PTvent, vent)
 * (Sciara.cpp:199-202) */
static void power_law(CALreal sol, CALreal vent, CALreal* k1, CALreal* k2)
{
	*k2 = (log10(vent) - log10(sol)) / (SC->PTvent - SC->PTsol);
	*k1 = log10(sol) - *k2 * SC->PTsol;
}

/* simulationInitialize (Sciara.cpp:254-285) */
static void sciara_simulation_init(struct CALModel2D* ca)
{
	int v;
	SC->elapsed_time = 0;
	SC->effusion_duration = (CALreal)(SC->emission_time * (unsigned int)SC->n_rates);
	make_border(ca);
	power_law(SC->Pr_Tsol, SC->Pr_Tvent, &SC->a, &SC->b);
	power_law(SC->Phc_Tsol, SC->Phc_Tvent, &SC->c, &SC->d);
	if (SC->active)
		for (v = 0; v < SC->n_vents; v++) calAddActiveCell2D(ca, SC->vent_rc[2 * v], SC->vent_rc[2 * v + 1]);
	calUpdate2D(ca);
}

/* steering (Sciara.cpp:185-197) */
static void sciara_steering(struct CALModel2D* ca)
{
	int i, j, k;
	for (k = 0; k < NUMBER_OF_OUTFLOWS; ++k) calInitSubstate2Dr(ca, SC->f[k], 0);
	for (i = 0; i < ca->rows; i++)
		for (j = 0; j < ca->columns; j++)
			if (calGet2Db(ca, SC->Mb, i, j) == CAL_TRUE) {
				calSet2Dr(ca, SC->Slt, i, j, 0);
				calSet2Dr(ca, SC->St, i, j, 0);
			}
	SC->elapsed_time += SC->Pclock;
}

/* stopCondition (Sciara.cpp:177-183) */
static CALbyte sciara_stop(struct CALModel2D* ca)
{
	(void)ca;
	return SC->elapsed_time >= SC->effusion_duration ? CAL_TRUE : CAL_FALSE;
}

/* Model of rows x cols cells of side cell_size [m] on the altitude z0.
 * params: Pclock, PTsol, PTvent, Pr(Tsol), Pr(Tvent), Phc(Tsol), Phc(Tvent),
 * Pcool, Prho, Pepsilon, Psigma, Pcv.  vents: n_vents (row, col); rates:
 * n_vents x n_rates [m^3/s], each held for emission_time seconds.
 * bind: 0 none, 1 device program with the steering twin, 2 device program
 * with the steering left to the host callback. */
void* sciara_create(int rows, int cols, const double* z0, double cell_size, const double* params, int n_vents,
                    const int* vent_rc, int n_rates, const double* rates, unsigned int emission_time, int steps,
                    int active, int bind)
{
	size_t n = (size_t)rows * cols, c;
	int k;
	sciara_t* s = (sciara_t*)calloc(1, sizeof(sciara_t));
	if (!s) return NULL;
	SC = s;
	s->active = active;
	s->model = calCADef2D(rows, cols, CAL_MOORE_NEIGHBORHOOD_2D, CAL_SPACE_FLAT,
	                      active ? CAL_OPT_ACTIVE_CELLS : CAL_NO_OPT);
	if (!s->model) return NULL;
	s->Sz = calAddSubstate2Dr(s->model);
	s->Slt = calAddSubstate2Dr(s->model);
	s->St = calAddSubstate2Dr(s->model);
	s->Mb = calAddSingleLayerSubstate2Db(s->model);
	s->Mv = calAddSingleLayerSubstate2Di(s->model);
	s->Msl = calAddSingleLayerSubstate2Dr(s->model);
	s->Sz_t0 = calAddSingleLayerSubstate2Dr(s->model);
	calInitSubstate2Dr(s->model, s->Sz, 0);
	calInitSubstate2Dr(s->model, s->Slt, 0);
	calInitSubstate2Dr(s->model, s->St, 0);
	for (c = 0; c < n; ++c) {
		s->Mb->current[c] = CAL_FALSE;
		s->Mv->current[c] = 0;
		s->Msl->current[c] = 0;
		s->Sz_t0->current[c] = 0;
	}
	for (k = 0; k < NUMBER_OF_OUTFLOWS; ++k) {
		s->f[k] = calAddSubstate2Dr(s->model);
		calInitSubstate2Dr(s->model, s->f[k], 0);
	}
	calAddElementaryProcess2D(s->model, sciara_vents);
	calAddElementaryProcess2D(s->model, sciara_flows);
	calAddElementaryProcess2D(s->model, sciara_width);
	calAddElementaryProcess2D(s->model, sciara_temperature);
	if (active) calAddElementaryProcess2D(s->model, sciara_remove);
	s->run = calRunDef2D(s->model, 1, steps, CAL_UPDATE_IMPLICIT);
	calRunAddInitFunc2D(s->run, sciara_simulation_init);
	calRunAddSteeringFunc2D(s->run, sciara_steering);
	calRunAddStopConditionFunc2D(s->run, sciara_stop);

	/* loadConfiguration: parameters, morphology, vents, emission rates */
	s->Pclock = params[0];
	s->PTsol = params[1];
	s->PTvent = params[2];
	s->Pr_Tsol = params[3];
	s->Pr_Tvent = params[4];
	s->Phc_Tsol = params[5];
	s->Phc_Tvent = params[6];
	s->Pcool = params[7];
	s->Prho = params[8];
	s->Pepsilon = params[9];
	s->Psigma = params[10];
	s->Pcv = params[11];
	s->Pc = cell_size;
	s->Pac = cell_size * cell_size;
	s->rad2 = sqrt(2.0);
	memcpy(s->Sz->current, z0, n * sizeof(double));
	memcpy(s->Sz->next, z0, n * sizeof(double));
	memcpy(s->Sz_t0->current, z0, n * sizeof(double));
	s->emission_time = emission_time;
	s->n_vents = n_vents;
	s->n_rates = n_rates;
	s->vent_rc = (int*)malloc(sizeof(int) * 2 * (size_t)(n_vents + 1));
	s->rates = (CALreal*)malloc(sizeof(CALreal) * ((size_t)n_vents * n_rates + 1));
	memcpy(s->vent_rc, vent_rc, sizeof(int) * 2 * (size_t)n_vents);
	memcpy(s->rates, rates, sizeof(CALreal) * (size_t)n_vents * n_rates);
	for (k = 0; k < n_vents; ++k) s->Mv->current[(size_t)vent_rc[2 * k] * cols + vent_rc[2 * k + 1]] = k + 1;

	if (bind) {
		struct calhipSciaraBinding b;
		memset(&b, 0, sizeof(b));
		b.Sz = s->Sz;
		b.Slt = s->Slt;
		b.St = s->St;
		for (k = 0; k < NUMBER_OF_OUTFLOWS; ++k) b.f[k] = s->f[k];
		b.Mb = s->Mb;
		b.Msl = s->Msl;
		b.vents = sciara_vents;
		b.flows = sciara_flows;
		b.width = sciara_width;
		b.temperature = sciara_temperature;
		b.remove = active ? sciara_remove : NULL;
		b.steering = bind == 1 ? sciara_steering : NULL;
		b.Pclock = &s->Pclock;
		b.PTsol = &s->PTsol;
		b.PTvent = &s->PTvent;
		b.Pc = &s->Pc;
		b.Pac = &s->Pac;
		b.a = &s->a;
		b.b = &s->b;
		b.c = &s->c;
		b.d = &s->d;
		b.rad2 = &s->rad2;
		b.Pcool = &s->Pcool;
		b.Prho = &s->Prho;
		b.Pepsilon = &s->Pepsilon;
		b.Psigma = &s->Psigma;
		b.Pcv = &s->Pcv;
		b.elapsed_time = &s->elapsed_time;
		b.emission_time = &s->emission_time;
		b.n_vents = n_vents;
		b.vent_rc = s->vent_rc;
		b.n_rates = n_rates;
		b.rates = s->rates;
		if (calhipBindSciara2D(s->model, &b)) return NULL;
	}
	return s;
}

/* calRun2D; returns the step counter after the run (CALRun2D.step). */
int sciara_run(void* handle)
{
	sciara_t* s = (sciara_t*)handle;
	SC = s;
	calRun2D(s->run);
	return calhipGetLastError() ? -1 : s->run->step;
}

/* calRunInitSimulation2D alone (for callers that step the model themselves) */
int sciara_init(void* handle)
{
	sciara_t* s = (sciara_t*)handle;
	SC = s;
	calRunInitSimulation2D(s->run);
	return calhipGetLastError() ? -1 : 0;
}

/* up to n calRunCAStep2D calls; stops early when the stop condition holds.
 * Returns the number of steps run, -1 on error. */
int sciara_steps(void* handle, int n)
{
	sciara_t* s = (sciara_t*)handle;
	int k;
	SC = s;
	for (k = 0; k < n; ++k) {
		CALbyte again = calRunCAStep2D(s->run);
		if (calhipGetLastError()) return -1;
		if (!again) return k + 1;
	}
	return n;
}

int sciara_sync(void* handle) { return calhipSynchronize2D(((sciara_t*)handle)->model); }

/* which: 0 altitude, 1 lava thickness, 2 temperature, 3 solidified lava */
int sciara_read(void* handle, int which, double* out)
{
	sciara_t* s = (sciara_t*)handle;
	struct CALSubstate2Dr* q = which == 0 ? s->Sz : which == 1 ? s->Slt : which == 2 ? s->St : s->Msl;
	size_t n = (size_t)s->model->rows * s->model->columns;
	if (calhipSyncHost2D(s->model)) return -1;
	memcpy(out, q->current, n * sizeof(double));
	return 0;
}

int sciara_active_count(void* handle)
{
	sciara_t* s = (sciara_t*)handle;
	if (calhipSyncHost2D(s->model)) return -1;
	return s->model->A.size_current;
}

double sciara_elapsed(void* handle) { return ((sciara_t*)handle)->elapsed_time; }

struct CALModel2D* sciara_model(void* handle) { return ((sciara_t*)handle)->model; }

void sciara_destroy(void* handle)
{
	sciara_t* s = (sciara_t*)handle;
	if (!s) return;
	calRunFinalize2D(s->run);
	calFinalize2D(s->model);
	free(s->vent_rc);
	free(s->rates);
	if (SC == s) SC = NULL;
	free(s);
}
