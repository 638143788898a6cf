// Harness around the reference SCIARA-fv2 transition functions
// (OpenCAL-OMP/examples/calomp_sciara-fv2/source/Sciara.cpp, linked unmodified
// against the OpenCAL-OMP engine).  It performs what initSciara()
// (Sciara.cpp:287-356) plus loadConfiguration() (io.cpp) do, but from
// caller-given arrays instead of GIS files, then runs calRun2D.
// Test infrastructure only.
#include "Sciara.h"
#include <string.h>

Sciara* sciara;
int active;

void updateVentsEmission(struct CALModel2D*, int, int);
void empiricalFlows(struct CALModel2D*, int, int);
void width_update(struct CALModel2D*, int, int);
void updateTemperature(struct CALModel2D*, int, int);
void removeActiveCells(struct CALModel2D*, int, int);
CALbyte stopCondition(struct CALModel2D*);
void steering(struct CALModel2D*);
void simulationInitialize(struct CALModel2D*);

extern "C" {

// params[12] = Pclock, PTsol, PTvent, Pr_Tsol, Pr_Tvent, Phc_Tsol, Phc_Tvent,
//              Pcool, Prho, Pepsilon, Psigma, Pcv;   cell_size = Pc.
// vents: n_vents (row, col) pairs; rates: n_vents x n_rates row-major [m^3/s].
// out: z, h, T, Msl (rows*cols each); *out_steps = steps actually run;
// *out_active = A.size_current at the end (active mode).
int ref_sciara_run(int rows, int cols, const double* z0, double cell_size,
                   const double* params, int n_vents, const int* vent_rc,
                   int n_rates, const double* rates, unsigned emission_time,
                   int steps, int use_active,
                   double* out_z, double* out_h, double* out_T, double* out_Msl,
                   int* out_steps, int* out_active)
{
	size_t n = (size_t)rows * cols;
	active = use_active;
	sciara = new Sciara;
	sciara->rows = rows;
	sciara->cols = cols;
	sciara->elapsed_time = 0.0;
	sciara->rad2 = sqrt(2.0);
	sciara->effusion_duration = 0;
	sciara->model = calCADef2D(rows, cols, CAL_MOORE_NEIGHBORHOOD_2D, CAL_SPACE_FLAT,
	                           active ? CAL_OPT_ACTIVE_CELLS : CAL_NO_OPT);
	sciara->substates = new SciaraSubstates();
	SciaraSubstates* S = sciara->substates;
	S->Sz = calAddSubstate2Dr(sciara->model);
	S->Slt = calAddSubstate2Dr(sciara->model);
	S->St = calAddSubstate2Dr(sciara->model);
	S->Mb = calAddSingleLayerSubstate2Db(sciara->model);
	S->Mv = calAddSingleLayerSubstate2Di(sciara->model);
	S->Msl = calAddSingleLayerSubstate2Dr(sciara->model);
	S->Sz_t0 = calAddSingleLayerSubstate2Dr(sciara->model);
	calInitSubstate2Dr(sciara->model, S->Sz, 0);
	calInitSubstate2Dr(sciara->model, S->Slt, 0);
	calInitSubstate2Dr(sciara->model, S->St, 0);
	for (size_t c = 0; c < n; ++c) {
		S->Mb->current[c] = CAL_FALSE;
		S->Mv->current[c] = 0;
		S->Msl->current[c] = 0;
		S->Sz_t0->current[c] = 0;
	}
	for (int k = 0; k < NUMBER_OF_OUTFLOWS; ++k) {
		S->f[k] = calAddSubstate2Dr(sciara->model);
		calInitSubstate2Dr(sciara->model, S->f[k], 0);
	}
	calAddElementaryProcess2D(sciara->model, updateVentsEmission);
	calAddElementaryProcess2D(sciara->model, empiricalFlows);
	calAddElementaryProcess2D(sciara->model, width_update);
	calAddElementaryProcess2D(sciara->model, updateTemperature);
	if (active)
		calAddElementaryProcess2D(sciara->model, removeActiveCells);
	sciara->run = calRunDef2D(sciara->model, 1, steps, CAL_UPDATE_IMPLICIT);
	calRunAddInitFunc2D(sciara->run, simulationInitialize);
	calRunAddSteeringFunc2D(sciara->run, steering);
	calRunAddStopConditionFunc2D(sciara->run, stopCondition);

	// loadConfiguration(): parameters, morphology, vents, emission rates
	sciara->Pclock = params[0];
	sciara->PTsol = params[1];
	sciara->PTvent = params[2];
	sciara->Pr_Tsol = params[3];
	sciara->Pr_Tvent = params[4];
	sciara->Phc_Tsol = params[5];
	sciara->Phc_Tvent = params[6];
	sciara->Pcool = params[7];
	sciara->Prho = params[8];
	sciara->Pepsilon = params[9];
	sciara->Psigma = params[10];
	sciara->Pcv = params[11];
	sciara->Pc = cell_size;
	sciara->Pac = cell_size * cell_size;
	memcpy(S->Sz->current, z0, n * sizeof(double));
	memcpy(S->Sz->next, z0, n * sizeof(double));
	memcpy(S->Sz_t0->current, z0, n * sizeof(double));
	sciara->emission_time = emission_time;
	sciara->emission_rate.clear();
	sciara->vent.clear();
	for (int v = 0; v < n_vents; ++v) {
		TEmissionRate er;
		er.set_vent_id(v + 1);
		for (int k = 0; k < n_rates; ++k)
			er.emission_rate().push_back(rates[(size_t)v * n_rates + k]);
		sciara->emission_rate.push_back(er);
	}
	for (int v = 0; v < n_vents; ++v) {
		TVent vent;
		vent.set_y(vent_rc[2 * v]);
		vent.set_x(vent_rc[2 * v + 1]);
		vent.set_vent_id(v + 1);
		vent.setEmissionRate(sciara->emission_rate, v + 1);
		sciara->vent.push_back(vent);
		S->Mv->current[(size_t)vent_rc[2 * v] * cols + vent_rc[2 * v + 1]] = v + 1;
	}

	calRun2D(sciara->run);

	memcpy(out_z, S->Sz->current, n * sizeof(double));
	memcpy(out_h, S->Slt->current, n * sizeof(double));
	memcpy(out_T, S->St->current, n * sizeof(double));
	memcpy(out_Msl, S->Msl->current, n * sizeof(double));
	if (out_steps) *out_steps = sciara->run->step;
	if (out_active) *out_active = sciara->model->A.size_current;
	calRunFinalize2D(sciara->run);
	calFinalize2D(sciara->model);
	delete sciara->substates;
	delete sciara;
	sciara = 0;
	return 0;
}

}  // extern "C"
