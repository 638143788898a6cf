// The binding a maintainer adds to the reference SCIARA-fv2 program
// (OpenCAL-OMP/examples/calomp_sciara-fv2/source/main.cpp) to run it on the
// MI355X engine: called once after loadConfiguration(), when the vents and
// emission rates are known.  Built only
// into the dropin_sciara program (oracle/ref/Makefile).
#include "Sciara.h"

#include <calhip2D.h>

#include <algorithm>
#include <vector>

void updateVentsEmission(struct CALModel2D*, int, int);
void empiricalFlows(struct CALModel2D*, int, int);
void width_update(struct CALModel2D*, int, int);
void updateTemperature(struct CALModel2D*, int, int);
void removeActiveCells(struct CALModel2D*, int, int);
void steering(struct CALModel2D*);

int calhip_bind_sciara()
{
	static std::vector<int> rc;          // vent (row, column) pairs
	static std::vector<double> rates;    // n_vents x n_rates, zero-padded
	unsigned n_rates = 0;
	for (auto& v : sciara->vent) n_rates = std::max(n_rates, v.size());
	for (auto& v : sciara->vent) {
		rc.push_back(v.y());
		rc.push_back(v.x());
		for (unsigned k = 0; k < n_rates; ++k) rates.push_back(k < v.size() ? v[k] : 0.0);
	}
	SciaraSubstates* S = sciara->substates;
	calhipSciaraBinding b = {};
	b.Sz = S->Sz;
	b.Slt = S->Slt;
	b.St = S->St;
	b.Mb = S->Mb;
	b.Msl = S->Msl;
	for (int k = 0; k < NUMBER_OF_OUTFLOWS; ++k) b.f[k] = S->f[k];
	b.vents = updateVentsEmission;
	b.flows = empiricalFlows;
	b.width = width_update;
	b.temperature = updateTemperature;
	b.remove = active ? removeActiveCells : NULL;
	b.steering = steering;
	b.Pclock = &sciara->Pclock;
	b.PTsol = &sciara->PTsol;
	b.PTvent = &sciara->PTvent;
	b.Pc = &sciara->Pc;
	b.Pac = &sciara->Pac;
	b.a = &sciara->a;
	b.b = &sciara->b;
	b.c = &sciara->c;
	b.d = &sciara->d;
	b.rad2 = &sciara->rad2;
	b.Pcool = &sciara->Pcool;
	b.Prho = &sciara->Prho;
	b.Pepsilon = &sciara->Pepsilon;
	b.Psigma = &sciara->Psigma;
	b.Pcv = &sciara->Pcv;
	b.elapsed_time = &sciara->elapsed_time;
	b.emission_time = &sciara->emission_time;
	b.n_vents = (int)sciara->vent.size();
	b.vent_rc = rc.data();
	b.n_rates = (int)n_rates;
	b.rates = rates.data();
	return calhipBindSciara2D(sciara->model, &b);
}
