// OpenCAL 2D run loop (include/cal2DRun.h), semantics of
// OpenCAL-OMP/source/cal2DRun.c:16-147.
#include <stdlib.h>

#include "engine.hpp"

using namespace calhip;

extern "C" struct CALRun2D* calRunDef2D(struct CALModel2D* ca2D, int initial_step, int final_step,
                                        enum CALUpdateMode mode)
{
	CALRun2D* r = (CALRun2D*)malloc(sizeof(CALRun2D));
	if (!r) return nullptr;
	r->ca2D = ca2D;
	r->step = 0;
	r->initial_step = initial_step;
	r->final_step = final_step;
	r->UPDATE_MODE = mode;
	r->init = nullptr;
	r->globalTransition = nullptr;
	r->steering = nullptr;
	r->stopCondition = nullptr;
	r->finalize = nullptr;
	return r;
}

extern "C" void calRunAddInitFunc2D(struct CALRun2D* r, void (*f)(struct CALModel2D*)) { r->init = f; }
extern "C" void calRunAddGlobalTransitionFunc2D(struct CALRun2D* r, void (*f)(struct CALModel2D*)) { r->globalTransition = f; }
extern "C" void calRunAddSteeringFunc2D(struct CALRun2D* r, void (*f)(struct CALModel2D*)) { r->steering = f; }
extern "C" void calRunAddStopConditionFunc2D(struct CALRun2D* r, CALbyte (*f)(struct CALModel2D*)) { r->stopCondition = f; }
extern "C" void calRunAddFinalizeFunc2D(struct CALRun2D* r, void (*f)(struct CALModel2D*)) { r->finalize = f; }

// cal2DRun.c:77-85
extern "C" void calRunInitSimulation2D(struct CALRun2D* r)
{
	if (r->init) {
		r->init(r->ca2D);
		if (r->UPDATE_MODE == CAL_UPDATE_IMPLICIT) calUpdate2D(r->ca2D);
	}
}

// Runs the steering callback, or its device twin when the bound program
// provides one for exactly this callback.
static void run_steering(struct CALRun2D* r)
{
	ModelX* x = X(r->ca2D);
	if (x->prog && x->prog_steering == r->steering) {
		if (ensure_device(x) == 0) {
			device_changed(x);
			x->prog->steer(x);
		}
	} else {
		r->steering(r->ca2D);
	}
}

// cal2DRun.c:88-112.  With the model's own transition and a steering that
// has a device twin, both go to the program in one call.
extern "C" CALbyte calRunCAStep2D(struct CALRun2D* r)
{
	ModelX* x = X(r->ca2D);
	if (!r->globalTransition && r->steering && x->prog && x->prog_steering == r->steering) {
		device_transition(x, true);
		if (r->UPDATE_MODE == CAL_UPDATE_IMPLICIT) calUpdate2D(r->ca2D);
	} else {
		if (r->globalTransition) {
			r->globalTransition(r->ca2D);
			if (r->UPDATE_MODE == CAL_UPDATE_IMPLICIT) calUpdate2D(r->ca2D);
		} else {
			calGlobalTransitionFunction2D(r->ca2D);
		}
		if (r->steering) {
			run_steering(r);
			if (r->UPDATE_MODE == CAL_UPDATE_IMPLICIT) calUpdate2D(r->ca2D);
		}
	}
	if (r->stopCondition && r->stopCondition(r->ca2D)) return CAL_FALSE;
	return CAL_TRUE;
}

// cal2DRun.c:114-122
extern "C" void calRunFinalizeSimulation2D(struct CALRun2D* r)
{
	if (r->finalize) {
		r->finalize(r->ca2D);
		if (r->UPDATE_MODE == CAL_UPDATE_IMPLICIT) calUpdate2D(r->ca2D);
	}
}

// cal2DRun.c:125-139.  When no host callback runs between steps (no user
// global transition, steering or stop condition) the remaining steps of a
// bounded run go to the device program in one call, which may batch them
// into fewer launches.  The model's state is in the host mirrors on return.
extern "C" void calRun2D(struct CALRun2D* r)
{
	calRunInitSimulation2D(r);
	ModelX* x = X(r->ca2D);
	bool batchable = !r->globalTransition && !r->steering && !r->stopCondition &&
	                 r->final_step != CAL_RUN_LOOP && x->prog;
	if (batchable && r->final_step >= r->initial_step) {
		int n = r->final_step - r->initial_step + 1;
		r->step = r->initial_step;
		calGlobalTransitionFunction2D(r->ca2D);   // binds + uploads, one step
		if (n > 1 && calhipGetLastError() == 0) {
			device_changed(x);
			if (x->prog->run_steps(x, n - 1) == 0) x->stats.steps += n - 1;
		}
		r->step = r->final_step + 1;
	} else {
		for (r->step = r->initial_step; r->step <= r->final_step || r->final_step == CAL_RUN_LOOP; r->step++) {
			if (!calRunCAStep2D(r)) break;
			if (calhipGetLastError()) break;
		}
	}
	calRunFinalizeSimulation2D(r);
	ensure_host(x);
}

// cal2DRun.c:142-147
extern "C" void calRunFinalize2D(struct CALRun2D* r) { free(r); }
