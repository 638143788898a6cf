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

// The step loop of calRun2D (cal2DRun.c:129-136) for n steps from r->step,
// which ends on the number after the last step run.  When no host callback
// runs between steps (no user global transition, no stop condition, and a
// steering that is either absent or has a device twin) the steps after the
// first go to the device program in one call, which may batch them into
// fewer launches (e.g. two SciddicaT steps per sweep).
extern "C" int calhipRunSteps2D(struct CALRun2D* r, int n)
{
	ModelX* x = X(r->ca2D);
	const bool twin = r->steering && x->prog && x->prog_steering == r->steering;
	const bool batchable = !r->globalTransition && !r->stopCondition && x->prog &&
	                       (!r->steering || (twin && r->UPDATE_MODE == CAL_UPDATE_IMPLICIT));
	int done = 0;
	while (done < n) {
		// The steps go to the program together once the model is on the
		// device: from the first step when it already is (and the program
		// covers its processes, which calRunCAStep2D would check), otherwise
		// after the first step, which uploads it.
		if (batchable && x->phase == Phase::Device && (done > 0 || program_covers_model(x))) {
			const int rest = n - done;
			device_changed(x);
			if (x->prog->run_steps(x, rest, twin)) return -1;
			x->stats.steps += rest;
			r->step += rest;
			return n;
		}
		const CALbyte again = calRunCAStep2D(r);
		if (calhipGetLastError()) return -1;
		done++;
		if (!again) return done;   // the reference leaves the step number at the stopping step
		r->step++;
	}
	return done;
}

// cal2DRun.c:125-139.  The model's state is in the host mirrors on return.
extern "C" void calRun2D(struct CALRun2D* r)
{
	calRunInitSimulation2D(r);
	ModelX* x = X(r->ca2D);
	r->step = r->initial_step;
	if (r->final_step == CAL_RUN_LOOP) {
		for (;; r->step++) {
			if (!calRunCAStep2D(r)) break;
			if (calhipGetLastError()) break;
		}
	} else if (r->final_step >= r->initial_step) {
		calhipRunSteps2D(r, r->final_step - r->initial_step + 1);
	}
	calRunFinalizeSimulation2D(r);
	ensure_host(x);
}

// cal2DRun.c:142-147
extern "C" void calRunFinalize2D(struct CALRun2D* r) { free(r); }
