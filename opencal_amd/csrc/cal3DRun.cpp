// OpenCAL 3D run loop (include/cal3DRun.h), semantics of
// OpenCAL-OMP/source/cal3DRun.c:16-148.
#include <stdlib.h>

#include "engine.hpp"

extern "C" {
#include "cal3DRun.h"
}

using namespace calhip;

extern "C" struct CALRun3D* calRunDef3D(struct CALModel3D* ca3D, int initial_step, int final_step,
                                        enum CALUpdateMode mode)
{
	CALRun3D* r = (CALRun3D*)malloc(sizeof(CALRun3D));
	if (!r) return nullptr;
	r->ca3D = ca3D;
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

extern "C" void calRunAddInitFunc3D(struct CALRun3D* r, void (*f)(struct CALModel3D*)) { r->init = f; }
extern "C" void calRunAddGlobalTransitionFunc3D(struct CALRun3D* r, void (*f)(struct CALModel3D*)) { r->globalTransition = f; }
extern "C" void calRunAddSteeringFunc3D(struct CALRun3D* r, void (*f)(struct CALModel3D*)) { r->steering = f; }
extern "C" void calRunAddStopConditionFunc3D(struct CALRun3D* r, CALbyte (*f)(struct CALModel3D*)) { r->stopCondition = f; }
extern "C" void calRunAddFinalizeFunc3D(struct CALRun3D* r, void (*f)(struct CALModel3D*)) { r->finalize = f; }

// cal3DRun.c:77-85
extern "C" void calRunInitSimulation3D(struct CALRun3D* r)
{
	if (r->init) {
		r->init(r->ca3D);
		if (r->UPDATE_MODE == CAL_UPDATE_IMPLICIT) calUpdate3D(r->ca3D);
	}
}

// cal3DRun.c:88-111
extern "C" CALbyte calRunCAStep3D(struct CALRun3D* r)
{
	if (r->globalTransition) {
		r->globalTransition(r->ca3D);
		if (r->UPDATE_MODE == CAL_UPDATE_IMPLICIT) calUpdate3D(r->ca3D);
	} else {
		calGlobalTransitionFunction3D(r->ca3D);
	}
	if (r->steering) {
		r->steering(r->ca3D);
		if (r->UPDATE_MODE == CAL_UPDATE_IMPLICIT) calUpdate3D(r->ca3D);
	}
	if (r->stopCondition && r->stopCondition(r->ca3D)) return CAL_FALSE;
	return CAL_TRUE;
}

// cal3DRun.c:114-122
extern "C" void calRunFinalizeSimulation3D(struct CALRun3D* r)
{
	if (r->finalize) {
		r->finalize(r->ca3D);
		if (r->UPDATE_MODE == CAL_UPDATE_IMPLICIT) calUpdate3D(r->ca3D);
	}
}

// cal3DRun.c:125-139; the model's state is in the host mirrors on return.
extern "C" void calRun3D(struct CALRun3D* r)
{
	calRunInitSimulation3D(r);
	for (r->step = r->initial_step; r->step <= r->final_step || r->final_step == CAL_RUN_LOOP; r->step++) {
		if (!calRunCAStep3D(r)) break;
		if (calhipGetLastError()) break;
	}
	calRunFinalizeSimulation3D(r);
	ensure_host(&X3(r->ca3D)->core);
}

// cal3DRun.c:143-148
extern "C" void calRunFinalize3D(struct CALRun3D* r) { free(r); }
