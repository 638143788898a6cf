/* OpenCAL 2D run loop (OpenCAL-OMP/include/cal2DRun.h:21-110), same struct
 * layout and semantics.  calRun2D keeps the reference order: init (+update),
 * then per step the global transition, steering (+update), stop condition;
 * finally the finalize callback (+update). */
#ifndef cal2DRun_h
#define cal2DRun_h

#include <cal2D.h>

#ifdef __cplusplus
extern "C" {
#endif

struct CALRun2D {
	struct CALModel2D* ca2D;
	int step;
	int initial_step;
	int final_step; /* CAL_RUN_LOOP (0): run until the stop condition holds */
	enum CALUpdateMode UPDATE_MODE;
	void (*init)(struct CALModel2D*);
	void (*globalTransition)(struct CALModel2D*);
	void (*steering)(struct CALModel2D*);
	CALbyte (*stopCondition)(struct CALModel2D*);
	void (*finalize)(struct CALModel2D*);
};

/* cal2DRun.c:16 */
struct CALRun2D* calRunDef2D(struct CALModel2D* ca2D, int initial_step, int final_step,
                             enum CALUpdateMode UPDATE_MODE);
/* cal2DRun.c:44-75 */
void calRunAddInitFunc2D(struct CALRun2D* simulation, void (*init)(struct CALModel2D*));
void calRunAddGlobalTransitionFunc2D(struct CALRun2D* simulation, void (*globalTransition)(struct CALModel2D*));
void calRunAddSteeringFunc2D(struct CALRun2D* simulation, void (*steering)(struct CALModel2D*));
void calRunAddStopConditionFunc2D(struct CALRun2D* simulation, CALbyte (*stopCondition)(struct CALModel2D*));
void calRunAddFinalizeFunc2D(struct CALRun2D* simulation, void (*finalize)(struct CALModel2D*));
/* cal2DRun.c:77 */
void calRunInitSimulation2D(struct CALRun2D* simulation);
/* cal2DRun.c:88 — returns CAL_FALSE when the stop condition holds. */
CALbyte calRunCAStep2D(struct CALRun2D* simulation);
/* cal2DRun.c:114 */
void calRunFinalizeSimulation2D(struct CALRun2D* simulation);
/* cal2DRun.c:125 — the whole run; the model's state is back in the host
 * mirrors when it returns. */
void calRun2D(struct CALRun2D* simulation);
/* cal2DRun.c:142 — frees the run, not the model. */
void calRunFinalize2D(struct CALRun2D* cal2DRun);

#ifdef __cplusplus
}
#endif
#endif
