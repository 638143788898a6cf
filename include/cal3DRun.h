/* OpenCAL 3D run loop (OpenCAL-OMP/include/cal3DRun.h:18-107), same struct
 * layout and semantics as cal3DRun.c:16-148: init (+update), then per step
 * the global transition, steering (+update), stop condition; finally the
 * finalize callback (+update). */
#ifndef cal3DRun_h
#define cal3DRun_h

#include <cal3D.h>

#ifdef __cplusplus
extern "C" {
#endif

struct CALRun3D {
	struct CALModel3D* ca3D;
	int step;
	int initial_step;
	int final_step; /* CAL_RUN_LOOP (0): run until the stop condition holds */
	enum CALUpdateMode UPDATE_MODE;
	void (*init)(struct CALModel3D*);
	void (*globalTransition)(struct CALModel3D*);
	void (*steering)(struct CALModel3D*);
	CALbyte (*stopCondition)(struct CALModel3D*);
	void (*finalize)(struct CALModel3D*);
};

/* cal3DRun.c:16 */
struct CALRun3D* calRunDef3D(struct CALModel3D* ca3D, int initial_step, int final_step,
                             enum CALUpdateMode UPDATE_MODE);
/* cal3DRun.c:44-75 */
void calRunAddInitFunc3D(struct CALRun3D* simulation, void (*init)(struct CALModel3D*));
void calRunAddGlobalTransitionFunc3D(struct CALRun3D* simulation, void (*globalTransition)(struct CALModel3D*));
void calRunAddSteeringFunc3D(struct CALRun3D* simulation, void (*steering)(struct CALModel3D*));
void calRunAddStopConditionFunc3D(struct CALRun3D* simulation, CALbyte (*stopCondition)(struct CALModel3D*));
void calRunAddFinalizeFunc3D(struct CALRun3D* simulation, void (*finalize)(struct CALModel3D*));
/* cal3DRun.c:77 */
void calRunInitSimulation3D(struct CALRun3D* simulation);
/* cal3DRun.c:88 — returns CAL_FALSE when the stop condition holds. */
CALbyte calRunCAStep3D(struct CALRun3D* simulation);
/* cal3DRun.c:114 */
void calRunFinalizeSimulation3D(struct CALRun3D* simulation);
/* cal3DRun.c:125 — the whole run; the model's state is back in the host
 * mirrors when it returns. */
void calRun3D(struct CALRun3D* simulation);
/* cal3DRun.c:143 — frees the run, not the model. */
void calRunFinalize3D(struct CALRun3D* cal3DRun);

#ifdef __cplusplus
}
#endif
#endif
