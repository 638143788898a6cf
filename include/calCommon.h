/* OpenCAL-compatible common types for the MI355X cellular-automata engine.
 *
 * Every type, enumerator value and macro here keeps the meaning and binary
 * layout of OpenCAL-OMP/include/calCommon.h:14-209 so that model code written
 * against OpenCAL compiles and links unchanged against libopencal_hip.so.
 */
#ifndef calCommon_h
#define calCommon_h

#define CAL_FALSE 0
#define CAL_TRUE 1

typedef char CALbyte;    /* 8-bit substate cell (calCommon.h:20) */
typedef int CALint;      /* integer substate cell (calCommon.h:21) */
typedef double CALreal;  /* real substate cell (calCommon.h:22) */

typedef CALbyte CALParameterb;
typedef CALint CALParameteri;
typedef CALreal CALParameterr;

/* Boundary of the cellular space (calCommon.h:32-35). */
enum CALSpaceBoundaryCondition {
	CAL_SPACE_FLAT = 0,
	CAL_SPACE_TOROIDAL
};

/* Whether the run loop calls calUpdate2D after each callback (calCommon.h:40-43). */
enum CALUpdateMode {
	CAL_UPDATE_EXPLICIT = 0,
	CAL_UPDATE_IMPLICIT
};

/* Full-grid sweeps or the active-cells optimisation (calCommon.h:48-51). */
enum CALOptimization {
	CAL_NO_OPT = 0,
	CAL_OPT_ACTIVE_CELLS
};

enum CALUnsafeState {
	CAL_UNSAFE_ACTIVE,
	CAL_UNSAFE_INACTIVE
};

/* Single toroidal wrap of a neighbour coordinate (calCommon.h:62). */
#define calGetToroidalX(index, size) \
	((index) < 0 ? ((size) + (index)) : ((index) > ((size)-1) ? ((index) - (size)) : (index)))

struct CALCell2D {
	int i; /* row */
	int j; /* column */
};

struct CALCell3D {
	int i;
	int j;
	int k;
};

/* Double-layer substates: `current` is read, `next` is written
 * (calCommon.h:105-140).  In this engine both pointers refer to host mirrors;
 * the device copies are private to the library. */
struct CALSubstate2Db {
	CALbyte* current;
	CALbyte* next;
};

struct CALSubstate2Di {
	CALint* current;
	CALint* next;
};

struct CALSubstate2Dr {
	CALreal* current;
	CALreal* next;
};

struct CALSubstate3Db {
	CALbyte* current;
	CALbyte* next;
};

struct CALSubstate3Di {
	CALint* current;
	CALint* next;
};

struct CALSubstate3Dr {
	CALreal* current;
	CALreal* next;
};

/* final_step value that turns calRun2D into a loop ended by the stop condition. */
#define CAL_RUN_LOOP 0

enum REDUCTION_OPERATION {
	REDUCTION_NONE = 0,
	REDUCTION_MAX,
	REDUCTION_MIN,
	REDUCTION_SUM,
	REDUCTION_PROD,
	REDUCTION_LOGICAL_AND,
	REDUCTION_BINARY_AND,
	REDUCTION_LOGICAL_OR,
	REDUCTION_BINARY_OR,
	REDUCTION_LOGICAL_XOR,
	REDUCTION_BINARY_XOR
};

#endif
