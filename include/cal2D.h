/* OpenCAL 2D model API, served by the MI355X engine (libopencal_hip.so).
 *
 * The structs keep the field order and types of OpenCAL-OMP/include/cal2D.h:38-88
 * (an OpenMP build, whose trailing `locks` member is pointer-sized), and every
 * function keeps the reference signature and meaning, cited per declaration.
 *
 * Where the engine differs (all of it documented at the declaration):
 *  - substates live on the GPU while a bound device program runs; the
 *    `current`/`next` pointers are host mirrors that the accessor functions
 *    keep coherent.  Code that dereferences them directly during a run calls
 *    calhipSyncHost2D() first (calhip2D.h);
 *  - an elementary process runs on the GPU through the device program bound to
 *    it with calhipBind*2D(); calling the global transition with an unbound
 *    process is an error (there is no CPU execution path).
 */
#ifndef cal2D_h
#define cal2D_h

#include <calCommon.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Neighbourhood ids (cal2D.h:38-44). */
enum CALNeighborhood2D {
	CAL_CUSTOM_NEIGHBORHOOD_2D,
	CAL_VON_NEUMANN_NEIGHBORHOOD_2D,
	CAL_MOORE_NEIGHBORHOOD_2D,
	CAL_HEXAGONAL_NEIGHBORHOOD_2D,
	CAL_HEXAGONAL_NEIGHBORHOOD_ALT_2D
};

/* Neighbour-index shift for odd columns (hex) / odd rows (alt hex) (cal2D.h:46). */
#define CAL_HEXAGONAL_SHIFT 7

/* Active-cell set (cal2D.h:47-53): flag grid, per-thread add/remove deltas,
 * the row-major list of active cells and its length. */
struct CALActiveCells2D {
	CALbyte* flags;
	int* size_next;
	struct CALCell2D* cells;
	int size_current;
	int num_threads;
};

struct CALModel2D;

/* Elementary process: called for cell (i, j) (cal2D.h:59). */
typedef void (*CALCallbackFunc2D)(struct CALModel2D* ca2D, int i, int j);

/* Cellular automaton (cal2D.h:64-88). */
struct CALModel2D {
	int rows;
	int columns;
	enum CALSpaceBoundaryCondition T;
	enum CALUnsafeState is_safe;

	enum CALOptimization OPTIMIZATION;
	struct CALActiveCells2D A;

	struct CALCell2D* X;
	int sizeof_X;
	enum CALNeighborhood2D X_id;

	struct CALSubstate2Db** pQb_array;
	struct CALSubstate2Di** pQi_array;
	struct CALSubstate2Dr** pQr_array;
	int sizeof_pQb_array;
	int sizeof_pQi_array;
	int sizeof_pQr_array;

	CALCallbackFunc2D* elementary_processes;
	int num_of_elementary_processes;

	void* locks; /* position of CAL_LOCKS_DEFINE(locks) in OpenMP builds; unused */
};

/* cal2D.c:252 — allocate the model, build the predefined neighbourhood. */
struct CALModel2D* calCADef2D(int rows, int columns, enum CALNeighborhood2D CAL_NEIGHBORHOOD_2D,
                              enum CALSpaceBoundaryCondition CAL_TOROIDALITY,
                              enum CALOptimization CAL_OPTIMIZATION);

/* cal2D.c:329 — marks the model as using the unsafe operations
 * (cal2DUnsafe.h).  They run on the host mirrors and, in process source, on
 * the GPU through the generic device program (calhip2D.h); the fused
 * programs do not use them. */
void calSetUnsafe2D(struct CALModel2D* ca2D);

/* cal2D.c:333-440 — active-cell set. */
void calAddActiveCell2D(struct CALModel2D* ca2D, int i, int j);
void calAddActiveCellX2D(struct CALModel2D* ca2D, int i, int j, int n);
void calRemoveActiveCell2D(struct CALModel2D* ca2D, int i, int j);
void calUpdateActiveCells2D(struct CALModel2D* ca2D);

/* cal2D.c:444 — append (i, j) to the neighbourhood. */
struct CALCell2D* calAddNeighbor2D(struct CALModel2D* ca2D, int i, int j);

/* cal2D.c:470-596 — substates (double- and single-layer). */
struct CALSubstate2Db* calAddSubstate2Db(struct CALModel2D* ca2D);
struct CALSubstate2Di* calAddSubstate2Di(struct CALModel2D* ca2D);
struct CALSubstate2Dr* calAddSubstate2Dr(struct CALModel2D* ca2D);
struct CALSubstate2Db* calAddSingleLayerSubstate2Db(struct CALModel2D* ca2D);
struct CALSubstate2Di* calAddSingleLayerSubstate2Di(struct CALModel2D* ca2D);
struct CALSubstate2Dr* calAddSingleLayerSubstate2Dr(struct CALModel2D* ca2D);

/* cal2D.c:600 — register an elementary process. */
CALCallbackFunc2D* calAddElementaryProcess2D(struct CALModel2D* ca2D, CALCallbackFunc2D elementary_process);

/* cal2D.c:625-691 — set a substate (both layers / next only) to a constant,
 * over the active cells when the active list exists, else over the grid. */
void calInitSubstate2Db(struct CALModel2D* ca2D, struct CALSubstate2Db* Q, CALbyte value);
void calInitSubstate2Di(struct CALModel2D* ca2D, struct CALSubstate2Di* Q, CALint value);
void calInitSubstate2Dr(struct CALModel2D* ca2D, struct CALSubstate2Dr* Q, CALreal value);
void calInitSubstateNext2Db(struct CALModel2D* ca2D, struct CALSubstate2Db* Q, CALbyte value);
void calInitSubstateNext2Di(struct CALModel2D* ca2D, struct CALSubstate2Di* Q, CALint value);
void calInitSubstateNext2Dr(struct CALModel2D* ca2D, struct CALSubstate2Dr* Q, CALreal value);

/* cal2D.c:694-713 — next -> current for one substate. */
void calUpdateSubstate2Db(struct CALModel2D* ca2D, struct CALSubstate2Db* Q);
void calUpdateSubstate2Di(struct CALModel2D* ca2D, struct CALSubstate2Di* Q);
void calUpdateSubstate2Dr(struct CALModel2D* ca2D, struct CALSubstate2Dr* Q);

/* cal2D.c:717 — one elementary process over the grid / active cells.  Runs
 * the process's device twin (generic, SciddicaT and Life programs: reads the
 * current layers, writes the next ones; the caller's update brings them in);
 * on the SCIARA program, whose kernels merge processes, it reports error 3
 * through calhipGetLastError(). */
void calApplyElementaryProcess2D(struct CALModel2D* ca2D, CALCallbackFunc2D elementary_process);

/* cal2D.c:738 — every elementary process in order, each followed by
 * calUpdate2D.  Executed on the GPU by the bound device program, fused into
 * as few sweeps as the processes' dependencies allow. */
void calGlobalTransitionFunction2D(struct CALModel2D* ca2D);

/* cal2D.c:758 — rebuild the active set (active mode), then next -> current
 * for every double-layer substate.  On the device: a buffer flip of each
 * written layer (full grid), or its copy at the active cells (active mode). */
void calUpdate2D(struct CALModel2D* ca2D);

/* cal2D.c:780-930 — per-cell accessors (host side; they bring the model's
 * state back from the GPU first when a device run left it there). */
void calInit2Db(struct CALModel2D* ca2D, struct CALSubstate2Db* Q, int i, int j, CALbyte value);
void calInit2Di(struct CALModel2D* ca2D, struct CALSubstate2Di* Q, int i, int j, CALint value);
void calInit2Dr(struct CALModel2D* ca2D, struct CALSubstate2Dr* Q, int i, int j, CALreal value);
CALbyte calGet2Db(struct CALModel2D* ca2D, struct CALSubstate2Db* Q, int i, int j);
CALint calGet2Di(struct CALModel2D* ca2D, struct CALSubstate2Di* Q, int i, int j);
CALreal calGet2Dr(struct CALModel2D* ca2D, struct CALSubstate2Dr* Q, int i, int j);
CALbyte calGetX2Db(struct CALModel2D* ca2D, struct CALSubstate2Db* Q, int i, int j, int n);
CALint calGetX2Di(struct CALModel2D* ca2D, struct CALSubstate2Di* Q, int i, int j, int n);
CALreal calGetX2Dr(struct CALModel2D* ca2D, struct CALSubstate2Dr* Q, int i, int j, int n);
void calSet2Db(struct CALModel2D* ca2D, struct CALSubstate2Db* Q, int i, int j, CALbyte value);
void calSet2Di(struct CALModel2D* ca2D, struct CALSubstate2Di* Q, int i, int j, CALint value);
void calSet2Dr(struct CALModel2D* ca2D, struct CALSubstate2Dr* Q, int i, int j, CALreal value);
void calSetCurrent2Db(struct CALModel2D* ca2D, struct CALSubstate2Db* Q, int i, int j, CALbyte value);
void calSetCurrent2Di(struct CALModel2D* ca2D, struct CALSubstate2Di* Q, int i, int j, CALint value);
void calSetCurrent2Dr(struct CALModel2D* ca2D, struct CALSubstate2Dr* Q, int i, int j, CALreal value);

/* cal2D.c:933 — release the model, its substates and its device memory. */
void calFinalize2D(struct CALModel2D* ca2D);

#ifdef __cplusplus
}
#endif
#endif
