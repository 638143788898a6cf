/* OpenCAL 3D model API, served by the MI355X engine (libopencal_hip.so).
 *
 * The structs keep the field order and types of OpenCAL-OMP/include/cal3D.h:33-87
 * (an OpenMP build, whose `locks` member is pointer-sized), and every
 * function keeps the reference signature and meaning, cited per declaration.
 * Cell (i, j, k) of a substate is element k*rows*columns + i*columns + j
 * (cal3DBuffer.h:131-136).
 *
 * As in 2D (cal2D.h): substates live on the GPU while a bound device program
 * runs, and the `current`/`next` pointers are host mirrors that the accessor
 * functions keep coherent (calhipSyncHost3D() before dereferencing them
 * directly during a run); elementary processes run on the GPU through the
 * generic device program (calhip3D.h) — there is no CPU execution path.
 */
#ifndef cal3D_h
#define cal3D_h

#include <calCommon.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Neighbourhood ids (cal3D.h:33-37). */
enum CALNeighborhood3D {
	CAL_CUSTOM_NEIGHBORHOOD_3D,
	CAL_VON_NEUMANN_NEIGHBORHOOD_3D,
	CAL_MOORE_NEIGHBORHOOD_3D
};

/* Active-cell set (cal3D.h:43-49): flag grid, per-thread add/remove deltas,
 * the list of active cells in (i, j, k) order and its length. */
struct CALActiveCells3D {
	CALbyte* flags;
	int* size_next;
	struct CALCell3D* cells;
	int size_current;
	int num_threads;
};

struct CALModel3D;

/* Elementary process: called for cell (i, j, k) (cal3D.h:87). */
typedef void (*CALCallbackFunc3D)(struct CALModel3D* ca3D, int i, int j, int k);

/* Cellular automaton (cal3D.h:55-81). */
struct CALModel3D {
	int rows;
	int columns;
	int slices;
	enum CALSpaceBoundaryCondition T;
	enum CALOptimization OPTIMIZATION;
	struct CALActiveCells3D A;

	struct CALCell3D* X;
	int sizeof_X;
	enum CALNeighborhood3D X_id;

	struct CALSubstate3Db** pQb_array;
	struct CALSubstate3Di** pQi_array;
	struct CALSubstate3Dr** pQr_array;
	int sizeof_pQb_array;
	int sizeof_pQi_array;
	int sizeof_pQr_array;

	CALCallbackFunc3D* elementary_processes;
	int num_of_elementary_processes;

	void* locks; /* position of CAL_LOCKS_DEFINE(locks) in OpenMP builds; unused */
	enum CALUnsafeState is_safe;
};

/* cal3D.c:188 — allocate the model, build the predefined neighbourhood
 * (von Neumann cal3D.c:29-52, Moore cal3D.c:58-101). */
struct CALModel3D* calCADef3D(int rows, int columns, int slices, enum CALNeighborhood3D CAL_NEIGHBORHOOD_3D,
                              enum CALSpaceBoundaryCondition CAL_TOROIDALITY,
                              enum CALOptimization CAL_OPTIMIZATION);

/* cal3D.c:262 — marks the model as using the unsafe operations (cal3DUnsafe.h). */
void calSetUnsafe3D(struct CALModel3D* ca3D);

/* cal3D.c:267-380 — active-cell set. */
void calAddActiveCell3D(struct CALModel3D* ca3D, int i, int j, int k);
void calAddActiveCellX3D(struct CALModel3D* ca3D, int i, int j, int k, int n);
void calRemoveActiveCell3D(struct CALModel3D* ca3D, int i, int j, int k);
void calUpdateActiveCells3D(struct CALModel3D* ca3D);

/* cal3D.c:382 — append (i, j, k) to the neighbourhood. */
struct CALCell3D* calAddNeighbor3D(struct CALModel3D* ca3D, int i, int j, int k);

/* cal3D.c:410-538 — substates (double- and single-layer). */
struct CALSubstate3Db* calAddSubstate3Db(struct CALModel3D* ca3D);
struct CALSubstate3Di* calAddSubstate3Di(struct CALModel3D* ca3D);
struct CALSubstate3Dr* calAddSubstate3Dr(struct CALModel3D* ca3D);
struct CALSubstate3Db* calAddSingleLayerSubstate3Db(struct CALModel3D* ca3D);
struct CALSubstate3Di* calAddSingleLayerSubstate3Di(struct CALModel3D* ca3D);
struct CALSubstate3Dr* calAddSingleLayerSubstate3Dr(struct CALModel3D* ca3D);

/* cal3D.c:540 — register an elementary process. */
CALCallbackFunc3D* calAddElementaryProcess3D(struct CALModel3D* ca3D, CALCallbackFunc3D elementary_process);

/* cal3D.c:565-631 — set a substate (both layers / next only) to a constant,
 * over the active cells when the active list exists, else over the grid. */
void calInitSubstate3Db(struct CALModel3D* ca3D, struct CALSubstate3Db* Q, CALbyte value);
void calInitSubstate3Di(struct CALModel3D* ca3D, struct CALSubstate3Di* Q, CALint value);
void calInitSubstate3Dr(struct CALModel3D* ca3D, struct CALSubstate3Dr* Q, CALreal value);
void calInitSubstateNext3Db(struct CALModel3D* ca3D, struct CALSubstate3Db* Q, CALbyte value);
void calInitSubstateNext3Di(struct CALModel3D* ca3D, struct CALSubstate3Di* Q, CALint value);
void calInitSubstateNext3Dr(struct CALModel3D* ca3D, struct CALSubstate3Dr* Q, CALreal value);

/* cal3D.c:635-654 — next -> current for one substate. */
void calUpdateSubstate3Db(struct CALModel3D* ca3D, struct CALSubstate3Db* Q);
void calUpdateSubstate3Di(struct CALModel3D* ca3D, struct CALSubstate3Di* Q);
void calUpdateSubstate3Dr(struct CALModel3D* ca3D, struct CALSubstate3Dr* Q);

/* cal3D.c:658 — one elementary process over the grid / active cells, on the
 * GPU through the generic device program. */
void calApplyElementaryProcess3D(struct CALModel3D* ca3D, CALCallbackFunc3D elementary_process);

/* cal3D.c:677 — every elementary process in order, each followed by
 * calUpdate3D, on the GPU. */
void calGlobalTransitionFunction3D(struct CALModel3D* ca3D);

/* cal3D.c:697 — rebuild the active set (active mode), then next -> current
 * for every double-layer substate. */
void calUpdate3D(struct CALModel3D* ca3D);

/* cal3D.c:718-857 — per-cell accessors (host side; reads copy the substate
 * back from the GPU once per device step, writes bring the whole model back
 * first when a device run left it there). */
void calInit3Db(struct CALModel3D* ca3D, struct CALSubstate3Db* Q, int i, int j, int k, CALbyte value);
void calInit3Di(struct CALModel3D* ca3D, struct CALSubstate3Di* Q, int i, int j, int k, CALint value);
void calInit3Dr(struct CALModel3D* ca3D, struct CALSubstate3Dr* Q, int i, int j, int k, CALreal value);
CALbyte calGet3Db(struct CALModel3D* ca3D, struct CALSubstate3Db* Q, int i, int j, int k);
CALint calGet3Di(struct CALModel3D* ca3D, struct CALSubstate3Di* Q, int i, int j, int k);
CALreal calGet3Dr(struct CALModel3D* ca3D, struct CALSubstate3Dr* Q, int i, int j, int k);
CALbyte calGetX3Db(struct CALModel3D* ca3D, struct CALSubstate3Db* Q, int i, int j, int k, int n);
CALint calGetX3Di(struct CALModel3D* ca3D, struct CALSubstate3Di* Q, int i, int j, int k, int n);
CALreal calGetX3Dr(struct CALModel3D* ca3D, struct CALSubstate3Dr* Q, int i, int j, int k, int n);
void calSet3Db(struct CALModel3D* ca3D, struct CALSubstate3Db* Q, int i, int j, int k, CALbyte value);
void calSet3Di(struct CALModel3D* ca3D, struct CALSubstate3Di* Q, int i, int j, int k, CALint value);
void calSet3Dr(struct CALModel3D* ca3D, struct CALSubstate3Dr* Q, int i, int j, int k, CALreal value);
void calSetCurrent3Db(struct CALModel3D* ca3D, struct CALSubstate3Db* Q, int i, int j, int k, CALbyte value);
void calSetCurrent3Di(struct CALModel3D* ca3D, struct CALSubstate3Di* Q, int i, int j, int k, CALint value);
void calSetCurrent3Dr(struct CALModel3D* ca3D, struct CALSubstate3Dr* Q, int i, int j, int k, CALreal value);

/* cal3D.c:860 — release the model, its substates and its device memory. */
void calFinalize3D(struct CALModel3D* ca3D);

#ifdef __cplusplus
}
#endif
#endif
