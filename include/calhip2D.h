/* MI355X engine extensions to the OpenCAL 2D API.
 *
 * These entry points have no counterpart in OpenCAL: they attach hand-written
 * HIP device programs to a model's elementary processes, move state between
 * the GPU and the host mirrors, report errors, and expose the two additions
 * that change the method's form (the batched ensemble and the SciddicaT
 * adjoint).  Plain C ABI: pointers, sizes, enums — no C++ or torch types.
 */
#ifndef calhip2D_h
#define calhip2D_h

#include <stddef.h>
#include <cal2D.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- errors */

/* 0 when no error happened since the last calhipClearError(); otherwise a
 * non-zero code (the first error is kept) and a message. */
int calhipGetLastError(void);
const char* calhipGetErrorString(void);
void calhipClearError(void);

/* --------------------------------------------------------------- residency */

/* Bring every substate and the active set of the model back into the host
 * mirrors (Q->current / Q->next, A.flags, A.cells, A.size_current).  Needed
 * only by code that reads those fields directly while a device run is in
 * progress; the accessor and IO functions do it themselves. */
int calhipSyncHost2D(struct CALModel2D* ca2D);

/* Wait for the model's queued GPU work. */
int calhipSynchronize2D(struct CALModel2D* ca2D);

/* 1 if the model's state currently lives on the GPU. */
int calhipIsOnDevice2D(struct CALModel2D* ca2D);

/* The step loop of calRun2D (cal2DRun.c:129-136) for n steps starting at
 * simulation->step: each step as calRunCAStep2D does it; afterwards
 * simulation->step is the number of the next step (the stopping step when
 * the stop condition held).  When no host callback runs between steps (no
 * user global transition, no stop condition, a steering that is absent or
 * has a device twin) the device program may batch the steps.  The model
 * stays on the GPU.  Returns the number of steps run, -1 on error.  Has no
 * counterpart in OpenCAL: it is calRun2D without its init, finalize and
 * read-back, for callers that keep a model resident between runs. */
struct CALRun2D;
int calhipRunSteps2D(struct CALRun2D* simulation, int n);

/* Counters since model creation. */
struct calhipStats {
	long long steps;           /* device program steps */
	long long launches;        /* kernel launches */
	long long uploads;         /* host -> device model transfers */
	long long downloads;       /* device -> host model transfers */
	long long rebuilds;        /* active-set compactions */
	long long materialisations;/* transient substates written out on demand */
	long long reruns;          /* steps run again through another path (a fused step that met a case it
	                              leaves to the general kernels) */
};
int calhipGetStats2D(struct CALModel2D* ca2D, struct calhipStats* out);

/* Kernel timing with HIP events recorded on the model's own stream: enable=1
 * around each launch of the program's main sweep kernel, enable=2 once when
 * timing starts and once at each query (the whole region, launches back to
 * back, over the engine's launches in it); 0 stops.  calhipGetKernelTime2D
 * returns the milliseconds and the launches. */
int calhipEnableKernelTiming2D(struct CALModel2D* ca2D, int enable);
int calhipGetKernelTime2D(struct CALModel2D* ca2D, double* total_ms, long long* launches);

/* ------------------------------------------------- generic device programs */

/* A model's own elementary processes, compiled for the GPU at run time.
 * The source is C written against OpenCAL's cell API (calGet2D*, calGetX2D*,
 * calSet2D*, calSetCurrent2D*, calInit2D*, the active-cell calls and the
 * unsafe operations of cal2DUnsafe.h), the same file a CPU build of the
 * model compiles, unchanged: its functions are compiled for both sides and
 * every file-scope variable also exists on the device.  The variables its
 * processes read (substate pointers, parameters) are mirrored from the host
 * with calhipBindGlobal2D.  A transition runs the processes in registration
 * order, each followed by OpenCAL's update (cal2D.c:738-776).  Reads of a
 * neighbour outside the grid's storage (flat boundary) give 0 and writes
 * there are dropped.  CALHIP_GLOBAL marks a variable for the device
 * explicitly; it is empty for host builds and optional. */
#ifndef CALHIP_GLOBAL
#define CALHIP_GLOBAL
#endif

/* Append C source to the model's process source (compiled on the first
 * transition; one compilation per distinct source for the process). */
int calhipAddProcessSource2D(struct CALModel2D* ca2D, const char* source);

/* Run `process` (registered with calAddElementaryProcess2D) as the device
 * function `device_name` of the source.  The model is stepped on the GPU
 * once every registered process is bound. */
int calhipBindProcess2D(struct CALModel2D* ca2D, CALCallbackFunc2D process, const char* device_name);

/* Compile process source for the GPU without running it (no GPU is
 * needed): 0 when it compiles with a kernel for each of the named device
 * functions; otherwise -1 and the compiler's log in log (log_bytes long). */
int calhipCheckProcessSource2D(const char* source, const char* const* names, int n_names, char* log,
                               size_t log_bytes);

/* 1 when the generic device program runs the process `device_name` of
 * `source` in colour phases, 0 when it runs every cell at once, -1 on a bad
 * argument (no GPU is needed; 2D and 3D sources alike).  A process runs in
 * phases when it, or a function of the source it calls, uses an operation
 * that reaches another cell's values: calSetX*, calSetCurrentX*, calInitX*,
 * calAddNextX*, calGetNextX* (OpenCAL-OMP/source/cal2DUnsafe.c,
 * cal3DUnsafe.c).  For a neighbourhood of radius r each axis is split into
 * colour classes of period 2r + 1 and the process is launched once per class
 * tuple, in row-major order of the tuples, so that no two cells that run
 * together reach a common cell: the result is deterministic and race-free;
 * it is the serial engine's visit reordered by phase, so sums that a cell
 * receives from several neighbours may round differently from the serial
 * row-major order.  No counterpart in the reference (its OpenMP engine
 * locks each access instead). */
int calhipProcessRunsInPhases(const char* source, const char* device_name);

/* Mirror the host variable at host_address (bytes long) into the
 * file-scope variable `name` of the source before each transition.  Words that
 * hold the address of one of this model's substates, or of the model, are
 * given the address of its device view. */
int calhipBindGlobal2D(struct CALModel2D* ca2D, const char* name, const void* host_address, size_t bytes);

/* ------------------------------------------------------- device programs */

/* Flow substates of SciddicaT are transient when the model's steering resets
 * them to zero every step (as OpenCAL's sciddicaT.c:133-146 does): the fused
 * program then never writes them to memory.  A steering that does not reset
 * them, or any read of them between steps, makes the program write them out
 * on demand. */

/* Attach the SciddicaT device program to a model built as the reference
 * example builds it (sciddicaT.c:143-184): von Neumann, toroidal, substates
 * z, h, f[4], processes `flows`, `width` and (active-cells variant) `remove`
 * registered in that order.  p_r / p_epsilon point to the model's parameter
 * variables and are read at every step.  remove may be NULL for CAL_NO_OPT. */
int calhipBindSciddicaT2D(struct CALModel2D* ca2D,
                          CALCallbackFunc2D flows, CALCallbackFunc2D width, CALCallbackFunc2D remove,
                          struct CALSubstate2Dr* z, struct CALSubstate2Dr* h,
                          struct CALSubstate2Dr* f[4],
                          const CALreal* p_r, const CALreal* p_epsilon);

/* Declare the model's steering callback to be the reference SciddicaT
 * steering (sciddicaT.c:133-146: every flow substate set to 0 with
 * calInitSubstate2Dr; the active-cells variant, cal_sciddicaT-activecells-glut
 * sciddicaT.c:137-144, with calInitSubstateNext2Dr).  The run loop then runs
 * its device twin instead of the callback, and calRun2D / calhipRunSteps2D
 * may advance the full-grid model two steps per sweep (the flows are never
 * stored).  The callback must do exactly that reset. */
int calhipBindSciddicaTSteering2D(struct CALModel2D* ca2D, void (*steering)(struct CALModel2D*));

/* Device check of the SciddicaT kernels' division of the level sum by the
 * count of remaining levels (sciddicaT.c:72-73), which they compute with a
 * multiply and two FMAs instead of the IEEE division sequence: for n seeded
 * doubles and every count 0..5, the number of results whose bits differ from
 * the IEEE quotient (0 expected), or -1 on error. */
long long calhipSciddicaTCheckDivision(long long n, unsigned long long seed);

/* Attach the Life device program (one process, Moore, toroidal) to a byte or
 * int substate.  Exactly one of Qb / Qi is non-NULL. */
int calhipBindLife2D(struct CALModel2D* ca2D, CALCallbackFunc2D transition,
                     struct CALSubstate2Db* Qb, struct CALSubstate2Di* Qi);

/* --------------------------------------------------------- SCIARA-fv2 */

/* What the SCIARA-fv2 device program needs from a model built as the
 * reference example builds it (OpenCAL-OMP/examples/calomp_sciara-fv2/
 * source/Sciara.cpp:287-356: Moore, CAL_SPACE_FLAT; double-layer Sz, Slt, St
 * and f[0..7]; single-layer Mb and Msl; processes vents, flows, width,
 * temperature and, with active cells, remove, registered in that order).
 * The parameter pointers point to the model's own variables and are read at
 * every step (the init callback may still compute a, b, c, d).  `steering`
 * is the model's steering callback (Sciara.cpp:185-197): when calRun2D or
 * calRunCAStep2D would call it, the engine runs its device twin instead,
 * which also advances *elapsed_time by *Pclock on the host.  Vents and their
 * emission-rate table are copied at bind time.
 * Out-of-grid neighbours (the reference reads outside its buffers there,
 * cal2D.c:859-870) read as altitude 0, thickness 0, temperature 0 and flow 0. */
struct calhipSciaraBinding {
	struct CALSubstate2Dr *Sz, *Slt, *St;
	struct CALSubstate2Dr* f[8];
	struct CALSubstate2Db* Mb;
	struct CALSubstate2Dr* Msl;
	CALCallbackFunc2D vents, flows, width, temperature, remove; /* remove NULL for CAL_NO_OPT */
	void (*steering)(struct CALModel2D*);
	const CALreal *Pclock, *PTsol, *PTvent, *Pc, *Pac, *a, *b, *c, *d, *rad2;
	const CALreal *Pcool, *Prho, *Pepsilon, *Psigma, *Pcv;
	CALreal* elapsed_time;
	const unsigned int* emission_time;
	int n_vents;
	const int* vent_rc;   /* n_vents (row, column) pairs */
	int n_rates;
	const CALreal* rates; /* n_vents x n_rates emission rates [m^3/s], row-major */
};

/* Attach the SCIARA-fv2 device program (replaces Sciara.cpp:3-197). */
int calhipBindSciara2D(struct CALModel2D* ca2D, const struct calhipSciaraBinding* b);

/* --------------------------------------------------- batched SciddicaT */

/* Run `steps` full-grid SciddicaT steps (von Neumann, toroidal, flows reset
 * every step) on `n` independent instances of uneven size and parameters;
 * each step of the whole ensemble is one launch over a flat tile list.
 *   rows[k], cols[k]   grid of instance k (each >= 3)
 *   p_r[k], p_eps[k]   parameters of instance k
 *   z_host[k], h_host[k]  host arrays (rows[k]*cols[k]) with altitude and
 *                      thickness before the model's init process (the init
 *                      z -= h where h > 0 of sciddicaT.c:104-130 is applied)
 *   h_out[k]           host arrays receiving the final thickness
 * timing_ms (optional): GPU time of the stepping launches only (inputs
 * resident).  Each instance's result equals its own single-model run. */
int calhipSciddicaTBatchRun(int n, const int* rows, const int* cols,
                            const double* p_r, const double* p_eps,
                            const double* const* z_host, const double* const* h_host,
                            int steps, double* const* h_out, double* timing_ms);

/* --------------------------------------------------- SciddicaT adjoint */

/* Device-pointer entry points (pointers from hipMalloc or a framework's
 * device allocator), used by the torch.autograd.Function wrapper.
 * All grids are rows x cols row-major doubles, toroidal.
 *
 * Forward: h_out = S^steps(h0) with checkpoints written every `ck` steps into
 * ckpt (capacity (steps/ck + 1) * rows*cols).  Backward: given lam = dL/dh_T
 * (overwritten with dL/dh_0), accumulates dL/dz into lam_z and returns
 * dL/dr, dL/deps; work buffers are allocated internally. */
int calhipSciddicaTForwardDev(int rows, int cols, const double* d_z, const double* d_h0,
                              double p_r, double p_eps, int steps, int ck,
                              double* d_ckpt, double* d_h_out, double* timing_ms);
int calhipSciddicaTAdjointDev(int rows, int cols, const double* d_z, const double* d_ckpt,
                              double p_r, double p_eps, int steps, int ck,
                              double* d_lam, double* d_lam_z, double* d_r, double* d_eps,
                              double* timing_ms);

/* As calhipSciddicaTForwardDev, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in OpenCAL.  It calls no hipMalloc,
 * hipFree, stream or event creation: its one working grid and the sweep's
 * spare row live in a fixed device array of the library (32 MiB + 1 KiB per
 * GPU, held until the process ends), so grids are limited to 4194304 cells
 * (e.g. 2048 x 2048; larger grids return -1 with error 51).  The sweeps run
 * on the current device's null stream; the call returns when they are done.
 * d_h_out may equal d_h0; d_ckpt may be NULL; no timing is reported.  Every
 * call on a GPU shares that GPU's array: calls that overlap on one GPU (from
 * several host threads) race. */
int calhipSciddicaTForwardNoAllocDev(int rows, int cols, const double* d_z, const double* d_h0,
                                     double p_r, double p_eps, int steps, int ck,
                                     double* d_ckpt, double* d_h_out);

/* As calhipSciddicaTAdjointDev, with dL/dz accumulated only for the cells of
 * the box [box_r0, box_r0 + box_rows) x [box_c0, box_c0 + box_cols) (the
 * source patch, whose altitude the init lowered by the initial thickness). */
int calhipSciddicaTAdjointBoxDev(int rows, int cols, const double* d_z, const double* d_ckpt,
                                 double p_r, double p_eps, int steps, int ck,
                                 double* d_lam, double* d_lam_z, double* d_r, double* d_eps,
                                 int box_r0, int box_c0, int box_rows, int box_cols,
                                 double* timing_ms);

#ifdef __cplusplus
}
#endif
#endif
