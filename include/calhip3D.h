/* MI355X engine extensions to the OpenCAL 3D API.
 *
 * These entry points have no counterpart in OpenCAL.  A 3D model's
 * elementary processes run on the GPU through the generic device program:
 * the model hands over the C source of its processes, which is compiled for
 * gfx950 at run time together with the device side of OpenCAL's 3D cell API
 * (calGet3D*, calGetX3D*, calSet3D*, calSetCurrent3D*, calInit3D*, the
 * active-cell calls and the unsafe operations of cal3DUnsafe.h).  Errors are
 * reported through calhipGetLastError() (calhip2D.h).  Plain C ABI.
 */
#ifndef calhip3D_h
#define calhip3D_h

#include <stddef.h>
#include <cal3D.h>
#include <calhip2D.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Bring every substate and the active set of the model back into the host
 * mirrors (Q->current / Q->next, A.flags, A.cells, A.size_current); needed
 * only by code that reads those fields directly during a device run. */
int calhipSyncHost3D(struct CALModel3D* ca3D);

/* 1 if the model's state currently lives on the GPU. */
int calhipIsOnDevice3D(struct CALModel3D* ca3D);

/* Counters since model creation (struct calhipStats of calhip2D.h). */
int calhipGetStats3D(struct CALModel3D* ca3D, struct calhipStats* out);

/* Kernel timing with HIP events around each process launch. */
int calhipEnableKernelTiming3D(struct CALModel3D* ca3D, int enable);
int calhipGetKernelTime3D(struct CALModel3D* ca3D, double* total_ms, long long* launches);

/* Append C source to the model's process source (compiled on the first
 * transition, once per distinct source in a process).  The variables its
 * processes read are declared CALHIP_GLOBAL (empty for host builds). */
int calhipAddProcessSource3D(struct CALModel3D* ca3D, const char* source);

/* Run `process` (registered with calAddElementaryProcess3D) as the device
 * function `device_name` of the source.  A transition runs the processes in
 * registration order, each followed by OpenCAL's update (cal3D.c:677-715). */
int calhipBindProcess3D(struct CALModel3D* ca3D, CALCallbackFunc3D process, const char* device_name);

/* Mirror the host variable at host_address (bytes long) into the
 * CALHIP_GLOBAL `name` of the source before each transition.  Words that
 * hold the address of one of this model's substates, or of the model, are
 * given the address of its device view. */
int calhipBindGlobal3D(struct CALModel3D* ca3D, const char* name, const void* host_address, size_t bytes);

/* Compile 3D process source without running it (no GPU needed): 0 when it
 * compiles with a kernel for each named device function, else -1 and the
 * compiler's log. */
int calhipCheckProcessSource3D(const char* source, const char* const* names, int n_names, char* log,
                               size_t log_bytes);

#ifdef __cplusplus
}
#endif
#endif
