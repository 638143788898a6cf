// Private state of the MI355X OpenCAL engine.
//
// A model is a ModelX whose first member is the public CALModel2D, and every
// substate is a SubX whose first member is the public CALSubstate2D{b,i,r};
// the public pointers handed to user code are the addresses of those first
// members.
//
// Residency.  A model is either in the HOST phase (the host mirrors
// Q->current / Q->next / A.* are authoritative and every API call runs the
// reference semantics on them) or in the DEVICE phase (the device buffers are
// authoritative).  The global transition of a bound device program moves the
// model to the DEVICE phase; whole-substate operations (init, update, active
// set rebuild) stay on the device; per-cell host accessors bring the model
// back to the HOST phase first.
//
// Device layers.  Each double-layer substate has two device buffers.  After a
// program step or a calUpdate2D the next layer is logically equal to current
// (OpenCAL copies next->current after every process) and is physically stale:
// `nxt_stale`.  A partial write to a stale next layer first copies current
// into it.  `nxt_dirty` marks a next layer written since the last update; the
// update then flips the two buffers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

extern "C" {
#include "cal2D.h"
#include "cal2DRun.h"
#include "cal3D.h"
#include "calhip2D.h"
}

namespace calhip {

enum class Phase { Host, Device };

// Transient substate states (flows of a fused program, see calhip2D.h).
enum class Transient {
	None,          // ordinary substate
	Zero,          // both layers logically 0 everywhere; buffers not written
	Derived,       // logically the flows of the last fused step (not stored)
	PendingZero,   // next layer reset to 0 by the steering, awaiting the update
	ZeroAtActive,  // physical buffer valid except: logically 0 at the active cells
	DerivedActive, // physical buffer holds the last flows (aliased layers)
	PendingActive  // next layer reset at the active cells, awaiting the update
};

struct ModelX;
struct SubX;

// A device program replaces the model's elementary processes (and may own
// transient substates).  Implementations live next to their kernels.
struct Program {
	virtual ~Program() {}
	virtual const char* name() const = 0;
	// calGlobalTransitionFunction2D on the device.
	virtual int step(ModelX* m) = 0;
	// n steps with no host callback in between; with_steering: each step is
	// followed by the steering twin and the update (calRunCAStep2D with a
	// twin steering, implicit update).  Default: that loop.
	virtual int run_steps(ModelX* m, int n, bool with_steering) {
		for (int k = 0; k < n; ++k) {
			if (int e = with_steering ? step_and_steer(m) : step(m)) return e;
			if (with_steering && device_update_hook(m)) return -1;
		}
		return 0;
	}
	// calUpdate2D on the device (runtime.cpp's device_update)
	static int device_update_hook(ModelX* m);
	// calInitSubstate(Next)2Dr on a substate the program marked transient.
	// Returns 1 when handled, 0 to fall back to an ordinary device fill (after
	// materialising), <0 on error.
	virtual int transient_init(ModelX*, SubX*, double, bool /*both*/) { return 0; }
	// calUpdate2D hook for transient substates.
	virtual void transient_update(ModelX*) {}
	// Write a transient substate's logical content into its device buffers.
	virtual int materialise(ModelX*, SubX*) { return -1; }
	// A non-program operation is about to modify `s` on the device.
	virtual int before_modify(ModelX*, SubX*) { return 0; }
	// The model has just been uploaded (host mirrors are the logical state).
	virtual int after_upload(ModelX*) { return 0; }
	// Device twin of the model's steering callback (ModelX::prog_steering).
	virtual int steer(ModelX*) { return -1; }
	// One transition followed by the steering twin; programs may fuse them.
	virtual int step_and_steer(ModelX* m)
	{
		if (int e = step(m)) return e;
		return steer(m);
	}
	// Called before the device active set is read back: restores the
	// row-major order of the list when the program keeps it unordered.
	virtual int sync_active(ModelX*) { return 0; }
	// Called before a substate's current device layer is read as it is (a
	// host read of one layer, a reduction): finishes deferred writes to it.
	virtual int before_read(ModelX*) { return 0; }
	// Launches every step the program has accepted but not yet launched
	// (before a stream synchronisation, timing or statistics query).
	virtual int flush(ModelX*) { return 0; }
	// calApplyElementaryProcess2D/3D: one process (a CALCallbackFunc2D or
	// CALCallbackFunc3D) over the grid (or the active list) without the
	// update; 1 when the program offers no single-process sweep, <0 on error.
	virtual int apply(ModelX*, const void* /*process*/) { return 1; }
};

struct SubX {
	union {
		CALSubstate2Db b;
		CALSubstate2Di i;
		CALSubstate2Dr r;
	} pub;
	char type = 'r';   // 'b', 'i' or 'r'
	int elem = 8;      // bytes per cell
	bool single = false;
	bool registered = false;  // listed in pQ*_array (double-layer)
	void* d_cur = nullptr;
	void* d_nxt = nullptr;
	bool nxt_stale = false;
	bool nxt_dirty = false;
	Transient tstate = Transient::None;
	// Never moved to the device: a single-layer substate the bound program
	// does not use.  Its host array stays authoritative in every phase, so
	// user code may even replace or free it as OpenCAL code does.
	bool host_only = false;
	// DEVICE phase: the host `current` mirror equals the device layer as of
	// ModelX::dev_epoch == host_epoch (read-only host accessors then need no
	// transfer; see host_read).
	unsigned long long host_epoch = 0;
	struct ModelX* owner = nullptr;

	void* h_cur() const { return (void*)pub.r.current; }
	void* h_nxt() const { return (void*)pub.r.next; }
};

struct Timer {
	bool enabled = false;
	double total_ms = 0.0;
	long long launches = 0;
	std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
	// region mode: one event when timing starts, one at each query
	bool region = false;
	hipEvent_t region_start = nullptr;
	long long region_launches0 = 0;
};

struct ModelX {
	CALModel2D pub;
	std::vector<SubX*> subs;   // every substate, creation order
	Phase phase = Phase::Host;
	hipStream_t stream = nullptr;
	int device = 0;

	// device active set: flags (rows*cols bytes), row-major list of linear
	// indices, count[0] = list length (0 <=> OpenCAL's A.cells == NULL).
	uint8_t* d_flags = nullptr;
	int* d_cells = nullptr;
	int* d_count = nullptr;
	void* d_block_counts = nullptr;   // rebuild: chunk status words and counter (util.hip)
	int n_count_blocks = 0;
	unsigned rebuild_epoch = 0;
	bool flags_dirty = false;  // flags changed since the last rebuild

	Program* prog = nullptr;
	std::vector<CALCallbackFunc2D> prog_eps;  // processes the program covers
	void (*prog_steering)(CALModel2D*) = nullptr;  // steering with a device twin

	calhipStats stats{};
	Timer timer;

	// reductions: block partials, and a staging copy of a host array
	void* d_red = nullptr;
	void* d_stage = nullptr;
	size_t stage_bytes = 0;

	// Incremented by every device operation that may change a substate while
	// the model is in the DEVICE phase (see SubX::host_epoch).
	unsigned long long dev_epoch = 1;

	// A 3D model keeps its engine state in a ModelX whose private 2D shape is
	// (slices*rows) x columns: the same linear index k*rows*columns +
	// i*columns + j as OpenCAL's 3D buffers (cal3DBuffer.h:131-136).  pub3 is
	// then the public 3D model; its active list (CALCell3D, (i,j,k) order) is
	// converted at every upload/download.
	CALModel3D* pub3 = nullptr;
	std::vector<CALCallbackFunc3D> prog_eps3;   // 3D processes the program covers

	size_t cells() const { return (size_t)pub.rows * (size_t)pub.columns; }
	bool active() const { return pub.OPTIMIZATION == CAL_OPT_ACTIVE_CELLS; }
};

inline ModelX* X(CALModel2D* m) { return reinterpret_cast<ModelX*>(m); }

// A 3D model: the public struct, then the engine state (see ModelX::pub3).
struct Model3X {
	CALModel3D pub;
	ModelX core;
};
inline Model3X* X3(CALModel3D* m) { return reinterpret_cast<Model3X*>(m); }
inline SubX* S(void* q) { return reinterpret_cast<SubX*>(q); }

// ---- error handling (never exits the process) ----
void set_error(int code, const std::string& msg);
int check_hip(hipError_t e, const char* what);
#define CALHIP_CHECK(call) ::calhip::check_hip((call), #call)

// ---- residency ----
int ensure_stream(ModelX* m);  // the model's stream (created on first use)
int ensure_host(ModelX* m);    // DEVICE -> HOST (no-op in HOST)
// Read-only host access to `current` of q: in the DEVICE phase only that
// layer is copied back (once per device epoch) and the model stays there.
int host_read(ModelX* m, void* q);
inline void device_changed(ModelX* m) { m->dev_epoch++; }
// ensure_host, unless the call touches only the host-only substate q
inline int ensure_host_for(ModelX* m, void* q) { return S(q)->host_only ? 0 : ensure_host(m); }
int ensure_device(ModelX* m);  // HOST -> DEVICE (no-op in DEVICE)
int ensure_device_buffers(ModelX* m, SubX* s);
int materialise_next(ModelX* m, SubX* s);  // stale next := current
int device_update_sub(ModelX* m, SubX* s);  // calUpdateSubstate of one written layer
int device_fill(ModelX* m, SubX* s, double v, bool both);
int device_update(ModelX* m);
int device_rebuild_active(ModelX* m);
// The bound program runs every elementary process of the 2D model, in order.
bool program_covers_model(ModelX* m);
void flip(SubX* s);

// calGlobalTransitionFunction2D on the device, optionally followed by the
// program's steering twin (cal2D.cpp); 0 on success.
int device_transition(ModelX* m, bool with_steering);

// ---- shared by the 2D and 3D APIs (cal2D.cpp) ----
SubX* new_substate(ModelX* x, char type, int elem, bool single);
void release_device(ModelX* x);   // device buffers, stream, timers

// ---- kernel launchers (defined in the .hip files) ----
namespace k {
int fill(ModelX* m, void* buf, int elem, double v, bool active_only);
int copy(ModelX* m, void* dst, const void* src, size_t bytes);
// next -> current at the active list's cells (the whole grid when it is empty)
int copy_active(ModelX* m, void* dst, const void* src, int elem);
int rebuild_active(ModelX* m);
int active_count_host(ModelX* m, int* out);  // synchronous read of count[0]
}  // namespace k

// kernel timing helpers
void timer_begin(ModelX* m, hipEvent_t* ev0);
void timer_end(ModelX* m, hipEvent_t ev0);
void timer_collect(ModelX* m);

}  // namespace calhip
