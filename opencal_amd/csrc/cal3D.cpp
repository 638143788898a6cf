// OpenCAL 3D model API (include/cal3D.h, cal3DUnsafe.h) on the MI355X engine.
//
// A 3D model is a Model3X: the public CALModel3D followed by a ModelX that
// holds the engine state for the (slices*rows) x columns grid with the same
// linear cell index (k*rows + i)*columns + j.  Residency, device buffers,
// fills, updates, the active set and the device program are the 2D engine's;
// this file adds the 3D neighbourhoods, accessors and the host-phase
// semantics of OpenCAL-OMP/source/cal3D.c (citations per function).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "engine.hpp"

extern "C" {
#include "cal3DBuffer.h"
#include "cal3DUnsafe.h"
#include "calhip3D.h"
}

using namespace calhip;

static inline ModelX* core(CALModel3D* m) { return &X3(m)->core; }

// ------------------------------------------------------ neighbourhoods

// cal3D.c:29-52: centre, N, W, E, S of slice 0, then slice -1, slice +1.
static void def_von_neumann3(CALModel3D* m)
{
	static const int d[7][3] = {{0, 0, 0}, {-1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {1, 0, 0}, {0, 0, -1}, {0, 0, 1}};
	for (auto& c : d) calAddNeighbor3D(m, c[0], c[1], c[2]);
}

// cal3D.c:58-101: the 2D Moore order (centre, N, W, E, S, NW, SW, SE, NE) on
// slice 0, then on slice -1, then on slice +1.
static void def_moore3(CALModel3D* m)
{
	static const int d[9][2] = {{0, 0}, {-1, 0}, {0, -1}, {0, 1}, {1, 0}, {-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
	for (int k : {0, -1, 1})
		for (auto& c : d) calAddNeighbor3D(m, c[0], c[1], k);
}

// ------------------------------------------------------------ model

extern "C" struct CALModel3D* calCADef3D(int rows, int columns, int slices, enum CALNeighborhood3D nb,
                                         enum CALSpaceBoundaryCondition tor, enum CALOptimization opt)
{
	if (rows <= 0 || columns <= 0 || slices <= 0 || (long long)rows * columns * slices > 0x7fffffffLL) {
		set_error(1, "calCADef3D: rows, columns and slices must be positive and the grid below 2^31 cells");
		return nullptr;
	}
	Model3X* x = new (std::nothrow) Model3X();
	if (!x) return nullptr;
	CALModel3D* m = &x->pub;
	memset(m, 0, sizeof(*m));
	m->rows = rows;
	m->columns = columns;
	m->slices = slices;
	m->T = tor;
	m->OPTIMIZATION = opt;
	if (opt == CAL_OPT_ACTIVE_CELLS) {
		m->A.flags = calAllocBuffer3Db(rows, columns, slices);
		if (!m->A.flags) {
			delete x;
			return nullptr;
		}
		memset(m->A.flags, CAL_FALSE, (size_t)rows * columns * slices);
	}
	m->A.num_threads = 1;   // one host thread issues all updates
	m->A.size_next = (int*)calloc(1, sizeof(int));
	m->X_id = nb;
	switch (nb) {
	case CAL_VON_NEUMANN_NEIGHBORHOOD_3D: def_von_neumann3(m); break;
	case CAL_MOORE_NEIGHBORHOOD_3D: def_moore3(m); break;
	default: break;
	}
	m->is_safe = CAL_UNSAFE_INACTIVE;
	m->locks = nullptr;

	// the engine's view: (slices*rows) x columns, sharing flags and deltas
	CALModel2D* c = &x->core.pub;
	memset(c, 0, sizeof(*c));
	c->rows = slices * rows;
	c->columns = columns;
	c->T = tor;
	c->OPTIMIZATION = opt;
	c->A.flags = m->A.flags;
	c->A.size_next = m->A.size_next;
	c->A.num_threads = 1;
	c->X_id = CAL_CUSTOM_NEIGHBORHOOD_2D;
	c->is_safe = CAL_UNSAFE_INACTIVE;
	x->core.pub3 = m;
	return m;
}

extern "C" void calSetUnsafe3D(struct CALModel3D* m) { m->is_safe = CAL_UNSAFE_ACTIVE; }

static inline size_t lin3(const CALModel3D* m, int i, int j, int k)
{
	return ((size_t)k * m->rows + i) * m->columns + j;
}

// Neighbour n of (i, j, k): toroidal single wrap per axis, or the raw
// coordinates for a flat space (cal3D.c:776-799).
static inline size_t nb3(const CALModel3D* m, int i, int j, int k, int n)
{
	int ii = i + m->X[n].i, jj = j + m->X[n].j, kk = k + m->X[n].k;
	if (m->T != CAL_SPACE_FLAT) {
		ii = calGetToroidalX(ii, m->rows);
		jj = calGetToroidalX(jj, m->columns);
		kk = calGetToroidalX(kk, m->slices);
	}
	return (size_t)(((long long)kk * m->rows + ii) * m->columns + jj);
}

// --------------------------------------------------------- active cells

// cal3D.c:267-279
extern "C" void calAddActiveCell3D(struct CALModel3D* m, int i, int j, int k)
{
	if (ensure_host(core(m))) return;
	CALbyte* f = &m->A.flags[lin3(m, i, j, k)];
	if (!*f) {
		*f = CAL_TRUE;
		m->A.size_next[0]++;
	}
}

// cal3D.c:281-302
extern "C" void calAddActiveCellX3D(struct CALModel3D* m, int i, int j, int k, int n)
{
	if (m->T == CAL_SPACE_FLAT)
		calAddActiveCell3D(m, i + m->X[n].i, j + m->X[n].j, k + m->X[n].k);
	else
		calAddActiveCell3D(m, calGetToroidalX(i + m->X[n].i, m->rows), calGetToroidalX(j + m->X[n].j, m->columns),
		                   calGetToroidalX(k + m->X[n].k, m->slices));
}

// cal3D.c:304-317
extern "C" void calRemoveActiveCell3D(struct CALModel3D* m, int i, int j, int k)
{
	if (ensure_host(core(m))) return;
	CALbyte* f = &m->A.flags[lin3(m, i, j, k)];
	if (*f) {
		*f = CAL_FALSE;
		m->A.size_next[0]--;
	}
}

// cal3D.c:319-378: size += deltas; empty -> NULL list; else the flags in
// (i, j, k) order.
static void host_rebuild_active3(CALModel3D* m)
{
	free(m->A.cells);
	m->A.cells = nullptr;
	int diff = 0;
	for (int t = 0; t < m->A.num_threads; ++t) {
		diff += m->A.size_next[t];
		m->A.size_next[t] = 0;
	}
	m->A.size_current += diff;
	if (m->A.size_current == 0) return;
	m->A.cells = (CALCell3D*)malloc(sizeof(CALCell3D) * (size_t)m->A.size_current);
	int n = 0;
	for (int i = 0; i < m->rows; ++i)
		for (int j = 0; j < m->columns; ++j)
			for (int k = 0; k < m->slices; ++k)
				if (m->A.flags[lin3(m, i, j, k)] && n < m->A.size_current) {
					m->A.cells[n].i = i;
					m->A.cells[n].j = j;
					m->A.cells[n].k = k;
					++n;
				}
}

extern "C" void calUpdateActiveCells3D(struct CALModel3D* m)
{
	ModelX* c = core(m);
	if (c->phase == Phase::Device) {
		device_rebuild_active(c);
		return;
	}
	host_rebuild_active3(m);
}

// cal3D.c:382-407
extern "C" struct CALCell3D* calAddNeighbor3D(struct CALModel3D* m, int i, int j, int k)
{
	CALCell3D* nx = (CALCell3D*)malloc(sizeof(CALCell3D) * (size_t)(m->sizeof_X + 1));
	if (!nx) return nullptr;
	for (int n = 0; n < m->sizeof_X; ++n) nx[n] = m->X[n];
	nx[m->sizeof_X].i = i;
	nx[m->sizeof_X].j = j;
	nx[m->sizeof_X].k = k;
	free(m->X);
	m->X = nx;
	m->sizeof_X++;
	return m->X;
}

// ------------------------------------------------------------ substates

template <typename Pub>
static void append(Pub*** arr, int* n, Pub* q)
{
	Pub** na = (Pub**)malloc(sizeof(Pub*) * (size_t)(*n + 1));
	for (int k = 0; k < *n; ++k) na[k] = (*arr)[k];
	na[*n] = q;
	free(*arr);
	*arr = na;
	(*n)++;
}

// cal3D.c:410-493 (double-layer, registered for updates)
#define ADD_SUBSTATE(SUF, TYPECH, ELEM)                                                         \
	extern "C" struct CALSubstate3D##SUF* calAddSubstate3D##SUF(struct CALModel3D* m)            \
	{                                                                                           \
		SubX* s = new_substate(core(m), TYPECH, ELEM, false);                                   \
		if (!s) return nullptr;                                                                 \
		s->registered = true;                                                                   \
		auto* q = reinterpret_cast<CALSubstate3D##SUF*>(&s->pub);                                \
		append(&m->pQ##SUF##_array, &m->sizeof_pQ##SUF##_array, q);                             \
		return q;                                                                               \
	}                                                                                           \
	extern "C" struct CALSubstate3D##SUF* calAddSingleLayerSubstate3D##SUF(struct CALModel3D* m) \
	{                                                                                           \
		SubX* s = new_substate(core(m), TYPECH, ELEM, true);                                    \
		return s ? reinterpret_cast<CALSubstate3D##SUF*>(&s->pub) : nullptr;                     \
	}

// cal3D.c:496-538 (single-layer: next == NULL, never updated)
ADD_SUBSTATE(b, 'b', 1)
ADD_SUBSTATE(i, 'i', 4)
ADD_SUBSTATE(r, 'r', 8)

// cal3D.c:540-562
extern "C" CALCallbackFunc3D* calAddElementaryProcess3D(struct CALModel3D* m, CALCallbackFunc3D ep)
{
	CALCallbackFunc3D* na =
	    (CALCallbackFunc3D*)malloc(sizeof(CALCallbackFunc3D) * (size_t)(m->num_of_elementary_processes + 1));
	if (!na) return nullptr;
	for (int n = 0; n < m->num_of_elementary_processes; ++n) na[n] = m->elementary_processes[n];
	na[m->num_of_elementary_processes] = ep;
	free(m->elementary_processes);
	m->elementary_processes = na;
	m->num_of_elementary_processes++;
	return m->elementary_processes;
}

// ------------------------------------------------- init / update (whole)

template <typename T>
static void host_set3(CALModel3D* m, T* M, T v)
{
	if (!M) return;
	if (m->A.cells) {
		for (int n = 0; n < m->A.size_current; ++n)
			M[lin3(m, m->A.cells[n].i, m->A.cells[n].j, m->A.cells[n].k)] = v;
	} else {
		const size_t N = (size_t)m->rows * m->columns * m->slices;
		for (size_t c = 0; c < N; ++c) M[c] = v;
	}
}

// Full-grid int fills are memset in the reference (cal3DBuffer.c:218-220):
// the value's low byte fills every byte of each cell.
template <typename T>
static T full_grid_value3(T v) { return v; }
template <>
CALint full_grid_value3<CALint>(CALint v)
{
	unsigned b = (unsigned)v & 0xffu;
	return (CALint)(b | (b << 8) | (b << 16) | (b << 24));
}

// cal3D.c:565-631
template <typename T>
static void init_sub3(CALModel3D* m, void* q, T v, bool both)
{
	ModelX* c = core(m);
	SubX* s = S(q);
	if (!m->A.cells && !(c->phase == Phase::Device && c->active())) v = full_grid_value3<T>(v);
	if (c->phase == Phase::Device) {
		device_fill(c, s, (double)v, both);
		return;
	}
	if (both) host_set3<T>(m, (T*)s->pub.r.current, v);
	host_set3<T>(m, (T*)s->pub.r.next, v);
}

extern "C" void calInitSubstate3Db(struct CALModel3D* m, struct CALSubstate3Db* Q, CALbyte v) { init_sub3<CALbyte>(m, Q, v, true); }
extern "C" void calInitSubstate3Di(struct CALModel3D* m, struct CALSubstate3Di* Q, CALint v) { init_sub3<CALint>(m, Q, v, true); }
extern "C" void calInitSubstate3Dr(struct CALModel3D* m, struct CALSubstate3Dr* Q, CALreal v) { init_sub3<CALreal>(m, Q, v, true); }
extern "C" void calInitSubstateNext3Db(struct CALModel3D* m, struct CALSubstate3Db* Q, CALbyte v) { init_sub3<CALbyte>(m, Q, v, false); }
extern "C" void calInitSubstateNext3Di(struct CALModel3D* m, struct CALSubstate3Di* Q, CALint v) { init_sub3<CALint>(m, Q, v, false); }
extern "C" void calInitSubstateNext3Dr(struct CALModel3D* m, struct CALSubstate3Dr* Q, CALreal v) { init_sub3<CALreal>(m, Q, v, false); }

// cal3D.c:635-654 / cal3DBuffer.c:43-163
static void host_update_sub3(CALModel3D* m, SubX* s)
{
	if (s->single) return;
	char* cur = (char*)s->pub.r.current;
	char* nxt = (char*)s->pub.r.next;
	if (m->A.cells) {
		for (int n = 0; n < m->A.size_current; ++n) {
			const size_t c = lin3(m, m->A.cells[n].i, m->A.cells[n].j, m->A.cells[n].k);
			memcpy(cur + c * s->elem, nxt + c * s->elem, (size_t)s->elem);
		}
	} else {
		memcpy(cur, nxt, (size_t)m->rows * m->columns * m->slices * s->elem);
	}
}

static void update_one3(CALModel3D* m, void* q)
{
	ModelX* c = core(m);
	SubX* s = S(q);
	if (c->phase == Phase::Device) {
		device_changed(c);
		device_update_sub(c, s);
		return;
	}
	host_update_sub3(m, s);
}

extern "C" void calUpdateSubstate3Db(struct CALModel3D* m, struct CALSubstate3Db* Q) { update_one3(m, Q); }
extern "C" void calUpdateSubstate3Di(struct CALModel3D* m, struct CALSubstate3Di* Q) { update_one3(m, Q); }
extern "C" void calUpdateSubstate3Dr(struct CALModel3D* m, struct CALSubstate3Dr* Q) { update_one3(m, Q); }

// cal3D.c:697-715
extern "C" void calUpdate3D(struct CALModel3D* m)
{
	ModelX* c = core(m);
	if (c->phase == Phase::Device) {
		device_update(c);
		return;
	}
	if (m->OPTIMIZATION == CAL_OPT_ACTIVE_CELLS) host_rebuild_active3(m);
	for (int k = 0; k < m->sizeof_pQb_array; ++k) host_update_sub3(m, S(m->pQb_array[k]));
	for (int k = 0; k < m->sizeof_pQi_array; ++k) host_update_sub3(m, S(m->pQi_array[k]));
	for (int k = 0; k < m->sizeof_pQr_array; ++k) host_update_sub3(m, S(m->pQr_array[k]));
}

// ------------------------------------------------ transition functions

static bool program_covers3(CALModel3D* m)
{
	ModelX* c = core(m);
	if (!c->prog) return false;
	if ((int)c->prog_eps3.size() != m->num_of_elementary_processes) return false;
	for (int k = 0; k < m->num_of_elementary_processes; ++k)
		if (c->prog_eps3[k] != m->elementary_processes[k]) return false;
	return true;
}

// cal3D.c:677-694
extern "C" void calGlobalTransitionFunction3D(struct CALModel3D* m)
{
	ModelX* c = core(m);
	if (!program_covers3(m)) {
		set_error(2, "calGlobalTransitionFunction3D: the model's elementary processes are not bound to "
		             "device source (calhipBindProcess3D); there is no CPU execution path");
		return;
	}
	if (ensure_device(c)) return;
	device_changed(c);
	if (c->prog->step(c) == 0) c->stats.steps++;
}

// cal3D.c:658-675
extern "C" void calApplyElementaryProcess3D(struct CALModel3D* m, CALCallbackFunc3D ep)
{
	ModelX* c = core(m);
	if (c->prog) {
		if (ensure_device(c)) return;
		device_changed(c);
		const int r = c->prog->apply(c, (const void*)ep);
		if (r <= 0) return;
	}
	set_error(3, "calApplyElementaryProcess3D: the process is not bound to device source (calhipBindProcess3D)");
}

// ------------------------------------------------------------ accessors

template <typename T>
static inline T* cur3(void* q) { return (T*)S(q)->pub.r.current; }
template <typename T>
static inline T* nxt3(void* q) { return (T*)S(q)->pub.r.next; }

// cal3D.c:718-857
#define CAL3_ACCESSORS(SUF, T)                                                                              \
	extern "C" void calInit3D##SUF(struct CALModel3D* m, struct CALSubstate3D##SUF* Q, int i, int j, int k, T v) \
	{                                                                                                        \
		if (ensure_host_for(core(m), Q)) return;                                                             \
		cur3<T>(Q)[lin3(m, i, j, k)] = v;                                                                    \
		nxt3<T>(Q)[lin3(m, i, j, k)] = v;                                                                    \
	}                                                                                                        \
	extern "C" T calGet3D##SUF(struct CALModel3D* m, struct CALSubstate3D##SUF* Q, int i, int j, int k)      \
	{                                                                                                        \
		if (host_read(core(m), Q)) return (T)0;                                                              \
		return cur3<T>(Q)[lin3(m, i, j, k)];                                                                 \
	}                                                                                                        \
	extern "C" T calGetX3D##SUF(struct CALModel3D* m, struct CALSubstate3D##SUF* Q, int i, int j, int k, int n) \
	{                                                                                                        \
		if (host_read(core(m), Q)) return (T)0;                                                              \
		return cur3<T>(Q)[nb3(m, i, j, k, n)];                                                               \
	}                                                                                                        \
	extern "C" void calSet3D##SUF(struct CALModel3D* m, struct CALSubstate3D##SUF* Q, int i, int j, int k, T v) \
	{                                                                                                        \
		if (ensure_host_for(core(m), Q)) return;                                                             \
		nxt3<T>(Q)[lin3(m, i, j, k)] = v;                                                                    \
	}                                                                                                        \
	extern "C" void calSetCurrent3D##SUF(struct CALModel3D* m, struct CALSubstate3D##SUF* Q, int i, int j, int k, \
	                                     T v)                                                                \
	{                                                                                                        \
		if (ensure_host_for(core(m), Q)) return;                                                             \
		cur3<T>(Q)[lin3(m, i, j, k)] = v;                                                                    \
	}

CAL3_ACCESSORS(b, CALbyte)
CAL3_ACCESSORS(i, CALint)
CAL3_ACCESSORS(r, CALreal)

// Unsafe operations (OpenCAL-OMP/source/cal3DUnsafe.c:18-244) on the host
// mirrors; the host phase is single-threaded, so the reference's cell locks
// have nothing to guard.  (On the GPU they are part of the generic
// program's device API.)
#define CAL3_UNSAFE(SUF, T)                                                                                 \
	extern "C" void calInitX3D##SUF(struct CALModel3D* m, struct CALSubstate3D##SUF* Q, int i, int j, int k, int n, \
	                                T v)                                                                     \
	{                                                                                                        \
		if (ensure_host_for(core(m), Q)) return;                                                             \
		const size_t c = nb3(m, i, j, k, n);                                                                 \
		cur3<T>(Q)[c] = v;                                                                                   \
		nxt3<T>(Q)[c] = v;                                                                                   \
	}                                                                                                        \
	extern "C" T calGetNext3D##SUF(struct CALModel3D* m, struct CALSubstate3D##SUF* Q, int i, int j, int k)  \
	{                                                                                                        \
		if (ensure_host_for(core(m), Q)) return (T)0;                                                        \
		return nxt3<T>(Q)[lin3(m, i, j, k)];                                                                 \
	}                                                                                                        \
	extern "C" T calGetNextX3D##SUF(struct CALModel3D* m, struct CALSubstate3D##SUF* Q, int i, int j, int k, int n) \
	{                                                                                                        \
		if (ensure_host_for(core(m), Q)) return (T)0;                                                        \
		return nxt3<T>(Q)[nb3(m, i, j, k, n)];                                                               \
	}                                                                                                        \
	extern "C" void calSetX3D##SUF(struct CALModel3D* m, struct CALSubstate3D##SUF* Q, int i, int j, int k, int n, \
	                               T v)                                                                      \
	{                                                                                                        \
		if (ensure_host_for(core(m), Q)) return;                                                             \
		nxt3<T>(Q)[nb3(m, i, j, k, n)] = v;                                                                  \
	}                                                                                                        \
	extern "C" void calSetCurrentX3D##SUF(struct CALModel3D* m, struct CALSubstate3D##SUF* Q, int i, int j, int k, \
	                                      int n, T v)                                                        \
	{                                                                                                        \
		if (ensure_host_for(core(m), Q)) return;                                                             \
		cur3<T>(Q)[nb3(m, i, j, k, n)] = v;                                                                  \
	}                                                                                                        \
	extern "C" void calAddNext3D##SUF(struct CALModel3D* m, struct CALSubstate3D##SUF* Q, int i, int j, int k, T v) \
	{                                                                                                        \
		if (ensure_host_for(core(m), Q)) return;                                                             \
		T* p = &nxt3<T>(Q)[lin3(m, i, j, k)];                                                                \
		*p = (T)(*p + v);                                                                                    \
	}                                                                                                        \
	extern "C" void calAddNextX3D##SUF(struct CALModel3D* m, struct CALSubstate3D##SUF* Q, int i, int j, int k, \
	                                   int n, T v)                                                           \
	{                                                                                                        \
		if (ensure_host_for(core(m), Q)) return;                                                             \
		T* p = &nxt3<T>(Q)[nb3(m, i, j, k, n)];                                                              \
		*p = (T)(*p + v);                                                                                    \
	}

CAL3_UNSAFE(b, CALbyte)
CAL3_UNSAFE(i, CALint)
CAL3_UNSAFE(r, CALreal)

// ------------------------------------------------------------- finalize

// cal3D.c:860-886 (also frees the substate records, which the reference leaks)
extern "C" void calFinalize3D(struct CALModel3D* m)
{
	if (!m) return;
	Model3X* x = X3(m);
	ModelX* c = &x->core;
	release_device(c);
	delete c->prog;
	c->prog = nullptr;
	free(m->A.flags);       // shared with the engine view
	free(m->A.size_next);   // shared with the engine view
	free(m->A.cells);
	free(c->pub.A.cells);
	free(m->X);
	for (SubX* s : c->subs) {
		free(s->pub.r.current);
		free(s->pub.r.next);
		delete s;
	}
	free(m->pQb_array);
	free(m->pQi_array);
	free(m->pQr_array);
	free(m->elementary_processes);
	delete x;
}

// ------------------------------------------------------- calhip3D.h

extern "C" int calhipSyncHost3D(struct CALModel3D* m) { return ensure_host(core(m)); }
extern "C" int calhipIsOnDevice3D(struct CALModel3D* m) { return core(m)->phase == Phase::Device; }
extern "C" int calhipGetStats3D(struct CALModel3D* m, struct calhipStats* out)
{
	*out = core(m)->stats;
	return 0;
}
extern "C" int calhipEnableKernelTiming3D(struct CALModel3D* m, int enable)
{
	return calhipEnableKernelTiming2D(&core(m)->pub, enable);
}
extern "C" int calhipGetKernelTime3D(struct CALModel3D* m, double* total_ms, long long* launches)
{
	return calhipGetKernelTime2D(&core(m)->pub, total_ms, launches);
}
