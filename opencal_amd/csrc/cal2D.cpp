// OpenCAL 2D model API (include/cal2D.h) on the MI355X engine.
//
// HOST phase: each function applies OpenCAL-OMP's semantics to the host
// mirrors (citations per function).  DEVICE phase: whole-substate operations
// run on the GPU, per-cell accessors first bring the model back to the host.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>

#include "engine.hpp"

extern "C" {
#include "cal2DUnsafe.h"
}

extern "C" {
#include "cal2DBuffer.h"
}

using namespace calhip;

// ------------------------------------------------------------------ errors

namespace calhip {
static std::mutex g_err_mu;
static int g_err_code = 0;
static std::string g_err_msg;

void set_error(int code, const std::string& msg)
{
	std::lock_guard<std::mutex> lk(g_err_mu);
	if (g_err_code == 0) {
		g_err_code = code;
		g_err_msg = msg;
	}
	fprintf(stderr, "[opencal-hip] error %d: %s\n", code, msg.c_str());
}

int check_hip(hipError_t e, const char* what)
{
	if (e == hipSuccess) return 0;
	set_error(1000 + (int)e, std::string(what) + ": " + hipGetErrorString(e));
	return -1;
}
}  // namespace calhip

extern "C" int calhipGetLastError(void)
{
	std::lock_guard<std::mutex> lk(g_err_mu);
	return g_err_code;
}

extern "C" const char* calhipGetErrorString(void)
{
	std::lock_guard<std::mutex> lk(g_err_mu);
	return g_err_code ? g_err_msg.c_str() : "no error";
}

extern "C" void calhipClearError(void)
{
	std::lock_guard<std::mutex> lk(g_err_mu);
	g_err_code = 0;
	g_err_msg.clear();
}

// ------------------------------------------------------ neighbourhoods

static void add_nb(CALModel2D* m, int i, int j) { calAddNeighbor2D(m, i, j); }

// cal2D.c:28-44
static void def_von_neumann(CALModel2D* m)
{
	add_nb(m, 0, 0);
	add_nb(m, -1, 0);
	add_nb(m, 0, -1);
	add_nb(m, 0, 1);
	add_nb(m, 1, 0);
}

// cal2D.c:50-70
static void def_moore(CALModel2D* m)
{
	def_von_neumann(m);
	add_nb(m, -1, -1);
	add_nb(m, 1, -1);
	add_nb(m, 1, 1);
	add_nb(m, -1, 1);
}

// cal2D.c:75-117: even-column table then odd-column table, sizeof_X = 7.
static void def_hex(CALModel2D* m)
{
	static const int even[7][2] = {{0, 0}, {-1, 1}, {-1, 0}, {-1, -1}, {0, -1}, {1, 0}, {0, 1}};
	static const int odd[7][2] = {{0, 0}, {0, 1}, {-1, 0}, {0, -1}, {1, -1}, {1, 0}, {1, 1}};
	for (auto& d : even) add_nb(m, d[0], d[1]);
	for (auto& d : odd) add_nb(m, d[0], d[1]);
	m->sizeof_X = 7;
}

// cal2D.c:121-166: even-row table then odd-row table, sizeof_X = 7.
static void def_hex_alt(CALModel2D* m)
{
	static const int even[7][2] = {{0, 0}, {-1, 0}, {-1, -1}, {0, -1}, {1, -1}, {1, 0}, {0, 1}};
	static const int odd[7][2] = {{0, 0}, {-1, 1}, {-1, 0}, {0, -1}, {1, 0}, {1, 1}, {0, 1}};
	for (auto& d : even) add_nb(m, d[0], d[1]);
	for (auto& d : odd) add_nb(m, d[0], d[1]);
	m->sizeof_X = 7;
}

// ------------------------------------------------------------ model

extern "C" struct CALModel2D* calCADef2D(int rows, int columns, enum CALNeighborhood2D nb,
                                         enum CALSpaceBoundaryCondition tor,
                                         enum CALOptimization opt)
{
	if (rows <= 0 || columns <= 0) {
		set_error(1, "calCADef2D: rows and columns must be positive");
		return nullptr;
	}
	ModelX* x = new (std::nothrow) ModelX();
	if (!x) return nullptr;
	CALModel2D* m = &x->pub;
	memset(m, 0, sizeof(*m));
	m->rows = rows;
	m->columns = columns;
	m->T = tor;
	m->OPTIMIZATION = opt;
	if (opt == CAL_OPT_ACTIVE_CELLS) {
		m->A.flags = calAllocBuffer2Db(rows, columns);
		if (!m->A.flags) {
			delete x;
			return nullptr;
		}
		calSetBuffer2Db(m->A.flags, rows, columns, CAL_FALSE);
	}
	m->A.num_threads = 1;   // one host thread issues all updates
	m->A.size_next = (int*)calloc(1, sizeof(int));
	m->A.cells = nullptr;
	m->A.size_current = 0;
	m->X = nullptr;
	m->sizeof_X = 0;
	m->X_id = nb;
	switch (nb) {
	case CAL_VON_NEUMANN_NEIGHBORHOOD_2D: def_von_neumann(m); break;
	case CAL_MOORE_NEIGHBORHOOD_2D: def_moore(m); break;
	case CAL_HEXAGONAL_NEIGHBORHOOD_2D: def_hex(m); break;
	case CAL_HEXAGONAL_NEIGHBORHOOD_ALT_2D: def_hex_alt(m); break;
	default: break;
	}
	m->is_safe = CAL_UNSAFE_INACTIVE;
	m->locks = nullptr;
	return m;
}

extern "C" void calSetUnsafe2D(struct CALModel2D* m) { m->is_safe = CAL_UNSAFE_ACTIVE; }

// --------------------------------------------------------- active cells

// cal2D.c:333-349
extern "C" void calAddActiveCell2D(struct CALModel2D* m, int i, int j)
{
	if (ensure_host(X(m))) return;
	CALbyte* f = &m->A.flags[(size_t)i * m->columns + j];
	if (!*f) {
		*f = CAL_TRUE;
		m->A.size_next[0]++;
	}
}

// cal2D.c:351-363
extern "C" void calAddActiveCellX2D(struct CALModel2D* m, int i, int j, int n)
{
	if ((m->X_id == CAL_HEXAGONAL_NEIGHBORHOOD_2D && j % 2 == 1) ||
	    (m->X_id == CAL_HEXAGONAL_NEIGHBORHOOD_ALT_2D && i % 2 == 1))
		n += CAL_HEXAGONAL_SHIFT;
	if (m->T == CAL_SPACE_FLAT)
		calAddActiveCell2D(m, i + m->X[n].i, j + m->X[n].j);
	else
		calAddActiveCell2D(m, calGetToroidalX(i + m->X[n].i, m->rows),
		                   calGetToroidalX(j + m->X[n].j, m->columns));
}

// cal2D.c:365-380
extern "C" void calRemoveActiveCell2D(struct CALModel2D* m, int i, int j)
{
	if (ensure_host(X(m))) return;
	CALbyte* f = &m->A.flags[(size_t)i * m->columns + j];
	if (*f) {
		*f = CAL_FALSE;
		m->A.size_next[0]--;
	}
}

// cal2D.c:382-440: size += deltas; empty -> NULL list; else row-major scan.
static void host_rebuild_active(CALModel2D* m)
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
	m->A.cells = (CALCell2D*)malloc(sizeof(CALCell2D) * (size_t)m->A.size_current);
	int n = 0;
	for (int i = 0; i < m->rows; ++i)
		for (int j = 0; j < m->columns; ++j)
			if (m->A.flags[(size_t)i * m->columns + j] && n < m->A.size_current) {
				m->A.cells[n].i = i;
				m->A.cells[n].j = j;
				++n;
			}
}

extern "C" void calUpdateActiveCells2D(struct CALModel2D* m)
{
	ModelX* x = X(m);
	if (x->phase == Phase::Device) {
		device_rebuild_active(x);
		return;
	}
	host_rebuild_active(m);
}

// cal2D.c:444-466
extern "C" struct CALCell2D* calAddNeighbor2D(struct CALModel2D* m, int i, int j)
{
	CALCell2D* nx = (CALCell2D*)malloc(sizeof(CALCell2D) * (size_t)(m->sizeof_X + 1));
	if (!nx) return nullptr;
	for (int n = 0; n < m->sizeof_X; ++n) nx[n] = m->X[n];
	nx[m->sizeof_X].i = i;
	nx[m->sizeof_X].j = j;
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

namespace calhip {
SubX* new_substate(ModelX* x, char type, int elem, bool single)
{
	if (x->phase == Phase::Device && ensure_host(x)) return nullptr;
	SubX* s = new (std::nothrow) SubX();
	if (!s) return nullptr;
	s->type = type;
	s->elem = elem;
	s->single = single;
	size_t bytes = x->cells() * (size_t)elem;
	s->pub.r.current = (CALreal*)malloc(bytes);
	s->pub.r.next = single ? nullptr : (CALreal*)malloc(bytes);
	if (!s->pub.r.current || (!single && !s->pub.r.next)) {
		free(s->pub.r.current);
		free(s->pub.r.next);
		delete s;
		return nullptr;
	}
	s->owner = x;
	x->subs.push_back(s);
	return s;
}
}  // namespace calhip

static SubX* new_sub(ModelX* x, char type, int elem, bool single) { return new_substate(x, type, elem, single); }

// cal2D.c:470-552 (double-layer, registered for updates)
extern "C" struct CALSubstate2Db* calAddSubstate2Db(struct CALModel2D* m)
{
	SubX* s = new_sub(X(m), 'b', 1, false);
	if (!s) return nullptr;
	s->registered = true;
	append(&m->pQb_array, &m->sizeof_pQb_array, &s->pub.b);
	return &s->pub.b;
}

extern "C" struct CALSubstate2Di* calAddSubstate2Di(struct CALModel2D* m)
{
	SubX* s = new_sub(X(m), 'i', 4, false);
	if (!s) return nullptr;
	s->registered = true;
	append(&m->pQi_array, &m->sizeof_pQi_array, &s->pub.i);
	return &s->pub.i;
}

extern "C" struct CALSubstate2Dr* calAddSubstate2Dr(struct CALModel2D* m)
{
	SubX* s = new_sub(X(m), 'r', 8, false);
	if (!s) return nullptr;
	s->registered = true;
	append(&m->pQr_array, &m->sizeof_pQr_array, &s->pub.r);
	return &s->pub.r;
}

// cal2D.c:556-596 (single-layer: next == NULL, never updated)
extern "C" struct CALSubstate2Db* calAddSingleLayerSubstate2Db(struct CALModel2D* m)
{
	SubX* s = new_sub(X(m), 'b', 1, true);
	return s ? &s->pub.b : nullptr;
}

extern "C" struct CALSubstate2Di* calAddSingleLayerSubstate2Di(struct CALModel2D* m)
{
	SubX* s = new_sub(X(m), 'i', 4, true);
	return s ? &s->pub.i : nullptr;
}

extern "C" struct CALSubstate2Dr* calAddSingleLayerSubstate2Dr(struct CALModel2D* m)
{
	SubX* s = new_sub(X(m), 'r', 8, true);
	return s ? &s->pub.r : nullptr;
}

// cal2D.c:600-621
extern "C" CALCallbackFunc2D* calAddElementaryProcess2D(struct CALModel2D* m, CALCallbackFunc2D ep)
{
	CALCallbackFunc2D* na =
	    (CALCallbackFunc2D*)malloc(sizeof(CALCallbackFunc2D) * (size_t)(m->num_of_elementary_processes + 1));
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
static void host_set(CALModel2D* m, T* M, T v)
{
	if (!M) return;
	if (m->A.cells) {
		for (int n = 0; n < m->A.size_current; ++n)
			M[(size_t)m->A.cells[n].i * m->columns + m->A.cells[n].j] = v;
	} else {
		size_t N = (size_t)m->rows * m->columns;
		for (size_t c = 0; c < N; ++c) M[c] = v;
	}
}

// cal2D.c:625-691.  Full-grid int fills go through memset in the reference
// (cal2DBuffer.c:227-230), which replicates the value's low byte into every
// byte of each cell; that result is reproduced here.
template <typename T>
static T full_grid_value(T v) { return v; }
template <>
CALint full_grid_value<CALint>(CALint v)
{
	unsigned b = (unsigned)v & 0xffu;
	return (CALint)(b | (b << 8) | (b << 16) | (b << 24));
}

template <typename T>
static void init_sub(CALModel2D* m, void* q, T v, bool both)
{
	ModelX* x = X(m);
	SubX* s = S(q);
	// a host-only substate of an active-cells model needs the current list
	if (s->host_only && x->active() && ensure_host(x)) return;
	if (!m->A.cells && !(x->phase == Phase::Device && x->active())) v = full_grid_value<T>(v);
	if (x->phase == Phase::Device && !s->host_only) {
		device_fill(x, s, (double)v, both);
		return;
	}
	if (both) host_set<T>(m, (T*)s->pub.r.current, v);
	host_set<T>(m, (T*)s->pub.r.next, v);
}

extern "C" void calInitSubstate2Db(struct CALModel2D* m, struct CALSubstate2Db* Q, CALbyte v) { init_sub<CALbyte>(m, Q, v, true); }
extern "C" void calInitSubstate2Di(struct CALModel2D* m, struct CALSubstate2Di* Q, CALint v) { init_sub<CALint>(m, Q, v, true); }
extern "C" void calInitSubstate2Dr(struct CALModel2D* m, struct CALSubstate2Dr* Q, CALreal v) { init_sub<CALreal>(m, Q, v, true); }
extern "C" void calInitSubstateNext2Db(struct CALModel2D* m, struct CALSubstate2Db* Q, CALbyte v) { init_sub<CALbyte>(m, Q, v, false); }
extern "C" void calInitSubstateNext2Di(struct CALModel2D* m, struct CALSubstate2Di* Q, CALint v) { init_sub<CALint>(m, Q, v, false); }
extern "C" void calInitSubstateNext2Dr(struct CALModel2D* m, struct CALSubstate2Dr* Q, CALreal v) { init_sub<CALreal>(m, Q, v, false); }

// cal2D.c:694-713 / cal2DBuffer.c:44-172
static void host_update_sub(CALModel2D* m, SubX* s)
{
	if (s->single) return;
	char* cur = (char*)s->pub.r.current;
	char* nxt = (char*)s->pub.r.next;
	if (m->A.cells) {
		for (int n = 0; n < m->A.size_current; ++n) {
			size_t c = (size_t)m->A.cells[n].i * m->columns + m->A.cells[n].j;
			memcpy(cur + c * s->elem, nxt + c * s->elem, (size_t)s->elem);
		}
	} else {
		memcpy(cur, nxt, (size_t)m->rows * m->columns * s->elem);
	}
}

static void update_one(CALModel2D* m, void* q)
{
	ModelX* x = X(m);
	SubX* s = S(q);
	if (x->phase == Phase::Device && !s->host_only) {
		device_changed(x);
		device_update_sub(x, s);
		return;
	}
	host_update_sub(m, s);
}

extern "C" void calUpdateSubstate2Db(struct CALModel2D* m, struct CALSubstate2Db* Q) { update_one(m, Q); }
extern "C" void calUpdateSubstate2Di(struct CALModel2D* m, struct CALSubstate2Di* Q) { update_one(m, Q); }
extern "C" void calUpdateSubstate2Dr(struct CALModel2D* m, struct CALSubstate2Dr* Q) { update_one(m, Q); }

// cal2D.c:758-776
extern "C" void calUpdate2D(struct CALModel2D* m)
{
	ModelX* x = X(m);
	if (x->phase == Phase::Device) {
		device_update(x);
		return;
	}
	if (m->OPTIMIZATION == CAL_OPT_ACTIVE_CELLS) host_rebuild_active(m);
	for (int k = 0; k < m->sizeof_pQb_array; ++k) host_update_sub(m, S(m->pQb_array[k]));
	for (int k = 0; k < m->sizeof_pQi_array; ++k) host_update_sub(m, S(m->pQi_array[k]));
	for (int k = 0; k < m->sizeof_pQr_array; ++k) host_update_sub(m, S(m->pQr_array[k]));
}

// ------------------------------------------------ transition functions

static bool program_covers(ModelX* x)
{
	if (!x->prog) return false;
	CALModel2D* m = &x->pub;
	if ((int)x->prog_eps.size() != m->num_of_elementary_processes) return false;
	for (int k = 0; k < m->num_of_elementary_processes; ++k)
		if (x->prog_eps[k] != m->elementary_processes[k]) return false;
	return true;
}

// cal2D.c:738-754
extern "C" void calGlobalTransitionFunction2D(struct CALModel2D* m)
{
	ModelX* x = X(m);
	if (!program_covers(x)) {
		set_error(2, "calGlobalTransitionFunction2D: the model's elementary processes are not bound "
		             "to a device program (calhipBind*2D); there is no CPU execution path");
		return;
	}
	if (ensure_device(x)) return;
	device_changed(x);
	if (x->prog->step(x) == 0) x->stats.steps++;
}

namespace calhip {
bool program_covers_model(ModelX* x) { return program_covers(x); }

int device_transition(ModelX* x, bool with_steering)
{
	if (!program_covers(x)) {
		set_error(2, "calGlobalTransitionFunction2D: the model's elementary processes are not bound "
		             "to a device program (calhipBind*2D); there is no CPU execution path");
		return -1;
	}
	if (ensure_device(x)) return -1;
	device_changed(x);
	int e = with_steering ? x->prog->step_and_steer(x) : x->prog->step(x);
	if (e == 0) x->stats.steps++;
	return e;
}
}  // namespace calhip

// cal2D.c:717-734
extern "C" void calApplyElementaryProcess2D(struct CALModel2D* m, CALCallbackFunc2D ep)
{
	ModelX* x = X(m);
	if (x->prog) {
		if (ensure_device(x)) return;
		device_changed(x);
		const int r = x->prog->apply(x, (const void*)ep);
		if (r <= 0) return;
	}
	set_error(3, std::string("calApplyElementaryProcess2D: single-process sweeps are not provided by the ") +
	                 (x->prog ? x->prog->name() : "(unbound)") +
	                 " device program; bind the processes to source (calhipBindProcess2D) or use "
	                 "calGlobalTransitionFunction2D");
}

// ------------------------------------------------------------ accessors

template <typename T>
static inline T* cur_of(void* q) { return (T*)S(q)->pub.r.current; }
template <typename T>
static inline T* nxt_of(void* q) { return (T*)S(q)->pub.r.next; }

static inline size_t idx(CALModel2D* m, int i, int j) { return (size_t)i * m->columns + j; }

// calGetX2D* neighbour index: hex shift, then toroidal wrap or raw (FLAT).
static inline size_t nb_index(CALModel2D* m, int i, int j, int n)
{
	if ((m->X_id == CAL_HEXAGONAL_NEIGHBORHOOD_2D && j % 2 == 1) ||
	    (m->X_id == CAL_HEXAGONAL_NEIGHBORHOOD_ALT_2D && i % 2 == 1))
		n += CAL_HEXAGONAL_SHIFT;
	int ii = i + m->X[n].i, jj = j + m->X[n].j;
	if (m->T != CAL_SPACE_FLAT) {
		ii = calGetToroidalX(ii, m->rows);
		jj = calGetToroidalX(jj, m->columns);
	}
	return (size_t)((long)ii * m->columns + jj);
}

#define CAL_ACCESSORS(SUF, T)                                                                       \
	extern "C" void calInit2D##SUF(struct CALModel2D* m, struct CALSubstate2D##SUF* Q, int i, int j, T v) \
	{                                                                                                \
		if (ensure_host_for(X(m), Q)) return;                                                        \
		cur_of<T>(Q)[idx(m, i, j)] = v;                                                              \
		nxt_of<T>(Q)[idx(m, i, j)] = v;                                                              \
	}                                                                                                \
	extern "C" T calGet2D##SUF(struct CALModel2D* m, struct CALSubstate2D##SUF* Q, int i, int j)     \
	{                                                                                                \
		if (host_read(X(m), Q)) return (T)0;                                                         \
		return cur_of<T>(Q)[idx(m, i, j)];                                                           \
	}                                                                                                \
	extern "C" T calGetX2D##SUF(struct CALModel2D* m, struct CALSubstate2D##SUF* Q, int i, int j, int n) \
	{                                                                                                \
		if (host_read(X(m), Q)) return (T)0;                                                         \
		return cur_of<T>(Q)[nb_index(m, i, j, n)];                                                   \
	}                                                                                                \
	extern "C" void calSet2D##SUF(struct CALModel2D* m, struct CALSubstate2D##SUF* Q, int i, int j, T v) \
	{                                                                                                \
		if (ensure_host_for(X(m), Q)) return;                                                        \
		nxt_of<T>(Q)[idx(m, i, j)] = v;                                                              \
	}                                                                                                \
	extern "C" void calSetCurrent2D##SUF(struct CALModel2D* m, struct CALSubstate2D##SUF* Q, int i, int j, T v) \
	{                                                                                                \
		if (ensure_host_for(X(m), Q)) return;                                                        \
		cur_of<T>(Q)[idx(m, i, j)] = v;                                                              \
	}

CAL_ACCESSORS(b, CALbyte)
CAL_ACCESSORS(i, CALint)
CAL_ACCESSORS(r, CALreal)

// Unsafe operations (OpenCAL-OMP/source/cal2DUnsafe.c:17-275) on the host
// mirrors: neighbour writes, reads of the next layer, additions to the next
// layer.  The host phase is single-threaded, so the reference's cell locks
// have nothing to guard.  (On the GPU these calls are part of the device API
// of the generic program.)
#define CAL_UNSAFE(SUF, T)                                                                              \
	extern "C" void calInitX2D##SUF(struct CALModel2D* m, struct CALSubstate2D##SUF* Q, int i, int j, int n, T v) \
	{                                                                                                   \
		if (ensure_host_for(X(m), Q)) return;                                                           \
		const size_t c = nb_index(m, i, j, n);                                                          \
		cur_of<T>(Q)[c] = v;                                                                            \
		nxt_of<T>(Q)[c] = v;                                                                            \
	}                                                                                                   \
	extern "C" T calGetNext2D##SUF(struct CALModel2D* m, struct CALSubstate2D##SUF* Q, int i, int j)    \
	{                                                                                                   \
		if (ensure_host_for(X(m), Q)) return (T)0;                                                      \
		return nxt_of<T>(Q)[idx(m, i, j)];                                                              \
	}                                                                                                   \
	extern "C" T calGetNextX2D##SUF(struct CALModel2D* m, struct CALSubstate2D##SUF* Q, int i, int j, int n) \
	{                                                                                                   \
		if (ensure_host_for(X(m), Q)) return (T)0;                                                      \
		return nxt_of<T>(Q)[nb_index(m, i, j, n)];                                                      \
	}                                                                                                   \
	extern "C" void calSetX2D##SUF(struct CALModel2D* m, struct CALSubstate2D##SUF* Q, int i, int j, int n, T v) \
	{                                                                                                   \
		if (ensure_host_for(X(m), Q)) return;                                                           \
		nxt_of<T>(Q)[nb_index(m, i, j, n)] = v;                                                         \
	}                                                                                                   \
	extern "C" void calSetCurrentX2D##SUF(struct CALModel2D* m, struct CALSubstate2D##SUF* Q, int i, int j, int n, \
	                                      T v)                                                          \
	{                                                                                                   \
		if (ensure_host_for(X(m), Q)) return;                                                           \
		cur_of<T>(Q)[nb_index(m, i, j, n)] = v;                                                         \
	}                                                                                                   \
	extern "C" void calAddNext2D##SUF(struct CALModel2D* m, struct CALSubstate2D##SUF* Q, int i, int j, T v) \
	{                                                                                                   \
		if (ensure_host_for(X(m), Q)) return;                                                           \
		T* p = &nxt_of<T>(Q)[idx(m, i, j)];                                                             \
		*p = (T)(*p + v);                                                                               \
	}                                                                                                   \
	extern "C" void calAddNextX2D##SUF(struct CALModel2D* m, struct CALSubstate2D##SUF* Q, int i, int j, int n, T v) \
	{                                                                                                   \
		if (ensure_host_for(X(m), Q)) return;                                                           \
		T* p = &nxt_of<T>(Q)[nb_index(m, i, j, n)];                                                     \
		*p = (T)(*p + v);                                                                               \
	}

CAL_UNSAFE(b, CALbyte)
CAL_UNSAFE(i, CALint)
CAL_UNSAFE(r, CALreal)

// ------------------------------------------------------------- finalize

namespace calhip {
void release_device(ModelX* x)
{
	if (x->stream) hipStreamSynchronize(x->stream);
	for (SubX* s : x->subs) {
		if (s->d_cur) hipFree(s->d_cur);
		if (s->d_nxt && s->d_nxt != s->d_cur) hipFree(s->d_nxt);
		s->d_cur = s->d_nxt = nullptr;
	}
	if (x->d_flags) hipFree(x->d_flags);
	if (x->d_cells) hipFree(x->d_cells);
	if (x->d_count) hipFree(x->d_count);
	if (x->d_block_counts) hipFree(x->d_block_counts);
	if (x->d_red) hipFree(x->d_red);
	if (x->d_stage) hipFree(x->d_stage);
	x->d_red = x->d_stage = nullptr;
	x->stage_bytes = 0;
	x->d_flags = nullptr;
	x->d_cells = x->d_count = nullptr;
	x->d_block_counts = nullptr;
	for (auto& p : x->timer.pending) {
		hipEventDestroy(p.first);
		hipEventDestroy(p.second);
	}
	x->timer.pending.clear();
	if (x->stream) hipStreamDestroy(x->stream);
	x->stream = nullptr;
}
}  // namespace calhip

// cal2D.c:933-961 (also frees the substate records, which the reference leaks)
extern "C" void calFinalize2D(struct CALModel2D* m)
{
	if (!m) return;
	ModelX* x = X(m);
	release_device(x);
	delete x->prog;
	x->prog = nullptr;
	free(m->A.flags);
	free(m->A.cells);
	free(m->A.size_next);
	free(m->X);
	for (SubX* s : x->subs) {
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
