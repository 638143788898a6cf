/* OpenCAL models run by the generic device program: their elementary
 * processes are ordinary C written against OpenCAL's cell API, compiled into
 * this file for the host (their addresses identify the processes, as
 * calAddElementaryProcess2D expects) and handed to the engine as source for
 * the GPU.  CALHIP_SOURCE(...) does both with one text: it emits the code and
 * a string holding it.  (A model keeping its processes in a separate file
 * would read that file instead.)
 *
 *   gsct_*   SciddicaT, full grid and active cells (OpenCAL/examples/
 *            cal_sciddicaT/source/sciddicaT.c:36-140, active variant
 *            cal_sciddicaT-activecells-glut/source/sciddicaT.c:20-100)
 *   glife_*  Conway Life on an int substate (OpenCAL-OMP/examples/
 *            calomp_life-glut/source/life.c:16-27)
 *   ggift_*  an integer model written with the unsafe operations: every cell
 *            adds its index weight to each neighbour's next value
 *            (calAddNextX2Di) and reads its own next value
 *            (calGetNext2Di) in a second process; any neighbourhood and
 *            boundary
 *   gred_*   a byte, an int and a real substate for the reductions of
 *            cal2DReduction.h, on the host arrays and on the GPU */
#include <cal2D.h>
#include <cal2DReduction.h>
#include <cal2DRun.h>
#include <cal2DUnsafe.h>
#include <calhip2D.h>
#include <stdlib.h>
#include <string.h>

#define CALHIP_SOURCE(name, ...) \
	__VA_ARGS__                  \
	static const char* name = #__VA_ARGS__;

/* ------------------------------------------------------------- SciddicaT */

CALHIP_SOURCE(gsct_source,
struct gsct_substates {
	struct CALSubstate2Dr* z;
	struct CALSubstate2Dr* h;
	struct CALSubstate2Dr* f[4];
};
struct gsct_parameters {
	CALreal epsilon;
	CALreal r;
	CALint active;
};
CALHIP_GLOBAL struct gsct_substates gsctQ;
CALHIP_GLOBAL struct gsct_parameters gsctP;

void gsct_flows(struct CALModel2D* ca, int i, int j)
{
	CALbyte eliminated[5] = {CAL_FALSE, CAL_FALSE, CAL_FALSE, CAL_FALSE, CAL_FALSE};
	CALbyte again;
	CALint cnt, n;
	CALreal average, m, u[5];
	if (calGet2Dr(ca, gsctQ.h, i, j) <= gsctP.epsilon) return;
	m = calGet2Dr(ca, gsctQ.h, i, j) - gsctP.epsilon;
	u[0] = calGet2Dr(ca, gsctQ.z, i, j) + gsctP.epsilon;
	for (n = 1; n < 5; n++) u[n] = calGetX2Dr(ca, gsctQ.z, i, j, n) + calGetX2Dr(ca, gsctQ.h, i, j, n);
	do {
		again = CAL_FALSE;
		average = m;
		cnt = 0;
		for (n = 0; n < 5; n++)
			if (!eliminated[n]) {
				average += u[n];
				cnt++;
			}
		if (cnt != 0) average /= cnt;
		for (n = 0; n < 5; n++)
			if (average <= u[n] && !eliminated[n]) {
				eliminated[n] = CAL_TRUE;
				again = CAL_TRUE;
			}
	} while (again);
	for (n = 1; n < 5; n++) {
		if (eliminated[n]) {
			calSet2Dr(ca, gsctQ.f[n - 1], i, j, 0.0);
		} else {
			calSet2Dr(ca, gsctQ.f[n - 1], i, j, (average - u[n]) * gsctP.r);
			if (gsctP.active) calAddActiveCellX2D(ca, i, j, n);
		}
	}
}

void gsct_width(struct CALModel2D* ca, int i, int j)
{
	CALreal h_next = calGet2Dr(ca, gsctQ.h, i, j);
	CALint n;
	for (n = 1; n < 5; n++)
		h_next += calGetX2Dr(ca, gsctQ.f[4 - n], i, j, n) - calGet2Dr(ca, gsctQ.f[n - 1], i, j);
	calSet2Dr(ca, gsctQ.h, i, j, h_next);
}

void gsct_remove(struct CALModel2D* ca, int i, int j)
{
	if (calGet2Dr(ca, gsctQ.h, i, j) <= gsctP.epsilon) calRemoveActiveCell2D(ca, i, j);
}
)

typedef struct {
	struct CALModel2D* model;
	struct CALRun2D* run;
	int steps;
} gsct_t;

static void gsct_init(struct CALModel2D* ca)
{
	CALint i, j, k;
	for (k = 0; k < 4; k++) calInitSubstate2Dr(ca, gsctQ.f[k], 0);
	for (i = 0; i < ca->rows; i++)
		for (j = 0; j < ca->columns; j++) {
			CALreal h = calGet2Dr(ca, gsctQ.h, i, j);
			if (h > 0.0) {
				calSet2Dr(ca, gsctQ.z, i, j, calGet2Dr(ca, gsctQ.z, i, j) - h);
				if (gsctP.active) calAddActiveCell2D(ca, i, j);
			}
		}
}

static void gsct_steering(struct CALModel2D* ca)
{
	CALint k;
	for (k = 0; k < 4; k++) {
		if (gsctP.active)
			calInitSubstateNext2Dr(ca, gsctQ.f[k], 0);
		else
			calInitSubstate2Dr(ca, gsctQ.f[k], 0);
	}
}

/* The model as the reference example builds it, its processes bound to the
 * source above. */
void* gsct_create(int rows, int cols, const double* z, const double* h, double p_r, double p_eps, int steps,
                  int active)
{
	size_t n = (size_t)rows * cols;
	int k;
	gsct_t* s = (gsct_t*)calloc(1, sizeof(gsct_t));
	if (!s) return NULL;
	s->steps = steps;
	s->model = calCADef2D(rows, cols, CAL_VON_NEUMANN_NEIGHBORHOOD_2D, CAL_SPACE_TOROIDAL,
	                      active ? CAL_OPT_ACTIVE_CELLS : CAL_NO_OPT);
	if (!s->model) return NULL;
	s->run = calRunDef2D(s->model, 1, steps, CAL_UPDATE_IMPLICIT);
	calAddElementaryProcess2D(s->model, gsct_flows);
	calAddElementaryProcess2D(s->model, gsct_width);
	if (active) calAddElementaryProcess2D(s->model, gsct_remove);
	gsctQ.z = calAddSubstate2Dr(s->model);
	gsctQ.h = calAddSubstate2Dr(s->model);
	for (k = 0; k < 4; k++) gsctQ.f[k] = calAddSubstate2Dr(s->model);
	memcpy(gsctQ.z->current, z, n * sizeof(double));
	memcpy(gsctQ.z->next, z, n * sizeof(double));
	memcpy(gsctQ.h->current, h, n * sizeof(double));
	memcpy(gsctQ.h->next, h, n * sizeof(double));
	gsctP.r = p_r;
	gsctP.epsilon = p_eps;
	gsctP.active = active;
	calRunAddInitFunc2D(s->run, gsct_init);
	calRunAddSteeringFunc2D(s->run, gsct_steering);
	if (calhipAddProcessSource2D(s->model, gsct_source) ||
	    calhipBindProcess2D(s->model, gsct_flows, "gsct_flows") ||
	    calhipBindProcess2D(s->model, gsct_width, "gsct_width") ||
	    (active && calhipBindProcess2D(s->model, gsct_remove, "gsct_remove")) ||
	    calhipBindGlobal2D(s->model, "gsctQ", &gsctQ, sizeof gsctQ) ||
	    calhipBindGlobal2D(s->model, "gsctP", &gsctP, sizeof gsctP))
		return NULL;
	return s;
}

int gsct_run(void* handle)
{
	gsct_t* s = (gsct_t*)handle;
	calRun2D(s->run);
	return calhipGetLastError();
}

int gsct_read(void* handle, double* z, double* h, int* active_count)
{
	gsct_t* s = (gsct_t*)handle;
	size_t n = (size_t)s->model->rows * s->model->columns;
	if (calhipSyncHost2D(s->model)) return -1;
	memcpy(z, gsctQ.z->current, n * sizeof(double));
	memcpy(h, gsctQ.h->current, n * sizeof(double));
	if (active_count) *active_count = s->model->A.cells ? s->model->A.size_current : 0;
	return calhipGetLastError();
}

void gsct_destroy(void* handle)
{
	gsct_t* s = (gsct_t*)handle;
	if (!s) return;
	calRunFinalize2D(s->run);
	calFinalize2D(s->model);
	free(s);
}

/* ------------------------------------------------------------------ Life */

CALHIP_SOURCE(glife_source,
CALHIP_GLOBAL struct CALSubstate2Di* glifeQ;

void glife_transition(struct CALModel2D* ca, int i, int j)
{
	int sum = 0, n;
	for (n = 1; n < ca->sizeof_X; n++) sum += calGetX2Di(ca, glifeQ, i, j, n);
	if (sum == 3 || (sum == 2 && calGet2Di(ca, glifeQ, i, j) == 1))
		calSet2Di(ca, glifeQ, i, j, 1);
	else
		calSet2Di(ca, glifeQ, i, j, 0);
}
)

typedef struct {
	struct CALModel2D* model;
	struct CALRun2D* run;
} glife_t;

void* glife_create(int rows, int cols, const unsigned char* alive, int steps)
{
	glife_t* s = (glife_t*)calloc(1, sizeof(glife_t));
	int i, j;
	if (!s) return NULL;
	s->model = calCADef2D(rows, cols, CAL_MOORE_NEIGHBORHOOD_2D, CAL_SPACE_TOROIDAL, CAL_NO_OPT);
	if (!s->model) return NULL;
	s->run = calRunDef2D(s->model, 1, steps, CAL_UPDATE_IMPLICIT);
	glifeQ = calAddSubstate2Di(s->model);
	calAddElementaryProcess2D(s->model, glife_transition);
	for (i = 0; i < rows; i++)
		for (j = 0; j < cols; j++) calInit2Di(s->model, glifeQ, i, j, alive[(size_t)i * cols + j]);
	if (calhipAddProcessSource2D(s->model, glife_source) ||
	    calhipBindProcess2D(s->model, glife_transition, "glife_transition") ||
	    calhipBindGlobal2D(s->model, "glifeQ", &glifeQ, sizeof glifeQ))
		return NULL;
	return s;
}

int glife_run(void* handle)
{
	glife_t* s = (glife_t*)handle;
	calRun2D(s->run);
	return calhipGetLastError();
}

int glife_read(void* handle, unsigned char* alive)
{
	glife_t* s = (glife_t*)handle;
	int i, j;
	for (i = 0; i < s->model->rows; i++)
		for (j = 0; j < s->model->columns; j++)
			alive[(size_t)i * s->model->columns + j] = (unsigned char)calGet2Di(s->model, glifeQ, i, j);
	return calhipGetLastError();
}

void glife_destroy(void* handle)
{
	glife_t* s = (glife_t*)handle;
	if (!s) return;
	calRunFinalize2D(s->run);
	calFinalize2D(s->model);
	free(s);
}

/* ------------------------------------------------------ unsafe operations */

CALHIP_SOURCE(ggift_source,
struct ggift_state {
	struct CALSubstate2Di* q;      /* the gifts received */
	struct CALSubstate2Di* seen;   /* a cell's own next value of q, read back */
	struct CALSubstate2Dr* acc;    /* a real accumulator of the same gifts */
};
CALHIP_GLOBAL struct ggift_state ggiftS;

void ggift_give(struct CALModel2D* ca, int i, int j)
{
	int n;
	const int w = 1 + (i * 7 + j * 3) % 5;
	for (n = 1; n < ca->sizeof_X; n++) {
		calAddNextX2Di(ca, ggiftS.q, i, j, n, w);
		calAddNextX2Dr(ca, ggiftS.acc, i, j, n, 0.25 * w);
	}
}

void ggift_look(struct CALModel2D* ca, int i, int j)
{
	calSet2Di(ca, ggiftS.seen, i, j, calGetNext2Di(ca, ggiftS.q, i, j) + calGetX2Di(ca, ggiftS.q, i, j, 1));
}
)

typedef struct {
	struct CALModel2D* model;
} ggift_t;

void* ggift_create(int rows, int cols, int neighborhood, int boundary)
{
	ggift_t* s = (ggift_t*)calloc(1, sizeof(ggift_t));
	if (!s) return NULL;
	s->model = calCADef2D(rows, cols, (enum CALNeighborhood2D)neighborhood, (enum CALSpaceBoundaryCondition)boundary,
	                      CAL_NO_OPT);
	if (!s->model) return NULL;
	ggiftS.q = calAddSubstate2Di(s->model);
	ggiftS.seen = calAddSubstate2Di(s->model);
	ggiftS.acc = calAddSubstate2Dr(s->model);
	calInitSubstate2Di(s->model, ggiftS.q, 0);
	calInitSubstate2Di(s->model, ggiftS.seen, 0);
	calInitSubstate2Dr(s->model, ggiftS.acc, 0.0);
	calAddElementaryProcess2D(s->model, ggift_give);
	calAddElementaryProcess2D(s->model, ggift_look);
	if (calhipAddProcessSource2D(s->model, ggift_source) ||
	    calhipBindProcess2D(s->model, ggift_give, "ggift_give") ||
	    calhipBindProcess2D(s->model, ggift_look, "ggift_look") ||
	    calhipBindGlobal2D(s->model, "ggiftS", &ggiftS, sizeof ggiftS))
		return NULL;
	return s;
}

/* n global transitions (each process followed by the update) */
int ggift_steps(void* handle, int n)
{
	ggift_t* s = (ggift_t*)handle;
	int k;
	for (k = 0; k < n; k++) calGlobalTransitionFunction2D(s->model);
	return calhipGetLastError();
}

/* calApplyElementaryProcess2D of the first process, then calUpdate2D */
int ggift_apply_give(void* handle)
{
	ggift_t* s = (ggift_t*)handle;
	calApplyElementaryProcess2D(s->model, ggift_give);
	calUpdate2D(s->model);
	return calhipGetLastError();
}

int ggift_read(void* handle, int* q, int* seen, double* acc)
{
	ggift_t* s = (ggift_t*)handle;
	int i, j;
	for (i = 0; i < s->model->rows; i++)
		for (j = 0; j < s->model->columns; j++) {
			size_t c = (size_t)i * s->model->columns + j;
			q[c] = calGet2Di(s->model, ggiftS.q, i, j);
			seen[c] = calGet2Di(s->model, ggiftS.seen, i, j);
			acc[c] = calGet2Dr(s->model, ggiftS.acc, i, j);
		}
	return calhipGetLastError();
}

/* the model's neighbourhood table (hexagonal: both tables) */
int ggift_table(void* handle, int* out, int max_entries)
{
	ggift_t* s = (ggift_t*)handle;
	struct CALModel2D* m = s->model;
	int hex = m->X_id == CAL_HEXAGONAL_NEIGHBORHOOD_2D || m->X_id == CAL_HEXAGONAL_NEIGHBORHOOD_ALT_2D;
	int nx = m->sizeof_X + (hex ? CAL_HEXAGONAL_SHIFT : 0), k;
	for (k = 0; k < nx && k < max_entries; k++) {
		out[2 * k] = m->X[k].i;
		out[2 * k + 1] = m->X[k].j;
	}
	return nx;
}

void ggift_destroy(void* handle)
{
	ggift_t* s = (ggift_t*)handle;
	if (!s) return;
	calFinalize2D(s->model);
	free(s);
}

/* ------------------------------------------------- active-set rebuild */

CALHIP_SOURCE(gmark_source,
struct gmark_state {
	struct CALSubstate2Di* mark;   /* cells with a non-zero mark join the set */
};
CALHIP_GLOBAL struct gmark_state gmarkS;

void gmark_select(struct CALModel2D* ca, int i, int j)
{
	if (calGet2Di(ca, gmarkS.mark, i, j))
		calAddActiveCell2D(ca, i, j);
	else
		calRemoveActiveCell2D(ca, i, j);
}
)

typedef struct {
	struct CALModel2D* model;
} gmark_t;

/* An active-cells model whose one process adds the marked cells to the set
 * (and removes the others); each transition ends with the device rebuild of
 * the row-major list. */
void* gmark_create(int rows, int cols, const int* mark)
{
	gmark_t* s = (gmark_t*)calloc(1, sizeof(gmark_t));
	if (!s) return NULL;
	s->model = calCADef2D(rows, cols, CAL_VON_NEUMANN_NEIGHBORHOOD_2D, CAL_SPACE_TOROIDAL, CAL_OPT_ACTIVE_CELLS);
	if (!s->model) return NULL;
	gmarkS.mark = calAddSubstate2Di(s->model);
	memcpy(gmarkS.mark->current, mark, sizeof(int) * (size_t)rows * cols);
	memcpy(gmarkS.mark->next, mark, sizeof(int) * (size_t)rows * cols);
	calAddElementaryProcess2D(s->model, gmark_select);
	if (calhipAddProcessSource2D(s->model, gmark_source) ||
	    calhipBindProcess2D(s->model, gmark_select, "gmark_select") ||
	    calhipBindGlobal2D(s->model, "gmarkS", &gmarkS, sizeof gmarkS))
		return NULL;
	return s;
}

int gmark_steps(void* handle, int n)
{
	gmark_t* s = (gmark_t*)handle;
	int k;
	for (k = 0; k < n; k++) calGlobalTransitionFunction2D(s->model);
	return calhipGetLastError();
}

/* the set as OpenCAL holds it: A.cells as linear indices, A.size_current */
int gmark_read(void* handle, int* cells, int max_cells, int* count)
{
	gmark_t* s = (gmark_t*)handle;
	struct CALModel2D* m = s->model;
	int k;
	if (calhipSyncHost2D(m)) return -1;
	*count = m->A.size_current;
	for (k = 0; k < m->A.size_current && k < max_cells; k++) cells[k] = m->A.cells[k].i * m->columns + m->A.cells[k].j;
	return calhipGetLastError();
}

void gmark_destroy(void* handle)
{
	gmark_t* s = (gmark_t*)handle;
	if (!s) return;
	calFinalize2D(s->model);
	free(s);
}

/* ------------------------------------------------------------ reductions */

CALHIP_SOURCE(gred_source,
struct gred_state {
	struct CALSubstate2Db* b;
	struct CALSubstate2Di* i;
	struct CALSubstate2Dr* r;
};
CALHIP_GLOBAL struct gred_state gredS;

void gred_keep(struct CALModel2D* ca, int i, int j)
{
	calSet2Db(ca, gredS.b, i, j, calGet2Db(ca, gredS.b, i, j));
	calSet2Di(ca, gredS.i, i, j, calGet2Di(ca, gredS.i, i, j));
	calSet2Dr(ca, gredS.r, i, j, calGet2Dr(ca, gredS.r, i, j));
}
)

typedef struct {
	struct CALModel2D* model;
} gred_t;

/* a model holding the three arrays in a byte, an int and a real substate */
void* gred_create(int rows, int cols, const signed char* b, const int* iv, const double* r)
{
	gred_t* s = (gred_t*)calloc(1, sizeof(gred_t));
	size_t n = (size_t)rows * cols;
	if (!s) return NULL;
	s->model = calCADef2D(rows, cols, CAL_VON_NEUMANN_NEIGHBORHOOD_2D, CAL_SPACE_TOROIDAL, CAL_NO_OPT);
	if (!s->model) return NULL;
	gredS.b = calAddSubstate2Db(s->model);
	gredS.i = calAddSubstate2Di(s->model);
	gredS.r = calAddSubstate2Dr(s->model);
	memcpy(gredS.b->current, b, n);
	memcpy(gredS.b->next, b, n);
	memcpy(gredS.i->current, iv, n * sizeof(int));
	memcpy(gredS.i->next, iv, n * sizeof(int));
	memcpy(gredS.r->current, r, n * sizeof(double));
	memcpy(gredS.r->next, r, n * sizeof(double));
	calAddElementaryProcess2D(s->model, gred_keep);
	if (calhipAddProcessSource2D(s->model, gred_source) || calhipBindProcess2D(s->model, gred_keep, "gred_keep") ||
	    calhipBindGlobal2D(s->model, "gredS", &gredS, sizeof gredS))
		return NULL;
	return s;
}

/* calReductionOperation2D{b,i,r} (which = 0, 1, 2) as a double */
double gred_reduce(void* handle, int which, int op)
{
	gred_t* s = (gred_t*)handle;
	if (which == 0) return (double)calReductionOperation2Db(s->model, gredS.b, (enum REDUCTION_OPERATION)op);
	if (which == 1) return (double)calReductionOperation2Di(s->model, gredS.i, (enum REDUCTION_OPERATION)op);
	return calReductionOperation2Dr(s->model, gredS.r, (enum REDUCTION_OPERATION)op);
}

/* calReductionComputeProd2D{b,i,r} (which = 0, 1, 2) as a double */
double gred_compute_prod(void* handle, int which)
{
	gred_t* s = (gred_t*)handle;
	if (which == 0) return (double)calReductionComputeProd2Db(s->model, gredS.b);
	if (which == 1) return (double)calReductionComputeProd2Di(s->model, gredS.i);
	return calReductionComputeProd2Dr(s->model, gredS.r);
}

/* a transition that keeps every value: the model moves to the GPU */
int gred_step(void* handle)
{
	calGlobalTransitionFunction2D(((gred_t*)handle)->model);
	return calhipGetLastError();
}

int gred_on_device(void* handle) { return calhipIsOnDevice2D(((gred_t*)handle)->model); }

void gred_destroy(void* handle)
{
	gred_t* s = (gred_t*)handle;
	if (!s) return;
	calFinalize2D(s->model);
	free(s);
}

/* Compile every source above for the GPU without running it (0 when all
 * compile; otherwise the compiler's log). */
int generic_check_sources(char* log, size_t log_bytes)
{
	const char* sct[] = {"gsct_flows", "gsct_width", "gsct_remove"};
	const char* life[] = {"glife_transition"};
	const char* gift[] = {"ggift_give", "ggift_look"};
	if (calhipCheckProcessSource2D(gsct_source, sct, 3, log, log_bytes)) return 1;
	if (calhipCheckProcessSource2D(glife_source, life, 1, log, log_bytes)) return 2;
	const char* red[] = {"gred_keep"};
	if (calhipCheckProcessSource2D(ggift_source, gift, 2, log, log_bytes)) return 3;
	if (calhipCheckProcessSource2D(gred_source, red, 1, log, log_bytes)) return 4;
	const char* mark[] = {"gmark_select"};
	if (calhipCheckProcessSource2D(gmark_source, mark, 1, log, log_bytes)) return 5;
	return 0;
}
