/* OpenCAL 3D models run by the generic device program, and the reductions.
 * As in generic_models.c, CALHIP_SOURCE(...) compiles each process source
 * into this file for the host (the processes' addresses identify them) and
 * keeps its text for the GPU.
 *
 *   gmod2_*   the parity automaton of OpenCAL/examples/cal_mod2CA3D-glut/
 *             source/mod2CA3D.c:20-28: each cell becomes the sum of its 27
 *             Moore cells (itself included) modulo 2, on a torus
 *   gheat_*   explicit heat diffusion on the von Neumann neighbourhood of a
 *             flat box (each cell checks its own borders), a byte "hot" flag
 *             and an int counter per cell, with the global reductions of
 *             cal3DReduction.h called from the steering and the stop
 *             condition while the model is on the GPU
 *   ginf_*    an active-cells model: infected cells activate their
 *             neighbours, susceptible active cells catch the infection from
 *             an infected neighbour, and cells `span` steps past infection
 *             leave the set and turn "removed" in the same process (on a
 *             flat box only neighbours inside the box are visited: the
 *             reference engine does not check its borders)
 *   ggift3_*  the unsafe operations in 3D: every cell adds a weight to each
 *             neighbour's next value (calAddNextX3Di/r) and reads its own
 *             next value back (calGetNext3Di) */
#include <cal3D.h>
#include <cal3DReduction.h>
#include <cal3DRun.h>
#include <cal3DUnsafe.h>
#include <stdlib.h>
#include <string.h>

#ifdef CALHIP_REFERENCE_BUILD
/* The same models built against the reference OpenCAL-OMP engine on the
 * CPU (oracle/ref/Makefile, the check of the GPU results): no device
 * program to bind. */
#define CALHIP_GLOBAL
#define CALHIP_BIND(calls) 0
static int calhipGetLastError(void) { return 0; }
static int calhipSyncHost3D(struct CALModel3D* m) { (void)m; return 0; }
static int calhipIsOnDevice3D(struct CALModel3D* m) { (void)m; return 0; }
#else
#include <calhip3D.h>
#define CALHIP_BIND(calls) (calls)
#endif

#define CALHIP_SOURCE(name, ...) \
	__VA_ARGS__                  \
	static const char* name = #__VA_ARGS__;

static size_t cells3(const struct CALModel3D* m) { return (size_t)m->rows * m->columns * m->slices; }

/* ------------------------------------------------------------ mod2 CA */

CALHIP_SOURCE(gmod2_source,
CALHIP_GLOBAL struct CALSubstate3Db* gmod2Q;

void gmod2_transition(struct CALModel3D* ca, int i, int j, int k)
{
	int sum = 0, n;
	for (n = 0; n < ca->sizeof_X; n++) sum += calGetX3Db(ca, gmod2Q, i, j, k, n);
	calSet3Db(ca, gmod2Q, i, j, k, sum % 2);
}
)

typedef struct {
	struct CALModel3D* model;
	struct CALRun3D* run;
} gmod2_t;

void* gmod2_create(int rows, int cols, int slices, const unsigned char* init, int steps)
{
	gmod2_t* s = (gmod2_t*)calloc(1, sizeof(gmod2_t));
	size_t c;
	if (!s) return NULL;
	s->model = calCADef3D(rows, cols, slices, CAL_MOORE_NEIGHBORHOOD_3D, CAL_SPACE_TOROIDAL, CAL_NO_OPT);
	if (!s->model) return NULL;
	s->run = calRunDef3D(s->model, 1, steps, CAL_UPDATE_IMPLICIT);
	calAddElementaryProcess3D(s->model, gmod2_transition);
	gmod2Q = calAddSubstate3Db(s->model);
	for (c = 0; c < cells3(s->model); c++) gmod2Q->current[c] = gmod2Q->next[c] = (CALbyte)init[c];
	if (CALHIP_BIND(calhipAddProcessSource3D(s->model, gmod2_source) ||
	                calhipBindProcess3D(s->model, gmod2_transition, "gmod2_transition") ||
	                calhipBindGlobal3D(s->model, "gmod2Q", &gmod2Q, sizeof gmod2Q)))
		return NULL;
	return s;
}

int gmod2_run(void* handle, unsigned char* out)
{
	gmod2_t* s = (gmod2_t*)handle;
	size_t c;
	calRun3D(s->run);
	for (c = 0; c < cells3(s->model); c++) out[c] = (unsigned char)gmod2Q->current[c];
	return calhipGetLastError();
}

void gmod2_destroy(void* handle)
{
	gmod2_t* s = (gmod2_t*)handle;
	if (!s) return;
	calRunFinalize3D(s->run);
	calFinalize3D(s->model);
	free(s);
}

/* ------------------------------------------------- heat + reductions */

CALHIP_SOURCE(gheat_source,
struct gheat_state {
	struct CALSubstate3Dr* T;
	struct CALSubstate3Db* hot;
	struct CALSubstate3Di* count;
	CALreal kappa;
	CALreal threshold;
};
CALHIP_GLOBAL struct gheat_state gheatS;

void gheat_diffuse(struct CALModel3D* ca, int i, int j, int k)
{
	CALreal t = calGet3Dr(ca, gheatS.T, i, j, k), d = 0.0;
	int n;
	for (n = 1; n < ca->sizeof_X; n++) {
		const int ii = i + ca->X[n].i, jj = j + ca->X[n].j, kk = k + ca->X[n].k;
		if (ii < 0 || jj < 0 || kk < 0 || ii >= ca->rows || jj >= ca->columns || kk >= ca->slices) continue;
		d += calGetX3Dr(ca, gheatS.T, i, j, k, n) - t;
	}
	t = t + gheatS.kappa * d;
	calSet3Dr(ca, gheatS.T, i, j, k, t);
	calSet3Db(ca, gheatS.hot, i, j, k, t > gheatS.threshold ? 1 : 0);
	if (t > gheatS.threshold) calSet3Di(ca, gheatS.count, i, j, k, calGet3Di(ca, gheatS.count, i, j, k) + 1);
}
)

/* per step, recorded by the steering: sum and max of T, hot cells, sum of
 * the counters; the run stops when max T falls to the stop level or after
 * max_steps */
typedef struct {
	struct CALModel3D* model;
	struct CALRun3D* run;
	int max_steps;
	double stop_level;
	int n_rec;
	double* rec;   /* 4 values per step */
} gheat_t;

static gheat_t* gheat_cur;

static void gheat_steering(struct CALModel3D* ca)
{
	gheat_t* s = gheat_cur;
	double* r = s->rec + 4 * s->n_rec;
	r[0] = calReductionComputeSum3Dr(ca, gheatS.T);
	r[1] = calReductionComputeMax3Dr(ca, gheatS.T);
	r[2] = (double)calReductionComputeSum3Di(ca, gheatS.count);
	r[3] = (double)calReductionComputeLogicalOr3Db(ca, gheatS.hot);
	s->n_rec++;
}

static CALbyte gheat_stop(struct CALModel3D* ca)
{
	gheat_t* s = gheat_cur;
	return (s->n_rec >= s->max_steps || calReductionComputeMax3Dr(ca, gheatS.T) <= s->stop_level) ? CAL_TRUE
	                                                                                               : CAL_FALSE;
}

void* gheat_create(int rows, int cols, int slices, const double* T0, double kappa, double threshold,
                   double stop_level, int max_steps, double* rec)
{
	gheat_t* s = (gheat_t*)calloc(1, sizeof(gheat_t));
	size_t c;
	if (!s) return NULL;
	s->model = calCADef3D(rows, cols, slices, CAL_VON_NEUMANN_NEIGHBORHOOD_3D, CAL_SPACE_FLAT, CAL_NO_OPT);
	if (!s->model) return NULL;
	s->run = calRunDef3D(s->model, 1, CAL_RUN_LOOP, CAL_UPDATE_IMPLICIT);
	s->max_steps = max_steps;
	s->stop_level = stop_level;
	s->rec = rec;
	calAddElementaryProcess3D(s->model, gheat_diffuse);
	gheatS.T = calAddSubstate3Dr(s->model);
	gheatS.hot = calAddSubstate3Db(s->model);
	gheatS.count = calAddSubstate3Di(s->model);
	gheatS.kappa = kappa;
	gheatS.threshold = threshold;
	for (c = 0; c < cells3(s->model); c++) gheatS.T->current[c] = gheatS.T->next[c] = T0[c];
	calInitSubstate3Db(s->model, gheatS.hot, 0);
	calInitSubstate3Di(s->model, gheatS.count, 0);
	calRunAddSteeringFunc3D(s->run, gheat_steering);
	calRunAddStopConditionFunc3D(s->run, gheat_stop);
	if (CALHIP_BIND(calhipAddProcessSource3D(s->model, gheat_source) ||
	                calhipBindProcess3D(s->model, gheat_diffuse, "gheat_diffuse") ||
	                calhipBindGlobal3D(s->model, "gheatS", &gheatS, sizeof gheatS)))
		return NULL;
	return s;
}

/* runs until the stop condition; returns the steps run (negative: error) */
int gheat_run(void* handle, double* T, unsigned char* hot, int* count)
{
	gheat_t* s = (gheat_t*)handle;
	size_t c;
	gheat_cur = s;
	s->n_rec = 0;
	calRun3D(s->run);
	if (calhipGetLastError()) return -1;
	for (c = 0; c < cells3(s->model); c++) {
		T[c] = gheatS.T->current[c];
		hot[c] = (unsigned char)gheatS.hot->current[c];
		count[c] = gheatS.count->current[c];
	}
	return s->n_rec;
}

/* one reduction of T (0), hot (1) or count (2), as a double */
double gheat_reduce(void* handle, int which, int op)
{
	gheat_t* s = (gheat_t*)handle;
	if (which == 0) return calReductionOperation3Dr(s->model, gheatS.T, (enum REDUCTION_OPERATION)op);
	if (which == 1) return (double)calReductionOperation3Db(s->model, gheatS.hot, (enum REDUCTION_OPERATION)op);
	return (double)calReductionOperation3Di(s->model, gheatS.count, (enum REDUCTION_OPERATION)op);
}

/* calReductionComputeProd3D of T (0), hot (1) or count (2), as a double */
double gheat_compute_prod(void* handle, int which)
{
	gheat_t* s = (gheat_t*)handle;
	if (which == 0) return calReductionComputeProd3Dr(s->model, gheatS.T);
	if (which == 1) return (double)calReductionComputeProd3Db(s->model, gheatS.hot);
	return (double)calReductionComputeProd3Di(s->model, gheatS.count);
}

/* overwrite the three substates from the host (the model returns to the host) */
int gheat_load(void* handle, const double* T, const unsigned char* hot, const int* count)
{
	gheat_t* s = (gheat_t*)handle;
	int i, j, k;
	struct CALModel3D* m = s->model;
	for (k = 0; k < m->slices; k++)
		for (i = 0; i < m->rows; i++)
			for (j = 0; j < m->columns; j++) {
				size_t c = ((size_t)k * m->rows + i) * m->columns + j;
				calInit3Dr(m, gheatS.T, i, j, k, T[c]);
				calInit3Db(m, gheatS.hot, i, j, k, (CALbyte)hot[c]);
				calInit3Di(m, gheatS.count, i, j, k, count[c]);
			}
	return calhipGetLastError();
}

int gheat_on_device(void* handle) { return calhipIsOnDevice3D(((gheat_t*)handle)->model); }

/* one global transition, leaving the model on the GPU */
int gheat_step(void* handle)
{
	calGlobalTransitionFunction3D(((gheat_t*)handle)->model);
	return calhipGetLastError();
}

void gheat_destroy(void* handle)
{
	gheat_t* s = (gheat_t*)handle;
	if (!s) return;
	calRunFinalize3D(s->run);
	calFinalize3D(s->model);
	free(s);
}

/* ------------------------------------------------------ active cells */

CALHIP_SOURCE(ginf_source,
struct ginf_state {
	struct CALSubstate3Db* s;     /* 0 susceptible, 1 infected, 2 removed */
	struct CALSubstate3Di* age;   /* steps since infection */
	CALint span;                  /* steps an infected cell stays active */
};
CALHIP_GLOBAL struct ginf_state ginfS;

/* neighbour n lies in the box (always, on a torus) */
int ginf_inside(struct CALModel3D* ca, int i, int j, int k, int n)
{
	const int ii = i + ca->X[n].i, jj = j + ca->X[n].j, kk = k + ca->X[n].k;
	return ca->T == CAL_SPACE_TOROIDAL ||
	       (ii >= 0 && jj >= 0 && kk >= 0 && ii < ca->rows && jj < ca->columns && kk < ca->slices);
}

void ginf_spread(struct CALModel3D* ca, int i, int j, int k)
{
	int n;
	const CALbyte st = calGet3Db(ca, ginfS.s, i, j, k);
	if (st == 1) {
		for (n = 1; n < ca->sizeof_X; n++)
			if (ginf_inside(ca, i, j, k, n)) calAddActiveCellX3D(ca, i, j, k, n);
		calSet3Di(ca, ginfS.age, i, j, k, calGet3Di(ca, ginfS.age, i, j, k) + 1);
	} else if (st == 0) {
		for (n = 1; n < ca->sizeof_X; n++)
			if (ginf_inside(ca, i, j, k, n) && calGetX3Db(ca, ginfS.s, i, j, k, n) == 1) {
				calSet3Db(ca, ginfS.s, i, j, k, 1);
				break;
			}
	}
}

void ginf_retire(struct CALModel3D* ca, int i, int j, int k)
{
	if (calGet3Di(ca, ginfS.age, i, j, k) >= ginfS.span) {
		calRemoveActiveCell3D(ca, i, j, k);
		calSet3Db(ca, ginfS.s, i, j, k, 2);
	}
}
)

typedef struct {
	struct CALModel3D* model;
} ginf_t;

void* ginf_create(int rows, int cols, int slices, int neighborhood, int boundary, const int* seeds, int n_seeds,
                  int span)
{
	ginf_t* s = (ginf_t*)calloc(1, sizeof(ginf_t));
	int q;
	if (!s) return NULL;
	s->model = calCADef3D(rows, cols, slices, (enum CALNeighborhood3D)neighborhood,
	                      (enum CALSpaceBoundaryCondition)boundary, CAL_OPT_ACTIVE_CELLS);
	if (!s->model) return NULL;
	calAddElementaryProcess3D(s->model, ginf_spread);
	calAddElementaryProcess3D(s->model, ginf_retire);
	ginfS.s = calAddSubstate3Db(s->model);
	ginfS.age = calAddSubstate3Di(s->model);
	ginfS.span = span;
	calInitSubstate3Db(s->model, ginfS.s, 0);
	calInitSubstate3Di(s->model, ginfS.age, 0);
	for (q = 0; q < n_seeds; q++) {
		calInit3Db(s->model, ginfS.s, seeds[3 * q], seeds[3 * q + 1], seeds[3 * q + 2], 1);
		calAddActiveCell3D(s->model, seeds[3 * q], seeds[3 * q + 1], seeds[3 * q + 2]);
	}
	calUpdate3D(s->model);
	if (CALHIP_BIND(calhipAddProcessSource3D(s->model, ginf_source) ||
	                calhipBindProcess3D(s->model, ginf_spread, "ginf_spread") ||
	                calhipBindProcess3D(s->model, ginf_retire, "ginf_retire") ||
	                calhipBindGlobal3D(s->model, "ginfS", &ginfS, sizeof ginfS)))
		return NULL;
	return s;
}

int ginf_steps(void* handle, int n)
{
	ginf_t* s = (ginf_t*)handle;
	int k;
	for (k = 0; k < n; k++) calGlobalTransitionFunction3D(s->model);
	return calhipGetLastError();
}

/* the same steps with each process run through calApplyElementaryProcess3D
 * and followed by calUpdate3D (the sequence of calGlobalTransitionFunction3D,
 * cal3D.c:677-695, written out by the model): ginf_retire writes a cell's
 * next layer and removes the cell, which the update then leaves out */
int ginf_apply_steps(void* handle, int n)
{
	ginf_t* s = (ginf_t*)handle;
	int k;
	for (k = 0; k < n; k++) {
		calApplyElementaryProcess3D(s->model, ginf_spread);
		calUpdate3D(s->model);
		calApplyElementaryProcess3D(s->model, ginf_retire);
		calUpdate3D(s->model);
	}
	return calhipGetLastError();
}

/* both layers of both substates, and the active list as (i, j, k) triples;
 * returns the list length (negative: error) */
int ginf_read(void* handle, unsigned char* s_cur, unsigned char* s_next, int* age_cur, int* age_next, int* cells,
              int max_cells)
{
	ginf_t* s = (ginf_t*)handle;
	struct CALModel3D* m = s->model;
	size_t c;
	int q, n;
	if (calhipSyncHost3D(m)) return -1;
	for (c = 0; c < cells3(m); c++) {
		s_cur[c] = (unsigned char)ginfS.s->current[c];
		s_next[c] = (unsigned char)ginfS.s->next[c];
		age_cur[c] = ginfS.age->current[c];
		age_next[c] = ginfS.age->next[c];
	}
	n = m->A.cells ? m->A.size_current : 0;
	for (q = 0; q < n && q < max_cells; q++) {
		cells[3 * q] = m->A.cells[q].i;
		cells[3 * q + 1] = m->A.cells[q].j;
		cells[3 * q + 2] = m->A.cells[q].k;
	}
	return n;
}

void ginf_destroy(void* handle)
{
	ginf_t* s = (ginf_t*)handle;
	if (!s) return;
	calFinalize3D(s->model);
	free(s);
}

/* ----------------------------------------------------- unsafe operations */

CALHIP_SOURCE(ggift3_source,
struct ggift3_state {
	struct CALSubstate3Di* q;
	struct CALSubstate3Di* seen;
	struct CALSubstate3Dr* acc;
};
CALHIP_GLOBAL struct ggift3_state ggift3S;

void ggift3_give(struct CALModel3D* ca, int i, int j, int k)
{
	int n;
	const int w = 1 + (i * 7 + j * 3 + k * 5) % 5;
	for (n = 1; n < ca->sizeof_X; n++) {
		calAddNextX3Di(ca, ggift3S.q, i, j, k, n, w);
		calAddNextX3Dr(ca, ggift3S.acc, i, j, k, n, 0.25 * w);
	}
}

void ggift3_look(struct CALModel3D* ca, int i, int j, int k)
{
	calSet3Di(ca, ggift3S.seen, i, j, k, calGetNext3Di(ca, ggift3S.q, i, j, k) + calGetX3Di(ca, ggift3S.q, i, j, k, 1));
}
)

typedef struct {
	struct CALModel3D* model;
} ggift3_t;

void* ggift3_create(int rows, int cols, int slices, int neighborhood, int boundary)
{
	ggift3_t* s = (ggift3_t*)calloc(1, sizeof(ggift3_t));
	if (!s) return NULL;
	s->model = calCADef3D(rows, cols, slices, (enum CALNeighborhood3D)neighborhood,
	                      (enum CALSpaceBoundaryCondition)boundary, CAL_NO_OPT);
	if (!s->model) return NULL;
	ggift3S.q = calAddSubstate3Di(s->model);
	ggift3S.seen = calAddSubstate3Di(s->model);
	ggift3S.acc = calAddSubstate3Dr(s->model);
	calInitSubstate3Di(s->model, ggift3S.q, 0);
	calInitSubstate3Di(s->model, ggift3S.seen, 0);
	calInitSubstate3Dr(s->model, ggift3S.acc, 0.0);
	calAddElementaryProcess3D(s->model, ggift3_give);
	calAddElementaryProcess3D(s->model, ggift3_look);
	if (CALHIP_BIND(calhipAddProcessSource3D(s->model, ggift3_source) ||
	                calhipBindProcess3D(s->model, ggift3_give, "ggift3_give") ||
	                calhipBindProcess3D(s->model, ggift3_look, "ggift3_look") ||
	                calhipBindGlobal3D(s->model, "ggift3S", &ggift3S, sizeof ggift3S)))
		return NULL;
	return s;
}

int ggift3_steps(void* handle, int n)
{
	ggift3_t* s = (ggift3_t*)handle;
	int k;
	for (k = 0; k < n; k++) calGlobalTransitionFunction3D(s->model);
	return calhipGetLastError();
}

/* calApplyElementaryProcess3D of the first process, then calUpdate3D */
int ggift3_apply_give(void* handle)
{
	ggift3_t* s = (ggift3_t*)handle;
	calApplyElementaryProcess3D(s->model, ggift3_give);
	calUpdate3D(s->model);
	return calhipGetLastError();
}

int ggift3_read(void* handle, int* q, int* seen, double* acc)
{
	ggift3_t* s = (ggift3_t*)handle;
	struct CALModel3D* m = s->model;
	int i, j, k;
	for (k = 0; k < m->slices; k++)
		for (i = 0; i < m->rows; i++)
			for (j = 0; j < m->columns; j++) {
				size_t c = ((size_t)k * m->rows + i) * m->columns + j;
				q[c] = calGet3Di(m, ggift3S.q, i, j, k);
				seen[c] = calGet3Di(m, ggift3S.seen, i, j, k);
				acc[c] = calGet3Dr(m, ggift3S.acc, i, j, k);
			}
	return calhipGetLastError();
}

/* the model's neighbourhood table as (i, j, k) triples */
int ggift3_table(void* handle, int* out, int max_entries)
{
	struct CALModel3D* m = ((ggift3_t*)handle)->model;
	int n;
	for (n = 0; n < m->sizeof_X && n < max_entries; n++) {
		out[3 * n] = m->X[n].i;
		out[3 * n + 1] = m->X[n].j;
		out[3 * n + 2] = m->X[n].k;
	}
	return m->sizeof_X;
}

void ggift3_destroy(void* handle)
{
	ggift3_t* s = (ggift3_t*)handle;
	if (!s) return;
	calFinalize3D(s->model);
	free(s);
}

#ifndef CALHIP_REFERENCE_BUILD
/* Compile every 3D source above for the GPU without running it. */
int generic3d_check_sources(char* log, size_t log_bytes)
{
	const char* mod2[] = {"gmod2_transition"};
	const char* heat[] = {"gheat_diffuse"};
	const char* inf[] = {"ginf_spread", "ginf_retire"};
	const char* gift[] = {"ggift3_give", "ggift3_look"};
	if (calhipCheckProcessSource3D(gmod2_source, mod2, 1, log, log_bytes)) return 1;
	if (calhipCheckProcessSource3D(gheat_source, heat, 1, log, log_bytes)) return 2;
	if (calhipCheckProcessSource3D(ginf_source, inf, 2, log, log_bytes)) return 3;
	if (calhipCheckProcessSource3D(ggift3_source, gift, 2, log, log_bytes)) return 4;
	return 0;
}
#endif
