/* Conway Life written against the OpenCAL C API as the reference example does
 * (OpenCAL-OMP/examples/calomp_life-glut/source/life.c:16-66), with a byte
 * substate, plus the binding to the Life device program. */
#include <cal2D.h>
#include <cal2DRun.h>
#include <calhip2D.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
	struct CALModel2D* model;
	struct CALRun2D* run;
	struct CALSubstate2Db* Q;
} life_t;

static life_t* L;

static void life_transition(struct CALModel2D* ca, int i, int j)
{
	int sum = 0, n;
	for (n = 1; n < ca->sizeof_X; n++) sum += calGetX2Db(ca, L->Q, i, j, n);
	if ((sum == 3) || (sum == 2 && calGet2Db(ca, L->Q, i, j) == 1))
		calSet2Db(ca, L->Q, i, j, 1);
	else
		calSet2Db(ca, L->Q, i, j, 0);
}

/* Full calRun2D of `steps` steps from the 0/1 grid `alive0`; no callbacks run
 * between steps, so the engine may batch them. */
void* life_create(int rows, int cols, const unsigned char* alive0, int steps)
{
	size_t n = (size_t)rows * cols;
	life_t* s = (life_t*)calloc(1, sizeof(life_t));
	if (!s) return NULL;
	L = s;
	s->model = calCADef2D(rows, cols, CAL_MOORE_NEIGHBORHOOD_2D, CAL_SPACE_TOROIDAL, CAL_NO_OPT);
	if (!s->model) return NULL;
	s->run = calRunDef2D(s->model, 1, steps, CAL_UPDATE_IMPLICIT);
	s->Q = calAddSubstate2Db(s->model);
	calAddElementaryProcess2D(s->model, life_transition);
	memcpy(s->Q->current, alive0, n);
	memcpy(s->Q->next, alive0, n);
	if (calhipBindLife2D(s->model, life_transition, s->Q, NULL)) return NULL;
	return s;
}

/* A global transition of the model's own: the process applied alone, then
 * calUpdateSubstate2Db (mode 1) or calUpdate2D (mode 2). */
static void life_explicit_substate(struct CALModel2D* ca)
{
	calApplyElementaryProcess2D(ca, life_transition);
	calUpdateSubstate2Db(ca, L->Q);
}

static void life_explicit_update(struct CALModel2D* ca)
{
	calApplyElementaryProcess2D(ca, life_transition);
	calUpdate2D(ca);
}

int life_set_explicit(void* h, int mode)
{
	life_t* s = (life_t*)h;
	if (mode == 1)
		calRunAddGlobalTransitionFunc2D(s->run, life_explicit_substate);
	else if (mode == 2)
		calRunAddGlobalTransitionFunc2D(s->run, life_explicit_update);
	else
		return -1;
	return 0;
}

int life_run(void* h)
{
	life_t* s = (life_t*)h;
	L = s;
	calRun2D(s->run);
	return calhipGetLastError();
}

/* n single steps through calRunCAStep2D (one launch per step). */
int life_steps(void* h, int n)
{
	life_t* s = (life_t*)h;
	int k;
	L = s;
	for (k = 0; k < n; k++) calRunCAStep2D(s->run);
	return calhipGetLastError();
}

int life_read(void* h, unsigned char* out)
{
	life_t* s = (life_t*)h;
	if (calhipSyncHost2D(s->model)) return -1;
	memcpy(out, s->Q->current, (size_t)s->model->rows * s->model->columns);
	return 0;
}

struct CALModel2D* life_model(void* h) { return ((life_t*)h)->model; }

void life_destroy(void* h)
{
	life_t* s = (life_t*)h;
	if (!s) return;
	calRunFinalize2D(s->run);
	calFinalize2D(s->model);
	if (L == s) L = NULL;
	free(s);
}
