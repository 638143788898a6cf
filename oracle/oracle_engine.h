/* ORACLE — test infrastructure, not product code.
 * Minimal restatement of the OpenCAL-OMP engine mechanics the models need:
 * double-layer substates, the active-cell flag grid with its row-major
 * rebuild, the next->current update, neighbourhood offsets and toroidal wrap.
 * Reference: OpenCAL-OMP/source/cal2D.c, cal2DBuffer.c, include/calCommon.h. */
#ifndef OPENCAL_ORACLE_ENGINE_H
#define OPENCAL_ORACLE_ENGINE_H
#include <omp.h>
#include <stdlib.h>
#include <string.h>

typedef struct { double* cur; double* nxt; } orc_sub_r;

typedef struct {
	int rows, cols;
	int active_mode;       /* CAL_OPT_ACTIVE_CELLS */
	unsigned char* flags;  /* A.flags */
	int* cells;            /* A.cells as linear indices; NULL when the set is empty or not built */
	int ncells;            /* A.size_current */
	long pending;          /* sum over threads of A.size_next */
} orc_grid;

/* calGetToroidalX (calCommon.h:62): one wrap only. */
static inline int orc_wrap(int x, int size)
{
	return x < 0 ? size + x : (x > size - 1 ? x - size : x);
}

/* Von Neumann (cal2D.c:28-44) and Moore (cal2D.c:50-70) offsets, index 0 = centre. */
static const int ORC_DI[9] = {0, -1, 0, 0, 1, -1, 1, 1, -1};
static const int ORC_DJ[9] = {0, 0, -1, 1, 0, -1, -1, 1, 1};

static inline int orc_nb_toroidal(const orc_grid* g, int i, int j, int n)
{
	return orc_wrap(i + ORC_DI[n], g->rows) * g->cols + orc_wrap(j + ORC_DJ[n], g->cols);
}

static inline void orc_sub_alloc(orc_sub_r* q, size_t n)
{
	q->cur = (double*)malloc(n * sizeof(double));
	q->nxt = (double*)malloc(n * sizeof(double));
}

static inline void orc_sub_free(orc_sub_r* q)
{
	free(q->cur);
	free(q->nxt);
}

/* calAddActiveCell2D / calRemoveActiveCell2D (cal2D.c:333-380). */
static inline void orc_add_active(orc_grid* g, int c)
{
	/* the reference takes a per-cell lock around test-and-set */
	if (!__atomic_exchange_n(&g->flags[c], (unsigned char)1, __ATOMIC_RELAXED))
		__atomic_fetch_add(&g->pending, 1, __ATOMIC_RELAXED);
}

static inline void orc_remove_active(orc_grid* g, int c)
{
	if (__atomic_exchange_n(&g->flags[c], (unsigned char)0, __ATOMIC_RELAXED))
		__atomic_fetch_sub(&g->pending, 1, __ATOMIC_RELAXED);
}

/* calUpdateActiveCells2D (cal2D.c:382-440): size += pending; empty set -> NULL
 * list; otherwise a row-major scan of the flags. */
static inline void orc_rebuild_active(orc_grid* g)
{
	free(g->cells);
	g->cells = NULL;
	g->ncells += (int)g->pending;
	g->pending = 0;
	if (g->ncells == 0)
		return;
	g->cells = (int*)malloc(sizeof(int) * (size_t)g->ncells);
	{
		int k = 0;
		size_t c, n = (size_t)g->rows * g->cols;
		for (c = 0; c < n; c++)
			if (g->flags[c])
				g->cells[k++] = (int)c;
	}
}

/* calUpdateSubstate2Dr (cal2D.c:708-713): next -> current, full grid or at the
 * listed active cells (cal2DBuffer.c:103-133, 160-172). */
static inline void orc_update_sub(const orc_grid* g, orc_sub_r* q)
{
	if (g->cells) {
		int k;
#pragma omp parallel for
		for (k = 0; k < g->ncells; k++) {
			int c = g->cells[k];
			if (q->cur[c] != q->nxt[c])
				q->cur[c] = q->nxt[c];
		}
	} else {
		/* chunked parallel memcpy, one chunk per thread (cal2DBuffer.c:103-133) */
		size_t n = (size_t)g->rows * g->cols;
#pragma omp parallel
		{
			int t = omp_get_thread_num(), nt = omp_get_num_threads();
			size_t chunk = n / (size_t)nt, start = chunk * (size_t)t;
			if (t == nt - 1) chunk = n - start;
			memcpy(q->cur + start, q->nxt + start, sizeof(double) * chunk);
		}
	}
}

/* calUpdate2D (cal2D.c:758-776). */
static inline void orc_update(orc_grid* g, orc_sub_r* subs, int nsubs)
{
	int s;
	if (g->active_mode)
		orc_rebuild_active(g);
	for (s = 0; s < nsubs; s++)
		orc_update_sub(g, &subs[s]);
}

/* calInitSubstate2Dr / calInitSubstateNext2Dr (cal2D.c:655-691). */
static inline void orc_init_sub(const orc_grid* g, orc_sub_r* q, double v, int both)
{
	size_t c, n = (size_t)g->rows * g->cols;
	if (g->cells) {
		int k;
		for (k = 0; k < g->ncells; k++) {
			if (both) q->cur[g->cells[k]] = v;
			q->nxt[g->cells[k]] = v;
		}
	} else {
		for (c = 0; c < n; c++) {
			if (both) q->cur[c] = v;
			q->nxt[c] = v;
		}
	}
}

static inline void orc_grid_init(orc_grid* g, int rows, int cols, int active)
{
	g->rows = rows;
	g->cols = cols;
	g->active_mode = active;
	g->flags = active ? (unsigned char*)calloc((size_t)rows * cols, 1) : NULL;
	g->cells = NULL;
	g->ncells = 0;
	g->pending = 0;
}

static inline void orc_grid_free(orc_grid* g)
{
	free(g->flags);
	free(g->cells);
}

#endif
