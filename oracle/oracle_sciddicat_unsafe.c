/* ORACLE — test infrastructure, not product code.
 *
 * The unsafe SciddicaT of OpenCAL/examples/cal_sciddicaT-unsafe-glut
 * (source/sciddicaTunsafe.c:21-136, ACTIVE_CELLS defined in
 * sciddicaTunsafe.h:17) on the serial OpenCAL engine
 * (OpenCAL/source/cal2D.cpp:311-374 active cells, :651-700 process sweep and
 * update; cal2DUnsafe.cpp:79-151 calGetNext2Dr / calGetNextX2Dr /
 * calSetX2Dr; cal2DRun.cpp:78-113 init and step), driven as its gui.c:84-92
 * drives it: the model's init and update, then `steps` calls of
 * calRunCAStep2D.
 *
 * One step: the transition visits the active cells (or every cell while
 * there are none) and, for each non-eliminated neighbour n, subtracts the
 * outflow f from the cell's next thickness and adds it to the neighbour's
 * (read-modify-writes of the next layer), marking the neighbour active;
 * then OpenCAL's update (row-major rebuild of the active list, next ->
 * current at its cells); then the removal of cells with h <= eps and the
 * update again.  z is single-layer (the init lowers it where h > 0).
 *
 * order 0 visits the cells in the list's row-major order, as the serial
 * engine does.  order 1 visits them in colour phases: (row class, column
 * class) pairs in row-major order of the classes, the class of row i being
 * i mod 3 for i < 3 floor(R/3) and 3 + (i - 3 floor(R/3)) for the last R mod 3
 * rows (columns alike), each phase in row-major order; cells of one phase are
 * at least three rows or three columns apart on the torus, so their
 * read-modify-writes never touch the same cell.  This is the order of the
 * device's colour-phased unsafe processes.
 * With track 0 the model is OpenCAL/examples/cal_sciddicaT-unsafe
 * (source/sciddicaT.c:36-130: the same processes, ACTIVE_CELLS not defined,
 * the model still declared with CAL_OPT_ACTIVE_CELLS, :139): no cell is ever
 * added, the list stays empty, so every sweep and every update covers the
 * whole grid (cal2D.cpp:356-361 and :651-663).
 * Pinned by: the shipped cal_sciddicaT-unsafe-glut/data/width_final.txt
 * (610 x 496, 4000 steps, "%f"), mass conservation on the torus, and the
 * reference programs of both examples built on the reference OpenCAL-OMP
 * engine (one thread: the serial order). */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

typedef struct {
	int rows, cols;
	const double* z;
	double *hc, *hn;
	unsigned char* flag;
	int* cells;
	int n_cells;
	int track; /* ACTIVE_CELLS defined: the processes add and remove cells */
	double r, eps;
} su_t;

static inline int wrap(int x, int n) { return x < 0 ? x + n : (x >= n ? x - n : x); }

/* sciddicaTunsafe.c:21-86 at cell (i, j) */
static void su_transition(su_t* s, int i, int j)
{
	static const int DI[5] = {0, -1, 0, 0, 1}, DJ[5] = {0, 0, -1, 1, 0};
	const int c = i * s->cols + j;
	unsigned char elim[5] = {0, 0, 0, 0, 0};
	int again, count, n, nb[5];
	double avg, m, u[5], f;
	m = s->hc[c] - s->eps;
	u[0] = s->z[c] + s->eps;
	nb[0] = c;
	for (n = 1; n < 5; n++) {
		nb[n] = wrap(i + DI[n], s->rows) * s->cols + wrap(j + DJ[n], s->cols);
		u[n] = s->z[nb[n]] + s->hc[nb[n]];
	}
	do {
		again = 0;
		avg = m;
		count = 0;
		for (n = 0; n < 5; n++)
			if (!elim[n]) {
				avg += u[n];
				count++;
			}
		if (count != 0) avg /= count;
		for (n = 0; n < 5; n++)
			if (avg <= u[n] && !elim[n]) {
				elim[n] = 1;
				again = 1;
			}
	} while (again);
	for (n = 1; n < 5; n++)
		if (!elim[n]) {
			f = (avg - u[n]) * s->r;
			s->hn[c] = s->hn[c] - f;
			s->hn[nb[n]] = s->hn[nb[n]] + f;
			if (s->track) s->flag[nb[n]] = 1;
		}
}

/* calUpdate2D: row-major rebuild (cal2D.cpp:352-374), next -> current at the
 * list's cells (or everywhere while it is empty) */
static void su_update(su_t* s)
{
	int c, n = s->rows * s->cols, k = 0;
	for (c = 0; c < n; c++) {
		/* eight clear flags at a time (the same cells, in the same order) */
		if ((c & 7) == 0 && c + 8 <= n) {
			unsigned long long w;
			memcpy(&w, s->flag + c, 8);
			if (!w) {
				c += 7;
				continue;
			}
		}
		if (s->flag[c]) s->cells[k++] = c;
	}
	s->n_cells = k;
	if (k)
		for (c = 0; c < k; c++) s->hc[s->cells[c]] = s->hn[s->cells[c]];
	else
		memcpy(s->hc, s->hn, (size_t)n * sizeof(double));
}

static int su_class(int x, int n)
{
	const int full = n / 3 * 3;
	return x < full ? x % 3 : 3 + (x - full);
}

static void su_sweep(su_t* s, int order)
{
	int k, n = s->rows * s->cols;
	const int all = s->n_cells == 0;
	const int cnt = all ? n : s->n_cells;
	if (order == 0) {
		for (k = 0; k < cnt; k++) {
			const int c = all ? k : s->cells[k];
			su_transition(s, c / s->cols, c % s->cols);
		}
		return;
	}
	{
		int ci, cj;
		const int nci = 3 + s->rows % 3, ncj = 3 + s->cols % 3;
		const int fi = s->rows / 3 * 3, fj = s->cols / 3 * 3;
		if (all) {
			/* the same phases, each class's rows and columns enumerated */
			int i, j;
			for (ci = 0; ci < nci; ci++)
				for (cj = 0; cj < ncj; cj++)
					for (i = ci < 3 ? ci : fi + ci - 3; i < (ci < 3 ? fi : fi + ci - 2); i += ci < 3 ? 3 : 1)
						for (j = cj < 3 ? cj : fj + cj - 3; j < (cj < 3 ? fj : fj + cj - 2); j += cj < 3 ? 3 : 1)
							su_transition(s, i, j);
			return;
		}
		for (ci = 0; ci < nci; ci++)
			for (cj = 0; cj < ncj; cj++)
				for (k = 0; k < cnt; k++) {
					const int c = all ? k : s->cells[k];
					const int i = c / s->cols, j = c % s->cols;
					if (su_class(i, s->rows) == ci && su_class(j, s->cols) == cj) su_transition(s, i, j);
				}
	}
}

int orc_sciddicat_unsafe_run_track(int rows, int cols, const double* z0, const double* h0, double p_r,
                                   double p_eps, int steps, int order, int track, double* out_h, int* out_active)
{
	su_t s;
	const int n = rows * cols;
	int c, t, k;
	double* z = (double*)malloc((size_t)n * sizeof(double));
	memset(&s, 0, sizeof(s));
	s.rows = rows;
	s.cols = cols;
	s.hc = (double*)malloc((size_t)n * sizeof(double));
	s.hn = (double*)malloc((size_t)n * sizeof(double));
	s.flag = (unsigned char*)calloc((size_t)n, 1);
	s.cells = (int*)malloc((size_t)n * sizeof(int));
	s.r = p_r;
	s.eps = p_eps;
	s.track = track != 0;
	if (!z || !s.hc || !s.hn || !s.flag || !s.cells) return -1;
	memcpy(z, z0, (size_t)n * sizeof(double));
	memcpy(s.hc, h0, (size_t)n * sizeof(double));
	memcpy(s.hn, h0, (size_t)n * sizeof(double));
	/* sciddicaTSimulationInit (:101-123) and its implicit update */
	for (c = 0; c < n; c++)
		if (s.hc[c] > 0.0) {
			z[c] = z[c] - s.hc[c];
			if (s.track) s.flag[c] = 1;
		}
	s.z = z;
	su_update(&s);
	for (t = 0; t < steps; t++) {
		su_sweep(&s, order);
		su_update(&s);
		/* sciddicaT_remove_inactive_cells (:88-94) over the list; an empty
		 * process without ACTIVE_CELLS */
		if (s.track) {
			const int all = s.n_cells == 0, cnt = all ? n : s.n_cells;
			for (k = 0; k < cnt; k++) {
				const int cc = all ? k : s.cells[k];
				if (s.hc[cc] <= s.eps) s.flag[cc] = 0;
			}
		}
		su_update(&s);
	}
	memcpy(out_h, s.hc, (size_t)n * sizeof(double));
	if (out_active) *out_active = s.n_cells;
	free(z);
	free(s.hc);
	free(s.hn);
	free(s.flag);
	free(s.cells);
	return 0;
}

int orc_sciddicat_unsafe_run(int rows, int cols, const double* z0, const double* h0, double p_r, double p_eps,
                             int steps, int order, double* out_h, int* out_active)
{
	return orc_sciddicat_unsafe_run_track(rows, cols, z0, h0, p_r, p_eps, steps, order, 1, out_h, out_active);
}
