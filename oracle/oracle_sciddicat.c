/* ORACLE — test infrastructure, not product code.
 * SciddicaT restated from
 *   OpenCAL/examples/cal_sciddicaT/source/sciddicaT.c:36-140 (full grid) and
 *   OpenCAL/examples/cal_sciddicaT-activecells-glut/source/sciddicaT.c:20-153
 *   (active cells), run by the loop of OpenCAL-OMP/source/cal2DRun.c:77-139
 *   and cal2D.c:717-776.
 * Pinned by: the reference build in oracle/_ref on seeded inputs (bit-exact),
 * the shipped OpenCAL-OMP/examples/calomp_sciddicaT/data/width_final.txt, and
 * mass conservation on the torus. */
#include "oracle.h"
#include "oracle_engine.h"
#include <omp.h>

enum { Z = 0, H = 1, F0 = 2 };   /* registration order: z, h, f[0..3] */

typedef struct {
	orc_grid g;
	orc_sub_r q[6];
	double r, eps;
} sciddicat_t;

/* sciddicaT_flows_computation (sciddicaT.c:36-88; active variant :20-79). */
static void ep_flows(sciddicat_t* s, int i, int j)
{
	const orc_grid* g = &s->g;
	int c = i * g->cols + j, n, cnt, again;
	unsigned char elim[5] = {0, 0, 0, 0, 0};
	double u[5], m, avg;
	int nb[5];
	if (s->q[H].cur[c] <= s->eps)
		return;
	m = s->q[H].cur[c] - s->eps;
	u[0] = s->q[Z].cur[c] + s->eps;
	for (n = 1; n < 5; n++) {
		nb[n] = orc_nb_toroidal(g, i, j, n);
		u[n] = s->q[Z].cur[nb[n]] + s->q[H].cur[nb[n]];
	}
	do {
		again = 0;
		avg = m;
		cnt = 0;
		for (n = 0; n < 5; n++)
			if (!elim[n]) {
				avg += u[n];
				cnt++;
			}
		if (cnt != 0)
			avg /= cnt;
		for (n = 0; n < 5; n++)
			if (avg <= u[n] && !elim[n]) {
				elim[n] = 1;
				again = 1;
			}
	} while (again);
	for (n = 1; n < 5; n++) {
		if (elim[n]) {
			s->q[F0 + n - 1].nxt[c] = 0.0;
		} else {
			s->q[F0 + n - 1].nxt[c] = (avg - u[n]) * s->r;
			if (g->active_mode)
				orc_add_active(&s->g, nb[n]);
		}
	}
}

/* sciddicaT_width_update (sciddicaT.c:91-101). */
static void ep_width(sciddicat_t* s, int i, int j)
{
	const orc_grid* g = &s->g;
	int c = i * g->cols + j, n;
	double h_next = s->q[H].cur[c];
	for (n = 1; n < 5; n++)
		h_next += s->q[F0 + 4 - n].cur[orc_nb_toroidal(g, i, j, n)] - s->q[F0 + n - 1].cur[c];
	s->q[H].nxt[c] = h_next;
}

/* sciddicaT_remove_inactive_cells (active variant :94-100). */
static void ep_remove(sciddicat_t* s, int i, int j)
{
	int c = i * s->g.cols + j;
	if (s->q[H].cur[c] <= s->eps)
		orc_remove_active(&s->g, c);
}

typedef void (*ep_fn)(sciddicat_t*, int, int);

/* calApplyElementaryProcess2D (cal2D.c:717-734): the listed active cells, or
 * the whole grid when the list is NULL. */
static void apply(sciddicat_t* s, ep_fn ep)
{
	orc_grid* g = &s->g;
	if (g->cells) {
		int k;
#pragma omp parallel for
		for (k = 0; k < g->ncells; k++)
			ep(s, g->cells[k] / g->cols, g->cells[k] % g->cols);
	} else {
		int i;
#pragma omp parallel for
		for (i = 0; i < g->rows; i++) {
			int j;
			for (j = 0; j < g->cols; j++)
				ep(s, i, j);
		}
	}
}

int orc_sciddicat_run(int rows, int cols, const double* z, const double* h,
                      double p_r, double p_eps, int steps, int active,
                      double* out_z, double* out_h, int* active_counts)
{
	sciddicat_t s;
	size_t n = (size_t)rows * cols, c;
	int k, step;
	ep_fn eps_list[3] = {ep_flows, ep_width, ep_remove};
	int n_eps = active ? 3 : 2;

	orc_grid_init(&s.g, rows, cols, active);
	for (k = 0; k < 6; k++)
		orc_sub_alloc(&s.q[k], n);
	/* calLoadSubstate2Dr: current, then copied into next (cal2DIO.c:46-51) */
	memcpy(s.q[Z].cur, z, n * sizeof(double));
	memcpy(s.q[Z].nxt, z, n * sizeof(double));
	memcpy(s.q[H].cur, h, n * sizeof(double));
	memcpy(s.q[H].nxt, h, n * sizeof(double));
	for (k = 2; k < 6; k++) {   /* malloc'd layers: contents defined by init below */
		memset(s.q[k].cur, 0xff, n * sizeof(double));
		memset(s.q[k].nxt, 0xff, n * sizeof(double));
	}

	/* init callback (sciddicaT.c:104-130 / active :106-137) + implicit update */
	for (k = 0; k < 4; k++)
		orc_init_sub(&s.g, &s.q[F0 + k], 0.0, 1);
	s.r = p_r;
	s.eps = p_eps;
	for (c = 0; c < n; c++) {
		double hv = s.q[H].cur[c];
		if (hv > 0.0) {
			s.q[Z].nxt[c] = s.q[Z].cur[c] - hv;
			if (active)
				orc_add_active(&s.g, (int)c);
		}
	}
	orc_update(&s.g, s.q, 6);

	/* calRun2D loop (cal2DRun.c:125-139) */
	for (step = 1; step <= steps; step++) {
		for (k = 0; k < n_eps; k++) {   /* calGlobalTransitionFunction2D */
			apply(&s, eps_list[k]);
			orc_update(&s.g, s.q, 6);
		}
		/* steering (sciddicaT.c:133-140 / active :139-146) */
		for (k = 0; k < 4; k++)
			orc_init_sub(&s.g, &s.q[F0 + k], 0.0, active ? 0 : 1);
		orc_update(&s.g, s.q, 6);
		if (active_counts)
			active_counts[step - 1] = s.g.ncells;
	}

	memcpy(out_z, s.q[Z].cur, n * sizeof(double));
	memcpy(out_h, s.q[H].cur, n * sizeof(double));
	for (k = 0; k < 6; k++)
		orc_sub_free(&s.q[k]);
	orc_grid_free(&s.g);
	return 0;
}

/* ---- single-layer step and its adjoint (for the gradient checks) ---- */

typedef struct {
	double u[5];
	double avg;
	int cnt;
	unsigned char elim[5];
	int nb[5];
	int wet;
} orc_cellflow;

static void cell_flow(int rows, int cols, const double* z, const double* h,
                      double eps, int i, int j, orc_cellflow* cf)
{
	orc_grid g;
	int c = i * cols + j, n, again;
	double m;
	g.rows = rows;
	g.cols = cols;
	memset(cf->elim, 0, 5);
	cf->wet = h[c] > eps;
	for (n = 1; n < 5; n++)
		cf->nb[n] = orc_nb_toroidal(&g, i, j, n);
	cf->nb[0] = c;
	if (!cf->wet)
		return;
	m = h[c] - eps;
	cf->u[0] = z[c] + eps;
	for (n = 1; n < 5; n++)
		cf->u[n] = z[cf->nb[n]] + h[cf->nb[n]];
	do {
		again = 0;
		cf->avg = m;
		cf->cnt = 0;
		for (n = 0; n < 5; n++)
			if (!cf->elim[n]) {
				cf->avg += cf->u[n];
				cf->cnt++;
			}
		if (cf->cnt != 0)
			cf->avg /= cf->cnt;
		for (n = 0; n < 5; n++)
			if (cf->avg <= cf->u[n] && !cf->elim[n]) {
				cf->elim[n] = 1;
				again = 1;
			}
	} while (again);
}

void orc_sciddicat_step(int rows, int cols, const double* z, const double* h,
                        double p_r, double p_eps, double* h_next)
{
	size_t n = (size_t)rows * cols;
	double* f = (double*)calloc(4 * n, sizeof(double));
	int i;
#pragma omp parallel for
	for (i = 0; i < rows; i++) {
		int j, k;
		orc_cellflow cf;
		for (j = 0; j < cols; j++) {
			cell_flow(rows, cols, z, h, p_eps, i, j, &cf);
			if (!cf.wet)
				continue;
			for (k = 1; k < 5; k++)
				f[(size_t)(k - 1) * n + i * cols + j] = cf.elim[k] ? 0.0 : (cf.avg - cf.u[k]) * p_r;
		}
	}
#pragma omp parallel for
	for (i = 0; i < rows; i++) {
		int j, k;
		orc_grid g;
		g.rows = rows;
		g.cols = cols;
		for (j = 0; j < cols; j++) {
			int c = i * cols + j;
			double hn = h[c];
			for (k = 1; k < 5; k++)
				hn += f[(size_t)(4 - k) * n + orc_nb_toroidal(&g, i, j, k)] - f[(size_t)(k - 1) * n + c];
			h_next[c] = hn;
		}
	}
	free(f);
}

int orc_sciddicat_adjoint(int rows, int cols, const double* z, const double* h0,
                          double p_r, double p_eps, int steps, const double* lam_T,
                          double* lam_h0, double* lam_z, double* d_r, double* d_eps)
{
	size_t n = (size_t)rows * cols;
	double* hs = (double*)malloc(sizeof(double) * n * (size_t)(steps + 1));
	double* lam = (double*)malloc(sizeof(double) * n);
	double* lam_prev = (double*)malloc(sizeof(double) * n);
	double dr = 0.0, de = 0.0;
	int t, i, j, k;
	memcpy(hs, h0, sizeof(double) * n);
	for (t = 0; t < steps; t++)
		orc_sciddicat_step(rows, cols, z, hs + n * t, p_r, p_eps, hs + n * (t + 1));
	memcpy(lam, lam_T, sizeof(double) * n);
	memset(lam_z, 0, sizeof(double) * n);
	for (t = steps - 1; t >= 0; t--) {
		const double* h = hs + n * t;
		memcpy(lam_prev, lam, sizeof(double) * n);   /* identity term of h' = h + ... */
		for (i = 0; i < rows; i++)
			for (j = 0; j < cols; j++) {
				orc_cellflow cf;
				int c = i * cols + j;
				double A = 0.0, B;
				cell_flow(rows, cols, z, h, p_eps, i, j, &cf);
				if (!cf.wet)
					continue;
				for (k = 1; k < 5; k++)
					if (!cf.elim[k]) {
						double g_k = lam[cf.nb[k]] - lam[c];
						A += p_r * g_k;
						dr += g_k * (cf.avg - cf.u[k]);
					}
				/* cnt == 0: rounding eliminated every level (h - eps below the
				 * resolution of the level sum), so no flow depends on avg */
				B = cf.cnt ? A / cf.cnt : 0.0;
				lam_prev[c] += B;                       /* m = h - eps */
				de -= B;
				if (!cf.elim[0]) {                      /* u0 = z + eps */
					lam_z[c] += B;
					de += B;
				}
				for (k = 1; k < 5; k++)
					if (!cf.elim[k]) {
						double du = -p_r * (lam[cf.nb[k]] - lam[c]) + B;
						lam_prev[cf.nb[k]] += du;       /* u_k = z_k + h_k */
						lam_z[cf.nb[k]] += du;
					}
			}
		memcpy(lam, lam_prev, sizeof(double) * n);
	}
	memcpy(lam_h0, lam, sizeof(double) * n);
	*d_r = dr;
	*d_eps = de;
	free(hs);
	free(lam);
	free(lam_prev);
	return 0;
}

int orc_set_threads(int nt)
{
	if (nt > 0)
		omp_set_num_threads(nt);
	return omp_get_max_threads();
}
