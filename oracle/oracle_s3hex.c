/* ORACLE — test infrastructure, not product code.
 *
 * The reference example cal_sciddicaS3hex-unsafe-explicit-glut restated
 * (OpenCAL/examples/cal_sciddicaS3hex-unsafe-explicit-glut/source/
 * sciddicaS3Hex.c): the debris flow S3 on OpenCAL's hexagonal neighbourhood
 * (OpenCAL-OMP/source/cal2D.c:75-117, CAL_HEXAGONAL_SHIFT on odd columns),
 * toroidal, with the model's own global transition of four elementary
 * processes, each followed by the explicit updates it lists
 * (sciddicaS3Hex.c:215-229):
 *   erosion (:41-68, with doErosion :21-38), update h, p;
 *   flows (:70-160: the outflows are written into the next layers of the
 *   cell and of its neighbours by read-modify-writes, cal2DUnsafe.c), update
 *   h, p;
 *   removal of inactive cells (:180-186: empty, ACTIVE_CELLS is not defined
 *   in sciddicaS3Hex.h:23);
 *   energy loss (:163-177), update p.
 * With ACTIVE_CELLS undefined the set stays empty, so every process covers
 * the grid.  Init (:192-213): h = p = 0 in both layers.
 *
 * order 0 visits the cells of each process in row-major order, as the
 * reference engine does with one thread.  order 1 visits the flows process
 * in the colour phases of the engine's generic device program: period 3 on
 * each axis (the neighbourhood's radius is 1); on an axis of n cells, cell x
 * < 3 floor(n/3) is in class x mod 3 and each remaining cell in a class of
 * its own; the phases are the class pairs in row-major order (cells of one
 * phase reach disjoint cells, so their order within the phase does not
 * matter).  The other processes touch their own cell only.
 *
 *   orc_s3hex_run(rows, cols, z, d, s, steps, order, h_out, p_out, z_out,
 *                 d_out, s_out)
 * z, d: altitude and regolith depth (single-layer reals); s: source
 * countdown (single-layer int); the outputs are the current layers after
 * `steps` steps of the run loop (the stop condition is the caller's). */
#include "oracle.h"
#include <stdlib.h>
#include <string.h>

/* sciddicaS3Hex.h:12-17 */
#define S3_ADH 0.01
#define S3_RL 0.6
#define S3_R 0.99
#define S3_F 0.1
#define S3_MT 3.5
#define S3_PEF 0.015

/* cal2D.c:92-114: even columns, then odd columns */
static const int HEX[2][7][2] = {
	{{0, 0}, {-1, 1}, {-1, 0}, {-1, -1}, {0, -1}, {1, 0}, {0, 1}},
	{{0, 0}, {0, 1}, {-1, 0}, {0, -1}, {1, -1}, {1, 0}, {1, 1}},
};

typedef struct {
	int rows, cols;
	double *z, *d, *hc, *hn, *pc, *pn;
	int* s;
} s3_t;

/* calGetToroidalX (calCommon.h:62): one wrap */
static int wrap1(int x, int n) { return x < 0 ? n + x : (x > n - 1 ? x - n : x); }

static size_t s3_nb(const s3_t* m, int i, int j, int n)
{
	const int (*t)[2] = HEX[j % 2 == 1];
	return (size_t)wrap1(i + t[n][0], m->rows) * m->cols + wrap1(j + t[n][1], m->cols);
}

/* sciddicaS3Hex.c:21-38 */
static void s3_do_erosion(s3_t* m, size_t c, double erosion_depth)
{
	const double z = m->z[c], d = m->d[c], h = m->hc[c], p = m->pc[c];
	double runup;
	if (h > 0)
		runup = p / h + erosion_depth;
	else
		runup = erosion_depth;
	m->z[c] = z - erosion_depth;
	m->d[c] = d - erosion_depth;
	m->hn[c] = h + erosion_depth;
	m->pn[c] = (h + erosion_depth) * runup;
}

/* sciddicaS3Hex.c:41-68 */
static void s3_erosion(s3_t* m, size_t c)
{
	const double d = m->d[c];
	if (d > 0) {
		const int s = m->s[c];
		if (s < -1) m->s[c] = s + 1;
		if (s == -1) {
			m->s[c] = 0;
			s3_do_erosion(m, c, d);
		}
		const double p = m->pc[c];
		if (p > S3_MT) {
			double erosion_depth = p * S3_PEF;
			if (erosion_depth > d) erosion_depth = d;
			s3_do_erosion(m, c, erosion_depth);
		}
	}
}

/* sciddicaS3Hex.c:70-160 */
static void s3_flows(s3_t* m, int i, int j)
{
	const size_t c = (size_t)i * m->cols + j;
	int elim[7] = {0, 0, 0, 0, 0, 0, 0};
	double u[7], dH[7];
	int n, again, cnt;
	double avg, sum;
	if (m->hc[c] <= S3_ADH) return;
	const double z0 = m->z[c], h0 = m->hc[c];
	const double runup0 = m->pc[c] / h0;
	const double zr0 = z0 + runup0;
	const double mm = runup0;
	u[0] = z0;
	dH[0] = 0;
	for (n = 1; n < 7; n++) {
		const size_t q = s3_nb(m, i, j, n);
		u[n] = m->z[q] + m->hc[q];
		dH[n] = zr0 - u[n];
	}
	for (n = 1; n < 7; n++) elim[n] = dH[n] < S3_F;
	do {
		again = 0;
		avg = mm;
		cnt = 0;
		for (n = 0; n < 7; n++)
			if (!elim[n]) {
				avg += u[n];
				cnt++;
			}
		if (cnt != 0) avg /= cnt;
		for (n = 0; n < 7; n++)
			if (avg <= u[n] && !elim[n]) {
				elim[n] = 1;
				again = 1;
			}
	} while (again);
	sum = 0;
	for (n = 0; n < 7; n++)
		if (!elim[n]) sum += avg - u[n];
	for (n = 1; n < 7; n++)
		if (!elim[n]) {
			const size_t q = s3_nb(m, i, j, n);
			const double f = h0 * ((avg - u[n]) / sum) * S3_R;
			m->hn[c] = m->hn[c] - f;
			m->hn[q] = m->hn[q] + f;
			m->pn[c] = m->pn[c] - runup0 * f;
			m->pn[q] = m->pn[q] + (zr0 - u[n]) * f;
		}
}

/* sciddicaS3Hex.c:163-177 */
static void s3_energy_loss(s3_t* m, size_t c)
{
	if (m->hc[c] <= S3_ADH) return;
	const double h = m->hc[c];
	if (h > S3_ADH) {
		double runup = m->pc[c] / h - S3_RL;
		if (runup < h) runup = h;
		m->pn[c] = h * runup;
	}
}

/* the engine's colour class of x on an axis of n cells, period p */
static int s3_class(int x, int n, int p)
{
	const int full = n / p * p;
	return x < full ? x % p : p + (x - full);
}

int orc_s3hex_run(int rows, int cols, const double* z, const double* d, const int* s, int steps, int order,
                  double* h_out, double* p_out, double* z_out, double* d_out, int* s_out)
{
	const size_t n = (size_t)rows * cols;
	s3_t m;
	int st, i, j;
	size_t c;
	if (rows < 1 || cols < 1 || steps < 0 || (order != 0 && order != 1)) return -1;
	m.rows = rows;
	m.cols = cols;
	m.z = malloc(n * sizeof(double));
	m.d = malloc(n * sizeof(double));
	m.hc = calloc(n, sizeof(double));
	m.hn = calloc(n, sizeof(double));
	m.pc = calloc(n, sizeof(double));
	m.pn = calloc(n, sizeof(double));
	m.s = malloc(n * sizeof(int));
	if (!m.z || !m.d || !m.hc || !m.hn || !m.pc || !m.pn || !m.s) return -2;
	memcpy(m.z, z, n * sizeof(double));
	memcpy(m.d, d, n * sizeof(double));
	memcpy(m.s, s, n * sizeof(int));
	for (st = 0; st < steps; st++) {
		for (c = 0; c < n; c++) s3_erosion(&m, c);
		memcpy(m.hc, m.hn, n * sizeof(double));
		memcpy(m.pc, m.pn, n * sizeof(double));
		if (order == 0) {
			for (i = 0; i < rows; i++)
				for (j = 0; j < cols; j++) s3_flows(&m, i, j);
		} else {
			const int P = 3;
			const int nci = P + rows % P, ncj = P + cols % P;
			int ci, cj;
			for (ci = 0; ci < nci; ci++)
				for (cj = 0; cj < ncj; cj++)
					for (i = 0; i < rows; i++) {
						if (s3_class(i, rows, P) != ci) continue;
						for (j = 0; j < cols; j++)
							if (s3_class(j, cols, P) == cj) s3_flows(&m, i, j);
					}
		}
		memcpy(m.hc, m.hn, n * sizeof(double));
		memcpy(m.pc, m.pn, n * sizeof(double));
		for (c = 0; c < n; c++) s3_energy_loss(&m, c);
		memcpy(m.pc, m.pn, n * sizeof(double));
	}
	memcpy(h_out, m.hc, n * sizeof(double));
	memcpy(p_out, m.pc, n * sizeof(double));
	memcpy(z_out, m.z, n * sizeof(double));
	memcpy(d_out, m.d, n * sizeof(double));
	memcpy(s_out, m.s, n * sizeof(int));
	free(m.z);
	free(m.d);
	free(m.hc);
	free(m.hn);
	free(m.pc);
	free(m.pn);
	free(m.s);
	return 0;
}
