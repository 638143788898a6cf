/* ORACLE — test infrastructure, not product code.
 * Conway Life restated from OpenCAL-OMP/examples/calomp_life-glut/source/life.c:16-27
 * (sum of the 8 Moore neighbours via calGetX2D, alive' = sum==3 || (sum==2 &&
 * alive)), Moore order cal2D.c:50-70, toroidal wrap calCommon.h:62, and the
 * single-EP step + full copy of cal2D.c:738-776.
 * Pinned by: the reference build in oracle/_ref (bit-exact), numpy roll-based
 * Life, and closed forms (blinker, block, glider). */
#include "oracle.h"
#include "oracle_engine.h"

int orc_life_run(int rows, int cols, const unsigned char* alive0, int steps,
                 unsigned char* out)
{
	size_t n = (size_t)rows * cols;
	unsigned char* cur = (unsigned char*)malloc(n);
	unsigned char* nxt = (unsigned char*)malloc(n);
	orc_grid g;
	int step;
	g.rows = rows;
	g.cols = cols;
	memcpy(cur, alive0, n);
	memcpy(nxt, alive0, n);
	for (step = 1; step <= steps; step++) {
		int i;
#pragma omp parallel for
		for (i = 0; i < rows; i++) {
			int j, k;
			for (j = 0; j < cols; j++) {
				int sum = 0;
				for (k = 1; k < 9; k++)
					sum += cur[orc_nb_toroidal(&g, i, j, k)];
				nxt[i * cols + j] = (sum == 3 || (sum == 2 && cur[i * cols + j] == 1)) ? 1 : 0;
			}
		}
		memcpy(cur, nxt, n);   /* calUpdate2D: full copy */
	}
	memcpy(out, cur, n);
	free(cur);
	free(nxt);
	return 0;
}
