/* Headless driver of the reference example
 * cal_sciddicaS3hex-unsafe-explicit-glut
 * (OpenCAL/examples/cal_sciddicaS3hex-unsafe-explicit-glut/source): the
 * model file sciddicaS3Hex.c, unchanged (hexagonal torus, the model's own
 * global transition of calApplyElementaryProcess2D and calUpdateSubstate2Dr
 * calls, a flows process that writes its neighbours' next layers), and the
 * step loop of its window program (gui.c:89-107: step++, calRunCAStep2D
 * until the stop condition, then the model's save) without the window.
 * Built twice by this directory's Makefile: on the reference OpenCAL-OMP
 * engine (calomp_s3hex; with OMP_NUM_THREADS=1 its sweeps visit the cells in
 * row-major order) and on this engine with the four elementary processes run
 * by the generic device program (dropin_s3hex, -DCALHIP_DROPIN: the bind
 * calls below, after the model is defined, and calhipSyncHost2D before the
 * model's fields are read directly).
 *
 *   <binary> [dump_prefix [steps]]
 *
 * Reads ./data/dem.txt, ./data/regolith.txt and ./data/source.txt (767 x
 * 925; the example ships the last two and no DEM), runs the example's 2500
 * steps or `steps` of them, writes ./data/width_final.txt as the example
 * does and prints the step count (or, on this engine, the engine's error and
 * exit status 1).  With a prefix it writes the final current layers of h, p,
 * z and d as raw doubles and s as int32, row-major, to
 * <prefix>_{h,p,z,d,s}.bin.  Test infrastructure: the product does not build
 * or run it. */
#include <stdio.h>
#include <stdlib.h>

#include "sciddicaS3Hex.h"

#ifdef CALHIP_DROPIN
#include <calhip2D.h>

extern const char calhip_process_source[];
void s3hexErosion(struct CALModel2D* s3hex, int i, int j);
void s3hexFlowsComputation(struct CALModel2D* s3hex, int i, int j);
void s3hexRomoveInactiveCells(struct CALModel2D* s3hex, int i, int j);
void s3hexEnergyLoss(struct CALModel2D* s3hex, int i, int j);

static void bind_device_program(void)
{
	calhipAddProcessSource2D(s3hex, calhip_process_source);
	calhipBindProcess2D(s3hex, s3hexErosion, "s3hexErosion");
	calhipBindProcess2D(s3hex, s3hexFlowsComputation, "s3hexFlowsComputation");
	calhipBindProcess2D(s3hex, s3hexRomoveInactiveCells, "s3hexRomoveInactiveCells");
	calhipBindProcess2D(s3hex, s3hexEnergyLoss, "s3hexEnergyLoss");
	calhipBindGlobal2D(s3hex, "Q", &Q, sizeof Q);
	calhipBindGlobal2D(s3hex, "P", &P, sizeof P);
}
#endif

static int dump_r(const char* prefix, const char* name, struct CALSubstate2Dr* q)
{
	char path[1024];
	FILE* f;
	int i, j;
	snprintf(path, sizeof path, "%s_%s.bin", prefix, name);
	f = fopen(path, "wb");
	if (!f) return 1;
	for (i = 0; i < s3hex->rows; i++)
		for (j = 0; j < s3hex->columns; j++) {
			const double v = calGet2Dr(s3hex, q, i, j);
			fwrite(&v, sizeof v, 1, f);
		}
	fclose(f);
	return 0;
}

int main(int argc, char** argv)
{
	const int limit = argc > 2 ? atoi(argv[2]) : 0;
	sciddicaTCADef();
#ifdef CALHIP_DROPIN
	bind_device_program();
#endif
	for (;;) {
		s3hexSimulation->step++;
		if (!calRunCAStep2D(s3hexSimulation)) break;
		if (limit > 0 && s3hexSimulation->step >= limit) break;
	}
#ifdef CALHIP_DROPIN
	if (calhipGetLastError()) {
		printf("engine error: %s\n", calhipGetErrorString());
		return 1;
	}
	calhipSyncHost2D(s3hex);
#endif
	sciddicaTSaveConfig();
	printf("steps %d\n", s3hexSimulation->step);
	if (argc > 1) {
		char path[1024];
		FILE* f;
		int i, j;
		if (dump_r(argv[1], "h", Q.h) || dump_r(argv[1], "p", Q.p) || dump_r(argv[1], "z", Q.z) ||
		    dump_r(argv[1], "d", Q.d))
			return 1;
		snprintf(path, sizeof path, "%s_s.bin", argv[1]);
		f = fopen(path, "wb");
		if (!f) return 1;
		for (i = 0; i < s3hex->rows; i++)
			for (j = 0; j < s3hex->columns; j++) {
				const int v = calGet2Di(s3hex, Q.s, i, j);
				fwrite(&v, sizeof v, 1, f);
			}
		fclose(f);
	}
	printf("Simulation terminated\n");
	sciddicaTExit();
	return 0;
}
