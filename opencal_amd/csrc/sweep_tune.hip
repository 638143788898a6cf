// Launch shape of the fused SciddicaT sweep and its autotuner.
//
// The sweep (step_fused2) gives every wave one 124-column strip of a chunk of
// `rpw` rows.  Short chunks make more waves (needed to fill 256 CUs on small
// grids, and they bound the tail of the few waves that meet wet cells) but
// re-read more halo rows; long chunks stream better on large grids.  The
// autotuner times each candidate chunk length on the caller's stream with HIP
// events (a few launches each, writing into a buffer the caller is about to
// overwrite anyway) and keeps the fastest per grid shape for the life of the
// process.  Every candidate computes the same bits, so the choice never
// changes a result.
#include <stdlib.h>

#include <map>
#include <mutex>
#include <utility>

#include "engine.hpp"
#include "device_common.hpp"
#include "sciddicat_device.hpp"

namespace calhip {
namespace sct {

SweepGeom sweep_geom(int rows, int cols, int rpw)
{
	SweepGeom g;
	g.rpw = rpw;
	g.n_strips = (cols + STRIP2_OUT - 1) / STRIP2_OUT;
	const int sgroups = (g.n_strips + FUSED_WAVES - 1) / FUSED_WAVES;
	g.n_logical = sgroups * ((rows + rpw - 1) / rpw);
	g.per_xcd = (g.n_logical + 7) / 8;
	g.blocks = g.per_xcd * 8;
	return g;
}

void launch_sweep(hipStream_t st, int rows, int cols, const double* z, const double* h, double* hn, double* trash,
                  double r, double eps, int rpw)
{
	const SweepGeom g = sweep_geom(rows, cols, rpw);
	if (cols % 2 == 0)
		hipLaunchKernelGGL((step_fused2<1, true>), dim3(g.blocks), dim3(64 * FUSED_WAVES), 0, st, z, h, hn, trash,
		                   rows, cols, g.n_strips, g.n_logical, g.per_xcd, g.rpw, r, eps);
	else
		hipLaunchKernelGGL((step_fused2<1, false>), dim3(g.blocks), dim3(64 * FUSED_WAVES), 0, st, z, h, hn, trash,
		                   rows, cols, g.n_strips, g.n_logical, g.per_xcd, g.rpw, r, eps);
}

namespace {
std::mutex tune_mu;
std::map<std::pair<int, int>, int> tuned;
}  // namespace

int sweep_rpw(hipStream_t st, int rows, int cols, const double* z, const double* h, double* scratch, double* trash,
              double r, double eps)
{
	if (const char* e = getenv("CALHIP_SCIDDICAT_ROWS_PER_WAVE")) {
		const int v = atoi(e);
		if (v > 0) return v;
	}
	std::lock_guard<std::mutex> lk(tune_mu);
	const auto key = std::make_pair(rows, cols);
	auto it = tuned.find(key);
	if (it != tuned.end()) return it->second;
	static const int cand[] = {2, 4, 8};
	hipEvent_t e0 = nullptr, e1 = nullptr;
	int best = 8;
	float best_ms = 0.f;
	if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
		if (e0) (void)hipEventDestroy(e0);
		return best;   // no timing: the default shape
	}
	for (int rpw : cand) {
		launch_sweep(st, rows, cols, z, h, scratch, trash, r, eps, rpw);   // warm
		(void)hipEventRecord(e0, st);
		for (int k = 0; k < 3; ++k) launch_sweep(st, rows, cols, z, h, scratch, trash, r, eps, rpw);
		(void)hipEventRecord(e1, st);
		float ms = 0.f;
		if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) continue;
		if (best_ms == 0.f || ms < best_ms) {
			best_ms = ms;
			best = rpw;
		}
	}
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	tuned[key] = best;
	return best;
}

}  // namespace sct
}  // namespace calhip
