// Batched SciddicaT ensemble.
#include "engine.hpp"

using namespace calhip;

extern "C" int calhipSciddicaTBatchRun(int, const int*, const int*, const double*, const double*,
                                       const double* const*, const double* const*, int, double* const*, double*)
{
	set_error(90, "calhipSciddicaTBatchRun: not available in this build");
	return -1;
}
