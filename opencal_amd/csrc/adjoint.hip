// SciddicaT adjoint.
#include "engine.hpp"

using namespace calhip;

extern "C" int calhipSciddicaTForwardDev(int, int, const double*, const double*, double, double, int, int, double*,
                                         double*, double*)
{
	set_error(91, "calhipSciddicaTForwardDev: not available in this build");
	return -1;
}

extern "C" int calhipSciddicaTAdjointDev(int, int, const double*, const double*, double, double, int, int, double*,
                                         double*, double*, double*, double*)
{
	set_error(92, "calhipSciddicaTAdjointDev: not available in this build");
	return -1;
}
