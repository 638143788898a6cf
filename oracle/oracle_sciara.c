/* ORACLE — test infrastructure, not product code.  SCIARA-fv2 restatement. */
#include "oracle.h"
#include "portable_math.h"

int orc_sciara_run(int rows, int cols, const double* z0, double cell_size,
                   const double* params, int n_vents, const int* vent_rc,
                   int n_rates, const double* rates, unsigned emission_time,
                   int steps, int use_active, int math_mode,
                   double* out_z, double* out_h, double* out_T, double* out_Msl,
                   int* out_steps, int* out_active)
{
	(void)rows; (void)cols; (void)z0; (void)cell_size; (void)params; (void)n_vents;
	(void)vent_rc; (void)n_rates; (void)rates; (void)emission_time; (void)steps;
	(void)use_active; (void)math_mode; (void)out_z; (void)out_h; (void)out_T;
	(void)out_Msl; (void)out_steps; (void)out_active;
	return -1;
}
