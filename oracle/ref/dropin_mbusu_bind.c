/* The binding a maintainer adds to run the reference mbusu example
 * (OpenCAL-OMP/examples/calomp_mbusu-glut/source/mbusu.c) on this engine:
 * its elementary process is handed over as source (the text of mbusu.h and
 * mbusu.c, unchanged) and run by the generic device program; the globals the
 * process reads are mirrored into the device program before every
 * transition.  Called once after mbusuCADef() (mbusu_main.c). */
#include <cal3D.h>
#include <calhip3D.h>

#include "mbusu.h"

extern const char calhip_process_source[];
extern float lato, delta_t, rain;
extern float tetas1, tetas2, tetas3, tetas4, tetar1, tetar2, tetar3, tetar4;
extern float alfa1, alfa2, alfa3, alfa4, n1, n2, n3, n4, ks1, ks2, ks3, ks4;
void mbusuTransitionFunction(struct CALModel3D* ca, int i, int j, int k);

#define BIND(v) calhipBindGlobal3D(mbusu, #v, &v, sizeof v)

void calhip_bind_mbusu(void)
{
	calhipAddProcessSource3D(mbusu, calhip_process_source);
	calhipBindProcess3D(mbusu, mbusuTransitionFunction, "mbusuTransitionFunction");
	BIND(mbusu);
	BIND(Q);
	BIND(lato);
	BIND(delta_t);
	BIND(rain);
	BIND(tetas1); BIND(tetas2); BIND(tetas3); BIND(tetas4);
	BIND(tetar1); BIND(tetar2); BIND(tetar3); BIND(tetar4);
	BIND(alfa1); BIND(alfa2); BIND(alfa3); BIND(alfa4);
	BIND(n1); BIND(n2); BIND(n3); BIND(n4);
	BIND(ks1); BIND(ks2); BIND(ks3); BIND(ks4);
}
