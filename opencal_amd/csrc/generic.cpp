// Generic device program: a model's own elementary processes, compiled at run
// time for gfx950.
//
// A model that has no hand-written program binds each process it registered
// with calAddElementaryProcess2D to a function of the same name in C source
// it hands over (calhipAddProcessSource2D, calhipBindProcess2D).  The source
// is written against OpenCAL's cell API (calGet2D*, calGetX2D*, calSet2D*,
// active cells, the unsafe operations) and is compiled with hipRTC together
// with the device side of that API (generic_api.inc) and one kernel per
// process that calls it for every cell of the grid, or every cell of the
// active list; a process that writes its neighbours (the unsafe operations)
// runs in colour phases, so that no two threads touch one cell (see
// processes_touching_neighbours below).  Every file-scope variable of the source is also a device
// variable; those the processes read are mirrored from the host
// (calhipBindGlobal2D/3D) before each transition, and pointer-sized words
// that hold the host address of one of the model's substates, or of the
// model, become the address of its device view.  3D models (CALModel3D,
// ModelX::pub3) use the device API of generic_api3.inc and 3D cell indices.
//
// A transition runs the processes in registration order, each followed by
// OpenCAL's update (OpenCAL-OMP/source/cal2D.c:738-776): the active set is
// rebuilt from the flags, then every double-layer substate's next layer is
// copied into current (next and current are equal again).  Unlike the fused
// programs nothing is known about which cells a process writes, so this is
// the reference's own per-process sequence on the device.
//
// Compilation: one code object per distinct source text, cached for the life
// of the process; each model loads its own module (its own globals).
#include <ctype.h>
#include <hip/hiprtc.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#include "engine.hpp"

namespace calhip {

namespace {

const char* kApiCommon =
#include "generic_api_common.inc"
    ;
const char* kApi2 =
#include "generic_api.inc"
    ;
const char* kApi3 =
#include "generic_api3.inc"
    ;

// Host image of the device model view (the struct CALModel2D of
// generic_api.inc): the leading fields of include/cal2D.h's CALModel2D.
struct DevModelView {
	int rows, columns, T, is_safe, OPTIMIZATION;
	CALActiveCells2D A;
	CALCell2D* X;
	int sizeof_X;
	int X_id;
	int nx;
	void** subs;
	int nsubs;
};

// Host image of the 3D device model view (struct CALModel3D of
// generic_api3.inc).
struct DevModelView3 {
	int rows, columns, slices, T, OPTIMIZATION;
	CALActiveCells3D A;
	CALCell3D* X;
	int sizeof_X;
	int X_id;
	int is_safe;
	int nx;
	void** subs;
	int nsubs;
};

// Headers a process source may include.  The cell API is declared by the
// device API above; these declare the host-side calls a model file also makes
// (model definition, run loop, IO, libc) so that the whole file parses.  Host
// code is never reached from a process kernel and is not compiled for the
// GPU.  (hipRTC rejects empty header contents.)
const char* kHdrCal2D = R"H(
struct CALModel2D* calCADef2D(int, int, enum CALNeighborhood2D, enum CALSpaceBoundaryCondition, enum CALOptimization);
void calSetUnsafe2D(struct CALModel2D*);
void calUpdateActiveCells2D(struct CALModel2D*);
struct CALCell2D* calAddNeighbor2D(struct CALModel2D*, int, int);
struct CALSubstate2Db* calAddSubstate2Db(struct CALModel2D*);
struct CALSubstate2Di* calAddSubstate2Di(struct CALModel2D*);
struct CALSubstate2Dr* calAddSubstate2Dr(struct CALModel2D*);
struct CALSubstate2Db* calAddSingleLayerSubstate2Db(struct CALModel2D*);
struct CALSubstate2Di* calAddSingleLayerSubstate2Di(struct CALModel2D*);
struct CALSubstate2Dr* calAddSingleLayerSubstate2Dr(struct CALModel2D*);
CALCallbackFunc2D* calAddElementaryProcess2D(struct CALModel2D*, CALCallbackFunc2D);
void calInitSubstate2Db(struct CALModel2D*, struct CALSubstate2Db*, CALbyte);
void calInitSubstate2Di(struct CALModel2D*, struct CALSubstate2Di*, CALint);
void calInitSubstate2Dr(struct CALModel2D*, struct CALSubstate2Dr*, CALreal);
void calInitSubstateNext2Db(struct CALModel2D*, struct CALSubstate2Db*, CALbyte);
void calInitSubstateNext2Di(struct CALModel2D*, struct CALSubstate2Di*, CALint);
void calInitSubstateNext2Dr(struct CALModel2D*, struct CALSubstate2Dr*, CALreal);
void calUpdateSubstate2Db(struct CALModel2D*, struct CALSubstate2Db*);
void calUpdateSubstate2Di(struct CALModel2D*, struct CALSubstate2Di*);
void calUpdateSubstate2Dr(struct CALModel2D*, struct CALSubstate2Dr*);
void calApplyElementaryProcess2D(struct CALModel2D*, CALCallbackFunc2D);
void calGlobalTransitionFunction2D(struct CALModel2D*);
void calUpdate2D(struct CALModel2D*);
void calFinalize2D(struct CALModel2D*);
)H";
const char* kHdrRun = R"H(
struct CALRun2D {
	struct CALModel2D* ca2D;
	int step;
	int initial_step;
	int final_step;
	enum CALUpdateMode UPDATE_MODE;
	void (*init)(struct CALModel2D*);
	void (*globalTransition)(struct CALModel2D*);
	void (*steering)(struct CALModel2D*);
	CALbyte (*stopCondition)(struct CALModel2D*);
	void (*finalize)(struct CALModel2D*);
};
#define CAL_RUN_LOOP 0
struct CALRun2D* calRunDef2D(struct CALModel2D*, int, int, enum CALUpdateMode);
void calRunAddInitFunc2D(struct CALRun2D*, void (*)(struct CALModel2D*));
void calRunAddGlobalTransitionFunc2D(struct CALRun2D*, void (*)(struct CALModel2D*));
void calRunAddSteeringFunc2D(struct CALRun2D*, void (*)(struct CALModel2D*));
void calRunAddStopConditionFunc2D(struct CALRun2D*, CALbyte (*)(struct CALModel2D*));
void calRunAddFinalizeFunc2D(struct CALRun2D*, void (*)(struct CALModel2D*));
void calRunInitSimulation2D(struct CALRun2D*);
CALbyte calRunCAStep2D(struct CALRun2D*);
void calRunFinalizeSimulation2D(struct CALRun2D*);
void calRun2D(struct CALRun2D*);
void calRunFinalize2D(struct CALRun2D*);
)H";
const char* kHdrIO = R"H(
CALbyte calLoadSubstate2Db(struct CALModel2D*, struct CALSubstate2Db*, char*);
CALbyte calLoadSubstate2Di(struct CALModel2D*, struct CALSubstate2Di*, char*);
CALbyte calLoadSubstate2Dr(struct CALModel2D*, struct CALSubstate2Dr*, char*);
CALbyte calSaveSubstate2Db(struct CALModel2D*, struct CALSubstate2Db*, char*);
CALbyte calSaveSubstate2Di(struct CALModel2D*, struct CALSubstate2Di*, char*);
CALbyte calSaveSubstate2Dr(struct CALModel2D*, struct CALSubstate2Dr*, char*);
)H";
const char* kHdrCalhip = R"H(
int calhipAddProcessSource2D(struct CALModel2D*, const char*);
int calhipBindProcess2D(struct CALModel2D*, CALCallbackFunc2D, const char*);
int calhipBindGlobal2D(struct CALModel2D*, const char*, const void*, size_t);
int calhipSyncHost2D(struct CALModel2D*);
int calhipGetLastError(void);
)H";
const char* kHdrTime = R"H(
typedef long time_t;
time_t time(time_t*);
)H";
const char* kHdrStdio = R"H(
typedef struct calhip_host_file FILE;
FILE* fopen(const char*, const char*);
int fclose(FILE*);
int fprintf(FILE*, const char*, ...);
int fscanf(FILE*, const char*, ...);
)H";
const char* kHdrStdlib = R"H(
void exit(int);
int atoi(const char*);
double atof(const char*);
int rand(void);
void srand(unsigned);
#define RAND_MAX 2147483647
)H";
const char* kHdrNone = "/* declared by the device API */\n";
const char* kHdrCal3D = R"H(
struct CALModel3D* calCADef3D(int, int, int, enum CALNeighborhood3D, enum CALSpaceBoundaryCondition, enum CALOptimization);
void calSetUnsafe3D(struct CALModel3D*);
void calUpdateActiveCells3D(struct CALModel3D*);
struct CALCell3D* calAddNeighbor3D(struct CALModel3D*, int, int, int);
struct CALSubstate3Db* calAddSubstate3Db(struct CALModel3D*);
struct CALSubstate3Di* calAddSubstate3Di(struct CALModel3D*);
struct CALSubstate3Dr* calAddSubstate3Dr(struct CALModel3D*);
struct CALSubstate3Db* calAddSingleLayerSubstate3Db(struct CALModel3D*);
struct CALSubstate3Di* calAddSingleLayerSubstate3Di(struct CALModel3D*);
struct CALSubstate3Dr* calAddSingleLayerSubstate3Dr(struct CALModel3D*);
CALCallbackFunc3D* calAddElementaryProcess3D(struct CALModel3D*, CALCallbackFunc3D);
void calInitSubstate3Db(struct CALModel3D*, struct CALSubstate3Db*, CALbyte);
void calInitSubstate3Di(struct CALModel3D*, struct CALSubstate3Di*, CALint);
void calInitSubstate3Dr(struct CALModel3D*, struct CALSubstate3Dr*, CALreal);
void calInitSubstateNext3Db(struct CALModel3D*, struct CALSubstate3Db*, CALbyte);
void calInitSubstateNext3Di(struct CALModel3D*, struct CALSubstate3Di*, CALint);
void calInitSubstateNext3Dr(struct CALModel3D*, struct CALSubstate3Dr*, CALreal);
void calUpdateSubstate3Db(struct CALModel3D*, struct CALSubstate3Db*);
void calUpdateSubstate3Di(struct CALModel3D*, struct CALSubstate3Di*);
void calUpdateSubstate3Dr(struct CALModel3D*, struct CALSubstate3Dr*);
void calApplyElementaryProcess3D(struct CALModel3D*, CALCallbackFunc3D);
void calGlobalTransitionFunction3D(struct CALModel3D*);
void calUpdate3D(struct CALModel3D*);
void calFinalize3D(struct CALModel3D*);
)H";
const char* kHdrRun3 = R"H(
struct CALRun3D {
	struct CALModel3D* ca3D;
	int step;
	int initial_step;
	int final_step;
	enum CALUpdateMode UPDATE_MODE;
	void (*init)(struct CALModel3D*);
	void (*globalTransition)(struct CALModel3D*);
	void (*steering)(struct CALModel3D*);
	CALbyte (*stopCondition)(struct CALModel3D*);
	void (*finalize)(struct CALModel3D*);
};
#define CAL_RUN_LOOP 0
struct CALRun3D* calRunDef3D(struct CALModel3D*, int, int, enum CALUpdateMode);
void calRunAddInitFunc3D(struct CALRun3D*, void (*)(struct CALModel3D*));
void calRunAddGlobalTransitionFunc3D(struct CALRun3D*, void (*)(struct CALModel3D*));
void calRunAddSteeringFunc3D(struct CALRun3D*, void (*)(struct CALModel3D*));
void calRunAddStopConditionFunc3D(struct CALRun3D*, CALbyte (*)(struct CALModel3D*));
void calRunAddFinalizeFunc3D(struct CALRun3D*, void (*)(struct CALModel3D*));
void calRunInitSimulation3D(struct CALRun3D*);
CALbyte calRunCAStep3D(struct CALRun3D*);
void calRunFinalizeSimulation3D(struct CALRun3D*);
void calRun3D(struct CALRun3D*);
void calRunFinalize3D(struct CALRun3D*);
)H";
const char* kHdrIO3 = R"H(
CALbyte calLoadSubstate3Db(struct CALModel3D*, struct CALSubstate3Db*, char*);
CALbyte calLoadSubstate3Di(struct CALModel3D*, struct CALSubstate3Di*, char*);
CALbyte calLoadSubstate3Dr(struct CALModel3D*, struct CALSubstate3Dr*, char*);
CALbyte calSaveSubstate3Db(struct CALModel3D*, struct CALSubstate3Db*, char*);
CALbyte calSaveSubstate3Di(struct CALModel3D*, struct CALSubstate3Di*, char*);
CALbyte calSaveSubstate3Dr(struct CALModel3D*, struct CALSubstate3Dr*, char*);
)H";
const char* kHdrCalhip3 = R"H(
int calhipAddProcessSource3D(struct CALModel3D*, const char*);
int calhipBindProcess3D(struct CALModel3D*, CALCallbackFunc3D, const char*);
int calhipBindGlobal3D(struct CALModel3D*, const char*, const void*, size_t);
int calhipSyncHost3D(struct CALModel3D*);
int calhipGetLastError(void);
)H";

// Host declarations of the reductions (cal2DReduction.h / cal3DReduction.h).
std::string reduction_header(int dims)
{
	static const char* ops[] = {"Max",       "Min",        "Sum",       "Prod",       "LogicalAnd",
	                            "BinaryAnd", "LogicalOr",  "BinaryOr",  "LogicalXor", "BinaryXor"};
	static const char* suf[] = {"b", "i", "r"};
	static const char* ty[] = {"CALbyte", "CALint", "CALreal"};
	const std::string d = std::to_string(dims);
	std::string h = "\nenum REDUCTION_OPERATION { REDUCTION_NONE = 0, REDUCTION_MAX, REDUCTION_MIN, REDUCTION_SUM, "
	                "REDUCTION_PROD, REDUCTION_LOGICAL_AND, REDUCTION_BINARY_AND, REDUCTION_LOGICAL_OR, "
	                "REDUCTION_BINARY_OR, REDUCTION_LOGICAL_XOR, REDUCTION_BINARY_XOR };\n";
	for (int t = 0; t < 3; ++t) {
		const std::string sig = std::string("(struct CALModel") + d + "D*, struct CALSubstate" + d + "D" + suf[t] + "*";
		h += std::string(ty[t]) + " calReductionOperation" + d + "D" + suf[t] + sig + ", enum REDUCTION_OPERATION);\n";
		for (const char* op : ops) h += std::string(ty[t]) + " calReductionCompute" + op + d + "D" + suf[t] + sig + ");\n";
	}
	return h;
}

struct HeaderSet {
	std::vector<const char*> names, texts;
};

const HeaderSet& headers(int dims)
{
	static const std::string red2 = reduction_header(2), red3 = reduction_header(3);
	static const HeaderSet h2 = {
	    {"cal2D.h", "calCommon.h", "cal2DUnsafe.h", "cal2DBuffer.h", "cal2DIO.h", "cal2DRun.h", "cal2DReduction.h",
	     "calhip2D.h", "math.h", "stdlib.h", "stdio.h", "string.h", "time.h"},
	    {kHdrCal2D, kHdrNone, kHdrNone, kHdrNone, kHdrIO, kHdrRun, red2.c_str(), kHdrCalhip, kHdrNone, kHdrStdlib,
	     kHdrStdio, kHdrNone, kHdrTime}};
	static const HeaderSet h3 = {
	    {"cal3D.h", "calCommon.h", "cal3DUnsafe.h", "cal3DBuffer.h", "cal3DIO.h", "cal3DRun.h", "cal3DReduction.h",
	     "calhip3D.h", "math.h", "stdlib.h", "stdio.h", "string.h", "time.h"},
	    {kHdrCal3D, kHdrNone, kHdrNone, kHdrNone, kHdrIO3, kHdrRun3, red3.c_str(), kHdrCalhip3, kHdrNone, kHdrStdlib,
	     kHdrStdio, kHdrNone, kHdrTime}};
	return dims == 3 ? h3 : h2;
}

std::mutex cache_mu;
std::map<std::string, std::vector<char>>& code_cache()
{
	static auto* c = new std::map<std::string, std::vector<char>>();   // lives with the process
	return *c;
}

int compile(const std::string& text, int dims, std::vector<char>& code)
{
	{
		std::lock_guard<std::mutex> lk(cache_mu);
		auto it = code_cache().find(text);
		if (it != code_cache().end()) {
			code = it->second;
			return 0;
		}
	}
	hiprtcProgram prog;
	const HeaderSet& hs = headers(dims);
	if (hiprtcCreateProgram(&prog, text.c_str(), "calhip_processes.hip", (int)hs.names.size(), hs.texts.data(),
	                        hs.names.data()) != HIPRTC_SUCCESS) {
		set_error(60, "calhip generic program: hiprtcCreateProgram failed");
		return -1;
	}
	// -ffp-contract=off: no fused multiply-add, the processes' arithmetic is
	// the reference's operation by operation.  -fgpu-defer-diag: host code of
	// the model file (which uses host-only globals such as a FILE*) is
	// diagnosed only if a kernel reaches it.
	const char* opts[] = {"--offload-arch=gfx950", "-O3", "-ffp-contract=off", "-std=c++17", "-Xclang",
	                      "-target-feature",        "-Xclang", "-packed-fp32-ops", "-Wno-everything",
	                      "-fgpu-defer-diag"};
	const hiprtcResult r = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
	if (r != HIPRTC_SUCCESS) {
		size_t ls = 0;
		hiprtcGetProgramLogSize(prog, &ls);
		std::string log(ls, '\0');
		if (ls) hiprtcGetProgramLog(prog, &log[0]);
		hiprtcDestroyProgram(&prog);
		set_error(61, "calhip generic program: compiling the process source failed:\n" + log);
		return -1;
	}
	size_t cs = 0;
	hiprtcGetCodeSize(prog, &cs);
	code.resize(cs);
	hiprtcGetCode(prog, code.data());
	hiprtcDestroyProgram(&prog);
	std::lock_guard<std::mutex> lk(cache_mu);
	code_cache()[text] = code;
	return 0;
}

// The text hipRTC compiles: the device API, the model's source (its
// functions become host-and-device functions), and one kernel per process
// calling it for every cell of the grid or of the active list.
std::string process_program_text(const std::string& source, const std::vector<std::string>& names, int dims)
{
	std::string t = kApiCommon;
	t += dims == 3 ? kApi3 : kApi2;
	// A model file's main() is host code: renamed so it is an ordinary
	// function.  Every function of the file becomes host-and-device, every
	// file-scope variable a device variable (its initial value, or the host
	// value mirrored by calhipBindGlobal*), so the file compiles unchanged.
	t += "\n#define main calhip_model_main\n#pragma clang force_cuda_host_device begin\n"
	     "#pragma clang attribute push (__attribute__((device)), apply_to = variable(is_global))\n"
	     "#line 1 \"process source\"\n";
	t += source;
	t += "\n#pragma clang attribute pop\n#pragma clang force_cuda_host_device end\n#undef main\n";
	// Every kernel takes a colour phase: period p (0: no phase, every cell)
	// and the phase's class on each axis.  A phase of the whole grid
	// enumerates its own cells; a phase of the active list visits the list's
	// cells of that class.
	for (size_t k = 0; k < names.size() && dims == 3; ++k) {
		t += "extern \"C\" __global__ void __launch_bounds__(256) calhip_proc_" + std::to_string(k) +
		     "(struct CALModel3D* m, const int* cells, const int* count, int ph_p, int ph_i, int ph_j, int ph_k)\n{\n"
		     "\tconst int cnt = count ? count[0] : 0;\n"
		     "\tconst unsigned long long cols = (unsigned long long)m->columns;\n"
		     "\tconst unsigned long long plane = (unsigned long long)m->rows * cols;\n"
		     "\tconst unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;\n"
		     "\tconst unsigned long long t0 = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;\n"
		     "\tif (ph_p > 0 && cnt == 0) {\n"
		     "\t\tconst unsigned long long ni = calhip_class_size(m->rows, ph_p, ph_i);\n"
		     "\t\tconst unsigned long long nj = calhip_class_size(m->columns, ph_p, ph_j);\n"
		     "\t\tconst unsigned long long nk = calhip_class_size(m->slices, ph_p, ph_k);\n"
		     "\t\tfor (unsigned long long t = t0; t < ni * nj * nk; t += stride) {\n"
		     "\t\t\tconst int i = calhip_class_at(m->rows, ph_p, ph_i, (int)(t / (nj * nk)));\n"
		     "\t\t\tconst int j = calhip_class_at(m->columns, ph_p, ph_j, (int)(t / nk % nj));\n"
		     "\t\t\tconst int kk = calhip_class_at(m->slices, ph_p, ph_k, (int)(t % nk));\n"
		     "\t\t\t" + names[k] + "(m, i, j, kk);\n\t\t}\n\t\treturn;\n\t}\n"
		     "\tconst unsigned long long total = cnt > 0 ? (unsigned long long)cnt : plane * m->slices;\n"
		     "\tfor (unsigned long long t = t0; t < total; t += stride) {\n"
		     "\t\tconst unsigned long long c = cnt > 0 ? (unsigned long long)cells[t] : t;\n"
		     "\t\tconst unsigned long long r = c % plane;\n"
		     "\t\tconst int i = (int)(r / cols), j = (int)(r % cols), kk = (int)(c / plane);\n"
		     "\t\tif (ph_p > 0 && (calhip_class(i, m->rows, ph_p) != ph_i || calhip_class(j, m->columns, ph_p) != ph_j ||"
		     " calhip_class(kk, m->slices, ph_p) != ph_k))\n\t\t\tcontinue;\n"
		     "\t\t" + names[k] + "(m, i, j, kk);\n\t}\n}\n";
	}
	for (size_t k = 0; k < names.size() && dims != 3; ++k) {
		t += "extern \"C\" __global__ void __launch_bounds__(256) calhip_proc_" + std::to_string(k) +
		     "(struct CALModel2D* m, const int* cells, const int* count, int ph_p, int ph_i, int ph_j, int)\n{\n"
		     "\tconst int cnt = count ? count[0] : 0;\n"
		     "\tconst int cols = m->columns;\n"
		     "\tconst unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;\n"
		     "\tconst unsigned long long t0 = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;\n"
		     "\tif (ph_p > 0 && cnt == 0) {\n"
		     "\t\tconst unsigned long long ni = calhip_class_size(m->rows, ph_p, ph_i);\n"
		     "\t\tconst unsigned long long nj = calhip_class_size(cols, ph_p, ph_j);\n"
		     "\t\tfor (unsigned long long t = t0; t < ni * nj; t += stride)\n"
		     "\t\t\t" + names[k] + "(m, calhip_class_at(m->rows, ph_p, ph_i, (int)(t / nj)),"
		     " calhip_class_at(cols, ph_p, ph_j, (int)(t % nj)));\n\t\treturn;\n\t}\n"
		     "\tconst unsigned long long total = cnt > 0 ? (unsigned long long)cnt : "
		     "(unsigned long long)m->rows * cols;\n"
		     "\tfor (unsigned long long t = t0; t < total; t += stride) {\n"
		     "\t\tconst unsigned long long c = cnt > 0 ? (unsigned long long)cells[t] : t;\n"
		     "\t\tconst int i = (int)(c / cols), j = (int)(c % cols);\n"
		     "\t\tif (ph_p > 0 && (calhip_class(i, m->rows, ph_p) != ph_i || calhip_class(j, cols, ph_p) != ph_j))"
		     " continue;\n"
		     "\t\t" + names[k] + "(m, i, j);\n\t}\n}\n";
	}
	return t;
}

// ---- processes that touch other cells ---------------------------------------
//
// The unsafe operations (OpenCAL-OMP/source/cal2DUnsafe.c, cal3DUnsafe.c)
// let a cell write its neighbours' next or current values and read their
// next values.  Run one thread per cell, such a process would race with its
// neighbours' threads (the reference's OpenMP engine takes a lock per access,
// which keeps each access whole but leaves the order, and a
// read-modify-write written as a get and a set, to the scheduler).  A process
// that reaches one of these operations therefore runs in colour phases: for
// a neighbourhood of radius r, colour period p = 2r + 1 on every axis (the
// classes of calhip_class), one launch per class tuple in row-major order of
// the tuples.  Two cells of one phase lie p or more apart on some axis, so
// the cells they reach (radius r each) are disjoint: within a phase no two
// threads touch the same cell, and the result does not depend on the
// scheduling.  It is the serial engine's result with the visits reordered
// by phase; the sums a cell receives from several neighbours may round
// differently from the serial row-major order, by the last bit.
//
// The analysis is textual: the process's function and every function of the
// source it calls, transitively, are searched for calls of the unsafe
// operations; a use of them outside every function body (a macro) marks
// every process.

bool ident_char(char c) { return (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || (c >= '0' && c <= '9') || c == '_'; }

// The source with comments and string and character literals blanked out
// (newlines kept).
std::string strip_comments_and_literals(const std::string& s)
{
	std::string o = s;
	const size_t n = s.size();
	size_t i = 0;
	while (i < n) {
		if (s[i] == '/' && i + 1 < n && s[i + 1] == '/') {
			while (i < n && s[i] != '\n') o[i++] = ' ';
		} else if (s[i] == '/' && i + 1 < n && s[i + 1] == '*') {
			o[i] = o[i + 1] = ' ';
			i += 2;
			while (i < n && !(s[i] == '*' && i + 1 < n && s[i + 1] == '/')) {
				if (s[i] != '\n') o[i] = ' ';
				++i;
			}
			if (i < n) {
				o[i] = o[i + 1] = ' ';
				i += 2;
			}
		} else if (s[i] == '"' || s[i] == '\'') {
			const char q = s[i];
			o[i++] = ' ';
			while (i < n && s[i] != q && s[i] != '\n') {
				if (s[i] == '\\' && i + 1 < n) o[i++] = ' ';
				if (s[i] != '\n') o[i] = ' ';
				++i;
			}
			if (i < n && s[i] == q) o[i++] = ' ';
		} else {
			++i;
		}
	}
	return o;
}

struct Body {
	size_t begin, end;   // the braces
};

// File-scope function definitions: an identifier, a parenthesised list and
// a brace at depth 0.
std::map<std::string, Body> function_bodies(const std::string& s)
{
	std::map<std::string, Body> out;
	int depth = 0;
	std::string pending;
	size_t start = 0;
	for (size_t i = 0; i < s.size(); ++i) {
		if (s[i] == '{') {
			if (depth == 0) {
				pending.clear();
				size_t k = i;
				while (k > 0 && isspace((unsigned char)s[k - 1])) --k;
				if (k > 0 && s[k - 1] == ')') {
					int pd = 0;
					size_t q = k - 1;
					for (;;) {
						if (s[q] == ')') ++pd;
						if (s[q] == '(' && --pd == 0) break;
						if (q == 0) break;
						--q;
					}
					if (pd == 0) {
						size_t e = q;
						while (e > 0 && isspace((unsigned char)s[e - 1])) --e;
						size_t b = e;
						while (b > 0 && ident_char(s[b - 1])) --b;
						if (b < e) pending = s.substr(b, e - b);
					}
				}
				start = i;
			}
			++depth;
		} else if (s[i] == '}' && depth > 0) {
			if (--depth == 0 && !pending.empty()) {
				out[pending] = {start, i};
				pending.clear();
			}
		}
	}
	return out;
}

// calSetX, calSetCurrentX, calInitX, calAddNextX, calGetNextX of 2D or 3D
// substates of any type.
bool unsafe_operation(const std::string& id)
{
	static const char* kPrefix[] = {"calSetX", "calSetCurrentX", "calInitX", "calAddNextX", "calGetNextX"};
	for (const char* p : kPrefix) {
		const size_t n = strlen(p);
		if (id.size() == n + 3 && id.compare(0, n, p) == 0 && (id[n] == '2' || id[n] == '3') && id[n + 1] == 'D' &&
		    (id[n + 2] == 'b' || id[n + 2] == 'i' || id[n + 2] == 'r'))
			return true;
	}
	return false;
}

template <class F>
void for_each_identifier(const std::string& s, size_t begin, size_t end, F f)
{
	size_t i = begin;
	while (i < end) {
		if (ident_char(s[i]) && !(s[i] >= '0' && s[i] <= '9') && (i == 0 || !ident_char(s[i - 1]))) {
			size_t e = i;
			while (e < end && ident_char(s[e])) ++e;
			f(s.substr(i, e - i));
			i = e;
		} else {
			++i;
		}
	}
}

// Per process name: true when it reaches an unsafe operation.
std::vector<char> processes_touching_neighbours(const std::string& source, const std::vector<std::string>& names)
{
	const std::string s = strip_comments_and_literals(source);
	const std::map<std::string, Body> bodies = function_bodies(s);
	// a use outside every body (a macro) marks every process
	bool outside = false, any = false;
	size_t at = 0;
	for_each_identifier(s, 0, s.size(), [&](const std::string& id) { any = any || unsafe_operation(id); });
	std::vector<Body> sorted;
	for (const auto& b : bodies) sorted.push_back(b.second);
	std::sort(sorted.begin(), sorted.end(), [](const Body& x, const Body& y) { return x.begin < y.begin; });
	for (const Body& b : sorted) {
		for_each_identifier(s, at, b.begin, [&](const std::string& id) { outside = outside || unsafe_operation(id); });
		at = b.end + 1;
	}
	for_each_identifier(s, at, s.size(), [&](const std::string& id) { outside = outside || unsafe_operation(id); });
	std::vector<char> out(names.size(), outside ? 1 : 0);
	for (size_t k = 0; k < names.size() && !outside; ++k) {
		std::vector<std::string> todo{names[k]};
		std::set<std::string> seen{names[k]};
		// a process not found as a plain definition may reach any use
		if (bodies.find(names[k]) == bodies.end()) {
			out[k] = any ? 1 : 0;
			continue;
		}
		bool hit = false;
		while (!todo.empty() && !hit) {
			const Body b = bodies.at(todo.back());
			todo.pop_back();
			for_each_identifier(s, b.begin, b.end, [&](const std::string& id) {
				if (unsafe_operation(id)) hit = true;
				else if (bodies.count(id) && seen.insert(id).second)
					todo.push_back(id);
			});
		}
		out[k] = hit ? 1 : 0;
	}
	return out;
}

int grid_for(size_t n)
{
	size_t b = (n + 255) / 256;
	return (int)(b > 16384 ? 16384 : (b < 1 ? 1 : b));
}

bool valid_identifier(const char* s)
{
	if (!s || !*s) return false;
	for (const char* p = s; *p; ++p) {
		const bool ok = (*p >= 'a' && *p <= 'z') || (*p >= 'A' && *p <= 'Z') || *p == '_' ||
		                (p != s && *p >= '0' && *p <= '9');
		if (!ok) return false;
	}
	return true;
}

}  // namespace

struct GenericProgram : Program {
	int dims = 2;   // 3 for a CALModel3D (ModelX::pub3)
	std::string source;
	struct Bound {
		const void* host;   // CALCallbackFunc2D or CALCallbackFunc3D
		std::string name;
	};
	std::vector<Bound> procs;
	struct Global {
		std::string name;
		const void* host;
		size_t bytes;
		void* dev = nullptr;
		std::vector<char> last;   // bytes last uploaded
	};
	std::vector<Global> globals;

	hipModule_t module = nullptr;
	std::vector<hipFunction_t> fns;   // per entry of procs
	std::vector<char> phased;         // per entry of procs: runs in colour phases
	int period = 1;                   // colour period: 2 x the neighbourhood's radius + 1
	std::string built_text;
	void* d_view = nullptr;
	void* d_X = nullptr;
	void** d_subs = nullptr;             // {current, next} per substate, in ModelX::subs order
	std::vector<void*> subs_uploaded;

	const char* name() const override { return "generic"; }

	~GenericProgram() override
	{
		if (module) (void)hipModuleUnload(module);
		if (d_view) (void)hipFree(d_view);
		if (d_X) (void)hipFree(d_X);
		if (d_subs) (void)hipFree(d_subs);
	}

	std::string program_text() const
	{
		std::vector<std::string> names;
		for (const Bound& b : procs) names.push_back(b.name);
		return process_program_text(source, names, dims);
	}

	size_t nsubs_built = 0;   // substates the device view was built for

	int build(ModelX* m)
	{
		const std::string text = program_text();
		if (module && text == built_text && m->subs.size() == nsubs_built) return 0;
		std::vector<char> code;
		if (compile(text, dims, code)) return -1;
		if (module) (void)hipModuleUnload(module);
		module = nullptr;
		fns.clear();
		if (CALHIP_CHECK(hipModuleLoadData(&module, code.data()))) return -1;
		for (size_t k = 0; k < procs.size(); ++k) {
			hipFunction_t f;
			const std::string kn = "calhip_proc_" + std::to_string(k);
			if (CALHIP_CHECK(hipModuleGetFunction(&f, module, kn.c_str()))) return -1;
			fns.push_back(f);
		}
		{
			std::vector<std::string> names;
			for (const Bound& b : procs) names.push_back(b.name);
			phased = processes_touching_neighbours(source, names);
		}
		for (Global& g : globals) {
			hipDeviceptr_t p = nullptr;
			size_t bytes = 0;
			if (hipModuleGetGlobal(&p, &bytes, module, g.name.c_str()) != hipSuccess || bytes != g.bytes) {
				(void)hipGetLastError();
				set_error(62, "calhip generic program: the process source declares no global '" + g.name +
				                  "' of " + std::to_string(g.bytes) + " bytes");
				return -1;
			}
			g.dev = p;
			g.last.clear();
		}
		built_text = text;
		if (d_subs) (void)hipFree(d_subs);
		d_subs = nullptr;
		if (CALHIP_CHECK(hipMalloc(&d_subs, sizeof(void*) * 2 * (m->subs.size() + 1)))) return -1;
		subs_uploaded.clear();
		nsubs_built = m->subs.size();
		if (d_X) (void)hipFree(d_X);
		d_X = nullptr;
		if (dims == 3) return build_view3(m);
		// model view and neighbourhood
		CALModel2D* p = &m->pub;
		// hexagonal tables hold the even and the odd table (cal2D.c:75-166)
		const bool hex = p->X_id == CAL_HEXAGONAL_NEIGHBORHOOD_2D || p->X_id == CAL_HEXAGONAL_NEIGHBORHOOD_ALT_2D;
		const int nx = p->sizeof_X + (hex ? CAL_HEXAGONAL_SHIFT : 0);
		if (CALHIP_CHECK(hipMalloc(&d_X, sizeof(CALCell2D) * (size_t)(nx > 0 ? nx : 1)))) return -1;
		if (nx > 0 && CALHIP_CHECK(hipMemcpy(d_X, p->X, sizeof(CALCell2D) * (size_t)nx, hipMemcpyHostToDevice)))
			return -1;
		int radius = 0;
		for (int n = 0; n < nx; ++n) radius = std::max(radius, std::max(std::abs(p->X[n].i), std::abs(p->X[n].j)));
		period = 2 * radius + 1;
		DevModelView v = {};
		v.rows = p->rows;
		v.columns = p->columns;
		v.T = (int)p->T;
		v.is_safe = (int)p->is_safe;
		v.OPTIMIZATION = (int)p->OPTIMIZATION;
		v.A.flags = (CALbyte*)m->d_flags;
		v.A.num_threads = 1;
		v.X = (CALCell2D*)d_X;
		v.sizeof_X = p->sizeof_X;
		v.X_id = (int)p->X_id;
		v.nx = nx;
		v.subs = d_subs;
		v.nsubs = (int)m->subs.size();
		if (!d_view && CALHIP_CHECK(hipMalloc(&d_view, sizeof(DevModelView)))) return -1;
		if (CALHIP_CHECK(hipMemcpy(d_view, &v, sizeof v, hipMemcpyHostToDevice))) return -1;
		return 0;
	}

	int build_view3(ModelX* m)
	{
		const CALModel3D* p = m->pub3;
		const int nx = p->sizeof_X;
		if (CALHIP_CHECK(hipMalloc(&d_X, sizeof(CALCell3D) * (size_t)(nx > 0 ? nx : 1)))) return -1;
		if (nx > 0 && CALHIP_CHECK(hipMemcpy(d_X, p->X, sizeof(CALCell3D) * (size_t)nx, hipMemcpyHostToDevice)))
			return -1;
		int radius = 0;
		for (int n = 0; n < nx; ++n)
			radius = std::max({radius, std::abs(p->X[n].i), std::abs(p->X[n].j), std::abs(p->X[n].k)});
		period = 2 * radius + 1;
		DevModelView3 v = {};
		v.rows = p->rows;
		v.columns = p->columns;
		v.slices = p->slices;
		v.T = (int)p->T;
		v.OPTIMIZATION = (int)p->OPTIMIZATION;
		v.A.flags = (CALbyte*)m->d_flags;
		v.A.num_threads = 1;
		v.X = (CALCell3D*)d_X;
		v.sizeof_X = p->sizeof_X;
		v.X_id = (int)p->X_id;
		v.is_safe = (int)p->is_safe;
		v.nx = nx;
		v.subs = d_subs;
		v.nsubs = (int)m->subs.size();
		if (!d_view && CALHIP_CHECK(hipMalloc(&d_view, sizeof(DevModelView3)))) return -1;
		if (CALHIP_CHECK(hipMemcpy(d_view, &v, sizeof v, hipMemcpyHostToDevice))) return -1;
		return 0;
	}

	// Substate views and globals, uploaded when they changed.
	int sync(ModelX* m)
	{
		std::vector<void*> t(2 * m->subs.size());
		for (size_t k = 0; k < m->subs.size(); ++k) {
			SubX* s = m->subs[k];
			t[2 * k] = s->d_cur;
			t[2 * k + 1] = s->single ? s->d_cur : s->d_nxt;   // single layer: writes go to its one buffer
		}
		if (t != subs_uploaded && !t.empty()) {
			if (CALHIP_CHECK(hipMemcpyAsync(d_subs, t.data(), sizeof(void*) * t.size(), hipMemcpyHostToDevice,
			                                m->stream)) ||
			    CALHIP_CHECK(hipStreamSynchronize(m->stream)))
				return -1;
			subs_uploaded = t;
		}
		for (Global& g : globals) {
			std::vector<char> b((const char*)g.host, (const char*)g.host + g.bytes);
			// translate host addresses of the model and its substates
			for (size_t off = 0; off + sizeof(void*) <= b.size(); off += sizeof(void*)) {
				uintptr_t w;
				memcpy(&w, b.data() + off, sizeof w);
				if (!w) continue;
				uintptr_t to = 0;
				if (w == (uintptr_t)&m->pub || (m->pub3 && w == (uintptr_t)m->pub3)) {
					to = (uintptr_t)d_view;
				} else {
					for (size_t k = 0; k < m->subs.size(); ++k)
						if (w == (uintptr_t)&m->subs[k]->pub) {
							to = (uintptr_t)(d_subs + 2 * k);
							break;
						}
				}
				if (to) memcpy(b.data() + off, &to, sizeof to);
			}
			if (b != g.last) {
				if (CALHIP_CHECK(hipMemcpyAsync(g.dev, b.data(), b.size(), hipMemcpyHostToDevice, m->stream)) ||
				    CALHIP_CHECK(hipStreamSynchronize(m->stream)))
					return -1;
				g.last.swap(b);
			}
		}
		return 0;
	}

	int prepare(ModelX* m)
	{
		if (build(m) || sync(m)) return -1;
		for (SubX* s : m->subs)
			if (!s->single && materialise_next(m, s)) return -1;   // next == current before a process
		return 0;
	}

	int index_of(const void* ep) const
	{
		for (size_t k = 0; k < procs.size(); ++k)
			if (procs[k].host == ep) return (int)k;
		return -1;
	}

	// Cells of colour class c on an axis of n cells (the device's
	// calhip_class_size).
	static long long class_size(int n, int p, int c) { return c < p ? n / p : 1; }

	int launch(ModelX* m, int k)
	{
		if (m->active() && device_rebuild_active(m)) return -1;
		const int* cells = m->active() ? m->d_cells : nullptr;
		const int* count = m->active() ? m->d_count : nullptr;
		void* view = d_view;
		int ph[4] = {0, 0, 0, 0};   // period, class of rows, columns, slices
		void* args[] = {&view, &cells, &count, &ph[0], &ph[1], &ph[2], &ph[3]};
		const int rows = m->pub3 ? m->pub3->rows : m->pub.rows;
		const int cols = m->pub3 ? m->pub3->columns : m->pub.columns;
		const int slices = m->pub3 ? m->pub3->slices : 1;
		const int p = phased[(size_t)k] ? period : 0;
		// colour classes per axis (one class and no phase for an unphased process)
		const int ni = p ? p + rows % p : 1, nj = p ? p + cols % p : 1, nk = p && m->pub3 ? p + slices % p : 1;
		hipEvent_t ev;
		timer_begin(m, &ev);
		for (int ci = 0; ci < ni; ++ci)
			for (int cj = 0; cj < nj; ++cj)
				for (int ck = 0; ck < nk; ++ck) {
					size_t n = m->cells();
					if (p && !m->active()) {
						n = (size_t)(class_size(rows, p, ci) * class_size(cols, p, cj) *
						             (m->pub3 ? class_size(slices, p, ck) : 1));
						if (n == 0) continue;
					}
					ph[0] = p;
					ph[1] = ci;
					ph[2] = cj;
					ph[3] = ck;
					if (CALHIP_CHECK(hipModuleLaunchKernel(fns[(size_t)k], (unsigned)grid_for(n), 1, 1, 256, 1, 1, 0,
					                                       m->stream, args, nullptr)))
						return -1;
					m->stats.launches++;
				}
		timer_end(m, ev);
		if (m->active()) m->flags_dirty = true;   // the process may have added or removed cells
		return 0;
	}

	// OpenCAL's update after a process (cal2D.c:758-776, cal3D.c:697-715):
	// rebuild the active set, then next -> current for every double-layer
	// substate, at the active cells in active mode (the whole grid while
	// the list is empty).  In active mode the layers may then differ outside
	// the set, as in the reference.
	int update(ModelX* m)
	{
		if (device_rebuild_active(m)) return -1;
		for (SubX* s : m->subs) {
			if (s->single) continue;
			const int e = m->active() ? k::copy_active(m, s->d_cur, s->d_nxt, s->elem)
			                          : k::copy(m, s->d_cur, s->d_nxt, m->cells() * (size_t)s->elem);
			if (e) return -1;
			s->nxt_stale = false;
			s->nxt_dirty = false;
		}
		return 0;
	}

	int step(ModelX* m) override
	{
		if (prepare(m)) return -1;
		const int n_ep = m->pub3 ? m->pub3->num_of_elementary_processes : m->pub.num_of_elementary_processes;
		for (int e = 0; e < n_ep; ++e) {
			const int k = index_of(m->pub3 ? (const void*)m->pub3->elementary_processes[e]
			                               : (const void*)m->pub.elementary_processes[e]);
			if (k < 0) {
				set_error(63, "calhip generic program: a process is not bound to device source");
				return -1;
			}
			if (launch(m, k) || update(m)) return -1;
		}
		return CALHIP_CHECK(hipGetLastError());
	}

	// calApplyElementaryProcess2D: the sweep alone; the caller's calUpdate2D
	// flips the layers (each next layer was equal to current before it).
	int apply(ModelX* m, const void* ep) override
	{
		const int k = index_of(ep);
		if (k < 0) {
			set_error(63, "calhip generic program: the process is not bound to device source");
			return -1;
		}
		if (prepare(m) || launch(m, k)) return -1;
		for (SubX* s : m->subs)
			if (!s->single) s->nxt_dirty = true;
		return CALHIP_CHECK(hipGetLastError());
	}
};

namespace {

GenericProgram* generic_of(ModelX* m, bool create)
{
	auto* g = dynamic_cast<GenericProgram*>(m->prog);
	if (g || !create) return g;
	if (ensure_host(m)) return nullptr;   // leaving another program: its state is written back first
	delete m->prog;
	g = new GenericProgram();
	g->dims = m->pub3 ? 3 : 2;
	m->prog = g;
	m->prog_eps.clear();
	m->prog_eps3.clear();
	m->prog_steering = nullptr;
	for (SubX* s : m->subs) s->host_only = false;
	return g;
}

}  // namespace

}  // namespace calhip

using namespace calhip;

namespace {

int add_source(ModelX* m, const char* source, const char* fn)
{
	if (!m || !source) {
		set_error(64, std::string(fn) + ": missing argument");
		return -1;
	}
	GenericProgram* g = generic_of(m, true);
	if (!g) return -1;
	g->source += source;
	g->source += "\n";
	return 0;
}

int bind_process(ModelX* m, const void* process, const char* device_name, const char* fn)
{
	if (!m || !process || !valid_identifier(device_name)) {
		set_error(64, std::string(fn) + ": give the process and the C name of its device function");
		return -1;
	}
	GenericProgram* g = generic_of(m, true);
	if (!g) return -1;
	const int k = g->index_of(process);
	if (k >= 0)
		g->procs[(size_t)k].name = device_name;
	else
		g->procs.push_back({process, device_name});
	// the program covers the model's processes in registration order
	m->prog_eps.clear();
	m->prog_eps3.clear();
	if (m->pub3) {
		const CALModel3D* p = m->pub3;
		for (int e = 0; e < p->num_of_elementary_processes; ++e)
			if (g->index_of((const void*)p->elementary_processes[e]) >= 0)
				m->prog_eps3.push_back(p->elementary_processes[e]);
	} else {
		const CALModel2D* p = &m->pub;
		for (int e = 0; e < p->num_of_elementary_processes; ++e)
			if (g->index_of((const void*)p->elementary_processes[e]) >= 0)
				m->prog_eps.push_back(p->elementary_processes[e]);
	}
	return 0;
}

int bind_global(ModelX* m, const char* name, const void* host_address, size_t bytes, const char* fn)
{
	if (!m || !valid_identifier(name) || !host_address || bytes == 0) {
		set_error(64, std::string(fn) + ": give the global's C name, its host address and its size");
		return -1;
	}
	GenericProgram* g = generic_of(m, true);
	if (!g) return -1;
	for (auto& gl : g->globals)
		if (gl.name == name) {
			gl.host = host_address;
			gl.bytes = bytes;
			gl.last.clear();
			return 0;
		}
	g->globals.push_back({name, host_address, bytes, nullptr, {}});
	g->built_text.clear();   // globals are looked up when the module loads
	return 0;
}

int check_source(const char* source, const char* const* names, int n_names, char* log, size_t log_bytes, int dims,
                 const char* fn)
{
	if (log && log_bytes) log[0] = '\0';
	if (!source || n_names < 0 || (n_names > 0 && !names)) {
		set_error(64, std::string(fn) + ": missing argument");
		return -1;
	}
	std::vector<std::string> nm;
	for (int k = 0; k < n_names; ++k) {
		if (!valid_identifier(names[k])) {
			set_error(64, std::string(fn) + ": process names must be C identifiers");
			return -1;
		}
		nm.push_back(names[k]);
	}
	std::vector<char> code;
	if (compile(process_program_text(std::string(source) + "\n", nm, dims), dims, code)) {
		if (log && log_bytes) {
			const char* e = calhipGetErrorString();
			strncpy(log, e ? e : "", log_bytes - 1);
			log[log_bytes - 1] = '\0';
		}
		return -1;
	}
	return 0;
}

}  // namespace

extern "C" int calhipAddProcessSource2D(struct CALModel2D* ca2D, const char* source)
{
	return add_source(ca2D ? X(ca2D) : nullptr, source, "calhipAddProcessSource2D");
}

extern "C" int calhipBindProcess2D(struct CALModel2D* ca2D, CALCallbackFunc2D process, const char* device_name)
{
	return bind_process(ca2D ? X(ca2D) : nullptr, (const void*)process, device_name, "calhipBindProcess2D");
}

extern "C" int calhipBindGlobal2D(struct CALModel2D* ca2D, const char* name, const void* host_address, size_t bytes)
{
	return bind_global(ca2D ? X(ca2D) : nullptr, name, host_address, bytes, "calhipBindGlobal2D");
}

extern "C" int calhipCheckProcessSource2D(const char* source, const char* const* names, int n_names, char* log,
                                          size_t log_bytes)
{
	return check_source(source, names, n_names, log, log_bytes, 2, "calhipCheckProcessSource2D");
}

extern "C" int calhipAddProcessSource3D(struct CALModel3D* ca3D, const char* source)
{
	return add_source(ca3D ? &X3(ca3D)->core : nullptr, source, "calhipAddProcessSource3D");
}

extern "C" int calhipBindProcess3D(struct CALModel3D* ca3D, CALCallbackFunc3D process, const char* device_name)
{
	return bind_process(ca3D ? &X3(ca3D)->core : nullptr, (const void*)process, device_name, "calhipBindProcess3D");
}

extern "C" int calhipBindGlobal3D(struct CALModel3D* ca3D, const char* name, const void* host_address, size_t bytes)
{
	return bind_global(ca3D ? &X3(ca3D)->core : nullptr, name, host_address, bytes, "calhipBindGlobal3D");
}

extern "C" int calhipCheckProcessSource3D(const char* source, const char* const* names, int n_names, char* log,
                                          size_t log_bytes)
{
	return check_source(source, names, n_names, log, log_bytes, 3, "calhipCheckProcessSource3D");
}

extern "C" int calhipProcessRunsInPhases(const char* source, const char* device_name)
{
	if (!source || !valid_identifier(device_name)) {
		set_error(64, "calhipProcessRunsInPhases: give the source and the C name of the process");
		return -1;
	}
	return processes_touching_neighbours(source, {device_name})[0] ? 1 : 0;
}
