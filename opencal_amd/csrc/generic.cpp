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
// active list.  Globals the processes read are declared CALHIP_GLOBAL and
// mirrored from the host (calhipBindGlobal2D); pointer-sized words that hold
// the host address of one of the model's substates, or of the model, become
// the address of its device view.
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
#include <hip/hiprtc.h>
#include <string.h>

#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "engine.hpp"

namespace calhip {

namespace {

const char* kApi =
#include "generic_api.inc"
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
const char* kHeaderNames[] = {"cal2D.h",    "calCommon.h", "cal2DUnsafe.h", "cal2DBuffer.h", "cal2DIO.h", "cal2DRun.h",
                              "calhip2D.h", "math.h",      "stdlib.h",      "stdio.h",       "string.h",  "time.h"};
const char* kHeaderText[] = {kHdrCal2D,  kHdrNone, kHdrNone,   kHdrNone,  kHdrIO,   kHdrRun,
                             kHdrCalhip, kHdrNone, kHdrStdlib, kHdrStdio, kHdrNone, kHdrTime};
constexpr int kNumHeaders = sizeof(kHeaderNames) / sizeof(kHeaderNames[0]);
static_assert(sizeof(kHeaderText) / sizeof(kHeaderText[0]) == kNumHeaders, "one text per header");

std::mutex cache_mu;
std::map<std::string, std::vector<char>>& code_cache()
{
	static auto* c = new std::map<std::string, std::vector<char>>();   // lives with the process
	return *c;
}

int compile(const std::string& text, std::vector<char>& code)
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
	if (hiprtcCreateProgram(&prog, text.c_str(), "calhip_processes.hip", kNumHeaders, kHeaderText, kHeaderNames) !=
	    HIPRTC_SUCCESS) {
		set_error(60, "calhip generic program: hiprtcCreateProgram failed");
		return -1;
	}
	// -ffp-contract=off: no fused multiply-add, the processes' arithmetic is
	// the reference's operation by operation
	const char* opts[] = {"--offload-arch=gfx950", "-O3", "-ffp-contract=off", "-std=c++17", "-Xclang",
	                      "-target-feature",        "-Xclang", "-packed-fp32-ops", "-Wno-everything"};
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
std::string process_program_text(const std::string& source, const std::vector<std::string>& names)
{
	std::string t = kApi;
	// a model file's main() is host code: renamed so it is an ordinary function
	t += "\n#define main calhip_model_main\n#pragma clang force_cuda_host_device begin\n#line 1 \"process source\"\n";
	t += source;
	t += "\n#pragma clang force_cuda_host_device end\n#undef main\n";
	for (size_t k = 0; k < names.size(); ++k) {
		t += "extern \"C\" __global__ void __launch_bounds__(256) calhip_proc_" + std::to_string(k) +
		     "(struct CALModel2D* m, const int* cells, const int* count)\n{\n"
		     "\tconst int cnt = count ? count[0] : 0;\n"
		     "\tconst int cols = m->columns;\n"
		     "\tconst unsigned long long total = cnt > 0 ? (unsigned long long)cnt : "
		     "(unsigned long long)m->rows * cols;\n"
		     "\tfor (unsigned long long t = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;"
		     " t += (unsigned long long)gridDim.x * blockDim.x) {\n"
		     "\t\tconst unsigned long long c = cnt > 0 ? (unsigned long long)cells[t] : t;\n"
		     "\t\t" + names[k] + "(m, (int)(c / cols), (int)(c % cols));\n\t}\n}\n";
	}
	return t;
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
	std::string source;
	struct Bound {
		CALCallbackFunc2D host;
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
	std::string built_text;
	void* d_view = nullptr;
	CALCell2D* d_X = nullptr;
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
		return process_program_text(source, names);
	}

	size_t nsubs_built = 0;   // substates the device view was built for

	int build(ModelX* m)
	{
		const std::string text = program_text();
		if (module && text == built_text && m->subs.size() == nsubs_built) return 0;
		std::vector<char> code;
		if (compile(text, code)) return -1;
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
		for (Global& g : globals) {
			hipDeviceptr_t p = nullptr;
			size_t bytes = 0;
			if (hipModuleGetGlobal(&p, &bytes, module, g.name.c_str()) != hipSuccess || bytes != g.bytes) {
				(void)hipGetLastError();
				set_error(62, "calhip generic program: the process source declares no CALHIP_GLOBAL '" + g.name +
				                  "' of " + std::to_string(g.bytes) + " bytes");
				return -1;
			}
			g.dev = p;
			g.last.clear();
		}
		built_text = text;
		// model view and neighbourhood
		CALModel2D* p = &m->pub;
		// hexagonal tables hold the even and the odd table (cal2D.c:75-166)
		const bool hex = p->X_id == CAL_HEXAGONAL_NEIGHBORHOOD_2D || p->X_id == CAL_HEXAGONAL_NEIGHBORHOOD_ALT_2D;
		const int nx = p->sizeof_X + (hex ? CAL_HEXAGONAL_SHIFT : 0);
		if (d_X) (void)hipFree(d_X);
		d_X = nullptr;
		if (CALHIP_CHECK(hipMalloc(&d_X, sizeof(CALCell2D) * (size_t)(nx > 0 ? nx : 1)))) return -1;
		if (nx > 0 && CALHIP_CHECK(hipMemcpy(d_X, p->X, sizeof(CALCell2D) * (size_t)nx, hipMemcpyHostToDevice)))
			return -1;
		DevModelView v = {};
		v.rows = p->rows;
		v.columns = p->columns;
		v.T = (int)p->T;
		v.is_safe = (int)p->is_safe;
		v.OPTIMIZATION = (int)p->OPTIMIZATION;
		v.A.flags = (CALbyte*)m->d_flags;
		v.A.num_threads = 1;
		v.X = d_X;
		v.sizeof_X = p->sizeof_X;
		v.X_id = (int)p->X_id;
		v.nx = nx;
		if (d_subs) (void)hipFree(d_subs);
		d_subs = nullptr;
		if (CALHIP_CHECK(hipMalloc(&d_subs, sizeof(void*) * 2 * (m->subs.size() + 1)))) return -1;
		subs_uploaded.clear();
		v.subs = d_subs;
		v.nsubs = (int)m->subs.size();
		nsubs_built = m->subs.size();
		if (!d_view && CALHIP_CHECK(hipMalloc(&d_view, sizeof(DevModelView)))) return -1;
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
				if (w == (uintptr_t)&m->pub) {
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

	int index_of(CALCallbackFunc2D ep) const
	{
		for (size_t k = 0; k < procs.size(); ++k)
			if (procs[k].host == ep) return (int)k;
		return -1;
	}

	int launch(ModelX* m, int k)
	{
		if (m->active() && device_rebuild_active(m)) return -1;
		const int* cells = m->active() ? m->d_cells : nullptr;
		const int* count = m->active() ? m->d_count : nullptr;
		void* view = d_view;
		void* args[] = {&view, &cells, &count};
		hipEvent_t ev;
		timer_begin(m, &ev);
		if (CALHIP_CHECK(hipModuleLaunchKernel(fns[(size_t)k], (unsigned)grid_for(m->cells()), 1, 1, 256, 1, 1, 0,
		                                       m->stream, args, nullptr)))
			return -1;
		timer_end(m, ev);
		m->stats.launches++;
		if (m->active()) m->flags_dirty = true;   // the process may have added or removed cells
		return 0;
	}

	// OpenCAL's update after a process: rebuild, then next -> current.
	int update(ModelX* m)
	{
		if (device_rebuild_active(m)) return -1;
		for (SubX* s : m->subs) {
			if (s->single) continue;
			if (k::copy(m, s->d_cur, s->d_nxt, m->cells() * (size_t)s->elem)) return -1;
			s->nxt_stale = false;
			s->nxt_dirty = false;
		}
		return 0;
	}

	int step(ModelX* m) override
	{
		if (prepare(m)) return -1;
		CALModel2D* p = &m->pub;
		for (int e = 0; e < p->num_of_elementary_processes; ++e) {
			const int k = index_of(p->elementary_processes[e]);
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
	int apply(ModelX* m, CALCallbackFunc2D ep) override
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
	m->prog = g;
	m->prog_eps.clear();
	m->prog_steering = nullptr;
	for (SubX* s : m->subs) s->host_only = false;
	return g;
}

}  // namespace

}  // namespace calhip

using namespace calhip;

extern "C" int calhipAddProcessSource2D(struct CALModel2D* ca2D, const char* source)
{
	if (!ca2D || !source) {
		set_error(64, "calhipAddProcessSource2D: missing argument");
		return -1;
	}
	GenericProgram* g = generic_of(X(ca2D), true);
	if (!g) return -1;
	g->source += source;
	g->source += "\n";
	return 0;
}

extern "C" int calhipBindProcess2D(struct CALModel2D* ca2D, CALCallbackFunc2D process, const char* device_name)
{
	if (!ca2D || !process || !valid_identifier(device_name)) {
		set_error(64, "calhipBindProcess2D: give the process and the C name of its device function");
		return -1;
	}
	ModelX* m = X(ca2D);
	GenericProgram* g = generic_of(m, true);
	if (!g) return -1;
	const int k = g->index_of(process);
	if (k >= 0)
		g->procs[(size_t)k].name = device_name;
	else
		g->procs.push_back({process, device_name});
	// the program covers the model's processes in registration order
	m->prog_eps.clear();
	for (int e = 0; e < ca2D->num_of_elementary_processes; ++e)
		if (g->index_of(ca2D->elementary_processes[e]) >= 0) m->prog_eps.push_back(ca2D->elementary_processes[e]);
	return 0;
}

extern "C" int calhipBindGlobal2D(struct CALModel2D* ca2D, const char* name, const void* host_address, size_t bytes)
{
	if (!ca2D || !valid_identifier(name) || !host_address || bytes == 0) {
		set_error(64, "calhipBindGlobal2D: give the global's C name, its host address and its size");
		return -1;
	}
	GenericProgram* g = generic_of(X(ca2D), true);
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

extern "C" int calhipCheckProcessSource2D(const char* source, const char* const* names, int n_names, char* log,
                                          size_t log_bytes)
{
	if (log && log_bytes) log[0] = '\0';
	if (!source || n_names < 0 || (n_names > 0 && !names)) {
		set_error(64, "calhipCheckProcessSource2D: missing argument");
		return -1;
	}
	std::vector<std::string> nm;
	for (int k = 0; k < n_names; ++k) {
		if (!valid_identifier(names[k])) {
			set_error(64, "calhipCheckProcessSource2D: process names must be C identifiers");
			return -1;
		}
		nm.push_back(names[k]);
	}
	std::vector<char> code;
	if (compile(process_program_text(std::string(source) + "\n", nm), code)) {
		if (log && log_bytes) {
			const char* e = calhipGetErrorString();
			strncpy(log, e ? e : "", log_bytes - 1);
			log[log_bytes - 1] = '\0';
		}
		return -1;
	}
	return 0;
}
