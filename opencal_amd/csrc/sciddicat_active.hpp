// SciddicaT active cells in one launch per K steps (K = 1, 2 or 4): the
// kernels of sciddicat_active.hip and their launchers.
//
// Reference: the active-cells SciddicaT
// (OpenCAL/examples/cal_sciddicaT-activecells-glut/source/sciddicaT.c:20-153)
// on OpenCAL-OMP's engine (cal2D.c:333-440 set, :686-713 init / update of
// the set's cells, :717-776 process sweeps and updates).  One step of the
// reference, with the steering's flow reset (calInitSubstateNext2Dr at the
// set's cells) after it, restated per cell.  A is the set at the start of
// the step, F(c) the stored flows of cell c (zero at the cells of A: the
// previous steering reset them there):
//   out(c)  = the minimisation's outflows if c is in A and h(c) > eps,
//             else F(c) (zero in A, otherwise what is stored);
//   recv(c) = a wet cell of A kept c in its minimisation;
//   h'(c)   = h + sum_n (out_{4-n}(neighbour n) - out_{n-1}(c)) for c in A
//             or recv, else h;
//   A'      = { c in A or recv : h'(c) > eps };
//   F'(c)   = 0 in A' (steering), out(c) for cells of A that left the set,
//             else F(c).
// A cell outside A has h <= eps (checked when the mode starts: the init
// adds every cell with h > 0, and a cell leaves only with h' <= eps), so a
// wet cell is always in A.  Flows stored at a cell outside A are almost
// always zero; the cells where they are not ("ghosts": a wet cell that left
// the set, with r >= 1 or at a rounding edge, or flows the host wrote) carry
// a ghost byte and keep their flows in the flow substates' buffers.
//
// Tiles of TILE x TILE cells.  A launch visits a list of tiles; each block
// loads its tile with a halo of 2K cells (the step's reach is two cells:
// the minimisation of a neighbour reads the neighbour's neighbours) and
// advances it K steps in LDS, one cell per thread.  Every state buffer has
// two copies (h and the flows in their substate's two layers, the set flags
// and ghost bytes in two byte arrays): a launch reads one and writes the
// other for every cell of its tiles, so no block reads what another block
// writes.  A tile the launch does not visit holds equal copies.  The next
// launch visits the tiles within KMAX cells of a cell of the new set and
// the tiles whose contents changed (their other copy is now stale).
//
// When the set becomes empty OpenCAL's list is NULL and every process and
// init covers the whole grid: the steering then resets every flow, and
// every later step adds 0.0 to each thickness (turning -0.0 into +0.0) and
// changes nothing else.  A launch that finds the previous one ended with an
// empty set applies that to the whole grid; each launch records in its mask
// whether the set was empty after its last and second-to-last steps, so a
// read of the state can finish a run that emptied within its last launch.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace calhip {
namespace actf {

constexpr int TILE = 16;   // tile rows and columns
constexpr int KMAX = 4;    // most steps per launch

// mask bits of a launch
constexpr unsigned MASK_SET_AFTER_LAST = 1u;    // the set is non-empty after the last step
constexpr unsigned MASK_SET_BEFORE_LAST = 2u;   // ... after the second-to-last step (the start for K = 1)
constexpr unsigned MASK_WHOLE_GRID = 4u;        // the launch found the set empty and covered the grid

// Buffers shared by every launch of the mode.
struct Lists {
	unsigned* tag;      // [n_tiles] epoch of the last list the tile joined
	int* list[2];       // [n_tiles] tiles to visit, by launch parity
	int* ctr;           // [3] list lengths, by launch index mod 3
	unsigned* mask;     // [3] launch masks, by launch index mod 3
	int* chk;           // [2] entry check: cells outside the set with h > eps, cells in the set
};

// One launch: what it reads (_in) and writes (_out).
struct Launch {
	const double* z;
	const double* h_in;
	double* h_out;
	const uint8_t* fl_in;
	uint8_t* fl_out;
	const uint8_t* gh_in;
	uint8_t* gh_out;
	const double* F_in[4];
	double* F_out[4];
	Lists L;
	int par;            // list parity of this launch (reads list[par], fills list[1 - par])
	int rot;            // launch index mod 3
	unsigned e1;        // tag of the tiles listed for the next launch
	int rows, cols, ty, tx, n_tiles;
	int km;             // marking reach (>= the steps of the next launch)
	double r, eps;
};

// Launchers (0 on success).  grid: blocks of the launch (<= n_tiles).
int launch_check(hipStream_t st, const double* h, const uint8_t* fl, const double* const F[4], bool f_zero,
                 uint8_t* gh, int* chk, size_t n, double eps);
int launch_setup(hipStream_t st, const Launch& p);
int launch_fused(hipStream_t st, int K, const Launch& p, int grid);
int launch_leave(hipStream_t st, double* h, const uint8_t* fl, const uint8_t* gh, double* const F[4],
                 const unsigned* mask_last, size_t n);
int launch_normalise(hipStream_t st, double* h, const unsigned* mask_last, size_t n);

// ---- full grid (CAL_NO_OPT) in launches of K steps ----
//
// The same tiles for the full-grid model with the steering's flow reset
// between steps: every cell is processed every step, a cell with h > eps
// (or NaN) sends its outflows, and a dry cell whose neighbours are all dry
// only adds 0.0 to its thickness (turning -0.0 into +0.0).  A launch
// advances the grid K steps: the listed tiles (every tile within reach of a
// wet cell) in LDS as above, every other tile by a streaming copy h + 0.0,
// both from one layer into the other, so each launch reads and writes every
// cell once.  The listed tiles carry the current launch's tag; the launch
// marks the next launch's tiles in the other tag array.
struct PlainLaunch {
	const double* z;
	const double* h_in;
	double* h_out;
	const unsigned* tag_in;    // tiles of this launch: tag_in[t] == e
	unsigned* tag_out;         // tiles of the next launch get tag e + 1
	const int* list_in;
	int* list_out;
	const int* cnt_in;
	int* cnt_out;              // zero on entry
	int* cnt_clear;            // zeroed by the launch (the next launch's cnt_out)
	unsigned* same;            // per tile: e of the first of its consecutive copies (0: none)
	unsigned e;
	int rows, cols, ty, tx, n_tiles;
	int km;                    // marking reach: the next launch's steps
	int grid_comp;             // blocks that visit the list; the rest copy
	double r, eps;
};

constexpr int PLAIN_KMAX = 4;   // most steps per full-grid launch (1 .. 4)
constexpr int PLAIN_PART = 8;   // side of a tile part

// threads per block of the K-step kernel and the column width of one copy block
int plain_threads(int K, bool parts);
int plain_copy_blocks(int K, bool parts, int rows, int cols);
int launch_plain_setup(hipStream_t st, const PlainLaunch& p);
// parts: each listed tile computed as (TILE / PLAIN_PART)^2 parts of
// PLAIN_PART x PLAIN_PART cells, one block each (for short lists: a tile's
// steps spread over several CUs); otherwise one block per tile.
int launch_plain(hipStream_t st, int K, bool parts, const PlainLaunch& p);

}  // namespace actf
}  // namespace calhip
