// Multigrid V-cycle preconditioner for the ghost-fluid Poisson system.
#pragma once

#include "garm_common.hpp"
#include "garm_pcg.hpp"

#include <string>
#include <vector>

namespace garm {

constexpr int kMgMaxLevels = 12;

// Face weights of the level operators in half precision: the fine weights
// are the face fractions (1 on every liquid face of a box domain, exact) and
// a coarse weight sums 4^l of them (exact up to 2048), so only the partial
// fractions of solid faces and very coarse sums round, to 11 bits -- inside
// a preconditioner whose result flexible CG judges in FP64.  Halving them
// takes 6 of the 28 bytes per cell out of every sweep and restriction.
typedef _Float16 mgw_t;

// Level grids over the bounding box of the unknowns.  Level 0 holds the
// fine operator; level l+1 is its Galerkin coarsening by 2x2x2 aggregates
// (piecewise-constant prolongation).  Per cell: diagonal and the weights
// of its -x, -y, -z faces (operator (A x)_c = diag_c x_c - sum_faces w x_nb);
// b, x are the level's right-hand side and iterate.  FP32 arithmetic; the
// weights are stored as mgw_t.
struct Multigrid {
	int levels = 0;
	int dims[kMgMaxLevels][3] = {};
	size_t cells[kMgMaxLevels] = {};
	float *diag[kMgMaxLevels] = {};
	mgw_t *wx[kMgMaxLevels] = {}, *wy[kMgMaxLevels] = {}, *wz[kMgMaxLevels] = {};
	float *b[kMgMaxLevels] = {}, *x[kMgMaxLevels] = {};
	float *xalt[kMgMaxLevels] = {};   // second iterate buffer: a sweep reads x, writes xalt, and the names swap
	int org[3] = {};          // box origin in cell coordinates
	int gn[3] = {};           // cell-grid size
	int D = 3;
	int tail_first = 0;       // first of the coarse levels run in one workgroup's LDS
	size_t tail_words = 0;    // cell-list words of those levels
	unsigned *tail_list = nullptr;   // their active cells per colour, listed once per operator
	int *tail_count = nullptr;       // active cells per level and colour
	unsigned *box_of = nullptr;      // level-0 cell of every unknown
	uint8_t *tile_active = nullptr;  // per level above the tail: 1 for each sweep tile holding an active cell
	size_t tile_off[kMgMaxLevels] = {};
	size_t tile_cap = 0;
	float *pool = nullptr;    // one allocation for every level array and the lists
	size_t pool_cap = 0;      // floats
	size_t level_floats = 0;  // floats of the level arrays at the start of the pool
	~Multigrid();
	// lays out the levels for a box of the given size holding `unknowns`
	// unknowns (multiple solves reuse the pool)
	bool layout(int D, const int box_org[3], const int box_n[3], const int grid_n[3], int unknowns, std::string &err);
};

// Level 0 from the FP32 rows of the n unknowns, then every coarser level.
void mgBuild(Multigrid &M, const Rows<float> &A, int n, const int *mat2grid, hipStream_t s);
// z = V(r): one symmetric V-cycle (red-black Gauss-Seidel, two sweeps down
// and up) applied to the compact residual r; z is compact, FP64.
void mgApply(Multigrid &M, int n, const double *r, double *z, hipStream_t s);
// The same cycle in two parts (the residual's load and the first level-0
// sweep, then the rest), so a caller can enqueue the head before it knows
// whether the cycle is needed; a head without its rest leaves nothing that
// the next head does not overwrite.  With `box`, r and z are laid out on
// level 0's box (its cells, zero outside the unknowns) instead of compact.
// With `loaded` (box only), level 0's right-hand side and iterate already
// hold r and 0 (the caller wrote them where mgLevel0 says) and the head only
// sweeps; with q (box only), the rest's store also writes the zqBlocks
// per-workgroup partials of z.q into zq, in the order of a grid-stride dot
// product over zqBlocks workgroups of kBlock threads, and z may be null (only
// the partials are formed; z is then level 0's iterate, where mgLevel0 says).
void mgApplyHead(Multigrid &M, int n, const double *r, hipStream_t s, bool box = false, bool loaded = false);
void mgApplyRest(Multigrid &M, int n, double *z, hipStream_t s, bool box = false, const double *q = nullptr, double *zq = nullptr,
	int zqBlocks = 0);
// Where the next cycle reads its right-hand side and starts its iterate on level 0.
void mgLevel0(const Multigrid &M, float **b, float **x);
// Algorithmic HBM bytes of one cycle (its arrays read or written once per launch).
double mgCycleBytes(const Multigrid &M);

} // namespace garm
