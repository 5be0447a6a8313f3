// Pressure projection state (unknown maps, operator rows, solver workspace).
#pragma once

#include "garm_internal.hpp"
#include "garm_pcg.hpp"
#include "garm_learned.hpp"
#include "../../include/garm_gpu.h"

#include <string>

namespace garm {

// Three face arrays with their grids, passed by value.
struct FaceFracs {
	const double *a[3];
	GridDev g[3];
};

struct Projection {
	uint8_t *flags = nullptr;                  // 1 where a cell is an unknown
	int *grid2mat = nullptr, *mat2grid = nullptr;
	int *tile_count = nullptr, *tile_base = nullptr;   // unknown numbering: per-tile counts, offsets (+ total, bounding box)
	int tiles = 0;
	int box_lo[3] = {0, 0, 0}, box_n[3] = {1, 1, 1};   // bounding box of the unknowns of the last numbering
	bool used_box = false;                               // the last Jacobi-PCG solve ran on that box
	BoxKrylov BK;
	Rows<double> rows{};
	Rows<float> rows32{};
	double *rhs = nullptr;
	Krylov K;
	Unet *net = nullptr;      // preconditioner of learned solves (not owned)
	LearnedWork LW;
	Multigrid MG;
	size_t cap = 0;
	int n = 0;

	~Projection();
	void release();
	bool reserveGrid(size_t N, std::string &err);
	bool reserveRows(int D, size_t n, bool mixed, std::string &err);
};

// The numbering passes over any flag array (tiles of kScanTile ids):
// tileBase[t] = flags set in tiles before t, *count = all of them; then
// grid2mat[id] = the id's number in id order (-1 where unflagged) and
// mat2grid its inverse.
void launchScanTiles(const int *tileCount, int tiles, int *tileBase, int *count, hipStream_t s);
void launchNumberFlags(const uint8_t *flags, const int *tileBase, int tiles, int *grid2mat, int *mat2grid, size_t N, hipStream_t s);

int runProjection(Fields &F, Projection &P, const garm_pcg_opts &o, garm_pcg_stats *st, hipStream_t s, std::string &err);
// What scatterCompact writes to every cell (0 where there is no unknown).
enum CompactField { kCompactPressure = 0, kCompactRhs = 1, kCompactIndicator = 2 };
void scatterCompact(const Fields &F, const Projection &P, int which, double *out, hipStream_t s);
// b = A x over whole cell fields with the operator of the last assembly
void applySystemCells(const Fields &F, const Projection &P, const double *x, double *b, hipStream_t s);

} // namespace garm
