// Multigrid V-cycle preconditioner for the ghost-fluid Poisson system
// (the operator of Projector::buildLinearSystem, Projector.cpp:30-68).
//
// No counterpart in the reference, which preconditions with Eigen's
// incomplete Cholesky (LinearSystem.h:10-39).  The hierarchy is algebraic:
// level 0 is the assembled operator itself laid out on the bounding box of
// the unknowns (diagonal plus the weights of each cell's -x, -y, -z faces),
// and level l+1 is the Galerkin product P^T A P with P the piecewise-constant
// prolongation from 2x2x2 aggregates, which keeps a 7-point stencil:
//   coarse face weight = sum of the fine face weights crossing it,
//   coarse diagonal    = sum of the fine diagonals - 2 * sum of the weights
//                        of the faces inside the aggregate.
// Free surfaces (ghost-fluid terms in the diagonal) and solid faces (zero
// weights) carry over to every level without special cases.  Smoothing is
// red-black Gauss-Seidel, two sweeps down and two up in mirrored colour order
// so the cycle is symmetric; the coarsest level (<= 512 cells) is relaxed
// 12 times.  The coarse levels that fit in 158 KiB of LDS together (from a
// few thousand cells down) run as one workgroup-wide kernel on LDS copies:
// each of their sweeps would otherwise be a launch of a few microseconds
// over a grid too small to fill the chip.  Everything is FP32 on structured
// grids: unit-stride, coalesced neighbours, no index arrays.
#include "garm_multigrid.hpp"

#include <algorithm>

namespace garm {

constexpr int kMgSweeps = 2;            // red-black sweep pairs before and after the coarse correction
constexpr int kMgSweeps0 = 2;           // ... on the finest level (at least 1: the cycle's head is one of them)
static_assert(kMgSweeps0 >= 1, "the head of a cycle is a level-0 sweep");
constexpr float kMgProlongScale = 1.9f; // over-correction of the coarse correction (aggregation underestimates it; 1.9 measured best at 256^3)
constexpr int kMgCoarsestCells = 512;
constexpr int kMgCoarsestSweeps = 12;

struct MgLevel {
	int n[3];
	float *diag;
	mgw_t *wx, *wy, *wz;
	float *b, *x;
};

// sweep tiles: large levels, small levels, 2D
constexpr int kTileL[3] = {32, 8, 8}, kTileS[3] = {16, 8, 4}, kTile2[3] = {32, 32, 1};

static MgLevel level(const Multigrid &M, int l)
{
	MgLevel L;
	for (int a = 0; a < 3; a++) L.n[a] = M.dims[l][a];
	L.diag = M.diag[l]; L.wx = M.wx[l]; L.wy = M.wy[l]; L.wz = M.wz[l]; L.b = M.b[l]; L.x = M.x[l];
	return L;
}

// Level boxes lie inside the cell grid, which holds fewer than 2^31 cells
// (checked at creation): 32-bit cell indices.
__device__ __forceinline__ unsigned cidx(const MgLevel &L, int i, int j, int k) { return (unsigned(k) * unsigned(L.n[1]) + unsigned(j)) * unsigned(L.n[0]) + unsigned(i); }

// The level kernels run on 3D grids (x over a row, y over rows, z over
// planes), so no thread divides its way from a linear index to (i, j, k).
struct Tid { int x, y, z; };
__device__ __forceinline__ Tid tid3() { return Tid{int(blockIdx.x * blockDim.x + threadIdx.x), int(blockIdx.y * blockDim.y + threadIdx.y), int(blockIdx.z)}; }

__global__ void kMgZero(float *p, size_t n)
{
	for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) p[i] = 0.0f;
}

__global__ void kMgZeroBytes(uint8_t *p, size_t n)
{
	for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) p[i] = 0;
}

struct MgBox {
	int org[3], gn[3];
};

__device__ __forceinline__ unsigned boxCell(const MgBox &B, const MgLevel &L, int g)
{
	const int i = g % B.gn[0] - B.org[0];
	const int j = (g / B.gn[0]) % B.gn[1] - B.org[1];
	const int k = g / (B.gn[0] * B.gn[1]) - B.org[2];
	return cidx(L, i, j, k);
}

// level 0: the operator entries of every unknown at its box cell
template <int D>
__global__ __launch_bounds__(kBlock) void kMgBuild0(Rows<float> A, int n, const int *mat2grid, MgBox B, MgLevel L, unsigned *boxOf)
{
	for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
		const unsigned c = boxCell(B, L, mat2grid[r]);
		boxOf[r] = c;
		L.diag[c] = A.diag[r];
		L.wx[c] = mgw_t(A.nb[0][r] >= 0 ? A.w[0][r] : 0.0f);
		L.wy[c] = mgw_t(A.nb[2][r] >= 0 ? A.w[2][r] : 0.0f);
		if constexpr (D == 3) L.wz[c] = mgw_t(A.nb[4][r] >= 0 ? A.w[4][r] : 0.0f);
	}
}

// Galerkin coarsening by 2x2x2 (2x2 in 2D) aggregates
template <int D>
__global__ __launch_bounds__(kBlock) void kMgCoarsen(MgLevel F, MgLevel C)
{
	const Tid t = tid3();
	const int I = t.x, J = t.y, K = t.z;
	if (I >= C.n[0] || J >= C.n[1]) return;
	const unsigned q = cidx(C, I, J, K);
	{
		float d = 0, inner = 0, sx = 0, sy = 0, sz = 0;
#pragma unroll
		for (int cz = 0; cz < (D == 3 ? 2 : 1); cz++)
#pragma unroll
			for (int cy = 0; cy < 2; cy++)
#pragma unroll
				for (int cx = 0; cx < 2; cx++) {
					const int i = 2 * I + cx, j = 2 * J + cy, k = 2 * K + cz;
					if (i >= F.n[0] || j >= F.n[1] || k >= F.n[2]) continue;
					const unsigned f = cidx(F, i, j, k);
					d += F.diag[f];
					const float wx = F.wx[f], wy = F.wy[f];
					if (cx) inner += wx; else sx += wx;
					if (cy) inner += wy; else sy += wy;
					if constexpr (D == 3) { const float wz = F.wz[f]; if (cz) inner += wz; else sz += wz; }
				}
		C.diag[q] = d - 2 * inner;
		C.wx[q] = mgw_t(sx);
		C.wy[q] = mgw_t(sy);
		if constexpr (D == 3) C.wz[q] = mgw_t(sz);
	}
}

// (A x)_c without the diagonal: the sum of face weight times neighbour
template <int D>
__device__ __forceinline__ float offSum(const MgLevel &L, unsigned c, int i, int j, int k)
{
	const unsigned sy = unsigned(L.n[0]), sz = unsigned(L.n[0]) * unsigned(L.n[1]);
	float s = 0;
	if (i > 0) s += float(L.wx[c]) * L.x[c - 1];
	if (i + 1 < L.n[0]) s += float(L.wx[c + 1]) * L.x[c + 1];
	if (j > 0) s += float(L.wy[c]) * L.x[c - sy];
	if (j + 1 < L.n[1]) s += float(L.wy[c + sy]) * L.x[c + sy];
	if constexpr (D == 3) {
		if (k > 0) s += float(L.wz[c]) * L.x[c - sz];
		if (k + 1 < L.n[2]) s += float(L.wz[c + sz]) * L.x[c + sz];
	}
	return s;
}

// The operator row of a cell: diagonal, right-hand side and its six face
// weights (-x, +x, -y, +y, -z, +z).  The +x, +y, +z weights of the level's
// last cells clamp to the last cell (they meet a 0 in the tile below).
struct MgRow { float d, b, w[6]; };

template <int D>
__device__ __forceinline__ MgRow loadRow(const MgLevel &L, unsigned c, unsigned last)
{
	const unsigned sy = unsigned(L.n[0]), sz = unsigned(L.n[0]) * unsigned(L.n[1]);
	MgRow r;
	r.d = L.diag[c];
	r.b = L.b[c];
	r.w[0] = L.wx[c];
	r.w[1] = L.wx[min(c + 1, last)];
	r.w[2] = L.wy[c];
	r.w[3] = L.wy[min(c + sy, last)];
	r.w[4] = D == 3 ? float(L.wz[c]) : 0.0f;
	r.w[5] = D == 3 ? float(L.wz[min(c + sz, last)]) : 0.0f;
	return r;
}

// Relaxation of a cell (tile index l) against the iterate in an LDS tile
// (row and plane strides SX, SXY).  The tile holds 0 outside the level box,
// and every face on the box boundary has weight 0 (the box bounds the
// unknowns; coarsening sums those zeros), so the six terms are added
// unconditionally: a missing neighbour adds +0 exactly, and s, which starts
// at +0 and adds products of non-negative weights, is never -0.  The result
// is offSum's, term by term.  An inactive cell (diag <= 0) keeps `keep`.
template <int D, int SX, int SXY>
__device__ __forceinline__ float relaxTile(const MgRow &r, const float *xs, int l, float keep)
{
	float s = 0;
	s += r.w[0] * xs[l - 1];
	s += r.w[1] * xs[l + 1];
	s += r.w[2] * xs[l - SX];
	s += r.w[3] * xs[l + SX];
	if constexpr (D == 3) {
		s += r.w[4] * xs[l - SXY];
		s += r.w[5] * xs[l + SXY];
	}
	return r.d > 0 ? (r.b + s) / r.d : keep;
}

// One red-black Gauss-Seidel sweep, colour `first` then the other, in one
// launch.  A block owns a TX x TY x TZ tile of the level: it stages the old
// iterate of the tile and two cells around it in LDS, relaxes the first
// colour on the tile and the ring of cells around it (the values the tile's
// second colour reads; the neighbouring blocks compute the same ring cells
// from the same inputs), then relaxes the second colour on the tile and
// writes the tile to the other iterate buffer.  Reading xin and writing xout
// keeps the blocks independent, and every cell gets exactly the operations
// of one single-colour pass after the other.  Each phase gives every thread
// a fixed count of cells, predicated (indices outside the level clamp to a
// valid cell, results are dropped), and a thread issues the global loads of
// all three phases before its first LDS store, so they are in flight
// together and the phases after the first barrier wait only on LDS.
template <int D, int TX, int TY, int TZ, int NTH>
__global__ __launch_bounds__(NTH) void kMgSweep(MgLevel L, const float *xin, float *xout, int first, const uint8_t *active)
{
	// a tile without active cells keeps x = 0 in both buffers: nothing to do
	const Tile3 tile = xcdTiles();
	if (!active[(size_t(tile.z) * gridDim.y + tile.y) * gridDim.x + tile.x]) return;
	static_assert(TX % 2 == 0 && (D == 3 || TZ == 1), "tile shape");
	constexpr int HZ = D == 3 ? 2 : 0;                    // halo of the staged tile along z
	constexpr int SX = TX + 4, SY = TY + 4, SZ = TZ + 2 * HZ, SXY = SX * SY, NS = SXY * SZ;
	constexpr int RY = TY + 2, RZ = D == 3 ? TZ + 2 : 1, HX = (TX + 2) / 2, NR = HX * RY * RZ, NT = TX * TY * TZ;
	constexpr int P0 = (NS + NTH - 1) / NTH, P1 = (NR + NTH - 1) / NTH, P2 = (NT + NTH - 1) / NTH;
	constexpr int kNone = SXY + SX + 1;                    // a tile index whose neighbours are all in the tile
	__shared__ float xs[NS];
	const int x0 = tile.x * TX, y0 = tile.y * TY, z0 = tile.z * TZ;
	const int n0 = L.n[0], n1 = L.n[1], n2 = L.n[2];
	const unsigned last = unsigned(n0) * unsigned(n1) * unsigned(n2) - 1;
	auto inBox = [&](int i, int j, int k) { return i >= 0 && i < n0 && j >= 0 && j < n1 && k >= 0 && k < n2; };
	// every global load of the three phases first: the staged iterate and the
	// operator rows of both colours' cells (the rows do not depend on the tile)
	float v0[P0];
	bool in0[P0];
#pragma unroll
	for (int r = 0; r < P0; r++) {
		const int t = int(threadIdx.x) + r * NTH;
		const int i = x0 - 2 + t % SX, j = y0 - 2 + (t / SX) % SY, k = z0 - HZ + t / SXY;
		in0[r] = t < NS && inBox(i, j, k);
		v0[r] = xin[in0[r] ? cidx(L, i, j, k) : 0u];
	}
	MgRow row1[P1];
	int l1[P1];
#pragma unroll
	for (int r = 0; r < P1; r++) {
		const int t = int(threadIdx.x) + r * NTH;
		const int j = y0 - 1 + (t / HX) % RY, k = z0 - (D == 3 ? 1 : 0) + t / (HX * RY);
		const int lo = x0 - 1, i = lo + 2 * (t % HX) + ((lo + j + k + first) & 1);   // (i + j + k) & 1 == first
		const bool in = t < NR && inBox(i, j, k);
		l1[r] = in ? ((k - z0 + HZ) * SY + j - y0 + 2) * SX + i - x0 + 2 : -1;
		row1[r] = loadRow<D>(L, in ? cidx(L, i, j, k) : 0u, last);
	}
	MgRow row2[P2];
	unsigned c2[P2];
	int l2[P2];
	bool in2[P2], upd[P2];
#pragma unroll
	for (int r = 0; r < P2; r++) {
		const int t = int(threadIdx.x) + r * NTH;
		const int ti = t % TX, tj = (t / TX) % TY, tk = t / (TX * TY);
		const int i = x0 + ti, j = y0 + tj, k = z0 + tk;
		in2[r] = t < NT && i < n0 && j < n1 && k < n2;
		upd[r] = ((i + j + k) & 1) != first;
		l2[r] = in2[r] ? ((tk + HZ) * SY + tj + 2) * SX + ti + 2 : kNone;
		c2[r] = in2[r] ? cidx(L, i, j, k) : 0u;
		row2[r] = loadRow<D>(L, upd[r] ? c2[r] : 0u, last);
	}
#pragma unroll
	for (int r = 0; r < P0; r++) {
		const int t = int(threadIdx.x) + r * NTH;
		if (t < NS) xs[t] = in0[r] ? v0[r] : 0.0f;
	}
	__syncthreads();
	// the first colour on the tile and one cell around it
	{
		float v1[P1];
#pragma unroll
		for (int r = 0; r < P1; r++) {
			const int l = l1[r] >= 0 ? l1[r] : kNone;
			v1[r] = relaxTile<D, SX, SXY>(row1[r], xs, l, xs[l]);
		}
#pragma unroll
		for (int r = 0; r < P1; r++)
			if (l1[r] >= 0) xs[l1[r]] = v1[r];
	}
	__syncthreads();
	// the second colour on the tile; the tile goes out
#pragma unroll
	for (int r = 0; r < P2; r++) {
		const float v = xs[l2[r]];
		const float nv = upd[r] ? relaxTile<D, SX, SXY>(row2[r], xs, l2[r], v) : v;
		if (in2[r]) xout[c2[r]] = nv;
	}
}

// Marks the sweep tiles (tx x ty x tz cells, ntx x nty per plane of tiles)
// that hold an active cell (diag > 0); every writer stores the same 1.
template <int D>
__global__ __launch_bounds__(kBlock) void kMgTileMask(MgLevel L, int tx, int ty, int tz, int ntx, int nty, uint8_t *mask)
{
	const Tid t = tid3();
	if (t.x >= L.n[0] || t.y >= L.n[1]) return;
	if (L.diag[cidx(L, t.x, t.y, t.z)] > 0) mask[(size_t(t.z / tz) * nty + t.y / ty) * ntx + t.x / tx] = 1;
}

// coarse right-hand side = sum of the fine residuals of the aggregate; coarse
// iterate = 0.  A block covers 64 x 2 x 2 fine cells (64 x 2 in 2D), one per
// thread, so each wave reads 64 consecutive cells of a row.  The residual
// takes the six neighbour terms unconditionally, in offSum's order: a face
// on the box boundary has weight 0 and an index outside the level reads an
// iterate of 0, so a missing neighbour adds +0 exactly, as in relaxTile.  The x pairs of an
// aggregate meet across adjacent lanes, the y and z pairs in LDS; the sum is
// ((r000 + r100) + (r010 + r110)) + ((r001 + r101) + (r011 + r111)).
constexpr int kRestrictX = 64;

template <int D>
__global__ __launch_bounds__(kBlock) void kMgRestrict(MgLevel F, MgLevel C)
{
	constexpr int BZ = D == 3 ? 2 : 1;
	__shared__ float pair[BZ][2][kRestrictX / 2];
	const int i = int(blockIdx.x) * kRestrictX + int(threadIdx.x), j = int(blockIdx.y) * 2 + int(threadIdx.y), k = int(blockIdx.z) * BZ + int(threadIdx.z);
	const int n0 = F.n[0], n1 = F.n[1], n2 = F.n[2];
	const bool in = i < n0 && j < n1 && k < n2;
	const unsigned last = unsigned(n0) * unsigned(n1) * unsigned(n2) - 1;
	const unsigned sy = unsigned(n0), sz = unsigned(n0) * unsigned(n1);
	const unsigned c = in ? cidx(F, i, j, k) : 0u;
	const float d = F.diag[c], b = F.b[c], xc = F.x[c];
	const float w0 = F.wx[c], w1 = F.wx[min(c + 1, last)], w2 = F.wy[c], w3 = F.wy[min(c + sy, last)];
	// a clamped +x, +y or +z index (past the last cell) stands for a missing
	// neighbour: its iterate reads as 0, so the clamped weight adds +0
	auto xAt = [&](unsigned e, bool ok) { const float v = F.x[ok ? e : 0u]; return ok ? v : 0.0f; };
	const float x0 = xAt(c - 1, c > 0), x1 = xAt(c + 1, c + 1 <= last), x2 = xAt(c - sy, c >= sy), x3 = xAt(c + sy, c + sy <= last);
	float s = 0;
	s += w0 * x0;
	s += w1 * x1;
	s += w2 * x2;
	s += w3 * x3;
	if constexpr (D == 3) {
		const float w4 = F.wz[c], w5 = F.wz[min(c + sz, last)];
		const float x4 = xAt(c - sz, c >= sz), x5 = xAt(c + sz, c + sz <= last);
		s += w4 * x4;
		s += w5 * x5;
	}
	float r = in && d > 0 ? b - (d * xc - s) : 0.0f;
	r += __shfl_xor(r, 1);
	if ((threadIdx.x & 1) == 0) pair[threadIdx.z][threadIdx.y][threadIdx.x / 2] = r;
	__syncthreads();
	if (threadIdx.y == 0 && threadIdx.z == 0 && (threadIdx.x & 1) == 0) {
		const int h = int(threadIdx.x) / 2, I = i / 2, J = j / 2, K = k / BZ;
		float sum = pair[0][0][h] + pair[0][1][h];
		if constexpr (D == 3) sum = sum + (pair[1][0][h] + pair[1][1][h]);
		if (I < C.n[0] && J < C.n[1] && K < C.n[2]) {
			const unsigned q = cidx(C, I, J, K);
			C.b[q] = sum;
			C.x[q] = 0.0f;
		}
	}
}

// fine iterate += coarse iterate of its aggregate
template <int D>
__global__ __launch_bounds__(kBlock) void kMgProlong(MgLevel F, MgLevel C)
{
	const Tid t = tid3();
	const int i = t.x, j = t.y, k = t.z;
	if (i >= F.n[0] || j >= F.n[1]) return;
	const unsigned q = cidx(F, i, j, k);
	if (!(F.diag[q] > 0)) return;
	F.x[q] += kMgProlongScale * C.x[cidx(C, i / 2, j / 2, D == 3 ? k / 2 : 0)];
}

// ---- the coarse levels in LDS -------------------------------------------
// Each tail level keeps its six arrays padded by one xy-plane of zeros on
// both sides: every cell at a box face already has a zero weight on that
// face (the box bounds the unknowns, and coarsening sums those zeros), so
// the stencil can read its six neighbours without bounds checks, the
// missing ones contributing 0 * finite = 0 in the same order as offSum.
// The active cells (diag > 0) of each colour are listed once per cycle
// (16-bit cell index, and for prolongation the coarse cell's index), so
// the sweeps do no index arithmetic.
constexpr int kMgTailFloats = 40448;    // 158 KiB of the 160 KiB of LDS: padded arrays and cell lists

struct MgTail {
	int first, levels;              // levels first .. levels-1 run in LDS
	int sweeps;                     // relaxation sweeps on the coarsest level
	MgLevel g[kMgMaxLevels];        // their global arrays
	int off[kMgMaxLevels];          // LDS offset (floats) of each level's padded arrays
	int list[kMgMaxLevels];         // LDS offset (words) of each level's cell list
	int list_words;                 // words of all the lists (contiguous from list[first])
	unsigned *glist;                // the lists in global memory, built once per operator
	int *gcount;                    // active cells per level and colour
};

// padded plane of level L
__device__ __forceinline__ int mgPad(const MgLevel &L) { return L.n[0] * L.n[1]; }
__device__ __forceinline__ int mgCells(const MgLevel &L) { return L.n[0] * L.n[1] * L.n[2]; }

__device__ __forceinline__ MgLevel ldsLevel(float *lds, const MgTail &T, int l)
{
	MgLevel V;
	for (int a = 0; a < 3; a++) V.n[a] = T.g[l].n[a];
	const int c = mgCells(V), pad = mgPad(V), span = c + 2 * pad;
	float *p = lds + T.off[l] + pad;
	// the weights keep float-sized slots in LDS (the layout's accounting)
	V.diag = p;
	V.wx = reinterpret_cast<mgw_t *>(p + span);
	V.wy = reinterpret_cast<mgw_t *>(p + 2 * span);
	V.wz = reinterpret_cast<mgw_t *>(p + 3 * span);
	V.b = p + 4 * span;
	V.x = p + 5 * span;
	return V;
}

template <int D>
__device__ __forceinline__ float offSumPadded(const MgLevel &L, int c)
{
	const int sy = L.n[0], sz = L.n[0] * L.n[1];
	float s = 0;
	s += float(L.wx[c]) * L.x[c - 1];
	s += float(L.wx[c + 1]) * L.x[c + 1];
	s += float(L.wy[c]) * L.x[c - sy];
	s += float(L.wy[c + sy]) * L.x[c + sy];
	if constexpr (D == 3) {
		s += float(L.wz[c]) * L.x[c - sz];
		s += float(L.wz[c + sz]) * L.x[c + sz];
	}
	return s;
}

// one colour of red-black Gauss-Seidel over the listed cells
template <int D>
__device__ __forceinline__ void tailSmooth(const MgLevel &L, const unsigned *list, int from, int to)
{
	for (int t = from + int(threadIdx.x); t < to; t += blockDim.x) {
		const int c = int(list[t] & 0xffffu);
		L.x[c] = (L.b[c] + offSumPadded<D>(L, c)) / L.diag[c];
	}
	__syncthreads();
}

template <int D>
__global__ __launch_bounds__(1024) void kMgTail(MgTail T)
{
	__shared__ float lds[kMgTailFloats];
	__shared__ int ncol[kMgMaxLevels][2];   // active cells per colour
	const int tid = int(threadIdx.x);
	if (tid < 2 * kMgMaxLevels) ncol[tid / 2][tid % 2] = 0;
	// operators of every tail level (zero padding), the first level's right-hand side, zero iterate
	for (int l = T.first; l < T.levels; l++) {
		const MgLevel G = T.g[l], V = ldsLevel(lds, T, l);
		const int c = mgCells(G), pad = mgPad(G);
		for (int q = tid - pad; q < c + pad; q += blockDim.x) {
			const bool in = q >= 0 && q < c;
			V.diag[q] = in ? G.diag[q] : 0.0f;
			V.wx[q] = in ? G.wx[q] : mgw_t(0);
			V.wy[q] = in ? G.wy[q] : mgw_t(0);
			V.wz[q] = in ? G.wz[q] : mgw_t(0);
			V.b[q] = in && l == T.first ? G.b[q] : 0.0f;
			V.x[q] = 0.0f;
		}
	}
	__syncthreads();
	// the active cells of each colour (red first), listed when the operator was built
	if (tid < 2 * kMgMaxLevels) ncol[tid / 2][tid % 2] = T.gcount[tid];
	{
		unsigned *list = reinterpret_cast<unsigned *>(lds + T.list[T.first]);
		for (int q = tid; q < T.list_words; q += blockDim.x) list[q] = T.glist[q];
	}
	__syncthreads();
	const int top = T.levels - 1;
	for (int l = T.first; l < top; l++) {
		const MgLevel F = ldsLevel(lds, T, l), C = ldsLevel(lds, T, l + 1);
		const unsigned *list = reinterpret_cast<const unsigned *>(lds + T.list[l]);
		const int nr = ncol[l][0], na = nr + ncol[l][1];
		for (int v = 0; v < kMgSweeps; v++) { tailSmooth<D>(F, list, 0, nr); tailSmooth<D>(F, list, nr, na); }
		// kMgRestrict
		const int cc = mgCells(C);
		for (int q = tid; q < cc; q += blockDim.x) {
			const int I = q % C.n[0], J = (q / C.n[0]) % C.n[1], K = q / (C.n[0] * C.n[1]);
			float s = 0;
			for (int cz = 0; cz < (D == 3 ? 2 : 1); cz++)
				for (int cy = 0; cy < 2; cy++)
					for (int cx = 0; cx < 2; cx++) {
						const int i = 2 * I + cx, j = 2 * J + cy, k = 2 * K + cz;
						if (i >= F.n[0] || j >= F.n[1] || k >= F.n[2]) continue;
						const int f = int(cidx(F, i, j, k));
						const float d = F.diag[f];
						if (d > 0) s += F.b[f] - (d * F.x[f] - offSumPadded<D>(F, f));
					}
			C.b[q] = s;
		}
		__syncthreads();
	}
	{
		// kMgCoarsest
		const MgLevel L = ldsLevel(lds, T, top);
		const unsigned *list = reinterpret_cast<const unsigned *>(lds + T.list[top]);
		const int nr = ncol[top][0], na = nr + ncol[top][1];
		for (int s = 0; s < 2 * T.sweeps; s++) {
			const int color = (s & 1) ^ (s >= T.sweeps ? 1 : 0);
			tailSmooth<D>(L, list, color ? nr : 0, color ? na : nr);
		}
	}
	for (int l = top - 1; l >= T.first; l--) {
		const MgLevel F = ldsLevel(lds, T, l), C = ldsLevel(lds, T, l + 1);
		const unsigned *list = reinterpret_cast<const unsigned *>(lds + T.list[l]);
		const int nr = ncol[l][0], na = nr + ncol[l][1];
		// kMgProlong
		for (int t = tid; t < na; t += blockDim.x) {
			const unsigned e = list[t];
			F.x[e & 0xffffu] += kMgProlongScale * C.x[e >> 16];
		}
		__syncthreads();
		for (int v = 0; v < kMgSweeps; v++) { tailSmooth<D>(F, list, nr, na); tailSmooth<D>(F, list, 0, nr); }
	}
	const MgLevel G = T.g[T.first], V = ldsLevel(lds, T, T.first);
	for (int q = tid; q < mgCells(G); q += blockDim.x) G.x[q] = V.x[q];
}

// Lists the active cells (diag > 0) of every tail level by colour, red
// first, as 16-bit cell index | coarse cell index << 16; the order within a
// colour is free (cells of one colour do not neighbour each other).
template <int D>
__global__ __launch_bounds__(1024) void kMgTailLists(MgTail T)
{
	__shared__ int ncol[kMgMaxLevels][2];
	__shared__ int fill[kMgMaxLevels][2];
	const int tid = int(threadIdx.x);
	if (tid < 2 * kMgMaxLevels) ncol[tid / 2][tid % 2] = 0;
	__syncthreads();
	for (int l = T.first; l < T.levels; l++) {
		const MgLevel G = T.g[l];
		for (int q = tid; q < mgCells(G); q += blockDim.x) {
			if (!(G.diag[q] > 0)) continue;
			const int i = q % G.n[0], j = (q / G.n[0]) % G.n[1], k = q / (G.n[0] * G.n[1]);
			atomicAdd(&ncol[l][(i + j + k) & 1], 1);
		}
	}
	__syncthreads();
	if (tid < 2 * kMgMaxLevels) {
		fill[tid / 2][tid % 2] = tid % 2 ? ncol[tid / 2][0] : 0;
		T.gcount[tid] = ncol[tid / 2][tid % 2];
	}
	__syncthreads();
	for (int l = T.first; l < T.levels; l++) {
		const MgLevel G = T.g[l];
		unsigned *list = T.glist + (T.list[l] - T.list[T.first]);
		const bool coarse = l + 1 < T.levels;
		const MgLevel C = coarse ? T.g[l + 1] : G;
		for (int q = tid; q < mgCells(G); q += blockDim.x) {
			if (!(G.diag[q] > 0)) continue;
			const int i = q % G.n[0], j = (q / G.n[0]) % G.n[1], k = q / (G.n[0] * G.n[1]);
			const unsigned cq = coarse ? unsigned(cidx(C, i / 2, j / 2, D == 3 ? k / 2 : 0)) : 0u;
			list[atomicAdd(&fill[l][(i + j + k) & 1], 1)] = unsigned(q) | cq << 16;
		}
	}
}

// compact FP64 residual -> level-0 right-hand side, zero iterate.  Only the
// unknowns' cells: the others keep b = x = 0 from the build (no kernel
// writes an inactive cell of level 0).
__global__ void kMgLoad(int n, const unsigned *boxOf, MgLevel L, const double *r)
{
	for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
		const unsigned c = boxOf[q];
		L.b[c] = float(r[q]);
		L.x[c] = 0.0f;
	}
}

__global__ void kMgStore(int n, const unsigned *boxOf, MgLevel L, double *z)
{
	for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) z[q] = double(L.x[boxOf[q]]);
}

// the same for residual and result laid out on level 0's box itself
__global__ void kMgLoadBox(int cells, MgLevel L, const double *r)
{
	for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += gridDim.x * blockDim.x) {
		L.b[c] = float(r[c]);
		L.x[c] = 0.0f;
	}
}

__global__ void kMgStoreBox(int cells, MgLevel L, double *z)
{
	for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += gridDim.x * blockDim.x) z[c] = double(L.x[c]);
}

// The same store with the per-workgroup partials of z.q (the flexible CG's
// next inner product): the launch's grid and each thread's order are those
// of the CG's own dot-product kernel, so the partials are the same.
__global__ __launch_bounds__(kBlock) void kMgStoreBoxDot(int cells, MgLevel L, double *z, const double *q, double *pa)
{
	double v[1] = {0};
	for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += gridDim.x * blockDim.x) {
		const double zc = double(L.x[c]);
		if (z) z[c] = zc;
		v[0] += zc * q[c];
	}
	double blk[1];
	blockSum<1>(v, blk);
	if (threadIdx.x == 0) pa[blockIdx.x] = blk[0];
}

// ------------------------------------------------------------------ host
static unsigned mgBlocks(size_t n) { return unsigned(std::max<size_t>(1, std::min<size_t>((n + kBlock - 1) / kBlock, 4096))); }

// 3D launch over nx x ny x nz threads: rows of bx threads (8 .. 64, the
// smallest power of two that covers nx), kBlock / bx rows per block
struct Launch3 { dim3 grid, block; };
static Launch3 mgLaunch3(size_t nx, int ny, int nz)
{
	unsigned bx = 8;
	while (bx < 64 && bx < nx) bx *= 2;
	const unsigned by = kBlock / bx;
	return Launch3{dim3(unsigned((nx + bx - 1) / bx), (unsigned(ny) + by - 1) / by, unsigned(nz)), dim3(bx, by, 1)};
}
static Launch3 mgLevelLaunch(const Multigrid &M, int l, size_t nx) { return mgLaunch3(nx, M.dims[l][1], M.dims[l][2]); }

// the sweep tile of level l
static const int *sweepTile(const Multigrid &M, int l)
{
	return M.D == 2 ? kTile2 : M.cells[l] >= size_t(1) << 20 ? kTileL : kTileS;
}

static size_t tileCount(const Multigrid &M, int l)
{
	const int *t = sweepTile(M, l), *n = M.dims[l];
	return size_t((n[0] + t[0] - 1) / t[0]) * ((n[1] + t[1] - 1) / t[1]) * ((n[2] + t[2] - 1) / t[2]);
}


Multigrid::~Multigrid()
{
	if (pool) (void)hipFree(pool);
	if (tile_active) (void)hipFree(tile_active);
}

bool Multigrid::layout(int dim, const int box_org[3], const int box_n[3], const int grid_n[3], int unknowns, std::string &err)
{
	D = dim;
	for (int a = 0; a < 3; a++) { org[a] = box_org[a]; gn[a] = grid_n[a]; dims[0][a] = box_n[a]; }
	levels = 1;
	for (int l = 0;; l++) {
		cells[l] = size_t(dims[l][0]) * dims[l][1] * dims[l][2];
		if (cells[l] <= size_t(kMgCoarsestCells) || l + 1 == kMgMaxLevels) break;
		for (int a = 0; a < 3; a++) dims[l + 1][a] = (a < D) ? (dims[l][a] + 1) / 2 : 1;
		levels++;
	}
	// the coarse levels that fit in the LDS of one workgroup together (six
	// padded arrays and a cell list each; 16-bit cell indices)
	auto tailWords = [&](int l) {
		const size_t pad = size_t(dims[l][0]) * dims[l][1];
		return 6 * (cells[l] + 2 * pad) + cells[l];
	};
	tail_first = levels - 1;
	size_t words = tailWords(levels - 1);
	while (tail_first > 0 && cells[tail_first - 1] <= 65535 && words + tailWords(tail_first - 1) <= size_t(kMgTailFloats)) {
		tail_first--;
		words += tailWords(tail_first);
	}
	tail_words = 0;
	for (int l = tail_first; l < levels; l++) tail_words += cells[l];
	// per level: diag, b, x, xalt in floats, the three weight arrays in halves
	auto levelFloats = [&](int l) { return 4 * cells[l] + 3 * ((cells[l] + 1) / 2); };
	level_floats = 0;
	for (int l = 0; l < levels; l++) level_floats += levelFloats(l);
	const size_t need = level_floats + tail_words + 2 * kMgMaxLevels + size_t(unknowns) + 64;
	if (need > pool_cap) {
		if (pool) (void)hipFree(pool);
		pool = nullptr;
		const size_t want = need + need / 4;
		if (hipMalloc(&pool, want * sizeof(float)) != hipSuccess) { err = "multigrid allocation failed"; pool_cap = 0; return false; }
		pool_cap = want;
	}
	float *p = pool;
	for (int l = 0; l < levels; l++) {
		float **arr[4] = {&diag[l], &b[l], &x[l], &xalt[l]};
		for (float **a : arr) { *a = p; p += cells[l]; }
		mgw_t **warr[3] = {&wx[l], &wy[l], &wz[l]};
		for (mgw_t **a : warr) { *a = reinterpret_cast<mgw_t *>(p); p += (cells[l] + 1) / 2; }
	}
	tail_list = reinterpret_cast<unsigned *>(p);
	tail_count = reinterpret_cast<int *>(p + tail_words);
	box_of = reinterpret_cast<unsigned *>(p + tail_words + 2 * kMgMaxLevels);
	// sweep-tile activity of the levels above the tail
	size_t tiles = 0;
	for (int l = 0; l < tail_first; l++) { tile_off[l] = tiles; tiles += tileCount(*this, l); }
	if (tiles > tile_cap || !tile_active) {
		if (tile_active) (void)hipFree(tile_active);
		tile_active = nullptr;
		const size_t want = std::max<size_t>(tiles + tiles / 4, 256);
		if (hipMalloc(&tile_active, want) != hipSuccess) { err = "multigrid allocation failed"; tile_cap = 0; return false; }
		tile_cap = want;
	}
	return true;
}

// the coarse levels run in LDS: LDS offsets of their padded arrays, then of their cell lists
static MgTail tailDesc(const Multigrid &M)
{
	MgTail T{};
	T.first = M.tail_first;
	T.levels = M.levels;
	T.sweeps = kMgCoarsestSweeps;
	int off = 0;
	for (int l = T.first; l < M.levels; l++) {
		T.g[l] = level(M, l);
		T.off[l] = off;
		off += int(6 * (M.cells[l] + 2 * size_t(M.dims[l][0]) * M.dims[l][1]));
	}
	for (int l = T.first; l < M.levels; l++) {
		T.list[l] = off;
		off += int(M.cells[l]);
	}
	T.list_words = int(M.tail_words);
	T.glist = M.tail_list;
	T.gcount = M.tail_count;
	return T;
}

void mgBuild(Multigrid &M, const Rows<float> &A, int n, const int *mat2grid, hipStream_t s)
{
	// every level array starts at zero (inactive cells keep zero entries)
	hipLaunchKernelGGL(kMgZero, dim3(mgBlocks(M.level_floats)), dim3(kBlock), 0, s, M.pool, M.level_floats);
	MgBox B;
	for (int a = 0; a < 3; a++) { B.org[a] = M.org[a]; B.gn[a] = M.gn[a]; }
	const MgLevel L0 = level(M, 0);
	if (M.D == 2) hipLaunchKernelGGL(kMgBuild0<2>, dim3(mgBlocks(size_t(n))), dim3(kBlock), 0, s, A, n, mat2grid, B, L0, M.box_of);
	else hipLaunchKernelGGL(kMgBuild0<3>, dim3(mgBlocks(size_t(n))), dim3(kBlock), 0, s, A, n, mat2grid, B, L0, M.box_of);
	for (int l = 0; l + 1 < M.levels; l++) {
		const MgLevel F = level(M, l), C = level(M, l + 1);
		const Launch3 g = mgLevelLaunch(M, l + 1, size_t(M.dims[l + 1][0]));
		if (M.D == 2) hipLaunchKernelGGL(kMgCoarsen<2>, g.grid, g.block, 0, s, F, C);
		else hipLaunchKernelGGL(kMgCoarsen<3>, g.grid, g.block, 0, s, F, C);
	}
	const MgTail T = tailDesc(M);
	if (M.D == 2) hipLaunchKernelGGL(kMgTailLists<2>, dim3(1), dim3(1024), 0, s, T);
	else hipLaunchKernelGGL(kMgTailLists<3>, dim3(1), dim3(1024), 0, s, T);
	if (M.tail_first > 0) {
		const size_t tiles = M.tile_off[M.tail_first - 1] + tileCount(M, M.tail_first - 1);
		hipLaunchKernelGGL(kMgZeroBytes, dim3(mgBlocks(tiles)), dim3(kBlock), 0, s, M.tile_active, tiles);
	}
	for (int l = 0; l < M.tail_first; l++) {
		const int *t = sweepTile(M, l), *n = M.dims[l];
		const int ntx = (n[0] + t[0] - 1) / t[0], nty = (n[1] + t[1] - 1) / t[1];
		const Launch3 g = mgLevelLaunch(M, l, size_t(n[0]));
		if (M.D == 2) hipLaunchKernelGGL(kMgTileMask<2>, g.grid, g.block, 0, s, level(M, l), t[0], t[1], t[2], ntx, nty, M.tile_active + M.tile_off[l]);
		else hipLaunchKernelGGL(kMgTileMask<3>, g.grid, g.block, 0, s, level(M, l), t[0], t[1], t[2], ntx, nty, M.tile_active + M.tile_off[l]);
	}
}

// one sweep of level l (colour `first` first) from x into xalt; the two swap names
template <int D>
static void sweep(Multigrid &M, int l, int first, hipStream_t s)
{
	const MgLevel L = level(M, l);
	const int *n = M.dims[l];
	const uint8_t *active = M.tile_active + M.tile_off[l];
	auto launch = [&](auto kernel, const int *tile, int threads) {
		const dim3 g(unsigned((n[0] + tile[0] - 1) / tile[0]), unsigned((n[1] + tile[1] - 1) / tile[1]), unsigned((n[2] + tile[2] - 1) / tile[2]));
		hipLaunchKernelGGL(kernel, g, dim3(unsigned(threads)), 0, s, L, (const float *)M.x[l], M.xalt[l], first, active);
	};
	if constexpr (D == 2) launch(kMgSweep<2, kTile2[0], kTile2[1], kTile2[2], 1024>, kTile2, 1024);
	else if (M.cells[l] >= size_t(1) << 20) launch(kMgSweep<3, kTileL[0], kTileL[1], kTileL[2], 1024>, kTileL, 1024);
	else launch(kMgSweep<3, kTileS[0], kTileS[1], kTileS[2], 256>, kTileS, 256);
	std::swap(M.x[l], M.xalt[l]);
}

// The V-cycle in two parts: the head (load of the residual and the first
// level-0 sweep) and the rest; together they are one cycle.
template <int D>
static void vcycleHead(Multigrid &M, int n, const double *r, bool box, bool loaded, hipStream_t s)
{
	if (box) {
		if (!loaded) hipLaunchKernelGGL(kMgLoadBox, dim3(mgBlocks(M.cells[0])), dim3(kBlock), 0, s, int(M.cells[0]), level(M, 0), r);
	} else hipLaunchKernelGGL(kMgLoad, dim3(mgBlocks(size_t(n))), dim3(kBlock), 0, s, n, (const unsigned *)M.box_of, level(M, 0), r);
	if (M.tail_first > 0) sweep<D>(M, 0, 0, s);
}

template <int D>
static void vcycleRest(Multigrid &M, int n, double *z, bool box, const double *q, double *zq, int zqBlocks, hipStream_t s)
{
	const int first = M.tail_first;
	for (int l = 0; l < first; l++) {
		for (int v = (l == 0 ? 1 : 0); v < (l == 0 ? kMgSweeps0 : kMgSweeps); v++) sweep<D>(M, l, 0, s);
		const int *f = M.dims[l];
		const dim3 gr(unsigned((f[0] + kRestrictX - 1) / kRestrictX), unsigned((f[1] + 1) / 2), unsigned(D == 3 ? (f[2] + 1) / 2 : 1));
		hipLaunchKernelGGL(kMgRestrict<D>, gr, dim3(kRestrictX, 2, D == 3 ? 2 : 1), 0, s, level(M, l), level(M, l + 1));
	}
	const MgTail T = tailDesc(M);
	hipLaunchKernelGGL(kMgTail<D>, dim3(1), dim3(1024), 0, s, T);
	for (int l = first - 1; l >= 0; l--) {
		const Launch3 gp = mgLevelLaunch(M, l, size_t(M.dims[l][0]));
		hipLaunchKernelGGL(kMgProlong<D>, gp.grid, gp.block, 0, s, level(M, l), level(M, l + 1));
		for (int v = 0; v < (l == 0 ? kMgSweeps0 : kMgSweeps); v++) sweep<D>(M, l, 1, s);
	}
	if (box && q) hipLaunchKernelGGL(kMgStoreBoxDot, dim3(unsigned(zqBlocks)), dim3(kBlock), 0, s, int(M.cells[0]), level(M, 0), z, q, zq);
	else if (box) hipLaunchKernelGGL(kMgStoreBox, dim3(mgBlocks(M.cells[0])), dim3(kBlock), 0, s, int(M.cells[0]), level(M, 0), z);
	else hipLaunchKernelGGL(kMgStore, dim3(mgBlocks(size_t(n))), dim3(kBlock), 0, s, n, (const unsigned *)M.box_of, level(M, 0), z);
}

void mgApplyHead(Multigrid &M, int n, const double *r, hipStream_t s, bool box, bool loaded)
{
	if (M.D == 2) vcycleHead<2>(M, n, r, box, loaded, s);
	else vcycleHead<3>(M, n, r, box, loaded, s);
}

void mgApplyRest(Multigrid &M, int n, double *z, hipStream_t s, bool box, const double *q, double *zq, int zqBlocks)
{
	if (M.D == 2) vcycleRest<2>(M, n, z, box, q, zq, zqBlocks, s);
	else vcycleRest<3>(M, n, z, box, q, zq, zqBlocks, s);
}

void mgLevel0(const Multigrid &M, float **b, float **x)
{
	*b = M.b[0];
	*x = M.x[0];
}

// Algorithmic HBM bytes of one V-cycle: the residual's load and the
// correction's store on level 0 (FP64 vector, FP32 level arrays: 16 + 12 B
// per cell), and per level above the LDS tail four sweeps (operator 10 B --
// FP32 diagonal, three half-precision weights --, b, x in, x out: 22 B per
// cell each), the restriction (22 B per fine cell) and the prolongation
// (8 B per fine cell).
double mgCycleBytes(const Multigrid &M)
{
	double b = 28.0 * double(M.cells[0]);
	for (int l = 0; l < M.tail_first; l++) b += (4 * 22.0 + 22.0 + 8.0) * double(M.cells[l]);
	return b;
}

void mgApply(Multigrid &M, int n, const double *r, double *z, hipStream_t s)
{
	mgApplyHead(M, n, r, s);
	mgApplyRest(M, n, z, s);
}

} // namespace garm
