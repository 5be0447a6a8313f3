// Ghost-fluid pressure projection for gfx950 (Projector.cpp:13-163).
//
// Unknowns are numbered in the reference's lexicographic order
// (Projector::setUnknowns, Projector.cpp:125-156) by a three-pass device
// prefix sum over tiles of consecutive cells.
// The 5/7-point operator is kept as one diagonal plus 2D (neighbour, weight)
// pairs per unknown, structure-of-arrays, and applied matrix-free: no sparse
// matrix is assembled and no factorisation is formed.  The conjugate-gradient
// scalars (alpha, beta, r.z, |r|^2) never leave the device; the host only
// polls a convergence flag every few iterations.
#include "garm_project.hpp"

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <vector>

namespace garm {

// ------------------------------------------------------------- unknowns
// Cells are numbered in tiles of kScanTile consecutive ids (x fastest), one
// workgroup per tile: pass 1 flags the unknowns and counts them per tile,
// pass 2 turns the tile counts into tile offsets (one workgroup), pass 3
// numbers each tile's unknowns in order.  Tiles and cells within a tile are
// taken in id order, so the numbering is Projector::setUnknowns' forEach order.
// box = (min i, min j, min k, -max i, -max j, -max k) of the unknowns
__global__ void kBoxInit(int *box)
{
	if (threadIdx.x < 6) box[threadIdx.x] = INT_MAX;
}

// open[a] holds 1 for each face of axis a whose fraction is below 1.
struct OpenFaces {
	GridDev g[3];
	const uint8_t *open[3];
};

template <int D>
__global__ __launch_bounds__(kBlock) void kUnknownFlags(GridDev g, OpenFaces fo, const double *phi, uint8_t *flags, int *tileCount, size_t N,
	int *box)
{
	const size_t base = size_t(blockIdx.x) * kScanTile;
	int cnt = 0;
	int bmin[6] = {INT_MAX, INT_MAX, INT_MAX, INT_MAX, INT_MAX, INT_MAX};
	// the thread's cell coordinates, divided out once and then carried
	// forward by kBlock ids per pass
	const size_t id0 = base + threadIdx.x;
	const size_t row = id0 / size_t(g.n[0]);
	int i = int(id0 - row * size_t(g.n[0]));
	int j = D == 2 ? int(row) : int(row % size_t(g.n[1]));
	int k = D == 2 ? 0 : int(row / size_t(g.n[1]));
	for (int sub = 0; sub < kScanSub; sub++) {
		const size_t id = base + size_t(sub) * kBlock + threadIdx.x;
		if (id >= N) break;
		int flag = 0;
		if (phi[id] <= 0) {
#pragma unroll
			for (int o = 0; o < 2 * D; o++) {
				const int ax = o >> 1, up = o & 1;
				const size_t f = gidx<D>(fo.g[ax], i + (ax == 0 ? up : 0), j + (ax == 1 ? up : 0), k + (ax == 2 ? up : 0));
				if (fo.open[ax][f]) { flag = 1; break; }
			}
		}
		flags[id] = uint8_t(flag);
		cnt += flag;
		if (flag) {
			bmin[0] = min(bmin[0], i); bmin[1] = min(bmin[1], j); bmin[2] = min(bmin[2], k);
			bmin[3] = min(bmin[3], -i); bmin[4] = min(bmin[4], -j); bmin[5] = min(bmin[5], -k);
		}
		i += kBlock;
		while (i >= g.n[0]) {
			i -= g.n[0];
			if (++j == g.n[1] && D == 3) { j = 0; k++; }
		}
	}
	int total;
	(void)blockExclusiveScan(cnt, &total);
	if (threadIdx.x == 0) tileCount[blockIdx.x] = total;
	// the tile's bounding box: wave minima, the workgroup's through LDS, then
	// one atomic per bound and tile holding unknowns (integer minima: the
	// result does not depend on the order)
	__shared__ int wmin[6][kBlock / kWave];
	const int lane = int(threadIdx.x) & (kWave - 1), wid = int(threadIdx.x) / kWave;
#pragma unroll
	for (int q = 0; q < 6; q++) {
		int v = bmin[q];
#pragma unroll
		for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_down(v, off, kWave));
		if (lane == 0) wmin[q][wid] = v;
	}
	__syncthreads();
	if (threadIdx.x < 6 && total > 0) {
		int v = INT_MAX;
		for (int w = 0; w < kBlock / kWave; w++) v = min(v, wmin[threadIdx.x][w]);
		atomicMin(&box[threadIdx.x], v);
	}
}

// tileBase[t] = unknowns in tiles before t; *count = all unknowns
__global__ __launch_bounds__(kBlock) void kScanTiles(const int *tileCount, int tiles, int *tileBase, int *count)
{
	int carry = 0;
	for (int t0 = 0; t0 < tiles; t0 += kBlock) {
		const int t = t0 + int(threadIdx.x);
		const int v = t < tiles ? tileCount[t] : 0;
		int total;
		const int ex = blockExclusiveScan(v, &total);
		if (t < tiles) tileBase[t] = carry + ex;
		carry += total;
	}
	if (threadIdx.x == 0) *count = carry;
}

__global__ __launch_bounds__(kBlock) void kNumberUnknowns(const uint8_t *flags, const int *tileBase, int *grid2mat, int *mat2grid, size_t N)
{
	const size_t base = size_t(blockIdx.x) * kScanTile;
	int carry = tileBase[blockIdx.x];
	for (int sub = 0; sub < kScanSub; sub++) {
		const size_t id = base + size_t(sub) * kBlock + threadIdx.x;
		const int flag = id < N ? flags[id] : 0;
		int total;
		const int idx = carry + blockExclusiveScan(flag, &total);
		if (id < N) {
			grid2mat[id] = flag ? idx : -1;
			if (flag) mat2grid[idx] = int(id);
		}
		carry += total;
	}
}

void launchScanTiles(const int *tileCount, int tiles, int *tileBase, int *count, hipStream_t s)
{
	hipLaunchKernelGGL(kScanTiles, dim3(1), dim3(kBlock), 0, s, tileCount, tiles, tileBase, count);
}

void launchNumberFlags(const uint8_t *flags, const int *tileBase, int tiles, int *grid2mat, int *mat2grid, size_t N, hipStream_t s)
{
	hipLaunchKernelGGL(kNumberUnknowns, dim3(tiles), dim3(kBlock), 0, s, flags, tileBase, grid2mat, mat2grid, N);
}

// Projector::buildLinearSystem (Projector.cpp:30-68), one unknown per thread.
template <int D>
__global__ __launch_bounds__(kBlock) void kAssemble(GridDev g, FaceFracs fr, FaceFracs uf, FaceFracs bv, const double *phi,
	const int *grid2mat, const int *mat2grid, int n, Rows<double> P, Rows<float> P32, double *rhs)
{
	const int r = int(blockIdx.x * blockDim.x + threadIdx.x);
	if (r >= n) return;
	const int cid = mat2grid[r];
	int c[3];
	c[0] = cid % g.n[0];
	if constexpr (D == 2) { c[1] = cid / g.n[0]; c[2] = 0; }
	else { c[1] = (cid / g.n[0]) % g.n[1]; c[2] = cid / g.n[0] / g.n[1]; }
	const double phic = phi[cid];
	double diag = 0, div = 0;
#pragma unroll
	for (int o = 0; o < 2 * D; o++) {
		const int ax = o >> 1, up = o & 1, side = up ? 1 : -1;
		int q[3] = {c[0], c[1], c[2]}, f[3] = {c[0], c[1], c[2]};
		q[ax] += side;
		f[ax] += up;
		const size_t fid = gidx<D>(fr.g[ax], f[0], f[1], f[2]);
		const double w = 1 - fr.a[ax][fid];
		int nbIdx = -1;
		double nbW = 0;
		// an unknown always has its 2D neighbours inside the grid when the
		// outermost faces are solid (fraction 1); an uploaded fraction field
		// that opens an outer face is treated as closed there
		const bool nbInside = q[ax] >= 0 && q[ax] < g.n[ax];
		if (w > 0 && nbInside) {
			const size_t qid = gidx<D>(g, q[0], q[1], q[2]);
			const int cc = grid2mat[qid];
			if (cc >= 0) {
				diag += w;
				nbIdx = cc;
				nbW = w;
			} else {
				const double theta = phic / (phic - phi[qid]);
				const double intf = 1 / fmax(theta, .001);
				diag += w * intf;
			}
			div += side * w * uf.a[ax][fid];
		}
		if (w < 1 || !nbInside) div += side * (w < 1 ? 1 - w : 1.0) * bv.a[ax][fid];
		P.nb[o][r] = nbIdx;
		P.w[o][r] = nbW;
		if (P32.w[0]) P32.w[o][r] = float(nbW);
	}
	rhs[r] = div;
	P.diag[r] = diag;
	const double inv = diag != 0 ? 1 / diag : 1.0;   // DiagonalPreconditioner (Eigen)
	P.invdiag[r] = inv;
	if (P32.w[0]) { P32.diag[r] = float(diag); P32.invdiag[r] = float(inv); }
}

// Projector::applyPressureGradient (Projector.cpp:91-123), no pressure jump.
template <int D>
__global__ __launch_bounds__(kBlock) void kApplyPressureGradient(GridDev fg, GridDev cg, int axis, double *u, const double *phi,
	const int *grid2mat, const double *x)
{
	const Coord c = threadCoord();
	if (!inGrid<D>(fg, c)) return;
	const int i0 = c.i - (axis == 0), j0 = c.j - (axis == 1), k0 = c.k - (axis == 2);
	if (i0 < 0 || j0 < 0 || k0 < 0) return;
	if (c.i >= cg.n[0] || c.j >= cg.n[1] || (D == 3 && c.k >= cg.n[2])) return;
	const size_t c0 = gidx<D>(cg, i0, j0, k0), c1 = gidx<D>(cg, c.i, c.j, c.k);
	const int id0 = grid2mat[c0], id1 = grid2mat[c1];
	if (id0 < 0 && id1 < 0) return;
	const double p0 = id0 >= 0 ? x[id0] : 0;
	const double p1 = id1 >= 0 ? x[id1] : 0;
	const size_t f = gidx<D>(fg, c.i, c.j, c.k);
	if (id0 >= 0 && id1 >= 0) u[f] += p1 - p0;
	else {
		const double phi0 = phi[c0], phi1 = phi[c1];
		if (phi0 <= 0 && phi1 <= 0) return;
		const double theta = phi0 / (phi0 - phi1);
		const double intf = 1 / fmax((phi0 <= 0 ? theta : 1 - theta), .001);
		u[f] += (p1 - p0) * intf;
	}
}

// b = A x on cell fields for the unknowns of the last assembly (0 elsewhere)
template <int D>
__global__ __launch_bounds__(kBlock) void kApplySystemCells(Rows<double> A, int n, const int *mat2grid, const double *x, double *b)
{
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
		double s = A.diag[i] * x[mat2grid[i]];
#pragma unroll
		for (int o = 0; o < 2 * D; o++) {
			const int c = A.nb[o][i];
			if (c >= 0) s += -A.w[o][i] * x[mat2grid[c]];
		}
		b[mat2grid[i]] = s;
	}
}

// x == nullptr writes the unknown indicator instead
__global__ void kScatterCompact(const int *grid2mat, const double *x, double *out, size_t n)
{
	for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
		const int m = grid2mat[i];
		out[i] = m >= 0 ? (x ? x[m] : 1.0) : 0.0;
	}
}

// ---------------------------------------------------------------- driver
static unsigned linBlocks(size_t n)
{
	const size_t b = (n + kBlock - 1) / kBlock;
	return unsigned(b < 8192 ? (b ? b : 1) : 8192);
}

Projection::~Projection() { release(); }

void Projection::release()
{
	auto fr = [](auto *&p) { if (p) { (void)hipFree(p); p = nullptr; } };
	fr(flags); fr(grid2mat); fr(mat2grid); fr(tile_count); fr(tile_base);
	for (int o = 0; o < 6; o++) { fr(rows.nb[o]); fr(rows.w[o]); fr(rows32.w[o]); }
	fr(rows.diag); fr(rows.invdiag); fr(rhs); fr(rows32.diag); fr(rows32.invdiag);
	for (int o = 0; o < 6; o++) rows32.nb[o] = nullptr;
	K.release();
	cap = 0;
}

bool Projection::reserveGrid(size_t N, std::string &err)
{
	if (flags) return true;
	tiles = int((N + kScanTile - 1) / kScanTile);
	if (hipMalloc(&flags, N) != hipSuccess || hipMalloc(&grid2mat, N * sizeof(int)) != hipSuccess ||
		hipMalloc(&tile_count, size_t(tiles) * sizeof(int)) != hipSuccess || hipMalloc(&tile_base, size_t(tiles + 8) * sizeof(int)) != hipSuccess) {
		err = "hipMalloc failed for unknown maps";
		return false;
	}
	return true;
}

bool Projection::reserveRows(int D, size_t n, bool mixed, std::string &err)
{
	auto fr = [](auto *&p) { if (p) { (void)hipFree(p); p = nullptr; } };
	if (n > cap || !mat2grid) {
		const size_t want = std::max<size_t>(n + n / 4, 1024);
		fr(mat2grid);
		for (int o = 0; o < 6; o++) { fr(rows.nb[o]); fr(rows.w[o]); fr(rows32.w[o]); }
		fr(rows.diag); fr(rows.invdiag); fr(rhs); fr(rows32.diag); fr(rows32.invdiag);
		bool ok = hipMalloc(&mat2grid, want * sizeof(int)) == hipSuccess;
		for (int o = 0; o < 2 * D && ok; o++)
			ok = hipMalloc(&rows.nb[o], want * sizeof(int)) == hipSuccess && hipMalloc(&rows.w[o], want * sizeof(double)) == hipSuccess;
		for (double **q : {&rows.diag, &rows.invdiag, &rhs}) ok = ok && hipMalloc(q, want * sizeof(double)) == hipSuccess;
		if (!ok) { err = "hipMalloc failed for the pressure system"; return false; }
		cap = want;
	}
	if (mixed && !rows32.diag) {
		bool ok = true;
		for (int o = 0; o < 2 * D && ok; o++) ok = hipMalloc(&rows32.w[o], cap * sizeof(float)) == hipSuccess;
		ok = ok && hipMalloc(&rows32.diag, cap * sizeof(float)) == hipSuccess && hipMalloc(&rows32.invdiag, cap * sizeof(float)) == hipSuccess;
		if (!ok) { err = "hipMalloc failed for the FP32 pressure system"; return false; }
	}
	for (int o = 0; o < 6; o++) rows32.nb[o] = rows.nb[o];
	return K.reserve(n, mixed, err);
}

static FaceFracs faceArrays(const Fields &F, double *const *a)
{
	FaceFracs f{};
	for (int ax = 0; ax < F.D; ax++) { f.a[ax] = a[ax]; f.g[ax] = F.face[ax]; }
	return f;
}

template <int D>
static int numberUnknowns(const Fields &F, Projection &P, bool mixed, hipStream_t s, std::string &err)
{
	// tile_base[tiles] receives the count, tile_base[tiles + 1 ..] the bounding box
	int *box = P.tile_base + P.tiles + 1;
	hipLaunchKernelGGL(kBoxInit, dim3(1), dim3(64), 0, s, box);
	OpenFaces fo{};
	for (int a = 0; a < D; a++) { fo.g[a] = F.face[a]; fo.open[a] = F.open_f[a]; }
	hipLaunchKernelGGL(kUnknownFlags<D>, dim3(P.tiles), dim3(kBlock), 0, s, F.cell, fo, (const double *)F.phi, P.flags, P.tile_count, F.N, box);
	hipLaunchKernelGGL(kScanTiles, dim3(1), dim3(kBlock), 0, s, (const int *)P.tile_count, P.tiles, P.tile_base, P.tile_base + P.tiles);
	int *cnt = F.host_ints;
	if (hipMemcpyAsync(cnt, P.tile_base + P.tiles, 7 * sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess || spinWait(s, F.wait) != hipSuccess) {
		err = "reading the unknown count failed";
		return -1;
	}
	const int n = cnt[0];
	for (int a = 0; a < 3; a++) {
		P.box_lo[a] = a < D ? cnt[1 + a] : 0;
		P.box_n[a] = a < D ? -cnt[4 + a] - cnt[1 + a] + 1 : 1;
	}
	if (!P.reserveRows(D, size_t(n), mixed, err)) return -1;
	hipLaunchKernelGGL(kNumberUnknowns, dim3(P.tiles), dim3(kBlock), 0, s, (const uint8_t *)P.flags, (const int *)P.tile_base, P.grid2mat,
		P.mat2grid, F.N);
	return n;
}

template <int D>
static void assemble(const Fields &F, Projection &P, int n, bool mixed, hipStream_t s)
{
	if (n == 0) return;
	Rows<float> r32 = P.rows32;
	if (!mixed) r32.w[0] = nullptr;
	hipLaunchKernelGGL(kAssemble<D>, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, F.cell, faceArrays(F, F.frac),
		faceArrays(F, F.u), faceArrays(F, F.bvel), (const double *)F.phi, (const int *)P.grid2mat, (const int *)P.mat2grid, n, P.rows, r32, P.rhs);
}

// the unknowns' bounding box of the last numbering
static BoxLayout boxLayout(const Fields &F, const Projection &P)
{
	BoxLayout L{};
	for (int a = 0; a < 3; a++) { L.lo[a] = P.box_lo[a]; L.n[a] = P.box_n[a]; L.gn[a] = F.cell.n[a]; }
	L.mat2grid = P.mat2grid;
	return L;
}

template <int D>
static int projectD(Fields &F, Projection &P, const garm_pcg_opts &o, garm_pcg_stats *st, hipStream_t s, std::string &err)
{
	const bool learned = o.mode == GARM_PCG_LEARNED, mg = o.mode == GARM_PCG_MULTIGRID;
	const bool mixed = o.mode == GARM_PCG_MIXED || learned || mg;
	if (learned && !P.net) { err = "no preconditioner network is attached"; return GARM_EINVAL; }
	if (!P.reserveGrid(F.N, err)) return GARM_ENOMEM;
	const int n = numberUnknowns<D>(F, P, mixed, s, err);
	if (n < 0) return GARM_EHIP;
	P.n = n;
	st->unknowns = n;
	st->iterations = 0;
	st->outer_iterations = 0;
	st->fell_back = 0;
	st->preconditioner_applications = 0;
	st->relative_residual = 0;
	st->hbm_bytes = 0;
	st->solver_cells = 0;
	int rc = GARM_OK;
	if (n > 0) {
		assemble<D>(F, P, n, mixed, s);
		if (learned || mg) {
			const BoxLayout L = boxLayout(F, P);
			rc = learned ? solveLearned(D, P.rows, P.rows32, P.rhs, n, P.grid2mat, L, o, P.K, P.BK, P.LW, *P.net, P.MG, s, st, err)
			             : solveMultigrid(D, P.rows, P.rows32, P.rhs, n, L, o, P.K, P.BK, P.LW, P.MG, s, st, err);
			if ((rc == GARM_OK || rc == GARM_ENOCONV) && hipMemcpyAsync(P.K.x, P.LW.x, size_t(n) * sizeof(double), hipMemcpyDeviceToDevice, s) != hipSuccess) {
				err = "copying the learned solution failed";
				rc = GARM_EHIP;
			}
		} else {
			rc = solveJacobi(D, mixed, P.rows, P.rows32, P.rhs, n, boxLayout(F, P), o, P.K, P.BK, s, st, err, &P.used_box);
		}
	}
	// the velocity is updated with whatever the solver reached (the reference
	// exits instead when its solver fails, LinearSystem.h:34-37)
	for (int a = 0; a < D && n > 0; a++)
		hipLaunchKernelGGL(kApplyPressureGradient<D>, gridDims(F.face[a], D), gridBlock(), 0, s, F.face[a], F.cell, a, F.u[a],
			(const double *)F.phi, (const int *)P.grid2mat, (const double *)P.K.x);
	return rc;
}

int runProjection(Fields &F, Projection &P, const garm_pcg_opts &o, garm_pcg_stats *st, hipStream_t s, std::string &err)
{
	return F.D == 2 ? projectD<2>(F, P, o, st, s, err) : projectD<3>(F, P, o, st, s, err);
}

void applySystemCells(const Fields &F, const Projection &P, const double *x, double *b, hipStream_t s)
{
	launchFill(b, 0.0, F.N, s);
	if (!P.grid2mat || P.n == 0) return;
	const dim3 g(std::min(8192, (P.n + kBlock - 1) / kBlock));
	if (F.D == 2) hipLaunchKernelGGL(kApplySystemCells<2>, g, dim3(kBlock), 0, s, P.rows, P.n, (const int *)P.mat2grid, x, b);
	else hipLaunchKernelGGL(kApplySystemCells<3>, g, dim3(kBlock), 0, s, P.rows, P.n, (const int *)P.mat2grid, x, b);
}

void scatterCompact(const Fields &F, const Projection &P, int which, double *out, hipStream_t s)
{
	if (!P.grid2mat || P.n == 0) { launchFill(out, 0.0, F.N, s); return; }
	const double *x = which == kCompactPressure ? P.K.x : which == kCompactRhs ? P.rhs : nullptr;
	hipLaunchKernelGGL(kScatterCompact, dim3(linBlocks(F.N)), dim3(kBlock), 0, s, (const int *)P.grid2mat, x, out, F.N);
}

} // namespace garm
