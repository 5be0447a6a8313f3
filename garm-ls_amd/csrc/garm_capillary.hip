// Surface tension for gfx950: SurfaceTension::solve (SurfaceTension.h:12-101)
// and LevelSetLiquid::trickLittleDrop (LevelSetLiquid.cpp:1037-1057), FP64
// with the reference's operations (compiled without multiply-add
// contraction).
//
// The interface faces (|phi| < band dx at the face centre under the cubic
// view of phi, fraction < 1) are numbered over the D face grids taken as one
// axis-major index space by the same three tile passes as the pressure
// unknowns, so the numbering is the reference's forEach(velocity.grids)
// order.  One thread per interface face then evaluates the normal, the
// curvature, the smoothed delta and the velocity's first and second
// differences and writes its row of the capillary operator as
// (diagonal, up to 2D neighbour weights); the system goes to the FP64 Jacobi
// PCG of the pressure solve and the solution replaces the velocity on those
// faces.
#include "garm_capillary.hpp"
#include "garm_views.hpp"

#include <climits>
#include <cmath>

namespace garm {

// The face grids as one index space: grid a holds ids off[a] .. off[a+1]-1.
struct FaceSpace {
	GridDev g[3];
	size_t off[4];
	const double *frac[3];
	double *u[3];
};

template <int D>
__device__ __forceinline__ int faceAxis(const FaceSpace &S, size_t id)
{
	if constexpr (D == 2) return id >= S.off[1] ? 1 : 0;
	else return id >= S.off[2] ? 2 : (id >= S.off[1] ? 1 : 0);
}

template <int D>
__device__ __forceinline__ void faceCoord(const GridDev &g, size_t local, int *c)
{
	c[0] = int(local % size_t(g.n[0]));
	if constexpr (D == 2) { c[1] = int(local / size_t(g.n[0])); c[2] = 0; }
	else {
		c[1] = int((local / size_t(g.n[0])) % size_t(g.n[1]));
		c[2] = int(local / (size_t(g.n[0]) * g.n[1]));
	}
}

template <int D>
__device__ __forceinline__ void facePosition(const GridDev &g, const int *c, double *p)
{
#pragma unroll
	for (int a = 0; a < D; a++) p[a] = g.o[a] + c[a] * g.dx;
}

// Interface faces (SurfaceTension.h:53-61), counted per tile.
template <int D>
__global__ __launch_bounds__(kBlock) void kCapFlags(FaceSpace S, GridDev cg, const double *phi, double eps, uint8_t *flags,
	int *tileCount, size_t N)
{
	const size_t base = size_t(blockIdx.x) * kScanTile;
	int cnt = 0;
	for (int sub = 0; sub < kScanSub; sub++) {
		const size_t id = base + size_t(sub) * kBlock + threadIdx.x;
		if (id >= N) break;
		const int ax = faceAxis<D>(S, id);
		const size_t local = id - S.off[ax];
		int flag = 0;
		if (S.frac[ax][local] < 1) {
			int c[3];
			faceCoord<D>(S.g[ax], local, c);
			double p[3] = {0, 0, 0};
			facePosition<D>(S.g[ax], c, p);
			const double v = cubicView<D>(cg, phi, p);
			flag = -eps < v && v < eps;
		}
		flags[id] = uint8_t(flag);
		cnt += flag;
	}
	int total;
	(void)blockExclusiveScan(cnt, &total);
	if (threadIdx.x == 0) tileCount[blockIdx.x] = total;
}

// Derivatives::getFirst<2> (Derivatives.h:47, 51-57) of a face field along b
// at the unclamped coordinate q, stencil points clamped
template <int D>
__device__ __forceinline__ double firstCentral(const GridDev &g, const double *u, const int *q, int b)
{
	const double hi = u[gat<D>(g, q[0] + (b == 0), q[1] + (b == 1), q[2] + (b == 2))];
	const double lo = u[gat<D>(g, q[0] - (b == 0), q[1] - (b == 1), q[2] - (b == 2))];
	return (hi - lo) / 2 * g.inv;
}

constexpr double kPi = 3.141592653589793;   // std::numbers::pi

struct CapCoef {
	double coeff, dt, eps, weight;   // weight = coeff dt^2 / dx^2
};

// One row of the capillary operator per interface face
// (SurfaceTension.h:67-90): diagonal 1/delta + weight per open neighbour,
// off-diagonal -weight to interface neighbours, the other neighbours'
// velocities moved to the right-hand side.
template <int D>
__global__ __launch_bounds__(kBlock) void kCapRows(FaceSpace S, GridDev cg, const double *phi, const int *grid2mat, const int *mat2grid,
	int n, CapCoef k, Rows<double> R, double *rhs)
{
	const int r = int(blockIdx.x * blockDim.x + threadIdx.x);
	if (r >= n) return;
	const size_t id = size_t(unsigned(mat2grid[r]));
	const int ax = faceAxis<D>(S, id);
	const size_t local = id - S.off[ax];
	const GridDev &fg = S.g[ax];
	int c[3];
	faceCoord<D>(fg, local, c);
	double p[3] = {0, 0, 0};
	facePosition<D>(fg, c, p);
	const double inv = cg.inv;

	// normal and curvature (getCurvature, SurfaceTension.h:12-23)
	double gr[3] = {0, 0, 0}, nrm[3] = {0, 0, 0};
	cubicViewGradient<D>(cg, phi, p, gr);
	normalizeD<D>(gr, nrm);
	double kappa = 0;
#pragma unroll
	for (int a = 0; a < D; a++) {
		double q[3] = {p[0], p[1], p[2]}, gp[3] = {0, 0, 0}, gm[3] = {0, 0, 0};
		q[a] = p[a] + cg.dx;
		cubicViewGradient<D>(cg, phi, q, gp);
		normalizeD<D>(gp, gp);
		q[a] = p[a] - cg.dx;
		cubicViewGradient<D>(cg, phi, q, gm);
		normalizeD<D>(gm, gm);
		kappa += gp[a] - gm[a];
	}
	kappa *= inv * .5;
	if (fabs(kappa) > inv) kappa = (kappa < 0 ? -1 : 1) * inv;

	// smoothed delta of the band (SurfaceTension.h:47-51, 74)
	const double v = cubicView<D>(cg, phi, p);
	double dirac = 0;
	if (!(v < -k.eps) && !(v > k.eps)) dirac = 0.5 * (1.0 + cos(kPi * v / k.eps)) / k.eps;
	const double delta = fmax(dirac, 1e-7 * inv);

	// first and second differences of this velocity component (Derivatives.h:59-87)
	const double *U = S.u[ax];
	double jac[3] = {0, 0, 0}, hes[3][3] = {};
#pragma unroll
	for (int b = 0; b < D; b++) jac[b] = firstCentral<D>(fg, U, c, b);
#pragma unroll
	for (int a1 = 0; a1 < D; a1++)
#pragma unroll
		for (int a2 = 0; a2 < D; a2++) {
			int qp[3] = {c[0], c[1], c[2]}, qm[3] = {c[0], c[1], c[2]};
			qp[a2] += 1;
			qm[a2] -= 1;
			hes[a1][a2] = (firstCentral<D>(fg, U, qp, a1) - firstCentral<D>(fg, U, qm, a1)) / 2 * fg.inv;
		}
	// n^T H n as a row vector times n, and J . n
	double nHn = 0, jn = 0;
#pragma unroll
	for (int j = 0; j < D; j++) {
		double t = 0;
#pragma unroll
		for (int i = 0; i < D; i++) t += nrm[i] * hes[i][j];
		nHn += t * nrm[j];
	}
#pragma unroll
	for (int i = 0; i < D; i++) jn += jac[i] * nrm[i];

	const size_t f = gidx<D>(fg, c[0], c[1], c[2]);
	double b = U[f] / delta - k.coeff * k.dt * (kappa * nrm[ax] + k.dt * (nHn + kappa * jn));
	double diag = 1 / delta;
#pragma unroll
	for (int o = 0; o < 2 * D; o++) {
		const int a = o >> 1;
		int q[3] = {c[0], c[1], c[2]};
		q[a] += (o & 1) ? 1 : -1;
		int m = -1;
		double w = 0;
		const bool inside = q[a] >= 0 && q[a] < fg.n[a];
		if (inside) {
			const size_t qf = gidx<D>(fg, q[0], q[1], q[2]);
			if (S.frac[ax][qf] != 1) {
				diag += k.weight;
				m = grid2mat[S.off[ax] + qf];
				if (m != -1) w = k.weight;
				else b += k.weight * U[qf];
			}
		}
		R.nb[o][r] = m;
		R.w[o][r] = w;
	}
	R.diag[r] = diag;
	R.invdiag[r] = 1 / diag;
	rhs[r] = b;
}

__global__ __launch_bounds__(kBlock) void kCapScatter3(FaceSpace S, int D, const int *mat2grid, const double *x, int n)
{
	for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
		const size_t id = size_t(unsigned(mat2grid[r]));
		const int ax = D == 3 && id >= S.off[2] ? 2 : (id >= S.off[1] ? 1 : 0);
		S.u[ax][id - S.off[ax]] = x[r];
	}
}

// LevelSetLiquid::trickLittleDrop with checkDrop (LevelSetLiquid.h:178-190):
// a liquid cell whose 2D faces are all open and whose neighbours are all air
// has its two y-faces pushed down by vt + v when it rises slowly.  Two such
// cells are never neighbours, so no face is written twice.
template <int D>
__global__ __launch_bounds__(kBlock) void kLittleDrop(GridDev cg, FaceSpace S, const double *phi, double vt)
{
	const Coord c = threadCoord();
	if (!inGrid<D>(cg, c)) return;
	if (phi[gidx<D>(cg, c.i, c.j, c.k)] > 0) return;
#pragma unroll
	for (int o = 0; o < 2 * D; o++) {
		const int a = o >> 1, up = o & 1;
		const int s = up ? 1 : -1;
		const size_t f = gidx<D>(S.g[a], c.i + (a == 0 ? up : 0), c.j + (a == 1 ? up : 0), c.k + (a == 2 ? up : 0));
		if (S.frac[a][f] != 0) return;
		if (phi[gat<D>(cg, c.i + (a == 0 ? s : 0), c.j + (a == 1 ? s : 0), c.k + (a == 2 ? s : 0))] <= 0) return;
	}
	double *u1 = S.u[1];
	const size_t f0 = gidx<D>(S.g[1], c.i, c.j, c.k), f1 = gidx<D>(S.g[1], c.i, c.j + 1, c.k);
	const double v = (u1[f0] + u1[f1]) * .5;
	if (v < vt * 1.25 && v > -vt && u1[f0] > 0 && u1[f1] < 0) {
		u1[f0] -= vt + v;
		u1[f1] -= vt + v;
	}
}

// ---------------------------------------------------------------- driver
Capillary::~Capillary() { release(); }

void Capillary::release()
{
	auto fr = [](auto *&p) { if (p) { (void)hipFree(p); p = nullptr; } };
	fr(flags); fr(grid2mat); fr(mat2grid); fr(tile_count); fr(tile_base); fr(rhs);
	for (int o = 0; o < 6; o++) { fr(rows.nb[o]); fr(rows.w[o]); }
	fr(rows.diag); fr(rows.invdiag);
	K.release();
	cap = 0;
	faces = 0;
}

static FaceSpace faceSpace(const Fields &F)
{
	FaceSpace S{};
	size_t off = 0;
	for (int a = 0; a < F.D; a++) {
		S.g[a] = F.face[a];
		S.off[a] = off;
		S.frac[a] = F.frac[a];
		S.u[a] = F.u[a];
		off += F.Nf[a];
	}
	for (int a = F.D; a < 4; a++) S.off[a] = off;
	return S;
}

static bool reserveCapillary(Capillary &C, int D, size_t N, size_t n, std::string &err)
{
	auto fr = [](auto *&p) { if (p) { (void)hipFree(p); p = nullptr; } };
	if (C.faces != N) {
		fr(C.flags); fr(C.grid2mat); fr(C.tile_count); fr(C.tile_base);
		C.tiles = int((N + kScanTile - 1) / kScanTile);
		if (hipMalloc(&C.flags, N) != hipSuccess || hipMalloc(&C.grid2mat, N * sizeof(int)) != hipSuccess ||
			hipMalloc(&C.tile_count, size_t(C.tiles) * sizeof(int)) != hipSuccess ||
			hipMalloc(&C.tile_base, size_t(C.tiles + 1) * sizeof(int)) != hipSuccess) {
			err = "hipMalloc failed for the capillary face maps";
			return false;
		}
		C.faces = N;
	}
	if (n > C.cap || !C.mat2grid) {
		const size_t want = std::max<size_t>(n + n / 4, 1024);
		fr(C.mat2grid); fr(C.rhs); fr(C.rows.diag); fr(C.rows.invdiag);
		for (int o = 0; o < 6; o++) { fr(C.rows.nb[o]); fr(C.rows.w[o]); }
		bool ok = hipMalloc(&C.mat2grid, want * sizeof(int)) == hipSuccess;
		for (int o = 0; o < 2 * D && ok; o++)
			ok = hipMalloc(&C.rows.nb[o], want * sizeof(int)) == hipSuccess && hipMalloc(&C.rows.w[o], want * sizeof(double)) == hipSuccess;
		for (double **q : {&C.rows.diag, &C.rows.invdiag, &C.rhs}) ok = ok && hipMalloc(q, want * sizeof(double)) == hipSuccess;
		if (!ok) { err = "hipMalloc failed for the capillary system"; return false; }
		C.cap = want;
	}
	return n == 0 || C.K.reserve(n, false, err);
}

template <int D>
static int capillaryD(Fields &F, Capillary &C, const CapillaryArgs &a, hipStream_t s, garm_pcg_stats *st, std::string &err)
{
	*st = garm_pcg_stats{};
	const FaceSpace S = faceSpace(F);
	const size_t N = S.off[D];
	if (N >= size_t(INT_MAX)) { err = "the face grids hold too many faces for the capillary numbering"; return GARM_EINVAL; }
	if (!reserveCapillary(C, D, N, 0, err)) return GARM_ENOMEM;
	const double dx = F.cell.dx;
	CapCoef k{a.coeff, a.dt, a.band * dx, a.coeff * a.dt * a.dt / (dx * dx)};
	hipLaunchKernelGGL(kCapFlags<D>, dim3(C.tiles), dim3(kBlock), 0, s, S, F.cell, (const double *)F.phi, k.eps, C.flags, C.tile_count, N);
	launchScanTiles(C.tile_count, C.tiles, C.tile_base, C.tile_base + C.tiles, s);
	int *cnt = F.host_ints;
	if (hipMemcpyAsync(cnt, C.tile_base + C.tiles, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess || spinWait(s, F.wait) != hipSuccess) {
		err = "reading the interface face count failed";
		return GARM_EHIP;
	}
	const int n = cnt[0];
	st->unknowns = n;
	if (n == 0) return GARM_OK;
	if (!reserveCapillary(C, D, N, size_t(n), err)) return GARM_ENOMEM;
	launchNumberFlags(C.flags, C.tile_base, C.tiles, C.grid2mat, C.mat2grid, N, s);
	hipLaunchKernelGGL(kCapRows<D>, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, S, F.cell, (const double *)F.phi,
		(const int *)C.grid2mat, (const int *)C.mat2grid, n, k, C.rows, C.rhs);
	garm_pcg_opts o{};
	o.mode = GARM_PCG_FP64;
	o.tolerance = a.tolerance;
	o.max_iterations = a.max_iterations;
	o.check_interval = a.check_interval;
	const int rc = solvePcg64(D, C.rows, C.rhs, n, o, C.K, s, st, err);
	// the velocity takes whatever the solver reached (the reference exits
	// instead when its solver fails, LinearSystem.h:34-37)
	if (rc == GARM_OK || rc == GARM_ENOCONV)
		hipLaunchKernelGGL(kCapScatter3, dim3(std::min(8192, (n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, S, D, (const int *)C.mat2grid,
			(const double *)C.K.x, n);
	return rc;
}

int runCapillary(Fields &F, Capillary &C, const CapillaryArgs &a, hipStream_t s, garm_pcg_stats *st, std::string &err)
{
	return F.D == 2 ? capillaryD<2>(F, C, a, s, st, err) : capillaryD<3>(F, C, a, s, st, err);
}

void launchLittleDrop(const Fields &F, double dt, hipStream_t s)
{
	const double v0 = F.cell.dx / dt;
	const double vt = v0 * .15;
	const FaceSpace S = faceSpace(F);
	if (F.D == 2) hipLaunchKernelGGL(kLittleDrop<2>, gridDims(F.cell, 2), gridBlock(), 0, s, F.cell, S, (const double *)F.phi, vt);
	else hipLaunchKernelGGL(kLittleDrop<3>, gridDims(F.cell, 3), gridBlock(), 0, s, F.cell, S, (const double *)F.phi, vt);
}

} // namespace garm
