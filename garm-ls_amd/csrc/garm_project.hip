// Ghost-fluid pressure projection for gfx950 (Projector.cpp:13-163).
//
// Unknowns are numbered in the reference's lexicographic order
// (Projector::setUnknowns, Projector.cpp:125-156) by a device prefix sum.
// The 5/7-point operator is kept as one diagonal plus 2D (neighbour, weight)
// pairs per unknown, structure-of-arrays, and applied matrix-free: no sparse
// matrix is assembled and no factorisation is formed.  The conjugate-gradient
// scalars (alpha, beta, r.z, |r|^2) never leave the device; the host only
// polls a convergence flag every few iterations.
#include "garm_project.hpp"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

namespace garm {

// ------------------------------------------------------------- unknowns
template <int D>
__global__ __launch_bounds__(kBlock) void kUnknownFlags(GridDev g, FaceFracs fr, const double *phi, int *flags)
{
	const Coord c = threadCoord();
	if (!inGrid<D>(g, c)) return;
	const size_t id = gidx<D>(g, c.i, c.j, c.k);
	int flag = 0;
	if (phi[id] <= 0) {
#pragma unroll
		for (int o = 0; o < 2 * D; o++) {
			const int ax = o >> 1, up = o & 1;
			const GridDev &fg = fr.g[ax];
			const size_t f = gidx<D>(fg, c.i + (ax == 0 ? up : 0), c.j + (ax == 1 ? up : 0), c.k + (ax == 2 ? up : 0));
			if (fr.a[ax][f] < 1) { flag = 1; break; }
		}
	}
	flags[id] = flag;
}

__global__ void kCompactUnknowns(const int *flags, int *grid2mat, int *mat2grid, size_t n)
{
	for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
		if (flags[i]) mat2grid[grid2mat[i]] = int(i);
		else grid2mat[i] = -1;
	}
}

// Projector::buildLinearSystem (Projector.cpp:30-68), one unknown per thread.
template <int D>
__global__ __launch_bounds__(kBlock) void kAssemble(GridDev g, FaceFracs fr, FaceFracs uf, FaceFracs bv, const double *phi,
	const int *grid2mat, const int *mat2grid, int n, PoissonRows P)
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
	}
	P.rhs[r] = div;
	P.diag[r] = diag;
	P.invdiag[r] = diag != 0 ? 1 / diag : 1.0;   // DiagonalPreconditioner (Eigen)
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

__global__ void kScatterPressure(const int *grid2mat, const double *x, double *out, size_t n)
{
	for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
		const int m = grid2mat[i];
		out[i] = m >= 0 ? x[m] : 0.0;
	}
}

// ------------------------------------------------------------------- PCG
// Device scalar slots (PcgScalars): see garm_project.hpp.

// y = A p (diag p - sum w p_nb) and the block partial of p.y.
template <int D>
__global__ __launch_bounds__(kBlock) void kSpmvDot(PoissonRows P, int n, const double *p, double *y, double *partials, const int *state)
{
	if (state[kStDone]) return;
	const int r = int(blockIdx.x * blockDim.x + threadIdx.x);
	double v[1] = {0};
	if (r < n) {
		const double pr = p[r];
		double s = P.diag[r] * pr;
#pragma unroll
		for (int o = 0; o < 2 * D; o++) {
			const int c = P.nb[o][r];
			if (c >= 0) s += -P.w[o][r] * p[c];
		}
		y[r] = s;
		v[0] = pr * s;
	}
	double out[1];
	blockSum<1>(v, out);
	if (threadIdx.x == 0) partials[blockIdx.x] = out[0];
}

// alpha = (r.z) / (p.Ap)
__global__ __launch_bounds__(kBlock) void kAlpha(const double *partials, int nb, double *sc, const int *state)
{
	if (state[kStDone]) return;
	double v[1] = {0};
	for (int i = threadIdx.x; i < nb; i += blockDim.x) v[0] += partials[i];
	double o[1];
	blockSum<1>(v, o);
	if (threadIdx.x == 0) sc[kScAlpha] = sc[kScAbsNew] / o[0];
}

// x += alpha p; r -= alpha Ap; partials of r.r and r.(M^-1 r)
__global__ __launch_bounds__(kBlock) void kUpdateXR(int n, double *x, double *r, const double *p, const double *y, const double *invdiag,
	double *partials, const double *sc, const int *state)
{
	if (state[kStDone]) return;
	const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
	double v[2] = {0, 0};
	if (i < n) {
		const double alpha = sc[kScAlpha];
		x[i] += alpha * p[i];
		const double ri = r[i] - alpha * y[i];
		r[i] = ri;
		v[0] = ri * ri;
		v[1] = ri * (invdiag[i] * ri);
	}
	double o[2];
	blockSum<2>(v, o);
	if (threadIdx.x == 0) { partials[2 * blockIdx.x] = o[0]; partials[2 * blockIdx.x + 1] = o[1]; }
}

// Convergence test ||r||^2 < threshold, then beta = (r.z)_new / (r.z)_old.
__global__ __launch_bounds__(kBlock) void kBeta(const double *partials, int nb, double *sc, int *state)
{
	if (state[kStDone]) return;
	double v[2] = {0, 0};
	for (int i = threadIdx.x; i < nb; i += blockDim.x) { v[0] += partials[2 * i]; v[1] += partials[2 * i + 1]; }
	double o[2];
	blockSum<2>(v, o);
	if (threadIdx.x == 0) {
		sc[kScRn2] = o[0];
		if (o[0] < sc[kScThreshold]) { state[kStDone] = 1; return; }
		const double absOld = sc[kScAbsNew];
		sc[kScAbsNew] = o[1];
		sc[kScBeta] = o[1] / absOld;
		const int it = state[kStIter] + 1;
		state[kStIter] = it;
		if (it >= state[kStMaxIter]) state[kStDone] = 1;
	}
}

// p = M^-1 r + beta p
__global__ void kUpdateP(int n, double *p, const double *r, const double *invdiag, const double *sc, const int *state)
{
	if (state[kStDone]) return;
	const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
	if (i < n) p[i] = invdiag[i] * r[i] + sc[kScBeta] * p[i];
}

// Initial state: x = 0, r = b, p = M^-1 r; partials of b.b and r.p.
__global__ __launch_bounds__(kBlock) void kPcgStart(int n, const double *b, const double *invdiag, double *x, double *r, double *p, double *partials)
{
	const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
	double v[2] = {0, 0};
	if (i < n) {
		const double bi = b[i];
		x[i] = 0;
		r[i] = bi;
		const double z = invdiag[i] * bi;
		p[i] = z;
		v[0] = bi * bi;
		v[1] = bi * z;
	}
	double o[2];
	blockSum<2>(v, o);
	if (threadIdx.x == 0) { partials[2 * blockIdx.x] = o[0]; partials[2 * blockIdx.x + 1] = o[1]; }
}

// Tolerance handling of LinearSystem::solve (LinearSystem.h:31) and the
// start-up tests of Eigen's conjugate_gradient.
__global__ __launch_bounds__(kBlock) void kPcgStartFinal(const double *partials, int nb, double tol, int maxIt, double *sc, int *state)
{
	double v[2] = {0, 0};
	for (int i = threadIdx.x; i < nb; i += blockDim.x) { v[0] += partials[2 * i]; v[1] += partials[2 * i + 1]; }
	double o[2];
	blockSum<2>(v, o);
	if (threadIdx.x == 0) {
		const double b2 = o[0];
		sc[kScBnorm2] = b2;
		sc[kScRn2] = b2;
		sc[kScAbsNew] = o[1];
		state[kStIter] = 0;
		state[kStMaxIter] = maxIt;
		state[kStDone] = 0;
		if (b2 == 0) { state[kStDone] = 1; sc[kScThreshold] = 0; sc[kScTol] = tol; return; }
		const double t = fmax(tol, DBL_EPSILON / sqrt(b2));
		sc[kScTol] = t;
		sc[kScThreshold] = fmax(t * t * b2, DBL_MIN);
		if (b2 < sc[kScThreshold] || maxIt <= 0) state[kStDone] = 1;
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
	fr(flags); fr(grid2mat); fr(mat2grid); fr(scan_tmp);
	for (int o = 0; o < 6; o++) { fr(rows.nb[o]); fr(rows.w[o]); }
	fr(rows.diag); fr(rows.invdiag); fr(rows.rhs);
	fr(x); fr(r); fr(p); fr(y); fr(partials); fr(sc); fr(state);
	if (host_state) { (void)hipHostFree(host_state); host_state = nullptr; }
	cap = 0;
}

bool Projection::reserveGrid(size_t N, std::string &err)
{
	if (flags) return true;
	if (hipMalloc(&flags, N * sizeof(int)) != hipSuccess || hipMalloc(&grid2mat, N * sizeof(int)) != hipSuccess) {
		err = "hipMalloc failed for unknown maps";
		return false;
	}
	size_t bytes = 0;
	if (hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, flags, grid2mat, int(N)) != hipSuccess) { err = "scan sizing failed"; return false; }
	scan_bytes = bytes;
	if (hipMalloc(&scan_tmp, bytes ? bytes : 16) != hipSuccess) { err = "hipMalloc failed for scan storage"; return false; }
	if (hipMalloc(&sc, 32 * sizeof(double)) != hipSuccess || hipMalloc(&state, 8 * sizeof(int)) != hipSuccess) { err = "hipMalloc failed"; return false; }
	if (hipHostMalloc(&host_state, 8 * sizeof(int)) != hipSuccess) { err = "hipHostMalloc failed"; return false; }
	return true;
}

bool Projection::reserveRows(int D, size_t n, std::string &err)
{
	if (n <= cap && mat2grid) return true;
	size_t want = std::max<size_t>(n + n / 4, 1024);
	auto fr = [](auto *&p) { if (p) { (void)hipFree(p); p = nullptr; } };
	fr(mat2grid);
	for (int o = 0; o < 6; o++) { fr(rows.nb[o]); fr(rows.w[o]); }
	fr(rows.diag); fr(rows.invdiag); fr(rows.rhs); fr(x); fr(r); fr(p); fr(y); fr(partials);
	bool ok = hipMalloc(&mat2grid, want * sizeof(int)) == hipSuccess;
	for (int o = 0; o < 2 * D && ok; o++)
		ok = hipMalloc(&rows.nb[o], want * sizeof(int)) == hipSuccess && hipMalloc(&rows.w[o], want * sizeof(double)) == hipSuccess;
	for (double **q : {&rows.diag, &rows.invdiag, &rows.rhs, &x, &r, &p, &y})
		ok = ok && hipMalloc(q, want * sizeof(double)) == hipSuccess;
	const size_t nb = (want + kBlock - 1) / kBlock + 1;
	ok = ok && hipMalloc(&partials, 2 * nb * sizeof(double)) == hipSuccess;
	if (!ok) { err = "hipMalloc failed for the pressure system"; return false; }
	cap = want;
	return true;
}

static FaceFracs faceArrays(const Fields &F, double *const *a)
{
	FaceFracs f{};
	for (int ax = 0; ax < F.D; ax++) { f.a[ax] = a[ax]; f.g[ax] = F.face[ax]; }
	return f;
}

template <int D>
static int numberUnknowns(const Fields &F, Projection &P, hipStream_t s, std::string &err)
{
	hipLaunchKernelGGL(kUnknownFlags<D>, gridDims(F.cell, D), gridBlock(), 0, s, F.cell, faceArrays(F, F.frac), (const double *)F.phi, P.flags);
	size_t bytes = P.scan_bytes;
	if (hipcub::DeviceScan::ExclusiveSum(P.scan_tmp, bytes, P.flags, P.grid2mat, int(F.N), s) != hipSuccess) { err = "prefix sum failed"; return -1; }
	int last[2];
	if (hipMemcpyAsync(&last[0], P.grid2mat + F.N - 1, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
		hipMemcpyAsync(&last[1], P.flags + F.N - 1, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
		hipStreamSynchronize(s) != hipSuccess) { err = "reading the unknown count failed"; return -1; }
	const int n = last[0] + last[1];
	if (!P.reserveRows(D, size_t(n), err)) return -1;
	hipLaunchKernelGGL(kCompactUnknowns, dim3(linBlocks(F.N)), dim3(kBlock), 0, s, (const int *)P.flags, P.grid2mat, P.mat2grid, F.N);
	return n;
}

template <int D>
static void assemble(const Fields &F, Projection &P, int n, hipStream_t s)
{
	if (n == 0) return;
	hipLaunchKernelGGL(kAssemble<D>, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, F.cell, faceArrays(F, F.frac),
		faceArrays(F, F.u), faceArrays(F, F.bvel), (const double *)F.phi, (const int *)P.grid2mat, (const int *)P.mat2grid, n, P.rows);
}

template <int D>
static int pcg64(Projection &P, int n, double tol, int maxIt, int check, hipStream_t s, garm_pcg_stats *st, std::string &err)
{
	const int nb = (n + kBlock - 1) / kBlock;
	const dim3 g1(nb > 0 ? nb : 1), b1(kBlock);
	hipLaunchKernelGGL(kPcgStart, g1, b1, 0, s, n, (const double *)P.rows.rhs, (const double *)P.rows.invdiag, P.x, P.r, P.p, P.partials);
	hipLaunchKernelGGL(kPcgStartFinal, dim3(1), b1, 0, s, (const double *)P.partials, nb, tol, maxIt, P.sc, P.state);
	int launched = 0;
	for (;;) {
		if (hipMemcpyAsync(P.host_state, P.state, 4 * sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
			hipStreamSynchronize(s) != hipSuccess) { err = "PCG state readback failed"; return GARM_EHIP; }
		if (P.host_state[kStDone] || launched >= maxIt) break;
		const int k = std::min(check, maxIt - launched);
		for (int q = 0; q < k; q++) {
			hipLaunchKernelGGL(kSpmvDot<D>, g1, b1, 0, s, P.rows, n, (const double *)P.p, P.y, P.partials, (const int *)P.state);
			hipLaunchKernelGGL(kAlpha, dim3(1), b1, 0, s, (const double *)P.partials, nb, P.sc, (const int *)P.state);
			hipLaunchKernelGGL(kUpdateXR, g1, b1, 0, s, n, P.x, P.r, (const double *)P.p, (const double *)P.y, (const double *)P.rows.invdiag,
				P.partials, (const double *)P.sc, (const int *)P.state);
			hipLaunchKernelGGL(kBeta, dim3(1), b1, 0, s, (const double *)P.partials, nb, P.sc, P.state);
			hipLaunchKernelGGL(kUpdateP, g1, b1, 0, s, n, P.p, (const double *)P.r, (const double *)P.rows.invdiag, (const double *)P.sc,
				(const int *)P.state);
		}
		launched += k;
	}
	double hs[8];
	if (hipMemcpyAsync(hs, P.sc, 8 * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
		hipMemcpyAsync(P.host_state, P.state, 4 * sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
		hipStreamSynchronize(s) != hipSuccess) { err = "PCG result readback failed"; return GARM_EHIP; }
	st->iterations = P.host_state[kStIter];
	st->relative_residual = hs[kScBnorm2] > 0 ? std::sqrt(hs[kScRn2] / hs[kScBnorm2]) : 0.0;
	if (st->relative_residual > hs[kScTol]) {
		err = "PCG did not reach the requested tolerance";
		return GARM_ENOCONV;
	}
	return GARM_OK;
}

template <int D>
static int projectD(Fields &F, Projection &P, const garm_pcg_opts &o, garm_pcg_stats *st, hipStream_t s, std::string &err)
{
	if (!P.reserveGrid(F.N, err)) return GARM_ENOMEM;
	const int n = numberUnknowns<D>(F, P, s, err);
	if (n < 0) return GARM_EHIP;
	P.n = n;
	st->unknowns = n;
	st->iterations = 0;
	st->outer_iterations = 0;
	st->fell_back = 0;
	st->relative_residual = 0;
	int rc = GARM_OK;
	if (n > 0) {
		assemble<D>(F, P, n, s);
		rc = pcg64<D>(P, n, o.tolerance, o.max_iterations, std::max(1, o.check_interval), s, st, err);
	}
	// the velocity is updated with whatever the solver reached, as the
	// reference does before deciding (it exits instead)
	for (int a = 0; a < D && n > 0; a++)
		hipLaunchKernelGGL(kApplyPressureGradient<D>, gridDims(F.face[a], D), gridBlock(), 0, s, F.face[a], F.cell, a, F.u[a],
			(const double *)F.phi, (const int *)P.grid2mat, (const double *)P.x);
	return rc;
}

int runProjection(Fields &F, Projection &P, const garm_pcg_opts &o, garm_pcg_stats *st, hipStream_t s, std::string &err)
{
	return F.D == 2 ? projectD<2>(F, P, o, st, s, err) : projectD<3>(F, P, o, st, s, err);
}

// Launch one PCG kernel `reps` times for timing with alpha forced to 0 so x
// and r keep their values; the solver scalars are restored afterwards.
int timePcgKernel(int D, Projection &P, int which, int reps, hipStream_t s, hipEvent_t e0, hipEvent_t e1, float *ms, double *bytes)
{
	const int n = P.n;
	if (n <= 0 || !P.sc) return GARM_EINVAL;
	double sc[32];
	int st[8];
	if (hipMemcpyAsync(sc, P.sc, sizeof(sc), hipMemcpyDeviceToHost, s) != hipSuccess ||
		hipMemcpyAsync(st, P.state, sizeof(st), hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
		return GARM_EHIP;
	double sc0[32];
	std::copy(sc, sc + 32, sc0);
	sc0[kScAlpha] = 0;
	int st0[8];
	std::copy(st, st + 8, st0);
	st0[kStDone] = 0;
	if (hipMemcpyAsync(P.sc, sc0, sizeof(sc0), hipMemcpyHostToDevice, s) != hipSuccess ||
		hipMemcpyAsync(P.state, st0, sizeof(st0), hipMemcpyHostToDevice, s) != hipSuccess) return GARM_EHIP;
	const int nb = (n + kBlock - 1) / kBlock;
	(void)hipEventRecord(e0, s);
	for (int q = 0; q < reps; q++) {
		if (which == 0) {
			if (D == 2) hipLaunchKernelGGL(kSpmvDot<2>, dim3(nb), dim3(kBlock), 0, s, P.rows, n, (const double *)P.p, P.y, P.partials, (const int *)P.state);
			else hipLaunchKernelGGL(kSpmvDot<3>, dim3(nb), dim3(kBlock), 0, s, P.rows, n, (const double *)P.p, P.y, P.partials, (const int *)P.state);
		} else {
			hipLaunchKernelGGL(kUpdateXR, dim3(nb), dim3(kBlock), 0, s, n, P.x, P.r, (const double *)P.p, (const double *)P.y,
				(const double *)P.rows.invdiag, P.partials, (const double *)P.sc, (const int *)P.state);
		}
	}
	(void)hipEventRecord(e1, s);
	if (hipMemcpyAsync(P.sc, sc, sizeof(sc), hipMemcpyHostToDevice, s) != hipSuccess ||
		hipMemcpyAsync(P.state, st, sizeof(st), hipMemcpyHostToDevice, s) != hipSuccess ||
		hipStreamSynchronize(s) != hipSuccess) return GARM_EHIP;
	(void)hipEventElapsedTime(ms, e0, e1);
	*bytes = which == 0 ? double(n) * (8 + 2 * D * (4 + 8) + 8 + 8) : double(n) * (5 * 8 + 2 * 8);
	return GARM_OK;
}

void scatterPressure(const Fields &F, const Projection &P, double *out, hipStream_t s)
{
	if (!P.grid2mat || P.n == 0) { launchFill(out, 0.0, F.N, s); return; }
	hipLaunchKernelGGL(kScatterPressure, dim3(linBlocks(F.N)), dim3(kBlock), 0, s, (const int *)P.grid2mat, (const double *)P.x, out, F.N);
}

} // namespace garm
