// Hybrid-Newton restart of the reference level set, the default build's
// LevelSetLiquid::restart with enable-fmm=n (LevelSetLiquid.cpp:1290-1303,
// HybridNewton.cpp:18-49, 64-79, 103-123, HybridNewton.h:119-136, 147-467
// with ENABLE_MODIFIED), in FP64 with the reference's operations (this file
// is compiled without multiply-add contraction).
//
// Cells next to a sign change of phi_ori, and one layer more, take their
// distance from the closest point on the Hermite interpolant of
// (phi_ori, grad phi) found by the modified Newton iteration, started from
// x - phi grad phi of the current level set; every other cell takes 20
// characteristic steps of 0.9 dx along the normalised gradient (RK3).  One
// thread per cell; each characteristic step rebuilds the Hermite records of
// the iterate and writes the next one to the other buffer.
#include "garm_internal.hpp"
#include "garm_views.hpp"

namespace garm {

// Hermite value and gradient of the scalar records `recs` at X, corners and
// orders in the reference's order (HermiteIntrpl.h:32-58)
template <int D>
__device__ __forceinline__ void hermiteBoth(const GridDev &g, const double *recs, const double *X, double *val, double *grad)
{
	constexpr int S = 1 << D;
	int l[3] = {0, 0, 0};
	double f[3] = {0, 0, 0};
	lowerFrac<D>(g, X, l, f);
	double B[3][2][2], Bg[3][2][2];
#pragma unroll
	for (int a = 0; a < D; a++) {
		const double x = f[a];
		B[a][0][0] = (x + x + 1) * (x - 1) * (x - 1); Bg[a][0][0] = 6 * x * (x - 1);
		B[a][1][0] = x * x * (3 - x - x);             Bg[a][1][0] = 6 * x * (1 - x);
		B[a][0][1] = x * (x - 1) * (x - 1);           Bg[a][0][1] = (3 * x - 1) * (x - 1);
		B[a][1][1] = x * x * (x - 1);                 Bg[a][1][1] = x * (3 * x - 2);
	}
	unsigned ox[2], oy[2], oz[2];
	cornerOffsets<D, 2>(g, l, ox, oy, oz);
	double v0 = 0, gr[3] = {0, 0, 0};
#pragma unroll
	for (int cn = 0; cn < S; cn++) {
		const int vx = cn & 1, vy = cn >> 1 & 1, vz = D == 3 ? cn >> 2 : 0;
		const double *r = recs + size_t(ox[vx] + oy[vy] + oz[vz]) * S;
#pragma unroll
		for (int ord = 0; ord < S; ord++) {
			const int ex = ord & 1, ey = ord >> 1 & 1, ez = D == 3 ? ord >> 2 : 0;
			double wt = B[0][vx][ex] * B[1][vy][ey];
			double wg[3];
			if constexpr (D == 2) {
				wg[0] = Bg[0][vx][ex] * B[1][vy][ey];
				wg[1] = B[0][vx][ex] * Bg[1][vy][ey];
			} else {
				wt = wt * B[2][vz][ez];
				wg[0] = Bg[0][vx][ex] * B[1][vy][ey] * B[2][vz][ez];
				wg[1] = B[0][vx][ex] * Bg[1][vy][ey] * B[2][vz][ez];
				wg[2] = B[0][vx][ex] * B[1][vy][ey] * Bg[2][vz][ez];
			}
			const double rv = r[ord];
			v0 += rv * wt;
#pragma unroll
			for (int a = 0; a < D; a++) gr[a] += rv * wg[a];
		}
	}
	*val = v0;
#pragma unroll
	for (int a = 0; a < D; a++) grad[a] = gr[a] * g.inv;
}

template <int D>
__device__ __forceinline__ void cellPosition(const GridDev &g, const Coord &c, double *p)
{
	p[0] = g.o[0] + c.i * g.dx;
	p[1] = g.o[1] + c.j * g.dx;
	if constexpr (D == 3) p[2] = g.o[2] + c.k * g.dx;
}

// points = x - phi(x) grad(x) with the cubic view of phi
// (ScalarFieldWrapper<CubicIntrpl>, ScalarField.h:23-47: CubicIntrpl value,
// linear weights of getFirst<2> at the corners)
template <int D>
__global__ __launch_bounds__(kBlock) void kNewtonPoints(GridDev g, const double *phi, double *p0, double *p1, double *p2)
{
	const Coord c = threadCoord();
	if (!inGrid<D>(g, c)) return;
	const size_t id = gidx<D>(g, c.i, c.j, c.k);
	double p[3] = {0, 0, 0};
	cellPosition<D>(g, c, p);
	const double v = cubicView<D>(g, phi, p);
	double gr[3] = {0, 0, 0};
	cubicViewGradient<D>(g, phi, p, gr);
	double *out[3] = {p0, p1, p2};
#pragma unroll
	for (int a = 0; a < D; a++) out[a][id] = p[a] - v * gr[a];
}

template <int D>
__device__ __forceinline__ bool validCell(const GridDev &g, int i, int j, int k)
{
	return i >= 0 && i < g.n[0] && j >= 0 && j < g.n[1] && (D == 2 || (k >= 0 && k < g.n[2]));
}

// HybridNewton::prepareInterface (HybridNewton.cpp:64-79)
template <int D>
__global__ __launch_bounds__(kBlock) void kNewtonMark(GridDev g, const double *f, uint8_t *mark)
{
	const Coord c = threadCoord();
	if (!inGrid<D>(g, c)) return;
	const size_t id = gidx<D>(g, c.i, c.j, c.k);
	const double v = f[id];
	uint8_t m = 0;
#pragma unroll
	for (int o = 0; o < 2 * D; o++) {
		const int s = (o & 1) ? 1 : -1, ax = o >> 1;
		const int i = c.i + (ax == 0 ? s : 0), j = c.j + (ax == 1 ? s : 0), k = c.k + (ax == 2 ? s : 0);
		if (validCell<D>(g, i, j, k) && v * f[gidx<D>(g, i, j, k)] <= 0) { m = 1; break; }
	}
	mark[id] = m;
}

// HybridNewton::extrapolate(1) (HybridNewton.cpp:103-123)
template <int D>
__global__ __launch_bounds__(kBlock) void kNewtonExtend(GridDev g, const uint8_t *mark0, uint8_t *mark)
{
	const Coord c = threadCoord();
	if (!inGrid<D>(g, c)) return;
	const size_t id = gidx<D>(g, c.i, c.j, c.k);
	uint8_t m = mark0[id];
	for (int o = 0; o < 2 * D && !m; o++) {
		const int s = (o & 1) ? 1 : -1, ax = o >> 1;
		const int i = c.i + (ax == 0 ? s : 0), j = c.j + (ax == 1 ? s : 0), k = c.k + (ax == 2 ? s : 0);
		if (validCell<D>(g, i, j, k) && mark0[gidx<D>(g, i, j, k)]) m = 1;
	}
	mark[id] = m;
}

template <int D>
struct NewtonArgs {
	GridDev g;
	const uint8_t *mark;
	const double *recs;       // Hermite records of (f, psi)
	const double *f, *psi[3];
	double *nf, *npsi[3];
	const double *pts[3];
};

// initializeInterface with the guesses (HybridNewton.h:119-136) and
// getClosestPosition (:147-467, ENABLE_MODIFIED, no FMM, epsl = 0); cells
// that are not interface cells copy their values
template <int D>
__global__ __launch_bounds__(kBlock) void kNewtonSolve(NewtonArgs<D> A)
{
	const GridDev &g = A.g;
	const Coord c = threadCoord();
	if (!inGrid<D>(g, c)) return;
	const size_t id = gidx<D>(g, c.i, c.j, c.k);
	if (!A.mark[id]) {
		A.nf[id] = A.f[id];
#pragma unroll
		for (int a = 0; a < D; a++) A.npsi[a][id] = A.psi[a][id];
		return;
	}
	double start[3] = {0, 0, 0}, pos[3] = {0, 0, 0};
	cellPosition<D>(g, c, start);
#pragma unroll
	for (int a = 0; a < D; a++) pos[a] = A.pts[a][id];
	const double dx = g.dx;
	const double tolerance = 1e-3 * dx * dx * (D == 2 ? 1 : dx);
	const double radius = .9 * dx;
	for (int iter = 0; iter < 20; iter++) {
		double phi, psi[3] = {0, 0, 0};
		hermiteBoth<D>(g, A.recs, pos, &phi, psi);
		phi = phi - 0.0;
		double sqn = 0, dot = 0, d[3] = {0, 0, 0}, delta[3] = {0, 0, 0};
#pragma unroll
		for (int a = 0; a < D; a++) sqn += psi[a] * psi[a];
#pragma unroll
		for (int a = 0; a < D; a++) d[a] = start[a] - pos[a];
#pragma unroll
		for (int a = 0; a < D; a++) dot += d[a] * psi[a];
#pragma unroll
		for (int a = 0; a < D; a++) {
			const double d1 = -phi * psi[a] / sqn;
			const double d2 = d[a] - psi[a] * dot / sqn;
			delta[a] = d1 + d2;
		}
		double n2 = 0;
#pragma unroll
		for (int a = 0; a < D; a++) n2 += delta[a] * delta[a];
		double norm = sqrt(n2);
		if (norm > radius) {
			const double sc = radius / norm;
#pragma unroll
			for (int a = 0; a < D; a++) delta[a] *= sc;
			norm = radius;
		}
#pragma unroll
		for (int a = 0; a < D; a++) pos[a] += delta[a];
		if (norm < tolerance) break;
	}
	double n2 = 0;
#pragma unroll
	for (int a = 0; a < D; a++) n2 += (start[a] - pos[a]) * (start[a] - pos[a]);
	A.nf[id] = sqrt(n2) * (A.f[id] < 0 ? -1 : 1);
	double v, gr[3] = {0, 0, 0}, nr[3] = {0, 0, 0};
	hermiteBoth<D>(g, A.recs, pos, &v, gr);
	normalizeD<D>(gr, nr);
#pragma unroll
	for (int a = 0; a < D; a++) A.npsi[a][id] = nr[a];
}

// HybridNewton::advect, one characteristic step (HybridNewton.cpp:30-49):
// RK3 trace (Advection.h:43-49) along Velocity (HybridNewton.h:23-38)
template <int D>
__global__ __launch_bounds__(kBlock) void kNewtonMarch(NewtonArgs<D> A, double dt)
{
	const GridDev &g = A.g;
	const Coord c = threadCoord();
	if (!inGrid<D>(g, c)) return;
	const size_t id = gidx<D>(g, c.i, c.j, c.k);
	if (A.mark[id]) {
		A.nf[id] = A.f[id];
#pragma unroll
		for (int a = 0; a < D; a++) A.npsi[a][id] = A.psi[a][id];
		return;
	}
	auto vel = [&](const double *x, double *v) {
		double val, gr[3] = {0, 0, 0};
		hermiteBoth<D>(g, A.recs, x, &val, gr);
		bool any = false;
#pragma unroll
		for (int a = 0; a < D; a++) any = any || gr[a] != 0;
		if (any) normalizeD<D>(gr, v);
		else {
#pragma unroll
			for (int a = 0; a < D; a++) v[a] = a == 0 ? 1 : 0;
		}
	};
	const int sgn = A.f[id] < 0 ? -1 : 1;
	double p[3] = {0, 0, 0}, v0[3] = {0, 0, 0}, v1[3] = {0, 0, 0}, v2[3] = {0, 0, 0}, q[3] = {0, 0, 0}, t[3] = {0, 0, 0};
	cellPosition<D>(g, c, p);
	const double h = -sgn * dt;
	vel(p, v0);
#pragma unroll
	for (int a = 0; a < D; a++) q[a] = p[a] + v0[a] * h;
	vel(q, v1);
#pragma unroll
	for (int a = 0; a < D; a++) q[a] = p[a] + (v0[a] + v1[a]) * h / 4;
	vel(q, v2);
#pragma unroll
	for (int a = 0; a < D; a++) t[a] = p[a] + (v0[a] + v1[a] + 4 * v2[a]) * h / 6;
	double val, gr[3] = {0, 0, 0}, nr[3] = {0, 0, 0};
	hermiteBoth<D>(g, A.recs, t, &val, gr);
	A.nf[id] = val + sgn * dt;
	normalizeD<D>(gr, nr);
#pragma unroll
	for (int a = 0; a < D; a++) A.npsi[a][id] = nr[a];
}

template <int D>
static void restartNewtonD(const Fields &F, hipStream_t s)
{
	constexpr int kSteps = 20;            // reinizer.perform(..., 20, 0.9) (LevelSetLiquid.cpp:1299)
	constexpr double kCfl = 0.9;
	const GridDev &g = F.cell;
	const dim3 gd = gridDims(g, D), bl = gridBlock();
	// iterate A: phi_ref / phi_ref_grad, iterate B: scratch[0] / J_new[0..D-1]
	double *fa = F.phi_ref, *fb = F.scratch[0];
	double *pa[3] = {F.phi_ref_grad[0], F.phi_ref_grad[1], D == 3 ? F.phi_ref_grad[2] : nullptr};
	double *pb[3] = {F.J_new[0], F.J_new[1], D == 3 ? F.J_new[2] : nullptr};
	launchCopy(fa, F.phi_ori, F.N, s);
	for (int a = 0; a < D; a++) launchCopy(pa[a], F.phi_grad[a], F.N, s);
	hipLaunchKernelGGL(kNewtonPoints<D>, gd, bl, 0, s, g, (const double *)F.phi, F.xi_new[0], F.xi_new[1], D == 3 ? F.xi_new[2] : nullptr);
	hipLaunchKernelGGL(kNewtonMark<D>, gd, bl, 0, s, g, (const double *)fa, F.layer);
	hipLaunchKernelGGL(kNewtonExtend<D>, gd, bl, 0, s, g, (const uint8_t *)F.layer, F.layer2);
	auto args = [&](double *f, double *const *psi, double *nf, double *const *npsi) {
		NewtonArgs<D> A{};
		A.g = g;
		A.mark = F.layer2;
		A.recs = F.ref_rec64;
		A.f = f;
		A.nf = nf;
		for (int a = 0; a < D; a++) { A.psi[a] = psi[a]; A.npsi[a] = npsi[a]; A.pts[a] = F.xi_new[a]; }
		return A;
	};
	launchLevelSetRecords(F, fa, pa, F.ref_rec64, F.ref_rec32, s);
	hipLaunchKernelGGL(kNewtonSolve<D>, gd, bl, 0, s, args(fa, pa, fb, pb));
	for (int it = 0; it < kSteps; it++) {
		// the current iterate is B after an even number of steps, A after an odd one
		const bool fromB = it % 2 == 0;
		double *f = fromB ? fb : fa, *nf = fromB ? fa : fb;
		double *const *psi = fromB ? pb : pa, *const *npsi = fromB ? pa : pb;
		launchLevelSetRecords(F, f, psi, F.ref_rec64, F.ref_rec32, s);
		hipLaunchKernelGGL(kNewtonMarch<D>, gd, bl, 0, s, args(f, psi, nf, npsi), g.dx * kCfl);
	}
	// 20 steps from B end in B: the result becomes the reference level set
	launchCopy(F.phi_ref, fb, F.N, s);
	for (int a = 0; a < D; a++) launchCopy(F.phi_ref_grad[a], pb[a], F.N, s);
	launchRefLevelSetRecords(F, s);
	launchResetRefMap(F, s);
}

void launchRestartNewton(const Fields &F, hipStream_t s)
{
	if (F.D == 2) restartNewtonD<2>(F, s);
	else restartNewtonD<3>(F, s);
}

} // namespace garm
