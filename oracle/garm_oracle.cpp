// ORACLE — test infrastructure, not product code.
//
// A clean-room C++/OpenMP FP64 restatement of the GARM-LS per-timestep grid
// path, used ONLY by tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline as the checker.  Nothing in the product links or calls it.
//
// Every function cites the reference file:line whose arithmetic it follows.
// Layout follows the reference: x-fastest cell indexing (Grid.h:40-44),
// array-of-structures vectors per cell, column-major D x D matrices
// (Types.h:51-53), clamped sampling (GridBasedData.h:47).
//
// Deviations from a default reference build, chosen for the GPU path and
// mirrored here so both sides compute the same thing:
//  - reinitialisation is WENO5 PDE (the --enable-fmm=n build), optionally
//    every k steps (k = 1 is the reference);
//  - the reference-map restart is the branch without Hybrid Newton
//    (LevelSetLiquid.cpp:1331-1352) unless Hybrid Newton is switched on;
//  - the pressure solve is a restated conjugate gradient with IC(0) (natural
//    ordering), Jacobi or no preconditioner in place of Eigen's
//    IncompleteCholesky (LinearSystem.h:10-39); the tolerance is a parameter.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace ora {

using u8 = std::uint8_t;

// ---------------------------------------------------------------- grid ----
// Grid<Dim> (Grid.h:10-64).
template <int D>
struct Grid {
	int n[3] = {1, 1, 1};
	double o[3] = {0, 0, 0};
	double dx = 1, inv = 1;

	size_t count() const { return size_t(n[0]) * n[1] * (D == 3 ? n[2] : 1); }
	// Grid::index (Grid.h:40-44)
	size_t index(const int *c) const
	{
		if constexpr (D == 2) return c[0] + size_t(n[0]) * c[1];
		else return c[0] + n[0] * (c[1] + size_t(n[1]) * c[2]);
	}
	// Grid::isValid (Grid.h:34-35)
	bool valid(const int *c) const
	{
		for (int a = 0; a < D; a++) if (c[a] < 0 || c[a] >= n[a]) return false;
		return true;
	}
	bool inside(const int *c, int off) const
	{
		for (int a = 0; a < D; a++) if (c[a] < off || c[a] >= n[a] - off) return false;
		return true;
	}
	// GridBasedData::at = index(clamp(coord)) (GridBasedData.h:47, Grid.h:38)
	size_t at(const int *c) const
	{
		int q[3];
		for (int a = 0; a < D; a++) q[a] = std::min(std::max(c[a], 0), n[a] - 1);
		return index(q);
	}
	void coord(size_t idx, int *c) const
	{
		c[0] = int(idx % n[0]);
		if constexpr (D == 2) c[1] = int(idx / n[0]);
		else { c[1] = int(idx / n[0] % n[1]); c[2] = int(idx / n[0] / n[1]); }
	}
	// Grid::position (Grid.h:37)
	void position(const int *c, double *p) const { for (int a = 0; a < D; a++) p[a] = o[a] + c[a] * dx; }
	// Grid::getLinearLower (Grid.h:52)
	void lower(const double *p, int *l) const { for (int a = 0; a < D; a++) l[a] = int(std::floor((p[a] - o[a]) * inv)); }
	// Grid::getLowerFrac (Grid.h:55)
	void frac(const double *p, const int *l, double *f) const { for (int a = 0; a < D; a++) f[a] = (p[a] - o[a] - l[a] * dx) * inv; }
};

// parallelForEach over a grid, OpenMP over the outermost axis (Grid.h:82-103).
template <int D, typename F>
void parFor(const Grid<D> &g, F f)
{
	const int outer = g.n[D - 1];
#pragma omp parallel for schedule(static)
	for (int t = 0; t < outer; t++) {
		int c[3] = {0, 0, 0};
		c[D - 1] = t;
		if constexpr (D == 2) {
			for (int i = 0; i < g.n[0]; i++) { c[0] = i; f(c); }
		} else {
			for (int j = 0; j < g.n[1]; j++)
				for (int i = 0; i < g.n[0]; i++) { c[0] = i; c[1] = j; f(c); }
		}
	}
}

// serial forEach (Grid.h:66-80)
template <int D, typename F>
void serFor(const Grid<D> &g, F f)
{
	int c[3] = {0, 0, 0};
	if constexpr (D == 2) {
		for (c[1] = 0; c[1] < g.n[1]; c[1]++)
			for (c[0] = 0; c[0] < g.n[0]; c[0]++) f(c);
	} else {
		for (c[2] = 0; c[2] < g.n[2]; c[2]++)
			for (c[1] = 0; c[1] < g.n[1]; c[1]++)
				for (c[0] = 0; c[0] < g.n[0]; c[0]++) f(c);
	}
}

static inline double sq(double x) { return x * x; }

// ------------------------------------------------------- interpolation ----
// LinearIntrpl::wtPoints + interpolate (LagrangeIntrpl.h:40-72).  C values
// per grid point (AoS); result accumulated corner by corner from zero.
template <int D, int C>
void lerp(const Grid<D> &g, const double *data, const double *pos, double *out)
{
	int l[3]; double f[3];
	g.lower(pos, l);
	g.frac(pos, l, f);
	double w[2][3];
	for (int a = 0; a < D; a++) { w[0][a] = 1 - f[a]; w[1][a] = f[a]; }
	for (int r = 0; r < C; r++) out[r] = 0;
	for (int k = 0; k < (D == 3 ? 2 : 1); k++)
		for (int j = 0; j < 2; j++)
			for (int i = 0; i < 2; i++) {
				int c[3] = {l[0] + i, l[1] + j, l[2] + k};
				double wt = D == 2 ? w[i][0] * w[j][1] : w[i][0] * w[j][1] * w[k][2];
				const double *v = data + g.at(c) * C;
				for (int r = 0; r < C; r++) out[r] += v[r] * wt;
			}
}

// CubicIntrpl::wtPoints + interpolate (LagrangeIntrpl.h:128-171).
template <int D>
double cubic(const Grid<D> &g, const double *data, const double *pos)
{
	int l[3]; double f[3];
	g.lower(pos, l);
	for (int a = 0; a < D; a++) l[a] -= 1;   // getCubicLower (Grid.h:54)
	g.frac(pos, l, f);
	double w[4][3];
	for (int a = 0; a < D; a++) {
		const double t = f[a];
		w[0][a] = (t - 1) * (t - 2) * (t - 3) / -6;
		w[1][a] = t * (t - 2) * (t - 3) / 2;
		w[2][a] = t * (t - 1) * (t - 3) / -2;
		w[3][a] = t * (t - 1) * (t - 2) / 6;
	}
	double val = 0;
	for (int k = 0; k < (D == 3 ? 4 : 1); k++)
		for (int j = 0; j < 4; j++)
			for (int i = 0; i < 4; i++) {
				int c[3] = {l[0] + i, l[1] + j, l[2] + k};
				double wt = D == 2 ? w[i][0] * w[j][1] : w[i][0] * w[j][1] * w[k][2];
				val += data[g.at(c)] * wt;
			}
	return val;
}

// HermiteIntrpl<Dim, Type> (HermiteIntrpl.h:7-125) for Type with C
// components (C = 1: double, C = D: VectorDd).  Derivative data is C x D
// column-major per point; column `axis` is getC(drv, axis) (Derivatives.h:37-43).
template <int D, int C>
struct Hermite {
	static constexpr int M = (1 << D) - D - 1;   // number of mixed-derivative arrays
	std::vector<double> mixed[M];

	// constructor (HermiteIntrpl.h:17-30)
	void build(const Grid<D> &g, const double *drv)
	{
		for (int m = 0; m < M; m++) mixed[m].assign(g.count() * C, 0.0);
		const double inv = g.inv;
		parFor<D>(g, [&](const int *c) {
			const size_t id = g.index(c);
			auto col = [&](int dx_, int dy_, int axis) {
				int q[3] = {c[0] + dx_, c[1] + dy_, D == 3 ? c[2] : 0};
				return drv + g.at(q) * C * D + axis * C;
			};
			if constexpr (D == 2) {
				const double *a = col(1, 0, 1), *b = col(-1, 0, 1);
				for (int r = 0; r < C; r++) mixed[0][id * C + r] = (a[r] - b[r]) / 2 * inv;
			} else {
				{ const double *a = col(1, 0, 1), *b = col(-1, 0, 1);
				  for (int r = 0; r < C; r++) mixed[0][id * C + r] = (a[r] - b[r]) / 2 * inv; }
				{ const double *a = col(1, 0, 2), *b = col(-1, 0, 2);
				  for (int r = 0; r < C; r++) mixed[1][id * C + r] = (a[r] - b[r]) / 2 * inv; }
				{ const double *a = col(0, 1, 2), *b = col(0, -1, 2);
				  for (int r = 0; r < C; r++) mixed[2][id * C + r] = (a[r] - b[r]) / 2 * inv; }
				{ const double *a = col(1, 1, 2), *b = col(-1, 1, 2), *e = col(1, -1, 2), *h = col(-1, -1, 2);
				  for (int r = 0; r < C; r++) mixed[3][id * C + r] = (a[r] - b[r] - e[r] + h[r]) / 4 * inv * inv; }
			}
		});
	}

	// weight (HermiteIntrpl.h:89-100)
	static double weight(const double *f, int vtx, int ord)
	{
		double wt = 1;
		for (int a = 0; a < D; a++) {
			const double x = f[a];
			if (ord >> a & 1) wt *= (vtx >> a & 1) ? x * x * (x - 1) : x * (x - 1) * (x - 1);
			else wt *= (vtx >> a & 1) ? x * x * (3 - x - x) : (x + x + 1) * (x - 1) * (x - 1);
		}
		return wt;
	}
	// weightGradient (HermiteIntrpl.h:102-125)
	static void weightGrad(const double *f, int vtx, int ord, double *out)
	{
		double wt[3], wg[3];
		for (int a = 0; a < D; a++) {
			const double x = f[a];
			if (ord >> a & 1) {
				if (vtx >> a & 1) wt[a] = x * x * (x - 1), wg[a] = x * (3 * x - 2);
				else wt[a] = x * (x - 1) * (x - 1), wg[a] = (3 * x - 1) * (x - 1);
			} else {
				if (vtx >> a & 1) wt[a] = x * x * (3 - x - x), wg[a] = 6 * x * (1 - x);
				else wt[a] = (x + x + 1) * (x - 1) * (x - 1), wg[a] = 6 * x * (x - 1);
			}
		}
		if constexpr (D == 2) { out[0] = wg[0] * wt[1]; out[1] = wt[0] * wg[1]; }
		else { out[0] = wg[0] * wt[1] * wt[2]; out[1] = wt[0] * wg[1] * wt[2]; out[2] = wt[0] * wt[1] * wg[2]; }
	}

	// getValues (HermiteIntrpl.h:62-87)
	void values(const Grid<D> &g, const double *val, const double *drv, const int *l, double (*v)[C]) const
	{
		const double dx = g.dx;
		for (int i = 0; i < (1 << D); i++) {
			int c[3] = {l[0] + (i & 1), l[1] + (i >> 1 & 1), D == 3 ? l[2] + (i >> 2) : 0};
			const size_t id = g.at(c);
			const double *dv = drv + id * C * D;
			for (int r = 0; r < C; r++) {
				v[i << D | 0][r] = val[id * C + r];
				v[i << D | 1][r] = dv[0 * C + r] * dx;
				v[i << D | 2][r] = dv[1 * C + r] * dx;
				v[i << D | 3][r] = mixed[0][id * C + r] * dx * dx;
				if constexpr (D == 3) {
					v[i << D | 4][r] = dv[2 * C + r] * dx;
					v[i << D | 5][r] = mixed[1][id * C + r] * dx * dx;
					v[i << D | 6][r] = mixed[2][id * C + r] * dx * dx;
					v[i << D | 7][r] = mixed[3][id * C + r] * dx * dx * dx;
				}
			}
		}
	}

	// interpolate (HermiteIntrpl.h:32-41)
	void interpolate(const Grid<D> &g, const double *val, const double *drv, const double *pos, double *out) const
	{
		int l[3]; double f[3];
		g.lower(pos, l);
		g.frac(pos, l, f);
		double v[1 << (2 * D)][C];
		values(g, val, drv, l, v);
		for (int r = 0; r < C; r++) out[r] = 0;
		for (int i = 0; i < (1 << (2 * D)); i++) {
			const double wt = weight(f, i >> D, i & ((1 << D) - 1));
			for (int r = 0; r < C; r++) out[r] += v[i][r] * wt;
		}
	}

	// interpolate + interpolateGradient (HermiteIntrpl.h:32-58); grad is
	// C x D column-major: grad(r, c) = d out_r / d x_c.
	void interpolateBoth(const Grid<D> &g, const double *val, const double *drv, const double *pos, double *out, double *grad) const
	{
		int l[3]; double f[3];
		g.lower(pos, l);
		g.frac(pos, l, f);
		double v[1 << (2 * D)][C];
		values(g, val, drv, l, v);
		for (int r = 0; r < C; r++) out[r] = 0;
		for (int e = 0; e < C * D; e++) grad[e] = 0;
		for (int i = 0; i < (1 << (2 * D)); i++) {
			const int vtx = i >> D, ord = i & ((1 << D) - 1);
			const double wt = weight(f, vtx, ord);
			double wg[3];
			weightGrad(f, vtx, ord, wg);
			for (int r = 0; r < C; r++) {
				out[r] += v[i][r] * wt;
				for (int c = 0; c < D; c++) grad[c * C + r] += v[i][r] * wg[c];
			}
		}
		for (int e = 0; e < C * D; e++) grad[e] *= g.inv;
	}
};

// ----------------------------------------------------------- surfaces ----
// ImplicitBox::signedDistance (ImplicitSurface.h:51-56)
template <int D>
double boxSdf(const double *center, const double *half, const double *p)
{
	double phi[3];
	bool allIn = true;
	for (int a = 0; a < D; a++) { phi[a] = std::abs(p[a] - center[a]) - half[a]; if (!(phi[a] <= 0)) allIn = false; }
	if (allIn) { double m = phi[0]; for (int a = 1; a < D; a++) m = std::max(m, phi[a]); return m; }
	double s = 0;
	for (int a = 0; a < D; a++) s += sq(std::max(phi[a], 0.0));
	return std::sqrt(s);
}

// ImplicitBox::closestNormal (ImplicitSurface.h:39-49), Eigen's first-max
// argmax and normalized() that leaves a zero vector unchanged.
template <int D>
void boxNormal(const double *center, const double *half, const double *p, double *nrm)
{
	double phi[3], n[3];
	bool allIn = true;
	for (int a = 0; a < D; a++) { phi[a] = std::abs(p[a] - center[a]) - half[a]; if (!(phi[a] <= 0)) allIn = false; }
	if (allIn) {
		int axis = 0;
		for (int a = 1; a < D; a++) if (phi[a] > phi[axis]) axis = a;
		for (int a = 0; a < D; a++) n[a] = a == axis ? 1.0 : 0.0;
	} else {
		for (int a = 0; a < D; a++) n[a] = std::max(phi[a], 0.0);
	}
	double s = 0;
	for (int a = 0; a < D; a++) {
		const double d = p[a] - center[a];
		const double sgn = d > 0 ? 1.0 : (d < 0 ? -1.0 : 0.0);
		n[a] *= sgn;
		s += n[a] * n[a];
	}
	if (s > 0) { const double r = std::sqrt(s); for (int a = 0; a < D; a++) n[a] /= r; }
	for (int a = 0; a < D; a++) nrm[a] = n[a];
}

// ------------------------------------------------------------- solver ----
template <int D>
struct Liquid {
	int bw = 2;
	double dx = 1, inv = 1;
	int res[3] = {1, 1, 1};
	double origin[3] = {0, 0, 0};

	Grid<D> cell, node, face[D];

	std::vector<double> u[D];
	std::vector<double> phi, phiGrad, xi, J, phiTmp, phiOri, phiRef, phiRefGrad;
	std::vector<u8> advRegion;
	Hermite<D, 1> refView;   // _refLevelSetView's mixed derivatives
	std::vector<uint8_t> newtonMark;   // HybridNewton::_mark of the last Newton restart (_reinitmark)
	bool newtonRestart = false;        // step() restarts with Hybrid Newton (the default build) instead of the plain branch

	std::vector<double> nodeDist, cellDist, frac[D], bvel[D], bnormal[D];

	std::vector<int> grid2mat, mat2grid;
	std::vector<double> pres;
	std::vector<double> lastX;         // Projector::_rdP of the last solve (compact)

	double elapsed = 0;                // _elapsedTime (LevelSetLiquid.h:83)
	// surface tension (LevelSetLiquid.h:86-91); step() applies it when on
	bool stOn = false, stLittleDrop = true;
	double stCoeff = 7.28e-2 / 1e3, stDeltaMin = 2, stDeltaMax = 8, stCritTime = 1, stTol = 1e-5;
	int stMaxIt = 2000;
	// the solvers' start and preconditioner: the reference's (IC(0) from the
	// velocity guess; the reprojection from the last pressure) unless changed
	int stPrecond = 2;
	bool stGuess = true, stWarm = true;
	int stFaces = 0, stIters = 0;      // last capillary solve
	double stErr = 0, stBand = 0;

	// StaggeredGrid (StaggeredGrid.h:27-36, 104-120) with an explicit origin.
	void setup(const int *r, double h, const double *org, int bw_)
	{
		bw = bw_; dx = h; inv = 1 / h;
		for (int a = 0; a < 3; a++) { res[a] = a < D ? r[a] : 1; origin[a] = a < D ? org[a] : 0; }
		auto mk = [&](Grid<D> &g, const int *n, const double *o) {
			for (int a = 0; a < 3; a++) { g.n[a] = n[a]; g.o[a] = o[a]; }
			g.dx = dx; g.inv = inv;
		};
		int n[3]; double o[3];
		for (int a = 0; a < 3; a++) { n[a] = res[a]; o[a] = origin[a] + (a < D ? dx / 2 : 0); }
		mk(cell, n, o);
		for (int a = 0; a < 3; a++) { n[a] = res[a] + (a < D ? 1 : 0); o[a] = origin[a]; }
		mk(node, n, o);
		for (int ax = 0; ax < D; ax++) {
			for (int a = 0; a < 3; a++) {
				n[a] = res[a] + (a == ax ? 1 : 0);
				o[a] = origin[a] + ((a < D && a != ax) ? dx / 2 : 0);
			}
			mk(face[ax], n, o);
		}
		const size_t N = cell.count();
		for (int a = 0; a < D; a++) {
			u[a].assign(face[a].count(), 0.0);
			frac[a].assign(face[a].count(), 0.0);
			bvel[a].assign(face[a].count(), 0.0);
			bnormal[a].assign(face[a].count(), 0.0);
		}
		phi.assign(N, std::numeric_limits<double>::infinity());   // LevelSetLiquid.cpp:50
		phiGrad.assign(N * D, 0.0);
		xi.assign(N * D, 0.0);
		J.assign(N * D * D, 0.0);
		phiTmp.assign(N, 0.0); phiOri.assign(N, 0.0); phiRef.assign(N, 0.0);
		phiRefGrad.assign(N * D, 0.0);
		advRegion.assign(N, 0);
		nodeDist.assign(node.count(), 0.0);
		cellDist.assign(N, 0.0);
		grid2mat.assign(N, -1);
		pres.assign(N, 0.0);
		buildBoundary();
		refView.build(cell, phiRefGrad.data());
	}

	// Boundary ctor + reset (Boundary.cpp:7-35) for the domain box, then
	// finish (Boundary.cpp:79-87) with getFaceFraction (Boundary.cpp:135-174).
	void buildBoundary()
	{
		resetBoundary();
		finishBoundary();
	}

	// The domain box with the solid half-space (x - pos).n < 0 added:
	// Boundary::reset, unions(ImplicitPlane) and finish (Boundary.cpp:19-56,
	// 79-87; ImplicitPlane, ImplicitSurface.h:111-126: normalised direction,
	// signed distance (x - pos).n, closest normal n).
	void solidPlane(const double *pos, const double *dir)
	{
		double n[3] = {0, 0, 0}, s = 0;
		for (int a = 0; a < D; a++) s += dir[a] * dir[a];
		s = std::sqrt(s);
		for (int a = 0; a < D; a++) n[a] = s > 0 ? dir[a] / s : dir[a];
		auto sd = [&](const double *p) {
			double d = 0;
			for (int a = 0; a < D; a++) d += (p[a] - pos[a]) * n[a];
			return d;
		};
		resetBoundary();
		parFor<D>(node, [&](const int *c) {
			double p[3]; node.position(c, p);
			const size_t id = node.index(c);
			nodeDist[id] = std::min(nodeDist[id], sd(p));
		});
		parFor<D>(cell, [&](const int *c) {
			double p[3]; cell.position(c, p);
			const size_t id = cell.index(c);
			cellDist[id] = std::min(cellDist[id], sd(p));
		});
		for (int ax = 0; ax < D; ax++) {
			parFor<D>(face[ax], [&](const int *c) {
				double p[3]; face[ax].position(c, p);
				const size_t id = face[ax].index(c);
				const double phi = sd(p);
				if (phi < frac[ax][id]) {   // fraction holds a signed distance until finish
					frac[ax][id] = phi;
					bnormal[ax][id] = n[ax];
				}
			});
		}
		finishBoundary();
	}

	void resetBoundary()
	{
		double dorg[3], dlen[3], center[3], half[3];
		for (int a = 0; a < D; a++) {
			dorg[a] = origin[a] + bw * dx;                         // domainOrigin (StaggeredGrid.h:40)
			dlen[a] = double(res[a] - bw * 2) * dx;               // domainLengths (StaggeredGrid.h:41)
			center[a] = dorg[a] + dlen[a] / 2;                     // ImplicitBox ctor (ImplicitSurface.h:37)
			half[a] = dlen[a] / 2;
		}
		parFor<D>(node, [&](const int *c) {
			double p[3]; node.position(c, p);
			nodeDist[node.index(c)] = -boxSdf<D>(center, half, p);
		});
		parFor<D>(cell, [&](const int *c) {
			double p[3]; cell.position(c, p);
			cellDist[cell.index(c)] = -boxSdf<D>(center, half, p);
		});
		for (int ax = 0; ax < D; ax++) {
			parFor<D>(face[ax], [&](const int *c) {
				double p[3], nrm[3]; face[ax].position(c, p);
				const size_t id = face[ax].index(c);
				frac[ax][id] = -boxSdf<D>(center, half, p);
				boxNormal<D>(center, half, p, nrm);
				bnormal[ax][id] = -nrm[ax];
			});
		}
	}

	void finishBoundary()
	{
		for (int ax = 0; ax < D; ax++) {
			parFor<D>(face[ax], [&](const int *c) {
				const size_t id = face[ax].index(c);
				frac[ax][id] = isBoundaryFace(ax, c) ? 1.0 : faceFraction(ax, c);
			});
		}
	}

	// StaggeredGrid::isBoundaryFace (StaggeredGrid.h:57)
	bool isBoundaryFace(int ax, const int *c) const
	{
		return c[ax] <= bw || c[ax] >= res[ax] - bw || !face[ax].inside(c, bw);
	}

	// Boundary::getFaceFraction (Boundary.cpp:135-174), StaggeredGrid::faceNode (StaggeredGrid.h:77-81)
	double faceFraction(int ax, const int *c) const
	{
		auto theta = [](double a, double b) { return a / (a - b); };
		double ff;
		if constexpr (D == 2) {
			int n0[3] = {c[0], c[1], 0}, n1[3] = {c[0], c[1], 0};
			n1[ax ^ 1] += 1;
			const double p0 = nodeDist[node.index(n0)], p1 = nodeDist[node.index(n1)];
			if (p0 <= 0 && p1 <= 0) ff = 1;
			else if (p0 <= 0 && p1 > 0) ff = theta(p0, p1);
			else if (p0 > 0 && p1 <= 0) ff = theta(p1, p0);
			else ff = 0;
		} else {
			auto tri = [&](double p0, double p1, double p2) -> double {
				if (p0 > p1) std::swap(p0, p1);
				if (p1 > p2) std::swap(p1, p2);
				if (p0 > p1) std::swap(p0, p1);
				if (p2 <= 0) return 1;
				else if (p1 <= 0) return 1 - theta(p2, p0) * theta(p2, p1);
				else if (p0 <= 0) return theta(p0, p1) * theta(p0, p2);
				else return 0;
			};
			double ph[4];
			for (int ord = 0; ord < 4; ord++) {
				int q[3] = {c[0], c[1], c[2]};
				q[(ax + 1) % 3] += ord & 1;
				q[(ax + 2) % 3] += ord >> 1 & 1;
				ph[ord] = nodeDist[node.index(q)];
			}
			const double cp = (ph[0] + ph[1] + ph[2] + ph[3]) * .25;
			ff = (tri(cp, ph[0], ph[1]) + tri(cp, ph[1], ph[3]) + tri(cp, ph[3], ph[2]) + tri(cp, ph[2], ph[0])) * .25;
		}
		return ff > .9 ? 1 : ff;
	}

	// ------------------------------------------------------ scene set-up ---
	// unionLevelSet (LevelSetLiquid.cpp:700-707)
	template <typename F>
	void unionLevelSet(F sdf)
	{
		parFor<D>(cell, [&](const int *c) {
			double p[3]; cell.position(c, p);
			const size_t id = cell.index(c);
			phi[id] = std::min(phi[id], sdf(p));
		});
	}

	// ----------------------------------------------------- step pieces ---
	// enforceBoundaryLevelSet(sdf, tol) (LevelSetLiquid.cpp:966-975)
	void enforceBoundaryLevelSet(std::vector<double> &sdf, double tol)
	{
		const size_t N = cell.count();
#pragma omp parallel for
		for (long long i = 0; i < (long long)N; i++)
			if (sdf[i] < -cellDist[i] - tol) sdf[i] = -cellDist[i] - tol;
	}

	// Derivatives::computeFirst<Order> for a scalar cell array into a
	// D-vector array (Derivatives.h:47-48, 51-57, 71-77, 98-105)
	template <int Order>
	void computeFirst(const Grid<D> &g, const double *val, double *drv)
	{
		parFor<D>(g, [&](const int *c) {
			const size_t id = g.index(c);
			for (int a = 0; a < D; a++) {
				double v[5];
				for (int i = 0; i <= Order; i++) {
					int q[3] = {c[0], c[1], c[2]};
					q[a] += i - (Order >> 1);
					v[i] = val[g.at(q)];
				}
				if constexpr (Order == 2) drv[id * D + a] = (v[2] - v[0]) / 2 * g.inv;
				else drv[id * D + a] = (v[0] - 8 * v[1] + 8 * v[3] - v[4]) / 12 * g.inv;
			}
		});
	}

	// VectorFieldWrapper<SGrid, Linear>::operator() (VectorField.h:66-72)
	void flowVel(const double *p, double *v) const
	{
		for (int a = 0; a < D; a++) lerp<D, 1>(face[a], u[a].data(), p, &v[a]);
	}
	// DrvVectorFieldWrapper<SGrid, Linear>::gradient (VectorField.h:129-135):
	// row a = multilinear sample of grad u_a; G is row-major here, G[a*D+b].
	void flowGrad(const std::vector<double> *gu, const double *p, double *G) const
	{
		for (int a = 0; a < D; a++) lerp<D, D>(face[a], gu[a].data(), p, &G[a * D]);
	}

	// advectFields, reference-map part (LevelSetLiquid.cpp:739-748):
	// computeFirst<2> on u, then Advection::solve<SemiLagrangian<D,2,Hermite>>
	// (Advection.h:24-31, 71-82 traceDrv<2>, 147-160).
	void advectRefMap(double dt)
	{
		std::vector<double> gu[D];
		for (int a = 0; a < D; a++) { gu[a].assign(face[a].count() * D, 0.0); computeFirst<2>(face[a], u[a].data(), gu[a].data()); }
		Hermite<D, D> H;
		H.build(cell, J.data());
		const size_t N = cell.count();
		std::vector<double> nxi(N * D), nJ(N * D * D);
		const double h = -dt;
		parFor<D>(cell, [&](const int *c) {
			const size_t id = cell.index(c);
			double x[3]; cell.position(c, x);
			// traceDrv<2>
			double v0[3], G0[9], p1[3], R1[9], v1[3], G1[9], X[3], R[9];
			flowVel(x, v0); flowGrad(gu, x, G0);
			for (int a = 0; a < D; a++) p1[a] = x[a] + v0[a] * h;
			for (int a = 0; a < D; a++) for (int b = 0; b < D; b++) R1[a * D + b] = (a == b ? 1.0 : 0.0) + G0[a * D + b] * h;
			flowVel(p1, v1); flowGrad(gu, p1, G1);
			for (int a = 0; a < D; a++) X[a] = x[a] + (v0[a] + v1[a]) * h / 2;
			for (int a = 0; a < D; a++)
				for (int b = 0; b < D; b++) {
					double gr = 0;
					for (int k = 0; k < D; k++) gr += G1[a * D + k] * R1[k * D + b];
					R[a * D + b] = (a == b ? 1.0 : 0.0) + (G0[a * D + b] + gr) * h / 2;
				}
			// xi' = H[xi](X); J' = grad H[xi](X) * R
			double val[3], Gh[9];   // Gh column-major (r + c*D)
			H.interpolateBoth(cell, xi.data(), J.data(), X, val, Gh);
			for (int r = 0; r < D; r++) nxi[id * D + r] = val[r];
			for (int r = 0; r < D; r++)
				for (int col = 0; col < D; col++) {
					double s = 0;
					for (int k = 0; k < D; k++) s += Gh[k * D + r] * R[k * D + col];
					nJ[id * D * D + col * D + r] = s;
				}
		});
		xi.swap(nxi);
		J.swap(nJ);
	}

	// Advection::trace<2> (Advection.h:37-43) through the multilinear velocity
	void trace2(const double *x, double h, double *X) const
	{
		double v0[3], p1[3], v1[3];
		flowVel(x, v0);
		for (int a = 0; a < D; a++) p1[a] = x[a] + v0[a] * h;
		flowVel(p1, v1);
		for (int a = 0; a < D; a++) X[a] = x[a] + (v0[a] + v1[a]) * h / 2;
	}

	// advectFields, velocity part (LevelSetLiquid.cpp:825; Advection.h:15-22,
	// 137-145): every face of every axis, backtraced with the old velocity.
	void advectVelocity(double dt)
	{
		std::vector<double> nu[D];
		for (int ax = 0; ax < D; ax++) {
			nu[ax].assign(face[ax].count(), 0.0);
			parFor<D>(face[ax], [&](const int *c) {
				double x[3], X[3];
				face[ax].position(c, x);
				trace2(x, -dt, X);
				nu[ax][face[ax].index(c)] = cubic<D>(face[ax], u[ax].data(), X);
			});
		}
		for (int ax = 0; ax < D; ax++) u[ax].swap(nu[ax]);
	}

	// generateCurrentFields (LevelSetLiquid.cpp:1218-1244), with
	// _levelSetGrad (Newton build) and prettified output.
	void compose()
	{
		const double band = 15.0;   // _airExtrapolationRadius (LevelSetLiquid.h:93)
		parFor<D>(cell, [&](const int *c) {
			const size_t id = cell.index(c);
			double val, g[3];
			refView.interpolateBoth(cell, phiRef.data(), phiRefGrad.data(), &xi[id * D], &val, g);
			phi[id] = val;
			// J^T * grad
			for (int col = 0; col < D; col++) {
				double s = 0;
				for (int r = 0; r < D; r++) s += J[id * D * D + col * D + r] * g[r];
				phiGrad[id * D + col] = s;
			}
			if (advRegion[id] == 0 && phi[id] < band * cell.dx) phi[id] = band * cell.dx;
		});
		enforceBoundaryLevelSet(phi, 3.0 * dx);
		phiTmp = phi;
		enforceBoundaryLevelSet(phiTmp, 0.0 * dx);
		phiOri = phi;
	}

	// Derivatives::weno5 (Derivatives.h:134-159)
	static void weno5(const double *v, double invDx, double &dm, double &dp)
	{
		for (int i = 0; i < 2; i++) {
			double dev[5];
			if (i) { dev[0] = v[6] - v[5]; dev[1] = v[5] - v[4]; dev[2] = v[4] - v[3]; dev[3] = v[3] - v[2]; dev[4] = v[2] - v[1]; }
			else { dev[0] = v[1] - v[0]; dev[1] = v[2] - v[1]; dev[2] = v[3] - v[2]; dev[3] = v[4] - v[3]; dev[4] = v[5] - v[4]; }
			double eps = 0;
			for (int j = 0; j < 5; j++) eps = std::max(eps, dev[j] * dev[j]);
			eps = eps * 1e-6 + 1e-99;
			const double phix1 = dev[0] / 3 - dev[1] * 7 / 6 + dev[2] * 11 / 6;
			const double phix2 = -dev[1] / 6 + dev[2] * 5 / 6 + dev[3] / 3;
			const double phix3 = dev[2] / 3 + dev[3] * 5 / 6 - dev[4] / 6;
			const double s1 = sq(dev[0] - 2 * dev[1] + dev[2]) * 13 / 12 + sq(dev[0] - 4 * dev[1] + 3 * dev[2]) / 4;
			const double s2 = sq(dev[1] - 2 * dev[2] + dev[3]) * 13 / 12 + sq(dev[1] - dev[3]) / 4;
			const double s3 = sq(dev[2] - 2 * dev[3] + dev[4]) * 13 / 12 + sq(3 * dev[2] - 4 * dev[3] + dev[4]) / 4;
			const double a1 = .1 / sq(s1 + eps);
			const double a2 = .6 / sq(s2 + eps);
			const double a3 = .3 / sq(s3 + eps);
			const double sum = a1 + a2 + a3;
			(i ? dp : dm) = (a1 * phix1 + a2 * phix2 + a3 * phix3) / sum * invDx;
		}
	}

	// Reinitialization::advance with weno5 (Reinitialization.h:20-48)
	void reinitAdvance(const std::vector<double> &f, std::vector<double> &nf, double dtau)
	{
		parFor<D>(cell, [&](const int *c) {
			const size_t id = cell.index(c);
			double dm[3], dp[3];
			for (int a = 0; a < D; a++) {
				double v[7];
				for (int i = 0; i < 7; i++) {
					int q[3] = {c[0], c[1], c[2]};
					q[a] += i - 3;
					v[i] = f[cell.at(q)];
				}
				weno5(v, cell.inv, dm[a], dp[a]);
			}
			double sum = 0;
			if (f[id] < 0) {
				for (int a = 0; a < D; a++) { sum += sq(std::min(dm[a], 0.0)); sum += sq(std::max(dp[a], 0.0)); }
			} else {
				for (int a = 0; a < D; a++) { sum += sq(std::max(dm[a], 0.0)); sum += sq(std::min(dp[a], 0.0)); }
			}
			sum = std::sqrt(sum);
			const double val = f[id], sp = cell.dx * sum;
			nf[id] = val + dtau * (val / std::sqrt(val * val + sp * sp)) * (1 - sum);
		});
	}

	// Reinitialization::solve<5> (Reinitialization.h:50-80)
	void reinit(int iters, double cfl)
	{
		const double dtau = cell.dx * cfl;
		const size_t N = cell.count();
		std::vector<double> p1(N), p2(N);
		for (int it = 0; it < iters; it++) {
			reinitAdvance(phi, p1, dtau);
			reinitAdvance(p1, p2, dtau);
#pragma omp parallel for
			for (long long i = 0; i < (long long)N; i++) p1[i] = (phi[i] * 3 + p2[i]) / 4;
			reinitAdvance(p1, p2, dtau);
#pragma omp parallel for
			for (long long i = 0; i < (long long)N; i++) phi[i] = (phi[i] + p2[i] * 2) / 3;
		}
	}

	// Extrapolation::solve for cell data with C doubles (Extrapolation.h:88-127
	// for C = 1, 178-218 for the reference map with `resetToPosition`).
	template <int C, typename Fixed>
	void extrapolateCells(std::vector<double> &g, double maxSteps, Fixed isFixed, bool resetToPosition)
	{
		const size_t N = cell.count();
		std::vector<double> ng = g;
		std::vector<u8> vis0(N), vis(N, 0);
		parFor<D>(cell, [&](const int *c) { vis0[cell.index(c)] = isFixed(c) ? 1 : 0; });
		for (int it = 0; it < maxSteps; it++) {
			parFor<D>(cell, [&](const int *c) {
				const size_t id = cell.index(c);
				if (!vis0[id]) {
					int cnt = 0;
					double sum[3] = {0, 0, 0};
					for (int i = 0; i < 2 * D; i++) {
						int q[3] = {c[0], c[1], c[2]};
						q[i >> 1] += (i & 1) ? 1 : -1;
						if (cell.valid(q) && vis0[cell.index(q)]) {
							const size_t nb = cell.index(q);
							for (int r = 0; r < C; r++) sum[r] += ng[nb * C + r];
							cnt++;
						}
					}
					if (cnt > 0) {
						for (int r = 0; r < C; r++) ng[id * C + r] = sum[r] / cnt;
						vis[id] = 1;
					} else if (resetToPosition) {
						double p[3]; cell.position(c, p);
						for (int r = 0; r < C; r++) ng[id * C + r] = p[r];
					}
				} else vis[id] = 1;
			});
			std::swap(vis0, vis);
		}
		g.swap(ng);
	}

	// Extrapolation::solve for uchar data (Extrapolation.h:130-175)
	template <typename Fixed>
	void extrapolateMask(std::vector<u8> &g, double maxSteps, Fixed isFixed)
	{
		const size_t N = cell.count();
		std::vector<u8> ng = g, vis0(N), vis(N, 0);
		parFor<D>(cell, [&](const int *c) { vis0[cell.index(c)] = isFixed(c) ? 1 : 0; });
		for (int it = 0; it < maxSteps; it++) {
			parFor<D>(cell, [&](const int *c) {
				const size_t id = cell.index(c);
				if (!vis0[id]) {
					int cnt = 0;
					for (int i = 0; i < 2 * D; i++) {
						int q[3] = {c[0], c[1], c[2]};
						q[i >> 1] += (i & 1) ? 1 : -1;
						if (cell.valid(q) && vis0[cell.index(q)]) cnt++;
					}
					if (cnt > 0) { ng[id] = 1; vis[id] = 1; }
				} else vis[id] = 1;
			});
			std::swap(vis0, vis);
		}
		for (size_t i = 0; i < N; i++) if (!vis0[i]) ng[i] = 0;
		g.swap(ng);
	}

	// Extrapolation::solve for face data (Extrapolation.h:44-84)
	template <typename Fixed>
	void extrapolateFaces(double maxSteps, Fixed isFixed)
	{
		std::vector<double> nu[D];
		std::vector<u8> vis0[D], vis[D];
		for (int ax = 0; ax < D; ax++) {
			nu[ax] = u[ax];
			vis0[ax].assign(face[ax].count(), 0);
			vis[ax].assign(face[ax].count(), 0);
			parFor<D>(face[ax], [&](const int *c) { vis0[ax][face[ax].index(c)] = isFixed(ax, c) ? 1 : 0; });
		}
		for (int it = 0; it < maxSteps; it++) {
			for (int ax = 0; ax < D; ax++) {
				const Grid<D> &g = face[ax];
				parFor<D>(g, [&](const int *c) {
					const size_t id = g.index(c);
					if (!vis0[ax][id]) {
						int cnt = 0;
						double sum = 0;
						for (int i = 0; i < 2 * D; i++) {
							int q[3] = {c[0], c[1], c[2]};
							q[i >> 1] += (i & 1) ? 1 : -1;
							if (g.valid(q) && vis0[ax][g.index(q)]) { sum += nu[ax][g.index(q)]; cnt++; }
						}
						if (cnt > 0) { nu[ax][id] = sum / cnt; vis[ax][id] = 1; }
						else nu[ax][id] = 0;
					} else vis[ax][id] = 1;
				});
			}
			for (int ax = 0; ax < D; ax++) std::swap(vis0[ax], vis[ax]);
		}
		for (int ax = 0; ax < D; ax++) u[ax].swap(nu[ax]);
	}

	// extrapolateFieldsToCollider (LevelSetLiquid.cpp:1086-1137)
	void extrapolateCollider()
	{
		const double radius = 6.0;   // _boundaryExtrapolationRadius (LevelSetLiquid.h:92)
		auto fixed = [&](const int *c) { return cellDist[cell.index(c)] > 0; };
		extrapolateCells<1>(phiOri, radius, fixed, false);
		extrapolateCells<1>(phi, radius, fixed, false);
	}

	// extrapolateFieldsToAir (LevelSetLiquid.cpp:1059-1084)
	void extrapolateAir()
	{
		const double radius = 15.0;  // _airExtrapolationRadius (LevelSetLiquid.h:93)
		extrapolateCells<D>(xi, radius, [&](const int *c) { return advRegion[cell.index(c)] > 0; }, true);
		const size_t N = cell.count();
		for (size_t i = 0; i < N; i++) advRegion[i] = phiOri[i] <= 0 ? 1 : 0;
		extrapolateMask(advRegion, radius, [&](const int *c) { return phiOri[cell.index(c)] <= 0; });
		extrapolateAirVelocity();
	}

	// extrapolateFieldsToAir(true): the velocity alone (LevelSetLiquid.cpp:1075-1083)
	void extrapolateAirVelocity()
	{
		const double radius = 15.0;
		extrapolateFaces(radius + 2, [&](int ax, const int *c) {
			const size_t id = face[ax].index(c);
			if (!(frac[ax][id] < 1)) return false;
			int c0[3] = {c[0], c[1], c[2]};
			c0[ax] -= 1;   // StaggeredGrid::faceAdjacentCell (StaggeredGrid.h:85)
			return phiOri[cell.index(c0)] <= 0 || phiOri[cell.index(c)] <= 0;
		});
	}

	// Boundary::enforce(velocity) (Boundary.cpp:89-120)
	void enforceBoundary()
	{
		std::vector<double> nu[D];
		for (int ax = 0; ax < D; ax++) nu[ax] = u[ax];
		for (int ax = 0; ax < D; ax++) {
			parFor<D>(face[ax], [&](const int *c) {
				const size_t id = face[ax].index(c);
				const double fr = frac[ax][id];
				if (fr > 0.1 && fr <= 1) {
					double p[3], n[3], vel[3], vel0[3];
					face[ax].position(c, p);
					for (int a = 0; a < D; a++) lerp<D, 1>(face[a], bnormal[a].data(), p, &n[a]);
					double s = 0;
					for (int a = 0; a < D; a++) s += n[a] * n[a];
					if (s > 0) { const double r = std::sqrt(s); for (int a = 0; a < D; a++) n[a] /= r; }
					bool any = false;
					for (int a = 0; a < D; a++) if (n[a] != 0) any = true;
					if (any) {
						for (int a = 0; a < D; a++) {
							lerp<D, 1>(face[a], u[a].data(), p, &vel[a]);
							lerp<D, 1>(face[a], bvel[a].data(), p, &vel0[a]);
						}
						double dv[3], vn = 0;
						for (int a = 0; a < D; a++) { dv[a] = vel[a] - vel0[a]; vn += dv[a] * n[a]; }
						if (vn < 0) {
							const double vt = dv[ax] - vn * n[ax];
							nu[ax][id] = (vel[ax] - vn * n[ax]) - 0.8 * vt;
						}
					}
				}
			});
		}
		for (int ax = 0; ax < D; ax++) u[ax].swap(nu[ax]);
	}

	// applyBodyForces, gravity (LevelSetLiquid.cpp:837-841; kGravity = 9.8, Constants.h:9)
	void gravity(double dt)
	{
		for (auto &v : u[1]) v -= 9.8 * dt;
	}

	// checkReferenceMap (LevelSetLiquid.cpp:1246-1283)
	bool checkRefMap(double threshold, double bandStep, double ratio, long long *totOut = nullptr, long long *cntOut = nullptr) const
	{
		long long tot = 0, cnt = 0;
		const size_t N = cell.count();
		for (size_t id = 0; id < N; id++) {
			if (std::abs(phiOri[id]) < bandStep * dx && cellDist[id] > 0) {
				tot++;
				const double *m = &J[id * D * D];
				auto dot = [&](int a, int b) { double s = 0; for (int r = 0; r < D; r++) s += m[a * D + r] * m[b * D + r]; return s; };
				if constexpr (D == 2) {
					const double vab = dot(0, 1) * dot(0, 1) / dot(0, 0) / dot(1, 1);
					if (vab > threshold) cnt++;
				} else {
					const double vab = sq(dot(0, 1)) / dot(0, 0) / dot(1, 1);
					const double vac = sq(dot(0, 2)) / dot(0, 0) / dot(2, 2);
					const double vbc = sq(dot(1, 2)) / dot(1, 1) / dot(2, 2);
					if (vab > threshold || vac > threshold || vbc > threshold) cnt++;
				}
			}
		}
		if (totOut) *totOut = tot;
		if (cntOut) *cntOut = cnt;
		return 1.0 * cnt <= ratio * tot;
	}

	// restart, branch without Hybrid Newton or marching cubes
	// (LevelSetLiquid.cpp:1331-1352)
	void restart()
	{
		phiRef = phi;
		computeFirst<4>(cell, phiRef.data(), phiRefGrad.data());
		refView.build(cell, phiRefGrad.data());
		resetRefMap();
	}

	// ------------------------------------------ Hybrid-Newton restart ---
	// LevelSetLiquid::restart of the default build (RM, GA, Newton) with
	// enable-fmm=n (LevelSetLiquid.cpp:1290-1303): closest-point guesses from
	// the cubic view of phi, then HybridNewton::perform on (phiOri, phiGrad)
	// (HybridNewton.cpp:18-28): interface cells by the modified Newton
	// iteration (HybridNewton.h:119-136, 147-467 with ENABLE_MODIFIED),
	// every other cell by 20 characteristic steps (HybridNewton.cpp:30-49);
	// the result becomes phiRef / phiRefGrad.  The restatement keeps the
	// reference's operation order (Eigen's coefficient-wise expressions).
	static void normalized(const double *v, double *out)
	{
		double s = 0;
		for (int a = 0; a < D; a++) s += v[a] * v[a];
		if (s > 0) { const double n = std::sqrt(s); for (int a = 0; a < D; a++) out[a] = v[a] / n; }
		else for (int a = 0; a < D; a++) out[a] = v[a];
	}

	// ScalarFieldWrapper<CubicIntrpl>::gradient (ScalarField.h:41-47):
	// linear weights of the corners times getFirst<2> at each (unclamped)
	// corner coordinate, whose stencil is clamped (Derivatives.h:51-57)
	void cubicViewGradient(const double *f, const double *pos, double *grad) const
	{
		int l[3] = {0, 0, 0}; double fr[3];
		cell.lower(pos, l);
		cell.frac(pos, l, fr);
		double w[2][3];
		for (int a = 0; a < D; a++) { w[0][a] = 1 - fr[a]; w[1][a] = fr[a]; }
		for (int a = 0; a < D; a++) grad[a] = 0;
		for (int k = 0; k < (D == 3 ? 2 : 1); k++)
			for (int j = 0; j < 2; j++)
				for (int i = 0; i < 2; i++) {
					const int c[3] = {l[0] + i, l[1] + j, l[2] + k};
					const double wt = D == 2 ? w[i][0] * w[j][1] : w[i][0] * w[j][1] * w[k][2];
					for (int a = 0; a < D; a++) {
						int cp[3] = {c[0], c[1], c[2]}, cm[3] = {c[0], c[1], c[2]};
						cp[a] += 1;
						cm[a] -= 1;
						grad[a] += (f[cell.at(cp)] - f[cell.at(cm)]) / 2 * cell.inv * wt;
					}
				}
	}

	// HybridNewton::getClosestPosition (HybridNewton.h:147-467): ENABLE_MODIFIED,
	// no FMM, epsl = 0; returns the position, *converged the flag
	void closestPosition(const Hermite<D, 1> &H, const double *val, const double *drv, const double *start, const double *init, double *pos,
		bool *converged) const
	{
		const double dx = cell.dx;
		const double tolerance = 1e-3 * dx * dx * (D == 2 ? 1 : dx);
		const double radius = .9 * dx;
		for (int a = 0; a < D; a++) pos[a] = init[a];
		*converged = false;
		for (int iter = 0; iter < 20; iter++) {
			double phi, psi[3];
			H.interpolateBoth(cell, val, drv, pos, &phi, psi);
			phi = phi - 0.0;
			double sqn = 0, dot = 0, d[3], delta[3];
			for (int a = 0; a < D; a++) sqn += psi[a] * psi[a];
			for (int a = 0; a < D; a++) d[a] = start[a] - pos[a];
			for (int a = 0; a < D; a++) dot += d[a] * psi[a];
			for (int a = 0; a < D; a++) {
				const double d1 = -phi * psi[a] / sqn;
				const double d2 = d[a] - psi[a] * dot / sqn;
				delta[a] = d1 + d2;
			}
			double n2 = 0;
			for (int a = 0; a < D; a++) n2 += delta[a] * delta[a];
			double norm = std::sqrt(n2);
			if (norm > radius) {
				const double sc = radius / norm;
				for (int a = 0; a < D; a++) delta[a] *= sc;
				norm = radius;
			}
			for (int a = 0; a < D; a++) pos[a] += delta[a];
			if (norm < tolerance) { *converged = true; break; }
		}
	}

	void restartNewton(int maxSteps = 20, double cfl = 0.9)
	{
		const size_t N = cell.count();
		// LevelSetLiquid.cpp:1292-1296: points = pos - phi(pos) grad(pos)
		std::vector<double> pts(N * D);
		parFor<D>(cell, [&](const int *c) {
			double p[3], g[3];
			cell.position(c, p);
			const double v = cubic<D>(cell, phi.data(), p);
			cubicViewGradient(phi.data(), p, g);
			const size_t id = cell.index(c);
			for (int a = 0; a < D; a++) pts[id * D + a] = p[a] - v * g[a];
		});
		std::vector<double> f = phiOri, psi = phiGrad;
		// prepareInterface + extrapolate(1) (HybridNewton.cpp:64-79, 103-123)
		std::vector<uint8_t> mark0(N, 0), mark(N, 0);
		parFor<D>(cell, [&](const int *c) {
			const size_t id = cell.index(c);
			for (int o = 0; o < 2 * D; o++) {
				int q[3] = {c[0], c[1], c[2]};
				q[o >> 1] += (o & 1) ? 1 : -1;
				if (cell.valid(q) && f[id] * f[cell.index(q)] <= 0) { mark0[id] = 1; break; }
			}
		});
		parFor<D>(cell, [&](const int *c) {
			const size_t id = cell.index(c);
			if (mark0[id]) { mark[id] = 1; return; }
			for (int o = 0; o < 2 * D; o++) {
				int q[3] = {c[0], c[1], c[2]};
				q[o >> 1] += (o & 1) ? 1 : -1;
				if (cell.valid(q) && mark0[cell.index(q)]) { mark[id] = 1; break; }
			}
		});
		newtonMark = mark;
		// initializeInterface with the guesses (HybridNewton.h:119-136)
		{
			Hermite<D, 1> H;
			H.build(cell, psi.data());
			std::vector<double> nf = f, npsi = psi;
			parFor<D>(cell, [&](const int *c) {
				const size_t id = cell.index(c);
				if (!mark[id]) return;
				double p[3], cp[3], g[3], v;
				cell.position(c, p);
				bool conv;
				closestPosition(H, f.data(), psi.data(), p, &pts[id * D], cp, &conv);
				double n2 = 0;
				for (int a = 0; a < D; a++) n2 += (p[a] - cp[a]) * (p[a] - cp[a]);
				nf[id] = std::sqrt(n2) * (f[id] < 0 ? -1 : 1);
				H.interpolateBoth(cell, f.data(), psi.data(), cp, &v, g);
				normalized(g, &npsi[id * D]);
			});
			f.swap(nf);
			psi.swap(npsi);
		}
		// advect (HybridNewton.cpp:30-49): characteristic steps of length
		// cfl dx along the normalised gradient, RK3 (Advection.h:43-49)
		const double dt = cell.dx * cfl;
		for (int iter = 0; iter < maxSteps; iter++) {
			Hermite<D, 1> H;
			H.build(cell, psi.data());
			std::vector<double> nf = f, npsi = psi;
			auto vel = [&](const double *x, double *v) {
				double val, g[3];
				H.interpolateBoth(cell, f.data(), psi.data(), x, &val, g);
				bool any = false;
				for (int a = 0; a < D; a++) any = any || g[a] != 0;
				if (any) normalized(g, v);
				else for (int a = 0; a < D; a++) v[a] = a == 0 ? 1 : 0;
			};
			parFor<D>(cell, [&](const int *c) {
				const size_t id = cell.index(c);
				if (mark[id]) return;
				const int sgn = f[id] < 0 ? -1 : 1;
				double p[3], v0[3], v1[3], v2[3], q[3], t[3];
				cell.position(c, p);
				const double h = -sgn * dt;
				vel(p, v0);
				for (int a = 0; a < D; a++) q[a] = p[a] + v0[a] * h;
				vel(q, v1);
				for (int a = 0; a < D; a++) q[a] = p[a] + (v0[a] + v1[a]) * h / 4;
				vel(q, v2);
				for (int a = 0; a < D; a++) t[a] = p[a] + (v0[a] + v1[a] + 4 * v2[a]) * h / 6;
				double val, g[3];
				H.interpolateBoth(cell, f.data(), psi.data(), t, &val, g);
				nf[id] = val + sgn * dt;
				normalized(g, &npsi[id * D]);
			});
			f.swap(nf);
			psi.swap(npsi);
		}
		phiRef = f;
		phiRefGrad = psi;
		refView.build(cell, phiRefGrad.data());
		resetRefMap();
	}

	void resetRefMap()
	{
		parFor<D>(cell, [&](const int *c) {
			const size_t id = cell.index(c);
			double p[3]; cell.position(c, p);
			for (int a = 0; a < D; a++) xi[id * D + a] = p[a];
			for (int e = 0; e < D * D; e++) J[id * D * D + e] = (e % (D + 1) == 0) ? 1.0 : 0.0;
		});
	}

	// ---------------------------------------------------------- project ---
	// Projector::setUnknowns (Projector.cpp:125-156)
	int setUnknowns()
	{
		std::fill(grid2mat.begin(), grid2mat.end(), -1);
		mat2grid.clear();
		int n = 0;
		serFor<D>(cell, [&](const int *c) {
			const size_t id = cell.index(c);
			if (phi[id] <= 0) {
				bool flag = false;
				for (int i = 0; i < 2 * D; i++) {
					const int ax = i >> 1;
					int f[3] = {c[0], c[1], c[2]};
					f[ax] += i & 1;
					if (frac[ax][face[ax].index(f)] < 1) { flag = true; break; }
				}
				if (flag) { grid2mat[id] = n++; mat2grid.push_back(int(id)); }
			}
		});
		return n;
	}

	struct System {
		int n = 0;
		std::vector<double> diag, off, rhs;   // off[r*2D + i] = A(r, nb_i) (negative weight) or 0
		std::vector<int> nb;                 // nb[r*2D + i] = unknown index or -1
	};

	// Projector::buildLinearSystem (Projector.cpp:30-68) without a pressure jump
	void buildSystem(System &S)
	{
		const int n = int(mat2grid.size());
		S.n = n;
		S.diag.assign(n, 0.0); S.rhs.assign(n, 0.0);
		S.off.assign(size_t(n) * 2 * D, 0.0); S.nb.assign(size_t(n) * 2 * D, -1);
		for (int r = 0; r < n; r++) {
			int c[3] = {0, 0, 0};
			cell.coord(mat2grid[r], c);
			const size_t id = cell.index(c);
			double diagCoeff = 0, div = 0;
			for (int i = 0; i < 2 * D; i++) {
				const int ax = i >> 1;
				const int side = (i & 1) ? 1 : -1;
				int q[3] = {c[0], c[1], c[2]};
				q[ax] += side;
				int f[3] = {c[0], c[1], c[2]};
				f[ax] += i & 1;
				const size_t fid = face[ax].index(f);
				const double w = 1 - frac[ax][fid];
				if (w > 0) {
					const int cc = grid2mat[cell.index(q)];
					if (cc >= 0) {
						diagCoeff += w;
						S.off[size_t(r) * 2 * D + i] = -w;
						S.nb[size_t(r) * 2 * D + i] = cc;
					} else {
						const double theta = phi[id] / (phi[id] - phi[cell.index(q)]);
						const double intf = 1 / std::max(theta, .001);
						diagCoeff += w * intf;
					}
					div += side * w * u[ax][fid];
				}
				if (w < 1) div += side * (1 - w) * bvel[ax][fid];
			}
			S.rhs[r] = div;
			S.diag[r] = diagCoeff;
		}
	}

	void spmv(const System &S, const std::vector<double> &p, std::vector<double> &out) const
	{
#pragma omp parallel for
		for (int r = 0; r < S.n; r++) {
			double s = S.diag[r] * p[r];
			for (int i = 0; i < 2 * D; i++) {
				const int c = S.nb[size_t(r) * 2 * D + i];
				if (c >= 0) s += S.off[size_t(r) * 2 * D + i] * p[c];
			}
			out[r] = s;
		}
	}

	// IC(0) in natural ordering: L(i,i)^2 = A(i,i) - sum_k L(i,k)^2 over lower
	// neighbours k; L(i,k) = A(i,k) / L(k,k) (a 5/7-point graph has no triangles).
	void icFactor(const System &S, std::vector<double> &Ld) const
	{
		Ld.assign(S.n, 0.0);
		for (int r = 0; r < S.n; r++) {
			double d = S.diag[r];
			for (int i = 0; i < 2 * D; i++) {
				const int c = S.nb[size_t(r) * 2 * D + i];
				if (c >= 0 && c < r) { const double l = S.off[size_t(r) * 2 * D + i] / Ld[c]; d -= l * l; }
			}
			Ld[r] = std::sqrt(d);
		}
	}
	void icSolve(const System &S, const std::vector<double> &Ld, const std::vector<double> &rr, std::vector<double> &z) const
	{
		// forward: L y = r
		for (int r = 0; r < S.n; r++) {
			double s = rr[r];
			for (int i = 0; i < 2 * D; i++) {
				const int c = S.nb[size_t(r) * 2 * D + i];
				if (c >= 0 && c < r) s -= (S.off[size_t(r) * 2 * D + i] / Ld[c]) * z[c];
			}
			z[r] = s / Ld[r];
		}
		// backward: L^T z = y
		for (int r = S.n - 1; r >= 0; r--) {
			double s = z[r];
			for (int i = 0; i < 2 * D; i++) {
				const int c = S.nb[size_t(r) * 2 * D + i];
				if (c >= 0 && c > r) s -= (S.off[size_t(r) * 2 * D + i] / Ld[r]) * z[c];
			}
			z[r] = s / Ld[r];
		}
	}

	// Conjugate gradient as published in Eigen's ConjugateGradient
	// (conjugate_gradient: residual test ||r||^2 < tol^2 ||b||^2, zero
	// initial guess), called as in LinearSystem::solve (LinearSystem.h:23-39).
	// precond: 0 identity, 1 Jacobi, 2 IC(0).  Returns 0 if converged.
	// guess: Eigen's solveWithGuess start (residual b - A x0); zero when null.
	int solveCG(const System &S, std::vector<double> &x, int precond, double tol, int maxIt, int *itersOut, double *errOut,
		const std::vector<double> *guess = nullptr) const
	{
		const int n = S.n;
		std::vector<double> r(S.rhs), z(n), p(n), t(n), Ld;
		if (guess) {
			x = *guess;
			spmv(S, x, t);
			for (int i = 0; i < n; i++) r[i] = S.rhs[i] - t[i];
		} else x.assign(n, 0.0);
		double bnorm2 = 0;
		for (int i = 0; i < n; i++) bnorm2 += S.rhs[i] * S.rhs[i];
		if (bnorm2 == 0) { x.assign(n, 0.0); *itersOut = 0; *errOut = 0; return 0; }
		tol = std::max(tol, std::numeric_limits<double>::epsilon() / std::sqrt(bnorm2));
		const double threshold = std::max(tol * tol * bnorm2, std::numeric_limits<double>::min());
		if (precond == 2) icFactor(S, Ld);
		// Jacobi as Eigen's DiagonalPreconditioner: stored reciprocal, then a product
		std::vector<double> invd(n);
		for (int i = 0; i < n; i++) invd[i] = S.diag[i] != 0 ? 1 / S.diag[i] : 1.0;
		auto applyM = [&](const std::vector<double> &in, std::vector<double> &out) {
			if (precond == 1) { for (int i = 0; i < n; i++) out[i] = invd[i] * in[i]; }
			else if (precond == 2) icSolve(S, Ld, in, out);
			else out = in;
		};
		double rn2 = 0;
		for (int i = 0; i < n; i++) rn2 += r[i] * r[i];
		if (rn2 < threshold) { *itersOut = 0; *errOut = std::sqrt(rn2 / bnorm2); return 0; }
		applyM(r, p);
		double absNew = 0;
		for (int i = 0; i < n; i++) absNew += r[i] * p[i];
		int it = 0;
		while (it < maxIt) {
			spmv(S, p, t);
			double pAp = 0;
			for (int i = 0; i < n; i++) pAp += p[i] * t[i];
			const double alpha = absNew / pAp;
			for (int i = 0; i < n; i++) { x[i] += alpha * p[i]; r[i] -= alpha * t[i]; }
			rn2 = 0;
			for (int i = 0; i < n; i++) rn2 += r[i] * r[i];
			if (rn2 < threshold) break;
			applyM(r, z);
			const double absOld = absNew;
			absNew = 0;
			for (int i = 0; i < n; i++) absNew += r[i] * z[i];
			const double beta = absNew / absOld;
			for (int i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
			it++;
		}
		*itersOut = it;
		*errOut = std::sqrt(rn2 / bnorm2);
		return *errOut <= tol ? 0 : 1;
	}

	// Projector::applyPressureGradient (Projector.cpp:91-123), no jump
	void applyPressureGradient(const std::vector<double> &x)
	{
		for (int ax = 0; ax < D; ax++) {
			parFor<D>(face[ax], [&](const int *f) {
				int c0[3] = {f[0], f[1], f[2]}, c1[3] = {f[0], f[1], f[2]};
				c0[ax] -= 1;
				if (!cell.valid(c0) || !cell.valid(c1)) return;
				const int id0 = grid2mat[cell.index(c0)], id1 = grid2mat[cell.index(c1)];
				if (id0 < 0 && id1 < 0) return;
				const double p0 = id0 >= 0 ? x[id0] : 0;
				const double p1 = id1 >= 0 ? x[id1] : 0;
				const size_t fid = face[ax].index(f);
				if (id0 >= 0 && id1 >= 0) u[ax][fid] += p1 - p0;
				else {
					const double phi0 = phi[cell.index(c0)], phi1 = phi[cell.index(c1)];
					if (phi0 <= 0 && phi1 <= 0) return;
					const double theta = phi0 / (phi0 - phi1);
					const double intf = 1 / std::max((phi0 <= 0 ? theta : 1 - theta), .001);
					u[ax][fid] += (p1 - p0) * intf;
				}
			});
		}
	}

	// Projector::project (Projector.cpp:13-20)
	int project(int precond, double tol, int maxIt, int *iters, double *err, int *unknowns)
	{
		const int n = setUnknowns();
		System S;
		buildSystem(S);
		std::vector<double> x;
		const int rc = solveCG(S, x, precond, tol, maxIt, iters, err);
		applyPressureGradient(x);
		std::fill(pres.begin(), pres.end(), 0.0);
		for (int r = 0; r < n; r++) pres[mat2grid[r]] = x[r];
		if (unknowns) *unknowns = n;
		lastX = x;
		return rc;
	}

	// Projector::reproject (Projector.cpp:22-28): the unknowns and operator of
	// the last projection (rebuilt from the unchanged phi and fractions), the
	// right-hand side of the current velocity (setRightHandSide,
	// Projector.cpp:70-89), solved from the last pressure.
	int reproject(int precond, double tol, int maxIt, int *iters, double *err, int *unknowns)
	{
		const int n = setUnknowns();
		System S;
		buildSystem(S);
		std::vector<double> x;
		std::vector<double> guess = lastX;
		guess.resize(n, 0.0);
		const int rc = solveCG(S, x, precond, tol, maxIt, iters, err, stWarm ? &guess : nullptr);
		applyPressureGradient(x);
		std::fill(pres.begin(), pres.end(), 0.0);
		for (int r = 0; r < n; r++) pres[mat2grid[r]] = x[r];
		if (unknowns) *unknowns = n;
		lastX = x;
		return rc;
	}

	// SurfaceTension::getCurvature (SurfaceTension.h:12-23) with the cubic view
	double curvature(const double *pos) const
	{
		const double invDx = 1. / dx;
		double kappa = 0;
		for (int a = 0; a < D; a++) {
			double q[3] = {pos[0], pos[1], D == 3 ? pos[2] : 0}, gp[3], gm[3];
			q[a] = pos[a] + dx;
			cubicViewGradient(phi.data(), q, gp);
			normalized(gp, gp);
			q[a] = pos[a] - dx;
			cubicViewGradient(phi.data(), q, gm);
			normalized(gm, gm);
			kappa += gp[a] - gm[a];
		}
		kappa *= invDx * .5;
		return std::abs(kappa) > invDx ? (kappa < 0 ? -1 : 1) * invDx : kappa;
	}

	// Derivatives::getFirst<2> along b at the unclamped face coordinate q of
	// face grid `ax`, stencil points clamped (Derivatives.h:47, 51-57)
	double faceFirst(int ax, const int *q, int b) const
	{
		int hi[3] = {q[0], q[1], q[2]}, lo[3] = {q[0], q[1], q[2]};
		hi[b] += 1;
		lo[b] -= 1;
		return (u[ax][face[ax].at(hi)] - u[ax][face[ax].at(lo)]) / 2 * face[ax].inv;
	}

	// SurfaceTension::solve (SurfaceTension.h:25-101): the capillary system
	// on the faces with |phi| < maxSteps dx (cubic view) and fraction < 1,
	// solved with the restated CG from the guess vel.
	int capillary(double coeff, double maxSteps, double dt, double tol, int maxIt, int precond)
	{
		const double invDx = 1. / dx;
		const double eps = maxSteps * dx;
		std::vector<int> g2m[D];
		std::vector<std::pair<int, size_t>> m2g;
		for (int ax = 0; ax < D; ax++) {
			g2m[ax].assign(face[ax].count(), -1);
			serFor<D>(face[ax], [&](const int *c) {
				double p[3] = {0, 0, 0};
				face[ax].position(c, p);
				const double v = cubic<D>(cell, phi.data(), p);
				const size_t id = face[ax].index(c);
				if (-eps < v && v < eps && frac[ax][id] < 1) {
					g2m[ax][id] = int(m2g.size());
					m2g.emplace_back(ax, id);
				}
			});
		}
		const int n = int(m2g.size());
		System S;
		S.n = n;
		S.diag.assign(n, 0.0); S.rhs.assign(n, 0.0);
		S.off.assign(size_t(n) * 2 * D, 0.0); S.nb.assign(size_t(n) * 2 * D, -1);
		std::vector<double> vel(n);
		const double weight = coeff * dt * dt / (dx * dx);
		for (int r = 0; r < n; r++) {
			const int ax = m2g[r].first;
			const Grid<D> &fg = face[ax];
			int c[3] = {0, 0, 0};
			fg.coord(m2g[r].second, c);
			double p[3] = {0, 0, 0};
			fg.position(c, p);
			double g[3] = {0, 0, 0}, nrm[3] = {0, 0, 0};
			cubicViewGradient(phi.data(), p, g);
			normalized(g, nrm);
			const double kappa = curvature(p);
			const double v = cubic<D>(cell, phi.data(), p);
			const double dirac = (v < -eps || v > eps) ? 0 : 0.5 * (1.0 + std::cos(3.141592653589793 * v / eps)) / eps;
			const double delta = std::max(dirac, 1e-7 * invDx);
			double jac[3] = {0, 0, 0}, hes[3][3] = {};
			for (int b = 0; b < D; b++) jac[b] = faceFirst(ax, c, b);
			for (int a1 = 0; a1 < D; a1++)
				for (int a2 = 0; a2 < D; a2++) {
					int qp[3] = {c[0], c[1], c[2]}, qm[3] = {c[0], c[1], c[2]};
					qp[a2] += 1;
					qm[a2] -= 1;
					hes[a1][a2] = (faceFirst(ax, qp, a1) - faceFirst(ax, qm, a1)) / 2 * fg.inv;
				}
			double nHn = 0, jn = 0;
			for (int j = 0; j < D; j++) {
				double t = 0;
				for (int i = 0; i < D; i++) t += nrm[i] * hes[i][j];
				nHn += t * nrm[j];
			}
			for (int i = 0; i < D; i++) jn += jac[i] * nrm[i];
			const size_t fid = fg.index(c);
			double diag = 1 / delta;
			vel[r] = S.rhs[r] = u[ax][fid] / delta - coeff * dt * (kappa * nrm[ax] + dt * (nHn + kappa * jn));
			for (int o = 0; o < 2 * D; o++) {
				int q[3] = {c[0], c[1], c[2]};
				q[o >> 1] += (o & 1) ? 1 : -1;
				if (!fg.valid(q)) continue;
				const size_t qf = fg.index(q);
				if (frac[ax][qf] == 1) continue;
				diag += weight;
				const int m = g2m[ax][qf];
				if (m != -1) {
					S.nb[size_t(r) * 2 * D + o] = m;
					S.off[size_t(r) * 2 * D + o] = -weight;
				} else S.rhs[r] += weight * u[ax][qf];
			}
			S.diag[r] = diag;
		}
		std::vector<double> x;
		const int rc = solveCG(S, x, precond, tol, maxIt, &stIters, &stErr, stGuess ? &vel : nullptr);
		for (int r = 0; r < n; r++) u[m2g[r].first][m2g[r].second] = x[r];
		stFaces = n;
		stBand = maxSteps;
		return rc;
	}

	// LevelSetLiquid::trickLittleDrop with checkDrop (LevelSetLiquid.cpp:1037-1057,
	// LevelSetLiquid.h:178-190)
	void littleDrop(double dt)
	{
		const double v0 = dx / dt;
		const double vt = v0 * .15;
		serFor<D>(cell, [&](const int *c) {
			if (phi[cell.index(c)] > 0) return;
			for (int o = 0; o < 2 * D; o++) {
				const int ax = o >> 1;
				int q[3] = {c[0], c[1], c[2]}, f[3] = {c[0], c[1], c[2]};
				q[ax] += (o & 1) ? 1 : -1;
				f[ax] += o & 1;
				if (frac[ax][face[ax].index(f)] != 0 || phi[cell.at(q)] <= 0) return;
			}
			int f1[3] = {c[0], c[1] + 1, c[2]};
			const size_t a = face[1].index(c), b = face[1].index(f1);
			const double v = (u[1][a] + u[1][b]) * .5;
			if (v < vt * 1.25 && v > -vt && u[1][a] > 0 && u[1][b] < 0) {
				u[1][a] -= vt + v;
				u[1][b] -= vt + v;
			}
		});
	}

	// LevelSetLiquid::applySurfaceTension + trickLittleDrop
	// (LevelSetLiquid.cpp:1026-1057); the capillary and pressure solves use
	// the given preconditioner (2: IC(0), as the reference's ICPCG)
	int surfaceTension(double dt, int precond, double tol, int maxIt, int *iters, double *err, int *unknowns)
	{
		extrapolateAirVelocity();
		enforceBoundary();
		stBand = stDeltaMin + (stDeltaMax - stDeltaMin) * 2 / (1 + std::exp(elapsed / stCritTime));
		int rc = capillary(stCoeff, stBand, dt, stTol, stMaxIt, stPrecond);
		const int rp = reproject(precond, tol, maxIt, iters, err, unknowns);
		if (stLittleDrop) littleDrop(dt);
		return rc ? rc : rp;
	}

	double maxSpeed() const
	{
		// SGridBasedData::normMax (SGridBasedData.h:48-52, GridBasedData.h:55-68)
		double m = 0;
		for (int a = 0; a < D; a++) {
			if (u[a].empty()) continue;
			auto mm = std::minmax_element(u[a].begin(), u[a].end());
			m = std::max(m, std::max(std::abs(*mm.first), std::abs(*mm.second)));
		}
		return m;
	}

	// LevelSetLiquid::initialize (LevelSetLiquid.cpp:528-598), Newton-build
	// gradient before reinitialisation, WENO5 reinitialisation.
	int initialize(int reinitIters, double cfl, int precond, double tol, int maxIt)
	{
		enforceBoundaryLevelSet(phi, 0);
		computeFirst<4>(cell, phi.data(), phiGrad.data());
		reinit(reinitIters, cfl);
		phiTmp = phi; phiRef = phi; phiOri = phi;
		phiRefGrad = phiGrad;
		refView.build(cell, phiRefGrad.data());
		resetRefMap();
		std::fill(advRegion.begin(), advRegion.end(), 1);
		int it, unk; double err;
		const int rc = project(precond, tol, maxIt, &it, &err, &unk);
		extrapolateAir();
		enforceBoundary();
		return rc;
	}

	// LevelSetLiquid::advance (LevelSetLiquid.cpp:608-698) with the
	// deviations listed at the top of this file.
	int step(double dt, int doReinit, int reinitIters, double cfl, int precond, double tol, int maxIt,
		int *restarted, int *iters, double *err, int *unknowns)
	{
		advectRefMap(dt);
		advectVelocity(dt);
		elapsed += dt * 0.5 * 3.141592653589793;   // LevelSetLiquid.cpp:619
		compose();
		if (doReinit) reinit(reinitIters, cfl);
		extrapolateCollider();
		*restarted = 0;
		if (!checkRefMap(0.883022221559489, 4, 1e-5)) {
			// the reference reinitialises every step, so a restart always
			// captures a reinitialised level set; with a longer cadence the
			// restart step reinitialises first
			if (!doReinit) reinit(reinitIters, cfl);
			if (newtonRestart) restartNewton();
			else restart();
			compose();
			extrapolateCollider();
			*restarted = 1;
		}
		gravity(dt);
		int rc = project(precond, tol, maxIt, iters, err, unknowns);
		if (stOn) {
			int it2, un2;
			double e2;
			const int r2 = surfaceTension(dt, precond, tol, maxIt, &it2, &e2, &un2);
			rc = rc ? rc : r2;
		}
		extrapolateAir();
		enforceBoundary();
		return rc;
	}
};

// ------------------------------------------------------ surface mesh ----
// Contourer<3>::contour at 0 (Contourer.cpp:9-97) and SurfaceMesh::
// computeNormals (SurfaceMesh.cpp:191-216) over a node field, the
// reference's numbering (StaggeredGrid.h:61-94: cube node k = (k&1, k>>1&1,
// k>>2&1), edge i along axis i>>2 offset (i&1, i>>1&1) along the next two
// axes) and vertex creation order.  The triangles of a cube type come from
// polygons made of directed face segments: walking each cube face's corners
// counter-clockwise as seen from outside the cube, a segment runs from each
// crossing entered from outside to the next crossing left to outside (on an
// ambiguous face the two around each inside corner); the segments chain
// into loops whose normals point from the inside corners out.  Each loop is
// started at its smallest edge ord and fanned; loops in the order of their
// smallest edge.
struct McPolys {
	std::vector<int> tri[256];   // edge ords, three per triangle
};

static int mcEdgeCorner(int i, int end)
{
	const int a = i >> 2;
	int p[3] = {0, 0, 0};
	p[(a + 1) % 3] = i & 1;
	p[(a + 2) % 3] = i >> 1 & 1;
	p[a] = end;
	return p[0] | p[1] << 1 | p[2] << 2;
}

static int mcEdgeBetween(int k0, int k1)
{
	for (int i = 0; i < 12; i++)
		if ((mcEdgeCorner(i, 0) == k0 && mcEdgeCorner(i, 1) == k1) || (mcEdgeCorner(i, 0) == k1 && mcEdgeCorner(i, 1) == k0)) return i;
	return -1;
}

static const McPolys &mcPolys()
{
	static McPolys P;
	static bool built = false;
	if (built) return P;
	for (int type = 0; type < 256; type++) {
		auto in = [&](int k) { return (type >> k & 1) != 0; };
		int next[12];
		for (int i = 0; i < 12; i++) next[i] = -1;
		for (int f = 0; f < 3; f++)
			for (int side = 0; side < 2; side++) {
				// the face's corners counter-clockwise around its outward normal
				const int u = (f + 1) % 3, v = (f + 2) % 3;   // (u, v, f) is right-handed
				int ring[4];
				const int uv[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
				for (int q = 0; q < 4; q++) {
					int p[3] = {0, 0, 0};
					p[f] = side;
					p[u] = uv[q][0];
					p[v] = uv[q][1];
					ring[q] = p[0] | p[1] << 1 | p[2] << 2;
				}
				if (side == 0) std::swap(ring[1], ring[3]);   // outward normal -f: reverse the walk
				// each crossing entered from outside (out -> in along the walk)
				// is followed by the next crossing left to outside
				for (int q = 0; q < 4; q++) {
					const int a = ring[q], b = ring[(q + 1) % 4];
					if (in(a) == in(b) || !in(b)) continue;
					const int from = mcEdgeBetween(a, b);
					for (int r = 1; r < 4; r++) {
						const int c = ring[(q + r) % 4], d = ring[(q + r + 1) % 4];
						if (in(c) && !in(d)) { next[from] = mcEdgeBetween(c, d); break; }
					}
				}
			}
		bool seen[12] = {};
		for (int s0 = 0; s0 < 12; s0++) {
			if (next[s0] < 0 || seen[s0]) continue;
			std::vector<int> loop;
			int cur = s0;
			while (!seen[cur]) { seen[cur] = true; loop.push_back(cur); cur = next[cur]; }
			std::rotate(loop.begin(), std::min_element(loop.begin(), loop.end()), loop.end());
			for (size_t q = 1; q + 1 < loop.size(); q++) {
				P.tri[type].push_back(loop[0]);
				P.tri[type].push_back(loop[q]);
				P.tri[type].push_back(loop[q + 1]);
			}
		}
	}
	built = true;
	return P;
}

struct MeshOut {
	std::vector<double> pos, nrm;
	std::vector<unsigned> idx;
};

// Contourer<2, true>::contour at 0 (Contourer.cpp:9-97, getCellType's
// diagonal rule :85-95): inside nodes in node order, then edge crossings in
// cell order and edge order (2D cell edge ord e: axis e >> 1, offset e & 1
// along the other axis, StaggeredGrid.h:87-94).  A cell's triangles: its
// boundary walked counter-clockwise from node 0 (nodes 0, 1, 3, 2 and the
// edges 0, 3, 1, 2 between them) lists the inside nodes and crossings of
// the inside polygon; the separated diagonal types (6, 9 with the cell
// mean > 0) give each inside node a triangle with the crossings beside it,
// every other type one polygon fanned from the walk's first element.
static void contour2(const int *n, const double *o, double dx, const double *f, MeshOut &M)
{
	const int nx = n[0], ny = n[1];
	M.pos.clear(); M.nrm.clear(); M.idx.clear();
	auto at = [&](int i, int j) { return f[size_t(j) * nx + i]; };
	std::vector<int> node(size_t(nx) * ny, -1);
	std::vector<int> ex(size_t(std::max(nx - 1, 0)) * ny, -1), ey(size_t(nx) * std::max(ny - 1, 0), -1);
	for (int j = 0; j < ny; j++)
		for (int i = 0; i < nx; i++)
			if (at(i, j) <= 0) {
				node[size_t(j) * nx + i] = int(M.pos.size() / 2);
				M.pos.push_back(o[0] + i * dx);
				M.pos.push_back(o[1] + j * dx);
			}
	auto crossing = [&](int i0, int j0, int i1, int j1, int &mk) {
		const double f0 = at(i0, j0), f1 = at(i1, j1);
		if ((f0 <= 0) != (f1 <= 0) && mk < 0) {
			const double theta = (f0 - 0) / (f0 - f1);
			mk = int(M.pos.size() / 2);
			M.pos.push_back((1 - theta) * (o[0] + i0 * dx) + theta * (o[0] + i1 * dx));
			M.pos.push_back((1 - theta) * (o[1] + j0 * dx) + theta * (o[1] + j1 * dx));
		}
		return mk;
	};
	for (int j = 0; j + 1 < ny; j++)
		for (int i = 0; i + 1 < nx; i++) {
			// nodes counter-clockwise: (i,j), (i+1,j), (i+1,j+1), (i,j+1)
			const int ci[4] = {i, i + 1, i + 1, i}, cj[4] = {j, j, j + 1, j + 1};
			bool in[4];
			for (int q = 0; q < 4; q++) in[q] = at(ci[q], cj[q]) <= 0;
			// edges in the reference's order 0..3 first (vertex creation order)
			int e[4];
			e[0] = crossing(i, j, i + 1, j, ex[size_t(j) * (nx - 1) + i]);
			e[1] = crossing(i, j + 1, i + 1, j + 1, ex[size_t(j + 1) * (nx - 1) + i]);
			e[2] = crossing(i, j, i, j + 1, ey[size_t(j) * nx + i]);
			e[3] = crossing(i + 1, j, i + 1, j + 1, ey[size_t(j) * nx + i + 1]);
			// the walk: node q, then the edge to node q + 1
			const int sideEdge[4] = {e[0], e[3], e[1], e[2]};
			std::vector<int> seq;
			std::vector<bool> isNode;
			for (int q = 0; q < 4; q++) {
				if (in[q]) { seq.push_back(node[size_t(cj[q]) * nx + ci[q]]); isNode.push_back(true); }
				if (in[q] != in[(q + 1) % 4]) { seq.push_back(sideEdge[q]); isNode.push_back(false); }
			}
			const bool diag = (in[0] && in[2] && !in[1] && !in[3]) || (in[1] && in[3] && !in[0] && !in[2]);
			const double centre = (at(i, j) + at(i + 1, j) + at(i, j + 1) + at(i + 1, j + 1)) / 4;
			const int m = int(seq.size());
			if (diag && !(centre <= 0)) {
				for (int p = 0; p < m; p++)
					if (isNode[p]) {
						M.idx.push_back(unsigned(seq[(p + m - 1) % m]));
						M.idx.push_back(unsigned(seq[p]));
						M.idx.push_back(unsigned(seq[(p + 1) % m]));
					}
			} else {
				for (int q = 1; q + 1 < m; q++) {
					M.idx.push_back(unsigned(seq[0]));
					M.idx.push_back(unsigned(seq[q]));
					M.idx.push_back(unsigned(seq[q + 1]));
				}
			}
		}
}

static void contour3(const int *n, const double *o, double dx, const double *f, MeshOut &M)
{
	const McPolys &P = mcPolys();
	const int m[3] = {n[0] - 1, n[1] - 1, n[2] - 1};
	auto node = [&](const int *c) { return size_t(c[0]) + size_t(n[0]) * (size_t(c[1]) + size_t(n[1]) * size_t(c[2])); };
	// edgeMark per axis over the edge grids (node grid minus one along the axis)
	std::vector<int> mark[3];
	int en[3][3];
	for (int a = 0; a < 3; a++) {
		for (int q = 0; q < 3; q++) en[a][q] = n[q] - (q == a);
		mark[a].assign(size_t(en[a][0]) * en[a][1] * en[a][2], -1);
	}
	auto edgeIndex = [&](int a, const int *e) { return size_t(e[0]) + size_t(en[a][0]) * (size_t(e[1]) + size_t(en[a][1]) * size_t(e[2])); };
	M.pos.clear(); M.nrm.clear(); M.idx.clear();
	if (m[0] < 1 || m[1] < 1 || m[2] < 1) return;
	int c[3];
	for (c[2] = 0; c[2] < m[2]; c[2]++)
		for (c[1] = 0; c[1] < m[1]; c[1]++)
			for (c[0] = 0; c[0] < m[0]; c[0]++) {
				int type = 0;
				for (int k = 0; k < 8; k++) {
					const int q[3] = {c[0] + (k & 1), c[1] + (k >> 1 & 1), c[2] + (k >> 2 & 1)};
					if (f[node(q)] <= 0) type |= 1 << k;
				}
				int eid[12];
				for (int i = 0; i < 12; i++) {
					const int a = i >> 2;
					int e[3] = {c[0], c[1], c[2]};
					e[(a + 1) % 3] += i & 1;
					e[(a + 2) % 3] += i >> 1 & 1;
					int e1[3] = {e[0], e[1], e[2]};
					e1[a] += 1;
					const bool cross = (f[node(e)] <= 0) != (f[node(e1)] <= 0);
					int &mk = mark[a][edgeIndex(a, e)];
					if (cross && mk < 0) {
						const double f0 = f[node(e)], f1 = f[node(e1)];
						const double theta = (f0 - 0) / (f0 - f1);
						mk = int(M.pos.size() / 3);
						for (int q = 0; q < 3; q++) {
							const double p0 = o[q] + e[q] * dx, p1 = o[q] + e1[q] * dx;
							M.pos.push_back((1 - theta) * p0 + theta * p1);
						}
					}
					eid[i] = mk;
				}
				for (int ord : P.tri[type]) M.idx.push_back(unsigned(eid[ord]));
			}
	// computeNormals: unit face normals summed per vertex in triangle order
	const size_t nv = M.pos.size() / 3;
	M.nrm.assign(nv * 3, 0.0);
	for (size_t t = 0; t < M.idx.size(); t += 3) {
		const double *p0 = &M.pos[3 * M.idx[t]], *p1 = &M.pos[3 * M.idx[t + 1]], *p2 = &M.pos[3 * M.idx[t + 2]];
		const double a[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
		const double b[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
		double nn[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
		const double s2 = nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2];
		if (s2 > 0) { const double r = std::sqrt(s2); for (double &x : nn) x /= r; }
		for (int j = 0; j < 3; j++)
			for (int q = 0; q < 3; q++) M.nrm[3 * M.idx[t + j] + q] += nn[q];
	}
	for (size_t v = 0; v < nv; v++) {
		double *x = &M.nrm[3 * v];
		const double s2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
		if (s2 > 0) { const double r = std::sqrt(s2); for (int q = 0; q < 3; q++) x[q] /= r; }
	}
}

} // namespace ora

// ------------------------------------------------------------------ C API --
namespace {

struct Handle {
	int dim;
	ora::Liquid<2> *l2 = nullptr;
	ora::Liquid<3> *l3 = nullptr;
};

template <int D>
size_t fieldCount(const ora::Liquid<D> &L, int f)
{
	const size_t N = L.cell.count();
	size_t F = 0;
	for (int a = 0; a < D; a++) F += L.face[a].count();
	switch (f) {
	case 0: case 4: case 5: case 6: case 8: case 10: case 15: return N;
	case 1: case 2: case 7: return N * D;
	case 3: return N * D * D;
	case 9: case 12: case 13: case 14: return F;
	case 11: return L.node.count();
	default: return 0;
	}
}

template <int D>
int access(ora::Liquid<D> &L, int f, void *data, size_t n, bool set)
{
	if (n != fieldCount(L, f) || n == 0) return -1;
	auto cp = [&](std::vector<double> &v) {
		if (set) std::memcpy(v.data(), data, n * sizeof(double));
		else std::memcpy(data, v.data(), n * sizeof(double));
	};
	auto cpFaces = [&](std::vector<double> *v) {
		size_t off = 0;
		for (int a = 0; a < D; a++) {
			const size_t m = v[a].size();
			if (set) std::memcpy(v[a].data(), (double *)data + off, m * sizeof(double));
			else std::memcpy((double *)data + off, v[a].data(), m * sizeof(double));
			off += m;
		}
	};
	switch (f) {
	case 0: cp(L.phi); break;
	case 1: cp(L.phiGrad); break;
	case 2: cp(L.xi); break;
	case 3: cp(L.J); break;
	case 4: cp(L.phiTmp); break;
	case 5: cp(L.phiOri); break;
	case 6: cp(L.phiRef); break;
	case 7: cp(L.phiRefGrad); if (set) L.refView.build(L.cell, L.phiRefGrad.data()); break;
	case 8:
		if (set) std::memcpy(L.advRegion.data(), data, n);
		else std::memcpy(data, L.advRegion.data(), n);
		break;
	case 9: cpFaces(L.u); break;
	case 10: cp(L.cellDist); break;
	case 11: cp(L.nodeDist); break;
	case 12: cpFaces(L.frac); break;
	case 13: cpFaces(L.bvel); break;
	case 14: cpFaces(L.bnormal); break;
	case 15: cp(L.pres); break;
	default: return -1;
	}
	return 0;
}

#define DISPATCH(h, expr) \
	do { if ((h)->dim == 2) { auto &L = *(h)->l2; expr; } else { auto &L = *(h)->l3; expr; } } while (0)

} // namespace

extern "C" {

void *ora_create(int dim, const int *res, double dx, const double *origin, int bw)
{
	if (dim != 2 && dim != 3) return nullptr;
	auto *h = new Handle;
	h->dim = dim;
	if (dim == 2) { h->l2 = new ora::Liquid<2>; h->l2->setup(res, dx, origin, bw); }
	else { h->l3 = new ora::Liquid<3>; h->l3->setup(res, dx, origin, bw); }
	return h;
}

void ora_destroy(void *p)
{
	auto *h = (Handle *)p;
	delete h->l2; delete h->l3; delete h;
}

void ora_set_threads(int n)
{
#ifdef _OPENMP
	if (n > 0) omp_set_num_threads(n);
#endif
}

int ora_get_threads(void)
{
#ifdef _OPENMP
	return omp_get_max_threads();
#else
	return 1;
#endif
}

size_t ora_field_count(void *p, int f) { auto *h = (Handle *)p; size_t r = 0; DISPATCH(h, r = fieldCount(L, f)); return r; }
int ora_set(void *p, int f, const void *data, size_t n) { auto *h = (Handle *)p; int r = 0; DISPATCH(h, r = access(L, f, (void *)data, n, true)); return r; }
int ora_get(void *p, int f, void *data, size_t n) { auto *h = (Handle *)p; int r = 0; DISPATCH(h, r = access(L, f, data, n, false)); return r; }

} // extern "C"

template <int D>
static void unionBox(ora::Liquid<D> &L, const double *mn, const double *len)
{
	double center[3], half[3];
	for (int a = 0; a < D; a++) { center[a] = mn[a] + len[a] / 2; half[a] = len[a] / 2; }   // ImplicitSurface.h:37
	L.unionLevelSet([&](const double *x) { return ora::boxSdf<D>(center, half, x); });
}

// ImplicitSphere::signedDistance (ImplicitSurface.h:23)
template <int D>
static void unionSphere(ora::Liquid<D> &L, const double *c, double r)
{
	L.unionLevelSet([&](const double *x) {
		double s = 0;
		for (int a = 0; a < D; a++) s += (x[a] - c[a]) * (x[a] - c[a]);
		return std::sqrt(s) - r;
	});
}

// ImplicitPlane::signedDistance (ImplicitSurface.h:132), normal normalised in the ctor
template <int D>
static void unionPlane(ora::Liquid<D> &L, const double *pos, const double *nrm)
{
	double n[3], s = 0;
	for (int a = 0; a < D; a++) s += nrm[a] * nrm[a];
	s = std::sqrt(s);
	for (int a = 0; a < D; a++) n[a] = s > 0 ? nrm[a] / s : nrm[a];
	L.unionLevelSet([&](const double *x) {
		double d = 0;
		for (int a = 0; a < D; a++) d += (x[a] - pos[a]) * n[a];
		return d;
	});
}

extern "C" {

void ora_union_box(void *p, const double *mn, const double *len) { auto *h = (Handle *)p; DISPATCH(h, unionBox(L, mn, len)); }
void ora_union_sphere(void *p, const double *c, double r) { auto *h = (Handle *)p; DISPATCH(h, unionSphere(L, c, r)); }
void ora_union_plane(void *p, const double *pos, const double *n) { auto *h = (Handle *)p; DISPATCH(h, unionPlane(L, pos, n)); }
void ora_solid_plane(void *p, const double *pos, const double *n) { auto *h = (Handle *)p; DISPATCH(h, L.solidPlane(pos, n)); }

int ora_initialize(void *p, int reinitIters, double cfl, int precond, double tol, int maxIt)
{
	auto *h = (Handle *)p; int r = 0;
	DISPATCH(h, r = L.initialize(reinitIters, cfl, precond, tol, maxIt));
	return r;
}
void ora_advect_refmap(void *p, double dt) { auto *h = (Handle *)p; DISPATCH(h, L.advectRefMap(dt)); }
void ora_advect_velocity(void *p, double dt) { auto *h = (Handle *)p; DISPATCH(h, L.advectVelocity(dt)); }
void ora_compose(void *p) { auto *h = (Handle *)p; DISPATCH(h, L.compose()); }
void ora_reinit(void *p, int iters, double cfl) { auto *h = (Handle *)p; DISPATCH(h, L.reinit(iters, cfl)); }
void ora_extrapolate_collider(void *p) { auto *h = (Handle *)p; DISPATCH(h, L.extrapolateCollider()); }
void ora_extrapolate_air(void *p) { auto *h = (Handle *)p; DISPATCH(h, L.extrapolateAir()); }
void ora_enforce_boundary(void *p) { auto *h = (Handle *)p; DISPATCH(h, L.enforceBoundary()); }
void ora_gravity(void *p, double dt) { auto *h = (Handle *)p; DISPATCH(h, L.gravity(dt)); }
int ora_check_refmap(void *p, double th, double band, double ratio, long long *tot, long long *cnt)
{
	auto *h = (Handle *)p; int r = 0;
	DISPATCH(h, r = L.checkRefMap(th, band, ratio, tot, cnt) ? 1 : 0);
	return r;
}
void ora_restart(void *p) { auto *h = (Handle *)p; DISPATCH(h, L.restart()); }
void ora_restart_newton(void *p) { auto *h = (Handle *)p; DISPATCH(h, L.restartNewton()); }
void ora_set_restart_newton(void *p, int on) { auto *h = (Handle *)p; DISPATCH(h, L.newtonRestart = on != 0); }
void ora_set_surface_tension(void *p, int on, double coeff, double dmin, double dmax, double crit, double tol, int maxIt, int littleDrop)
{
	auto *h = (Handle *)p;
	DISPATCH(h, (L.stOn = on != 0, L.stCoeff = coeff, L.stDeltaMin = dmin, L.stDeltaMax = dmax, L.stCritTime = crit, L.stTol = tol,
		L.stMaxIt = maxIt, L.stLittleDrop = littleDrop != 0));
}
void ora_set_surface_tension_solvers(void *p, int capillaryPrecond, int capillaryGuess, int reprojectWarm)
{
	auto *h = (Handle *)p;
	DISPATCH(h, (L.stPrecond = capillaryPrecond, L.stGuess = capillaryGuess != 0, L.stWarm = reprojectWarm != 0));
}
int ora_apply_surface_tension(void *p, double dt, int precond, double tol, int maxIt, int *iters, double *err, int *unknowns)
{
	auto *h = (Handle *)p;
	int r = 0;
	DISPATCH(h, r = L.surfaceTension(dt, precond, tol, maxIt, iters, err, unknowns));
	return r;
}
// last capillary solve: faces, iterations, relative residual, band
void ora_capillary_stats(void *p, int *faces, int *iters, double *err, double *band)
{
	auto *h = (Handle *)p;
	DISPATCH(h, (*faces = L.stFaces, *iters = L.stIters, *err = L.stErr, *band = L.stBand));
}
int ora_capillary(void *p, double coeff, double band, double dt, double tol, int maxIt, int precond)
{
	auto *h = (Handle *)p;
	int r = 0;
	DISPATCH(h, r = L.capillary(coeff, band, dt, tol, maxIt, precond));
	return r;
}
double ora_curvature(void *p, const double *pos) { auto *h = (Handle *)p; double r = 0; DISPATCH(h, r = L.curvature(pos)); return r; }
void ora_little_drop(void *p, double dt) { auto *h = (Handle *)p; DISPATCH(h, L.littleDrop(dt)); }
void ora_set_elapsed_time(void *p, double t) { auto *h = (Handle *)p; DISPATCH(h, L.elapsed = t); }
double ora_get_elapsed_time(void *p) { auto *h = (Handle *)p; double r = 0; DISPATCH(h, r = L.elapsed); return r; }
int ora_project(void *p, int precond, double tol, int maxIt, int *iters, double *err, int *unknowns)
{
	auto *h = (Handle *)p; int r = 0;
	DISPATCH(h, r = L.project(precond, tol, maxIt, iters, err, unknowns));
	return r;
}
int ora_step(void *p, double dt, int doReinit, int reinitIters, double cfl, int precond, double tol, int maxIt,
	int *restarted, int *iters, double *err, int *unknowns)
{
	auto *h = (Handle *)p; int r = 0;
	DISPATCH(h, r = L.step(dt, doReinit, reinitIters, cfl, precond, tol, maxIt, restarted, iters, err, unknowns));
	return r;
}
double ora_max_speed(void *p) { auto *h = (Handle *)p; double r = 0; DISPATCH(h, r = L.maxSpeed()); return r; }


double ora_mc_fraction(const double *c);
double ora_filled_area(const double *p);

// LevelSetLiquid::calculateVolume (LevelSetLiquid.cpp:1162-1190) on phi_tmp,
// summed serially in the reference's x-fastest cell order.  corrected = 0
// passes the lexicographic stencil as the reference does; 1 reorders it to
// the fraction's circular corner order first.
// Contourer<3>::contour of a node field (n[0] x n[1] x n[2], x fastest,
// origin o, spacing dx): counts always; arrays (float positions and normals,
// 3 per vertex; uint indices) when the capacities suffice.  Returns 0, or -1
// when a capacity is short.
// pos64 (optional): the FP64 positions SurfaceMesh::writeOBJ prints.
int ora_contour3(const int *n, const double *o, double dx, const double *f, float *pos, float *nrm, size_t vcap, unsigned *idx,
	size_t icap, size_t *nv, size_t *ni, double *pos64)
{
	ora::MeshOut M;
	ora::contour3(n, o, dx, f, M);
	*nv = M.pos.size() / 3;
	*ni = M.idx.size();
	if (*nv > vcap || *ni > icap) return -1;
	for (size_t i = 0; i < M.pos.size(); i++) { pos[i] = float(M.pos[i]); nrm[i] = float(M.nrm[i]); }
	if (pos64) std::copy(M.pos.begin(), M.pos.end(), pos64);
	std::copy(M.idx.begin(), M.idx.end(), idx);
	return 0;
}
// Contourer<2, true>::contour of a node field (n[0] x n[1], x fastest):
// FP64 positions (2 per vertex) and uint indices; counts always, arrays when
// the capacities suffice (returns -1 otherwise).
int ora_contour2(const int *n, const double *o, double dx, const double *f, double *pos, size_t vcap, unsigned *idx, size_t icap,
	size_t *nv, size_t *ni)
{
	ora::MeshOut M;
	ora::contour2(n, o, dx, f, M);
	*nv = M.pos.size() / 2;
	*ni = M.idx.size();
	if (*nv > vcap || *ni > icap) return -1;
	std::copy(M.pos.begin(), M.pos.end(), pos);
	std::copy(M.idx.begin(), M.idx.end(), idx);
	return 0;
}
// the triangle table: ntri[type] and up to 8 triangles (24 edge ords) per type
int ora_mc_table(int *ntri, int *tri)
{
	const ora::McPolys &P = ora::mcPolys();
	for (int t = 0; t < 256; t++) {
		const size_t k = P.tri[t].size();
		if (k > 24) return -1;
		ntri[t] = int(k / 3);
		for (int q = 0; q < 24; q++) tri[t * 24 + q] = q < int(k) ? P.tri[t][q] : -1;
	}
	return 0;
}

double ora_volume(void *p, int corrected)
{
	auto *h = (Handle *)p;
	double res = 0.0;
	if (h->dim == 2) {
		const auto &L = *h->l2;
		const int nx = L.res[0], ny = L.res[1];
		for (int j = 0; j + 1 < ny; j++)
			for (int i = 0; i + 1 < nx; i++) {
				const size_t b = size_t(j) * nx + i;
				const double v[4] = {L.phiTmp[b], L.phiTmp[b + 1], L.phiTmp[b + nx], L.phiTmp[b + nx + 1]};
				res += ora_filled_area(v) * L.dx * L.dx;
			}
		return res;
	}
	const auto &L = *h->l3;
	const int nx = L.res[0], ny = L.res[1], nz = L.res[2];
	const double dx3 = L.dx * L.dx * L.dx;
	const size_t sy = size_t(nx), sz = size_t(nx) * ny;
	for (int k = 0; k + 1 < nz; k++)
		for (int j = 0; j + 1 < ny; j++)
			for (int i = 0; i + 1 < nx; i++) {
				const size_t b = (size_t(k) * ny + j) * nx + i;
				const double *f = L.phiTmp.data();
				const double lex[8] = {f[b], f[b + 1], f[b + sy], f[b + sy + 1], f[b + sz], f[b + sz + 1], f[b + sz + sy], f[b + sz + sy + 1]};
				if (corrected) {
					const double c[8] = {lex[0], lex[1], lex[3], lex[2], lex[4], lex[5], lex[7], lex[6]};
					res += ora_mc_fraction(c) * dx3;
				} else {
					res += ora_mc_fraction(lex) * dx3;
				}
			}
	return res;
}

} // extern "C"
