// Gradient-augmented reference-map advection, semi-Lagrangian velocity
// advection and the phi / grad-phi composition for gfx950.
//
// One thread per cell (or face); the velocity gradient and the Hermite mixed
// derivatives of the reference map are formed on the fly from the clamped
// neighbours instead of being materialised as the reference does
// (LevelSetLiquid.cpp:739-742, HermiteIntrpl.h:17-30), so each launch reads
// u, xi and J once from HBM and writes xi', J' once.
#include "garm_internal.hpp"

#include <type_traits>

namespace garm {

// ---------------------------------------------------------------- velocity
// VectorFieldWrapper<SGrid, Linear>::operator() (VectorField.h:66-72):
// per axis, the multilinear sample of u_a on face grid a (LagrangeIntrpl.h:40-72).
template <int D, typename T>
__device__ __forceinline__ void sampleVel(const FaceSet &U, const T *p, T *v)
{
#pragma unroll
	for (int a = 0; a < D; a++) {
		const GridDev &g = U.g[a];
		const double *ua = U.a[a];
		int l[3]; T f[3];
		lowerFracT<D, T>(g, p, l, f);
		T w[2][3];
#pragma unroll
		for (int b = 0; b < D; b++) { w[0][b] = 1 - f[b]; w[1][b] = f[b]; }
		T val = 0;
#pragma unroll
		for (int k = 0; k < (D == 3 ? 2 : 1); k++)
#pragma unroll
			for (int j = 0; j < 2; j++)
#pragma unroll
				for (int i = 0; i < 2; i++) {
					const T wt = D == 2 ? w[i][0] * w[j][1] : w[i][0] * w[j][1] * w[k][2];
					val += T(ua[gat<D>(g, l[0] + i, l[1] + j, D == 3 ? l[2] + k : 0)]) * wt;
				}
		v[a] = val;
	}
}

// DrvVectorFieldWrapper<SGrid, Linear> (VectorField.h:106-138) with the
// face-gradient of Derivatives::computeFirst<2> (Derivatives.h:47,51-57,98-105)
// evaluated at each clamped corner: row a of G is grad u_a.  G row-major.
template <int D, typename T>
__device__ __forceinline__ void sampleVelGrad(const FaceSet &U, const T *p, T *v, T *G)
{
#pragma unroll
	for (int a = 0; a < D; a++) {
		const GridDev &g = U.g[a];
		const double *ua = U.a[a];
		int l[3]; T f[3];
		lowerFracT<D, T>(g, p, l, f);
		T w[2][3];
#pragma unroll
		for (int b = 0; b < D; b++) { w[0][b] = 1 - f[b]; w[1][b] = f[b]; }
		const T inv = T(g.inv);
		T val = 0, gr[3] = {0, 0, 0};
#pragma unroll
		for (int k = 0; k < (D == 3 ? 2 : 1); k++)
#pragma unroll
			for (int j = 0; j < 2; j++)
#pragma unroll
				for (int i = 0; i < 2; i++) {
					const T wt = D == 2 ? w[i][0] * w[j][1] : w[i][0] * w[j][1] * w[k][2];
					const int qi = clampi(l[0] + i, g.n[0] - 1), qj = clampi(l[1] + j, g.n[1] - 1);
					const int qk = D == 3 ? clampi(l[2] + k, g.n[2] - 1) : 0;
					val += T(ua[gidx<D>(g, qi, qj, qk)]) * wt;
					// neighbour differences taken on the FP64 data, then rounded
					gr[0] += (T(ua[gat<D>(g, qi + 1, qj, qk)] - ua[gat<D>(g, qi - 1, qj, qk)]) / 2 * inv) * wt;
					gr[1] += (T(ua[gat<D>(g, qi, qj + 1, qk)] - ua[gat<D>(g, qi, qj - 1, qk)]) / 2 * inv) * wt;
					if constexpr (D == 3)
						gr[2] += (T(ua[gat<D>(g, qi, qj, qk + 1)] - ua[gat<D>(g, qi, qj, qk - 1)]) / 2 * inv) * wt;
				}
		v[a] = val;
#pragma unroll
		for (int b = 0; b < D; b++) G[a * D + b] = gr[b];
	}
}

// ----------------------------------------------------------------- Hermite
// 1-D cubic Hermite basis (HermiteIntrpl.h:89-125) for one axis:
// B[vtx][ord] value weights, Bg[vtx][ord] their derivatives.
template <typename T>
__device__ __forceinline__ void hermiteBasis(T x, T B[2][2], T Bg[2][2])
{
	B[0][0] = (x + x + 1) * (x - 1) * (x - 1); Bg[0][0] = 6 * x * (x - 1);
	B[1][0] = x * x * (3 - x - x);             Bg[1][0] = 6 * x * (1 - x);
	B[0][1] = x * (x - 1) * (x - 1);           Bg[0][1] = (3 * x - 1) * (x - 1);
	B[1][1] = x * x * (x - 1);                 Bg[1][1] = x * (3 * x - 2);
}

// HermiteIntrpl<D, VectorDd>::interpolate + interpolateGradient
// (HermiteIntrpl.h:32-58, getValues :62-87) of the reference map with the
// mixed derivatives of its constructor (:17-30) formed from J at the clamped
// neighbours of each clamped corner.  Gh is column-major: Gh[c*D + r].
// In float the map is interpolated as the displacement xi - x: the data at a
// corner become xi - x_corner (x_corner of the unclamped corner) and J - I,
// taken in FP64 before rounding, and X and I are added back.  Cubic Hermite
// reproduces the identity map exactly, so this is the same interpolant
// without the cancellation of positions of size 1 in gradients of size 1/dx.
template <int D, typename T>
__device__ __forceinline__ void hermiteRefMap(const GridDev &g, const double *const *xi, const double *const *J,
	const T *X, T *val, T *Gh)
{
	int l[3] = {0, 0, 0}; T f[3];
	lowerFracT<D, T>(g, X, l, f);
	T B[3][2][2], Bg[3][2][2];
#pragma unroll
	for (int a = 0; a < D; a++) hermiteBasis(f[a], B[a], Bg[a]);
#pragma unroll
	for (int r = 0; r < D; r++) val[r] = 0;
#pragma unroll
	for (int e = 0; e < D * D; e++) Gh[e] = 0;
	const T dx = T(g.dx), inv = T(g.inv);
	constexpr bool kDisp = !std::is_same<T, double>::value;

#pragma unroll
	for (int cn = 0; cn < (1 << D); cn++) {
		const int vx = cn & 1, vy = cn >> 1 & 1, vz = D == 3 ? cn >> 2 : 0;
		const int qi = clampi(l[0] + vx, g.n[0] - 1), qj = clampi(l[1] + vy, g.n[1] - 1);
		const int qk = D == 3 ? clampi(l[2] + vz, g.n[2] - 1) : 0;
		const size_t q = gidx<D>(g, qi, qj, qk);
#pragma unroll
		for (int ord = 0; ord < (1 << D); ord++) {
			const int ox = ord & 1, oy = ord >> 1 & 1, oz = D == 3 ? ord >> 2 : 0;
			T wt = B[0][vx][ox] * B[1][vy][oy];
			T wg[3];
			if constexpr (D == 2) {
				wg[0] = Bg[0][vx][ox] * B[1][vy][oy];
				wg[1] = B[0][vx][ox] * Bg[1][vy][oy];
			} else {
				wt = wt * B[2][vz][oz];
				wg[0] = Bg[0][vx][ox] * B[1][vy][oy] * B[2][vz][oz];
				wg[1] = B[0][vx][ox] * Bg[1][vy][oy] * B[2][vz][oz];
				wg[2] = B[0][vx][ox] * B[1][vy][oy] * Bg[2][vz][oz];
			}
#pragma unroll
			for (int r = 0; r < D; r++) {
				T v;
				// getValues ordering: 0 value, 1 d/dx, 2 d/dy, 3 dxy, 4 d/dz, 5 dxz, 6 dyz, 7 dxyz
				if constexpr (kDisp) {
					const int lv[3] = {l[0] + vx, l[1] + vy, l[2] + vz};
					if (ord == 0) v = T(xi[r][q] - (g.o[r] + lv[r] * g.dx));
					else if (ord == 1) v = T(J[0 * D + r][q] - (r == 0)) * dx;
					else if (ord == 2) v = T(J[1 * D + r][q] - (r == 1)) * dx;
					else if (ord == 4) v = T(J[2 * D + r][q] - (r == 2)) * dx;
				} else {
					if (ord == 0) v = T(xi[r][q]);
					else if (ord == 1) v = T(J[0 * D + r][q]) * dx;
					else if (ord == 2) v = T(J[1 * D + r][q]) * dx;
					else if (ord == 4) v = T(J[2 * D + r][q]) * dx;
				}
				if (ord == 3) {
					const T m = T(J[1 * D + r][gat<D>(g, qi + 1, qj, qk)] - J[1 * D + r][gat<D>(g, qi - 1, qj, qk)]) / 2 * inv;
					v = m * dx * dx;
				} else if (ord == 5) {
					const T m = T(J[2 * D + r][gat<D>(g, qi + 1, qj, qk)] - J[2 * D + r][gat<D>(g, qi - 1, qj, qk)]) / 2 * inv;
					v = m * dx * dx;
				} else if (ord == 6) {
					const T m = T(J[2 * D + r][gat<D>(g, qi, qj + 1, qk)] - J[2 * D + r][gat<D>(g, qi, qj - 1, qk)]) / 2 * inv;
					v = m * dx * dx;
				} else if (ord == 7) {
					const double *c2 = J[2 * D + r];
					const T m = T(c2[gat<D>(g, qi + 1, qj + 1, qk)] - c2[gat<D>(g, qi - 1, qj + 1, qk)]
						- c2[gat<D>(g, qi + 1, qj - 1, qk)] + c2[gat<D>(g, qi - 1, qj - 1, qk)]) / 4 * inv * inv;
					v = m * dx * dx * dx;
				}
				val[r] += v * wt;
#pragma unroll
				for (int c = 0; c < D; c++) Gh[c * D + r] += v * wg[c];
			}
		}
	}
#pragma unroll
	for (int e = 0; e < D * D; e++) Gh[e] *= inv;
	if constexpr (kDisp) {
#pragma unroll
		for (int r = 0; r < D; r++) { val[r] += X[r]; Gh[r * D + r] += T(1); }
	}
}

// ------------------------------------------------------------------ kernels
template <int D>
struct RefMapArgs {
	GridDev cell;
	FaceSet U;
	const double *xi[3];
	const double *J[9];
	double *xi_new[3];
	double *J_new[9];
	double h;   // -dt
};

// SemiLagrangian<D, 2, HermiteIntrpl<D, VectorDd>>::advect (drv overload,
// Advection.h:147-160) with Advection::traceDrv<2> (Advection.h:71-82).
template <int D, typename T>
__global__ __launch_bounds__(kBlock) void kAdvectRefMap(RefMapArgs<D> A)
{
	const Coord c = threadCoord();
	if (!inGrid<D>(A.cell, c)) return;
	const size_t id = gidx<D>(A.cell, c.i, c.j, c.k);
	const T h = T(A.h);
	T x[3] = {T(A.cell.o[0] + c.i * A.cell.dx), T(A.cell.o[1] + c.j * A.cell.dx), T(A.cell.o[2] + c.k * A.cell.dx)};

	T v0[3], G0[9], p1[3], R1[9], v1[3], G1[9], X[3], R[9];
	sampleVelGrad<D, T>(A.U, x, v0, G0);
#pragma unroll
	for (int a = 0; a < D; a++) p1[a] = x[a] + v0[a] * h;
#pragma unroll
	for (int a = 0; a < D; a++)
#pragma unroll
		for (int b = 0; b < D; b++) R1[a * D + b] = (a == b ? T(1) : T(0)) + G0[a * D + b] * h;
	sampleVelGrad<D, T>(A.U, p1, v1, G1);
#pragma unroll
	for (int a = 0; a < D; a++) X[a] = x[a] + (v0[a] + v1[a]) * h / 2;
#pragma unroll
	for (int a = 0; a < D; a++)
#pragma unroll
		for (int b = 0; b < D; b++) {
			T gr = 0;
#pragma unroll
			for (int k = 0; k < D; k++) gr += G1[a * D + k] * R1[k * D + b];
			R[a * D + b] = (a == b ? T(1) : T(0)) + (G0[a * D + b] + gr) * h / 2;
		}

	T val[3], Gh[9];
	hermiteRefMap<D, T>(A.cell, A.xi, A.J, X, val, Gh);
#pragma unroll
	for (int r = 0; r < D; r++) A.xi_new[r][id] = val[r];
#pragma unroll
	for (int r = 0; r < D; r++)
#pragma unroll
		for (int col = 0; col < D; col++) {
			T s = 0;
#pragma unroll
			for (int k = 0; k < D; k++) s += Gh[k * D + r] * R[k * D + col];
			A.J_new[col * D + r][id] = s;
		}
}

// SemiLagrangian<D, 2, CubicIntrpl<D>>::advect on face grid `axis`
// (Advection.h:15-22, 37-43, 137-145; CubicIntrpl LagrangeIntrpl.h:128-171).
template <int D, typename T>
__global__ __launch_bounds__(kBlock) void kAdvectVelocity(FaceSet U, int axis, double *out, double hd)
{
	const GridDev &g = U.g[axis];
	const Coord c = threadCoord();
	if (!inGrid<D>(g, c)) return;
	const T h = T(hd);
	T x[3] = {T(g.o[0] + c.i * g.dx), T(g.o[1] + c.j * g.dx), T(g.o[2] + c.k * g.dx)};
	T v0[3], p1[3], v1[3], X[3];
	sampleVel<D, T>(U, x, v0);
#pragma unroll
	for (int a = 0; a < D; a++) p1[a] = x[a] + v0[a] * h;
	sampleVel<D, T>(U, p1, v1);
#pragma unroll
	for (int a = 0; a < D; a++) X[a] = x[a] + (v0[a] + v1[a]) * h / 2;

	int l[3]; T f[3];
#pragma unroll
	for (int a = 0; a < D; a++) {
		const T o = T(g.o[a]), inv = T(g.inv);
		l[a] = int(floor((X[a] - o) * inv)) - 1;   // getCubicLower (Grid.h:54)
		f[a] = (X[a] - o - l[a] * T(g.dx)) * inv;
	}
	T w[4][3];
#pragma unroll
	for (int a = 0; a < D; a++) {
		const T t = f[a];
		w[0][a] = (t - 1) * (t - 2) * (t - 3) / -6;
		w[1][a] = t * (t - 2) * (t - 3) / 2;
		w[2][a] = t * (t - 1) * (t - 3) / -2;
		w[3][a] = t * (t - 1) * (t - 2) / 6;
	}
	const double *ua = U.a[axis];
	T val = 0;
#pragma unroll
	for (int k = 0; k < (D == 3 ? 4 : 1); k++)
#pragma unroll
		for (int j = 0; j < 4; j++)
#pragma unroll
			for (int i = 0; i < 4; i++) {
				const T wt = D == 2 ? w[i][0] * w[j][1] : w[i][0] * w[j][1] * w[k][2];
				val += T(ua[gat<D>(g, l[0] + i, l[1] + j, D == 3 ? l[2] + k : 0)]) * wt;
			}
	out[gidx<D>(g, c.i, c.j, c.k)] = val;
}

template <int D>
struct ComposeArgs {
	GridDev cell;
	const double *xi[3];
	const double *J[9];
	const double *phi_ref;
	const double *phi_ref_grad[3];
	const double *mixed[4];
	const uint8_t *adv;
	const double *cell_dist;
	double *phi, *phi_grad[3], *phi_tmp, *phi_ori;
	double band;   // 15 dx (LevelSetLiquid.h:93)
	double tol;    // 3 dx (LevelSetLiquid.cpp:1234)
};

// generateCurrentFields (LevelSetLiquid.cpp:1218-1244): phi = H[phi_ref](xi),
// grad phi = J^T grad H[phi_ref](xi) (DrvScalarFieldWrapper with
// HermiteIntrpl<D, double>, ScalarField.h:50-69), air clamp,
// enforceBoundaryLevelSet at 3 dx and 0 (:966-975), phi_tmp, phi_ori.
template <int D, typename T>
__global__ __launch_bounds__(kBlock) void kCompose(ComposeArgs<D> A)
{
	const GridDev &g = A.cell;
	const Coord c = threadCoord();
	if (!inGrid<D>(g, c)) return;
	const size_t id = gidx<D>(g, c.i, c.j, c.k);
	T X[3];
#pragma unroll
	for (int a = 0; a < D; a++) X[a] = T(A.xi[a][id]);
	int l[3]; T f[3];
	lowerFracT<D, T>(g, X, l, f);
	T B[3][2][2], Bg[3][2][2];
#pragma unroll
	for (int a = 0; a < D; a++) hermiteBasis(f[a], B[a], Bg[a]);
	const T dx = T(g.dx);
	T val = 0, gr[3] = {0, 0, 0};
#pragma unroll
	for (int cn = 0; cn < (1 << D); cn++) {
		const int vx = cn & 1, vy = cn >> 1 & 1, vz = D == 3 ? cn >> 2 : 0;
		const size_t q = gat<D>(g, l[0] + vx, l[1] + vy, D == 3 ? l[2] + vz : 0);
		T v[8];
		v[0] = T(A.phi_ref[q]);
		v[1] = T(A.phi_ref_grad[0][q]) * dx;
		v[2] = T(A.phi_ref_grad[1][q]) * dx;
		v[3] = T(A.mixed[0][q]) * dx * dx;
		if constexpr (D == 3) {
			v[4] = T(A.phi_ref_grad[2][q]) * dx;
			v[5] = T(A.mixed[1][q]) * dx * dx;
			v[6] = T(A.mixed[2][q]) * dx * dx;
			v[7] = T(A.mixed[3][q]) * dx * dx * dx;
		}
#pragma unroll
		for (int ord = 0; ord < (1 << D); ord++) {
			const int ox = ord & 1, oy = ord >> 1 & 1, oz = D == 3 ? ord >> 2 : 0;
			T wt = B[0][vx][ox] * B[1][vy][oy];
			T wg[3];
			if constexpr (D == 2) {
				wg[0] = Bg[0][vx][ox] * B[1][vy][oy];
				wg[1] = B[0][vx][ox] * Bg[1][vy][oy];
			} else {
				wt = wt * B[2][vz][oz];
				wg[0] = Bg[0][vx][ox] * B[1][vy][oy] * B[2][vz][oz];
				wg[1] = B[0][vx][ox] * Bg[1][vy][oy] * B[2][vz][oz];
				wg[2] = B[0][vx][ox] * B[1][vy][oy] * Bg[2][vz][oz];
			}
			val += v[ord] * wt;
#pragma unroll
			for (int a = 0; a < D; a++) gr[a] += v[ord] * wg[a];
		}
	}
#pragma unroll
	for (int a = 0; a < D; a++) gr[a] *= T(g.inv);
#pragma unroll
	for (int col = 0; col < D; col++) {
		T s = 0;
#pragma unroll
		for (int r = 0; r < D; r++) s += T(A.J[col * D + r][id]) * gr[r];
		A.phi_grad[col][id] = s;
	}
	double phi = val;
	if (A.adv[id] == 0 && phi < A.band) phi = A.band;
	const double cd = A.cell_dist[id];
	if (phi < -cd - A.tol) phi = -cd - A.tol;
	A.phi[id] = phi;
	A.phi_ori[id] = phi;
	A.phi_tmp[id] = phi < -cd - 0.0 ? -cd - 0.0 : phi;
}

// HermiteIntrpl<D, double> constructor (HermiteIntrpl.h:17-30) for the
// reference level set; rebuilt whenever _refLevelSetGrad changes.
template <int D>
__global__ __launch_bounds__(kBlock) void kScalarMixed(GridDev g, const double *g0, const double *g1, const double *g2,
	double *m0, double *m1, double *m2, double *m3)
{
	const Coord c = threadCoord();
	if (!inGrid<D>(g, c)) return;
	const size_t id = gidx<D>(g, c.i, c.j, c.k);
	const double inv = g.inv;
	m0[id] = (g1[gat<D>(g, c.i + 1, c.j, c.k)] - g1[gat<D>(g, c.i - 1, c.j, c.k)]) / 2 * inv;
	if constexpr (D == 3) {
		m1[id] = (g2[gat<D>(g, c.i + 1, c.j, c.k)] - g2[gat<D>(g, c.i - 1, c.j, c.k)]) / 2 * inv;
		m2[id] = (g2[gat<D>(g, c.i, c.j + 1, c.k)] - g2[gat<D>(g, c.i, c.j - 1, c.k)]) / 2 * inv;
		m3[id] = (g2[gat<D>(g, c.i + 1, c.j + 1, c.k)] - g2[gat<D>(g, c.i - 1, c.j + 1, c.k)]
			- g2[gat<D>(g, c.i + 1, c.j - 1, c.k)] + g2[gat<D>(g, c.i - 1, c.j - 1, c.k)]) / 4 * inv * inv;
	}
}

// Derivatives::computeFirst<4> (Derivatives.h:48, 51-57, 98-105).
template <int D>
__global__ __launch_bounds__(kBlock) void kGradient4(GridDev g, const double *src, double *d0, double *d1, double *d2)
{
	const Coord c = threadCoord();
	if (!inGrid<D>(g, c)) return;
	const size_t id = gidx<D>(g, c.i, c.j, c.k);
	double *dst[3] = {d0, d1, d2};
#pragma unroll
	for (int a = 0; a < D; a++) {
		double v[5];
#pragma unroll
		for (int t = 0; t < 5; t++) {
			const int o = t - 2;
			v[t] = src[gat<D>(g, c.i + (a == 0 ? o : 0), c.j + (a == 1 ? o : 0), c.k + (a == 2 ? o : 0))];
		}
		dst[a][id] = (v[0] - 8 * v[1] + 8 * v[3] - v[4]) / 12 * g.inv;
	}
}

// ---------------------------------------------------------------- launchers
// This file is compiled twice: once without fused multiply-add contraction
// for the FP64 kernels (the IEEE operations of the oracle, so the plain path
// reproduces it bit for bit) and once, with GARM_ADVECT_FP32_UNIT defined and
// contraction on, for the FP32 kernels.  Each object launches only its own
// instantiations; the dispatchers live in the FP64 object.
static FaceSet faceSet(const Fields &F)
{
	FaceSet U{};
	for (int a = 0; a < F.D; a++) { U.a[a] = F.u[a]; U.g[a] = F.face[a]; }
	return U;
}

#ifdef GARM_ADVECT_FP32_UNIT
using Real = float;
#else
using Real = double;
#endif

template <int D>
static void advectRefMapD(const Fields &F, double dt, hipStream_t s)
{
	RefMapArgs<D> A{};
	A.cell = F.cell;
	A.U = faceSet(F);
	for (int a = 0; a < D; a++) { A.xi[a] = F.xi[a]; A.xi_new[a] = F.xi_new[a]; }
	for (int e = 0; e < D * D; e++) { A.J[e] = F.J[e]; A.J_new[e] = F.J_new[e]; }
	A.h = -dt;
	hipLaunchKernelGGL((kAdvectRefMap<D, Real>), gridDims(F.cell, D), gridBlock(), 0, s, A);
}

template <int D>
static void advectVelocityD(const Fields &F, double dt, hipStream_t s)
{
	const FaceSet U = faceSet(F);
	for (int a = 0; a < D; a++) hipLaunchKernelGGL((kAdvectVelocity<D, Real>), gridDims(F.face[a], D), gridBlock(), 0, s, U, a, F.u_new[a], -dt);
}

template <int D>
static void composeD(const Fields &F, hipStream_t s)
{
	ComposeArgs<D> A{};
	A.cell = F.cell;
	for (int a = 0; a < D; a++) { A.xi[a] = F.xi[a]; A.phi_ref_grad[a] = F.phi_ref_grad[a]; A.phi_grad[a] = F.phi_grad[a]; }
	for (int e = 0; e < D * D; e++) A.J[e] = F.J[e];
	for (int m = 0; m < 4; m++) A.mixed[m] = F.phi_ref_mixed[m];
	A.phi_ref = F.phi_ref;
	A.adv = F.adv_region;
	A.cell_dist = F.cell_dist;
	A.phi = F.phi; A.phi_tmp = F.phi_tmp; A.phi_ori = F.phi_ori;
	A.band = 15.0 * F.cell.dx;
	A.tol = 3.0 * F.cell.dx;
	hipLaunchKernelGGL((kCompose<D, Real>), gridDims(F.cell, D), gridBlock(), 0, s, A);
}

#ifdef GARM_ADVECT_FP32_UNIT

void launchAdvectRefMap32(const Fields &F, double dt, hipStream_t s)
{
	if (F.D == 2) advectRefMapD<2>(F, dt, s); else advectRefMapD<3>(F, dt, s);
}

void launchAdvectVelocity32(const Fields &F, double dt, hipStream_t s)
{
	if (F.D == 2) advectVelocityD<2>(F, dt, s); else advectVelocityD<3>(F, dt, s);
}

void launchCompose32(const Fields &F, hipStream_t s)
{
	if (F.D == 2) composeD<2>(F, s); else composeD<3>(F, s);
}

#else

void launchAdvectRefMap32(const Fields &F, double dt, hipStream_t s);
void launchAdvectVelocity32(const Fields &F, double dt, hipStream_t s);
void launchCompose32(const Fields &F, hipStream_t s);

void launchAdvectRefMap(const Fields &F, double dt, hipStream_t s, bool fp32)
{
	if (fp32) launchAdvectRefMap32(F, dt, s);
	else if (F.D == 2) advectRefMapD<2>(F, dt, s);
	else advectRefMapD<3>(F, dt, s);
}

void launchAdvectVelocity(const Fields &F, double dt, hipStream_t s, bool fp32)
{
	if (fp32) launchAdvectVelocity32(F, dt, s);
	else if (F.D == 2) advectVelocityD<2>(F, dt, s);
	else advectVelocityD<3>(F, dt, s);
}

void launchCompose(const Fields &F, hipStream_t s, bool fp32)
{
	if (fp32) launchCompose32(F, s);
	else if (F.D == 2) composeD<2>(F, s);
	else composeD<3>(F, s);
}

void launchRefLevelSetMixed(const Fields &F, hipStream_t s)
{
	const double *g = F.phi_ref_grad[0];
	if (F.D == 2)
		hipLaunchKernelGGL(kScalarMixed<2>, gridDims(F.cell, 2), gridBlock(), 0, s, F.cell, g, F.phi_ref_grad[1], g,
			F.phi_ref_mixed[0], nullptr, nullptr, nullptr);
	else
		hipLaunchKernelGGL(kScalarMixed<3>, gridDims(F.cell, 3), gridBlock(), 0, s, F.cell, g, F.phi_ref_grad[1], F.phi_ref_grad[2],
			F.phi_ref_mixed[0], F.phi_ref_mixed[1], F.phi_ref_mixed[2], F.phi_ref_mixed[3]);
}

void launchGradient4(const Fields &F, const double *src, double *const *dst, hipStream_t s)
{
	if (F.D == 2) hipLaunchKernelGGL(kGradient4<2>, gridDims(F.cell, 2), gridBlock(), 0, s, F.cell, src, dst[0], dst[1], nullptr);
	else hipLaunchKernelGGL(kGradient4<3>, gridDims(F.cell, 3), gridBlock(), 0, s, F.cell, src, dst[0], dst[1], dst[2]);
}

#endif

} // namespace garm
