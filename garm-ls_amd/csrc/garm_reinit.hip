// WENO5 / TVD-RK3 level-set reinitialisation for gfx950
// (Reinitialization::solve<5>, Reinitialization.h:20-80).
//
// One launch per Runge-Kutta stage; the stage combinations of solve<5>
// (:70-78) are fused into the stage that produces them, so an iteration is
// three passes:  s0 = L(phi);  s1 = (3 phi + L(s0)) / 4;  phi = (phi + 2 L(s1)) / 3.
// The FP32 form computes in float (constant quotients as products with
// rounded reciprocals, the weight and normalisation quotients with the
// hardware reciprocal and reciprocal square root, fused multiply-adds) and
// keeps s0, s1 in float; phi stays FP64.
// Its WENO epsilon floor is 1e-18 instead of 1e-99 (which is 0 in float and
// would give 0/0 on flat data); in FP64 it is the reference's 1e-99.
#include "garm_internal.hpp"

#include <type_traits>

namespace garm {

template <typename T>
__device__ __forceinline__ T sq(T x) { return x * x; }

__device__ __forceinline__ float rcpT(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float rsqrtT(float x) { return __builtin_amdgcn_rsqf(x); }

template <typename T> __device__ __forceinline__ T wenoEpsFloor();
template <> __device__ __forceinline__ double wenoEpsFloor<double>() { return 1e-99; }
template <> __device__ __forceinline__ float wenoEpsFloor<float>() { return 1e-18f; }

// Derivatives::weno5 (Derivatives.h:134-159): (minus, plus) one-sided derivatives.
template <typename T>
__device__ __forceinline__ void weno5(const T *v, T invDx, T &dm, T &dp)
{
#pragma unroll
	for (int i = 0; i < 2; i++) {
		T dev[5];
		if (i) { dev[0] = v[6] - v[5]; dev[1] = v[5] - v[4]; dev[2] = v[4] - v[3]; dev[3] = v[3] - v[2]; dev[4] = v[2] - v[1]; }
		else { dev[0] = v[1] - v[0]; dev[1] = v[2] - v[1]; dev[2] = v[3] - v[2]; dev[3] = v[4] - v[3]; dev[4] = v[5] - v[4]; }
		T eps = 0;
#pragma unroll
		for (int j = 0; j < 5; j++) eps = fmax(eps, dev[j] * dev[j]);
		eps = eps * T(1e-6) + wenoEpsFloor<T>();
		// the constant quotients of the reference formulas as products with
		// rounded reciprocals (within one ulp of the quotient; twelve
		// divisions fewer per side)
		constexpr T k3 = T(1.0 / 3), k6 = T(1.0 / 6), k56 = T(5.0 / 6), k76 = T(7.0 / 6), k116 = T(11.0 / 6), k1312 = T(13.0 / 12);
		const T phix1 = dev[0] * k3 - dev[1] * k76 + dev[2] * k116;
		const T phix2 = -dev[1] * k6 + dev[2] * k56 + dev[3] * k3;
		const T phix3 = dev[2] * k3 + dev[3] * k56 - dev[4] * k6;
		const T s1 = sq(dev[0] - 2 * dev[1] + dev[2]) * k1312 + sq(dev[0] - 4 * dev[1] + 3 * dev[2]) * T(0.25);
		const T s2 = sq(dev[1] - 2 * dev[2] + dev[3]) * k1312 + sq(dev[1] - dev[3]) * T(0.25);
		const T s3 = sq(dev[2] - 2 * dev[3] + dev[4]) * k1312 + sq(3 * dev[2] - 4 * dev[3] + dev[4]) * T(0.25);
		// v_rcp_f32 (1 ulp); s + eps >= 1e-18 keeps every weight below 1e36
		const T a1 = T(.1) * rcpT(sq(s1 + eps));
		const T a2 = T(.6) * rcpT(sq(s2 + eps));
		const T a3 = T(.3) * rcpT(sq(s3 + eps));
		const T sum = a1 + a2 + a3;
		(i ? dp : dm) = (a1 * phix1 + a2 * phix2 + a3 * phix3) * rcpT(sum) * invDx;
	}
}

// The FP64 form evaluates the reference's expressions in the reference's
// order (divisions by the constants, left to right) without contracting
// a * b + c into fused multiply-adds, i.e. the IEEE operations of the CPU
// code: the front where the reinitialisation meets the flat air band
// amplifies the last-bit differences of any other order by eight orders of
// magnitude.
__device__ __forceinline__ double sqExact(double x)
{
#pragma clang fp contract(off)
	return x * x;
}

__device__ __forceinline__ void weno5Exact(const double *v, double invDx, double &dm, double &dp)
{
#pragma clang fp contract(off)
#pragma unroll
	for (int i = 0; i < 2; i++) {
		double dev[5];
		if (i) { dev[0] = v[6] - v[5]; dev[1] = v[5] - v[4]; dev[2] = v[4] - v[3]; dev[3] = v[3] - v[2]; dev[4] = v[2] - v[1]; }
		else { dev[0] = v[1] - v[0]; dev[1] = v[2] - v[1]; dev[2] = v[3] - v[2]; dev[3] = v[4] - v[3]; dev[4] = v[5] - v[4]; }
		double eps = 0;
#pragma unroll
		for (int j = 0; j < 5; j++) eps = fmax(eps, dev[j] * dev[j]);
		eps = eps * 1e-6 + 1e-99;
		const double phix1 = dev[0] / 3 - dev[1] * 7 / 6 + dev[2] * 11 / 6;
		const double phix2 = -dev[1] / 6 + dev[2] * 5 / 6 + dev[3] / 3;
		const double phix3 = dev[2] / 3 + dev[3] * 5 / 6 - dev[4] / 6;
		const double s1 = sqExact(dev[0] - 2 * dev[1] + dev[2]) * 13 / 12 + sqExact(dev[0] - 4 * dev[1] + 3 * dev[2]) / 4;
		const double s2 = sqExact(dev[1] - 2 * dev[2] + dev[3]) * 13 / 12 + sqExact(dev[1] - dev[3]) / 4;
		const double s3 = sqExact(dev[2] - 2 * dev[3] + dev[4]) * 13 / 12 + sqExact(3 * dev[2] - 4 * dev[3] + dev[4]) / 4;
		const double a1 = .1 / sqExact(s1 + eps);
		const double a2 = .6 / sqExact(s2 + eps);
		const double a3 = .3 / sqExact(s3 + eps);
		const double sum = a1 + a2 + a3;
		(i ? dp : dm) = (a1 * phix1 + a2 * phix2 + a3 * phix3) / sum * invDx;
	}
}

// The 7-point stencils of the D axes around cell c, clamped into the grid
// (GridBasedData::at).  Cells three or more away from every face of the grid
// (all but the outer layers) take plain offsets from their own index.
template <int D, typename T, typename TI>
__device__ __forceinline__ void gatherStencils(const GridDev &g, const TI *f, const Coord &c, size_t id, T (&v)[3][7])
{
	const bool inner = c.i >= 3 && c.i < g.n[0] - 3 && c.j >= 3 && c.j < g.n[1] - 3 && (D == 2 || (c.k >= 3 && c.k < g.n[2] - 3));
	if (inner) {
		const long sx = 1, sy = g.n[0], sz = long(g.n[0]) * g.n[1];
		const TI *p = f + id;
#pragma unroll
		for (int a = 0; a < D; a++) {
			const long st = a == 0 ? sx : a == 1 ? sy : sz;
#pragma unroll
			for (int t = 0; t < 7; t++) v[a][t] = T(p[(t - 3) * st]);
		}
	} else {
#pragma unroll
		for (int a = 0; a < D; a++)
#pragma unroll
			for (int t = 0; t < 7; t++) {
				const int o = t - 3;
				v[a][t] = T(f[gat<D>(g, c.i + (a == 0 ? o : 0), c.j + (a == 1 ? o : 0), c.k + (a == 2 ? o : 0))]);
			}
	}
}

template <int D>
__device__ __forceinline__ double advanceCellExact(const GridDev &g, const double *f, const Coord &c, size_t id, double dtau)
{
#pragma clang fp contract(off)
	double v[3][7];
	gatherStencils<D, double, double>(g, f, c, id, v);
	double dm[3], dp[3];
#pragma unroll
	for (int a = 0; a < D; a++) weno5Exact(v[a], g.inv, dm[a], dp[a]);
	double sum = 0;
	const double val = v[0][3];
	if (val < 0) {
#pragma unroll
		for (int a = 0; a < D; a++) { sum += sqExact(fmin(dm[a], 0.0)); sum += sqExact(fmax(dp[a], 0.0)); }
	} else {
#pragma unroll
		for (int a = 0; a < D; a++) { sum += sqExact(fmax(dm[a], 0.0)); sum += sqExact(fmin(dp[a], 0.0)); }
	}
	sum = sqrt(sum);
	const double sp = g.dx * sum;
	return val + dtau * (val / sqrt(val * val + sp * sp)) * (1 - sum);
}

// Reinitialization::advance (Reinitialization.h:20-48) at one cell, in float,
// reading the stage input from storage of type TI.
template <int D, typename TI>
__device__ __forceinline__ float advanceCell(const GridDev &g, const TI *f, const Coord &c, size_t id, float dtau)
{
	float v[3][7];
	gatherStencils<D, float, TI>(g, f, c, id, v);
	float dm[3], dp[3];
#pragma unroll
	for (int a = 0; a < D; a++) weno5<float>(v[a], float(g.inv), dm[a], dp[a]);
	const float val = v[0][3];
	float sum = 0;
	if (val < 0) {
#pragma unroll
		for (int a = 0; a < D; a++) { sum += sq(fminf(dm[a], 0.f)); sum += sq(fmaxf(dp[a], 0.f)); }
	} else {
#pragma unroll
		for (int a = 0; a < D; a++) { sum += sq(fmaxf(dm[a], 0.f)); sum += sq(fminf(dp[a], 0.f)); }
	}
	sum = sqrtf(sum);
	const float sp = float(g.dx) * sum;
	return val + dtau * (val * rsqrtT(val * val + sp * sp)) * (1 - sum);
}

// mode 0: out = L(in);  mode 1: out = (3 phi + L(in)) / 4;  mode 2: phi = (phi + 2 L(in)) / 3
template <int D, int MODE, typename T, typename TI, typename TO>
__global__ __launch_bounds__(kBlock) void kReinitStage(GridDev g, const TI *in, TO *out, double *phi, double dtau)
{
	const Coord c = threadCoord();
	if (!inGrid<D>(g, c)) return;
	const size_t id = gidx<D>(g, c.i, c.j, c.k);
	if constexpr (std::is_same<T, double>::value) {
#pragma clang fp contract(off)
		const double l = advanceCellExact<D>(g, in, c, id, dtau);
		if constexpr (MODE == 0) out[id] = l;
		else if constexpr (MODE == 1) out[id] = (phi[id] * 3 + l) / 4;
		else phi[id] = (phi[id] + l * 2) / 3;
	} else {
		const T l = advanceCell<D, TI>(g, in, c, id, T(dtau));
		if constexpr (MODE == 0) out[id] = TO(l);
		else if constexpr (MODE == 1) out[id] = TO((T(phi[id]) * 3 + l) / 4);
		else phi[id] = (T(phi[id]) + l * 2) / 3;
	}
}

template <int D, typename T>
static void reinitD(const Fields &F, int iterations, double cfl, hipStream_t s)
{
	const double dtau = F.cell.dx * cfl;
	const dim3 gd = gridDims(F.cell, D), bd = gridBlock();
	// stage buffers in the compute type (scratch holds N doubles each)
	T *s0 = reinterpret_cast<T *>(F.scratch[0]), *s1 = reinterpret_cast<T *>(F.scratch[1]);
	for (int it = 0; it < iterations; it++) {
		hipLaunchKernelGGL((kReinitStage<D, 0, T, double, T>), gd, bd, 0, s, F.cell, (const double *)F.phi, s0, F.phi, dtau);
		hipLaunchKernelGGL((kReinitStage<D, 1, T, T, T>), gd, bd, 0, s, F.cell, (const T *)s0, s1, F.phi, dtau);
		hipLaunchKernelGGL((kReinitStage<D, 2, T, T, T>), gd, bd, 0, s, F.cell, (const T *)s1, (T *)nullptr, F.phi, dtau);
	}
}

void launchReinitStage(const Fields &F, double dtau, hipStream_t s, bool fp32)
{
	const dim3 gd = gridDims(F.cell, F.D), bd = gridBlock();
	const double *phi = F.phi;
	if (fp32) {
		float *s0 = reinterpret_cast<float *>(F.scratch[0]);
		if (F.D == 2) hipLaunchKernelGGL((kReinitStage<2, 0, float, double, float>), gd, bd, 0, s, F.cell, phi, s0, F.phi, dtau);
		else hipLaunchKernelGGL((kReinitStage<3, 0, float, double, float>), gd, bd, 0, s, F.cell, phi, s0, F.phi, dtau);
	} else {
		if (F.D == 2) hipLaunchKernelGGL((kReinitStage<2, 0, double, double, double>), gd, bd, 0, s, F.cell, phi, F.scratch[0], F.phi, dtau);
		else hipLaunchKernelGGL((kReinitStage<3, 0, double, double, double>), gd, bd, 0, s, F.cell, phi, F.scratch[0], F.phi, dtau);
	}
}

void launchReinit(const Fields &F, int iterations, double cfl, hipStream_t s, bool fp32)
{
	if (F.D == 2) { if (fp32) reinitD<2, float>(F, iterations, cfl, s); else reinitD<2, double>(F, iterations, cfl, s); }
	else { if (fp32) reinitD<3, float>(F, iterations, cfl, s); else reinitD<3, double>(F, iterations, cfl, s); }
}

} // namespace garm
