// WENO5 / TVD-RK3 level-set reinitialisation for gfx950
// (Reinitialization::solve<5>, Reinitialization.h:20-80).
//
// One launch per Runge-Kutta stage; the stage combinations of solve<5>
// (:70-78) are fused into the stage that produces them, so an iteration is
// three passes:  s0 = L(phi);  s1 = (3 phi + L(s0)) / 4;  phi = (phi + 2 L(s1)) / 3.
#include "garm_internal.hpp"

namespace garm {

__device__ __forceinline__ double sq(double x) { return x * x; }

// Derivatives::weno5 (Derivatives.h:134-159): (minus, plus) one-sided derivatives.
__device__ __forceinline__ void weno5(const double *v, double invDx, double &dm, double &dp)
{
#pragma unroll
	for (int i = 0; i < 2; i++) {
		double dev[5];
		if (i) { dev[0] = v[6] - v[5]; dev[1] = v[5] - v[4]; dev[2] = v[4] - v[3]; dev[3] = v[3] - v[2]; dev[4] = v[2] - v[1]; }
		else { dev[0] = v[1] - v[0]; dev[1] = v[2] - v[1]; dev[2] = v[3] - v[2]; dev[3] = v[4] - v[3]; dev[4] = v[5] - v[4]; }
		double eps = 0;
#pragma unroll
		for (int j = 0; j < 5; j++) eps = fmax(eps, dev[j] * dev[j]);
		eps = eps * 1e-6 + 1e-99;
		// the constant quotients of the reference formulas as products with
		// rounded reciprocals (within one ulp of the quotient; twelve FP64
		// divisions fewer per side)
		constexpr double k3 = 1.0 / 3, k6 = 1.0 / 6, k56 = 5.0 / 6, k76 = 7.0 / 6, k116 = 11.0 / 6, k1312 = 13.0 / 12;
		const double phix1 = dev[0] * k3 - dev[1] * k76 + dev[2] * k116;
		const double phix2 = -dev[1] * k6 + dev[2] * k56 + dev[3] * k3;
		const double phix3 = dev[2] * k3 + dev[3] * k56 - dev[4] * k6;
		const double s1 = sq(dev[0] - 2 * dev[1] + dev[2]) * k1312 + sq(dev[0] - 4 * dev[1] + 3 * dev[2]) * 0.25;
		const double s2 = sq(dev[1] - 2 * dev[2] + dev[3]) * k1312 + sq(dev[1] - dev[3]) * 0.25;
		const double s3 = sq(dev[2] - 2 * dev[3] + dev[4]) * k1312 + sq(3 * dev[2] - 4 * dev[3] + dev[4]) * 0.25;
		const double a1 = .1 / sq(s1 + eps);
		const double a2 = .6 / sq(s2 + eps);
		const double a3 = .3 / sq(s3 + eps);
		const double sum = a1 + a2 + a3;
		(i ? dp : dm) = (a1 * phix1 + a2 * phix2 + a3 * phix3) / sum * invDx;
	}
}

// Reinitialization::advance (Reinitialization.h:20-48) at one cell.
template <int D>
__device__ __forceinline__ double advanceCell(const GridDev &g, const double *f, const Coord &c, size_t id, double dtau)
{
	double dm[3], dp[3];
#pragma unroll
	for (int a = 0; a < D; a++) {
		double v[7];
#pragma unroll
		for (int t = 0; t < 7; t++) {
			const int o = t - 3;
			v[t] = f[gat<D>(g, c.i + (a == 0 ? o : 0), c.j + (a == 1 ? o : 0), c.k + (a == 2 ? o : 0))];
		}
		weno5(v, g.inv, dm[a], dp[a]);
	}
	const double val = f[id];
	double sum = 0;
	if (val < 0) {
#pragma unroll
		for (int a = 0; a < D; a++) { sum += sq(fmin(dm[a], 0.0)); sum += sq(fmax(dp[a], 0.0)); }
	} else {
#pragma unroll
		for (int a = 0; a < D; a++) { sum += sq(fmax(dm[a], 0.0)); sum += sq(fmin(dp[a], 0.0)); }
	}
	sum = sqrt(sum);
	const double sp = g.dx * sum;
	return val + dtau * (val / sqrt(val * val + sp * sp)) * (1 - sum);
}

// mode 0: out = L(in);  mode 1: out = (3 phi + L(in)) / 4;  mode 2: phi = (phi + 2 L(in)) / 3
template <int D, int MODE>
__global__ __launch_bounds__(kBlock) void kReinitStage(GridDev g, const double *in, double *out, double *phi, double dtau)
{
	const Coord c = threadCoord();
	if (!inGrid<D>(g, c)) return;
	const size_t id = gidx<D>(g, c.i, c.j, c.k);
	const double l = advanceCell<D>(g, in, c, id, dtau);
	if constexpr (MODE == 0) out[id] = l;
	else if constexpr (MODE == 1) out[id] = (phi[id] * 3 + l) / 4;
	else phi[id] = (phi[id] + l * 2) / 3;
}

template <int D>
static void reinitD(const Fields &F, int iterations, double cfl, hipStream_t s)
{
	const double dtau = F.cell.dx * cfl;
	const dim3 gd = gridDims(F.cell, D), bd = gridBlock();
	double *s0 = F.scratch[0], *s1 = F.scratch[1];
	for (int it = 0; it < iterations; it++) {
		hipLaunchKernelGGL((kReinitStage<D, 0>), gd, bd, 0, s, F.cell, F.phi, s0, F.phi, dtau);
		hipLaunchKernelGGL((kReinitStage<D, 1>), gd, bd, 0, s, F.cell, s0, s1, F.phi, dtau);
		hipLaunchKernelGGL((kReinitStage<D, 2>), gd, bd, 0, s, F.cell, s1, nullptr, F.phi, dtau);
	}
}

void launchReinitStage(const Fields &F, double dtau, hipStream_t s)
{
	if (F.D == 2) hipLaunchKernelGGL((kReinitStage<2, 0>), gridDims(F.cell, 2), gridBlock(), 0, s, F.cell, (const double *)F.phi, F.scratch[0], F.phi, dtau);
	else hipLaunchKernelGGL((kReinitStage<3, 0>), gridDims(F.cell, 3), gridBlock(), 0, s, F.cell, (const double *)F.phi, F.scratch[0], F.phi, dtau);
}

void launchReinit(const Fields &F, int iterations, double cfl, hipStream_t s)
{
	if (F.D == 2) reinitD<2>(F, iterations, cfl, s); else reinitD<3>(F, iterations, cfl, s);
}

} // namespace garm
