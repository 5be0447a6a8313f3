// Sampling views of a cell field shared by the restart and the surface
// tension kernels: the cubic Lagrange view of the level set and its
// gradient (ScalarFieldWrapper<CubicIntrpl>, ScalarField.h:23-47), and
// Eigen's normalized().  FP64, the reference's operations in its order.
#pragma once

#include "garm_common.hpp"

namespace garm {

// CubicIntrpl::wtPoints + interpolate (LagrangeIntrpl.h:128-171): lower
// corner getCubicLower (Grid.h:54), clamped corners (GridBasedData::at).
template <int D>
__device__ __forceinline__ double cubicView(const GridDev &g, const double *phi, const double *p)
{
	int l[3] = {0, 0, 0};
	double f[3] = {0, 0, 0};
#pragma unroll
	for (int a = 0; a < D; a++) l[a] = int(floor((p[a] - g.o[a]) * g.inv)) - 1;
#pragma unroll
	for (int a = 0; a < D; a++) f[a] = (p[a] - g.o[a] - l[a] * g.dx) * g.inv;
	double w[4][3];
#pragma unroll
	for (int a = 0; a < D; a++) {
		const double t = f[a];
		w[0][a] = (t - 1) * (t - 2) * (t - 3) / -6;
		w[1][a] = t * (t - 2) * (t - 3) / 2;
		w[2][a] = t * (t - 1) * (t - 3) / -2;
		w[3][a] = t * (t - 1) * (t - 2) / 6;
	}
	double v = 0;
	for (int k = 0; k < (D == 3 ? 4 : 1); k++)
		for (int j = 0; j < 4; j++)
#pragma unroll
			for (int i = 0; i < 4; i++) {
				const double wt = D == 2 ? w[i][0] * w[j][1] : w[i][0] * w[j][1] * w[k][2];
				v += phi[gat<D>(g, l[0] + i, l[1] + j, l[2] + k)] * wt;
			}
	return v;
}

// ScalarFieldWrapper<CubicIntrpl>::gradient (ScalarField.h:41-47): the
// linear weights of the corners times getFirst<2> at each (unclamped)
// corner, whose stencil points are clamped (Derivatives.h:47, 51-57).
template <int D>
__device__ __forceinline__ void cubicViewGradient(const GridDev &g, const double *phi, const double *p, double *gr)
{
	int l[3] = {0, 0, 0};
	double f[3] = {0, 0, 0};
	lowerFrac<D>(g, p, l, f);
	double w[2][3];
#pragma unroll
	for (int a = 0; a < D; a++) { w[0][a] = 1 - f[a]; w[1][a] = f[a]; gr[a] = 0; }
	for (int k = 0; k < (D == 3 ? 2 : 1); k++)
#pragma unroll
		for (int j = 0; j < 2; j++)
#pragma unroll
			for (int i = 0; i < 2; i++) {
				const int q[3] = {l[0] + i, l[1] + j, l[2] + k};
				const double wt = D == 2 ? w[i][0] * w[j][1] : w[i][0] * w[j][1] * w[k][2];
#pragma unroll
				for (int a = 0; a < D; a++) {
					const double hi = phi[gat<D>(g, q[0] + (a == 0), q[1] + (a == 1), q[2] + (a == 2))];
					const double lo = phi[gat<D>(g, q[0] - (a == 0), q[1] - (a == 1), q[2] - (a == 2))];
					gr[a] += (hi - lo) / 2 * g.inv * wt;
				}
			}
}

// Eigen's normalized(): unchanged when the norm is 0
template <int D>
__device__ __forceinline__ void normalizeD(const double *v, double *out)
{
	double s = 0;
#pragma unroll
	for (int a = 0; a < D; a++) s += v[a] * v[a];
	if (s > 0) {
		const double n = sqrt(s);
#pragma unroll
		for (int a = 0; a < D; a++) out[a] = v[a] / n;
	} else {
#pragma unroll
		for (int a = 0; a < D; a++) out[a] = v[a];
	}
}

} // namespace garm
