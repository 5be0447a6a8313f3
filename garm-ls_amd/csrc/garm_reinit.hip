// WENO5 / TVD-RK3 level-set reinitialisation for gfx950
// (Reinitialization::solve<5>, Reinitialization.h:20-80).
//
// One launch per Runge-Kutta stage; the stage combinations of solve<5>
// (:70-78) are fused into the stage that produces them, so an iteration is
// three passes:  s0 = L(phi);  s1 = (3 phi + L(s0)) / 4;  phi = (phi + 2 L(s1)) / 3.
// The FP32 form computes in float (constant quotients as products with
// rounded reciprocals, the weight and normalisation quotients with the
// hardware reciprocal, square root and reciprocal square root, fused
// multiply-adds, the WENO weights up to common factors that cancel) and keeps
// s0, s1 and a float copy of phi in float; phi stays FP64.
// Its WENO epsilon floor is 1e-18 instead of 1e-99 (which is 0 in float and
// would give 0/0 on flat data); in FP64 it is the reference's 1e-99.
#include "garm_internal.hpp"

#include <algorithm>
#include <type_traits>

namespace garm {

template <typename T>
__device__ __forceinline__ T sq(T x) { return x * x; }

__device__ __forceinline__ float rcpT(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float rsqrtT(float x) { return __builtin_amdgcn_rsqf(x); }

// Derivatives::weno5 (Derivatives.h:134-159): (minus, plus) one-sided
// derivatives, in float and times dx (the caller scales by 1/dx once).
// The two sides share the differences d_k = v[k+1] - v[k], the squares behind
// their epsilon and two of their second differences; every quantity is
// evaluated once.  The weights are the reference's up to common factors,
// which cancel in the normalisation: smoothness indicators and epsilon are
// scaled by 4 (13/3 e^2 + f^2 instead of 13/12 e^2 + 1/4 f^2), the linear
// weights 0.1 : 0.6 : 0.3 are 1 : 6 : 3 with the 6 and the 3 folded into the
// candidate polynomials 2 and 3.  The nonlinear weights 1 / (s_k + eps)^2 are
// multiplied through by the product of all three, (s_k / eps + 1)^2 taken
// with one reciprocal of eps: weight k becomes the product of the other two
// (between 1 and 1e31, since eps is 4e-6 of the largest d^2 and s_k at most
// about 140 times it), so a side takes two reciprocals instead of four.
__device__ __forceinline__ void weno5f(const float *v, float &dm, float &dp)
{
	float d[6];
#pragma unroll
	for (int k = 0; k < 6; k++) d[k] = v[k + 1] - v[k];
	float q[6];
#pragma unroll
	for (int k = 0; k < 6; k++) q[k] = d[k] * d[k];
	// 4 (1e-6 max d^2 + 1e-18); the floor keeps 0/0 away on flat data
	const float mid = fmaxf(fmaxf(q[1], q[2]), fmaxf(q[3], q[4]));
	const float epsM = fmaxf(mid, q[0]) * 4e-6f + 4e-18f;
	const float epsP = fmaxf(mid, q[5]) * 4e-6f + 4e-18f;
	// second differences e_k = d_{k-1} - 2 d_k + d_{k+1}, k = 1..4
	float ee[5];
#pragma unroll
	for (int k = 1; k < 5; k++) { const float e = d[k - 1] - 2 * d[k] + d[k + 1]; ee[k] = e * e; }
	constexpr float k133 = 13.f / 3, k3 = 1.f / 3, k76 = 7.f / 6, k116 = 11.f / 6;
	// minus side: the stencil d0..d4
	{
		const float s1 = ee[1] * k133 + sq(d[0] - 4 * d[1] + 3 * d[2]);
		const float s2 = ee[2] * k133 + sq(d[1] - d[3]);
		const float s3 = ee[3] * k133 + sq(3 * d[2] - 4 * d[3] + d[4]);
		const float p1 = d[0] * k3 - d[1] * k76 + d[2] * k116;       // candidate 1
		const float p2 = 5 * d[2] - d[1] + 2 * d[3];                  // 6 x candidate 2
		const float p3 = d[2] + 2.5f * d[3] - 0.5f * d[4];            // 3 x candidate 3
		const float r = rcpT(epsM), t1 = sq(s1 * r + 1), t2 = sq(s2 * r + 1), t3 = sq(s3 * r + 1);
		const float a1 = t2 * t3, a2 = t1 * t3, a3 = t1 * t2;
		dm = (a1 * p1 + a2 * p2 + a3 * p3) * rcpT(a1 + 6 * a2 + 3 * a3);
	}
	// plus side: the stencil d5, d4, d3, d2, d1
	{
		const float s1 = ee[4] * k133 + sq(d[5] - 4 * d[4] + 3 * d[3]);
		const float s2 = ee[3] * k133 + sq(d[4] - d[2]);
		const float s3 = ee[2] * k133 + sq(3 * d[3] - 4 * d[2] + d[1]);
		const float p1 = d[5] * k3 - d[4] * k76 + d[3] * k116;
		const float p2 = 5 * d[3] - d[4] + 2 * d[2];
		const float p3 = d[3] + 2.5f * d[2] - 0.5f * d[1];
		const float r = rcpT(epsP), t1 = sq(s1 * r + 1), t2 = sq(s2 * r + 1), t3 = sq(s3 * r + 1);
		const float a1 = t2 * t3, a2 = t1 * t3, a3 = t1 * t2;
		dp = (a1 * p1 + a2 * p2 + a3 * p3) * rcpT(a1 + 6 * a2 + 3 * a3);
	}
}

// One WENO5 window of five differences w0..w4 (times dx) along an axis: the
// minus-side derivative of the cell whose stencil starts the window
// (forward) and the plus-side derivative of the cell before that one (the
// window reversed).  The two share the window's epsilon, its three
// smoothness indicators and the products behind the nonlinear weights --
// reversed, the indicators swap ends, so the weights (a1, a2, a3) become
// (a3, a2, a1) -- with the formulas of weno5f's minus side; only the
// candidate polynomials and the normalisation are per side.  A thread that
// walks a column evaluates one window per cell instead of two.
__device__ __forceinline__ void weno5Window(const float *w, const float *q, float &fwd, float &rev)
{
	const float eps = fmaxf(fmaxf(fmaxf(q[1], q[2]), fmaxf(q[3], q[4])), q[0]) * 4e-6f + 4e-18f;
	constexpr float k133 = 13.f / 3, k3 = 1.f / 3, k76 = 7.f / 6, k116 = 11.f / 6;
	const float e1 = w[0] - 2 * w[1] + w[2], e2 = w[1] - 2 * w[2] + w[3], e3 = w[2] - 2 * w[3] + w[4];
	const float sa = e1 * e1 * k133 + sq(w[0] - 4 * w[1] + 3 * w[2]);
	const float sb = e2 * e2 * k133 + sq(w[1] - w[3]);
	const float sc = e3 * e3 * k133 + sq(3 * w[2] - 4 * w[3] + w[4]);
	const float r = rcpT(eps), ta = sq(sa * r + 1), tb = sq(sb * r + 1), tc = sq(sc * r + 1);
	const float a1 = tb * tc, a2 = ta * tc, a3 = ta * tb;
	{
		const float p1 = w[0] * k3 - w[1] * k76 + w[2] * k116;
		const float p2 = 5 * w[2] - w[1] + 2 * w[3];
		const float p3 = w[2] + 2.5f * w[3] - 0.5f * w[4];
		fwd = (a1 * p1 + a2 * p2 + a3 * p3) * rcpT(a1 + 6 * a2 + 3 * a3);
	}
	{
		const float p1 = w[4] * k3 - w[3] * k76 + w[2] * k116;
		const float p2 = 5 * w[2] - w[3] + 2 * w[1];
		const float p3 = w[2] + 2.5f * w[1] - 0.5f * w[0];
		rev = (a3 * p1 + a2 * p2 + a1 * p3) * rcpT(a3 + 6 * a2 + 3 * a1);
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

// Load of base[i] as base + 4 i (or 8 i) bytes with a 32-bit offset: the
// scalar base plus a vector offset, no 64-bit vector address arithmetic.
template <typename TI>
__device__ __forceinline__ TI ldOff(const TI *base, unsigned byteOff)
{
	return *reinterpret_cast<const TI *>(reinterpret_cast<const char *>(base) + byteOff);
}

// Offsets of the cell c of a stage kernel.  A wavefront is one x row of 64
// cells (blocks are 64 x 4), so j and k are the same across it: the row
// offsets are scalar values and the wave's loads differ only in i.
struct StageRow {
	int i, j, k;
	unsigned row;      // index of cell (0, j, k)
	unsigned ib;       // i in bytes of the element type
};

template <int D, typename TI>
__device__ __forceinline__ StageRow stageRow(const GridDev &g, const Coord &c)
{
	StageRow r;
	r.i = c.i;
	r.j = __builtin_amdgcn_readfirstlane(c.j);
	r.k = D == 3 ? __builtin_amdgcn_readfirstlane(c.k) : 0;
	r.row = (unsigned(r.k) * unsigned(g.n[1]) + unsigned(r.j)) * unsigned(g.n[0]);   // grids hold < 2^31 cells
	r.ib = unsigned(c.i) * unsigned(sizeof(TI));
	return r;
}

// The 7-point stencils of the D axes around cell c, clamped into the grid
// (GridBasedData::at).  The y and z offsets are clamped in scalar
// arithmetic; the x offsets take immediate displacements, except in the
// waves that touch the first or last three columns, which clamp per lane.
template <int D, typename T, typename TI>
__device__ __forceinline__ void gatherStencils(const GridDev &g, const TI *f, const StageRow &r, T (&v)[3][7])
{
	const int nx = g.n[0], ny = g.n[1];
	const TI *rp = f + r.row;
	const int i0 = int(blockIdx.x * blockDim.x);   // the wave's first column
	if (i0 >= 3 && i0 + 64 + 3 <= nx) {
#pragma unroll
		for (int t = 0; t < 7; t++) v[0][t] = T(ldOff(rp + (t - 3), r.ib));
	} else {
#pragma unroll
		for (int t = 0; t < 7; t++) v[0][t] = T(rp[clampi(r.i + t - 3, nx - 1)]);
	}
#pragma unroll
	for (int t = 0; t < 7; t++) {
		if (t == 3) { v[1][3] = v[0][3]; continue; }
		const unsigned jr = (unsigned(r.k) * unsigned(ny) + unsigned(clampi(r.j + t - 3, ny - 1))) * unsigned(nx);
		v[1][t] = T(ldOff(f + jr, r.ib));
	}
	if constexpr (D == 3) {
#pragma unroll
		for (int t = 0; t < 7; t++) {
			if (t == 3) { v[2][3] = v[0][3]; continue; }
			const unsigned kr = (unsigned(clampi(r.k + t - 3, g.n[2] - 1)) * unsigned(ny) + unsigned(r.j)) * unsigned(nx);
			v[2][t] = T(ldOff(f + kr, r.ib));
		}
	}
}

template <int D>
__device__ __forceinline__ double advanceCellExact(const GridDev &g, const double *f, const StageRow &r, double dtau)
{
#pragma clang fp contract(off)
	double v[3][7];
	gatherStencils<D, double, double>(g, f, r, v);
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
__device__ __forceinline__ float advanceCell(const GridDev &g, const TI *f, const StageRow &r, float dtau)
{
	float v[3][7];
	gatherStencils<D, float, TI>(g, f, r, v);
	float dm[3], dp[3];
#pragma unroll
	for (int a = 0; a < D; a++) weno5f(v[a], dm[a], dp[a]);
	const float val = v[0][3];
	float sum = 0;
	if (val < 0) {
#pragma unroll
		for (int a = 0; a < D; a++) { sum += sq(fminf(dm[a], 0.f)); sum += sq(fmaxf(dp[a], 0.f)); }
	} else {
#pragma unroll
		for (int a = 0; a < D; a++) { sum += sq(fmaxf(dm[a], 0.f)); sum += sq(fminf(dp[a], 0.f)); }
	}
	// |grad phi| = sqrt(sum) / dx (v_sqrt_f32, 1 ulp)
	sum = __builtin_amdgcn_sqrtf(sum) * float(g.inv);
	const float sp = float(g.dx) * sum;
	return val + dtau * (val * rsqrtT(val * val + sp * sp)) * (1 - sum);
}

// mode 0: out = L(in);  mode 1: out = (3 phi + L(in)) / 4;  mode 2: phi = (phi + 2 L(in)) / 3
// The float form reads phi as a float copy (phi32, the values it would round
// FP64 phi to) and mode 2 writes both phi and phi32.
template <int D, int MODE, typename T, typename TI, typename TO>
__global__ __launch_bounds__(kBlock) void kReinitStage(GridDev g, const TI *in, TO *out, double *phi, float *phi32, double dtau)
{
	const Coord c = threadCoord();
	if (!inGrid<D>(g, c)) return;
	const StageRow r = stageRow<D, TI>(g, c);
	const size_t id = size_t(r.row) + unsigned(r.i);
	if constexpr (std::is_same<T, double>::value) {
#pragma clang fp contract(off)
		(void)phi32;
		const double l = advanceCellExact<D>(g, in, r, dtau);
		if constexpr (MODE == 0) out[id] = l;
		else if constexpr (MODE == 1) out[id] = (phi[id] * 3 + l) / 4;
		else phi[id] = (phi[id] + l * 2) / 3;
	} else {
		const T l = advanceCell<D, TI>(g, in, r, T(dtau));
		const unsigned i = unsigned(r.i);
		if constexpr (MODE == 0) (out + r.row)[i] = TO(l);
		else if constexpr (MODE == 1) (out + r.row)[i] = TO(((phi32 + r.row)[i] * 3 + l) / 4);
		else {
			const T x = ((phi32 + r.row)[i] + l * 2) / 3;
			(phi + r.row)[i] = x;
			(phi32 + r.row)[i] = x;
		}
	}
}

// The FP32 3D stage with the z axis walked: a thread owns kReinitZ cells of
// a z-column, loads the column's 7-point z stencils once (kReinitZ + 6
// values) and evaluates one z window per cell (weno5Window: the minus side
// of cell k and the plus side of cell k - 1 from the same window) instead of
// two; x and y are as in kReinitStage.  Blocks are 64 x 4 threads over x, y
// and kReinitZ cells in z, so a wavefront is one x row at fixed j.
constexpr int kReinitZ = 8;

template <int MODE>
__global__ __launch_bounds__(kBlock) void kReinitStageZ(GridDev g, const float *in, float *out, double *phi, float *phi32, double dtau)
{
	constexpr int KZ = kReinitZ;
	const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
	const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
	const int j = __builtin_amdgcn_readfirstlane(int(blockIdx.y * blockDim.y + threadIdx.y));
	const int k0 = int(blockIdx.z) * KZ;
	if (i >= nx || j >= ny) return;
	const unsigned ib = unsigned(i) * 4u;
	float zv[KZ + 6];
#pragma unroll
	for (int t = 0; t < KZ + 6; t++) {
		const unsigned kr = (unsigned(clampi(k0 - 3 + t, nz - 1)) * unsigned(ny) + unsigned(j)) * unsigned(nx);
		zv[t] = ldOff(in + kr, ib);
	}
	float dz[KZ + 5], qz[KZ + 5];
#pragma unroll
	for (int t = 0; t < KZ + 5; t++) { dz[t] = zv[t + 1] - zv[t]; qz[t] = dz[t] * dz[t]; }
	const float dt = float(dtau);
	const int i0 = int(blockIdx.x * blockDim.x);   // the wave's first column
	const bool interiorX = i0 >= 3 && i0 + 64 + 3 <= nx;
	float dmPrev = 0;
#pragma unroll
	for (int w = 0; w <= KZ; w++) {
		float fwd, rev;
		weno5Window(dz + w, qz + w, fwd, rev);
		if (w > 0) {
			const int kk = w - 1, k = k0 + kk;
			if (k < nz) {
				const unsigned row = (unsigned(k) * unsigned(ny) + unsigned(j)) * unsigned(nx);
				float v[2][7];
				const float *rp = in + row;
				if (interiorX) {
#pragma unroll
					for (int t = 0; t < 7; t++) v[0][t] = ldOff(rp + (t - 3), ib);
				} else {
#pragma unroll
					for (int t = 0; t < 7; t++) v[0][t] = rp[clampi(i + t - 3, nx - 1)];
				}
#pragma unroll
				for (int t = 0; t < 7; t++) {
					if (t == 3) { v[1][3] = v[0][3]; continue; }
					const unsigned jr = (unsigned(k) * unsigned(ny) + unsigned(clampi(j + t - 3, ny - 1))) * unsigned(nx);
					v[1][t] = ldOff(in + jr, ib);
				}
				float dm[3], dp[3];
				weno5f(v[0], dm[0], dp[0]);
				weno5f(v[1], dm[1], dp[1]);
				dm[2] = dmPrev;
				dp[2] = rev;
				const float val = zv[kk + 3];
				float sum = 0;
				if (val < 0) {
#pragma unroll
					for (int a = 0; a < 3; a++) { sum += sq(fminf(dm[a], 0.f)); sum += sq(fmaxf(dp[a], 0.f)); }
				} else {
#pragma unroll
					for (int a = 0; a < 3; a++) { sum += sq(fmaxf(dm[a], 0.f)); sum += sq(fminf(dp[a], 0.f)); }
				}
				sum = __builtin_amdgcn_sqrtf(sum) * float(g.inv);
				const float sp = float(g.dx) * sum;
				const float l = val + dt * (val * rsqrtT(val * val + sp * sp)) * (1 - sum);
				if constexpr (MODE == 0) (out + row)[i] = l;
				else if constexpr (MODE == 1) (out + row)[i] = ((phi32 + row)[i] * 3 + l) / 4;
				else {
					const float x = ((phi32 + row)[i] + l * 2) / 3;
					(phi + row)[i] = x;
					(phi32 + row)[i] = x;
				}
			}
		}
		dmPrev = fwd;
	}
}

static dim3 gridDimsZ(const GridDev &g)
{
	return dim3((g.n[0] + 63) / 64, (g.n[1] + 3) / 4, (g.n[2] + kReinitZ - 1) / kReinitZ);
}

__global__ void kPhiToFloat(const double *phi, float *out, size_t n)
{
	for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) out[i] = float(phi[i]);
}

// float stage buffers: s0 in the first half of scratch[0], the float copy of
// phi in its second half, s1 in scratch[1] (each scratch array holds N doubles)
static float *phi32Of(const Fields &F) { return reinterpret_cast<float *>(F.scratch[0]) + F.N; }

static void phiToFloat(const Fields &F, hipStream_t s)
{
	const unsigned nb = unsigned(std::min<size_t>((F.N + kBlock - 1) / kBlock, 8192));
	hipLaunchKernelGGL(kPhiToFloat, dim3(nb), dim3(kBlock), 0, s, (const double *)F.phi, phi32Of(F), F.N);
}

template <int D, typename T>
static void reinitD(const Fields &F, int iterations, double cfl, hipStream_t s)
{
	const double dtau = F.cell.dx * cfl;
	const dim3 gd = gridDims(F.cell, D), bd = gridBlock();
	if constexpr (std::is_same<T, double>::value) {
		double *s0 = F.scratch[0], *s1 = F.scratch[1];
		for (int it = 0; it < iterations; it++) {
			hipLaunchKernelGGL((kReinitStage<D, 0, T, double, T>), gd, bd, 0, s, F.cell, (const double *)F.phi, s0, F.phi, (float *)nullptr, dtau);
			hipLaunchKernelGGL((kReinitStage<D, 1, T, T, T>), gd, bd, 0, s, F.cell, (const T *)s0, s1, F.phi, (float *)nullptr, dtau);
			hipLaunchKernelGGL((kReinitStage<D, 2, T, T, T>), gd, bd, 0, s, F.cell, (const T *)s1, (T *)nullptr, F.phi, (float *)nullptr, dtau);
		}
	} else {
		float *s0 = reinterpret_cast<float *>(F.scratch[0]), *s1 = reinterpret_cast<float *>(F.scratch[1]), *p32 = phi32Of(F);
		phiToFloat(F, s);
		if constexpr (D == 3) {
			const dim3 gz = gridDimsZ(F.cell);
			for (int it = 0; it < iterations; it++) {
				hipLaunchKernelGGL(kReinitStageZ<0>, gz, bd, 0, s, F.cell, (const float *)p32, s0, F.phi, p32, dtau);
				hipLaunchKernelGGL(kReinitStageZ<1>, gz, bd, 0, s, F.cell, (const float *)s0, s1, F.phi, p32, dtau);
				hipLaunchKernelGGL(kReinitStageZ<2>, gz, bd, 0, s, F.cell, (const float *)s1, (float *)nullptr, F.phi, p32, dtau);
			}
		} else {
			for (int it = 0; it < iterations; it++) {
				hipLaunchKernelGGL((kReinitStage<D, 0, float, float, float>), gd, bd, 0, s, F.cell, (const float *)p32, s0, F.phi, p32, dtau);
				hipLaunchKernelGGL((kReinitStage<D, 1, float, float, float>), gd, bd, 0, s, F.cell, (const float *)s0, s1, F.phi, p32, dtau);
				hipLaunchKernelGGL((kReinitStage<D, 2, float, float, float>), gd, bd, 0, s, F.cell, (const float *)s1, (float *)nullptr, F.phi, p32, dtau);
			}
		}
	}
}

void launchReinitPrepare(const Fields &F, hipStream_t s, bool fp32)
{
	if (fp32) phiToFloat(F, s);
}

void launchReinitStage(const Fields &F, double dtau, hipStream_t s, bool fp32)
{
	const dim3 gd = gridDims(F.cell, F.D), bd = gridBlock();
	const double *phi = F.phi;
	if (fp32) {
		float *s0 = reinterpret_cast<float *>(F.scratch[0]), *p32 = phi32Of(F);
		if (F.D == 2) hipLaunchKernelGGL((kReinitStage<2, 0, float, float, float>), gd, bd, 0, s, F.cell, (const float *)p32, s0, F.phi, p32, dtau);
		else hipLaunchKernelGGL(kReinitStageZ<0>, gridDimsZ(F.cell), bd, 0, s, F.cell, (const float *)p32, s0, F.phi, p32, dtau);
	} else {
		if (F.D == 2) hipLaunchKernelGGL((kReinitStage<2, 0, double, double, double>), gd, bd, 0, s, F.cell, phi, F.scratch[0], F.phi, (float *)nullptr, dtau);
		else hipLaunchKernelGGL((kReinitStage<3, 0, double, double, double>), gd, bd, 0, s, F.cell, phi, F.scratch[0], F.phi, (float *)nullptr, dtau);
	}
}

void launchReinit(const Fields &F, int iterations, double cfl, hipStream_t s, bool fp32)
{
	if (F.D == 2) { if (fp32) reinitD<2, float>(F, iterations, cfl, s); else reinitD<2, double>(F, iterations, cfl, s); }
	else { if (fp32) reinitD<3, float>(F, iterations, cfl, s); else reinitD<3, double>(F, iterations, cfl, s); }
}

} // namespace garm
