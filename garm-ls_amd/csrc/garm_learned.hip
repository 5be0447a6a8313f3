// Learned pressure solve and the network's C-ABI (no reference counterpart;
// the reference solves with Eigen's IC-PCG, LinearSystem.h:10-39).
//
// Flexible conjugate gradient (Notay's FCG with one direction kept):
//   alpha = r.p / p.Ap,  x += alpha p,  r -= alpha Ap,
//   z = M(r),  beta = -z.Ap / p.Ap,  p = z + beta p,
// which stays a descent method for a nonlinear preconditioner M.  M is a
// two-level cycle with the U-Net as the coarse correction:
//   z0 = w D^-1 r                      (damped Jacobi pre-smoothing)
//   z1 = z0 + N(r - A z0)              (network on the smoothed residual)
//   z  = z1 + w D^-1 (r - A z1)        (post-smoothing)
// with w = 2/3 and N the U-Net applied on the bounding box of the unknowns
// (two cells of context, sizes padded to multiples of four) to its input
// scaled to unit RMS, the output scaled back.  The smoother removes the rough
// error components a network trained on pressures does not resolve.  When the
// network stops paying (the A-norm error no longer halving per iteration)
// or the iteration budget runs out, the solve falls back to the mixed-precision
// Jacobi PCG on the current residual, or continues with the multigrid
// V-cycle; the FP64 residual b - A x judges the result either
// way.  While the bounding box of the unknowns is not much larger than their
// count, the Krylov vectors live on that box (the multigrid's level-0
// layout) with the FP64 operator on the same box; otherwise they are compact.
#include "garm_learned.hpp"
#include "garm_multigrid.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <functional>
#include <random>
#include <vector>

namespace garm {

constexpr int kLGrid = 1024;

static int lblocks(int n) { return std::max(1, std::min(kLGrid, (n + kBlock - 1) / kBlock)); }

template <int NV>
__device__ __forceinline__ void sumPart(const double *partials, int nb, double (&tot)[NV])
{
	double s[NV];
#pragma unroll
	for (int q = 0; q < NV; q++) s[q] = 0;
	for (int i = threadIdx.x; i < nb; i += blockDim.x)
#pragma unroll
		for (int q = 0; q < NV; q++) s[q] += partials[NV * i + q];
	blockSum<NV>(s, tot);
}

template <int NV>
__device__ __forceinline__ void bcast(double (&v)[NV])
{
	__shared__ double b[NV];
	__syncthreads();
	if (threadIdx.x == 0)
#pragma unroll
		for (int q = 0; q < NV; q++) b[q] = v[q];
	__syncthreads();
#pragma unroll
	for (int q = 0; q < NV; q++) v[q] = b[q];
}

template <int NV>
__device__ __forceinline__ void putPart(double (&v)[NV], double *partials)
{
	double blk[NV];
	blockSum<NV>(v, blk);
	if (threadIdx.x == 0)
#pragma unroll
		for (int q = 0; q < NV; q++) partials[NV * blockIdx.x + q] = blk[q];
}

enum { kLsPq = 0, kLsRr = 1, kLsBb = 2, kLsTrue = 3, kLsEnergy = 4, kLsSmooth = 5, kLsTrainIn = 6, kLsGamma = 7, kLsTrainOut = 8,
	kLsWords = 16 };

constexpr double kOmega = 2.0 / 3.0;
// a preconditioner stalls when the residual norm has not fallen by this
// factor over its last window of iterations
constexpr double kMinReduction = 0.5;

// z = w D^-1 r
__global__ void kJacobi(const double *invdiag, int n, const double *r, double *z)
{
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) z[i] = kOmega * invdiag[i] * r[i];
}

// out = z + w D^-1 (r - A z)
template <int D>
__global__ __launch_bounds__(kBlock) void kPostSmooth(Rows<double> A, int n, const double *r, const double *z, double *out)
{
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
		const double zi = z[i];
		double s = A.diag[i] * zi;
#pragma unroll
		for (int o = 0; o < 2 * D; o++) {
			const int c = A.nb[o][i];
			if (c >= 0) s += -A.w[o][i] * z[c];
		}
		out[i] = zi + kOmega * A.invdiag[i] * (r[i] - s);
	}
}

// q = A p; partials of p.q and r.p
template <int D>
__global__ __launch_bounds__(kBlock) void kFcgSpmv(Rows<double> A, int n, const double *p, double *q, const double *r, double *pa)
{
	double v[2] = {0, 0};
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
		const double pi = p[i];
		double s = A.diag[i] * pi;
#pragma unroll
		for (int o = 0; o < 2 * D; o++) {
			const int c = A.nb[o][i];
			if (c >= 0) s += -A.w[o][i] * p[c];
		}
		q[i] = s;
		v[0] += pi * s;
		v[1] += r[i] * pi;
	}
	putPart<2>(v, pa);
}

// alpha = (r.p) / (p.q); x += alpha p; r -= alpha q; partial |r|^2.  With
// b32 (box layout, a V-cycle next), the new residual also goes to the
// cycle's level-0 right-hand side and its iterate to 0 (the cycle's load).
__global__ __launch_bounds__(kBlock) void kFcgUpdate(int n, double *x, double *r, const double *p, const double *q, const double *pa, double *pb, double *ls,
	float *b32, float *x32)
{
	double t[2];
	sumPart<2>(pa, gridDim.x, t);
	double a[1] = {0};
	if (threadIdx.x == 0) a[0] = t[0] != 0 ? t[1] / t[0] : 0.0;
	bcast<1>(a);
	// the step's decrease of the A-norm error energy e^T A e: (r.p)^2 / (p.Ap)
	if (blockIdx.x == 0 && threadIdx.x == 0) { ls[kLsPq] = t[0]; ls[kLsEnergy] = t[0] != 0 ? t[1] * t[1] / t[0] : 0.0; }
	double v[1] = {0};
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
		x[i] += a[0] * p[i];
		const double ri = r[i] - a[0] * q[i];
		r[i] = ri;
		if (b32) {
			b32[i] = float(ri);
			x32[i] = 0.0f;
		}
		v[0] += ri * ri;
	}
	putPart<1>(v, pb);
}

__global__ __launch_bounds__(kBlock) void kFcgReduce1(const double *pb, int nb, double *dst)
{
	double t[1];
	sumPart<1>(pb, nb, t);
	if (threadIdx.x == 0) *dst = t[0];
}

// partial z.q
__global__ __launch_bounds__(kBlock) void kFcgDot(int n, const double *z, const double *q, double *pa)
{
	double v[1] = {0};
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) v[0] += z[i] * q[i];
	putPart<1>(v, pa);
}

// beta = -(z.q) / (p.q); p = z + beta p
__global__ __launch_bounds__(kBlock) void kFcgP(int n, double *p, const double *z, const double *pa, const double *ls)
{
	double t[1];
	sumPart<1>(pa, gridDim.x, t);
	double b[1] = {0};
	if (threadIdx.x == 0) b[0] = ls[kLsPq] != 0 ? -t[0] / ls[kLsPq] : 0.0;
	bcast<1>(b);
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = z[i] + b[0] * p[i];
}

__global__ void kCopyD(double *dst, const double *src, int n)
{
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = src[i];
}

__global__ void kZeroD(double *dst, int n)
{
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = 0;
}

__global__ void kAddD(double *dst, const double *src, int n)
{
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] += src[i];
}

// rt = b - A x (FP64), partial |rt|^2
template <int D>
__global__ __launch_bounds__(kBlock) void kTrueResidual(Rows<double> A, int n, const double *b, const double *x, double *rt, double *pb)
{
	double v[1] = {0};
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
		double s = A.diag[i] * x[i];
#pragma unroll
		for (int o = 0; o < 2 * D; o++) {
			const int c = A.nb[o][i];
			if (c >= 0) s += -A.w[o][i] * x[c];
		}
		const double ri = b[i] - s;
		rt[i] = ri;
		v[0] += ri * ri;
	}
	putPart<1>(v, pb);
}

// bounding box of the unknowns' cells
__global__ __launch_bounds__(kBlock) void kBBox(const int *mat2grid, int n, int nx, int ny, int *box)
{
	int mn[3] = {INT_MAX, INT_MAX, INT_MAX}, mx[3] = {INT_MIN, INT_MIN, INT_MIN};
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
		const int c = mat2grid[i];
		const int x = c % nx, y = (c / nx) % ny, z = c / nx / ny;
		mn[0] = min(mn[0], x); mn[1] = min(mn[1], y); mn[2] = min(mn[2], z);
		mx[0] = max(mx[0], x); mx[1] = max(mx[1], y); mx[2] = max(mx[2], z);
	}
	// reduce in the workgroup (wave shuffles, then LDS), one atomic per corner component
	__shared__ int red[6][kBlock / kWave];
	const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
#pragma unroll
	for (int a = 0; a < 3; a++) {
		for (int off = kWave / 2; off > 0; off >>= 1) {
			mn[a] = min(mn[a], __shfl_xor(mn[a], off));
			mx[a] = max(mx[a], __shfl_xor(mx[a], off));
		}
		if (lane == 0) { red[a][wid] = mn[a]; red[3 + a][wid] = mx[a]; }
	}
	__syncthreads();
	if (threadIdx.x < 6) {
		int v = red[threadIdx.x][0];
		for (int w = 1; w < kBlock / kWave; w++) v = threadIdx.x < 3 ? min(v, red[threadIdx.x][w]) : max(v, red[threadIdx.x][w]);
		if (threadIdx.x < 3) atomicMin(&box[threadIdx.x], v);
		else atomicMax(&box[threadIdx.x], v);
	}
}

__global__ void kBBoxInit(int *box)
{
	if (threadIdx.x < 3) box[threadIdx.x] = INT_MAX;
	else if (threadIdx.x < 6) box[threadIdx.x] = INT_MIN;
}

// ---------------------------------------------------------- box layout
// The flexible solves keep their Krylov vectors on the bounding box of the
// unknowns (cell order (k ny + j) nx + i, the multigrid's level-0 layout,
// zero outside the unknowns) with the FP64 operator laid out on the same
// box: diagonal, inverse diagonal and the weights of each cell's -x, -y, -z
// faces.  A product then reads its neighbours at fixed offsets instead of
// through six index arrays.  A face on the box boundary has weight 0, so a
// missing neighbour contributes 0 x (a finite value); past either end of the
// box the neighbour reads as 0.
struct BoxOp {
	int n, sy, sz;               // cells, row and plane strides
	const double *diag, *idiag, *wx, *wy, *wz;
};

// box cell of every unknown, for a box with origin org and sizes bn
__global__ void kBoxIndex(const int *mat2grid, int n, int gx, int gy, int ox, int oy, int oz, int bx, int by, unsigned *boxOf)
{
	for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
		const int g = mat2grid[r];
		const int i = g % gx - ox, j = (g / gx) % gy - oy, k = g / (gx * gy) - oz;
		boxOf[r] = (unsigned(k) * unsigned(by) + unsigned(j)) * unsigned(bx) + unsigned(i);
	}
}

template <int D>
__global__ void kBoxOpBuild(Rows<double> A, int n, const unsigned *boxOf, double *diag, double *idiag, double *wx, double *wy, double *wz)
{
	for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
		const unsigned c = boxOf[r];
		diag[c] = A.diag[r];
		idiag[c] = A.invdiag[r];
		wx[c] = A.nb[0][r] >= 0 ? A.w[0][r] : 0.0;
		wy[c] = A.nb[2][r] >= 0 ? A.w[2][r] : 0.0;
		if constexpr (D == 3) wz[c] = A.nb[4][r] >= 0 ? A.w[4][r] : 0.0;
	}
}

__global__ void kScatterToBox(int n, const unsigned *boxOf, const double *src, double *dst)
{
	for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) dst[boxOf[r]] = src[r];
}

__global__ void kGatherFromBox(int n, const unsigned *boxOf, const double *src, double *dst)
{
	for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) dst[r] = src[boxOf[r]];
}

// (A p)_c on the box: diag p_c minus the face terms, in the compact rows'
// order -x, +x, -y, +y, -z, +z
template <int D>
__device__ __forceinline__ double boxApply(const BoxOp &O, const double *p, int c)
{
	const int last = O.n - 1;
	auto P = [&](int e) { return e >= 0 && e <= last ? p[e] : 0.0; };
	auto Wp = [&](const double *w, int e) { return e <= last ? w[e] : 0.0; };
	double s = O.diag[c] * p[c];
	s += -O.wx[c] * P(c - 1);
	s += -Wp(O.wx, c + 1) * P(c + 1);
	s += -O.wy[c] * P(c - O.sy);
	s += -Wp(O.wy, c + O.sy) * P(c + O.sy);
	if constexpr (D == 3) {
		s += -O.wz[c] * P(c - O.sz);
		s += -Wp(O.wz, c + O.sz) * P(c + O.sz);
	}
	return s;
}

// q = A p; partials of p.q and r.p
template <int D>
__global__ __launch_bounds__(kBlock) void kBoxSpmv(BoxOp O, const double *p, double *q, const double *r, double *pa)
{
	double v[2] = {0, 0};
	for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < O.n; c += gridDim.x * blockDim.x) {
		const double s = boxApply<D>(O, p, c);
		q[c] = s;
		v[0] += p[c] * s;
		v[1] += r[c] * p[c];
	}
	putPart<2>(v, pa);
}

// The V-cycle's direction update fused into the next product: p = z + beta p
// with z the cycle's level-0 iterate (its FP64 value is what a store of the
// cycle would write) and beta = -(z.q) / (p.q) from the cycle's nzq
// partials of z.q and the previous p.q, then q = A p and the partials of p.q
// and r.p -- kFcgP's and kBoxSpmv's values.  The old direction is read at the
// neighbours, so the new one goes to the other buffer.
template <int D>
__global__ __launch_bounds__(kBlock) void kBoxSpmvP(BoxOp O, const float *z, const double *pIn, double *pOut, double *q, const double *r,
	const double *zq, int nzq, const double *ls, double *pa)
{
	double t[1];
	sumPart<1>(zq, nzq, t);
	double b[1] = {0};
	if (threadIdx.x == 0) b[0] = ls[kLsPq] != 0 ? -t[0] / ls[kLsPq] : 0.0;
	bcast<1>(b);
	const double beta = b[0];
	auto pv = [&](int e) { return double(z[e]) + beta * pIn[e]; };
	const int last = O.n - 1;
	auto P = [&](int e) { return e >= 0 && e <= last ? pv(e) : 0.0; };
	auto Wp = [&](const double *w, int e) { return e <= last ? w[e] : 0.0; };
	double v[2] = {0, 0};
	for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < O.n; c += gridDim.x * blockDim.x) {
		const double pc = pv(c);
		double s = O.diag[c] * pc;
		s += -O.wx[c] * P(c - 1);
		s += -Wp(O.wx, c + 1) * P(c + 1);
		s += -O.wy[c] * P(c - O.sy);
		s += -Wp(O.wy, c + O.sy) * P(c + O.sy);
		if constexpr (D == 3) {
			s += -O.wz[c] * P(c - O.sz);
			s += -Wp(O.wz, c + O.sz) * P(c + O.sz);
		}
		pOut[c] = pc;
		q[c] = s;
		v[0] += pc * s;
		v[1] += r[c] * pc;
	}
	putPart<2>(v, pa);
}

// rt = b - A x, partial |rt|^2
template <int D>
__global__ __launch_bounds__(kBlock) void kBoxResidual(BoxOp O, const double *b, const double *x, double *rt, double *pb)
{
	double v[1] = {0};
	for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < O.n; c += gridDim.x * blockDim.x) {
		const double ri = b[c] - boxApply<D>(O, x, c);
		rt[c] = ri;
		v[0] += ri * ri;
	}
	putPart<1>(v, pb);
}

__global__ void kBoxJacobi(BoxOp O, const double *r, double *z)
{
	for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < O.n; c += gridDim.x * blockDim.x) z[c] = kOmega * O.idiag[c] * r[c];
}

// out = z + w D^-1 (r - A z)
template <int D>
__global__ __launch_bounds__(kBlock) void kBoxPostSmooth(BoxOp O, const double *r, const double *z, double *out)
{
	for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < O.n; c += gridDim.x * blockDim.x)
		out[c] = z[c] + kOmega * O.idiag[c] * (r[c] - boxApply<D>(O, z, c));
}

// The network's correction u enters with the step length that minimises
// the A-norm of the error left after the pre-smoothing: with t = r - A z0,
// gamma = (u.t) / (u.A u).  w = A u and the partials of u.t, u.w:
template <int D>
__global__ __launch_bounds__(kBlock) void kBoxCorrectionDots(BoxOp O, const double *u, const double *t, double *w, double *pa)
{
	double v[2] = {0, 0};
	for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < O.n; c += gridDim.x * blockDim.x) {
		const double s = boxApply<D>(O, u, c);
		w[c] = s;
		v[0] += u[c] * t[c];
		v[1] += u[c] * s;
	}
	putPart<2>(v, pa);
}

template <int D>
__global__ __launch_bounds__(kBlock) void kRowsCorrectionDots(Rows<double> A, int n, const double *u, const double *t, double *w, double *pa)
{
	double v[2] = {0, 0};
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
		double s = A.diag[i] * u[i];
#pragma unroll
		for (int o = 0; o < 2 * D; o++) {
			const int c = A.nb[o][i];
			if (c >= 0) s += -A.w[o][i] * u[c];
		}
		w[i] = s;
		v[0] += u[i] * t[i];
		v[1] += u[i] * s;
	}
	putPart<2>(v, pa);
}

// ls[kLsGamma] = (u.t) / (u.A u), 0 when u.A u is not positive
__global__ __launch_bounds__(kBlock) void kGamma(const double *pa, int nb, double *ls)
{
	double t[2];
	sumPart<2>(pa, nb, t);
	if (threadIdx.x == 0) ls[kLsGamma] = (t[1] > 0 && isfinite(t[0] / t[1])) ? t[0] / t[1] : 0.0;
}

// z = z0 + gamma u, then a Jacobi post-smoothing step, whose residual
// r - A z = t - gamma w needs no further product:
// out = z0 + gamma u + w D^-1 (t - gamma w)
__global__ void kCombine(int n, const double *idiag, const double *z0, const double *u, const double *t, const double *w, const double *ls,
	double *out)
{
	const double g = ls[kLsGamma];
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
		out[i] = z0[i] + g * u[i] + kOmega * idiag[i] * (t[i] - g * w[i]);
}

// ------------------------------------------------------------------ work
LearnedWork::~LearnedWork()
{
	auto fr = [](auto *&q) { if (q) { (void)hipFree(q); q = nullptr; } };
	fr(x); fr(r); fr(p); fr(q); fr(z); fr(t); fr(b2); fr(u); fr(w); fr(pa); fr(pb); fr(ls); fr(box);
	if (host) { (void)hipHostFree(host); host = nullptr; }
	if (ready) { (void)hipEventDestroy(ready); ready = nullptr; }
	if (bpool) { (void)hipFree(bpool); bpool = nullptr; }
	if (boxOf) { (void)hipFree(boxOf); boxOf = nullptr; }
}

bool LearnedWork::reserveBox(size_t cells, size_t unknowns, std::string &err)
{
	const size_t need = 13 * cells;
	if (need > bcap) {
		if (bpool) (void)hipFree(bpool);
		bpool = nullptr;
		const size_t want = need + need / 4;
		if (hipMalloc(&bpool, want * sizeof(double)) != hipSuccess) { err = "flexible-solver box allocation failed"; bcap = 0; return false; }
		bcap = want;
	}
	if (unknowns > ocap) {
		if (boxOf) (void)hipFree(boxOf);
		boxOf = nullptr;
		const size_t want = unknowns + unknowns / 4;
		if (hipMalloc(&boxOf, want * sizeof(unsigned)) != hipSuccess) { err = "flexible-solver box allocation failed"; ocap = 0; return false; }
		ocap = want;
	}
	return true;
}

bool LearnedWork::reserve(size_t n, std::string &err)
{
	bool ok = true;
	if (!ls) {
		ok = hipMalloc(&pa, 2 * kLGrid * sizeof(double)) == hipSuccess && hipMalloc(&pb, kLGrid * sizeof(double)) == hipSuccess
			&& hipMalloc(&ls, kLsWords * sizeof(double)) == hipSuccess && hipMalloc(&box, 8 * sizeof(int)) == hipSuccess
			&& hipHostMalloc(&host, kLsWords * sizeof(double)) == hipSuccess;
		if (!ok) { err = "learned-solver scalar allocation failed"; return false; }
	}
	if (!ready && hipEventCreateWithFlags(&ready, hipEventDisableTiming) != hipSuccess) { err = "learned-solver event creation failed"; return false; }
	if (n > cap) {
		auto fr = [](auto *&q) { if (q) { (void)hipFree(q); q = nullptr; } };
		fr(x); fr(r); fr(p); fr(q); fr(z); fr(t); fr(b2); fr(u); fr(w);
		const size_t want = std::max<size_t>(n + n / 4, 1024);
		for (double **v : {&x, &r, &p, &q, &z, &t, &b2, &u, &w}) ok = ok && hipMalloc(v, want * sizeof(double)) == hipSuccess;
		if (!ok) { err = "learned-solver vector allocation failed"; cap = 0; return false; }
		cap = want;
	}
	return true;
}

// The host polls an event behind the copy instead of blocking in a stream
// synchronisation: the solver reads a few scalars every iteration, and a
// blocking wait wakes the host tens of microseconds after the copy lands,
// time the GPU spends idle before the next iteration is launched.
static int readStart(LearnedWork &W, const double *src, int count, hipStream_t s, std::string &err)
{
	if (hipMemcpyAsync(W.host, src, count * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess || hipEventRecord(W.ready, s) != hipSuccess) {
		err = "learned-solver readback failed";
		return GARM_EHIP;
	}
	return GARM_OK;
}

static int readWait(LearnedWork &W, std::string &err)
{
	if (pollEvent(W.ready) != hipSuccess) { err = "learned-solver readback failed"; return GARM_EHIP; }
	return GARM_OK;
}

static int readHost(LearnedWork &W, const double *src, int count, hipStream_t s, std::string &err)
{
	int rc = readStart(W, src, count, s, err);
	return rc ? rc : readWait(W, err);
}

// Bounding box of the unknowns' cells (inclusive corners).
static int unknownBox(const int *mat2grid, int n, const GridDev &cell, LearnedWork &W, hipStream_t s, int lo[3], int hi[3], std::string &err)
{
	hipLaunchKernelGGL(kBBoxInit, dim3(1), dim3(64), 0, s, W.box);
	hipLaunchKernelGGL(kBBox, dim3(lblocks(n)), dim3(kBlock), 0, s, mat2grid, n, cell.n[0], cell.n[1], W.box);
	// six ints through the pinned readback buffer
	const int *hb = reinterpret_cast<const int *>(W.host);
	if (hipMemcpyAsync(W.host, W.box, 6 * sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess || hipEventRecord(W.ready, s) != hipSuccess) {
		err = "bounding-box readback failed";
		return GARM_EHIP;
	}
	if (int rc = readWait(W, err)) return rc;
	for (int a = 0; a < 3; a++) { lo[a] = hb[a]; hi[a] = hb[3 + a]; }
	return GARM_OK;
}

// The network's grid for the unknowns in [lo, hi]: their bounding box with
// two cells of context, sizes padded to multiples of four.
static UnetBox networkBox(int D, const int lo[3], const int hi[3], const int gn[3])
{
	UnetBox B{};
	for (int a = 0; a < 3; a++) {
		const bool live = a < D;
		const int l = live ? lo[a] - 2 : 0, h = live ? hi[a] + 2 : 0;
		B.org[a] = l;
		B.n[a] = live ? (h - l + 1 + 3) / 4 * 4 : 4;
		B.gn[a] = gn[a];
	}
	return B;
}

// A preconditioner application z = M(r) as an optional head and the rest:
// the head may be enqueued before the solver knows whether the application
// is needed (it then only writes scratch the next head overwrites).
struct Precond {
	std::function<void()> head, rest;
	// box layout, V-cycle: where the x/r update can write the cycle's
	// load itself, the head that then only sweeps, and the rest that forms
	// the partials of z.q instead of storing z (the next product forms the
	// direction from the level-0 iterate; each unset otherwise)
	std::function<void(float **, float **)> loadTargets;
	std::function<void()> headLoaded, restDot;
	std::function<const float *()> level0;   // the level-0 iterate restDot leaves (z, FP32)
	double bytes = 0;   // algorithmic HBM bytes of one application
	void full() const { if (head) head(); rest(); }
};

// Algorithmic HBM bytes of one flexible-CG iteration per vector element:
// product (p and the operator in, the residual for r.p, q out), x/r update,
// z.q, p update; on the box the operator is diag + three face weights (40 B),
// on compact rows diag + six (index, weight) pairs (80 B).
constexpr double kFcgBoxBytes = 56 + 48 + 16 + 24;
// with the V-cycle on the box: the direction formed inside the product
// (z 4, p in and out 16, the operator 32, r 8, q 8), x/r update
constexpr double kFcgBoxFusedBytes = 68 + 48;
constexpr double kFcgRowBytes = 96 + 48 + 16 + 24;
// One network application's vector work per element (the network's own
// activations excluded): Jacobi 24, pre-smoothed residual 56, zeroing the
// correction 8, gather 16 and scatter 20, A u with two dots 56, combination 48.
constexpr double kNetBoxBytes = 24 + 56 + 8 + 16 + 20 + 56 + 48;
constexpr double kNetRowBytes = kNetBoxBytes + 2 * 40;

// The Krylov vectors of a flexible solve and how A acts on them: compact
// (the unknowns' numbering, the FP64 rows), or on the bounding box of the
// unknowns (the multigrid's level-0 layout, the FP64 operator on the box),
// which a solve uses while the box holds at most kBoxMaxRatio cells per
// unknown: its products read neighbours at fixed offsets instead of through
// six index arrays, but every vector operation also covers the box's cells
// that are not unknowns (a splash's bounding box holds mostly air).
constexpr double kBoxMaxRatio = 1.5;

struct FcgSpace {
	bool box = false;
	int nv = 0;                        // vector length
	double *x = nullptr, *r = nullptr, *p = nullptr, *q = nullptr, *z = nullptr, *t = nullptr;
	double *u = nullptr, *w = nullptr; // network correction and A times it
	BoxOp O{};                         // box: the operator
	const unsigned *boxOf = nullptr;   // box: box cell of every unknown
};

static void compactSpace(LearnedWork &W, int n, FcgSpace &V)
{
	V.box = false;
	V.nv = n;
	V.x = W.x; V.r = W.r; V.p = W.p; V.q = W.q; V.z = W.z; V.t = W.t;
	V.u = W.u; V.w = W.w;
}

// Lays out the box of the unknowns (origin lo, sizes hi - lo + 1) in W:
// the operator from the FP64 rows and zeroed vectors.
template <int D>
static int boxSpace(const Rows<double> &A64, int n, const int *mat2grid, const int gn[3], const int lo[3], const int hi[3],
	LearnedWork &W, hipStream_t s, FcgSpace &V, std::string &err)
{
	int bn[3];
	for (int a = 0; a < 3; a++) bn[a] = a < D ? hi[a] - lo[a] + 1 : 1;
	const size_t cells = size_t(bn[0]) * bn[1] * bn[2];
	if (cells > size_t(INT_MAX)) { err = "flexible-solver box too large"; return GARM_EINVAL; }
	if (!W.reserveBox(cells, size_t(n), err)) return GARM_ENOMEM;
	const int N = int(cells);
	V.box = true;
	V.nv = N;
	double *q = W.bpool;
	double *arr[13];
	for (double *&e : arr) { e = q; q += cells; }
	V.O.n = N;
	V.O.sy = bn[0];
	V.O.sz = bn[0] * bn[1];
	V.O.diag = arr[0]; V.O.idiag = arr[1]; V.O.wx = arr[2]; V.O.wy = arr[3]; V.O.wz = arr[4];
	V.x = arr[5]; V.r = arr[6]; V.p = arr[7]; V.q = arr[8]; V.z = arr[9]; V.t = arr[10];
	V.u = arr[11]; V.w = arr[12];
	V.boxOf = W.boxOf;
	// the operator is zero outside the unknowns; the vectors are written
	// whole before they are read
	const size_t opWords = 5 * cells;
	if (opWords > size_t(INT_MAX)) { err = "flexible-solver box too large"; return GARM_EINVAL; }
	const dim3 gu(lblocks(n)), b1(kBlock);
	hipLaunchKernelGGL(kZeroD, dim3(unsigned(std::min<size_t>((opWords + kBlock - 1) / kBlock, 8192))), b1, 0, s, W.bpool, int(opWords));
	hipLaunchKernelGGL(kBoxIndex, gu, b1, 0, s, mat2grid, n, gn[0], gn[1], lo[0], lo[1], D == 3 ? lo[2] : 0, bn[0], bn[1], W.boxOf);
	hipLaunchKernelGGL(kBoxOpBuild<D>, gu, b1, 0, s, A64, n, (const unsigned *)W.boxOf, arr[0], arr[1], arr[2], arr[3], arr[4]);
	return GARM_OK;
}

// The V-cycle's hooks for the fused box-layout iteration (flexibleCg).
static void fuseCycle(Precond &P, Multigrid &M, int n, FcgSpace &V, LearnedWork &W, hipStream_t s)
{
	if (!V.box) return;
	P.loadTargets = [&M](float **b32, float **x32) { mgLevel0(M, b32, x32); };
	P.headLoaded = [&M, n, &V, s]() { mgApplyHead(M, n, V.r, s, true, true); };
	P.restDot = [&M, n, &V, &W, s]() { mgApplyRest(M, n, nullptr, s, true, V.q, W.pb, lblocks(V.nv)); };
	P.level0 = [&M]() { float *b32, *x32; mgLevel0(M, &b32, &x32); return (const float *)x32; };
}

// Flexible CG on the FP64 system, on the vectors of V, with preconditioner
// `precondition` (writes V.z = M(V.r)).  Stall rules (CG minimises the
// error in the A-norm, not the residual, so a single iteration may leave
// the residual where it was):
//  - with `energyRule`, the first preconditioner is judged by the A-norm:
//    each iteration lowers the error energy e^T A e by (r.p)^2 / (p.Ap),
//    and the decrease of an iteration must be at most kEnergyRatio of the
//    decrease of the iteration before (the A-norm error halving per
//    iteration, as a V-cycle does and more), from the second iteration on;
//  - otherwise, and for the second preconditioner, the residual norm must
//    halve over every kStallWindow iterations.
// After the loop the solution goes to the compact W.x, the FP64 residual
// b - A x of the compact rows judges it, and what is left goes to the
// mixed-precision Jacobi solver (solveJacobi: on the box of the unknowns
// where that form applies).
constexpr int kStallWindow = 10;
constexpr int kEnergyWindow = 1;
constexpr double kEnergyRatio = 1.0 / 4.0;

template <int D>
static int flexibleCg(const Rows<double> &A64, const Rows<float> &A32, const double *b, int n, FcgSpace &V, const BoxLayout &L,
	const garm_pcg_opts &o, Krylov &K, BoxKrylov &BK, LearnedWork &W, hipStream_t s, garm_pcg_stats *st, std::string &err, const Precond &first, bool energyRule,
	int firstBudget, const Precond *second, const char *what)
{
	const Precond *precondition = &first;
	const int budget = second ? o.max_iterations : firstBudget;
	int switchedAt = -1;
	const int N = V.nv;
	const int nb = lblocks(n), nbb = lblocks(N);
	const dim3 g1(nb), gb(nbb), b1(kBlock);
	int rc;
	// x = 0, r = b, |b|^2
	hipLaunchKernelGGL(kZeroD, gb, b1, 0, s, V.x, N);
	if (V.box) {
		hipLaunchKernelGGL(kZeroD, gb, b1, 0, s, V.r, N);
		hipLaunchKernelGGL(kScatterToBox, g1, b1, 0, s, n, V.boxOf, b, V.r);
	} else hipLaunchKernelGGL(kCopyD, g1, b1, 0, s, V.r, b, n);
	hipLaunchKernelGGL(kFcgDot, g1, b1, 0, s, n, b, b, W.pb);
	hipLaunchKernelGGL(kFcgReduce1, dim3(1), b1, 0, s, (const double *)W.pb, nb, W.ls + kLsBb);
	// |b|^2 is read once the first iteration is enqueued (everything before
	// the first convergence check is finite for b = 0: the updates' step
	// lengths are 0 where p.Ap is, and a zero right-hand side returns x = 0)
	if ((rc = readStart(W, W.ls, 4, s, err))) return rc;
	st->iterations = 0;
	st->outer_iterations = 0;
	st->fell_back = 0;
	st->preconditioner_applications = 0;
	st->hbm_bytes = 0;
	st->solver_cells = N;
	double bb = -1, tol = 0, thr = 0;
	std::vector<double> hist, energy;
	auto readNorm = [&]() -> int {
		if (int e = readWait(W, err)) return e;
		bb = W.host[kLsBb];
		tol = std::max(o.tolerance, 2.220446049250313e-16 / std::sqrt(bb));
		thr = tol * tol * bb;
		hist.assign(1, bb);
		return GARM_OK;
	};
	precondition->full();
	hipLaunchKernelGGL(kCopyD, gb, b1, 0, s, V.p, (const double *)V.z, N);
	int it = 0, fusedIts = 0;
	// p = z + beta p still to be formed, from the V-cycle's level-0 iterate
	// and the z.q partials in W.pb, inside the next product
	const float *zPending = nullptr;
	double *p = V.p;
	for (; it < budget; it++) {
		if (zPending) {
			// the old direction and the new one alternate between p and z
			// (z itself is no longer stored once the fused cycle runs)
			double *pNew = p == V.p ? V.z : V.p;
			hipLaunchKernelGGL(kBoxSpmvP<D>, gb, b1, 0, s, V.O, zPending, (const double *)p, pNew, V.q, (const double *)V.r,
				(const double *)W.pb, nbb, (const double *)W.ls, W.pa);
			p = pNew;
			zPending = nullptr;
			fusedIts++;
		} else if (V.box) hipLaunchKernelGGL(kBoxSpmv<D>, gb, b1, 0, s, V.O, (const double *)p, V.q, (const double *)V.r, W.pa);
		else hipLaunchKernelGGL(kFcgSpmv<D>, gb, b1, 0, s, A64, n, (const double *)p, V.q, (const double *)V.r, W.pa);
		// a V-cycle on the box takes its load from the update and forms z.q
		// in its store (the same values as its own load and the dot product)
		const Precond *fused = V.box && precondition->loadTargets ? precondition : nullptr;
		float *b32 = nullptr, *x32 = nullptr;
		if (fused) fused->loadTargets(&b32, &x32);
		hipLaunchKernelGGL(kFcgUpdate, gb, b1, 0, s, N, V.x, V.r, (const double *)p, (const double *)V.q, (const double *)W.pa, W.pb, W.ls,
			b32, x32);
		hipLaunchKernelGGL(kFcgReduce1, dim3(1), b1, 0, s, (const double *)W.pb, nbb, W.ls + kLsRr);
		if (bb < 0) {
			// the norm's readback shares the slot with the iteration's
			if ((rc = readNorm())) return rc;
			if (bb == 0) break;
		}
		if ((rc = readStart(W, W.ls, kLsEnergy + 1, s, err))) return rc;
		// the next application's head keeps the GPU busy while the host
		// reads the residual and decides
		if (fused) fused->headLoaded();
		else if (precondition->head) precondition->head();
		if ((rc = readWait(W, err))) return rc;
		const double rr = W.host[kLsRr];
		if (!std::isfinite(rr)) break;
		hist.push_back(rr);
		energy.push_back(W.host[kLsEnergy]);
		if (rr < thr) { it++; break; }
		const bool onFirst = switchedAt < 0;
		bool stalled;
		if (onFirst && energyRule)
			stalled = energy.size() > size_t(kEnergyWindow) && energy.back() > kEnergyRatio * energy[energy.size() - 1 - kEnergyWindow];
		else {
			const size_t from = onFirst ? 0 : size_t(switchedAt) + 1;
			stalled = hist.size() > from + kStallWindow && rr > kMinReduction * kMinReduction * hist[hist.size() - 1 - kStallWindow];
		}
		if (stalled || (onFirst && it + 1 >= firstBudget)) {
			if (!onFirst || !second) { it++; break; }
			// the first preconditioner has stopped paying: the same iteration
			// continues with the second one (whole: the head already enqueued
			// was the first one's)
			precondition = second;
			switchedAt = int(hist.size()) - 1;
			st->fell_back = 1;
			precondition->full();
			fused = nullptr;
		} else if (fused) fused->restDot();
		else precondition->rest();
		if (fused) zPending = fused->level0();
		else {
			hipLaunchKernelGGL(kFcgDot, gb, b1, 0, s, N, (const double *)V.z, (const double *)V.q, W.pa);
			hipLaunchKernelGGL(kFcgP, gb, b1, 0, s, N, p, (const double *)V.z, (const double *)W.pa, (const double *)W.ls);
		}
	}
	if (bb < 0 && (rc = readNorm())) return rc;
	if (bb == 0) {
		hipLaunchKernelGGL(kZeroD, g1, b1, 0, s, W.x, n);
		st->relative_residual = 0;
		return GARM_OK;
	}
	st->iterations = it;
	st->preconditioner_applications = switchedAt < 0 ? it : switchedAt;
	{
		// the application before the loop and one per iteration but the last
		const int apps = std::max(it, 1);
		const int firstApps = switchedAt < 0 ? apps : switchedAt + 1;
		st->hbm_bytes = (double(it - fusedIts) * (V.box ? kFcgBoxBytes : kFcgRowBytes) + double(fusedIts) * kFcgBoxFusedBytes) * double(N) +
			firstApps * first.bytes +
			(second ? double(apps - firstApps) * second->bytes : 0.0);
	}
	// the solution in the compact unknowns; the FP64 residual judges it, and
	// plain PCG takes what is left
	if (V.box) hipLaunchKernelGGL(kGatherFromBox, g1, b1, 0, s, n, V.boxOf, (const double *)V.x, W.x);
	if (!std::isfinite(W.host[kLsRr])) hipLaunchKernelGGL(kZeroD, g1, b1, 0, s, W.x, n);
	hipLaunchKernelGGL(kTrueResidual<D>, g1, b1, 0, s, A64, n, b, (const double *)W.x, W.b2, W.pb);
	hipLaunchKernelGGL(kFcgReduce1, dim3(1), b1, 0, s, (const double *)W.pb, nb, W.ls + kLsTrue);
	if ((rc = readHost(W, W.ls, 4, s, err))) return rc;
	double rt = W.host[kLsTrue];
	if (!(rt < thr)) {
		st->fell_back = 1;
		garm_pcg_opts fo = o;
		fo.mode = GARM_PCG_MIXED;
		fo.tolerance = std::sqrt(thr / rt) * 0.999;
		fo.max_iterations = std::max(1, o.max_iterations - it);
		garm_pcg_stats fs{};
		bool boxed;
		rc = solveJacobi(D, true, A64, A32, W.b2, n, L, fo, K, BK, s, &fs, err, &boxed);
		if (rc != GARM_OK && rc != GARM_ENOCONV) return rc;
		hipLaunchKernelGGL(kAddD, g1, b1, 0, s, W.x, (const double *)K.x, n);
		st->iterations += fs.iterations;
		st->outer_iterations = fs.outer_iterations;
		st->hbm_bytes += fs.hbm_bytes;
		hipLaunchKernelGGL(kTrueResidual<D>, g1, b1, 0, s, A64, n, b, (const double *)W.x, W.b2, W.pb);
		hipLaunchKernelGGL(kFcgReduce1, dim3(1), b1, 0, s, (const double *)W.pb, nb, W.ls + kLsTrue);
		if ((rc = readHost(W, W.ls, 4, s, err))) return rc;
		rt = W.host[kLsTrue];
	}
	st->relative_residual = std::sqrt(rt / bb);
	if (!(rt < thr) && st->relative_residual > tol) { err = std::string(what) + " and its fallback did not reach the tolerance"; return GARM_ENOCONV; }
	return GARM_OK;
}

template <int D>
static int learned(const Rows<double> &A64, const Rows<float> &A32, const double *b, int n, const int *grid2mat, const BoxLayout &L,
	const garm_pcg_opts &o, Krylov &K, BoxKrylov &BK, LearnedWork &W, Unet &U, Multigrid &M, hipStream_t s, garm_pcg_stats *st,
	std::string &err)
{
	if (!W.reserve(size_t(n), err)) return GARM_ENOMEM;
	const dim3 b1(kBlock);
	const int *mat2grid = L.mat2grid;
	int rc, lo[3], hi[3];
	for (int a = 0; a < 3; a++) { lo[a] = L.lo[a]; hi[a] = L.lo[a] + L.n[a] - 1; }
	const UnetBox B = networkBox(D, lo, hi, L.gn);
	const size_t boxVox = size_t(B.n[0]) * B.n[1] * B.n[2];
	// the buffers are reallocated only when the box outgrows them, then for
	// inference only (no training gradients) and with half again as much room
	if (boxVox > U.buf.cap_vox && !U.buf.reserve(boxVox + boxVox / 2, false, err)) return GARM_ENOMEM;
	if (!U.buf.setDims(B.n[0], B.n[1], B.n[2], err)) return GARM_ENOMEM;
	FcgSpace V;
	const size_t boxCells = size_t(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (D == 3 ? hi[2] - lo[2] + 1 : 1);
	if (double(boxCells) <= kBoxMaxRatio * n) {
		if ((rc = boxSpace<D>(A64, n, mat2grid, L.gn, lo, hi, W, s, V, err))) return rc;
	} else compactSpace(W, n, V);
	VecBox VB{};
	for (int a = 0; a < 3; a++) {
		VB.org[a] = a < D ? lo[a] : 0;
		VB.n[a] = a < D ? hi[a] - lo[a] + 1 : 1;
	}
	const dim3 gb(lblocks(V.nv));
	const int nbb = lblocks(V.nv);
	// z = M(r): two-level cycle, network on the pre-smoothed residual, its
	// correction scaled to minimise the A-norm of the error
	Precond net;
	net.rest = [&]() {
		if (V.box) {
			hipLaunchKernelGGL(kBoxJacobi, gb, b1, 0, s, V.O, (const double *)V.r, V.z);
			hipLaunchKernelGGL(kBoxResidual<D>, gb, b1, 0, s, V.O, (const double *)V.r, (const double *)V.z, V.t, W.pb);
		} else {
			hipLaunchKernelGGL(kJacobi, gb, b1, 0, s, (const double *)A64.invdiag, n, (const double *)V.r, V.z);
			hipLaunchKernelGGL(kTrueResidual<D>, gb, b1, 0, s, A64, n, (const double *)V.r, (const double *)V.z, V.t, W.pb);
		}
		hipLaunchKernelGGL(kFcgReduce1, dim3(1), b1, 0, s, (const double *)W.pb, nbb, W.ls + kLsSmooth);
		hipLaunchKernelGGL(kZeroD, gb, b1, 0, s, V.u, V.nv);
		if (V.box) {
			unetApplyVecBox(U, B, VB, grid2mat, V.t, W.ls + kLsSmooth, n, V.u, s);
			hipLaunchKernelGGL(kBoxCorrectionDots<D>, gb, b1, 0, s, V.O, (const double *)V.u, (const double *)V.t, V.w, W.pa);
		} else {
			unetApplyCompact(U, B, grid2mat, V.t, W.ls + kLsSmooth, n, V.u, s);
			hipLaunchKernelGGL(kRowsCorrectionDots<D>, gb, b1, 0, s, A64, n, (const double *)V.u, (const double *)V.t, V.w, W.pa);
		}
		hipLaunchKernelGGL(kGamma, dim3(1), b1, 0, s, (const double *)W.pa, nbb, W.ls);
		hipLaunchKernelGGL(kCombine, gb, b1, 0, s, V.nv, V.box ? V.O.idiag : (const double *)A64.invdiag, (const double *)V.z,
			(const double *)V.u, (const double *)V.t, (const double *)V.w, (const double *)W.ls, V.t);
		std::swap(V.z, V.t);
	};
	net.bytes = (V.box ? kNetBoxBytes : kNetRowBytes) * double(V.nv);
	Precond mg;
	const Precond *second = nullptr;
	if (o.fallback_mode == GARM_PCG_MULTIGRID) {
		int org[3], dims[3], gn[3];
		for (int a = 0; a < 3; a++) {
			const bool live = a < D;
			org[a] = live ? lo[a] : 0;
			dims[a] = live ? hi[a] - lo[a] + 1 : 1;
			gn[a] = L.gn[a];
		}
		if (!M.layout(D, org, dims, gn, n, err)) return GARM_ENOMEM;
		mgBuild(M, A32, n, mat2grid, s);
		mg.head = [&]() { mgApplyHead(M, n, V.r, s, V.box); };
		mg.rest = [&]() { mgApplyRest(M, n, V.z, s, V.box); };
		fuseCycle(mg, M, n, V, W, s);
		mg.bytes = mgCycleBytes(M);
		second = &mg;
	}
	return flexibleCg<D>(A64, A32, b, n, V, L, o, K, BK, W, s, st, err, net, true, std::min(o.max_iterations, 400), second, "learned solve");
}

template <int D>
static int multigrid(const Rows<double> &A64, const Rows<float> &A32, const double *b, int n, const BoxLayout &L, const garm_pcg_opts &o,
	Krylov &K, BoxKrylov &BK, LearnedWork &W, Multigrid &M, hipStream_t s, garm_pcg_stats *st, std::string &err)
{
	if (!W.reserve(size_t(n), err)) return GARM_ENOMEM;
	const int *mat2grid = L.mat2grid;
	int rc, lo[3], hi[3];
	for (int a = 0; a < 3; a++) { lo[a] = L.lo[a]; hi[a] = L.lo[a] + L.n[a] - 1; }
	int org[3], dims[3], gn[3];
	for (int a = 0; a < 3; a++) {
		const bool live = a < D;
		org[a] = live ? lo[a] : 0;
		dims[a] = live ? hi[a] - lo[a] + 1 : 1;
		gn[a] = L.gn[a];
	}
	if (!M.layout(D, org, dims, gn, n, err)) return GARM_ENOMEM;
	mgBuild(M, A32, n, mat2grid, s);
	FcgSpace V;
	if (double(M.cells[0]) <= kBoxMaxRatio * n) {
		if ((rc = boxSpace<D>(A64, n, mat2grid, L.gn, lo, hi, W, s, V, err))) return rc;
	} else compactSpace(W, n, V);
	Precond mg;
	mg.head = [&]() { mgApplyHead(M, n, V.r, s, V.box); };
	mg.rest = [&]() { mgApplyRest(M, n, V.z, s, V.box); };
	fuseCycle(mg, M, n, V, W, s);
	mg.bytes = mgCycleBytes(M);
	return flexibleCg<D>(A64, A32, b, n, V, L, o, K, BK, W, s, st, err, mg, false, o.max_iterations, nullptr, "multigrid solve");
}

int solveLearned(int D, const Rows<double> &A64, const Rows<float> &A32, const double *b, int n, const int *grid2mat, const BoxLayout &L,
	const garm_pcg_opts &o, Krylov &K, BoxKrylov &BK, LearnedWork &W, Unet &U, Multigrid &M, hipStream_t s, garm_pcg_stats *st,
	std::string &err)
{
	return D == 2 ? learned<2>(A64, A32, b, n, grid2mat, L, o, K, BK, W, U, M, s, st, err)
	              : learned<3>(A64, A32, b, n, grid2mat, L, o, K, BK, W, U, M, s, st, err);
}

int solveMultigrid(int D, const Rows<double> &A64, const Rows<float> &A32, const double *b, int n, const BoxLayout &L,
	const garm_pcg_opts &o, Krylov &K, BoxKrylov &BK, LearnedWork &W, Multigrid &M, hipStream_t s, garm_pcg_stats *st, std::string &err)
{
	return D == 2 ? multigrid<2>(A64, A32, b, n, L, o, K, BK, W, M, s, st, err) : multigrid<3>(A64, A32, b, n, L, o, K, BK, W, M, s, st, err);
}

// ------------------------------------------------------------- training
__global__ __launch_bounds__(kBlock) void kLossGrad(const float *y, const float *target, const uint8_t *mask, float scale, float invCount,
	float *gy, size_t vox, float *lossPart)
{
	float acc = 0;
	for (size_t v = blockIdx.x * size_t(blockDim.x) + threadIdx.x; v < vox; v += size_t(gridDim.x) * blockDim.x) {
		const float d = mask[v] ? y[v] - target[v] * scale : 0.0f;
		gy[v] = 2.0f * d * invCount;
		acc += d * d;
	}
	double a[1] = {acc};
	double o[1];
	blockSum<1>(a, o);
	if (threadIdx.x == 0) lossPart[blockIdx.x] = float(o[0]);
}

__global__ void kLossReduce(const float *part, int nb, float invCount, float *dst)
{
	if (threadIdx.x != 0) return;
	float s = 0;
	for (int i = 0; i < nb; i++) s += part[i];
	*dst = s * invCount;
}

__global__ void kLoadScaled(const float *r, const uint8_t *mask, float scale, bf16 *in, size_t vox)
{
	for (size_t v = blockIdx.x * size_t(blockDim.x) + threadIdx.x; v < vox; v += size_t(gridDim.x) * blockDim.x) {
		const bool m = mask[v] != 0;
		in[v * 4 + 0] = f2bf(m ? r[v] * scale : 0.0f);
		in[v * 4 + 1] = f2bf(m ? 1.0f : 0.0f);
		in[v * 4 + 2] = 0;
		in[v * 4 + 3] = 0;
	}
}

// -------------------------------------------- training on a solve's system
// Pairs drawn from the pressure system of the last projection, made on the
// device: (A e, e) for random smooth e on the unknowns -- the residual and
// error of the smooth error component the Jacobi smoothing of the
// preconditioner leaves to the network -- and (b, x), the projection's own
// right-hand side and solution, with a random sign.
__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
	z += 0x9e3779b97f4a7c15ull;
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
	return z ^ (z >> 31);
}

__device__ __forceinline__ double unit01(uint64_t h) { return double(h >> 11) * 0x1.0p-53; }

constexpr int kModes = 12;
// The network sees about 20 cells around a voxel (two 3x3x3 convolutions per
// level at spacings 1, 2 and 4, going down and up), so it cannot invert the
// operator on wavelengths much longer than twice that; the operator pairs
// stop there (with the unknowns' whole extent the loss stayed near 1 and the
// trained network slowed the flexible solve down).
constexpr double kMaxWavelength = 32.0;

// modes[m] = (kx, ky, kz, phase, amplitude): wavelengths log-uniform in
// [4, L] cells, directions uniform (on the sphere in 3D, the circle in 2D),
// amplitude (lambda / L)^slope with slope uniform in [0.5, 2]
__global__ void kSmoothModes(int D, uint64_t key, double L, double *modes)
{
	const int m = int(threadIdx.x);
	if (m >= kModes) return;
	const double slope = 0.5 + 1.5 * unit01(mix64(key));
	const uint64_t h = mix64(key ^ (0x632be59bd9b4e019ull * uint64_t(m + 1)));
	const double twoPi = 6.283185307179586;
	const double lam = exp(log(4.0) + unit01(mix64(h + 1)) * (log(fmax(L, 4.0)) - log(4.0)));
	double d[3];
	const double az = twoPi * unit01(mix64(h + 2));
	if (D == 3) {
		const double z = 2.0 * unit01(mix64(h + 3)) - 1.0, rr = sqrt(fmax(0.0, 1.0 - z * z));
		d[0] = rr * cos(az); d[1] = rr * sin(az); d[2] = z;
	} else {
		d[0] = cos(az); d[1] = sin(az); d[2] = 0;
	}
	const double k = twoPi / lam;
	modes[5 * m + 0] = k * d[0];
	modes[5 * m + 1] = k * d[1];
	modes[5 * m + 2] = k * d[2];
	modes[5 * m + 3] = twoPi * unit01(mix64(h + 4));
	modes[5 * m + 4] = pow(lam / fmax(L, 4.0), slope);
}

__global__ __launch_bounds__(kBlock) void kSmoothField(const int *mat2grid, int n, int nx, int ny, const double *modes, double *e)
{
	__shared__ double md[5 * kModes];
	for (int q = int(threadIdx.x); q < 5 * kModes; q += blockDim.x) md[q] = modes[q];
	__syncthreads();
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
		const int c = mat2grid[i];
		const double x = double(c % nx), y = double((c / nx) % ny), z = double(c / nx / ny);
		double v = 0;
#pragma unroll
		for (int m = 0; m < kModes; m++) v += md[5 * m + 4] * sin(md[5 * m] * x + md[5 * m + 1] * y + md[5 * m + 2] * z + md[5 * m + 3]);
		e[i] = v;
	}
}

// r = A e; partials of |r|^2
template <int D>
__global__ __launch_bounds__(kBlock) void kOperatorPair(Rows<double> A, int n, const double *e, double *r, double *pb)
{
	double v[1] = {0};
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
		double t = A.diag[i] * e[i];
#pragma unroll
		for (int o = 0; o < 2 * D; o++) {
			const int c = A.nb[o][i];
			if (c >= 0) t += -A.w[o][i] * e[c];
		}
		r[i] = t;
		v[0] += t * t;
	}
	putPart<1>(v, pb);
}

__global__ void kSignedCopies(int n, double sign, const double *b, const double *x, double *r, double *t)
{
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
		r[i] = sign * b[i];
		t[i] = sign * x[i];
	}
}

// masked squared error of the network output y (box) against the target
// (compact unknowns, scaled to unit RMS: the solver scales the network's
// correction itself) and its gradient
__global__ __launch_bounds__(kBlock) void kLossGradBox(UnetBox B, const int *grid2mat, const float *y, const double *target, const double *tt,
	int n, float *gy, float *part)
{
	const double scale = tt[0] > 0 ? sqrt(double(n) / tt[0]) : 1.0;
	const float inv = 1.0f / float(n);
	const size_t vox = size_t(B.n[0]) * B.n[1] * B.n[2];
	float acc = 0;
	for (size_t v = blockIdx.x * size_t(blockDim.x) + threadIdx.x; v < vox; v += size_t(gridDim.x) * blockDim.x) {
		const int x = int(v % B.n[0]), yy = int((v / B.n[0]) % B.n[1]), z = int(v / (size_t(B.n[0]) * B.n[1]));
		const int gx = B.org[0] + x, gy_ = B.org[1] + yy, gz = B.org[2] + z;
		int m = -1;
		if (gx >= 0 && gx < B.gn[0] && gy_ >= 0 && gy_ < B.gn[1] && gz >= 0 && gz < B.gn[2])
			m = grid2mat[(size_t(gz) * B.gn[1] + gy_) * B.gn[0] + gx];
		const float d = m >= 0 ? y[v] - float(target[m] * scale) : 0.0f;
		gy[v] = 2.0f * d * inv;
		acc += d * d;
	}
	double a[1] = {acc};
	double o[1];
	blockSum<1>(a, o);
	if (threadIdx.x == 0) part[blockIdx.x] = float(o[0]);
}

int trainOnSystem(int D, Unet &U, const Rows<double> &A64, const double *rhs, const double *solution, int n, const int *grid2mat,
	const int *mat2grid, const GridDev &cell, LearnedWork &W, int steps, float lr, double operatorFraction, unsigned long long seed,
	float *lossHistory, hipStream_t s, std::string &err)
{
	if (n <= 0) { err = "the last projection has no unknowns"; return GARM_EINVAL; }
	if (!W.reserve(size_t(n), err)) return GARM_ENOMEM;
	int rc, lo[3], hi[3];
	if ((rc = unknownBox(mat2grid, n, cell, W, s, lo, hi, err))) return rc;
	const UnetBox B = networkBox(D, lo, hi, cell.n);
	const size_t vox = size_t(B.n[0]) * B.n[1] * B.n[2];
	if (!U.buf.reserve(vox, true, err) || !U.buf.setDims(B.n[0], B.n[1], B.n[2], err)) return GARM_ENOMEM;
	float *loss = nullptr, *part = nullptr;
	const int nbl = 1024;
	if (hipMalloc(&loss, size_t(std::max(steps, 1)) * sizeof(float)) != hipSuccess || hipMalloc(&part, nbl * sizeof(float)) != hipSuccess) {
		if (loss) (void)hipFree(loss);
		err = "training scratch allocation failed";
		return GARM_ENOMEM;
	}
	// longest wavelength: the unknowns' extent, at most kMaxWavelength cells
	double L = 0;
	for (int a = 0; a < D; a++) L = std::max(L, double(hi[a] - lo[a] + 1));
	L = std::min(L, kMaxWavelength);
	const dim3 g1(lblocks(n)), b1(kBlock);
	const int nb = lblocks(n);
	double *rr = W.ls + kLsTrainIn, *tt = W.ls + kLsTrainOut, *modes = W.pa;
	std::mt19937_64 rng(seed);
	for (int k = 0; k < steps; k++) {
		const uint64_t draw = rng();
		if (double(draw >> 11) * 0x1.0p-53 < operatorFraction) {
			hipLaunchKernelGGL(kSmoothModes, dim3(1), dim3(64), 0, s, D, uint64_t(rng()), L, modes);
			hipLaunchKernelGGL(kSmoothField, g1, b1, 0, s, mat2grid, n, cell.n[0], cell.n[1], (const double *)modes, W.t);
			if (D == 2) hipLaunchKernelGGL(kOperatorPair<2>, g1, b1, 0, s, A64, n, (const double *)W.t, W.r, W.pb);
			else hipLaunchKernelGGL(kOperatorPair<3>, g1, b1, 0, s, A64, n, (const double *)W.t, W.r, W.pb);
		} else {
			hipLaunchKernelGGL(kSignedCopies, g1, b1, 0, s, n, (draw & 1) ? -1.0 : 1.0, rhs, solution, W.r, W.t);
			hipLaunchKernelGGL(kFcgDot, g1, b1, 0, s, n, (const double *)W.r, (const double *)W.r, W.pb);
		}
		hipLaunchKernelGGL(kFcgReduce1, dim3(1), b1, 0, s, (const double *)W.pb, nb, rr);
		hipLaunchKernelGGL(kFcgDot, g1, b1, 0, s, n, (const double *)W.t, (const double *)W.t, W.pb);
		hipLaunchKernelGGL(kFcgReduce1, dim3(1), b1, 0, s, (const double *)W.pb, nb, tt);
		unetGatherCompact(U, B, grid2mat, W.r, rr, n, s);
		unetForward(U, s);
		hipLaunchKernelGGL(kLossGradBox, dim3(nbl), b1, 0, s, B, grid2mat, (const float *)U.buf.y, (const double *)W.t, (const double *)tt, n,
			U.buf.gy, part);
		hipLaunchKernelGGL(kLossReduce, dim3(1), dim3(64), 0, s, (const float *)part, nbl, 1.0f / float(n), loss + k);
		if (!unetBackward(U, s, err)) { (void)hipFree(loss); (void)hipFree(part); return GARM_ENOMEM; }
		unetAdam(U, lr, 0.9f, 0.999f, 1e-8f, s);
	}
	bool ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(s) == hipSuccess;
	if (ok && lossHistory && steps > 0) ok = hipMemcpy(lossHistory, loss, size_t(steps) * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess;
	(void)hipFree(loss);
	(void)hipFree(part);
	if (!ok) { err = "training on the pressure system failed"; return GARM_EHIP; }
	return GARM_OK;
}

} // namespace garm

// ==================================================================== C-ABI
using namespace garm;

struct garm_unet {
	Unet U;
	hipStream_t s = nullptr;
	int device = 0;
	std::string err;
};

Unet *garm::unetOf(garm_unet *net, int *device)
{
	*device = net->device;
	return &net->U;
}

namespace {
thread_local std::string g_unet_create_error;

int ufail(garm_unet *u, int code, const std::string &m)
{
	u->err = m;
	return code;
}

template <typename T>
struct DevBuf {
	T *p = nullptr;
	explicit DevBuf(size_t n) { if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) p = nullptr; }
	~DevBuf() { if (p) (void)hipFree(p); }
};

bool dims3(const int *d, size_t &vox)
{
	if (!d || d[0] < 4 || d[1] < 4 || d[2] < 4 || d[0] % 4 || d[1] % 4 || d[2] % 4 || d[0] > 1024 || d[1] > 1024 || d[2] > 1024) return false;
	vox = size_t(d[0]) * d[1] * d[2];
	return true;
}
} // namespace

extern "C" {

const char *garm_unet_create_error(void) { return g_unet_create_error.c_str(); }
const char *garm_unet_last_error(const garm_unet *u) { return u ? u->err.c_str() : "null network"; }
size_t garm_unet_param_count(void) { return unetParamCount(); }

int garm_unet_create(garm_unet **out, unsigned long long seed)
{
	if (!out) return GARM_EINVAL;
	*out = nullptr;
	int ndev = 0;
	if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) { g_unet_create_error = "no HIP device"; return GARM_ENODEV; }
	int dev = 0;
	hipDeviceProp_t prop{};
	if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) { g_unet_create_error = "device query failed"; return GARM_ENODEV; }
	if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) { g_unet_create_error = std::string("device is not gfx950: ") + prop.gcnArchName; return GARM_ENODEV; }
	auto *u = new garm_unet;
	u->device = dev;
	std::string err;
	if (hipStreamCreateWithFlags(&u->s, hipStreamNonBlocking) != hipSuccess || !unetCreate(u->U, seed, u->s, err)) {
		g_unet_create_error = err.empty() ? "stream creation failed" : err;
		garm_unet_destroy(u);
		return GARM_ENOMEM;
	}
	*out = u;
	return GARM_OK;
}

void garm_unet_destroy(garm_unet *u)
{
	if (!u) return;
	if (u->s) (void)hipStreamSynchronize(u->s);
	hipStream_t s = u->s;
	delete u;
	if (s) (void)hipStreamDestroy(s);
}

int garm_unet_get_params(garm_unet *u, float *out, size_t n)
{
	if (!u || !out || n != unetParamCount()) return GARM_EINVAL;
	return unetGetParams(u->U, out, false, u->err) ? GARM_OK : GARM_EHIP;
}

int garm_unet_set_params(garm_unet *u, const float *in, size_t n)
{
	if (!u || !in || n != unetParamCount()) return GARM_EINVAL;
	return unetSetParams(u->U, in, u->err) ? GARM_OK : GARM_EHIP;
}

int garm_unet_forward(garm_unet *u, const int dims[3], const float *r, const uint8_t *mask, float *y)
{
	size_t vox;
	if (!u || !r || !mask || !y || !dims3(dims, vox)) return u ? ufail(u, GARM_EINVAL, "bad network grid or buffers") : GARM_EINVAL;
	if (!u->U.buf.reserve(vox, u->U.buf.has_grad, u->err) || !u->U.buf.setDims(dims[0], dims[1], dims[2], u->err)) return GARM_ENOMEM;
	DevBuf<float> dr(vox);
	DevBuf<uint8_t> dm(vox);
	if (!dr.p || !dm.p) return ufail(u, GARM_ENOMEM, "device allocation failed");
	if (hipMemcpy(dr.p, r, vox * sizeof(float), hipMemcpyHostToDevice) != hipSuccess || hipMemcpy(dm.p, mask, vox, hipMemcpyHostToDevice) != hipSuccess)
		return ufail(u, GARM_EHIP, "upload failed");
	unetLoadInput(u->U, dr.p, dm.p, u->s);
	unetForward(u->U, u->s);
	if (hipGetLastError() != hipSuccess || hipStreamSynchronize(u->s) != hipSuccess) return ufail(u, GARM_EHIP, "network forward failed");
	return hipMemcpy(y, u->U.buf.y, vox * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess ? GARM_OK : ufail(u, GARM_EHIP, "download failed");
}

int garm_unet_loss_grad(garm_unet *u, const int dims[3], const float *r, const uint8_t *mask, const float *target, double *loss, float *grad, size_t n)
{
	size_t vox;
	if (!u || !r || !mask || !target || !loss || !grad || n != unetParamCount() || !dims3(dims, vox)) return GARM_EINVAL;
	if (!u->U.buf.reserve(vox, true, u->err) || !u->U.buf.setDims(dims[0], dims[1], dims[2], u->err)) return GARM_ENOMEM;
	DevBuf<float> dr(vox), dt(vox), part(1024), dl(1);
	DevBuf<uint8_t> dm(vox);
	if (!dr.p || !dt.p || !dm.p || !part.p || !dl.p) return ufail(u, GARM_ENOMEM, "device allocation failed");
	size_t m = 0;
	for (size_t v = 0; v < vox; v++) m += mask[v] != 0;
	if (m == 0) return ufail(u, GARM_EINVAL, "empty mask");
	if (hipMemcpy(dr.p, r, vox * sizeof(float), hipMemcpyHostToDevice) != hipSuccess || hipMemcpy(dt.p, target, vox * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
		hipMemcpy(dm.p, mask, vox, hipMemcpyHostToDevice) != hipSuccess) return ufail(u, GARM_EHIP, "upload failed");
	hipLaunchKernelGGL(kLoadScaled, dim3(1024), dim3(kBlock), 0, u->s, (const float *)dr.p, (const uint8_t *)dm.p, 1.0f, u->U.buf.in, vox);
	unetForward(u->U, u->s);
	const int nb = 1024;
	hipLaunchKernelGGL(kLossGrad, dim3(nb), dim3(kBlock), 0, u->s, (const float *)u->U.buf.y, (const float *)dt.p, (const uint8_t *)dm.p, 1.0f,
		float(1.0 / double(m)), u->U.buf.gy, vox, part.p);
	hipLaunchKernelGGL(kLossReduce, dim3(1), dim3(64), 0, u->s, (const float *)part.p, nb, float(1.0 / double(m)), dl.p);
	if (!unetBackward(u->U, u->s, u->err)) return GARM_ENOMEM;
	float hl = 0;
	if (hipGetLastError() != hipSuccess || hipStreamSynchronize(u->s) != hipSuccess || hipMemcpy(&hl, dl.p, sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
		return ufail(u, GARM_EHIP, "network gradient failed");
	*loss = hl;
	return unetGetParams(u->U, grad, true, u->err) ? GARM_OK : GARM_EHIP;
}

// synthetic network input for timing: a smooth residual and a full mask
__global__ void kSyntheticInput(bf16 *in, int nx, int ny, int nz)
{
	const size_t vox = size_t(nx) * ny * nz;
	for (size_t v = blockIdx.x * size_t(blockDim.x) + threadIdx.x; v < vox; v += size_t(gridDim.x) * blockDim.x) {
		const int i = int(v % nx), j = int((v / nx) % ny), k = int(v / (size_t(nx) * ny));
		in[v * 4 + 0] = f2bf(__sinf(0.11f * i) * __cosf(0.07f * j + 0.05f * k));
		in[v * 4 + 1] = f2bf(1.0f);
		in[v * 4 + 2] = 0;
		in[v * 4 + 3] = 0;
	}
}

int garm_unet_time_forward(garm_unet *u, const int dims[3], int reps, double *avg_ms, double *flops)
{
	size_t vox;
	if (!u || !avg_ms || !flops || reps < 1 || !dims3(dims, vox)) return u ? ufail(u, GARM_EINVAL, "bad network grid or arguments") : GARM_EINVAL;
	if (!u->U.buf.reserve(vox, u->U.buf.has_grad, u->err) || !u->U.buf.setDims(dims[0], dims[1], dims[2], u->err)) return GARM_ENOMEM;
	hipLaunchKernelGGL(kSyntheticInput, dim3(1024), dim3(kBlock), 0, u->s, u->U.buf.in, dims[0], dims[1], dims[2]);
	unetForward(u->U, u->s);   // warm-up
	hipEvent_t e0 = nullptr, e1 = nullptr;
	if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
		if (e0) (void)hipEventDestroy(e0);
		return ufail(u, GARM_EHIP, "event creation failed");
	}
	(void)hipEventRecord(e0, u->s);
	for (int q = 0; q < reps; q++) unetForward(u->U, u->s);
	(void)hipEventRecord(e1, u->s);
	float ms = 0;
	const bool ok = hipGetLastError() == hipSuccess && hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	if (!ok) return ufail(u, GARM_EHIP, "network timing failed");
	*avg_ms = double(ms) / reps;
	// multiply-adds of the true channel counts: 2 * 27 * cin * cout per voxel of each layer's level, plus the 1x1 output
	double f = 0;
	for (int l = 0; l < kUnetLayers; l++) {
		const UnetLayer &L = u->U.L[l];
		f += 2.0 * 27 * L.cin * L.cout * double(u->U.buf.vox[L.level]);
	}
	*flops = f + 2.0 * 16 * double(vox);
	return GARM_OK;
}

int garm_unet_train(garm_unet *u, const int dims[3], int n_pairs, const float *rhs, const float *pressure, const uint8_t *mask, int steps,
	float lr, unsigned long long seed, float *loss_history)
{
	size_t vox;
	if (!u || !rhs || !pressure || !mask || n_pairs < 1 || steps < 0 || !(lr > 0) || !dims3(dims, vox)) return GARM_EINVAL;
	if (!u->U.buf.reserve(vox, true, u->err) || !u->U.buf.setDims(dims[0], dims[1], dims[2], u->err)) return GARM_ENOMEM;
	const size_t total = vox * size_t(n_pairs);
	DevBuf<float> dr(total), dp(total), part(1024), dl(std::max(steps, 1));
	DevBuf<uint8_t> dm(total);
	if (!dr.p || !dp.p || !dm.p || !part.p || !dl.p) return ufail(u, GARM_ENOMEM, "training set does not fit");
	if (hipMemcpy(dr.p, rhs, total * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
		hipMemcpy(dp.p, pressure, total * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
		hipMemcpy(dm.p, mask, total, hipMemcpyHostToDevice) != hipSuccess) return ufail(u, GARM_EHIP, "upload failed");
	// per pair: unit-RMS scale of the right-hand side over its unknowns
	std::vector<float> scale(n_pairs);
	std::vector<double> invCount(n_pairs);
	for (int i = 0; i < n_pairs; i++) {
		double ss = 0;
		size_t m = 0;
		for (size_t v = 0; v < vox; v++)
			if (mask[i * vox + v]) { ss += double(rhs[i * vox + v]) * rhs[i * vox + v]; m++; }
		scale[i] = (ss > 0 && m > 0) ? float(std::sqrt(double(m) / ss)) : 1.0f;
		invCount[i] = m > 0 ? 1.0 / double(m) : 0.0;
	}
	std::mt19937_64 rng(seed);
	const int nb = 1024;
	for (int k = 0; k < steps; k++) {
		const uint64_t draw = rng();
		const int i = int((draw >> 1) % uint64_t(n_pairs));
		// the pair or its negative: the preconditioner must follow the sign of
		// its input rather than predict a typical pressure from the mask
		const float sc = (draw & 1) ? -scale[i] : scale[i];
		const size_t off = size_t(i) * vox;
		hipLaunchKernelGGL(kLoadScaled, dim3(1024), dim3(kBlock), 0, u->s, (const float *)(dr.p + off), (const uint8_t *)(dm.p + off), sc, u->U.buf.in, vox);
		unetForward(u->U, u->s);
		hipLaunchKernelGGL(kLossGrad, dim3(nb), dim3(kBlock), 0, u->s, (const float *)u->U.buf.y, (const float *)(dp.p + off), (const uint8_t *)(dm.p + off),
			sc, float(invCount[i]), u->U.buf.gy, vox, part.p);
		hipLaunchKernelGGL(kLossReduce, dim3(1), dim3(64), 0, u->s, (const float *)part.p, nb, float(invCount[i]), dl.p + k);
		if (!unetBackward(u->U, u->s, u->err)) return GARM_ENOMEM;
		unetAdam(u->U, lr, 0.9f, 0.999f, 1e-8f, u->s);
	}
	if (hipGetLastError() != hipSuccess || hipStreamSynchronize(u->s) != hipSuccess) return ufail(u, GARM_EHIP, "training failed");
	if (loss_history && steps > 0 && hipMemcpy(loss_history, dl.p, steps * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
		return ufail(u, GARM_EHIP, "download failed");
	return GARM_OK;
}

} // extern "C"
