// Matrix-free Jacobi-preconditioned CG on the bounding box of the unknowns,
// one launch per iteration.
//
// The ghost-fluid Poisson operator (Projector.cpp:30-68) is held per box
// cell as its diagonal and the weights of the cell's -x, -y, -z faces (zero
// where a face does not join two unknowns), so (A v)_c = diag_c v_c minus the
// six face-weighted neighbours at fixed strides: no neighbour indices, no
// rows.  Cells of the box that are not unknowns have a zero operator and stay
// zero in every vector; the Jacobi preconditioner is u = r / diag there and
// 0 where diag = 0.
//
// The iteration is the single-reduction form of preconditioned CG
// (Chronopoulos and Gear, 1989), with u = D^-1 r and w = A u kept beside r:
//   p = u + beta p,   s = w + beta s   (so s = A p),
//   x += alpha p,     r' = r - alpha s,  u' = D^-1 r',  w' = A u',
//   gamma' = r'.u',   delta' = u'.w',    rho' = r'.r',
//   beta' = gamma' / gamma,  alpha' = gamma' / (delta' - beta' gamma' / alpha).
// All inner products of an iteration come out of the same pass, so one
// launch per iteration suffices: every workgroup re-reduces the previous
// launch's per-workgroup partials in the same order (the same scalars
// everywhere, no fences or atomics), updates its cells, recomputes u' of the
// six neighbours from their r, w, s and diagonal (its stencil must not wait
// for another workgroup), and writes its own partials.  r, w and s alternate
// between two buffers (launch k reads set k & 1 and writes the other), p and
// x are updated in place.  3D boxes are walked in z-columns with each plane
// of u' shared through LDS, one workgroup per column tile that holds
// unknowns (a splash's bounding box is mostly empty; the other tiles stay
// zero in every vector and are not launched).
//
// The recursive residual of this form can drift further from b - A x than
// plain CG's, so a solve ends only on the true residual: when the recursive
// one meets its threshold, r is replaced by b - A x (FP64 operator), w and
// s = A p are recomputed, and the solve stops or continues from the previous
// direction: the first iteration after a replacement conjugates the new
// direction against it (beta = -u.Ap / p.Ap) and takes p.Ap from three extra
// inner products instead of the recurrences, which hold again from there.
//
// The mixed-precision form keeps the operator and p, s, r, w in FP32 and
// accumulates the iteration's correction e in FP32; the FP64 solution takes
// e at every replacement, which comes whenever the FP32 residual has fallen
// by kBoxReplaceDrop.  Inner products and norms accumulate in FP64 in both.
#include "garm_pcg.hpp"

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>

namespace garm {

constexpr int kBoxGrid = 1024;
constexpr double kBoxReplaceDrop = 1e-3;
constexpr int kBoxReplaceMax = 200;
// most cells of active column tiles per unknown for the box solver
constexpr double kBoxActiveRatio64 = 1.6, kBoxActiveRatio32 = 2.0;

constexpr int kColPartialsMax = 8192;   // most column tiles of one box

// scalar slots of the single-reduction iteration (by iteration parity)
enum { kScGam0 = 10, kScAlp0 = 12, kScPrevTrue = 14 };

static int boxBlocks(size_t n) { return int(std::max<size_t>(1, std::min<size_t>(kBoxGrid, (n + kBlock - 1) / kBlock))); }

// Sum of NV values per thread over a workgroup of NW wave64s in a fixed
// order; the result is valid in thread 0.
template <int NW, int NV>
__device__ __forceinline__ void wgSum(double (&v)[NV], double (&tot)[NV])
{
	__shared__ double red[NV][NW];
	const int tid = int(threadIdx.x + blockDim.x * threadIdx.y), lane = tid & (kWave - 1), wid = tid / kWave;
#pragma unroll
	for (int q = 0; q < NV; q++) {
		const double s = waveSum(v[q]);
		if (lane == 0) red[q][wid] = s;
	}
	__syncthreads();
	if (tid == 0)
#pragma unroll
		for (int q = 0; q < NV; q++) {
			double s = 0;
			for (int w = 0; w < NW; w++) s += red[q][w];
			tot[q] = s;
		}
}

// Sum of NV interleaved partials of `nb` workgroups; valid in thread 0.
template <int NW, int NV>
__device__ __forceinline__ void wgSumParts(const double *partials, int nb, double (&tot)[NV])
{
	double s[NV];
#pragma unroll
	for (int q = 0; q < NV; q++) s[q] = 0;
	const int tid = int(threadIdx.x + blockDim.x * threadIdx.y);
	for (int i = tid; i < nb; i += NW * kWave)
#pragma unroll
		for (int q = 0; q < NV; q++) s[q] += partials[NV * i + q];
	wgSum<NW, NV>(s, tot);
}

template <int NW, int NV>
__device__ __forceinline__ void wgPutParts(double (&v)[NV], double *partials, int slot)
{
	double t[NV];
	wgSum<NW, NV>(v, t);
	if (threadIdx.x == 0 && threadIdx.y == 0)
#pragma unroll
		for (int q = 0; q < NV; q++) partials[NV * slot + q] = t[q];
}

template <typename T>
struct BoxSys {
	int n, sy, sz;                      // cells, row and plane strides
	const T *diag, *wx, *wy, *wz;
};

// Jacobi: r / diag, 0 on cells that are not unknowns
template <typename T>
__device__ __forceinline__ T jacobi(T r, T d) { return d != T(0) ? r / d : T(0); }

// (A v)_c with v given by a functor of the cell index
template <int D, typename T, typename V>
__device__ __forceinline__ T boxRow(const BoxSys<T> &O, int c, const V &v)
{
	T s = O.diag[c] * v(c);
	s -= O.wx[c] * v(c - 1);
	s -= O.wx[c + 1] * v(c + 1);
	s -= O.wy[c] * v(c - O.sy);
	s -= O.wy[c + O.sy] * v(c + O.sy);
	if constexpr (D == 3) {
		s -= O.wz[c] * v(c - O.sz);
		s -= O.wz[c + O.sz] * v(c + O.sz);
	}
	return s;
}

// Index arithmetic: a cell's -face weight is zero on the box's low faces and
// the +neighbour's weight reads one past the high face, so the arrays carry a
// zero margin of one plane (plus one cell) on each side: element c of a
// vector lives at base + margin + c.

// box cell of each unknown
__global__ void kBoxCells(const int *mat2grid, int nu, int gx, int gy, int ox, int oy, int oz, int bx, int by, unsigned *boxOf)
{
	for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nu; r += gridDim.x * blockDim.x) {
		const int g = mat2grid[r];
		const int i = g % gx - ox, j = (g / gx) % gy - oy, k = g / (gx * gy) - oz;
		boxOf[r] = (unsigned(k) * unsigned(by) + unsigned(j)) * unsigned(bx) + unsigned(i);
	}
}

// Column tiles holding unknowns: a byte per tile (same-value stores), then
// their ordered list, compacted by one workgroup (the list order fixes the
// order of the partial sums, so it must not depend on timing); the list's
// length and the count of box cells its tiles cover follow it, at
// kColPartialsMax and kColPartialsMax + 1.
__global__ void kTileMark(const unsigned *boxOf, int nu, int bx, int by, int ty, int zch, int gx, int gy, unsigned char *mask)
{
	for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nu; r += gridDim.x * blockDim.x) {
		const unsigned c = boxOf[r];
		const int i = int(c % unsigned(bx)), j = int((c / unsigned(bx)) % unsigned(by)), k = int(c / (unsigned(bx) * unsigned(by)));
		mask[i / 64 + gx * (j / ty + gy * (k / zch))] = 1;
	}
}

__global__ __launch_bounds__(kBlock) void kTileList(const unsigned char *mask, int nt, int gx, int gy, int bx, int by, int bz, int ty, int zch,
	int *tiles)
{
	const int per = (nt + kBlock - 1) / kBlock, t0 = int(threadIdx.x) * per, t1 = min(nt, t0 + per);
	int cnt = 0, cells = 0;
	for (int t = t0; t < t1; t++)
		if (mask[t]) {
			cnt++;
			// the tile's cells inside the box
			const int x = t % gx, y = (t / gx) % gy, z = t / (gx * gy);
			cells += min(64, bx - 64 * x) * min(ty, by - ty * y) * min(zch, bz - zch * z);
		}
	int total, allCells;
	int at = blockExclusiveScan(cnt, &total);
	(void)blockExclusiveScan(cells, &allCells);
	for (int t = t0; t < t1; t++)
		if (mask[t]) tiles[at++] = t;
	if (threadIdx.x == 0) { tiles[kColPartialsMax] = total; tiles[kColPartialsMax + 1] = allCells; }
}

__global__ void kZeroWords(double *p, size_t words)
{
	for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < words; i += size_t(gridDim.x) * blockDim.x) p[i] = 0;
}

// operator on the box from the rows (ordinal o: -x +x -y +y -z +z)
template <int D, typename T>
__global__ void kBoxOperator(Rows<T> A, int nu, const unsigned *boxOf, T *diag, T *wx, T *wy, T *wz)
{
	for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nu; r += gridDim.x * blockDim.x) {
		const unsigned c = boxOf[r];
		diag[c] = A.diag[r];
		wx[c] = A.nb[0][r] >= 0 ? A.w[0][r] : T(0);
		wy[c] = A.nb[2][r] >= 0 ? A.w[2][r] : T(0);
		if constexpr (D == 3) wz[c] = A.nb[4][r] >= 0 ? A.w[4][r] : T(0);
	}
}

// r = T(b) on the unknowns' cells, b kept on the box in FP64 for the
// replacements; partials of |b|^2
template <typename T>
__global__ __launch_bounds__(kBlock) void kBoxStart(int nu, const unsigned *boxOf, const double *b, double *bbox, T *r, double *pb)
{
	double v[1] = {0};
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nu; i += gridDim.x * blockDim.x) {
		const unsigned c = boxOf[i];
		const double bi = b[i];
		bbox[c] = bi;
		r[c] = T(bi);
		v[0] += bi * bi;
	}
	wgPutParts<kBlock / kWave, 1>(v, pb, int(blockIdx.x));
}

__global__ __launch_bounds__(kBlock) void kBoxStartFinish(const double *pb, int nb, double tol, int maxIt, bool mixed, double *sc, int *st)
{
	double t[1];
	wgSumParts<kBlock / kWave, 1>(pb, nb, t);
	if (threadIdx.x != 0) return;
	const double b2 = t[0];
	sc[kScBnorm2] = b2;
	sc[kScRn2] = b2;
	sc[kScOuterRn2] = b2;
	// the first iteration is a restart: p = s = 0 gives beta = 0
	sc[kScGam0] = sc[kScGam0 + 1] = 0.0;
	sc[kScAlp0] = sc[kScAlp0 + 1] = 1.0;
	sc[kScPrevTrue] = DBL_MAX;
	st[kStIter] = 0;
	st[kStMaxIter] = maxIt;
	st[kStDone] = 0;
	st[kStOuterDone] = 0;
	st[kStSweep] = 0;
	const double tt = b2 > 0 ? fmax(tol, DBL_EPSILON / sqrt(b2)) : tol;
	sc[kScTol] = tt;
	const double thr = fmax(tt * tt * b2, DBL_MIN);
	sc[kScOuterThreshold] = thr;
	sc[kScThreshold] = mixed ? fmax(thr, kBoxReplaceDrop * kBoxReplaceDrop * b2) : thr;
	if (b2 == 0 || b2 < thr || maxIt <= 0) { st[kStDone] = 1; st[kStOuterDone] = 1; }
}

// Start of a run of iterations (after the start or a replacement): w = A u
// with u = D^-1 r and, after a replacement, s = A p afresh (the recurrence
// for s drifts from A p as r's does from b - A x); partials of r.u, u.w, r.r
// and, for p.Ap of the first iteration, u.s, p.w, p.s
template <int D, typename T, bool REFRESH>
__global__ __launch_bounds__(kBlock) void kCgStart(BoxSys<T> O, const T *r, T *w, T *s, const T *p, double *part)
{
	auto u = [&](int c) { return jacobi(r[c], O.diag[c]); };
	auto pv = [&](int c) { return p[c]; };
	double v[6] = {0, 0, 0, 0, 0, 0};
	for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < O.n; c += gridDim.x * blockDim.x) {
		const T uc = u(c), rc = r[c];
		const T q = boxRow<D, T>(O, c, u);
		w[c] = q;
		T sc;
		if constexpr (REFRESH) {
			sc = boxRow<D, T>(O, c, pv);
			s[c] = sc;
		} else sc = s[c];
		const double pc = double(p[c]);
		v[0] += double(rc) * double(uc);
		v[1] += double(uc) * double(q);
		v[2] += double(rc) * double(rc);
		v[3] += double(uc) * double(sc);
		v[4] += pc * double(q);
		v[5] += pc * double(sc);
	}
	wgPutParts<kBlock / kWave, 6>(v, part, int(blockIdx.x));
}

// Vectors and scalars of one iteration launch.
template <typename T, typename TX>
struct CgArgs {
	BoxSys<T> O;
	int bx, by, bz, zch;
	const T *rIn, *wIn, *sIn;
	T *rOut, *wOut, *sOut, *p;
	TX *x;                       // solution (FP64) or correction (mixed)
	const double *partIn;        // partials of the previous launch
	int nbIn;
	double *partOut;
	double *sc;
	int *st;
	const int *tiles;            // column kernel: the tiles holding unknowns
	int gx, gy;                  // column kernel: tiles per row and per plane
};

// alpha, beta of iteration k from the previous launch's partials (every
// workgroup computes the same values); the designated workgroup records the
// state.  Returns false when the solve stops here.
template <int NW, bool RESTART, typename T, typename TX>
__device__ __forceinline__ bool cgScalars(const CgArgs<T, TX> &A, int k, bool designated, double &alpha, double &beta)
{
	constexpr int NV = RESTART ? 6 : 3;
	double t[NV];
	wgSumParts<NW, NV>(A.partIn, A.nbIn, t);
	__shared__ double bc[3];
	const int tid = int(threadIdx.x + blockDim.x * threadIdx.y);
	if (tid == 0) {
		double al = 0, be = 0, stop = 0;
		const int q = k & 1;
		if (t[2] < A.sc[kScThreshold] || k >= A.st[kStMaxIter]) stop = 1;
		else {
			double pap;
			if constexpr (RESTART) {
				// the new residual is not the recurrence's, so beta = gamma /
				// gamma_prev would not make p A-conjugate to the previous
				// direction (and the recurrences after it would break down):
				// take the conjugating beta = -u.Ap_prev / p_prev.Ap_prev
				be = t[5] > 0 ? -t[3] / t[5] : 0.0;
				pap = t[1] + be * (t[3] + t[4]) + be * be * t[5];
			} else {
				be = t[0] / A.sc[kScGam0 + (q ^ 1)];
				pap = t[1] - be * t[0] / A.sc[kScAlp0 + (q ^ 1)];
			}
			if (!(pap > 0)) stop = 1;   // breakdown (p.Ap must be positive)
			else al = t[0] / pap;
		}
		bc[0] = al;
		bc[1] = be;
		bc[2] = stop;
		if (designated) {
			A.sc[kScRn2] = t[2];
			if (stop != 0) A.st[kStDone] = 1;
			else {
				A.sc[kScGam0 + q] = t[0];
				A.sc[kScAlp0 + q] = al;
				A.st[kStIter] = k + 1;
			}
		}
	}
	__syncthreads();
	alpha = bc[0];
	beta = bc[1];
	return bc[2] == 0;
}

// One iteration over the box cells (2D, and 3D boxes too wide for the
// column kernel's partials): grid-stride, u' of the neighbours recomputed
// from global memory.
template <int D, typename T, typename TX, bool RESTART>
__global__ __launch_bounds__(kBlock) void kCgIter(CgArgs<T, TX> A, int k)
{
	if (A.st[kStDone]) return;
	double al, be;
	if (!cgScalars<kBlock / kWave, RESTART>(A, k, blockIdx.x == 0, al, be)) return;
	const T alpha = T(al), beta = T(be);
	const TX alphaX = TX(al);
	auto uNew = [&](int c) {
		const T sn = A.wIn[c] + beta * A.sIn[c];
		return jacobi(A.rIn[c] - alpha * sn, A.O.diag[c]);
	};
	double v[3] = {0, 0, 0};
	for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < A.O.n; c += gridDim.x * blockDim.x) {
		const T r0 = A.rIn[c], d = A.O.diag[c];
		const T sn = A.wIn[c] + beta * A.sIn[c];
		const T rn = r0 - alpha * sn;
		const T un = jacobi(rn, d);
		const T q = boxRow<D, T>(A.O, c, uNew);
		const T pn = jacobi(r0, d) + beta * A.p[c];
		A.p[c] = pn;
		A.x[c] += alphaX * TX(pn);
		A.rOut[c] = rn;
		A.sOut[c] = sn;
		A.wOut[c] = q;
		v[0] += double(rn) * double(un);
		v[1] += double(un) * double(q);
		v[2] += double(rn) * double(rn);
	}
	wgPutParts<kBlock / kWave, 3>(v, A.partOut, int(blockIdx.x));
}

// The same iteration for 3D boxes, walking z: a workgroup of 64 x TY threads
// owns a 64 x TY column of cells over `zch` planes.  A thread keeps u' of its
// cell at planes k - 1, k, k + 1 in registers (the next plane's r, w, s,
// diagonal and -z weight loaded once, then carried) and shares plane k with
// its x and y neighbours through LDS.  The plane tile's one-cell ring is
// computed by the first 2 (64 + TY) threads, one cell each (whole waves take
// the two rows), so no wave carries more than one ring cell per plane: with
// the ring on the edge threads, the waves of the first and last rows did
// three evaluations per plane and the barrier made every wave wait for them
// (48.6 against 54.8 us per FP64 iteration at 256^3).  Cells outside the box
// read as 0 (their face weights are 0).
constexpr int kColX = 64;
constexpr int kColPartials = kColPartialsMax;   // most workgroups of one launch

template <typename T, typename TX, bool RESTART, int TY>
__global__ __launch_bounds__(kColX * TY) void kCgIterCol(CgArgs<T, TX> A, int k)
{
	constexpr int NW = kColX * TY / kWave;
	if (A.st[kStDone]) return;
	__shared__ T plane[TY + 2][kColX + 2];
	const int tx = int(threadIdx.x), ty = int(threadIdx.y);
	// one workgroup per tile holding unknowns (the others stay zero in every
	// vector), dealt to the XCDs in contiguous runs of the list
	const int slot = xcdIndex();
	const int tl = A.tiles[slot];
	const Tile3 tile{tl % A.gx, (tl / A.gx) % A.gy, tl / (A.gx * A.gy)};
	const int bx = A.bx, by = A.by, bz = A.bz;
	const int i = tile.x * kColX + tx, j = tile.y * TY + ty;
	const int k0 = tile.z * A.zch, k1 = min(bz, k0 + A.zch);
	const bool inside = i < bx && j < by;
	const int sy = A.O.sy, sz = A.O.sz;
	const T *__restrict__ diag = A.O.diag;
	// the own column's loads of planes k0 - 1 and k0 do not depend on alpha
	// and beta: issued before the previous launch's partials are summed, so
	// their latency overlaps that reduction
	int c = (k0 * by + j) * bx + i;
	T rm = 0, wm = 0, sm = 0, dm = 0, r0 = 0, w0 = 0, s0 = 0, d0 = 0, wz0 = 0;
	if (inside) {
		rm = A.rIn[c - sz]; wm = A.wIn[c - sz]; sm = A.sIn[c - sz]; dm = diag[c - sz];   // k0 - 1 >= -1: the zero margin
		r0 = A.rIn[c]; w0 = A.wIn[c]; s0 = A.sIn[c]; d0 = diag[c]; wz0 = A.O.wz[c];
	}
	double al, be;
	if (!cgScalars<NW, RESTART>(A, k, blockIdx.x == 0, al, be)) return;
	const T alpha = T(al), beta = T(be);
	const TX alphaX = TX(al);
	auto uAt = [&](int ii, int jj, int kk) -> T {
		if (ii < 0 || ii >= bx || jj < 0 || jj >= by) return T(0);
		const int c = (kk * by + jj) * bx + ii;   // kk in [-1, bz]: the zero margins
		const T sn = A.wIn[c] + beta * A.sIn[c];
		return jacobi(A.rIn[c] - alpha * sn, diag[c]);
	};
	// ring cell of this thread (if any): grid (rx, ry), LDS (lx, ly)
	const int flat = tx + kColX * ty;
	int rx, ry, lx, ly;
	if (flat < 2 * kColX) { lx = 1 + (flat & (kColX - 1)); ly = flat < kColX ? 0 : TY + 1; }
	else { const int t = flat - 2 * kColX; lx = t < TY ? 0 : kColX + 1; ly = 1 + (t < TY ? t : t - TY); }
	rx = tile.x * kColX + lx - 1;
	ry = tile.y * TY + ly - 1;
	// own cell: plane k0 - 1 (u' only) and plane k0, as uAt forms them
	T um = 0, sn0 = 0, rn0 = 0, un0 = 0;
	if (inside) {
		um = jacobi(rm - alpha * (wm + beta * sm), dm);
		sn0 = w0 + beta * s0;
		rn0 = r0 - alpha * sn0;
		un0 = jacobi(rn0, d0);
	}
	double v[3] = {0, 0, 0};
	for (int kk = k0; kk < k1; kk++, c += sz) {
		T r1 = 0, d1 = 0, sn1 = 0, rn1 = 0, un1 = 0, wz1 = 0;
		if (inside) {   // plane kk + 1 (the upper margin past the box)
			const int cn = c + sz;
			r1 = A.rIn[cn];
			d1 = diag[cn];
			sn1 = A.wIn[cn] + beta * A.sIn[cn];
			rn1 = r1 - alpha * sn1;
			un1 = jacobi(rn1, d1);
			wz1 = A.O.wz[cn];
		}
		// the ring of the plane tile, one cell per thread of the first
		// 2 (64 + TY) threads: whole waves take the two rows
		T ring = 0;
		if (flat < 2 * kColX + 2 * TY) ring = uAt(rx, ry, kk);
		__syncthreads();
		plane[ty + 1][tx + 1] = un0;
		if (flat < 2 * kColX + 2 * TY) plane[ly][lx] = ring;
		__syncthreads();
		if (inside) {
			T q = d0 * un0;
			q -= A.O.wx[c] * plane[ty + 1][tx];
			q -= A.O.wx[c + 1] * plane[ty + 1][tx + 2];
			q -= A.O.wy[c] * plane[ty][tx + 1];
			q -= A.O.wy[c + sy] * plane[ty + 2][tx + 1];
			q -= wz0 * um;
			q -= wz1 * un1;
			const T pn = jacobi(r0, d0) + beta * A.p[c];
			A.p[c] = pn;
			A.x[c] += alphaX * TX(pn);
			A.rOut[c] = rn0;
			A.sOut[c] = sn0;
			A.wOut[c] = q;
			v[0] += double(rn0) * double(un0);
			v[1] += double(un0) * double(q);
			v[2] += double(rn0) * double(rn0);
		}
		um = un0;
		r0 = r1;
		d0 = d1;
		sn0 = sn1;
		rn0 = rn1;
		un0 = un1;
		wz0 = wz1;
	}
	wgPutParts<NW, 3>(v, A.partOut, slot);
}

// mixed form: x += e, e = 0
__global__ void kCgFold(int n, double *x, float *e)
{
	for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
		x[c] += double(e[c]);
		e[c] = 0.f;
	}
}

// Residual replacement: r = T(b - A x) with the FP64 operator; partials of
// |b - A x|^2 (FP64)
template <int D, typename T>
__global__ __launch_bounds__(kBlock) void kCgReplace(BoxSys<double> O64, const double *bbox, const double *x, T *r, double *pb)
{
	double v[1] = {0};
	auto xv = [&](int c) { return x[c]; };
	for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < O64.n; c += gridDim.x * blockDim.x) {
		const double rc = bbox[c] - boxRow<D, double>(O64, c, xv);
		r[c] = T(rc);
		v[0] += rc * rc;
	}
	wgPutParts<kBlock / kWave, 1>(v, pb, int(blockIdx.x));
}

// Accept when the FP64 residual meets the tolerance, or when it has reached
// the rounding floor of b - A x (it fell by less than half since the last
// replacement, within kBoxFloorSlack of the tolerance: b - A x cannot be
// formed more accurately than about eps |A| |x|); else aim the iteration at
// the next reduction (mixed: by kBoxReplaceDrop) and continue with the same
// direction.
constexpr double kBoxFloorSlack = 1e4;   // on squared norms: 100 x the tolerance
__global__ __launch_bounds__(kBlock) void kCgReplaceFinish(const double *pb, int nb, bool mixed, double *sc, int *st)
{
	double t[1];
	wgSumParts<kBlock / kWave, 1>(pb, nb, t);
	if (threadIdx.x != 0) return;
	sc[kScOuterRn2] = t[0];
	sc[kScRn2] = t[0];
	const double prev = sc[kScPrevTrue];
	sc[kScPrevTrue] = t[0];
	st[kStSweep] = st[kStSweep] + 1;
	const bool floor = t[0] > 0.25 * prev && t[0] < kBoxFloorSlack * sc[kScOuterThreshold];
	if (t[0] < sc[kScOuterThreshold] || floor) { st[kStOuterDone] = 1; st[kStDone] = 1; return; }
	sc[kScThreshold] = mixed ? fmax(sc[kScOuterThreshold], kBoxReplaceDrop * kBoxReplaceDrop * t[0]) : sc[kScOuterThreshold];
	st[kStDone] = st[kStIter] >= st[kStMaxIter] ? 1 : 0;
}

__global__ void kBoxGather(int nu, const unsigned *boxOf, const double *xbox, double *x)
{
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nu; i += gridDim.x * blockDim.x) x[i] = xbox[boxOf[i]];
}

// ----------------------------------------------------------------- driver
BoxKrylov::~BoxKrylov()
{
	if (pool) (void)hipFree(pool);
	if (partials) (void)hipFree(partials);
	if (boxOf) (void)hipFree(boxOf);
	if (tiles) (void)hipFree(tiles);
	if (tileMask) (void)hipFree(tileMask);
}

static int readBoxState(Krylov &K, hipStream_t s, std::string &err)
{
	if (hipMemcpyAsync(K.host_state, K.state, 8 * sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
		hipMemcpyAsync(K.host_sc, K.sc, 16 * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
		spinWait(s, K.ready) != hipSuccess) { err = "solver state readback failed"; return GARM_EHIP; }
	return GARM_OK;
}

template <int D, typename T>
static int boxSolve(const Rows<double> &A64, const Rows<T> &A, const double *b, int nu, const BoxLayout &L, const garm_pcg_opts &o, Krylov &K,
	BoxKrylov &B, hipStream_t s, garm_pcg_stats *st, std::string &err)
{
	constexpr bool mixed = sizeof(T) == 4;
	using TX = T;   // FP64: the solution itself; mixed: the FP32 correction
	const int bx = L.n[0], by = L.n[1], bz = D == 3 ? L.n[2] : 1;
	const size_t cells = size_t(bx) * by * bz;
	const size_t margin = size_t(bx) * by + 64;   // one plane and a little more
	const size_t span = (cells + 2 * margin + 63) / 64 * 64;
	if (cells + 2 * margin >= size_t(INT_MAX)) { err = "solver box too large"; return GARM_EINVAL; }
	// T arrays: diag wx wy wz, r w s (two sets), p, e (mixed) ; FP64: x, b,
	// and the FP64 operator when mixed (4)
	constexpr int nT = mixed ? 12 : 11;
	const size_t tWords = (nT * span * sizeof(T) + 7) / 8;
	const size_t words = tWords + 2 * span + (mixed ? 4 * span : 0);
	constexpr size_t kPartWords = 6 * kColPartials;   // one parity's partials
	if (!B.partials && hipMalloc(&B.partials, 2 * kPartWords * sizeof(double)) != hipSuccess) { err = "solver partials allocation failed"; return GARM_ENOMEM; }
	if (!B.tiles && (hipMalloc(&B.tiles, (kColPartialsMax + 2) * sizeof(int)) != hipSuccess ||
			hipMalloc(&B.tileMask, kColPartialsMax) != hipSuccess)) { err = "solver tile list allocation failed"; return GARM_ENOMEM; }
	if (size_t(nu) > B.ocap) {
		if (B.boxOf) (void)hipFree(B.boxOf);
		B.boxOf = nullptr;
		const size_t want = size_t(nu) + size_t(nu) / 4;
		if (hipMalloc(&B.boxOf, want * sizeof(unsigned)) != hipSuccess) { B.ocap = 0; err = "solver box allocation failed"; return GARM_ENOMEM; }
		B.ocap = want;
	}
	const int nbU = boxBlocks(size_t(nu)), nbC = boxBlocks(cells);
	const dim3 b1(kBlock);
	hipLaunchKernelGGL(kBoxCells, dim3(nbU), b1, 0, s, L.mat2grid, nu, L.gn[0], L.gn[1], L.lo[0], L.lo[1], D == 3 ? L.lo[2] : 0, bx, by, B.boxOf);
	// 3D: the z-walking kernel over 64 x ty tiles and columns of zmin planes
	// (more planes per column when that would exceed kColPartials
	// workgroups); 2D and very wide boxes: the grid-stride one.  Measured at
	// 256^3 over 64 x {4, 8, 16} tiles and {4, 8, 16}-plane columns: FP64
	// 64 x 4 x 16 (48.6 us per iteration; 51.6-61.5 for the others), FP32
	// 64 x 8 x 8 (28.7 us; 29.7-37.7)
	constexpr int ty = mixed ? 8 : 4, zmin = mixed ? 8 : 16;
	const int gx = (bx + kColX - 1) / kColX, gy = (by + ty - 1) / ty;
	int zch = 0, gz = 0;
	if (D == 3 && gx * gy <= kColPartials) {
		zch = std::max(zmin, (bz + kColPartials / (gx * gy) - 1) / (kColPartials / (gx * gy)));
		gz = (bz + zch - 1) / zch;
	}
	// the tiles holding unknowns, in order; only they are launched (a
	// splash's box is mostly empty)
	int ntiles = 0, activeCells = int(cells);
	if (zch) {
		const int nt = gx * gy * gz;
		if (hipMemsetAsync(B.tileMask, 0, nt, s) != hipSuccess) { err = "solver tile mask clear failed"; return GARM_EHIP; }
		hipLaunchKernelGGL(kTileMark, dim3(nbU), b1, 0, s, (const unsigned *)B.boxOf, nu, bx, by, ty, zch, gx, gy, B.tileMask);
		hipLaunchKernelGGL(kTileList, dim3(1), b1, 0, s, (const unsigned char *)B.tileMask, nt, gx, gy, bx, by, bz, ty, zch, B.tiles);
		if (hipMemcpyAsync(K.host_state + 8, B.tiles + kColPartialsMax, 2 * sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
			spinWait(s, K.ready) != hipSuccess) { err = "solver tile count readback failed"; return GARM_EHIP; }
		ntiles = K.host_state[8];
		activeCells = K.host_state[9];
		// too sparse for the box: the compact rows cost less per iteration
		// (per box cell of an active tile against per unknown, at 256^3:
		// FP64 22.6 against 37 ns, FP32 13.2 against 27 ns)
		if (double(activeCells) > (mixed ? kBoxActiveRatio32 : kBoxActiveRatio64) * nu) return kBoxTooSparse;
	}
	if (words > B.cap) {
		if (B.pool) (void)hipFree(B.pool);
		B.pool = nullptr;
		const size_t want = words + words / 4;
		if (hipMalloc(&B.pool, want * sizeof(double)) != hipSuccess) { B.cap = 0; err = "solver box allocation failed"; return GARM_ENOMEM; }
		B.cap = want;
	}
	T *tp = reinterpret_cast<T *>(B.pool);
	auto carveT = [&]() { T *q = tp + margin; tp += span; return q; };
	T *diag = carveT(), *wx = carveT(), *wy = carveT(), *wz = carveT();
	T *rv[2], *wv[2], *sv[2];
	for (int q = 0; q < 2; q++) { rv[q] = carveT(); wv[q] = carveT(); sv[q] = carveT(); }
	T *p = carveT();
	T *e = mixed ? carveT() : nullptr;
	double *dp = B.pool + tWords;
	auto carveD = [&]() { double *q = dp + margin; dp += span; return q; };
	double *bbox = carveD();
	double *x64 = carveD();
	double *d64 = nullptr, *w64[3] = {nullptr, nullptr, nullptr};
	TX *xIt;
	if constexpr (mixed) {
		d64 = carveD();
		for (int a = 0; a < 3; a++) w64[a] = carveD();
		xIt = e;
	} else xIt = x64;
	double *part[2] = {B.partials, B.partials + kPartWords};

	// everything zero (margins, non-unknown cells, p, s, e), then the operator and b
	hipLaunchKernelGGL(kZeroWords, dim3(8192), b1, 0, s, B.pool, words);
	hipLaunchKernelGGL((kBoxOperator<D, T>), dim3(nbU), b1, 0, s, A, nu, (const unsigned *)B.boxOf, diag, wx, wy, wz);
	if (mixed)
		hipLaunchKernelGGL((kBoxOperator<D, double>), dim3(nbU), b1, 0, s, A64, nu, (const unsigned *)B.boxOf, d64, w64[0], w64[1], w64[2]);
	hipLaunchKernelGGL((kBoxStart<T>), dim3(nbU), b1, 0, s, nu, (const unsigned *)B.boxOf, b, bbox, rv[0], K.pb);
	hipLaunchKernelGGL(kBoxStartFinish, dim3(1), b1, 0, s, (const double *)K.pb, nbU, o.tolerance, o.max_iterations, mixed, K.sc, K.state);
	const BoxSys<T> O{int(cells), bx, bx * by, diag, wx, wy, wz};
	BoxSys<double> O64;
	if constexpr (mixed) O64 = BoxSys<double>{int(cells), bx, bx * by, d64, w64[0], w64[1], w64[2]};
	else O64 = BoxSys<double>{int(cells), bx, bx * by, diag, wx, wy, wz};
	// the run of iterations starts from w = A u, its partials where
	// iteration `it` reads them
	auto startRun = [&](int it, bool refresh) {
		const int cur = it & 1;
		if (refresh)
			hipLaunchKernelGGL((kCgStart<D, T, true>), dim3(nbC), b1, 0, s, O, (const T *)rv[cur], wv[cur], sv[cur], (const T *)p, part[cur ^ 1]);
		else
			hipLaunchKernelGGL((kCgStart<D, T, false>), dim3(nbC), b1, 0, s, O, (const T *)rv[cur], wv[cur], sv[cur], (const T *)p, part[cur ^ 1]);
	};
	startRun(0, false);   // p = s = 0
	const int check = std::max(1, o.check_interval);
	const int nbIter = zch ? ntiles : nbC;
	int rc, replacements = 0;
	bool restart = true;
	for (;;) {
		// chunks of `check` iterations until the device stops them
		for (;;) {
			if ((rc = readBoxState(K, s, err))) return rc;
			if (K.host_state[kStDone]) break;
			const int base = K.host_state[kStIter];
			for (int q = 0; q < check; q++) {
				const int k = base + q, cur = k & 1;
				const bool first = restart && q == 0;
				CgArgs<T, TX> a{O, bx, by, bz, zch, rv[cur], wv[cur], sv[cur], rv[cur ^ 1], wv[cur ^ 1], sv[cur ^ 1], p, xIt,
					part[cur ^ 1], first ? nbC : nbIter, part[cur], K.sc, K.state, B.tiles, gx, gy};
				if (zch) {
					const dim3 g(ntiles), bk(kColX, ty);
					if (first) hipLaunchKernelGGL((kCgIterCol<T, TX, true, ty>), g, bk, 0, s, a, k);
					else hipLaunchKernelGGL((kCgIterCol<T, TX, false, ty>), g, bk, 0, s, a, k);
				} else {
					if (first) hipLaunchKernelGGL((kCgIter<D, T, TX, true>), dim3(nbC), b1, 0, s, a, k);
					else hipLaunchKernelGGL((kCgIter<D, T, TX, false>), dim3(nbC), b1, 0, s, a, k);
				}
			}
			restart = false;
		}
		if (K.host_state[kStOuterDone] || K.host_state[kStIter] >= o.max_iterations || replacements >= kBoxReplaceMax) break;
		// the recursive residual met its target: replace it by b - A x
		const int it = K.host_state[kStIter], cur = it & 1;
		if constexpr (mixed) hipLaunchKernelGGL(kCgFold, dim3(nbC), b1, 0, s, int(cells), x64, e);
		hipLaunchKernelGGL((kCgReplace<D, T>), dim3(nbC), b1, 0, s, O64, (const double *)bbox, (const double *)x64, rv[cur], K.pb);
		hipLaunchKernelGGL(kCgReplaceFinish, dim3(1), b1, 0, s, (const double *)K.pb, nbC, mixed, K.sc, K.state);
		startRun(it, true);
		restart = true;
		replacements++;
	}
	if constexpr (mixed) hipLaunchKernelGGL(kCgFold, dim3(nbC), b1, 0, s, int(cells), x64, e);
	hipLaunchKernelGGL(kBoxGather, dim3(nbU), b1, 0, s, nu, (const unsigned *)B.boxOf, (const double *)x64, K.x);
	st->iterations = K.host_state[kStIter];
	st->outer_iterations = replacements;
	const double b2 = K.host_sc[kScBnorm2];
	st->relative_residual = b2 > 0 ? std::sqrt(K.host_sc[kScOuterRn2] / b2) : 0.0;
	// algorithmic bytes per box cell: an iteration reads p, s, x (or e), r,
	// w, diag and the three weights and writes p, s, x, r, w; the start reads
	// r, diag, s, p and writes w; a replacement reads x, b and the FP64
	// operator and writes r (mixed: plus the fold, x and e read and written),
	// then reads r, p and the operator and writes w and s
	// (iterations: the cells of the active tiles)
	const double ts = double(sizeof(T));
	const double iterBytes = 14 * ts;
	const double startBytes = 5 * ts;
	const double replBytes = 48 + ts + (mixed ? 16 + 2 * ts : 0.0) + 8 * ts;
	st->hbm_bytes = double(activeCells) * iterBytes * st->iterations + double(cells) * (startBytes + replBytes * replacements);
	st->solver_cells = activeCells;
	if (!K.host_state[kStOuterDone]) {
		err = std::string(mixed ? "mixed-precision solve" : "PCG") + " did not reach the requested tolerance (" + std::to_string(st->iterations) +
			" iterations, " + std::to_string(replacements) + " replacements, relative residual " + std::to_string(st->relative_residual) + ")";
		return GARM_ENOCONV;
	}
	return GARM_OK;
}

int solveBoxPcg(int D, bool mixed, const Rows<double> &A64, const Rows<float> &A32, const double *b, int n, const BoxLayout &L,
	const garm_pcg_opts &o, Krylov &K, BoxKrylov &B, hipStream_t s, garm_pcg_stats *st, std::string &err)
{
	if (mixed) return D == 2 ? boxSolve<2, float>(A64, A32, b, n, L, o, K, B, s, st, err) : boxSolve<3, float>(A64, A32, b, n, L, o, K, B, s, st, err);
	return D == 2 ? boxSolve<2, double>(A64, A64, b, n, L, o, K, B, s, st, err) : boxSolve<3, double>(A64, A64, b, n, L, o, K, B, s, st, err);
}

int solveJacobi(int D, bool mixed, const Rows<double> &A64, const Rows<float> &A32, const double *b, int n, const BoxLayout &L,
	const garm_pcg_opts &o, Krylov &K, BoxKrylov &B, hipStream_t s, garm_pcg_stats *st, std::string &err, bool *usedBox)
{
	const double cells = double(L.n[0]) * L.n[1] * L.n[2];
	int rc = kBoxTooSparse;
	if (cells <= (D == 3 ? kBoxSolverRatio3 : kBoxSolverRatio2) * n) rc = solveBoxPcg(D, mixed, A64, A32, b, n, L, o, K, B, s, st, err);
	*usedBox = rc != kBoxTooSparse;
	if (*usedBox) return rc;
	return mixed ? solveMixed(D, A64, A32, b, n, o, K, s, st, err) : solvePcg64(D, A64, b, n, o, K, s, st, err);
}

} // namespace garm
