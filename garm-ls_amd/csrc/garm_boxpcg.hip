// Matrix-free Jacobi-preconditioned CG on the bounding box of the unknowns.
//
// The ghost-fluid Poisson operator (Projector.cpp:30-68) is held per box
// cell as its diagonal and the weights of the cell's -x, -y, -z faces (zero
// where a face does not join two unknowns), so (A p)_c = diag_c p_c minus the
// six face-weighted neighbours at fixed strides: no neighbour indices, no
// rows.  Cells of the box that are not unknowns have a zero operator and stay
// zero in every vector.  One iteration is two launches:
//   kDirProduct  beta = (r.z)_k / (r.z)_{k-1} (convergence and iteration
//                limit tested first), p_k = z + beta p_{k-1} -- recomputed for
//                the six neighbours, so no launch waits for p --, q = A p_k;
//                partials of p.q
//   kUpdate      alpha = (r.z)_k / (p.q); x += alpha p; r -= alpha q;
//                z = D^-1 r; partials of |r|^2 and r.z
// p alternates between two buffers (a launch reads the previous direction
// from one and writes the new one to the other, so no workgroup reads what
// another writes).  Every workgroup re-reduces the producer's partials in
// the same order, so all hold the same scalars without fences or atomics;
// dot products and norms accumulate in FP64.  The arithmetic per element is
// that of the row-based solver (the same alpha, beta, updates), so the two
// agree to rounding.
//
// The mixed-precision form keeps the operator and p, r, z, q in FP32 and x in
// FP64; when the FP32 residual has fallen by kReplaceDrop it is replaced by
// the FP64 b - A x on the FP64 box operator, and the solve continues with the
// same direction.
#include "garm_pcg.hpp"

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>

namespace garm {

constexpr int kBoxGrid = 1024;
constexpr double kBoxReplaceDrop = 1e-3;

static int boxBlocks(size_t n) { return int(std::max<size_t>(1, std::min<size_t>(kBoxGrid, (n + kBlock - 1) / kBlock))); }

// Sum NV interleaved partials of `nb` workgroups over all threads of a
// (possibly two-dimensional) workgroup; the result is valid in thread 0.
template <int NV>
__device__ __forceinline__ void sumParts(const double *partials, int nb, double (&tot)[NV])
{
	double s[NV];
#pragma unroll
	for (int q = 0; q < NV; q++) s[q] = 0;
	const int tid = int(threadIdx.x + blockDim.x * threadIdx.y), nth = int(blockDim.x * blockDim.y);
	for (int i = tid; i < nb; i += nth)
#pragma unroll
		for (int q = 0; q < NV; q++) s[q] += partials[NV * i + q];
	blockSum<NV>(s, tot);
}

template <int NV>
__device__ __forceinline__ void bcastAll(double (&v)[NV])
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
__device__ __forceinline__ void putParts(double (&v)[NV], double *partials)
{
	double blk[NV];
	blockSum<NV>(v, blk);
	if (threadIdx.x == 0)
#pragma unroll
		for (int q = 0; q < NV; q++) partials[NV * blockIdx.x + q] = blk[q];
}

template <typename T>
struct BoxSys {
	int n, sy, sz;                      // cells, row and plane strides
	const T *diag, *idiag, *wx, *wy, *wz;
};

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

__global__ void kZeroWords(double *p, size_t words)
{
	for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < words; i += size_t(gridDim.x) * blockDim.x) p[i] = 0;
}

// operator on the box from the rows (ordinal o: -x +x -y +y -z +z); idiag
// may be null
template <int D, typename T>
__global__ void kBoxOperator(Rows<T> A, int nu, const unsigned *boxOf, T *diag, T *idiag, T *wx, T *wy, T *wz)
{
	for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nu; r += gridDim.x * blockDim.x) {
		const unsigned c = boxOf[r];
		diag[c] = A.diag[r];
		if (idiag) idiag[c] = A.invdiag[r];
		wx[c] = A.nb[0][r] >= 0 ? A.w[0][r] : T(0);
		wy[c] = A.nb[2][r] >= 0 ? A.w[2][r] : T(0);
		if constexpr (D == 3) wz[c] = A.nb[4][r] >= 0 ? A.w[4][r] : T(0);
	}
}

// r = T(b), z = D^-1 r on the unknowns, b kept on the box in FP64 for the
// replacements; partials of |b|^2 and r.z
template <typename T>
__global__ __launch_bounds__(kBlock) void kBoxStart(int nu, const unsigned *boxOf, const double *b, const T *idiag, double *bbox, T *r, T *z,
	double *pb)
{
	double v[2] = {0, 0};
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nu; i += gridDim.x * blockDim.x) {
		const unsigned c = boxOf[i];
		const double bi = b[i];
		bbox[c] = bi;
		const T ri = T(bi);
		const T zi = idiag[c] * ri;
		r[c] = ri;
		z[c] = zi;
		v[0] += bi * bi;
		v[1] += double(ri) * double(zi);
	}
	putParts<2>(v, pb);
}

__global__ __launch_bounds__(kBlock) void kBoxStartFinish(const double *pb, int nb, double tol, int maxIt, bool mixed, double *sc, int *st)
{
	double t[2];
	sumParts<2>(pb, nb, t);
	if (threadIdx.x != 0) return;
	const double b2 = t[0];
	sc[kScBnorm2] = b2;
	sc[kScRn2] = b2;
	sc[kScOuterRn2] = b2;
	sc[kScRz1] = 1.0;   // the previous direction is zero: any beta gives p_0 = z_0
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

// Iteration k, first launch.  pb holds the partials of |r_k|^2 and r_k.z_k
// (from kBoxUpdate of iteration k-1, the start, or a replacement).
template <int D, typename T>
__global__ __launch_bounds__(kBlock) void kDirProduct(BoxSys<T> O, const T *z, const T *pold, T *pnew, T *q, const double *pb, int nb,
	double *pa, double *sc, int *st, int k)
{
	if (st[kStDone]) return;
	double t[2];
	sumParts<2>(pb, nb, t);
	double b[2] = {0, 0};   // beta, stop
	if (threadIdx.x == 0) {
		if (t[0] < sc[kScThreshold] || k >= st[kStMaxIter]) b[1] = 1;
		else b[0] = t[1] / sc[kScRz0 + ((k & 1) ^ 1)];
	}
	bcastAll<2>(b);
	if (blockIdx.x == 0 && threadIdx.x == 0) {
		sc[kScRn2] = t[0];
		if (b[1] != 0) st[kStDone] = 1;
		else {
			sc[kScRz0 + (k & 1)] = t[1];
			st[kStIter] = k + 1;
		}
	}
	if (b[1] != 0) return;
	const T beta = T(b[0]);
	auto dir = [&](int c) { return z[c] + beta * pold[c]; };
	double v[1] = {0};
	for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < O.n; c += gridDim.x * blockDim.x) {
		const T pc = dir(c);
		const T s = boxRow<D, T>(O, c, dir);
		pnew[c] = pc;
		q[c] = s;
		v[0] += double(pc) * double(s);
	}
	putParts<1>(v, pa);
}

// The same launch for 3D boxes, walking z: a workgroup of 64 x 4 threads
// owns a 64 x 4 column of cells over `zch` planes.  Each thread keeps the new
// direction of its cell at planes k - 1, k, k + 1 in registers (one load of
// z and p per cell) and shares plane k with its x and y neighbours through
// LDS, whose one-cell ring the edge threads load.  Cells outside the box
// read as 0 (their face weights are 0).
constexpr int kColX = 64, kColY = 4, kColZ = 16;
constexpr int kColPartials = 8192;   // most workgroups of one launch (partials of p.q)

template <typename T>
__global__ __launch_bounds__(kColX * kColY) void kDirProductCol(BoxSys<T> O, int bx, int by, int bz, int zch, const T *z, const T *pold,
	T *pnew, T *q, const double *pb, int nb, double *pa, double *sc, int *st, int k)
{
	if (st[kStDone]) return;
	double t[2];
	sumParts<2>(pb, nb, t);
	double b[2] = {0, 0};   // beta, stop
	if (threadIdx.x == 0 && threadIdx.y == 0) {
		if (t[0] < sc[kScThreshold] || k >= st[kStMaxIter]) b[1] = 1;
		else b[0] = t[1] / sc[kScRz0 + ((k & 1) ^ 1)];
	}
	{
		__shared__ double bb[2];
		__syncthreads();
		if (threadIdx.x == 0 && threadIdx.y == 0) { bb[0] = b[0]; bb[1] = b[1]; }
		__syncthreads();
		b[0] = bb[0];
		b[1] = bb[1];
	}
	const int flat = int(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z));
	if (flat == 0 && threadIdx.x == 0 && threadIdx.y == 0) {
		sc[kScRn2] = t[0];
		if (b[1] != 0) st[kStDone] = 1;
		else {
			sc[kScRz0 + (k & 1)] = t[1];
			st[kStIter] = k + 1;
		}
	}
	if (b[1] != 0) return;
	const T beta = T(b[0]);
	__shared__ T plane[kColY + 2][kColX + 2];
	const int tx = int(threadIdx.x), ty = int(threadIdx.y);
	const Tile3 tile = xcdTiles();
	const int i = tile.x * kColX + tx, j = tile.y * kColY + ty;
	const int k0 = tile.z * zch, k1 = min(bz, k0 + zch);
	const bool inside = i < bx && j < by;
	const int sy = O.sy, sz = O.sz;
	auto dirAt = [&](int ii, int jj, int kk) -> T {
		if (ii < 0 || ii >= bx || jj < 0 || jj >= by) return T(0);
		const int c = (kk * by + jj) * bx + ii;   // kk in [-1, bz]: the zero margins
		return z[c] + beta * pold[c];
	};
	// ring cells of the tile: threads of rows 0 and 3 load rows -1 and 4,
	// threads of columns 0 and 63 load columns -1 and 64
	T dm = inside ? dirAt(i, j, k0 - 1) : T(0);
	T d0 = inside ? dirAt(i, j, k0) : T(0);
	double v = 0;
	for (int kk = k0; kk < k1; kk++) {
		const T dp = inside ? dirAt(i, j, kk + 1) : T(0);
		__syncthreads();
		plane[ty + 1][tx + 1] = d0;
		if (ty == 0) plane[0][tx + 1] = dirAt(i, j - 1, kk);
		if (ty == kColY - 1) plane[kColY + 1][tx + 1] = dirAt(i, j + 1, kk);
		if (tx == 0) plane[ty + 1][0] = dirAt(i - 1, j, kk);
		if (tx == kColX - 1) plane[ty + 1][kColX + 1] = dirAt(i + 1, j, kk);
		__syncthreads();
		if (inside) {
			const int c = (kk * by + j) * bx + i;
			T s = O.diag[c] * d0;
			s -= O.wx[c] * plane[ty + 1][tx];
			s -= O.wx[c + 1] * plane[ty + 1][tx + 2];
			s -= O.wy[c] * plane[ty][tx + 1];
			s -= O.wy[c + sy] * plane[ty + 2][tx + 1];
			s -= O.wz[c] * dm;
			s -= O.wz[c + sz] * dp;
			pnew[c] = d0;
			q[c] = s;
			v += double(d0) * double(s);
		}
		dm = d0;
		d0 = dp;
	}
	// workgroup partial of p.q (256 threads as one flat block)
	__shared__ double red[kColX * kColY / kWave];
	const int tid = tx + kColX * ty;
	double w = waveSum(v);
	if ((tid & (kWave - 1)) == 0) red[tid / kWave] = w;
	__syncthreads();
	if (tid == 0) {
		double tot = 0;
		for (int q2 = 0; q2 < kColX * kColY / kWave; q2++) tot += red[q2];
		pa[tile.x + int(gridDim.x) * (tile.y + int(gridDim.y) * tile.z)] = tot;   // the tile's slot: the sum order of the partials is the launch's
	}
}

// Iteration k, second launch.
template <typename T, typename TX>
__global__ __launch_bounds__(kBlock) void kBoxUpdate(int n, TX *x, T *r, T *z, const T *p, const T *q, const T *idiag, const double *pa,
	int na, double *pb, const double *sc, const int *st, int k)
{
	if (st[kStDone]) return;
	double a[1];
	sumParts<1>(pa, na, a);
	if (threadIdx.x == 0) a[0] = sc[kScRz0 + (k & 1)] / a[0];
	bcastAll<1>(a);
	const TX alphaX = TX(a[0]);
	const T alpha = T(a[0]);
	double v[2] = {0, 0};
	for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
		x[c] += alphaX * TX(p[c]);
		const T rc = r[c] - alpha * q[c];
		const T zc = idiag[c] * rc;
		r[c] = rc;
		z[c] = zc;
		v[0] += double(rc) * double(rc);
		v[1] += double(rc) * double(zc);
	}
	putParts<2>(v, pb);
}

// Residual replacement: r = T(b - A x) with the FP64 operator, z = D^-1 r;
// partials of |b - A x|^2 (FP64) and r.z
template <int D, typename T>
__global__ __launch_bounds__(kBlock) void kBoxReplace(BoxSys<double> O64, const double *bbox, const double *x, const T *idiag, T *r, T *z,
	double *pb)
{
	double v[2] = {0, 0};
	auto xv = [&](int c) { return x[c]; };
	for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < O64.n; c += gridDim.x * blockDim.x) {
		const double rc = bbox[c] - boxRow<D, double>(O64, c, xv);
		const T rt = T(rc);
		const T zt = idiag[c] * rt;
		r[c] = rt;
		z[c] = zt;
		v[0] += rc * rc;
		v[1] += double(rt) * double(zt);
	}
	putParts<2>(v, pb);
}

// accept when the FP64 residual meets the tolerance, else aim the inner
// iterations at the next 1e3 reduction and continue with the same direction
__global__ __launch_bounds__(kBlock) void kBoxReplaceFinish(const double *pb, int nb, double *sc, int *st)
{
	double t[2];
	sumParts<2>(pb, nb, t);
	if (threadIdx.x != 0) return;
	sc[kScOuterRn2] = t[0];
	sc[kScRn2] = t[0];
	st[kStSweep] = st[kStSweep] + 1;
	if (t[0] < sc[kScOuterThreshold]) { st[kStOuterDone] = 1; st[kStDone] = 1; return; }
	sc[kScThreshold] = fmax(sc[kScOuterThreshold], kBoxReplaceDrop * kBoxReplaceDrop * t[0]);
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
}

static int readBoxState(Krylov &K, hipStream_t s, std::string &err)
{
	if (hipMemcpyAsync(K.host_state, K.state, 8 * sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
		hipMemcpyAsync(K.host_sc, K.sc, 16 * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
		spinWait(s, K.ready) != hipSuccess) { err = "solver state readback failed"; return GARM_EHIP; }
	return GARM_OK;
}

// Vectors of one solve carved from the pool: every array has `margin` zero
// words before and after its `cells` (the operator's neighbour reads).
template <typename T>
struct BoxVecs {
	T *diag, *idiag, *wx, *wy, *wz, *r, *z, *p[2], *q;
	double *x, *bbox;
	double *d64, *w64[3];   // FP64 operator of the replacements (mixed form)
};

template <int D, typename T>
static int boxSolve(const Rows<double> &A64, const Rows<T> &A, const double *b, int nu, const BoxLayout &L, const garm_pcg_opts &o, Krylov &K,
	BoxKrylov &B, hipStream_t s, garm_pcg_stats *st, std::string &err)
{
	constexpr bool mixed = sizeof(T) == 4;
	const int bx = L.n[0], by = L.n[1], bz = D == 3 ? L.n[2] : 1;
	const size_t cells = size_t(bx) * by * bz;
	const size_t margin = size_t(bx) * by + 64;   // one plane and a little more, rounded below
	const size_t span = (cells + 2 * margin + 63) / 64 * 64;
	if (cells + 2 * margin >= size_t(INT_MAX)) { err = "solver box too large"; return GARM_EINVAL; }
	// words: T arrays (diag idiag wx wy wz r z p0 p1 q) + FP64 (x, b) + FP64 operator when mixed (4)
	const size_t tWords = (10 * span * sizeof(T) + 7) / 8;
	const size_t words = tWords + 2 * span + (mixed ? 4 * span : 0);
	if (words > B.cap) {
		if (B.pool) (void)hipFree(B.pool);
		B.pool = nullptr;
		const size_t want = words + words / 4;
		if (hipMalloc(&B.pool, want * sizeof(double)) != hipSuccess) { B.cap = 0; err = "solver box allocation failed"; return GARM_ENOMEM; }
		B.cap = want;
	}
	if (!B.partials && hipMalloc(&B.partials, kColPartials * sizeof(double)) != hipSuccess) { err = "solver partials allocation failed"; return GARM_ENOMEM; }
	if (size_t(nu) > B.ocap) {
		if (B.boxOf) (void)hipFree(B.boxOf);
		B.boxOf = nullptr;
		const size_t want = size_t(nu) + size_t(nu) / 4;
		if (hipMalloc(&B.boxOf, want * sizeof(unsigned)) != hipSuccess) { B.ocap = 0; err = "solver box allocation failed"; return GARM_ENOMEM; }
		B.ocap = want;
	}
	BoxVecs<T> V{};
	T *tp = reinterpret_cast<T *>(B.pool);
	T **tarr[10] = {&V.diag, &V.idiag, &V.wx, &V.wy, &V.wz, &V.r, &V.z, &V.p[0], &V.p[1], &V.q};
	for (int a = 0; a < 10; a++) { *tarr[a] = tp + margin; tp += span; }
	double *dp = B.pool + tWords;
	V.x = dp + margin; dp += span;
	V.bbox = dp + margin; dp += span;
	if (mixed) {
		V.d64 = dp + margin; dp += span;
		for (int a = 0; a < 3; a++) { V.w64[a] = dp + margin; dp += span; }
	}
	const int nbU = boxBlocks(size_t(nu)), nbC = boxBlocks(cells);
	const dim3 b1(kBlock);
	// everything zero (margins, non-unknown cells), then the operator and b
	hipLaunchKernelGGL(kZeroWords, dim3(8192), b1, 0, s, B.pool, words);
	hipLaunchKernelGGL(kBoxCells, dim3(nbU), b1, 0, s, L.mat2grid, nu, L.gn[0], L.gn[1], L.lo[0], L.lo[1], D == 3 ? L.lo[2] : 0, bx, by, B.boxOf);
	hipLaunchKernelGGL((kBoxOperator<D, T>), dim3(nbU), b1, 0, s, A, nu, (const unsigned *)B.boxOf, V.diag, V.idiag, V.wx, V.wy, V.wz);
	if (mixed)
		hipLaunchKernelGGL((kBoxOperator<D, double>), dim3(nbU), b1, 0, s, A64, nu, (const unsigned *)B.boxOf, V.d64, (double *)nullptr, V.w64[0],
			V.w64[1], V.w64[2]);
	hipLaunchKernelGGL((kBoxStart<T>), dim3(nbU), b1, 0, s, nu, (const unsigned *)B.boxOf, b, (const T *)V.idiag, V.bbox, V.r, V.z, K.pb);
	hipLaunchKernelGGL(kBoxStartFinish, dim3(1), b1, 0, s, (const double *)K.pb, nbU, o.tolerance, o.max_iterations, mixed, K.sc, K.state);
	BoxSys<T> O{int(cells), bx, bx * by, V.diag, V.idiag, V.wx, V.wy, V.wz};
	const int check = std::max(1, o.check_interval);
	int rc, replacements = 0, nbPrev = nbU;
	// 3D FP64: the z-walking product over columns of kColZ planes (more
	// planes per column when that would exceed kColPartials workgroups);
	// otherwise, and for FP32 and in 2D, the grid-stride one (measured at
	// 256^3: FP64 33.7 against 35.4 us per launch; FP32 slower walking)
	const int gx = (bx + kColX - 1) / kColX, gy = (by + kColY - 1) / kColY;
	int zch = 0, gz = 0;
	if (D == 3 && !mixed && gx * gy <= kColPartials) {
		zch = std::max(kColZ, (bz + kColPartials / (gx * gy) - 1) / (kColPartials / (gx * gy)));
		gz = (bz + zch - 1) / zch;
	}
	const int nbDir = zch ? gx * gy * gz : nbC;
	double *pa = zch ? B.partials : K.pa;
	for (;;) {
		// chunks of `check` iterations until the device stops them
		for (;;) {
			if ((rc = readBoxState(K, s, err))) return rc;
			if (K.host_state[kStDone]) break;
			const int base = K.host_state[kStIter];
			for (int q = 0; q < check; q++) {
				const int k = base + q;
				if (zch)
					hipLaunchKernelGGL((kDirProductCol<T>), dim3(gx, gy, gz), dim3(kColX, kColY), 0, s, O, bx, by, bz, zch, (const T *)V.z,
						(const T *)V.p[(k & 1) ^ 1], V.p[k & 1], V.q, (const double *)K.pb, nbPrev, pa, K.sc, K.state, k);
				else
					hipLaunchKernelGGL((kDirProduct<D, T>), dim3(nbC), b1, 0, s, O, (const T *)V.z, (const T *)V.p[(k & 1) ^ 1], V.p[k & 1], V.q,
						(const double *)K.pb, nbPrev, K.pa, K.sc, K.state, k);
				hipLaunchKernelGGL((kBoxUpdate<T, double>), dim3(nbC), b1, 0, s, int(cells), V.x, V.r, V.z, (const T *)V.p[k & 1], (const T *)V.q,
					(const T *)V.idiag, (const double *)pa, nbDir, K.pb, (const double *)K.sc, (const int *)K.state, k);
				nbPrev = nbC;
			}
		}
		if (!mixed || K.host_state[kStOuterDone] || K.host_state[kStIter] >= o.max_iterations || replacements > 200) break;
		// the FP32 residual met its target: replace it by the FP64 one
		BoxSys<double> O64{int(cells), bx, bx * by, V.d64, nullptr, V.w64[0], V.w64[1], V.w64[2]};
		hipLaunchKernelGGL((kBoxReplace<D, T>), dim3(nbC), b1, 0, s, O64, (const double *)V.bbox, (const double *)V.x, (const T *)V.idiag, V.r, V.z,
			K.pb);
		hipLaunchKernelGGL(kBoxReplaceFinish, dim3(1), b1, 0, s, (const double *)K.pb, nbC, K.sc, K.state);
		nbPrev = nbC;
		replacements++;
	}
	hipLaunchKernelGGL(kBoxGather, dim3(nbU), b1, 0, s, nu, (const unsigned *)B.boxOf, (const double *)V.x, K.x);
	st->iterations = K.host_state[kStIter];
	st->outer_iterations = mixed ? replacements : 1;
	const double b2 = K.host_sc[kScBnorm2];
	const double rn2 = mixed ? K.host_sc[kScOuterRn2] : K.host_sc[kScRn2];
	st->relative_residual = b2 > 0 ? std::sqrt(rn2 / b2) : 0.0;
	// algorithmic bytes per box cell and iteration: direction + product (z, p
	// in; operator diag + 3 weights; p, q out), update (x rw, r rw, z out,
	// p, q, idiag in); replacement (x, b, FP64 operator in, r, z out)
	const double ts = double(sizeof(T));
	const double iterBytes = (2 * ts + 4 * ts + 2 * ts) + (16 + 2 * ts + ts + 2 * ts + ts);
	st->hbm_bytes = double(cells) * (iterBytes * st->iterations + (mixed ? (16 + 32 + 2 * ts) * replacements : 0.0));
	if (mixed ? !K.host_state[kStOuterDone] : st->relative_residual > K.host_sc[kScTol]) {
		err = mixed ? "mixed-precision solve did not reach the requested tolerance" : "PCG did not reach the requested tolerance";
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

} // namespace garm
