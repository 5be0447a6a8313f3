// Per-step glue of LevelSetLiquid::advance on gfx950: scene level sets,
// boundary enforcement, gravity, |u|max and the reference-map distortion
// check.  (Extrapolation lives in garm_extrap.hip.)
#include "garm_internal.hpp"

namespace garm {

// Nine component arrays passed by value (J = d xi / dx, column-major).
struct Mat9 { double *p[9]; };

// ----------------------------------------------------------------- scene
template <int D>
__global__ __launch_bounds__(kBlock) void kUnionBox(GridDev g, double *phi, double c0, double c1, double c2, double h0, double h1, double h2)
{
	const Coord c = threadCoord();
	if (!inGrid<D>(g, c)) return;
	const size_t id = gidx<D>(g, c.i, c.j, c.k);
	const double p[3] = {g.o[0] + c.i * g.dx, g.o[1] + c.j * g.dx, g.o[2] + c.k * g.dx};
	const double ce[3] = {c0, c1, c2}, hl[3] = {h0, h1, h2};
	// ImplicitBox::signedDistance (ImplicitSurface.h:51-56)
	double ph[3];
	bool allIn = true;
#pragma unroll
	for (int a = 0; a < D; a++) { ph[a] = fabs(p[a] - ce[a]) - hl[a]; if (!(ph[a] <= 0)) allIn = false; }
	double sd;
	if (allIn) { sd = ph[0]; for (int a = 1; a < D; a++) sd = fmax(sd, ph[a]); }
	else { double s = 0; for (int a = 0; a < D; a++) { const double m = fmax(ph[a], 0.0); s += m * m; } sd = sqrt(s); }
	phi[id] = fmin(phi[id], sd);
}

template <int D>
__global__ __launch_bounds__(kBlock) void kUnionSphere(GridDev g, double *phi, double c0, double c1, double c2, double r)
{
	const Coord c = threadCoord();
	if (!inGrid<D>(g, c)) return;
	const size_t id = gidx<D>(g, c.i, c.j, c.k);
	const double p[3] = {g.o[0] + c.i * g.dx, g.o[1] + c.j * g.dx, g.o[2] + c.k * g.dx};
	const double ce[3] = {c0, c1, c2};
	double s = 0;
#pragma unroll
	for (int a = 0; a < D; a++) s += (p[a] - ce[a]) * (p[a] - ce[a]);
	phi[id] = fmin(phi[id], sqrt(s) - r);   // ImplicitSphere (ImplicitSurface.h:23)
}

template <int D>
__global__ __launch_bounds__(kBlock) void kUnionPlane(GridDev g, double *phi, double p0, double p1, double p2, double n0, double n1, double n2)
{
	const Coord c = threadCoord();
	if (!inGrid<D>(g, c)) return;
	const size_t id = gidx<D>(g, c.i, c.j, c.k);
	const double p[3] = {g.o[0] + c.i * g.dx, g.o[1] + c.j * g.dx, g.o[2] + c.k * g.dx};
	const double q[3] = {p0, p1, p2}, n[3] = {n0, n1, n2};
	double d = 0;
#pragma unroll
	for (int a = 0; a < D; a++) d += (p[a] - q[a]) * n[a];
	phi[id] = fmin(phi[id], d);   // ImplicitPlane (ImplicitSurface.h:132)
}

// _refMap = cell centre, _refMapGrad = I (LevelSetLiquid.cpp:563-569, 1345-1351)
template <int D>
__global__ __launch_bounds__(kBlock) void kResetRefMap(GridDev g, double *x0, double *x1, double *x2, Mat9 J)
{
	const Coord c = threadCoord();
	if (!inGrid<D>(g, c)) return;
	const size_t id = gidx<D>(g, c.i, c.j, c.k);
	x0[id] = g.o[0] + c.i * g.dx;
	x1[id] = g.o[1] + c.j * g.dx;
	if constexpr (D == 3) x2[id] = g.o[2] + c.k * g.dx;
#pragma unroll
	for (int e = 0; e < D * D; e++) J.p[e][id] = (e % (D + 1) == 0) ? 1.0 : 0.0;
}

__global__ void kFill(double *p, double v, size_t n)
{
	for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) p[i] = v;
}
__global__ void kFillU8(uint8_t *p, uint8_t v, size_t n)
{
	for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) p[i] = v;
}
__global__ void kCopy(double *dst, const double *src, size_t n)
{
	for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) dst[i] = src[i];
}

static unsigned linearBlocks(size_t n)
{
	const size_t b = (n + kBlock - 1) / kBlock;
	return unsigned(b < 8192 ? (b ? b : 1) : 8192);
}

// enforceBoundaryLevelSet(sdf, tol) (LevelSetLiquid.cpp:966-975)
__global__ void kEnforceLevelSet(double *sdf, const double *cellDist, double tol, size_t n)
{
	for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
		if (sdf[i] < -cellDist[i] - tol) sdf[i] = -cellDist[i] - tol;
}

// applyBodyForces (LevelSetLiquid.cpp:837-841), kGravity = 9.8 (Constants.h:9)
__global__ void kGravity(double *uy, double dv, size_t n)
{
	for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) uy[i] -= dv;
}

// Analytic velocity fields of the reference's advection tests, sampled at
// the face positions of axis `axis`.  Enright deformation on the unit domain
// (TestAdvect3D.cpp:152-190 with the domain mapped to [0,1]^3):
//   u = 2 sin^2(pi x) sin(2 pi y) sin(2 pi z) c,  v = -sin^2(pi y) sin(2 pi x) sin(2 pi z) c,
//   w = -sin^2(pi z) sin(2 pi x) sin(2 pi y) c,   c = cos(pi t / period).
__global__ void kEnrightFaces(GridDev g, int axis, double coeff, double *u)
{
	const Coord c = threadCoord();
	if (!inGrid<3>(g, c)) return;
	const double pi = 3.14159265358979323846;
	const double x = g.o[0] + c.i * g.dx, y = g.o[1] + c.j * g.dx, z = g.o[2] + c.k * g.dx;
	double v;
	if (axis == 0) v = 2 * sin(pi * x) * sin(pi * x) * sin(2 * pi * y) * sin(2 * pi * z);
	else if (axis == 1) v = -sin(pi * y) * sin(pi * y) * sin(2 * pi * x) * sin(2 * pi * z);
	else v = -sin(pi * z) * sin(pi * z) * sin(2 * pi * x) * sin(2 * pi * y);
	u[gidx<3>(g, c.i, c.j, c.k)] = v * coeff;
}

void launchEnrightVelocity(const Fields &F, double coeff, hipStream_t s)
{
	for (int a = 0; a < 3; a++) hipLaunchKernelGGL(kEnrightFaces, gridDims(F.face[a], 3), gridBlock(), 0, s, F.face[a], a, coeff, F.u[a]);
}

// ---------------------------------------------------------- enforce (u)
template <int D>
__device__ __forceinline__ double lerpFace(const GridDev &g, const double *a, const double *p)
{
	int l[3]; double f[3];
	lowerFrac<D>(g, p, l, f);
	double w[2][3];
#pragma unroll
	for (int b = 0; b < D; b++) { w[0][b] = 1 - f[b]; w[1][b] = f[b]; }
	double val = 0;
#pragma unroll
	for (int k = 0; k < (D == 3 ? 2 : 1); k++)
#pragma unroll
		for (int j = 0; j < 2; j++)
#pragma unroll
			for (int i = 0; i < 2; i++) {
				const double wt = D == 2 ? w[i][0] * w[j][1] : w[i][0] * w[j][1] * w[k][2];
				val += a[gat<D>(g, l[0] + i, l[1] + j, D == 3 ? l[2] + k : 0)] * wt;
			}
	return val;
}

template <int D>
struct EnforceArgs {
	GridDev g[3];
	const double *u[3], *bvel[3], *bnormal[3], *frac[3];
	double *out[3];
};

// Boundary::enforce(velocity) (Boundary.cpp:89-120) touches only the faces
// with 0.1 < fraction <= 1 (:95); its `fraction == 1` branch (:115) is
// unreachable and omitted.  Those faces, and the solid's interpolated unit
// normal and velocity at them (fixed unless the boundary fields are
// replaced), are listed once; each enforce then computes the listed faces
// into a side buffer and writes them back in place, so the rest of the
// velocity is neither read nor copied.
template <int D>
__device__ __forceinline__ void decodeFace(const EnforceArgs<D> &A, const int *off, int gid, int &a, int &i, int &j, int &k)
{
	a = 0;
	while (a + 1 < D && gid >= off[a + 1]) a++;
	const int loc = gid - off[a];
	const GridDev &g = A.g[a];
	i = loc % g.n[0];
	const int rest = loc / g.n[0];
	j = rest % g.n[1];
	k = rest / g.n[1];
}

struct FaceOffsets { int off[4]; };

template <int D>
__global__ __launch_bounds__(kBlock) void kBoundaryFaces(EnforceArgs<D> A, int axis, FaceOffsets O, int *ids, int *cnt)
{
	const GridDev &g = A.g[axis];
	const Coord c = threadCoord();
	bool take = false;
	int gid = 0;
	if (inGrid<D>(g, c)) {
		const size_t id = gidx<D>(g, c.i, c.j, c.k);
		const double fr = A.frac[axis][id];
		take = fr > 0.1 && fr <= 1;
		gid = O.off[axis] + int(id);
	}
	const unsigned long long m = __ballot(take);
	if (m == 0) return;
	const int lane = int(__lane_id()), leader = __ffsll((long long)m) - 1;
	int base = 0;
	if (lane == leader) base = atomicAdd(cnt, __popcll(m));
	base = __shfl(base, leader, kWave);
	if (take) ids[base + __popcll(m & ((1ull << lane) - 1ull))] = gid;
}

__device__ __forceinline__ void facePos(const GridDev &g, int i, int j, int k, double *p)
{
	p[0] = g.o[0] + i * g.dx; p[1] = g.o[1] + j * g.dx; p[2] = g.o[2] + k * g.dx;
}

// record per listed face: unit normal n[0..2], solid velocity v0[0..2], out, pad
constexpr int kFaceRec = 8;

template <int D>
__global__ __launch_bounds__(kBlock) void kBoundaryRecords(EnforceArgs<D> A, FaceOffsets O, const int *ids, int n, double *rec)
{
	const int t = int(blockIdx.x * blockDim.x + threadIdx.x);
	if (t >= n) return;
	int ax, i, j, k;
	decodeFace<D>(A, O.off, ids[t], ax, i, j, k);
	double p[3];
	facePos(A.g[ax], i, j, k, p);
	double nr[3] = {0, 0, 0}, v0[3] = {0, 0, 0};
#pragma unroll
	for (int a = 0; a < D; a++) nr[a] = lerpFace<D>(A.g[a], A.bnormal[a], p);
	double s = 0;
#pragma unroll
	for (int a = 0; a < D; a++) s += nr[a] * nr[a];
	if (s > 0) { const double r = sqrt(s); for (int a = 0; a < D; a++) nr[a] /= r; }
#pragma unroll
	for (int a = 0; a < D; a++) v0[a] = lerpFace<D>(A.g[a], A.bvel[a], p);
	double *q = rec + size_t(t) * kFaceRec;
#pragma unroll
	for (int a = 0; a < 3; a++) { q[a] = nr[a]; q[3 + a] = v0[a]; }
}

template <int D>
__global__ __launch_bounds__(kBlock) void kEnforceFaces(EnforceArgs<D> A, FaceOffsets O, const int *ids, int n, double *rec)
{
	const int t = int(blockIdx.x * blockDim.x + threadIdx.x);
	if (t >= n) return;
	int axis, i, j, k;
	decodeFace<D>(A, O.off, ids[t], axis, i, j, k);
	const GridDev &g = A.g[axis];
	double *q = rec + size_t(t) * kFaceRec;
	double out = A.u[axis][gidx<D>(g, i, j, k)];
	bool any = false;
#pragma unroll
	for (int a = 0; a < D; a++) any = any || q[a] != 0;
	if (any) {
		double p[3];
		facePos(g, i, j, k, p);
		double dv[3], vel_ax = 0, vn = 0;
#pragma unroll
		for (int a = 0; a < D; a++) {
			const double vel = lerpFace<D>(A.g[a], A.u[a], p);
			dv[a] = vel - q[3 + a];
			if (a == axis) vel_ax = vel;
			vn += dv[a] * q[a];
		}
		if (vn < 0) {
			const double vt = dv[axis] - vn * q[axis];
			out = (vel_ax - vn * q[axis]) - 0.8 * vt;
		}
	}
	q[6] = out;
}

template <int D>
__global__ __launch_bounds__(kBlock) void kEnforceWrite(EnforceArgs<D> A, FaceOffsets O, const int *ids, int n, const double *rec)
{
	const int t = int(blockIdx.x * blockDim.x + threadIdx.x);
	if (t >= n) return;
	int axis, i, j, k;
	decodeFace<D>(A, O.off, ids[t], axis, i, j, k);
	A.out[axis][gidx<D>(A.g[axis], i, j, k)] = rec[size_t(t) * kFaceRec + 6];
}

// ------------------------------------------------------------- reductions
__global__ __launch_bounds__(kBlock) void kMaxAbsPartial(const double *a, size_t n, double *partials, int slot0)
{
	// four loads in flight per thread; the maximum does not depend on the order
	const size_t st = size_t(gridDim.x) * blockDim.x;
	size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
	double m = 0, m1 = 0;
	for (; i + 3 * st < n; i += 4 * st) {
		m = fmax(m, fmax(fabs(a[i]), fabs(a[i + st])));
		m1 = fmax(m1, fmax(fabs(a[i + 2 * st]), fabs(a[i + 3 * st])));
	}
	for (; i < n; i += st) m = fmax(m, fabs(a[i]));
	m = fmax(m, m1);
	__shared__ double red[kBlock / kWave];
	m = waveMax(m);
	const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
	if (lane == 0) red[wid] = m;
	__syncthreads();
	if (threadIdx.x == 0) {
		double r = 0;
		for (int w = 0; w < kBlock / kWave; w++) r = fmax(r, red[w]);
		partials[slot0 + blockIdx.x] = r;
	}
}

__global__ __launch_bounds__(kBlock) void kMaxFinal(const double *partials, int n, double *out)
{
	double m = 0;
	for (int i = threadIdx.x; i < n; i += blockDim.x) m = fmax(m, partials[i]);
	__shared__ double red[kBlock / kWave];
	m = waveMax(m);
	const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
	if (lane == 0) red[wid] = m;
	__syncthreads();
	if (threadIdx.x == 0) {
		double r = 0;
		for (int w = 0; w < kBlock / kWave; w++) r = fmax(r, red[w]);
		*out = r;
	}
}

// checkReferenceMap (LevelSetLiquid.cpp:1246-1283): per-block (tot, cnt).
template <int D>
__global__ __launch_bounds__(kBlock) void kCheckRefMap(GridDev g, const double *phiOri, const double *cellDist, Mat9 J,
	double threshold, double band, double *partials)
{
	const Coord c = threadCoord();
	double v[2] = {0, 0};
	if (inGrid<D>(g, c)) {
		const size_t id = gidx<D>(g, c.i, c.j, c.k);
		if (fabs(phiOri[id]) < band && cellDist[id] > 0) {
			v[0] = 1;
			double m[9];
#pragma unroll
			for (int e = 0; e < D * D; e++) m[e] = J.p[e][id];
			auto dot = [&](int a, int b) { double s = 0; for (int r = 0; r < D; r++) s += m[a * D + r] * m[b * D + r]; return s; };
			if constexpr (D == 2) {
				const double vab = dot(0, 1) * dot(0, 1) / dot(0, 0) / dot(1, 1);
				if (vab > threshold) v[1] = 1;
			} else {
				const double ab = dot(0, 1), ac = dot(0, 2), bc = dot(1, 2);
				const double vab = ab * ab / dot(0, 0) / dot(1, 1);
				const double vac = ac * ac / dot(0, 0) / dot(2, 2);
				const double vbc = bc * bc / dot(1, 1) / dot(2, 2);
				if (vab > threshold || vac > threshold || vbc > threshold) v[1] = 1;
			}
		}
	}
	double outp[2];
	blockSum<2>(v, outp);
	const int tid = threadIdx.x + blockDim.x * threadIdx.y;
	if (tid == 0) {
		const int b = flatBlock();
		partials[2 * b] = outp[0];
		partials[2 * b + 1] = outp[1];
	}
}

// The two partials per workgroup are counts (whole numbers in double, far
// below 2^53), so any summation order gives the same totals: one wide
// workgroup with two loads in flight per thread.
constexpr int kSumThreads = 1024;

__global__ __launch_bounds__(kSumThreads) void kSumFinal2(const double *partials, int nblocks, double *out)
{
	double a0 = 0, a1 = 0, b0 = 0, b1 = 0;
	int i = int(threadIdx.x);
	for (; i + kSumThreads < nblocks; i += 2 * kSumThreads) {
		a0 += partials[2 * i]; a1 += partials[2 * i + 1];
		b0 += partials[2 * (i + kSumThreads)]; b1 += partials[2 * (i + kSumThreads) + 1];
	}
	if (i < nblocks) { a0 += partials[2 * i]; a1 += partials[2 * i + 1]; }
	a0 = waveSum(a0 + b0);
	a1 = waveSum(a1 + b1);
	__shared__ double red[2][kSumThreads / kWave];
	const int lane = int(threadIdx.x) & (kWave - 1), wid = int(threadIdx.x) / kWave;
	if (lane == 0) { red[0][wid] = a0; red[1][wid] = a1; }
	__syncthreads();
	if (threadIdx.x == 0) {
		double s0 = 0, s1 = 0;
		for (int w = 0; w < kSumThreads / kWave; w++) { s0 += red[0][w]; s1 += red[1][w]; }
		out[0] = s0;
		out[1] = s1;
	}
}

// --------------------------------------------------------------- launchers
template <typename K, typename... Args>
static void launchCells(int D, const GridDev &g, hipStream_t s, K k2, K k3, Args... args)
{
	if (D == 2) hipLaunchKernelGGL(k2, gridDims(g, 2), gridBlock(), 0, s, args...);
	else hipLaunchKernelGGL(k3, gridDims(g, 3), gridBlock(), 0, s, args...);
}

void launchUnionBox(const Fields &F, const double *ce, const double *hl, hipStream_t s)
{
	launchCells(F.D, F.cell, s, kUnionBox<2>, kUnionBox<3>, F.cell, F.phi, ce[0], ce[1], ce[2], hl[0], hl[1], hl[2]);
}
void launchUnionSphere(const Fields &F, const double *ce, double r, hipStream_t s)
{
	launchCells(F.D, F.cell, s, kUnionSphere<2>, kUnionSphere<3>, F.cell, F.phi, ce[0], ce[1], ce[2], r);
}
void launchUnionPlane(const Fields &F, const double *p, const double *n, hipStream_t s)
{
	launchCells(F.D, F.cell, s, kUnionPlane<2>, kUnionPlane<3>, F.cell, F.phi, p[0], p[1], p[2], n[0], n[1], n[2]);
}

void launchResetRefMap(const Fields &F, hipStream_t s)
{
	Mat9 J{};
	for (int e = 0; e < F.D * F.D; e++) J.p[e] = F.J[e];
	launchCells(F.D, F.cell, s, kResetRefMap<2>, kResetRefMap<3>, F.cell, F.xi[0], F.xi[1], F.xi[2], J);
}

void launchFill(double *p, double v, size_t n, hipStream_t s) { hipLaunchKernelGGL(kFill, dim3(linearBlocks(n)), dim3(kBlock), 0, s, p, v, n); }
void launchFillU8(uint8_t *p, uint8_t v, size_t n, hipStream_t s) { hipLaunchKernelGGL(kFillU8, dim3(linearBlocks(n)), dim3(kBlock), 0, s, p, v, n); }

__global__ void kOpenFaces(const double *frac, uint8_t *open, size_t n)
{
	for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) open[i] = frac[i] < 1 ? 1 : 0;
}

void launchOpenFaces(const Fields &F, hipStream_t s)
{
	for (int a = 0; a < F.D; a++)
		hipLaunchKernelGGL(kOpenFaces, dim3(linearBlocks(F.Nf[a])), dim3(kBlock), 0, s, (const double *)F.frac[a], F.open_f[a], F.Nf[a]);
}
void launchCopy(double *d, const double *src, size_t n, hipStream_t s) { hipLaunchKernelGGL(kCopy, dim3(linearBlocks(n)), dim3(kBlock), 0, s, d, src, n); }

void launchEnforceLevelSet(const Fields &F, double *sdf, double tol, hipStream_t s)
{
	hipLaunchKernelGGL(kEnforceLevelSet, dim3(linearBlocks(F.N)), dim3(kBlock), 0, s, sdf, (const double *)F.cell_dist, tol, F.N);
}

void launchGravity(const Fields &F, double dt, hipStream_t s)
{
	hipLaunchKernelGGL(kGravity, dim3(linearBlocks(F.Nf[1])), dim3(kBlock), 0, s, F.u[1], 9.8 * dt, F.Nf[1]);
}

template <int D>
static EnforceArgs<D> enforceArgs(const Fields &F, FaceOffsets &O)
{
	EnforceArgs<D> A{};
	int off = 0;
	for (int a = 0; a < D; a++) {
		A.g[a] = F.face[a]; A.u[a] = F.u[a]; A.bvel[a] = F.bvel[a]; A.bnormal[a] = F.bnormal[a]; A.frac[a] = F.frac[a]; A.out[a] = F.u[a];
		O.off[a] = off;
		off += int(F.Nf[a]);
	}
	O.off[D] = off;
	return A;
}

template <int D>
static void boundaryFacesD(const Fields &F, int *ids, int *cnt, hipStream_t s)
{
	FaceOffsets O{};
	const EnforceArgs<D> A = enforceArgs<D>(F, O);
	for (int a = 0; a < D; a++) hipLaunchKernelGGL(kBoundaryFaces<D>, gridDims(F.face[a], D), gridBlock(), 0, s, A, a, O, ids, cnt);
}

void launchBoundaryFaces(const Fields &F, int *ids, int *cnt, hipStream_t s)
{
	if (F.D == 2) boundaryFacesD<2>(F, ids, cnt, s); else boundaryFacesD<3>(F, ids, cnt, s);
}

template <int D>
static void boundaryRecordsD(const Fields &F, hipStream_t s)
{
	FaceOffsets O{};
	const EnforceArgs<D> A = enforceArgs<D>(F, O);
	const unsigned nb = unsigned((F.nbface + kBlock - 1) / kBlock);
	hipLaunchKernelGGL(kBoundaryRecords<D>, dim3(nb), dim3(kBlock), 0, s, A, O, (const int *)F.bface, F.nbface, F.bface_rec);
}

void launchBoundaryRecords(const Fields &F, hipStream_t s)
{
	if (F.nbface == 0) return;
	if (F.D == 2) boundaryRecordsD<2>(F, s); else boundaryRecordsD<3>(F, s);
}

template <int D>
static void enforceD(const Fields &F, hipStream_t s)
{
	FaceOffsets O{};
	const EnforceArgs<D> A = enforceArgs<D>(F, O);
	const unsigned nb = unsigned((F.nbface + kBlock - 1) / kBlock);
	hipLaunchKernelGGL(kEnforceFaces<D>, dim3(nb), dim3(kBlock), 0, s, A, O, (const int *)F.bface, F.nbface, F.bface_rec);
	hipLaunchKernelGGL(kEnforceWrite<D>, dim3(nb), dim3(kBlock), 0, s, A, O, (const int *)F.bface, F.nbface, (const double *)F.bface_rec);
}

void launchEnforceBoundary(const Fields &F, hipStream_t s)
{
	if (F.nbface == 0) return;
	if (F.D == 2) enforceD<2>(F, s); else enforceD<3>(F, s);
}

void launchMaxSpeed(const Fields &F, hipStream_t s)
{
	int slot = 0;
	for (int a = 0; a < F.D; a++) {
		const unsigned nb = 1024;
		hipLaunchKernelGGL(kMaxAbsPartial, dim3(nb), dim3(kBlock), 0, s, (const double *)F.u[a], F.Nf[a], F.partials, slot);
		slot += nb;
	}
	hipLaunchKernelGGL(kMaxFinal, dim3(1), dim3(kBlock), 0, s, (const double *)F.partials, slot, F.dscalars);
}

void launchCheckRefMap(const Fields &F, double threshold, double band, hipStream_t s)
{
	const dim3 gd = gridDims(F.cell, F.D);
	const int nb = int(gd.x * gd.y * gd.z);
	Mat9 J{};
	for (int e = 0; e < F.D * F.D; e++) J.p[e] = F.J[e];
	launchCells(F.D, F.cell, s, kCheckRefMap<2>, kCheckRefMap<3>, F.cell, (const double *)F.phi_ori, (const double *)F.cell_dist,
		J, threshold, band, F.partials);
	hipLaunchKernelGGL(kSumFinal2, dim3(1), dim3(kSumThreads), 0, s, (const double *)F.partials, nb, F.dscalars + 1);
}

} // namespace garm
