// Multigrid V-cycle preconditioner for the ghost-fluid Poisson system
// (the operator of Projector::buildLinearSystem, Projector.cpp:30-68).
//
// No counterpart in the reference, which preconditions with Eigen's
// incomplete Cholesky (LinearSystem.h:10-39).  The hierarchy is algebraic:
// level 0 is the assembled operator itself laid out on the bounding box of
// the unknowns (diagonal plus the weights of each cell's -x, -y, -z faces),
// and level l+1 is the Galerkin product P^T A P with P the piecewise-constant
// prolongation from 2x2x2 aggregates, which keeps a 7-point stencil:
//   coarse face weight = sum of the fine face weights crossing it,
//   coarse diagonal    = sum of the fine diagonals - 2 * sum of the weights
//                        of the faces inside the aggregate.
// Free surfaces (ghost-fluid terms in the diagonal) and solid faces (zero
// weights) carry over to every level without special cases.  Smoothing is
// red-black Gauss-Seidel, two sweeps down and two up in mirrored colour order
// so the cycle is symmetric; the coarsest level (<= 512 cells) is relaxed by
// one workgroup.  Everything is FP32 on structured grids: unit-stride,
// coalesced neighbours, no index arrays.
#include "garm_multigrid.hpp"

#include <algorithm>

namespace garm {

constexpr int kMgSweeps = 2;            // red-black sweep pairs before and after the coarse correction
constexpr float kMgProlongScale = 1.9f; // over-correction of the coarse correction (aggregation underestimates it; 1.9 measured best at 256^3)
constexpr int kMgCoarsestCells = 512;
constexpr int kMgCoarsestSweeps = 30;

struct MgLevel {
	int n[3];
	float *diag, *wx, *wy, *wz, *b, *x;
};

static MgLevel level(const Multigrid &M, int l)
{
	MgLevel L;
	for (int a = 0; a < 3; a++) L.n[a] = M.dims[l][a];
	L.diag = M.diag[l]; L.wx = M.wx[l]; L.wy = M.wy[l]; L.wz = M.wz[l]; L.b = M.b[l]; L.x = M.x[l];
	return L;
}

__device__ __forceinline__ size_t cidx(const MgLevel &L, int i, int j, int k) { return (size_t(k) * L.n[1] + j) * L.n[0] + i; }

__global__ void kMgZero(float *p, size_t n)
{
	for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) p[i] = 0.0f;
}

struct MgBox {
	int org[3], gn[3];
};

__device__ __forceinline__ size_t boxCell(const MgBox &B, const MgLevel &L, int g)
{
	const int i = g % B.gn[0] - B.org[0];
	const int j = (g / B.gn[0]) % B.gn[1] - B.org[1];
	const int k = g / (B.gn[0] * B.gn[1]) - B.org[2];
	return cidx(L, i, j, k);
}

// level 0: the operator entries of every unknown at its box cell
template <int D>
__global__ __launch_bounds__(kBlock) void kMgBuild0(Rows<float> A, int n, const int *mat2grid, MgBox B, MgLevel L)
{
	for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
		const size_t c = boxCell(B, L, mat2grid[r]);
		L.diag[c] = A.diag[r];
		L.wx[c] = A.nb[0][r] >= 0 ? A.w[0][r] : 0.0f;
		L.wy[c] = A.nb[2][r] >= 0 ? A.w[2][r] : 0.0f;
		if constexpr (D == 3) L.wz[c] = A.nb[4][r] >= 0 ? A.w[4][r] : 0.0f;
	}
}

// Galerkin coarsening by 2x2x2 (2x2 in 2D) aggregates
template <int D>
__global__ __launch_bounds__(kBlock) void kMgCoarsen(MgLevel F, MgLevel C)
{
	const size_t cells = size_t(C.n[0]) * C.n[1] * C.n[2];
	for (size_t q = blockIdx.x * size_t(blockDim.x) + threadIdx.x; q < cells; q += size_t(gridDim.x) * blockDim.x) {
		const int I = int(q % C.n[0]), J = int((q / C.n[0]) % C.n[1]), K = int(q / (size_t(C.n[0]) * C.n[1]));
		float d = 0, inner = 0, sx = 0, sy = 0, sz = 0;
#pragma unroll
		for (int cz = 0; cz < (D == 3 ? 2 : 1); cz++)
#pragma unroll
			for (int cy = 0; cy < 2; cy++)
#pragma unroll
				for (int cx = 0; cx < 2; cx++) {
					const int i = 2 * I + cx, j = 2 * J + cy, k = 2 * K + cz;
					if (i >= F.n[0] || j >= F.n[1] || k >= F.n[2]) continue;
					const size_t f = cidx(F, i, j, k);
					d += F.diag[f];
					if (cx) inner += F.wx[f]; else sx += F.wx[f];
					if (cy) inner += F.wy[f]; else sy += F.wy[f];
					if constexpr (D == 3) { if (cz) inner += F.wz[f]; else sz += F.wz[f]; }
				}
		C.diag[q] = d - 2 * inner;
		C.wx[q] = sx;
		C.wy[q] = sy;
		if constexpr (D == 3) C.wz[q] = sz;
	}
}

// (A x)_c without the diagonal: the sum of face weight times neighbour
template <int D>
__device__ __forceinline__ float offSum(const MgLevel &L, size_t c, int i, int j, int k)
{
	const size_t sy = L.n[0], sz = size_t(L.n[0]) * L.n[1];
	float s = 0;
	if (i > 0) s += L.wx[c] * L.x[c - 1];
	if (i + 1 < L.n[0]) s += L.wx[c + 1] * L.x[c + 1];
	if (j > 0) s += L.wy[c] * L.x[c - sy];
	if (j + 1 < L.n[1]) s += L.wy[c + sy] * L.x[c + sy];
	if constexpr (D == 3) {
		if (k > 0) s += L.wz[c] * L.x[c - sz];
		if (k + 1 < L.n[2]) s += L.wz[c + sz] * L.x[c + sz];
	}
	return s;
}

// one colour of red-black Gauss-Seidel; threads cover the cells of that colour only
template <int D>
__global__ __launch_bounds__(kBlock) void kMgSmooth(MgLevel L, int color)
{
	const int hx = (L.n[0] + 1) / 2;
	const size_t half = size_t(hx) * L.n[1] * L.n[2];
	for (size_t q = blockIdx.x * size_t(blockDim.x) + threadIdx.x; q < half; q += size_t(gridDim.x) * blockDim.x) {
		const int ii = int(q % hx), j = int((q / hx) % L.n[1]), k = int(q / (size_t(hx) * L.n[1]));
		const int i = 2 * ii + ((j + k + color) & 1);
		if (i >= L.n[0]) continue;
		const size_t c = cidx(L, i, j, k);
		const float d = L.diag[c];
		if (!(d > 0)) continue;
		L.x[c] = (L.b[c] + offSum<D>(L, c, i, j, k)) / d;
	}
}

// coarse right-hand side = sum of the fine residuals of the aggregate; coarse iterate = 0
template <int D>
__global__ __launch_bounds__(kBlock) void kMgRestrict(MgLevel F, MgLevel C)
{
	const size_t cells = size_t(C.n[0]) * C.n[1] * C.n[2];
	for (size_t q = blockIdx.x * size_t(blockDim.x) + threadIdx.x; q < cells; q += size_t(gridDim.x) * blockDim.x) {
		const int I = int(q % C.n[0]), J = int((q / C.n[0]) % C.n[1]), K = int(q / (size_t(C.n[0]) * C.n[1]));
		float s = 0;
#pragma unroll
		for (int cz = 0; cz < (D == 3 ? 2 : 1); cz++)
#pragma unroll
			for (int cy = 0; cy < 2; cy++)
#pragma unroll
				for (int cx = 0; cx < 2; cx++) {
					const int i = 2 * I + cx, j = 2 * J + cy, k = 2 * K + cz;
					if (i >= F.n[0] || j >= F.n[1] || k >= F.n[2]) continue;
					const size_t f = cidx(F, i, j, k);
					const float d = F.diag[f];
					if (d > 0) s += F.b[f] - (d * F.x[f] - offSum<D>(F, f, i, j, k));
				}
		C.b[q] = s;
		C.x[q] = 0.0f;
	}
}

// fine iterate += coarse iterate of its aggregate
template <int D>
__global__ __launch_bounds__(kBlock) void kMgProlong(MgLevel F, MgLevel C)
{
	const size_t cells = size_t(F.n[0]) * F.n[1] * F.n[2];
	for (size_t q = blockIdx.x * size_t(blockDim.x) + threadIdx.x; q < cells; q += size_t(gridDim.x) * blockDim.x) {
		if (!(F.diag[q] > 0)) continue;
		const int i = int(q % F.n[0]), j = int((q / F.n[0]) % F.n[1]), k = int(q / (size_t(F.n[0]) * F.n[1]));
		F.x[q] += kMgProlongScale * C.x[cidx(C, i / 2, j / 2, D == 3 ? k / 2 : 0)];
	}
}

// the coarsest level, relaxed by one workgroup (its iterate is zero on entry)
template <int D>
__global__ __launch_bounds__(1024) void kMgCoarsest(MgLevel L, int sweeps)
{
	const size_t cells = size_t(L.n[0]) * L.n[1] * L.n[2];
	for (int s = 0; s < 2 * sweeps; s++) {
		const int color = (s & 1) ^ (s >= sweeps ? 1 : 0);   // red-black down, black-red up
		for (size_t q = threadIdx.x; q < cells; q += blockDim.x) {
			const int i = int(q % L.n[0]), j = int((q / L.n[0]) % L.n[1]), k = int(q / (size_t(L.n[0]) * L.n[1]));
			if (((i + j + k) & 1) != color) continue;
			const float d = L.diag[q];
			if (d > 0) L.x[q] = (L.b[q] + offSum<D>(L, q, i, j, k)) / d;
		}
		__syncthreads();
	}
}

// compact FP64 residual -> level-0 right-hand side (the level's b and x were zeroed)
__global__ void kMgLoad(int n, const int *mat2grid, MgBox B, MgLevel L, const double *r)
{
	for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) L.b[boxCell(B, L, mat2grid[q])] = float(r[q]);
}

__global__ void kMgStore(int n, const int *mat2grid, MgBox B, MgLevel L, double *z)
{
	for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) z[q] = double(L.x[boxCell(B, L, mat2grid[q])]);
}

// ------------------------------------------------------------------ host
static unsigned mgBlocks(size_t n) { return unsigned(std::max<size_t>(1, std::min<size_t>((n + kBlock - 1) / kBlock, 4096))); }

Multigrid::~Multigrid()
{
	if (pool) (void)hipFree(pool);
}

bool Multigrid::layout(int dim, const int box_org[3], const int box_n[3], const int grid_n[3], std::string &err)
{
	D = dim;
	for (int a = 0; a < 3; a++) { org[a] = box_org[a]; gn[a] = grid_n[a]; dims[0][a] = box_n[a]; }
	levels = 1;
	size_t total = 0;
	for (int l = 0;; l++) {
		cells[l] = size_t(dims[l][0]) * dims[l][1] * dims[l][2];
		total += cells[l];
		if (cells[l] <= size_t(kMgCoarsestCells) || l + 1 == kMgMaxLevels) break;
		for (int a = 0; a < 3; a++) dims[l + 1][a] = (a < D) ? (dims[l][a] + 1) / 2 : 1;
		levels++;
	}
	const size_t need = 6 * total + 64;
	if (need > pool_cap) {
		if (pool) (void)hipFree(pool);
		pool = nullptr;
		const size_t want = need + need / 4;
		if (hipMalloc(&pool, want * sizeof(float)) != hipSuccess) { err = "multigrid allocation failed"; pool_cap = 0; return false; }
		pool_cap = want;
	}
	float *p = pool;
	for (int l = 0; l < levels; l++) {
		float **arr[6] = {&diag[l], &wx[l], &wy[l], &wz[l], &b[l], &x[l]};
		for (float **a : arr) { *a = p; p += cells[l]; }
	}
	return true;
}

void mgBuild(Multigrid &M, const Rows<float> &A, int n, const int *mat2grid, hipStream_t s)
{
	// every level array starts at zero (inactive cells keep zero entries)
	size_t total = 0;
	for (int l = 0; l < M.levels; l++) total += 6 * M.cells[l];
	hipLaunchKernelGGL(kMgZero, dim3(mgBlocks(total)), dim3(kBlock), 0, s, M.pool, total);
	MgBox B;
	for (int a = 0; a < 3; a++) { B.org[a] = M.org[a]; B.gn[a] = M.gn[a]; }
	const MgLevel L0 = level(M, 0);
	if (M.D == 2) hipLaunchKernelGGL(kMgBuild0<2>, dim3(mgBlocks(size_t(n))), dim3(kBlock), 0, s, A, n, mat2grid, B, L0);
	else hipLaunchKernelGGL(kMgBuild0<3>, dim3(mgBlocks(size_t(n))), dim3(kBlock), 0, s, A, n, mat2grid, B, L0);
	for (int l = 0; l + 1 < M.levels; l++) {
		const MgLevel F = level(M, l), C = level(M, l + 1);
		if (M.D == 2) hipLaunchKernelGGL(kMgCoarsen<2>, dim3(mgBlocks(M.cells[l + 1])), dim3(kBlock), 0, s, F, C);
		else hipLaunchKernelGGL(kMgCoarsen<3>, dim3(mgBlocks(M.cells[l + 1])), dim3(kBlock), 0, s, F, C);
	}
}

template <int D>
static void vcycle(Multigrid &M, const int *mat2grid, int n, const double *r, double *z, hipStream_t s)
{
	MgBox B;
	for (int a = 0; a < 3; a++) { B.org[a] = M.org[a]; B.gn[a] = M.gn[a]; }
	const MgLevel L0 = level(M, 0);
	hipLaunchKernelGGL(kMgZero, dim3(mgBlocks(2 * M.cells[0])), dim3(kBlock), 0, s, L0.b, 2 * M.cells[0]);   // b and x are adjacent
	hipLaunchKernelGGL(kMgLoad, dim3(mgBlocks(size_t(n))), dim3(kBlock), 0, s, n, mat2grid, B, L0, r);
	const int top = M.levels - 1;
	for (int l = 0; l < top; l++) {
		const MgLevel L = level(M, l);
		const unsigned g = mgBlocks((M.cells[l] + 1) / 2);
		for (int v = 0; v < kMgSweeps; v++) {
			hipLaunchKernelGGL(kMgSmooth<D>, dim3(g), dim3(kBlock), 0, s, L, 0);
			hipLaunchKernelGGL(kMgSmooth<D>, dim3(g), dim3(kBlock), 0, s, L, 1);
		}
		hipLaunchKernelGGL(kMgRestrict<D>, dim3(mgBlocks(M.cells[l + 1])), dim3(kBlock), 0, s, L, level(M, l + 1));
	}
	hipLaunchKernelGGL(kMgCoarsest<D>, dim3(1), dim3(1024), 0, s, level(M, top), kMgCoarsestSweeps);
	for (int l = top - 1; l >= 0; l--) {
		const MgLevel L = level(M, l);
		hipLaunchKernelGGL(kMgProlong<D>, dim3(mgBlocks(M.cells[l])), dim3(kBlock), 0, s, L, level(M, l + 1));
		const unsigned g = mgBlocks((M.cells[l] + 1) / 2);
		for (int v = 0; v < kMgSweeps; v++) {
			hipLaunchKernelGGL(kMgSmooth<D>, dim3(g), dim3(kBlock), 0, s, L, 1);
			hipLaunchKernelGGL(kMgSmooth<D>, dim3(g), dim3(kBlock), 0, s, L, 0);
		}
	}
	hipLaunchKernelGGL(kMgStore, dim3(mgBlocks(size_t(n))), dim3(kBlock), 0, s, n, mat2grid, B, L0, z);
}

void mgApply(Multigrid &M, const int *mat2grid, int n, const double *r, double *z, hipStream_t s)
{
	if (M.D == 2) vcycle<2>(M, mat2grid, n, r, z, s);
	else vcycle<3>(M, mat2grid, n, r, z, s);
}

} // namespace garm
