// Shared device-side definitions for the gfx950 GARM-LS kernels.
//
// Layout: every field lives in HBM as structure-of-arrays, one contiguous
// x-fastest array per scalar component (the reference stores vectors and
// matrices per cell, GridBasedData.h:22-24; the index order is Grid::index,
// Grid.h:40-44).  Sampling clamps the coordinate into the grid exactly like
// GridBasedData::at (GridBasedData.h:47).
#pragma once

#include <hip/hip_runtime.h>
#include <chrono>
#include <cstddef>
#include <cstdint>

namespace garm {

// One regular grid (Grid.h:10-64): sizes, origin, spacing.  Passed by value.
struct GridDev {
	int n[3];
	double o[3];
	double dx, inv;
};

constexpr int kBlock = 256;       // threads per workgroup (4 wave64)

// Host wait for a recorded event: polls it for up to kSpinMicros, so the
// wait for a small readback ends within microseconds of its copy where a
// blocking wait can leave the thread asleep for tens of microseconds, then
// blocks in hipEventSynchronize so a long wait does not hold a host core.
constexpr long kSpinMicros = 2000;
inline hipError_t pollEvent(hipEvent_t ev)
{
	const auto t0 = std::chrono::steady_clock::now();
	for (unsigned n = 1;; n++) {
		const hipError_t e = hipEventQuery(ev);
		if (e != hipErrorNotReady) return e;
		if ((n & 255) == 0 &&
			std::chrono::duration_cast<std::chrono::microseconds>(std::chrono::steady_clock::now() - t0).count() > kSpinMicros)
			return hipEventSynchronize(ev);
	}
}

// Host wait for the work enqueued on s so far (ev created with
// hipEventDisableTiming).
inline hipError_t spinWait(hipStream_t s, hipEvent_t ev)
{
	const hipError_t e = hipEventRecord(ev, s);
	if (e != hipSuccess) return e;
	return pollEvent(ev);
}
constexpr int kWave = 64;

template <int D>
__device__ __forceinline__ size_t gidx(const GridDev &g, int i, int j, int k)
{
	if constexpr (D == 2) return size_t(i) + size_t(g.n[0]) * size_t(j);
	else return size_t(i) + size_t(g.n[0]) * (size_t(j) + size_t(g.n[1]) * size_t(k));
}

__device__ __forceinline__ int clampi(int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); }

// GridBasedData::at: index of the clamped coordinate.
template <int D>
__device__ __forceinline__ size_t gat(const GridDev &g, int i, int j, int k)
{
	if constexpr (D == 2) return gidx<2>(g, clampi(i, g.n[0] - 1), clampi(j, g.n[1] - 1), 0);
	else return gidx<3>(g, clampi(i, g.n[0] - 1), clampi(j, g.n[1] - 1), clampi(k, g.n[2] - 1));
}

// Offsets of the clamped corner coordinates l + w, w = 0..W-1, per axis
// (GridBasedData::at): the index of corner (i, j, k) of the W^D box is
// ox[i] + oy[j] + oz[k], W clamps per axis instead of W^D clamped indices.
// Grids hold fewer than 2^31 elements (checked at creation).
template <int D, int W>
__device__ __forceinline__ void cornerOffsets(const GridDev &g, const int *l, unsigned (&ox)[W], unsigned (&oy)[W], unsigned (&oz)[W])
{
	const unsigned sy = unsigned(g.n[0]), sz = unsigned(g.n[0]) * unsigned(g.n[1]);
#pragma unroll
	for (int w = 0; w < W; w++) {
		ox[w] = unsigned(clampi(l[0] + w, g.n[0] - 1));
		oy[w] = unsigned(clampi(l[1] + w, g.n[1] - 1)) * sy;
		oz[w] = D == 3 ? unsigned(clampi(l[2] + w, g.n[2] - 1)) * sz : 0u;
	}
}

// Cell/face coordinate of this thread in a (64, 4, 1) x (.., .., n2) launch.
struct Coord { int i, j, k; };

__device__ __forceinline__ Coord threadCoord()
{
	return Coord{int(blockIdx.x * blockDim.x + threadIdx.x), int(blockIdx.y * blockDim.y + threadIdx.y), int(blockIdx.z)};
}

template <int D>
__device__ __forceinline__ bool inGrid(const GridDev &g, const Coord &c)
{
	return c.i < g.n[0] && c.j < g.n[1] && (D == 2 || c.k < g.n[2]);
}

// Grid::getLinearLower and Grid::getLowerFrac (Grid.h:52,55).
template <int D>
__device__ __forceinline__ void lowerFrac(const GridDev &g, const double *p, int *l, double *f)
{
#pragma unroll
	for (int a = 0; a < D; a++) {
		l[a] = int(floor((p[a] - g.o[a]) * g.inv));
		f[a] = (p[a] - g.o[a] - l[a] * g.dx) * g.inv;
	}
}

// The same in compute type T (double reproduces lowerFrac exactly).
template <int D, typename T>
__device__ __forceinline__ void lowerFracT(const GridDev &g, const T *p, int *l, T *f)
{
#pragma unroll
	for (int a = 0; a < D; a++) {
		const T o = T(g.o[a]), inv = T(g.inv);
		l[a] = int(floor((p[a] - o) * inv));
		f[a] = (p[a] - o - l[a] * T(g.dx)) * inv;
	}
}

// Launch geometry for one grid: x is the fastest axis and one wave64 covers
// 64 consecutive x-cells so every component array is read in 512-byte runs.
inline dim3 gridBlock() { return dim3(64, 4, 1); }
inline dim3 gridDims(const GridDev &g, int D)
{
	return dim3((g.n[0] + 63) / 64, (g.n[1] + 3) / 4, D == 3 ? g.n[2] : 1);
}

// ---- deterministic reductions ---------------------------------------------
// Block sum of a double: wave64 shuffle tree, then the 4 wave partials in a
// fixed order.  No atomics, so FP64 results repeat bit for bit.
__device__ __forceinline__ double waveSum(double v)
{
#pragma unroll
	for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
	return v;
}

__device__ __forceinline__ double waveMax(double v)
{
#pragma unroll
	for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, kWave));
	return v;
}

template <int NV>
__device__ __forceinline__ void blockSum(double (&v)[NV], double *out /* block's NV partials */)
{
	__shared__ double red[NV][kBlock / kWave];
	const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
	const int lane = tid & (kWave - 1), wid = tid / kWave;
#pragma unroll
	for (int q = 0; q < NV; q++) {
		const double s = waveSum(v[q]);
		if (lane == 0) red[q][wid] = s;
	}
	__syncthreads();
	if (tid == 0) {
#pragma unroll
		for (int q = 0; q < NV; q++) {
			double s = 0;
			for (int w = 0; w < kBlock / kWave; w++) s += red[q][w];
			out[q] = s;
		}
	}
}

// ---- tile scans --------------------------------------------------------------
// Index spaces are numbered in tiles of kScanTile consecutive ids, one
// workgroup per tile (kScanSub passes of kBlock ids).
constexpr int kScanSub = 16;                   // passes of kBlock cells per tile
constexpr int kScanTile = kBlock * kScanSub;

// Exclusive scan of one int per thread over a 256-thread workgroup; *total
// receives the workgroup sum.
__device__ __forceinline__ int blockExclusiveScan(int v, int *total)
{
	__shared__ int waveSum[kBlock / kWave];
	const int tid = int(threadIdx.x), lane = tid & (kWave - 1), wid = tid / kWave;
	int x = v;
#pragma unroll
	for (int off = 1; off < kWave; off <<= 1) {
		const int y = __shfl_up(x, off, kWave);
		if (lane >= off) x += y;
	}
	if (lane == kWave - 1) waveSum[wid] = x;
	__syncthreads();
	int before = 0, all = 0;
#pragma unroll
	for (int w = 0; w < kBlock / kWave; w++) {
		const int t = waveSum[w];
		before += w < wid ? t : 0;
		all += t;
	}
	__syncthreads();   // waveSum is reused by the next call
	*total = all;
	return before + x - v;
}

__device__ __forceinline__ int flatBlock() { return int(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)); }

// Workgroups are handed to the eight XCDs round-robin in launch order (x
// fastest), so neighbouring tiles of a plain launch land on different XCDs,
// each with its own L2, and the halo they share is fetched once per XCD.
// xcdTiles renumbers a launch of nx x ny x nz tiles so that XCD x takes one
// contiguous run of tiles in x-fastest order (runs differ in length by at
// most one): neighbours along x and y, and along z inside a run, share an L2.
struct Tile3 { int x, y, z; };
constexpr unsigned kXcds = 8;
__device__ __forceinline__ Tile3 xcdTiles()
{
	const unsigned nx = gridDim.x, ny = gridDim.y, total = nx * ny * gridDim.z;
	const unsigned b = blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z);
	const unsigned q = total / kXcds, r = total % kXcds, x = b % kXcds, k = b / kXcds;
	const unsigned t = x * q + min(x, r) + k;
	return Tile3{int(t % nx), int((t / nx) % ny), int(t / (nx * ny))};
}

// The same for a one-dimensional launch over a list: the index in [0,
// gridDim.x) this workgroup takes, contiguous runs per XCD.
__device__ __forceinline__ int xcdIndex()
{
	const unsigned total = gridDim.x, b = blockIdx.x;
	const unsigned q = total / kXcds, r = total % kXcds, x = b % kXcds, k = b / kXcds;
	return int(x * q + min(x, r) + k);
}

} // namespace garm
