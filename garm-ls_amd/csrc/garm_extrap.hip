// Layered extrapolation (Extrapolation.h:44-218) by frontier lists on gfx950.
//
// The reference sweeps the whole grid once per layer, each not-yet-reached
// element taking the mean of its reached neighbours.  Here each element
// carries the layer at which it is reached (0 fixed, 255 not yet).  One
// full-grid pass marks the fixed elements and lists those with an unfixed
// neighbour; every later layer L walks only the list of layer L-1: each
// unreached neighbour is taken by exactly one element of that list (its
// first neighbour at layer L-1 in the reference's ordinal order -x, +x, -y,
// +y, -z, +z), which gives it the mean of its neighbours below layer L (in
// that order, so the value is the reference's exactly) and lists it for
// layer L+1.  Work per layer is proportional to the front, not to the grid,
// and takes no atomics but the list appends.  Lists are
// appended a workgroup at a time (a prefix sum in LDS and one atomic on the
// list's counter per workgroup and pass): atomics on one address serialise
// in L2, and a front of 10^5 elements appended one by one costs more than
// the rest of the layer.
#include "garm_internal.hpp"

namespace garm {

constexpr uint8_t kFar = 255;
constexpr int kFrontGrid = 1024;

// One to three grids whose elements are numbered consecutively.
struct LayerSpace {
	int ng;
	GridDev g[3];
	uint8_t *layer[3];
	double *v[3][3];       // v[grid][component]
	int nv;                // components carried per element
	int off[4];
};

enum FixRule { kFixCellDistPositive = 0, kFixAdvRegion = 1, kFixPhiOriNonPositive = 2, kFixLiquidFace = 3 };

struct FixArgs {
	int rule;
	GridDev cell;
	const double *cellDist;
	const uint8_t *adv;
	const double *phiOri;
	const uint8_t *open[3];    // 1 where the face fraction is < 1
	const uint8_t *liquid;     // if set: 0 exactly where phi_ori <= 0 (the marked advRegion front)
};

template <int D>
__device__ __forceinline__ bool validIn(const GridDev &g, int i, int j, int k)
{
	return i >= 0 && i < g.n[0] && j >= 0 && j < g.n[1] && (D == 2 || (k >= 0 && k < g.n[2]));
}

template <int D, int RULE>
__device__ __forceinline__ bool isFixed(const FixArgs &P, int a, const GridDev &g, int i, int j, int k)
{
	const size_t id = gidx<D>(g, i, j, k);
	if constexpr (RULE == kFixCellDistPositive) return P.cellDist[id] > 0;   // LevelSetLiquid.cpp:1090-1101
	else if constexpr (RULE == kFixAdvRegion) return P.adv[id] > 0;           // LevelSetLiquid.cpp:1062-1064
	else if constexpr (RULE == kFixPhiOriNonPositive) return P.phiOri[id] <= 0;   // LevelSetLiquid.cpp:1068-1070
	else {                                                                    // LevelSetLiquid.cpp:1072-1083
		if (!P.open[a][id]) return false;   // fraction < 1
		const int i0 = i - (a == 0), j0 = j - (a == 1), k0 = k - (a == 2);
		if (!validIn<D>(P.cell, i0, j0, k0) || !validIn<D>(P.cell, i, j, k)) return false;
		const size_t c0 = gidx<D>(P.cell, i0, j0, k0), c1 = gidx<D>(P.cell, i, j, k);
		if (P.liquid) return P.liquid[c0] == 0 || P.liquid[c1] == 0;
		return P.phiOri[c0] <= 0 || P.phiOri[c1] <= 0;
	}
}

// Workgroup-wide append: every thread of the (256-thread) workgroup calls
// this with the number of entries it will write; the return value is where
// its entries start.  One atomic on *cnt per workgroup.
__device__ __forceinline__ int blockAppendBase(int *cnt, int count)
{
	__shared__ int waveOff[kBlock / kWave];
	__shared__ int blockBase;
	const int tid = int(threadIdx.x + blockDim.x * threadIdx.y);
	const int lane = tid & (kWave - 1), wid = tid / kWave;
	int x = count;   // inclusive scan over the wave
#pragma unroll
	for (int off = 1; off < kWave; off <<= 1) {
		const int y = __shfl_up(x, off, kWave);
		if (lane >= off) x += y;
	}
	if (lane == kWave - 1) waveOff[wid] = x;
	__syncthreads();
	if (tid == 0) {
		int t = 0;
		for (int w = 0; w < kBlock / kWave; w++) { const int v = waveOff[w]; waveOff[w] = t; t += v; }
		blockBase = t > 0 ? atomicAdd(cnt, t) : 0;
	}
	__syncthreads();
	const int r = blockBase + waveOff[wid] + x - count;
	__syncthreads();   // the LDS words are reused by the next call
	return r;
}

__device__ __forceinline__ void decode(const LayerSpace &S, int gid, int &a, int &i, int &j, int &k)
{
	a = 0;
	while (a + 1 < S.ng && gid >= S.off[a + 1]) a++;
	const int loc = gid - S.off[a];
	const GridDev &g = S.g[a];
	i = loc % g.n[0];
	const int rest = loc / g.n[0];
	j = rest % g.n[1];
	k = rest / g.n[1];
}

// Start of a front on grid `a`, in two passes: kFrontMark gives every element
// layer 0 if fixed and 255 otherwise (the fixing rule's loads once per
// element); kFrontEdges lists the fixed elements with an unfixed neighbour,
// reading only layer bytes.  Each kFrontEdges thread takes kInitRun elements
// along the slowest axis (3D: k = blockIdx.z * kInitRun + r; 2D:
// j = blockIdx.y * 4 * kInitRun + 4 r + threadIdx.y), so a workgroup appends
// up to 2048 elements with one atomic.
constexpr int kInitRun = 8;

template <int D>
static dim3 frontInitDims(const GridDev &g)
{
	if constexpr (D == 3) return dim3((g.n[0] + 63) / 64, (g.n[1] + 3) / 4, (g.n[2] + kInitRun - 1) / kInitRun);
	else return dim3((g.n[0] + 63) / 64, (g.n[1] + 4 * kInitRun - 1) / (4 * kInitRun), 1);
}

template <int D, int RULE>
__global__ __launch_bounds__(kBlock) void kFrontMark(LayerSpace S, int a, FixArgs P)
{
	const GridDev &g = S.g[a];
	const Coord c = threadCoord();
	if (!inGrid<D>(g, c)) return;
	S.layer[a][gidx<D>(g, c.i, c.j, c.k)] = isFixed<D, RULE>(P, a, g, c.i, c.j, c.k) ? 0 : kFar;
}

template <int D>
__global__ __launch_bounds__(kBlock) void kFrontEdges(LayerSpace S, int a, int *list, int *cnt)
{
	const GridDev &g = S.g[a];
	const uint8_t *lay = S.layer[a];
	const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
	unsigned edges = 0;
#pragma unroll
	for (int r = 0; r < kInitRun; r++) {
		const int j = D == 3 ? int(blockIdx.y * blockDim.y + threadIdx.y) : int(blockIdx.y * blockDim.y * kInitRun + r * blockDim.y + threadIdx.y);
		const int k = D == 3 ? int(blockIdx.z) * kInitRun + r : 0;
		if (!validIn<D>(g, i, j, k) || lay[gidx<D>(g, i, j, k)] != 0) continue;
		bool edge = false;
#pragma unroll
		for (int o = 0; o < 2 * D; o++) {
			const int ax = o >> 1, sd = (o & 1) ? 1 : -1;
			const int qi = i + (ax == 0 ? sd : 0), qj = j + (ax == 1 ? sd : 0), qk = k + (ax == 2 ? sd : 0);
			if (validIn<D>(g, qi, qj, qk) && lay[gidx<D>(g, qi, qj, qk)] == kFar) edge = true;
		}
		if (edge) edges |= 1u << r;
	}
	int at = blockAppendBase(cnt, __popc(edges));
#pragma unroll
	for (int r = 0; r < kInitRun; r++) {
		if (!(edges >> r & 1u)) continue;
		const int j = D == 3 ? int(blockIdx.y * blockDim.y + threadIdx.y) : int(blockIdx.y * blockDim.y * kInitRun + r * blockDim.y + threadIdx.y);
		const int k = D == 3 ? int(blockIdx.z) * kInitRun + r : 0;
		list[at++] = S.off[a] + int(gidx<D>(g, i, j, k));
	}
}

// Layer L of one front: one grid-stride element t of its layer L-1 list
// looks at each unreached neighbour q.  q is reached at layer L, and its
// value is the mean of its neighbours below layer L (Extrapolation.h:101-124,
// in the ordinal order, so the value is the reference's exactly); it is
// taken by its owner, the first of its neighbours in the ordinal order that
// is at layer L-1, which writes the value and the layer byte and lists q for
// layer L+1.  Every other element that sees q stops at that test.  The test
// and the mean read only bytes that are final (layers below L, written by
// earlier launches) or move from 255 to L during this launch, which both
// reads treat alike, so no atomics are needed and each q is listed once.
// Every thread of the workgroup calls this for the same workgroup-uniform t0,
// as blockAppendBase needs.
template <int D>
__device__ __forceinline__ void frontStepElements(const LayerSpace &S, int L, const int *listIn, int n, int t0, int *listOut, int *cntOut)
{
	const int t = t0 + int(threadIdx.x);
	int a = 0, i = 0, j = 0, k = 0;
	unsigned took = 0;
	if (t < n) {
		decode(S, listIn[t], a, i, j, k);
		const GridDev &g = S.g[a];
		uint8_t *lay = S.layer[a];
		for (int o = 0; o < 2 * D; o++) {
			const int ax = o >> 1, sd = (o & 1) ? 1 : -1;
			const int qi = i + (ax == 0 ? sd : 0), qj = j + (ax == 1 ? sd : 0), qk = k + (ax == 2 ? sd : 0);
			if (!validIn<D>(g, qi, qj, qk)) continue;
			const size_t q = gidx<D>(g, qi, qj, qk);
			if (lay[q] != kFar) continue;
			// this element is q's neighbour in direction o ^ 1 (seen from q)
			bool owner = true, seen = false;
			int cnt = 0;
			double s0 = 0, s1 = 0, s2 = 0;
			for (int p = 0; p < 2 * D; p++) {
				const int bx = p >> 1, bs = (p & 1) ? 1 : -1;
				const int ri = qi + (bx == 0 ? bs : 0), rj = qj + (bx == 1 ? bs : 0), rk = qk + (bx == 2 ? bs : 0);
				if (!validIn<D>(g, ri, rj, rk)) continue;
				const size_t rr = gidx<D>(g, ri, rj, rk);
				const uint8_t lr = lay[rr];
				if (!seen && lr == uint8_t(L - 1)) {
					seen = true;
					if (p != (o ^ 1)) { owner = false; break; }
				}
				if (lr < L) {
					if (S.nv > 0) s0 += S.v[a][0][rr];
					if (S.nv > 1) s1 += S.v[a][1][rr];
					if (S.nv > 2) s2 += S.v[a][2][rr];
					cnt++;
				}
			}
			if (!owner) continue;
			took |= 1u << o;
			if (S.nv > 0) S.v[a][0][q] = s0 / cnt;
			if (S.nv > 1) S.v[a][1][q] = s1 / cnt;
			if (S.nv > 2) S.v[a][2][q] = s2 / cnt;
			lay[q] = uint8_t(L);
		}
	}
	int at = blockAppendBase(cntOut, __popc(took));
	for (int o = 0; o < 2 * D; o++) {
		if (!(took >> o & 1u)) continue;
		const int ax = o >> 1, sd = (o & 1) ? 1 : -1;
		listOut[at++] = S.off[a] + int(gidx<D>(S.g[a], i + (ax == 0 ? sd : 0), j + (ax == 1 ? sd : 0), k + (ax == 2 ? sd : 0)));
	}
}

// Layer L: claim the unreached neighbours of the layer L-1 front.
template <int D>
__global__ __launch_bounds__(kBlock) void kFrontStep(LayerSpace S, int L, const int *listIn, const int *cntIn, int *listOut, int *cntOut, int *cntZero)
{
	if (blockIdx.x == 0 && threadIdx.x == 0) *cntZero = 0;
	const int n = *cntIn;
	for (int t0 = blockIdx.x * blockDim.x; t0 < n; t0 += gridDim.x * blockDim.x)
		frontStepElements<D>(S, L, listIn, n, t0, listOut, cntOut);
}

// Several independent fronts (their own layer arrays, lists and counters)
// advanced by one layer in one launch: each workgroup-sized chunk of the
// concatenated lists belongs to one front.  Fronts past their last layer
// take no work.
struct FrontJob {
	LayerSpace S;
	const int *listIn, *cntIn;
	int *listOut, *cntOut, *cntZero;
	int active;
};

constexpr int kMaxJobs = 3;
struct FrontJobs { FrontJob j[kMaxJobs]; int n; };

template <int D>
__global__ __launch_bounds__(kBlock) void kFrontStepJobs(FrontJobs J, int L)
{
	int chunks[kMaxJobs], n[kMaxJobs], total = 0;
#pragma unroll
	for (int q = 0; q < kMaxJobs; q++) {
		n[q] = q < J.n && J.j[q].active ? *J.j[q].cntIn : 0;
		chunks[q] = (n[q] + kBlock - 1) / kBlock;
		total += chunks[q];
	}
	if (blockIdx.x == 0 && threadIdx.x == 0)
		for (int q = 0; q < J.n; q++) if (J.j[q].active) *J.j[q].cntZero = 0;
	for (int c = blockIdx.x; c < total; c += gridDim.x) {
		int q = 0, first = 0;
		while (q + 1 < kMaxJobs && c >= first + chunks[q]) { first += chunks[q]; q++; }
		const FrontJob &F = J.j[q];
		frontStepElements<D>(F.S, L, F.listIn, n[q], (c - first) * kBlock, F.listOut, F.cntOut);
	}
}

__global__ void kZeroInts(int *p, int n)
{
	if (int(threadIdx.x) < n) p[threadIdx.x] = 0;
}

// Reference map not reached: the cell centre (Extrapolation.h:210-212).
template <int D>
__global__ __launch_bounds__(kBlock) void kFrontFinishRefMap(GridDev g, const uint8_t *layer, double *x0, double *x1, double *x2)
{
	const Coord c = threadCoord();
	if (!inGrid<D>(g, c)) return;
	const size_t id = gidx<D>(g, c.i, c.j, c.k);
	if (layer[id] != kFar) return;
	x0[id] = g.o[0] + c.i * g.dx;
	x1[id] = g.o[1] + c.j * g.dx;
	if constexpr (D == 3) x2[id] = g.o[2] + c.k * g.dx;
}

// _advRegion = reached (Extrapolation.h:130-175 on the phi_ori <= 0 indicator).
__global__ void kFrontToMask(const uint8_t *layer, uint8_t *adv, size_t n)
{
	for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
		adv[i] = layer[i] != kFar ? 1 : 0;
}

// Faces never reached are zeroed ("for velocity only", Extrapolation.h:76-78).
__global__ void kFrontFinishFaces(const uint8_t *layer, double *u, size_t n)
{
	for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
		if (layer[i] == kFar) u[i] = 0;
}

static unsigned linBlocks(size_t n)
{
	const size_t b = (n + kBlock - 1) / kBlock;
	return unsigned(b < 8192 ? (b ? b : 1) : 8192);
}

template <int D>
static void launchFrontInit(const LayerSpace &S, int a, const FixArgs &P, int *list, int *cnt, hipStream_t s)
{
	const dim3 gm = gridDims(S.g[a], D), bd = gridBlock();
	switch (P.rule) {
	case kFixCellDistPositive: hipLaunchKernelGGL((kFrontMark<D, kFixCellDistPositive>), gm, bd, 0, s, S, a, P); break;
	case kFixAdvRegion: hipLaunchKernelGGL((kFrontMark<D, kFixAdvRegion>), gm, bd, 0, s, S, a, P); break;
	case kFixPhiOriNonPositive: hipLaunchKernelGGL((kFrontMark<D, kFixPhiOriNonPositive>), gm, bd, 0, s, S, a, P); break;
	default: hipLaunchKernelGGL((kFrontMark<D, kFixLiquidFace>), gm, bd, 0, s, S, a, P); break;
	}
	hipLaunchKernelGGL(kFrontEdges<D>, frontInitDims<D>(S.g[a]), bd, 0, s, S, a, list, cnt);
}

// Run `steps` layers over S with fixed elements given by P.
static void runFront(const Fields &F, const LayerSpace &S, const FixArgs &P, int steps, hipStream_t s)
{
	int *cnt = F.front_cnt;
	hipLaunchKernelGGL(kZeroInts, dim3(1), dim3(64), 0, s, cnt, 3);
	for (int a = 0; a < S.ng; a++) {
		if (F.D == 2) launchFrontInit<2>(S, a, P, F.front_list[0], cnt, s);
		else launchFrontInit<3>(S, a, P, F.front_list[0], cnt, s);
	}
	for (int L = 1; L <= steps; L++) {
		const int *in = F.front_list[(L - 1) & 1];
		int *out = F.front_list[L & 1];
		const int *cin = cnt + (L - 1) % 3;
		int *cout = cnt + L % 3, *cz = cnt + (L + 1) % 3;
		if (F.D == 2) hipLaunchKernelGGL(kFrontStep<2>, dim3(kFrontGrid), dim3(kBlock), 0, s, S, L, in, cin, out, cout, cz);
		else hipLaunchKernelGGL(kFrontStep<3>, dim3(kFrontGrid), dim3(kBlock), 0, s, S, L, in, cin, out, cout, cz);
	}
}

static LayerSpace cellSpace(const Fields &F, int nv, double *v0, double *v1, double *v2)
{
	LayerSpace S{};
	S.ng = 1;
	S.g[0] = F.cell;
	S.layer[0] = F.layer;
	S.v[0][0] = v0; S.v[0][1] = v1; S.v[0][2] = v2;
	S.nv = nv;
	S.off[0] = 0;
	S.off[1] = int(F.N);
	return S;
}

static FixArgs fixArgs(const Fields &F, int rule)
{
	FixArgs P{};
	P.rule = rule;
	P.cell = F.cell;
	P.cellDist = F.cell_dist;
	P.adv = F.adv_region;
	P.phiOri = F.phi_ori;
	for (int a = 0; a < F.D; a++) P.open[a] = F.open_f[a];
	P.liquid = nullptr;
	return P;
}

// extrapolateFieldsToCollider (LevelSetLiquid.cpp:1086-1137): phi_ori and phi,
// 6 layers (_boundaryExtrapolationRadius, LevelSetLiquid.h:92), fixed where
// cellDist > 0; both share the fixed set, so one front carries both.  The
// fixed set depends on cellDist alone, which steps do not change: the layers
// are found once by the claiming fronts (buildColliderLayers, kept as one
// list per layer) and each step only averages, layer by layer.
template <int D>
__global__ __launch_bounds__(kBlock) void kColliderLayer(GridDev g, const uint8_t *layer, const int *list, int n, int L, double *v0, double *v1)
{
	const int t = int(blockIdx.x * blockDim.x + threadIdx.x);
	if (t >= n) return;
	const int id = list[t];
	const int qi = id % g.n[0], qj = (id / g.n[0]) % g.n[1], qk = D == 3 ? id / (g.n[0] * g.n[1]) : 0;
	int cnt = 0;
	double s0 = 0, s1 = 0;
	for (int p = 0; p < 2 * D; p++) {
		const int bx = p >> 1, bs = (p & 1) ? 1 : -1;
		const int ri = qi + (bx == 0 ? bs : 0), rj = qj + (bx == 1 ? bs : 0), rk = qk + (bx == 2 ? bs : 0);
		if (!validIn<D>(g, ri, rj, rk)) continue;
		const size_t rr = gidx<D>(g, ri, rj, rk);
		if (layer[rr] < L) {
			s0 += v0[rr];
			s1 += v1[rr];
			cnt++;
		}
	}
	v0[id] = s0 / cnt;
	v1[id] = s1 / cnt;
}

bool buildColliderLayers(Fields &F, hipStream_t s)
{
	LayerSpace S = cellSpace(F, 0, nullptr, nullptr, nullptr);
	S.layer[0] = F.coll_layer;
	int *cnt = F.front_cnt + 10;   // [0] the fixed boundary, [L] layer L, [7] scratch
	hipLaunchKernelGGL(kZeroInts, dim3(1), dim3(64), 0, s, cnt, 8);
	if (F.D == 2) launchFrontInit<2>(S, 0, fixArgs(F, kFixCellDistPositive), F.front_list[0], cnt, s);
	else launchFrontInit<3>(S, 0, fixArgs(F, kFixCellDistPositive), F.front_list[0], cnt, s);
	F.coll_off[1] = 0;
	for (int L = 1; L <= kColliderLayers; L++) {
		const int *in = L == 1 ? F.front_list[0] : F.coll_list + F.coll_off[L - 1];
		if (F.D == 2) hipLaunchKernelGGL(kFrontStep<2>, dim3(kFrontGrid), dim3(kBlock), 0, s, S, L, in, (const int *)cnt + L - 1, F.coll_list + F.coll_off[L], cnt + L, cnt + 7);
		else hipLaunchKernelGGL(kFrontStep<3>, dim3(kFrontGrid), dim3(kBlock), 0, s, S, L, in, (const int *)cnt + L - 1, F.coll_list + F.coll_off[L], cnt + L, cnt + 7);
		int n = 0;
		if (hipMemcpyAsync(&n, cnt + L, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess) return false;
		F.coll_cnt[L] = n;
		F.coll_off[L + 1] = F.coll_off[L] + n;
	}
	return hipGetLastError() == hipSuccess;
}

void launchExtrapolateCollider(const Fields &F, hipStream_t s)
{
	for (int L = 1; L <= kColliderLayers; L++) {
		const int n = F.coll_cnt[L];
		if (n == 0) break;
		const dim3 gd((n + kBlock - 1) / kBlock);
		if (F.D == 2) hipLaunchKernelGGL(kColliderLayer<2>, gd, dim3(kBlock), 0, s, F.cell, (const uint8_t *)F.coll_layer, (const int *)(F.coll_list + F.coll_off[L]), n, L, F.phi_ori, F.phi);
		else hipLaunchKernelGGL(kColliderLayer<3>, gd, dim3(kBlock), 0, s, F.cell, (const uint8_t *)F.coll_layer, (const int *)(F.coll_list + F.coll_off[L]), n, L, F.phi_ori, F.phi);
	}
}

// extrapolateFieldsToAir (LevelSetLiquid.cpp:1059-1084): xi from the old
// advRegion (15 layers, unreached reset to the cell centre), advRegion
// rebuilt from phi_ori <= 0 (15 layers), u on every face grid (17 layers,
// unreached zeroed).  The three fronts are independent (their own layer
// arrays and lists), so each layer of all three is one launch.
void launchExtrapolateAir(const Fields &F, hipStream_t s, bool velocityOnly)
{
	LayerSpace Sx = cellSpace(F, F.D, F.xi[0], F.xi[1], F.D == 3 ? F.xi[2] : nullptr);
	LayerSpace Sa = cellSpace(F, 0, nullptr, nullptr, nullptr);
	Sa.layer[0] = F.layer2;
	LayerSpace Sf{};
	Sf.ng = F.D;
	Sf.nv = 1;
	int off = 0;
	for (int a = 0; a < F.D; a++) {
		Sf.g[a] = F.face[a];
		Sf.layer[a] = F.layer_f[a];
		Sf.v[a][0] = F.u[a];
		Sf.off[a] = off;
		off += int(F.Nf[a]);
	}
	Sf.off[F.D] = off;
	// the velocity-only form (extrapolateFieldsToAir(true)) runs the face front alone
	const LayerSpace *all[3] = {&Sx, &Sa, &Sf};
	const int allRules[3] = {kFixAdvRegion, kFixPhiOriNonPositive, kFixLiquidFace};
	const int allSteps[3] = {15, 15, 17};
	const int nq = velocityOnly ? 1 : 3, q0 = 3 - nq;
	const LayerSpace *spaces[3] = {all[q0], nq > 1 ? all[q0 + 1] : nullptr, nq > 2 ? all[q0 + 2] : nullptr};
	const int rules[3] = {allRules[q0], nq > 1 ? allRules[q0 + 1] : 0, nq > 2 ? allRules[q0 + 2] : 0};
	const int steps[3] = {allSteps[q0], nq > 1 ? allSteps[q0 + 1] : 0, nq > 2 ? allSteps[q0 + 2] : 0};
	int *cnt = F.front_cnt;   // three counters per front
	hipLaunchKernelGGL(kZeroInts, dim3(1), dim3(64), 0, s, cnt, 9);
	for (int q = 0; q < nq; q++)
		for (int a = 0; a < spaces[q]->ng; a++) {
			const int f = q0 + q;   // the front's own lists and counters
			FixArgs P = fixArgs(F, rules[q]);
			// the face front's rule reads the advRegion front's marks (marked
			// just before, 0 exactly where phi_ori <= 0) instead of phi_ori
			if (rules[q] == kFixLiquidFace && !velocityOnly) P.liquid = F.layer2;
			if (F.D == 2) launchFrontInit<2>(*spaces[q], a, P, F.front_list[2 * f], cnt + 3 * f, s);
			else launchFrontInit<3>(*spaces[q], a, P, F.front_list[2 * f], cnt + 3 * f, s);
		}
	for (int L = 1; L <= 17; L++) {
		FrontJobs J{};
		J.n = nq;
		for (int q = 0; q < nq; q++) {
			FrontJob &j = J.j[q];
			const int f = q0 + q;
			j.S = *spaces[q];
			j.listIn = F.front_list[2 * f + ((L - 1) & 1)];
			j.listOut = F.front_list[2 * f + (L & 1)];
			j.cntIn = cnt + 3 * f + (L - 1) % 3;
			j.cntOut = cnt + 3 * f + L % 3;
			j.cntZero = cnt + 3 * f + (L + 1) % 3;
			j.active = L <= steps[q] ? 1 : 0;
		}
		if (F.D == 2) hipLaunchKernelGGL(kFrontStepJobs<2>, dim3(kFrontGrid), dim3(kBlock), 0, s, J, L);
		else hipLaunchKernelGGL(kFrontStepJobs<3>, dim3(kFrontGrid), dim3(kBlock), 0, s, J, L);
	}
	if (!velocityOnly) {
	if (F.D == 2) hipLaunchKernelGGL(kFrontFinishRefMap<2>, gridDims(F.cell, 2), gridBlock(), 0, s, F.cell, (const uint8_t *)F.layer, F.xi[0], F.xi[1], F.xi[2]);
	else hipLaunchKernelGGL(kFrontFinishRefMap<3>, gridDims(F.cell, 3), gridBlock(), 0, s, F.cell, (const uint8_t *)F.layer, F.xi[0], F.xi[1], F.xi[2]);
	hipLaunchKernelGGL(kFrontToMask, dim3(linBlocks(F.N)), dim3(kBlock), 0, s, (const uint8_t *)F.layer2, F.adv_region, F.N);
	}
	for (int a = 0; a < F.D; a++)
		hipLaunchKernelGGL(kFrontFinishFaces, dim3(linBlocks(F.Nf[a])), dim3(kBlock), 0, s, (const uint8_t *)F.layer_f[a], F.u[a], F.Nf[a]);
}

} // namespace garm
