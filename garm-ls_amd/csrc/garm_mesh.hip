// Surface mesh of a cell field for gfx950: Contourer<3>::contour at 0
// (Contourer.cpp:9-97) with SurfaceMesh::computeNormals (SurfaceMesh.cpp:
// 191-216), FP64 with the reference's operations (compiled without
// multiply-add contraction): positions in FP64 (SurfaceMesh::writeOBJ
// prints them, SurfaceMesh.cpp:126-144; SurfaceMesh::write rounds them to
// float, :103-124), normals rounded to float on the device.
//
// The field's grid points are the marching cubes' nodes; a cube is a cell of
// that node grid.  Numbering follows the reference exactly: node ord k of a
// cube is (k&1, k>>1&1, k>>2&1), edge ord i runs along axis i>>2 with offsets
// (i&1, i>>1&1) along the next two axes (StaggeredGrid.h:61-94), a cube's
// type has bit k set where the field is <= 0 (getCellType), and a vertex is
// created, in cube order and then edge order, by the first cube that holds
// its edge -- so vertex v is the (owner cube, edge) pair at rank v, which
// the GPU finds with a scan of per-cube counts.  The triangles of a cube come
// from a table built here (not the reference's table): on each cube face the
// crossing edges are paired, an ambiguous face (two diagonal inside corners)
// cutting off each inside corner, the pairs chain into closed polygons, each
// oriented so its normal runs from the inside corners to the outside ones,
// started at its smallest edge ord and fanned from there; polygons in the
// order of their smallest edge.  A shared face is paired the same way from
// both cubes, so the mesh of a closed surface is watertight.
//
// Passes (one workgroup per tile of kScanTile cubes): per-tile vertex and
// triangle counts; two tile scans; vertices (positions, each cube's first
// vertex id); triangles (three vertex ids each, found from the owner cube of
// each edge); vertex normals gathered per vertex from the up to four cubes
// around its edge in cube order, triangle order within a cube.
#include "garm_internal.hpp"
#include "garm_project.hpp"

#include <algorithm>
#include <climits>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace garm {

constexpr int kMcMaxTri = 8;   // checked when the table is built

struct McTable {
	int8_t tri[256][kMcMaxTri * 3];   // edge ords, three per triangle
	uint8_t ntri[256];
	uint16_t cross[256];              // crossing edges
};

__constant__ McTable c_mc;

// corner ord of end `end` (0: the edge's origin) of edge ord i
static int edgeCorner(int i, int end)
{
	const int a = i >> 2, b = (a + 1) % 3, c = (a + 2) % 3;
	int p[3] = {0, 0, 0};
	p[b] = i & 1;
	p[c] = i >> 1 & 1;
	p[a] = end;
	return p[0] + 2 * p[1] + 4 * p[2];
}

static bool buildMcTable(McTable &T, std::string &err)
{
	std::memset(&T, -1, sizeof(T.tri));
	for (int type = 0; type < 256; type++) {
		auto inside = [&](int k) { return (type >> k & 1) != 0; };
		uint16_t cross = 0;
		for (int i = 0; i < 12; i++)
			if (inside(edgeCorner(i, 0)) != inside(edgeCorner(i, 1))) cross |= uint16_t(1u << i);
		T.cross[type] = cross;
		// pairings on the six faces
		int nbr[12][2], deg[12] = {};
		for (int f = 0; f < 3; f++)
			for (int s = 0; s < 2; s++) {
				int fe[4], ne = 0;
				for (int i = 0; i < 12; i++) {
					const int a = i >> 2;
					if (a == f || !(cross >> i & 1)) continue;
					const int b = (a + 1) % 3;
					const int coord = f == b ? (i & 1) : (i >> 1 & 1);
					if (coord == s) fe[ne++] = i;
				}
				auto link = [&](int x, int y) { nbr[x][deg[x]++] = y; nbr[y][deg[y]++] = x; };
				if (ne == 2) link(fe[0], fe[1]);
				else if (ne == 4) {
					// each inside corner of the face is cut off by its two edges
					for (int k = 0; k < 8; k++) {
						const int pk[3] = {k & 1, k >> 1 & 1, k >> 2 & 1};
						if (pk[f] != s || !inside(k)) continue;
						int two[2], m = 0;
						for (int q = 0; q < 4; q++)
							if (edgeCorner(fe[q], 0) == k || edgeCorner(fe[q], 1) == k) two[m++] = fe[q];
						if (m != 2) { err = "marching-cubes table: bad ambiguous face"; return false; }
						link(two[0], two[1]);
					}
				} else if (ne != 0) { err = "marching-cubes table: odd crossings on a face"; return false; }
			}
		int nt = 0;
		bool seen[12] = {};
		for (int start = 0; start < 12; start++) {
			if (!(cross >> start & 1) || seen[start]) continue;
			if (deg[start] != 2) { err = "marching-cubes table: open polygon"; return false; }
			int cyc[12], len = 0, prev = -1, cur = start;
			do {
				seen[cur] = true;
				cyc[len++] = cur;
				const int next = nbr[cur][0] != prev ? nbr[cur][0] : nbr[cur][1];
				prev = cur;
				cur = next;
			} while (cur != start && len < 12);
			if (cur != start) { err = "marching-cubes table: polygon does not close"; return false; }
			// orientation: area vector of the edge midpoints against the
			// inside-to-outside directions of its edges
			double N[3] = {0, 0, 0}, R[3] = {0, 0, 0};
			for (int q = 0; q < len; q++) {
				double m0[3], m1[3];
				for (int w = 0; w < 2; w++) {
					const int e = cyc[(q + w) % len];
					const int c0 = edgeCorner(e, 0), c1 = edgeCorner(e, 1);
					double *m = w ? m1 : m0;
					for (int d = 0; d < 3; d++) m[d] = 0.5 * ((c0 >> d & 1) + (c1 >> d & 1));
				}
				N[0] += m0[1] * m1[2] - m0[2] * m1[1];
				N[1] += m0[2] * m1[0] - m0[0] * m1[2];
				N[2] += m0[0] * m1[1] - m0[1] * m1[0];
				const int e = cyc[q], in = inside(edgeCorner(e, 0)) ? 0 : 1;
				R[e >> 2] += in == 0 ? 1.0 : -1.0;
			}
			const double dot = N[0] * R[0] + N[1] * R[1] + N[2] * R[2];
			if (dot == 0) { err = "marching-cubes table: unoriented polygon"; return false; }
			if (dot < 0) std::reverse(cyc, cyc + len);
			// canonical start: the polygon's smallest edge ord (polygons are
			// found in the order of their smallest edge)
			std::rotate(cyc, std::min_element(cyc, cyc + len), cyc + len);
			for (int q = 1; q + 1 < len; q++) {
				if (nt == kMcMaxTri) { err = "marching-cubes table: too many triangles"; return false; }
				T.tri[type][3 * nt] = int8_t(cyc[0]);
				T.tri[type][3 * nt + 1] = int8_t(cyc[q]);
				T.tri[type][3 * nt + 2] = int8_t(cyc[q + 1]);
				nt++;
			}
		}
		T.ntri[type] = uint8_t(nt);
	}
	return true;
}

// ------------------------------------------------------------- device side
struct McGrid {
	GridDev g;      // the field's grid (nodes)
	int m[3];       // cubes per axis
	size_t M;       // cubes
};

__device__ __forceinline__ void cubeCoord(const McGrid &G, size_t id, int *c)
{
	c[0] = int(id % size_t(G.m[0]));
	c[1] = int((id / size_t(G.m[0])) % size_t(G.m[1]));
	c[2] = int(id / (size_t(G.m[0]) * G.m[1]));
}

__device__ __forceinline__ size_t cubeId(const McGrid &G, const int *c)
{
	return size_t(c[0]) + size_t(G.m[0]) * (size_t(c[1]) + size_t(G.m[1]) * size_t(c[2]));
}

__device__ __forceinline__ int cubeType(const McGrid &G, const double *f, const int *c)
{
	int t = 0;
#pragma unroll
	for (int k = 0; k < 8; k++)
		if (f[gidx<3>(G.g, c[0] + (k & 1), c[1] + (k >> 1 & 1), c[2] + (k >> 2 & 1))] <= 0) t |= 1 << k;
	return t;
}

// edges this cube creates the vertex of: the first cube holding an edge is
// the one shifted back along every axis where that is possible
__device__ __forceinline__ unsigned ownedEdges(const int *c)
{
	unsigned m = 0;
#pragma unroll
	for (int i = 0; i < 12; i++) {
		const int a = i >> 2, b = (a + 1) % 3, d = (a + 2) % 3;
		if (((i & 1) || c[b] == 0) && ((i >> 1 & 1) || c[d] == 0)) m |= 1u << i;
	}
	return m;
}

// vertex position on edge ord i of cube c (Contourer.cpp:40-45)
__device__ __forceinline__ void edgeVertex(const McGrid &G, const double *f, const int *c, int i, double *p)
{
	const int a = i >> 2, b = (a + 1) % 3, d = (a + 2) % 3;
	int n0[3] = {c[0], c[1], c[2]};
	n0[b] += i & 1;
	n0[d] += i >> 1 & 1;
	int n1[3] = {n0[0], n0[1], n0[2]};
	n1[a] += 1;
	const double f0 = f[gidx<3>(G.g, n0[0], n0[1], n0[2])], f1 = f[gidx<3>(G.g, n1[0], n1[1], n1[2])];
	const double theta = (f0 - 0) / (f0 - f1);
#pragma unroll
	for (int q = 0; q < 3; q++) {
		const double p0 = G.g.o[q] + n0[q] * G.g.dx, p1 = G.g.o[q] + n1[q] * G.g.dx;
		p[q] = (1 - theta) * p0 + theta * p1;
	}
}

__global__ __launch_bounds__(kBlock) void kMeshTileCounts(McGrid G, const double *f, int *tileV, int *tileT)
{
	const size_t base = size_t(blockIdx.x) * kScanTile;
	int nv = 0, nt = 0;
	for (int sub = 0; sub < kScanSub; sub++) {
		const size_t id = base + size_t(sub) * kBlock + threadIdx.x;
		if (id >= G.M) break;
		int c[3];
		cubeCoord(G, id, c);
		const int t = cubeType(G, f, c);
		nv += __popc(unsigned(c_mc.cross[t]) & ownedEdges(c));
		nt += c_mc.ntri[t];
	}
	int tv, tt;
	(void)blockExclusiveScan(nv, &tv);
	(void)blockExclusiveScan(nt, &tt);
	if (threadIdx.x == 0) { tileV[blockIdx.x] = tv; tileT[blockIdx.x] = tt; }
}

__global__ __launch_bounds__(kBlock) void kMeshVertices(McGrid G, const double *f, const int *tileVBase, int *vbase, double *pos)
{
	const size_t base = size_t(blockIdx.x) * kScanTile;
	int carry = tileVBase[blockIdx.x];
	for (int sub = 0; sub < kScanSub; sub++) {
		const size_t id = base + size_t(sub) * kBlock + threadIdx.x;
		int c[3] = {0, 0, 0};
		unsigned mine = 0;
		if (id < G.M) {
			cubeCoord(G, id, c);
			mine = unsigned(c_mc.cross[cubeType(G, f, c)]) & ownedEdges(c);
		}
		int total;
		int v = carry + blockExclusiveScan(__popc(mine), &total);
		if (id < G.M) {
			vbase[id] = v;
			for (int i = 0; i < 12; i++) {
				if (!(mine >> i & 1)) continue;
				double p[3];
				edgeVertex(G, f, c, i, p);
				for (int q = 0; q < 3; q++) pos[size_t(v) * 3 + q] = p[q];
				v++;
			}
		}
		carry += total;
	}
}

// the vertex id of edge ord i of cube c: rank of the edge in its owner cube
__device__ __forceinline__ int vertexOf(const McGrid &G, const double *f, const int *vbase, const int *c, int i)
{
	const int a = i >> 2, b = (a + 1) % 3, d = (a + 2) % 3;
	int o[3] = {c[0], c[1], c[2]};
	int bb = i & 1, bd = i >> 1 & 1;
	if (!bb && c[b] > 0) { o[b] -= 1; bb = 1; }
	if (!bd && c[d] > 0) { o[d] -= 1; bd = 1; }
	const int j = a << 2 | bb | bd << 1;
	const unsigned mine = unsigned(c_mc.cross[cubeType(G, f, o)]) & ownedEdges(o);
	return vbase[cubeId(G, o)] + __popc(mine & ((1u << j) - 1));
}

__global__ __launch_bounds__(kBlock) void kMeshTriangles(McGrid G, const double *f, const int *tileTBase, const int *vbase, unsigned *idx)
{
	const size_t base = size_t(blockIdx.x) * kScanTile;
	int carry = tileTBase[blockIdx.x];
	for (int sub = 0; sub < kScanSub; sub++) {
		const size_t id = base + size_t(sub) * kBlock + threadIdx.x;
		int c[3] = {0, 0, 0};
		int t = 0, nt = 0;
		if (id < G.M) {
			cubeCoord(G, id, c);
			t = cubeType(G, f, c);
			nt = c_mc.ntri[t];
		}
		int total;
		const int first = carry + blockExclusiveScan(nt, &total);
		for (int q = 0; q < nt * 3; q++) idx[size_t(first) * 3 + q] = unsigned(vertexOf(G, f, vbase, c, c_mc.tri[t][q]));
		carry += total;
	}
}

// Eigen's normalized() / normalize(): unchanged when the norm is 0
__device__ __forceinline__ void unit3(double *v)
{
	const double s = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
	if (s > 0) {
		const double n = sqrt(s);
		v[0] /= n; v[1] /= n; v[2] /= n;
	}
}

// computeNormals: per vertex the sum of the unit normals (p1 - p0) x (p2 - p0)
// of the triangles using it, in triangle order, normalised
__global__ __launch_bounds__(kBlock) void kMeshNormals(McGrid G, const double *f, const int *vbase, float *nrm)
{
	const size_t id = size_t(blockIdx.x) * kBlock + threadIdx.x;
	if (id >= G.M) return;
	int c[3];
	cubeCoord(G, id, c);
	const unsigned mine = unsigned(c_mc.cross[cubeType(G, f, c)]) & ownedEdges(c);
	if (!mine) return;
	int v = vbase[id];
	for (int i = 0; i < 12; i++) {
		if (!(mine >> i & 1)) continue;
		const int a = i >> 2, b = (a + 1) % 3, d = (a + 2) % 3;
		int e[3] = {c[0], c[1], c[2]};   // the edge's origin node
		e[b] += i & 1;
		e[d] += i >> 1 & 1;
		// the cubes holding the edge, in cube order
		int cand[4][3], nc = 0;
		size_t key[4];
		for (int s1 = 0; s1 < 2; s1++)
			for (int s2 = 0; s2 < 2; s2++) {
				int q[3] = {e[0], e[1], e[2]};
				q[b] -= s1;
				q[d] -= s2;
				if (q[b] < 0 || q[d] < 0 || q[b] >= G.m[b] || q[d] >= G.m[d]) continue;
				cand[nc][0] = q[0]; cand[nc][1] = q[1]; cand[nc][2] = q[2];
				key[nc] = cubeId(G, q);
				nc++;
			}
		for (int x = 1; x < nc; x++)
			for (int y = x; y > 0 && key[y] < key[y - 1]; y--) {
				const size_t tk = key[y]; key[y] = key[y - 1]; key[y - 1] = tk;
				for (int z = 0; z < 3; z++) { const int tc = cand[y][z]; cand[y][z] = cand[y - 1][z]; cand[y - 1][z] = tc; }
			}
		double sum[3] = {0, 0, 0};
		for (int k = 0; k < nc; k++) {
			const int *q = cand[k];
			const int j = a << 2 | (e[b] - q[b]) | (e[d] - q[d]) << 1;   // the edge's ord in cube q
			const int t = cubeType(G, f, q);
			for (int tr = 0; tr < c_mc.ntri[t]; tr++) {
				const int8_t *te = c_mc.tri[t] + 3 * tr;
				if (te[0] != j && te[1] != j && te[2] != j) continue;
				double p0[3], p1[3], p2[3];
				edgeVertex(G, f, q, te[0], p0);
				edgeVertex(G, f, q, te[1], p1);
				edgeVertex(G, f, q, te[2], p2);
				double u[3], w[3];
				for (int z = 0; z < 3; z++) { u[z] = p1[z] - p0[z]; w[z] = p2[z] - p0[z]; }
				double n[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
				unit3(n);
				for (int z = 0; z < 3; z++) sum[z] += n[z];
			}
		}
		unit3(sum);
		for (int z = 0; z < 3; z++) nrm[size_t(v) * 3 + z] = float(sum[z]);
		v++;
	}
}

// ---------------------------------------------------------------- driver
static std::mutex g_table_mutex;
static unsigned long long g_table_set = 0;   // devices whose constant table is set, as a bit set

static int ensureTable(std::string &err)
{
	int dev = 0;
	if (hipGetDevice(&dev) != hipSuccess) { err = "hipGetDevice failed"; return GARM_EHIP; }
	std::lock_guard<std::mutex> lock(g_table_mutex);
	if (dev < 64 && (g_table_set >> dev & 1)) return GARM_OK;
	static McTable T;
	static bool built = false;
	if (!built) {
		if (!buildMcTable(T, err)) return GARM_EINVAL;
		built = true;
	}
	if (hipMemcpyToSymbol(HIP_SYMBOL(c_mc), &T, sizeof(T)) != hipSuccess) { err = "uploading the marching-cubes table failed"; return GARM_EHIP; }
	if (dev < 64) g_table_set |= 1ull << dev;
	return GARM_OK;
}

// ------------------------------------------------------------- 2D, host
// Contourer<2, true>::contour at 0 (Contourer.cpp:9-97), the filled contour
// the reference writes for 2D frames: vertices are the inside nodes (field
// <= 0) in node order, then the edge crossings in cell order and edge order
// (first cell to meet a crossing edge creates it); cell types 6 and 9 (two
// diagonal inside nodes) become 16 and 17 when the mean of the four nodes
// is <= 0 (getCellType).  Triangles come from a table built here: walking a
// cell's boundary counter-clockwise from node 0 lists its inside nodes and
// crossings in order, which is the inside polygon (convex: a square less
// corner triangles); a separated diagonal type gives each inside node its
// own triangle with the crossings beside it, every other type one polygon
// fanned from the first element of the walk, so triangles are
// counter-clockwise.  Table entries: < 4 an edge ord, 4 + k node ord k (the
// reference's encoding, Contourer.cpp:59-71).  Mesh export is off the step,
// so this runs on the host on a copy of the field.
struct FilledTable2 {
	int8_t e[18][12];
	uint8_t n[18];   // entries (three per triangle)
};

static void buildFilledTable2(FilledTable2 &T)
{
	const int ring[4] = {0, 1, 3, 2};       // nodes counter-clockwise
	const int side[4] = {0, 3, 1, 2};       // the edge from ring[q] to ring[q + 1]
	for (int t = 0; t < 18; t++) {
		const int base = t < 16 ? t : (t == 16 ? 6 : 9);
		auto inside = [&](int k) { return (base >> k & 1) != 0; };
		int seq[8], m = 0;
		for (int q = 0; q < 4; q++) {
			const int c = ring[q], nx = ring[(q + 1) % 4];
			if (inside(c)) seq[m++] = 4 + c;
			if (inside(c) != inside(nx)) seq[m++] = side[q];
		}
		int k = 0;
		if ((base == 6 || base == 9) && t < 16) {
			for (int p = 0; p < m; p++)
				if (seq[p] >= 4) {
					T.e[t][k++] = int8_t(seq[(p + m - 1) % m]);
					T.e[t][k++] = int8_t(seq[p]);
					T.e[t][k++] = int8_t(seq[(p + 1) % m]);
				}
		} else {
			for (int q = 1; q + 1 < m; q++) {
				T.e[t][k++] = int8_t(seq[0]);
				T.e[t][k++] = int8_t(seq[q]);
				T.e[t][k++] = int8_t(seq[q + 1]);
			}
		}
		T.n[t] = uint8_t(k);
	}
}

static int contourFilled2(const Fields &F, const double *field, std::vector<double> &pos, std::vector<unsigned> &idx, hipStream_t s,
	std::string &err)
{
	static FilledTable2 T;
	static std::once_flag once;
	std::call_once(once, [] { buildFilledTable2(T); });
	const int nx = F.cell.n[0], ny = F.cell.n[1];
	std::vector<double> f(size_t(nx) * ny);
	if (hipMemcpyAsync(f.data(), field, f.size() * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
		hipStreamSynchronize(s) != hipSuccess) {
		err = "reading the field failed";
		return GARM_EHIP;
	}
	pos.clear();
	idx.clear();
	const double *o = F.cell.o, dx = F.cell.dx;
	auto at = [&](int i, int j) { return f[size_t(j) * nx + i]; };
	std::vector<int> nodeMark(f.size(), -1), mark[2];
	mark[0].assign(size_t(nx - 1 > 0 ? nx - 1 : 0) * ny, -1);   // x edges: (nx - 1) x ny
	mark[1].assign(size_t(nx) * (ny - 1 > 0 ? ny - 1 : 0), -1); // y edges: nx x (ny - 1)
	for (int j = 0; j < ny; j++)
		for (int i = 0; i < nx; i++)
			if (at(i, j) <= 0) {
				nodeMark[size_t(j) * nx + i] = int(pos.size() / 2);
				pos.push_back(o[0] + i * dx);
				pos.push_back(o[1] + j * dx);
			}
	for (int j = 0; j + 1 < ny; j++)
		for (int i = 0; i + 1 < nx; i++) {
			const double v[4] = {at(i, j), at(i + 1, j), at(i, j + 1), at(i + 1, j + 1)};
			int type = 0;
			for (int k = 0; k < 4; k++)
				if (v[k] <= 0) type |= 1 << k;
			if (type == 6 || type == 9) {
				const double centre = (v[0] + v[1] + v[2] + v[3]) / 4;
				if (centre <= 0) type = 16 + (type > 8);
			}
			int eid[4];
			for (int e = 0; e < 4; e++) {
				const int axis = e >> 1;
				const int ei = i + (axis == 1 ? (e & 1) : 0), ej = j + (axis == 0 ? (e & 1) : 0);
				const int fi = ei + (axis == 0), fj = ej + (axis == 1);
				const double f0 = at(ei, ej), f1 = at(fi, fj);
				int &mk = mark[axis][axis == 0 ? size_t(ej) * (nx - 1) + ei : size_t(ej) * nx + ei];
				if ((f0 <= 0) != (f1 <= 0) && mk < 0) {
					const double theta = (f0 - 0) / (f0 - f1);
					mk = int(pos.size() / 2);
					pos.push_back((1 - theta) * (o[0] + ei * dx) + theta * (o[0] + fi * dx));
					pos.push_back((1 - theta) * (o[1] + ej * dx) + theta * (o[1] + fj * dx));
				}
				eid[e] = mk;
			}
			for (int q = 0; q < T.n[type]; q++) {
				const int c = T.e[type][q];
				if (c < 4) idx.push_back(unsigned(eid[c]));
				else {
					const int k = c & 3;
					idx.push_back(unsigned(nodeMark[size_t(j + (k >> 1)) * nx + i + (k & 1)]));
				}
			}
		}
	return GARM_OK;
}

int contourMesh(const Fields &F, const double *field, std::vector<double> &pos, std::vector<float> &nrm, std::vector<unsigned> &idx,
	hipStream_t s, std::string &err)
{
	if (F.D == 2) {
		nrm.clear();
		return contourFilled2(F, field, pos, idx, s, err);
	}
	int rc = ensureTable(err);
	if (rc) return rc;
	McGrid G{};
	G.g = F.cell;
	for (int a = 0; a < 3; a++) G.m[a] = F.cell.n[a] - 1;
	if (G.m[0] < 1 || G.m[1] < 1 || G.m[2] < 1) { pos.clear(); nrm.clear(); idx.clear(); return GARM_OK; }
	G.M = size_t(G.m[0]) * G.m[1] * G.m[2];
	const int tiles = int((G.M + kScanTile - 1) / kScanTile);
	int *ibuf = nullptr;
	const size_t nI = size_t(tiles) * 4 + 4 + G.M;
	if (hipMalloc(&ibuf, nI * sizeof(int)) != hipSuccess) { err = "hipMalloc failed for the mesh counts"; return GARM_ENOMEM; }
	int *tileV = ibuf, *tileT = ibuf + tiles, *baseV = ibuf + 2 * tiles, *baseT = ibuf + 3 * tiles + 1, *vbase = ibuf + 4 * tiles + 4;
	double *dpos = nullptr;
	float *dnrm = nullptr;
	unsigned *didx = nullptr;
	int counts[2] = {0, 0};
	auto done = [&](int code) {
		(void)hipFree(ibuf);
		if (dpos) (void)hipFree(dpos);
		if (dnrm) (void)hipFree(dnrm);
		if (didx) (void)hipFree(didx);
		return code;
	};
	hipLaunchKernelGGL(kMeshTileCounts, dim3(tiles), dim3(kBlock), 0, s, G, field, tileV, tileT);
	launchScanTiles(tileV, tiles, baseV, baseV + tiles, s);
	launchScanTiles(tileT, tiles, baseT, baseT + tiles, s);
	if (hipMemcpyAsync(&counts[0], baseV + tiles, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
		hipMemcpyAsync(&counts[1], baseT + tiles, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
		hipStreamSynchronize(s) != hipSuccess) {
		err = "reading the mesh counts failed";
		return done(GARM_EHIP);
	}
	const size_t nv = size_t(counts[0]), nt = size_t(counts[1]);
	pos.assign(nv * 3, 0.0);
	nrm.assign(nv * 3, 0.f);
	idx.assign(nt * 3, 0u);
	if (nv == 0) return done(GARM_OK);
	if (hipMalloc(&dpos, nv * 3 * sizeof(double)) != hipSuccess || hipMalloc(&dnrm, nv * 3 * sizeof(float)) != hipSuccess ||
		(nt && hipMalloc(&didx, nt * 3 * sizeof(unsigned)) != hipSuccess)) {
		err = "hipMalloc failed for the mesh";
		return done(GARM_ENOMEM);
	}
	hipLaunchKernelGGL(kMeshVertices, dim3(tiles), dim3(kBlock), 0, s, G, field, (const int *)baseV, vbase, dpos);
	if (nt) hipLaunchKernelGGL(kMeshTriangles, dim3(tiles), dim3(kBlock), 0, s, G, field, (const int *)baseT, (const int *)vbase, didx);
	hipLaunchKernelGGL(kMeshNormals, dim3(unsigned((G.M + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, G, field, (const int *)vbase, dnrm);
	if (hipGetLastError() != hipSuccess ||
		hipMemcpyAsync(pos.data(), dpos, nv * 3 * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
		hipMemcpyAsync(nrm.data(), dnrm, nv * 3 * sizeof(float), hipMemcpyDeviceToHost, s) != hipSuccess ||
		(nt && hipMemcpyAsync(idx.data(), didx, nt * 3 * sizeof(unsigned), hipMemcpyDeviceToHost, s) != hipSuccess) ||
		hipStreamSynchronize(s) != hipSuccess) {
		err = "building the mesh failed";
		return done(GARM_EHIP);
	}
	return done(GARM_OK);
}

} // namespace garm
