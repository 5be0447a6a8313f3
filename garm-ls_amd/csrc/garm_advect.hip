// Gradient-augmented reference-map advection, semi-Lagrangian velocity
// advection and the phi / grad-phi composition for gfx950.
//
// One thread per cell (or face).  The interpolants gather from the 2^D
// corners of the cell that holds the traced point, so the data each corner
// contributes are first packed per cell / face into one contiguous record
// (array of structures, 16-byte vector loads): the velocity with its
// central-difference gradient per face, the Hermite data of the reference
// map (value, first and mixed derivatives per component) per cell, and the
// same for the reference level set.  A gather then costs one clamped index
// and a few vector loads per corner instead of a dozen scalar loads with
// their own clamped neighbour indices.  The float reference-map records are
// planar instead (six planes of one float4 per cell in 3D): a record is 96
// bytes, and neighbouring lanes reading it whole would spread each wave's
// 16-byte loads over 96-byte strides, six times the cache lines.  The packed values are the products
// the reference's interpolants form (HermiteIntrpl.h:17-30, 62-87;
// Derivatives.h:98-105), computed once per cell in the same operations.
#include "garm_internal.hpp"

#include <algorithm>
#include <type_traits>

namespace garm {

// ---------------------------------------------------------------- velocity
// Face velocity arrays of element type S (double, or a float copy for the
// FP32 kernels) on their face grids.
template <typename S>
struct FaceArrays {
	const S *a[3];
	GridDev g[3];
};

// VectorFieldWrapper<SGrid, Linear>::operator() (VectorField.h:66-72):
// per axis, the multilinear sample of u_a on face grid a (LagrangeIntrpl.h:40-72).
// In double the corner weights are formed and summed as the reference does;
// in float the lerps are nested along x, y, z.
template <int D, typename T, typename S>
__device__ __forceinline__ void sampleVel(const FaceArrays<S> &U, const T *p, T *v)
{
#pragma unroll
	for (int a = 0; a < D; a++) {
		const GridDev &g = U.g[a];
		const S *ua = U.a[a];
		int l[3]; T f[3];
		lowerFracT<D, T>(g, p, l, f);
		unsigned ox[2], oy[2], oz[2];
		cornerOffsets<D, 2>(g, l, ox, oy, oz);
		if constexpr (std::is_same<T, float>::value) {
			T yz[2];
#pragma unroll
			for (int k = 0; k < (D == 3 ? 2 : 1); k++) {
				T row[2];
#pragma unroll
				for (int j = 0; j < 2; j++) {
					const T u0 = T(ua[ox[0] + oy[j] + oz[k]]), u1 = T(ua[ox[1] + oy[j] + oz[k]]);
					row[j] = u0 + (u1 - u0) * f[0];
				}
				yz[k] = row[0] + (row[1] - row[0]) * f[1];
			}
			v[a] = D == 3 ? yz[0] + (yz[1] - yz[0]) * f[2] : yz[0];
		} else {
			T w[2][3];
#pragma unroll
			for (int b = 0; b < D; b++) { w[0][b] = 1 - f[b]; w[1][b] = f[b]; }
			T val = 0;
#pragma unroll
			for (int k = 0; k < (D == 3 ? 2 : 1); k++)
#pragma unroll
				for (int j = 0; j < 2; j++)
#pragma unroll
					for (int i = 0; i < 2; i++) {
						const T wt = D == 2 ? w[i][0] * w[j][1] : w[i][0] * w[j][1] * w[k][2];
						val += T(ua[ox[i] + oy[j] + oz[k]]) * wt;
					}
			v[a] = val;
		}
	}
}

// ------------------------------------------------------------------ packing
template <typename T> struct Vec4;
template <> struct Vec4<float> { using type = float4; };
template <> struct Vec4<double> { using type = double4; };

// S consecutive values of record q (S a multiple of four)
template <typename T, int S>
__device__ __forceinline__ void loadRecord(const T *base, size_t q, T (&v)[S])
{
	using V = typename Vec4<T>::type;
	const V *p = reinterpret_cast<const V *>(base) + q * (S / 4);
#pragma unroll
	for (int i = 0; i < S / 4; i++) {
		const V t = p[i];
		v[4 * i] = t.x; v[4 * i + 1] = t.y; v[4 * i + 2] = t.z; v[4 * i + 3] = t.w;
	}
}

// Record q of a planar float layout: S/4 planes of n float4, plane i holding
// values 4i..4i+3 of every record.  Lanes whose records are neighbours then
// read one contiguous run per plane.
template <int S>
__device__ __forceinline__ void loadRecordPlanar(const float *base, size_t n, size_t q, float (&v)[S])
{
	const float4 *p = reinterpret_cast<const float4 *>(base) + q;
#pragma unroll
	for (int i = 0; i < S / 4; i++) {
		const float4 t = p[i * n];
		v[4 * i] = t.x; v[4 * i + 1] = t.y; v[4 * i + 2] = t.z; v[4 * i + 3] = t.w;
	}
}

template <typename T, int S>
__device__ __forceinline__ void storeRecord(T *base, size_t q, const T (&v)[S])
{
	using V = typename Vec4<T>::type;
	V *p = reinterpret_cast<V *>(base) + q * (S / 4);
#pragma unroll
	for (int i = 0; i < S / 4; i++) p[i] = V{v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3]};
}

// Face record of axis a: u_a and its central differences along x, y, z
// (0 for a missing axis) -- what sampleVelGrad reads at one corner.
template <int D, typename T>
__global__ __launch_bounds__(kBlock) void kPackVelocity(GridDev g, const double *ua, T *out, float *copy)
{
	const Coord c = threadCoord();
	if (!inGrid<D>(g, c)) return;
	const size_t id = gidx<D>(g, c.i, c.j, c.k);
	const T inv = T(g.inv);
	T v[4];
	v[0] = T(ua[id]);
	v[1] = T(ua[gat<D>(g, c.i + 1, c.j, c.k)] - ua[gat<D>(g, c.i - 1, c.j, c.k)]) / 2 * inv;
	v[2] = T(ua[gat<D>(g, c.i, c.j + 1, c.k)] - ua[gat<D>(g, c.i, c.j - 1, c.k)]) / 2 * inv;
	v[3] = D == 3 ? T(ua[gat<D>(g, c.i, c.j, c.k + 1)] - ua[gat<D>(g, c.i, c.j, c.k - 1)]) / 2 * inv : T(0);
	storeRecord<T, 4>(out, id, v);
	if (copy) copy[id] = float(ua[id]);
}

// kPackVelocity in 3D for a column of kPackColumn z-planes per thread: the
// values at k - 1, k, k + 1 stay in registers from plane to plane, so a face
// loads its x and y neighbours and one new z value instead of seven values
// (the z re-reads otherwise come from other XCDs' rows).  Same operations,
// same records.
constexpr int kPackColumn = 16;

template <typename T>
__global__ __launch_bounds__(kBlock) void kPackVelocityColumn(GridDev g, const double *ua, T *out, float *copy)
{
	const int i = int(blockIdx.x * blockDim.x + threadIdx.x), j = int(blockIdx.y * blockDim.y + threadIdx.y);
	if (i >= g.n[0] || j >= g.n[1]) return;
	const int z0 = int(blockIdx.z) * kPackColumn, z1 = min(z0 + kPackColumn, g.n[2]);
	const T inv = T(g.inv);
	double um = ua[gat<3>(g, i, j, z0 - 1)], uc = ua[gat<3>(g, i, j, z0)];
	for (int k = z0; k < z1; k++) {
		const double up = ua[gat<3>(g, i, j, k + 1)];
		const size_t id = gidx<3>(g, i, j, k);
		T v[4];
		v[0] = T(uc);
		v[1] = T(ua[gat<3>(g, i + 1, j, k)] - ua[gat<3>(g, i - 1, j, k)]) / 2 * inv;
		v[2] = T(ua[gat<3>(g, i, j + 1, k)] - ua[gat<3>(g, i, j - 1, k)]) / 2 * inv;
		v[3] = T(up - um) / 2 * inv;
		storeRecord<T, 4>(out, id, v);
		if (copy) copy[id] = float(uc);
		um = uc;
		uc = up;
	}
}

// DrvVectorFieldWrapper<SGrid, Linear> (VectorField.h:106-138): per axis a,
// the multilinear sample of u_a and of its face gradient
// (Derivatives::computeFirst<2>, Derivatives.h:47,51-57,98-105) taken at each
// clamped corner, from the face records.  Row a of G is grad u_a (row-major).
template <int D, typename T>
__device__ __forceinline__ void sampleVelGradPacked(const FaceSet &U, const T *const *rec, const T *p, T *v, T *G)
{
#pragma unroll
	for (int a = 0; a < D; a++) {
		const GridDev &g = U.g[a];
		int l[3]; T f[3];
		lowerFracT<D, T>(g, p, l, f);
		T w[2][3];
#pragma unroll
		for (int b = 0; b < D; b++) { w[0][b] = 1 - f[b]; w[1][b] = f[b]; }
		unsigned ox[2], oy[2], oz[2];
		cornerOffsets<D, 2>(g, l, ox, oy, oz);
		T val = 0, gr[3] = {0, 0, 0};
#pragma unroll
		for (int k = 0; k < (D == 3 ? 2 : 1); k++)
#pragma unroll
			for (int j = 0; j < 2; j++)
#pragma unroll
				for (int i = 0; i < 2; i++) {
					const T wt = D == 2 ? w[i][0] * w[j][1] : w[i][0] * w[j][1] * w[k][2];
					T d[4];
					loadRecord<T, 4>(rec[a], ox[i] + oy[j] + oz[k], d);
					val += d[0] * wt;
#pragma unroll
					for (int b = 0; b < D; b++) gr[b] += d[1 + b] * wt;
				}
		v[a] = val;
#pragma unroll
		for (int b = 0; b < D; b++) G[a * D + b] = gr[b];
	}
}

// sampleVelGradPacked at a cell centre, in float: along axis a the centre
// lies halfway between faces i and i + 1 of face grid a and on that grid's
// points along the other axes, so the multilinear sample is the mean of those
// two face records (the other corners carry weight 0).
template <int D>
__device__ __forceinline__ void sampleVelGradCentre(const FaceSet &U, const float *const *rec, const Coord &c, float *v, float *G)
{
#pragma unroll
	for (int a = 0; a < D; a++) {
		const GridDev &g = U.g[a];
		float d0[4], d1[4];
		loadRecord<float, 4>(rec[a], gidx<D>(g, c.i, c.j, c.k), d0);
		loadRecord<float, 4>(rec[a], gidx<D>(g, c.i + (a == 0), c.j + (a == 1), c.k + (a == 2)), d1);
		v[a] = (d0[0] + d1[0]) * 0.5f;
#pragma unroll
		for (int b = 0; b < D; b++) G[a * D + b] = (d0[1 + b] + d1[1 + b]) * 0.5f;
	}
}

// ----------------------------------------------------------------- Hermite
// 1-D cubic Hermite basis (HermiteIntrpl.h:89-125) for one axis:
// B[vtx][ord] value weights, Bg[vtx][ord] their derivatives.
template <typename T>
__device__ __forceinline__ void hermiteBasis(T x, T B[2][2], T Bg[2][2])
{
	B[0][0] = (x + x + 1) * (x - 1) * (x - 1); Bg[0][0] = 6 * x * (x - 1);
	B[1][0] = x * x * (3 - x - x);             Bg[1][0] = 6 * x * (1 - x);
	B[0][1] = x * (x - 1) * (x - 1);           Bg[0][1] = (3 * x - 1) * (x - 1);
	B[1][1] = x * x * (x - 1);                 Bg[1][1] = x * (3 * x - 2);
}

// Hermite record of the reference map at one cell (S = 2^D orders x D
// components, rec[ord * D + r]), in the getValues order of HermiteIntrpl
// (HermiteIntrpl.h:62-87): 0 value, 1 d/dx, 2 d/dy, 3 dxy, 4 d/dz, 5 dxz,
// 6 dyz, 7 dxyz, each scaled by dx per derivative order; the mixed
// derivatives are the central differences of J of the constructor
// (:17-30).  In float the map is stored as the displacement: xi - x_cell and
// J - I, taken in FP64 before rounding (cubic Hermite reproduces the
// identity map exactly, so the interpolant of the displacement plus the
// traced point is the same interpolant without the cancellation of
// positions of size 1 against gradients of size 1/dx).
template <int D>
struct RefMapFields {
	const double *xi[3];
	const double *J[9];
};

template <int D, typename T>
__global__ __launch_bounds__(kBlock) void kPackRefMap(GridDev g, RefMapFields<D> A, T *out)
{
	constexpr int S = (1 << D) * D;
	constexpr bool kDisp = !std::is_same<T, double>::value;
	__shared__ T stage[kBlock / kWave][kWave * S];
	const Coord c = threadCoord();
	const bool live = inGrid<D>(g, c);
	const int lane = int(threadIdx.x);   // the block is one wave wide in x
	T rec[S];
	if (live) {
		const size_t q = gidx<D>(g, c.i, c.j, c.k);
		const int qi = c.i, qj = c.j, qk = c.k;
		const int pos[3] = {qi, qj, qk};
		const T dx = T(g.dx), inv = T(g.inv);
		// x / y neighbours (clamped) shared by every component
		size_t xp, xm, yp, ym, pp, mp, pm, mm;
		if (qi >= 1 && qi + 1 < g.n[0] && qj >= 1 && qj + 1 < g.n[1]) {
			const size_t sy = size_t(g.n[0]);
			xp = q + 1; xm = q - 1; yp = q + sy; ym = q - sy;
			pp = q + 1 + sy; mp = q - 1 + sy; pm = q + 1 - sy; mm = q - 1 - sy;
		} else {
			xp = gat<D>(g, qi + 1, qj, qk); xm = gat<D>(g, qi - 1, qj, qk);
			yp = gat<D>(g, qi, qj + 1, qk); ym = gat<D>(g, qi, qj - 1, qk);
			pp = gat<D>(g, qi + 1, qj + 1, qk); mp = gat<D>(g, qi - 1, qj + 1, qk);
			pm = gat<D>(g, qi + 1, qj - 1, qk); mm = gat<D>(g, qi - 1, qj - 1, qk);
		}
#pragma unroll
		for (int r = 0; r < D; r++) {
			T v[8];
			if constexpr (kDisp) {
				v[0] = T(A.xi[r][q] - (g.o[r] + pos[r] * g.dx));
				v[1] = T(A.J[0 * D + r][q] - (r == 0)) * dx;
				v[2] = T(A.J[1 * D + r][q] - (r == 1)) * dx;
			} else {
				v[0] = T(A.xi[r][q]);
				v[1] = T(A.J[0 * D + r][q]) * dx;
				v[2] = T(A.J[1 * D + r][q]) * dx;
			}
			{
				const double *c1 = A.J[1 * D + r];
				const T m = T(c1[xp] - c1[xm]) / 2 * inv;
				v[3] = m * dx * dx;
			}
			if constexpr (D == 3) {
				const double *c2 = A.J[2 * D + r];
				v[4] = kDisp ? T(c2[q] - (r == 2)) * dx : T(c2[q]) * dx;
				T m = T(c2[xp] - c2[xm]) / 2 * inv;
				v[5] = m * dx * dx;
				m = T(c2[yp] - c2[ym]) / 2 * inv;
				v[6] = m * dx * dx;
				m = T(c2[pp] - c2[mp] - c2[pm] + c2[mm]) / 4 * inv * inv;
				v[7] = m * dx * dx * dx;
			}
#pragma unroll
			for (int ord = 0; ord < (1 << D); ord++) rec[ord * D + r] = v[ord];
		}
		if constexpr (kDisp) {
			// planar: S/4 planes of one float4 per cell, each wave's stores one contiguous run
			const size_t n = size_t(g.n[0]) * size_t(g.n[1]) * size_t(D == 3 ? g.n[2] : 1);
			float4 *p = reinterpret_cast<float4 *>(out) + q;
#pragma unroll
			for (int e = 0; e < S / 4; e++) p[e * n] = make_float4(rec[4 * e], rec[4 * e + 1], rec[4 * e + 2], rec[4 * e + 3]);
		} else {
#pragma unroll
			for (int e = 0; e < S; e++) stage[threadIdx.y][lane * S + e] = rec[e];
		}
	}
	if constexpr (!kDisp) {
		__syncthreads();
		// the wave's records are one run in memory: write it as consecutive 16-byte words
		const int i0 = int(blockIdx.x * blockDim.x);
		const bool row = c.j < g.n[1] && (D == 2 || c.k < g.n[2]);
		const int count = row ? min(kWave, g.n[0] - i0) : 0;
		using V = typename Vec4<T>::type;
		const V *src = reinterpret_cast<const V *>(stage[threadIdx.y]);
		V *dst = reinterpret_cast<V *>(out + gidx<D>(g, row ? i0 : 0, row ? c.j : 0, row ? c.k : 0) * S);
		for (int e = lane; e < count * S / 4; e += kWave) dst[e] = src[e];
	}
}

// HermiteIntrpl<D, VectorDd>::interpolate + interpolateGradient
// (HermiteIntrpl.h:32-58) of the reference map from the cell records.
// Gh is column-major: Gh[c*D + r].  In float a corner clamped into the grid
// keeps its own position in the displacement (the reference reads the
// clamped corner's data at the unclamped corner).
template <int D, typename T>
__device__ __forceinline__ void hermiteRefMap(const GridDev &g, const T *recs, const T *X, T *val, T *Gh)
{
	constexpr int S = (1 << D) * D;
	constexpr bool kDisp = !std::is_same<T, double>::value;
	int l[3] = {0, 0, 0}; T f[3];
	lowerFracT<D, T>(g, X, l, f);
	T B[3][2][2], Bg[3][2][2];
#pragma unroll
	for (int a = 0; a < D; a++) hermiteBasis(f[a], B[a], Bg[a]);
#pragma unroll
	for (int r = 0; r < D; r++) val[r] = 0;
#pragma unroll
	for (int e = 0; e < D * D; e++) Gh[e] = 0;
	const T inv = T(g.inv);
	unsigned ox[2], oy[2], oz[2];
	cornerOffsets<D, 2>(g, l, ox, oy, oz);

#pragma unroll
	for (int cn = 0; cn < (1 << D); cn++) {
		const int vx = cn & 1, vy = cn >> 1 & 1, vz = D == 3 ? cn >> 2 : 0;
		const int lv[3] = {l[0] + vx, l[1] + vy, l[2] + vz};
		const int qv[3] = {clampi(lv[0], g.n[0] - 1), clampi(lv[1], g.n[1] - 1), D == 3 ? clampi(lv[2], g.n[2] - 1) : 0};
		T rec[S];
		loadRecord<T, S>(recs, ox[vx] + oy[vy] + oz[vz], rec);
		if constexpr (kDisp) {
#pragma unroll
			for (int r = 0; r < D; r++)
				if (qv[r] != lv[r]) rec[r] += T((qv[r] - lv[r]) * g.dx);
		}
#pragma unroll
		for (int ord = 0; ord < (1 << D); ord++) {
			const int ox = ord & 1, oy = ord >> 1 & 1, oz = D == 3 ? ord >> 2 : 0;
			T wt = B[0][vx][ox] * B[1][vy][oy];
			T wg[3];
			if constexpr (D == 2) {
				wg[0] = Bg[0][vx][ox] * B[1][vy][oy];
				wg[1] = B[0][vx][ox] * Bg[1][vy][oy];
			} else {
				wt = wt * B[2][vz][oz];
				wg[0] = Bg[0][vx][ox] * B[1][vy][oy] * B[2][vz][oz];
				wg[1] = B[0][vx][ox] * Bg[1][vy][oy] * B[2][vz][oz];
				wg[2] = B[0][vx][ox] * B[1][vy][oy] * Bg[2][vz][oz];
			}
#pragma unroll
			for (int r = 0; r < D; r++) {
				const T v = rec[ord * D + r];
				val[r] += v * wt;
#pragma unroll
				for (int c = 0; c < D; c++) Gh[c * D + r] += v * wg[c];
			}
		}
	}
#pragma unroll
	for (int e = 0; e < D * D; e++) Gh[e] *= inv;
	if constexpr (kDisp) {
#pragma unroll
		for (int r = 0; r < D; r++) { val[r] += X[r]; Gh[r * D + r] += T(1); }
	}
}

// The same interpolant in float as a tensor product: the corners in
// (vy, vz) pairs, each pair's two records contracted along x first (value
// and x-derivative basis), then weighted in y and z.  About 600 operations
// per cell instead of 1800; the rounding differs from the corner-by-corner
// sum only in float.
template <int D>
__device__ __forceinline__ void hermiteRefMapFast(const GridDev &g, const float *recs, const float *X, float *val, float *Gh)
{
	constexpr int S = (1 << D) * D;
	const size_t n = size_t(g.n[0]) * size_t(g.n[1]) * size_t(D == 3 ? g.n[2] : 1);
	int l[3] = {0, 0, 0}; float f[3];
	lowerFracT<D, float>(g, X, l, f);
	float B[3][2][2], Bg[3][2][2];
#pragma unroll
	for (int a = 0; a < D; a++) hermiteBasis(f[a], B[a], Bg[a]);
	unsigned ox[2], oy[2], oz[2];
	cornerOffsets<D, 2>(g, l, ox, oy, oz);
	// a corner clamped into the grid keeps its own position in the displacement
	float cx[2], cy[2], cz[2];
#pragma unroll
	for (int w = 0; w < 2; w++) {
		const int lx = l[0] + w, ly = l[1] + w, lz = l[2] + w;
		cx[w] = float((clampi(lx, g.n[0] - 1) - lx) * g.dx);
		cy[w] = float((clampi(ly, g.n[1] - 1) - ly) * g.dx);
		cz[w] = D == 3 ? float((clampi(lz, g.n[2] - 1) - lz) * g.dx) : 0.f;
	}
	float v[3] = {0, 0, 0}, gx[3] = {0, 0, 0}, gy[3] = {0, 0, 0}, gz[3] = {0, 0, 0};
#pragma unroll
	for (int vz = 0; vz < (D == 3 ? 2 : 1); vz++)
#pragma unroll
		for (int vy = 0; vy < 2; vy++) {
			float r0[S], r1[S];
			loadRecordPlanar<S>(recs, n, ox[0] + oy[vy] + oz[vz], r0);
			loadRecordPlanar<S>(recs, n, ox[1] + oy[vy] + oz[vz], r1);
			r0[0] += cx[0]; r1[0] += cx[1];
			r0[1] += cy[vy]; r1[1] += cy[vy];
			if constexpr (D == 3) { r0[2] += cz[vz]; r1[2] += cz[vz]; }
#pragma unroll
			for (int wz = 0; wz < (D == 3 ? 2 : 1); wz++)
#pragma unroll
				for (int wy = 0; wy < 2; wy++) {
					const int o0 = (wy << 1) | (wz << 2), o1 = o0 | 1;
					const float by = B[1][vy][wy], bgy = Bg[1][vy][wy];
					const float bz = D == 3 ? B[2][vz][wz] : 1.f, bgz = D == 3 ? Bg[2][vz][wz] : 0.f;
					const float wv = by * bz, wgy = bgy * bz, wgz = by * bgz;
#pragma unroll
					for (int r = 0; r < D; r++) {
						const float a0 = r0[o0 * D + r], a1 = r0[o1 * D + r], b0 = r1[o0 * D + r], b1 = r1[o1 * D + r];
						const float tx = a0 * B[0][0][0] + a1 * B[0][0][1] + b0 * B[0][1][0] + b1 * B[0][1][1];
						const float tgx = a0 * Bg[0][0][0] + a1 * Bg[0][0][1] + b0 * Bg[0][1][0] + b1 * Bg[0][1][1];
						v[r] += tx * wv;
						gx[r] += tgx * wv;
						gy[r] += tx * wgy;
						if constexpr (D == 3) gz[r] += tx * wgz;
					}
				}
		}
	const float inv = float(g.inv);
#pragma unroll
	for (int r = 0; r < D; r++) {
		val[r] = v[r] + X[r];
		Gh[0 * D + r] = gx[r] * inv;
		Gh[1 * D + r] = gy[r] * inv;
		if constexpr (D == 3) Gh[2 * D + r] = gz[r] * inv;
		Gh[r * D + r] += 1.f;
	}
}

// ------------------------------------------------------------------ kernels
template <int D, typename T>
struct RefMapArgs {
	GridDev cell;
	FaceSet U;
	const T *vrec[3];   // face records of u
	const T *hrec;      // Hermite records of the reference map
	double *xi_new[3];
	double *J_new[9];
	double h;   // -dt
};

// SemiLagrangian<D, 2, HermiteIntrpl<D, VectorDd>>::advect (drv overload,
// Advection.h:147-160) with Advection::traceDrv<2> (Advection.h:71-82).
template <int D, typename T>
__global__ __launch_bounds__(kBlock) void kAdvectRefMap(RefMapArgs<D, T> A)
{
	const Coord c = threadCoord();
	if (!inGrid<D>(A.cell, c)) return;
	const size_t id = gidx<D>(A.cell, c.i, c.j, c.k);
	const T h = T(A.h);
	T x[3] = {T(A.cell.o[0] + c.i * A.cell.dx), T(A.cell.o[1] + c.j * A.cell.dx), T(A.cell.o[2] + c.k * A.cell.dx)};

	T v0[3], G0[9], p1[3], R1[9], v1[3], G1[9], X[3], R[9];
	if constexpr (std::is_same<T, float>::value) sampleVelGradCentre<D>(A.U, A.vrec, c, v0, G0);
	else sampleVelGradPacked<D, T>(A.U, A.vrec, x, v0, G0);
#pragma unroll
	for (int a = 0; a < D; a++) p1[a] = x[a] + v0[a] * h;
#pragma unroll
	for (int a = 0; a < D; a++)
#pragma unroll
		for (int b = 0; b < D; b++) R1[a * D + b] = (a == b ? T(1) : T(0)) + G0[a * D + b] * h;
	sampleVelGradPacked<D, T>(A.U, A.vrec, p1, v1, G1);
#pragma unroll
	for (int a = 0; a < D; a++) X[a] = x[a] + (v0[a] + v1[a]) * h / 2;
#pragma unroll
	for (int a = 0; a < D; a++)
#pragma unroll
		for (int b = 0; b < D; b++) {
			T gr = 0;
#pragma unroll
			for (int k = 0; k < D; k++) gr += G1[a * D + k] * R1[k * D + b];
			R[a * D + b] = (a == b ? T(1) : T(0)) + (G0[a * D + b] + gr) * h / 2;
		}

	T val[3], Gh[9];
	if constexpr (std::is_same<T, float>::value) hermiteRefMapFast<D>(A.cell, A.hrec, X, val, Gh);
	else hermiteRefMap<D, T>(A.cell, A.hrec, X, val, Gh);
#pragma unroll
	for (int r = 0; r < D; r++) A.xi_new[r][id] = val[r];
#pragma unroll
	for (int r = 0; r < D; r++)
#pragma unroll
		for (int col = 0; col < D; col++) {
			T s = 0;
#pragma unroll
			for (int k = 0; k < D; k++) s += Gh[k * D + r] * R[k * D + col];
			A.J_new[col * D + r][id] = s;
		}
}

// SemiLagrangian<D, 2, CubicIntrpl<D>>::advect on face grid `axis`
// (Advection.h:15-22, 37-43, 137-145; CubicIntrpl LagrangeIntrpl.h:128-171).
template <int D, typename T, typename S>
__global__ __launch_bounds__(kBlock) void kAdvectVelocity(FaceArrays<S> U, int axis, double *out, double hd)
{
	const GridDev &g = U.g[axis];
	const Coord c = threadCoord();
	if (!inGrid<D>(g, c)) return;
	const T h = T(hd);
	T x[3] = {T(g.o[0] + c.i * g.dx), T(g.o[1] + c.j * g.dx), T(g.o[2] + c.k * g.dx)};
	T v0[3], p1[3], v1[3], X[3];
	sampleVel<D, T, S>(U, x, v0);
#pragma unroll
	for (int a = 0; a < D; a++) p1[a] = x[a] + v0[a] * h;
	sampleVel<D, T, S>(U, p1, v1);
#pragma unroll
	for (int a = 0; a < D; a++) X[a] = x[a] + (v0[a] + v1[a]) * h / 2;

	int l[3]; T f[3];
#pragma unroll
	for (int a = 0; a < D; a++) {
		const T o = T(g.o[a]), inv = T(g.inv);
		l[a] = int(floor((X[a] - o) * inv)) - 1;   // getCubicLower (Grid.h:54)
		f[a] = (X[a] - o - l[a] * T(g.dx)) * inv;
	}
	T w[4][3];
#pragma unroll
	for (int a = 0; a < D; a++) {
		const T t = f[a];
		w[0][a] = (t - 1) * (t - 2) * (t - 3) / -6;
		w[1][a] = t * (t - 2) * (t - 3) / 2;
		w[2][a] = t * (t - 1) * (t - 3) / -2;
		w[3][a] = t * (t - 1) * (t - 2) / 6;
	}
	const S *ua = U.a[axis];
	unsigned ox[4], oy[4], oz[4];
	cornerOffsets<D, 4>(g, l, ox, oy, oz);
	T val = 0;
	if constexpr (std::is_same<T, float>::value) {
		// separable: rows along x, then y, then z
#pragma unroll
		for (int k = 0; k < (D == 3 ? 4 : 1); k++) {
			T plane = 0;
#pragma unroll
			for (int j = 0; j < 4; j++) {
				T row = 0;
#pragma unroll
				for (int i = 0; i < 4; i++) row += T(ua[ox[i] + oy[j] + oz[k]]) * w[i][0];
				plane += row * w[j][1];
			}
			val += D == 3 ? plane * w[k][2] : plane;
		}
	} else {
#pragma unroll
		for (int k = 0; k < (D == 3 ? 4 : 1); k++)
#pragma unroll
			for (int j = 0; j < 4; j++)
#pragma unroll
				for (int i = 0; i < 4; i++) {
					const T wt = D == 2 ? w[i][0] * w[j][1] : w[i][0] * w[j][1] * w[k][2];
					val += T(ua[ox[i] + oy[j] + oz[k]]) * wt;
				}
	}
	out[gidx<D>(g, c.i, c.j, c.k)] = val;
}

// The float form with the cubic's input staged in LDS: the workgroup's 64 x 4
// faces of one z-plane and three faces of halo on every side (70 x 10 x 7
// floats in 3D, clamped into the grid as GridBasedData::at clamps), which
// covers every cubic stencil of a backtrace shorter than one cell.  A face
// whose stencil leaves the tile (a longer backtrace) reads the global copy;
// the multilinear samples of the trace read the global copies (tiling them
// too was measured slower: the larger fill and LDS footprint cost more than
// the gathers).  Same operations in the same order as
// kAdvectVelocity<D, float, float>, so the results are the same.
constexpr int kVelHalo = 3;

template <int D>
__global__ __launch_bounds__(kBlock) void kAdvectVelocityTile(FaceArrays<float> U, int axis, double *out, double hd)
{
	constexpr int TXT = 64 + 2 * kVelHalo, TYT = 4 + 2 * kVelHalo, TZT = D == 3 ? 1 + 2 * kVelHalo : 1;
	constexpr int kTile = TZT * TYT * TXT, kFill = (kTile + kBlock - 1) / kBlock;
	__shared__ float tile[kTile];
	const GridDev &g = U.g[axis];
	const float *ua = U.a[axis];
	const int tx0 = int(blockIdx.x * blockDim.x) - kVelHalo, ty0 = int(blockIdx.y * blockDim.y) - kVelHalo;
	const int tz0 = D == 3 ? int(blockIdx.z) - kVelHalo : 0;
	const int tid = int(threadIdx.x + blockDim.x * threadIdx.y);
	{
		// all of a thread's fill loads in flight together
		float v[kFill];
#pragma unroll
		for (int r = 0; r < kFill; r++) {
			const int q = tid + r * kBlock;
			const int tx = q % TXT, ty = (q / TXT) % TYT, tz = q / (TXT * TYT);
			v[r] = q < kTile ? ua[gat<D>(g, tx0 + tx, ty0 + ty, tz0 + tz)] : 0.0f;
		}
#pragma unroll
		for (int r = 0; r < kFill; r++) {
			const int q = tid + r * kBlock;
			if (q < kTile) tile[q] = v[r];
		}
	}
	__syncthreads();
	const Coord c = threadCoord();
	if (!inGrid<D>(g, c)) return;
	const float h = float(hd);
	float x[3] = {float(g.o[0] + c.i * g.dx), float(g.o[1] + c.j * g.dx), float(g.o[2] + c.k * g.dx)};
	float v0[3], p1[3], v1[3], X[3];
	sampleVel<D, float, float>(U, x, v0);
#pragma unroll
	for (int a = 0; a < D; a++) p1[a] = x[a] + v0[a] * h;
	sampleVel<D, float, float>(U, p1, v1);
#pragma unroll
	for (int a = 0; a < D; a++) X[a] = x[a] + (v0[a] + v1[a]) * h / 2;
	int l[3] = {0, 0, 0}; float f[3];
#pragma unroll
	for (int a = 0; a < D; a++) {
		const float o = float(g.o[a]), inv = float(g.inv);
		l[a] = int(floor((X[a] - o) * inv)) - 1;   // getCubicLower (Grid.h:54)
		f[a] = (X[a] - o - l[a] * float(g.dx)) * inv;
	}
	float w[4][3];
#pragma unroll
	for (int a = 0; a < D; a++) {
		const float t = f[a];
		w[0][a] = (t - 1) * (t - 2) * (t - 3) / -6;
		w[1][a] = t * (t - 2) * (t - 3) / 2;
		w[2][a] = t * (t - 1) * (t - 3) / -2;
		w[3][a] = t * (t - 1) * (t - 2) / 6;
	}
	const int lx = l[0] - tx0, ly = l[1] - ty0, lz = D == 3 ? l[2] - tz0 : 0;
	const bool inside = lx >= 0 && lx + 3 < TXT && ly >= 0 && ly + 3 < TYT && (D == 2 || (lz >= 0 && lz + 3 < TZT));
	float val = 0;
	if (inside) {
#pragma unroll
		for (int k = 0; k < (D == 3 ? 4 : 1); k++) {
			float plane = 0;
#pragma unroll
			for (int j = 0; j < 4; j++) {
				const float *row = tile + ((lz + k) * TYT + ly + j) * TXT + lx;
				float r = 0;
#pragma unroll
				for (int i = 0; i < 4; i++) r += row[i] * w[i][0];
				plane += r * w[j][1];
			}
			val += D == 3 ? plane * w[k][2] : plane;
		}
	} else {
		unsigned ox[4], oy[4], oz[4];
		cornerOffsets<D, 4>(g, l, ox, oy, oz);
#pragma unroll
		for (int k = 0; k < (D == 3 ? 4 : 1); k++) {
			float plane = 0;
#pragma unroll
			for (int j = 0; j < 4; j++) {
				float r = 0;
#pragma unroll
				for (int i = 0; i < 4; i++) r += ua[ox[i] + oy[j] + oz[k]] * w[i][0];
				plane += r * w[j][1];
			}
			val += D == 3 ? plane * w[k][2] : plane;
		}
	}
	out[gidx<D>(g, c.i, c.j, c.k)] = val;
}

// The 3D float form for a column of kVelColumn z-planes per workgroup: the
// 64 x 4 faces of each plane in turn, with a window of the seven planes
// k - 3 .. k + 3 (70 x 10 faces each, clamped into the grid) in LDS as a
// ring.  Moving to the next plane loads one plane of the window instead of
// seven, so each face costs about four staged loads instead of nineteen.
// Stencils that leave the window read the global copy.  The first sample of
// the trace, at the face itself, takes the two or four faces its multilinear
// weights do not zero; the rest is kAdvectVelocityTile's arithmetic.
constexpr int kVelColumn = 16;

template <int D>   // 3 (a template so that each compile of this file has its own instance)
__global__ __launch_bounds__(kBlock) void kAdvectVelocityColumn(FaceArrays<float> U, int axis, double *out, double hd)
{
	static_assert(D == 3, "3D only");
	constexpr int TXT = 64 + 2 * kVelHalo, TYT = 4 + 2 * kVelHalo, W = 1 + 2 * kVelHalo, kPlane = TXT * TYT;
	constexpr int kFill = (kPlane + kBlock - 1) / kBlock;
	__shared__ float win[W][kPlane];
	const GridDev &g = U.g[axis];
	const float *ua = U.a[axis];
	const int tx0 = int(blockIdx.x * blockDim.x) - kVelHalo, ty0 = int(blockIdx.y * blockDim.y) - kVelHalo;
	const int z0 = int(blockIdx.z) * kVelColumn;
	const int tid = int(threadIdx.x + blockDim.x * threadIdx.y);
	const int i = int(blockIdx.x * blockDim.x + threadIdx.x), j = int(blockIdx.y * blockDim.y + threadIdx.y);
	const bool live = i < g.n[0] && j < g.n[1];
	const float h = float(hd);
	// plane p of the grid (clamped) into its ring slot
	auto fillPlane = [&](int p) {
		float v[kFill];
#pragma unroll
		for (int r = 0; r < kFill; r++) {
			const int q = tid + r * kBlock;
			v[r] = q < kPlane ? ua[gat<3>(g, tx0 + q % TXT, ty0 + q / TXT, p)] : 0.0f;
		}
		float *slot = win[(p + W) % W];
#pragma unroll
		for (int r = 0; r < kFill; r++) {
			const int q = tid + r * kBlock;
			if (q < kPlane) slot[q] = v[r];
		}
	};
	for (int p = z0 - kVelHalo; p < z0 + kVelHalo; p++) fillPlane(p);
	for (int k = z0; k < z0 + kVelColumn && k < g.n[2]; k++) {
		fillPlane(k + kVelHalo);
		__syncthreads();
		if (live) {
			float x[3] = {float(g.o[0] + i * g.dx), float(g.o[1] + j * g.dx), float(g.o[2] + k * g.dx)};
			float v0[3], p1[3], v1[3], X[3];
			// the multilinear sample at the face itself: its own component is
			// the face value; component a lies halfway between points
			// idx - 1, idx along `axis` and idx, idx + 1 along a of face
			// grid a (the other corners carry weight 0)
#pragma unroll
			for (int a = 0; a < 3; a++) {
				if (a == axis) { v0[a] = win[(k + W) % W][(j - ty0) * TXT + i - tx0]; continue; }
				const GridDev &ga = U.g[a];
				const float *u = U.a[a];
				const int e0[3] = {axis == 0, axis == 1, axis == 2}, e1[3] = {a == 0, a == 1, a == 2};
				const int b0 = i - e0[0], b1 = j - e0[1], b2 = k - e0[2];
				const float u00 = u[gat<3>(ga, b0, b1, b2)], u10 = u[gat<3>(ga, b0 + e0[0], b1 + e0[1], b2 + e0[2])];
				const float u01 = u[gat<3>(ga, b0 + e1[0], b1 + e1[1], b2 + e1[2])];
				const float u11 = u[gat<3>(ga, b0 + e0[0] + e1[0], b1 + e0[1] + e1[1], b2 + e0[2] + e1[2])];
				v0[a] = ((u00 + u10) + (u01 + u11)) * 0.25f;
			}
#pragma unroll
			for (int a = 0; a < 3; a++) p1[a] = x[a] + v0[a] * h;
			sampleVel<3, float, float>(U, p1, v1);
#pragma unroll
			for (int a = 0; a < 3; a++) X[a] = x[a] + (v0[a] + v1[a]) * h / 2;
			int l[3]; float f[3];
#pragma unroll
			for (int a = 0; a < 3; a++) {
				const float o = float(g.o[a]), inv = float(g.inv);
				l[a] = int(floor((X[a] - o) * inv)) - 1;   // getCubicLower (Grid.h:54)
				f[a] = (X[a] - o - l[a] * float(g.dx)) * inv;
			}
			float w[4][3];
#pragma unroll
			for (int a = 0; a < 3; a++) {
				const float t = f[a];
				w[0][a] = (t - 1) * (t - 2) * (t - 3) / -6;
				w[1][a] = t * (t - 2) * (t - 3) / 2;
				w[2][a] = t * (t - 1) * (t - 3) / -2;
				w[3][a] = t * (t - 1) * (t - 2) / 6;
			}
			const int lx = l[0] - tx0, ly = l[1] - ty0;
			const bool inside = lx >= 0 && lx + 3 < TXT && ly >= 0 && ly + 3 < TYT && l[2] >= k - kVelHalo && l[2] + 3 <= k + kVelHalo;
			float val = 0;
			if (inside) {
#pragma unroll
				for (int kk = 0; kk < 4; kk++) {
					const float *pl = win[(l[2] + kk + W) % W];
					float plane = 0;
#pragma unroll
					for (int jj = 0; jj < 4; jj++) {
						const float *row = pl + (ly + jj) * TXT + lx;
						float r = 0;
#pragma unroll
						for (int ii = 0; ii < 4; ii++) r += row[ii] * w[ii][0];
						plane += r * w[jj][1];
					}
					val += plane * w[kk][2];
				}
			} else {
				unsigned ox[4], oy[4], oz[4];
				cornerOffsets<3, 4>(g, l, ox, oy, oz);
#pragma unroll
				for (int kk = 0; kk < 4; kk++) {
					float plane = 0;
#pragma unroll
					for (int jj = 0; jj < 4; jj++) {
						float r = 0;
#pragma unroll
						for (int ii = 0; ii < 4; ii++) r += ua[ox[ii] + oy[jj] + oz[kk]] * w[ii][0];
						plane += r * w[jj][1];
					}
					val += plane * w[kk][2];
				}
			}
			out[gidx<3>(g, i, j, k)] = val;
		}
		__syncthreads();
	}
}

// float copy of a face array (the FP32 velocity advection's input)
template <typename S>
__global__ void kFaceToFloat(const double *u, S *out, size_t n)
{
	for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) out[i] = float(u[i]);
}

template <int D, typename T>
struct ComposeArgs {
	GridDev cell;
	const double *xi[3];
	const double *J[9];
	const T *rrec;         // Hermite records of the reference level set
	const uint8_t *adv;
	const double *cell_dist;
	double *phi, *phi_grad[3], *phi_tmp, *phi_ori;
	double band;   // 15 dx (LevelSetLiquid.h:93)
	double tol;    // 3 dx (LevelSetLiquid.cpp:1234)
};

// generateCurrentFields (LevelSetLiquid.cpp:1218-1244): phi = H[phi_ref](xi),
// grad phi = J^T grad H[phi_ref](xi) (DrvScalarFieldWrapper with
// HermiteIntrpl<D, double>, ScalarField.h:50-69), air clamp,
// enforceBoundaryLevelSet at 3 dx and 0 (:966-975), phi_tmp, phi_ori.
template <int D, typename T>
__global__ __launch_bounds__(kBlock) void kCompose(ComposeArgs<D, T> A)
{
	const GridDev &g = A.cell;
	const Coord c = threadCoord();
	if (!inGrid<D>(g, c)) return;
	const size_t id = gidx<D>(g, c.i, c.j, c.k);
	T X[3];
#pragma unroll
	for (int a = 0; a < D; a++) X[a] = T(A.xi[a][id]);
	int l[3]; T f[3];
	lowerFracT<D, T>(g, X, l, f);
	T B[3][2][2], Bg[3][2][2];
#pragma unroll
	for (int a = 0; a < D; a++) hermiteBasis(f[a], B[a], Bg[a]);
	unsigned ox[2], oy[2], oz[2];
	cornerOffsets<D, 2>(g, l, ox, oy, oz);
	T val = 0, gr[3] = {0, 0, 0};
	if constexpr (std::is_same<T, float>::value) {
		// tensor product as in hermiteRefMapFast: x contracted per (vy, vz) pair
#pragma unroll
		for (int vz = 0; vz < (D == 3 ? 2 : 1); vz++)
#pragma unroll
			for (int vy = 0; vy < 2; vy++) {
				T r0[1 << D], r1[1 << D];
				loadRecord<T, (1 << D)>(A.rrec, ox[0] + oy[vy] + oz[vz], r0);
				loadRecord<T, (1 << D)>(A.rrec, ox[1] + oy[vy] + oz[vz], r1);
#pragma unroll
				for (int wz = 0; wz < (D == 3 ? 2 : 1); wz++)
#pragma unroll
					for (int wy = 0; wy < 2; wy++) {
						const int o0 = (wy << 1) | (wz << 2), o1 = o0 | 1;
						const T by = B[1][vy][wy], bgy = Bg[1][vy][wy];
						const T bz = D == 3 ? B[2][vz][wz] : T(1), bgz = D == 3 ? Bg[2][vz][wz] : T(0);
						const T tx = r0[o0] * B[0][0][0] + r0[o1] * B[0][0][1] + r1[o0] * B[0][1][0] + r1[o1] * B[0][1][1];
						const T tgx = r0[o0] * Bg[0][0][0] + r0[o1] * Bg[0][0][1] + r1[o0] * Bg[0][1][0] + r1[o1] * Bg[0][1][1];
						const T wv = by * bz;
						val += tx * wv;
						gr[0] += tgx * wv;
						gr[1] += tx * (bgy * bz);
						if constexpr (D == 3) gr[2] += tx * (by * bgz);
					}
			}
	} else {
#pragma unroll
	for (int cn = 0; cn < (1 << D); cn++) {
		const int vx = cn & 1, vy = cn >> 1 & 1, vz = D == 3 ? cn >> 2 : 0;
		T v[1 << D];
		loadRecord<T, (1 << D)>(A.rrec, ox[vx] + oy[vy] + oz[vz], v);
#pragma unroll
		for (int ord = 0; ord < (1 << D); ord++) {
			const int ox = ord & 1, oy = ord >> 1 & 1, oz = D == 3 ? ord >> 2 : 0;
			T wt = B[0][vx][ox] * B[1][vy][oy];
			T wg[3];
			if constexpr (D == 2) {
				wg[0] = Bg[0][vx][ox] * B[1][vy][oy];
				wg[1] = B[0][vx][ox] * Bg[1][vy][oy];
			} else {
				wt = wt * B[2][vz][oz];
				wg[0] = Bg[0][vx][ox] * B[1][vy][oy] * B[2][vz][oz];
				wg[1] = B[0][vx][ox] * Bg[1][vy][oy] * B[2][vz][oz];
				wg[2] = B[0][vx][ox] * B[1][vy][oy] * Bg[2][vz][oz];
			}
			val += v[ord] * wt;
#pragma unroll
			for (int a = 0; a < D; a++) gr[a] += v[ord] * wg[a];
		}
	}
	}
#pragma unroll
	for (int a = 0; a < D; a++) gr[a] *= T(g.inv);
#pragma unroll
	for (int col = 0; col < D; col++) {
		T s = 0;
#pragma unroll
		for (int r = 0; r < D; r++) s += T(A.J[col * D + r][id]) * gr[r];
		A.phi_grad[col][id] = s;
	}
	double phi = val;
	if (A.adv[id] == 0 && phi < A.band) phi = A.band;
	const double cd = A.cell_dist[id];
	if (phi < -cd - A.tol) phi = -cd - A.tol;
	A.phi[id] = phi;
	A.phi_ori[id] = phi;
	A.phi_tmp[id] = phi < -cd - 0.0 ? -cd - 0.0 : phi;
}

// Hermite records of the reference level set for kCompose, in FP64 and
// FP32: HermiteIntrpl<D, double> constructor (HermiteIntrpl.h:17-30: mixed
// derivatives as central differences of the stored gradient) and getValues
// (:62-87) order 0 value, 1 d/dx, 2 d/dy, 3 dxy, 4 d/dz, 5 dxz, 6 dyz, 7 dxyz,
// scaled by dx per derivative order.  Rebuilt whenever _refLevelSet or
// _refLevelSetGrad changes.
template <int D>
__global__ __launch_bounds__(kBlock) void kPackRefLevelSet(GridDev g, const double *phi_ref, const double *g0, const double *g1,
	const double *g2, double *out64, float *out32)
{
	const Coord c = threadCoord();
	if (!inGrid<D>(g, c)) return;
	const size_t id = gidx<D>(g, c.i, c.j, c.k);
	const double inv = g.inv, dx = g.dx;
	double m[4] = {0, 0, 0, 0};
	m[0] = (g1[gat<D>(g, c.i + 1, c.j, c.k)] - g1[gat<D>(g, c.i - 1, c.j, c.k)]) / 2 * inv;
	if constexpr (D == 3) {
		m[1] = (g2[gat<D>(g, c.i + 1, c.j, c.k)] - g2[gat<D>(g, c.i - 1, c.j, c.k)]) / 2 * inv;
		m[2] = (g2[gat<D>(g, c.i, c.j + 1, c.k)] - g2[gat<D>(g, c.i, c.j - 1, c.k)]) / 2 * inv;
		m[3] = (g2[gat<D>(g, c.i + 1, c.j + 1, c.k)] - g2[gat<D>(g, c.i - 1, c.j + 1, c.k)]
			- g2[gat<D>(g, c.i + 1, c.j - 1, c.k)] + g2[gat<D>(g, c.i - 1, c.j - 1, c.k)]) / 4 * inv * inv;
	}
	constexpr int S = 1 << D;
	double r64[S];
	float r32[S];
	const float fdx = float(dx);
	r64[0] = phi_ref[id];        r32[0] = float(phi_ref[id]);
	r64[1] = g0[id] * dx;        r32[1] = float(g0[id]) * fdx;
	r64[2] = g1[id] * dx;        r32[2] = float(g1[id]) * fdx;
	r64[3] = m[0] * dx * dx;     r32[3] = float(m[0]) * fdx * fdx;
	if constexpr (D == 3) {
		r64[4] = g2[id] * dx;        r32[4] = float(g2[id]) * fdx;
		r64[5] = m[1] * dx * dx;     r32[5] = float(m[1]) * fdx * fdx;
		r64[6] = m[2] * dx * dx;     r32[6] = float(m[2]) * fdx * fdx;
		r64[7] = m[3] * dx * dx * dx; r32[7] = float(m[3]) * fdx * fdx * fdx;
	}
	storeRecord<double, S>(out64, id, r64);
	storeRecord<float, S>(out32, id, r32);
}

// Derivatives::computeFirst<4> (Derivatives.h:48, 51-57, 98-105).
template <int D>
__global__ __launch_bounds__(kBlock) void kGradient4(GridDev g, const double *src, double *d0, double *d1, double *d2)
{
	const Coord c = threadCoord();
	if (!inGrid<D>(g, c)) return;
	const size_t id = gidx<D>(g, c.i, c.j, c.k);
	double *dst[3] = {d0, d1, d2};
#pragma unroll
	for (int a = 0; a < D; a++) {
		double v[5];
#pragma unroll
		for (int t = 0; t < 5; t++) {
			const int o = t - 2;
			v[t] = src[gat<D>(g, c.i + (a == 0 ? o : 0), c.j + (a == 1 ? o : 0), c.k + (a == 2 ? o : 0))];
		}
		dst[a][id] = (v[0] - 8 * v[1] + 8 * v[3] - v[4]) / 12 * g.inv;
	}
}

// ---------------------------------------------------------------- launchers
// This file is compiled twice: once without fused multiply-add contraction
// for the FP64 kernels (the IEEE operations of the oracle, so the plain path
// reproduces it bit for bit) and once, with GARM_ADVECT_FP32_UNIT defined and
// contraction on, for the FP32 kernels.  Each object launches only its own
// instantiations; the dispatchers live in the FP64 object.
static FaceSet faceSet(const Fields &F)
{
	FaceSet U{};
	for (int a = 0; a < F.D; a++) { U.a[a] = F.u[a]; U.g[a] = F.face[a]; }
	return U;
}

#ifdef GARM_ADVECT_FP32_UNIT
using Real = float;
#else
using Real = double;
#endif

// Layout of F.adv_rec (sized for FP64 records): the cell records, the face
// records of each axis, then (float kernels) a float copy of each face
// array of u; each part 256-byte aligned.
struct RecordLayout {
	size_t face[3], copy[3];
};

static size_t align256(size_t b) { return (b + 255) / 256 * 256; }

template <int D>
static RecordLayout recordLayout(const Fields &F, size_t real)
{
	RecordLayout L{};
	size_t off = align256(F.N * ((1 << D) * D) * real);
	for (int a = 0; a < D; a++) { L.face[a] = off; off += align256(F.Nf[a] * 4 * real); }
	for (int a = 0; a < D; a++) { L.copy[a] = off; off += align256(F.Nf[a] * sizeof(float)); }
	return L;
}

// Face records of u (and its float copy), then the reference-map records,
// then the trace.
template <int D>
static void advectRefMapD(const Fields &F, double dt, hipStream_t s)
{
	RefMapArgs<D, Real> A{};
	A.cell = F.cell;
	A.U = faceSet(F);
	char *base = static_cast<char *>(F.adv_rec);
	const RecordLayout Ly = recordLayout<D>(F, sizeof(Real));
	constexpr bool fp32 = std::is_same<Real, float>::value;
	for (int a = 0; a < D; a++) {
		Real *rec = reinterpret_cast<Real *>(base + Ly.face[a]);
		float *copy = fp32 ? reinterpret_cast<float *>(base + Ly.copy[a]) : nullptr;
		if constexpr (D == 3) {
			const GridDev &g = F.face[a];
			const dim3 gd(unsigned((g.n[0] + 63) / 64), unsigned((g.n[1] + 3) / 4), unsigned((g.n[2] + kPackColumn - 1) / kPackColumn));
			hipLaunchKernelGGL((kPackVelocityColumn<Real>), gd, gridBlock(), 0, s, g, (const double *)F.u[a], rec, copy);
		} else hipLaunchKernelGGL((kPackVelocity<D, Real>), gridDims(F.face[a], D), gridBlock(), 0, s, F.face[a], (const double *)F.u[a], rec, copy);
		A.vrec[a] = rec;
	}
	RefMapFields<D> R{};
	for (int a = 0; a < D; a++) R.xi[a] = F.xi[a];
	for (int e = 0; e < D * D; e++) R.J[e] = F.J[e];
	Real *hrec = reinterpret_cast<Real *>(base);
	hipLaunchKernelGGL((kPackRefMap<D, Real>), gridDims(F.cell, D), gridBlock(), 0, s, F.cell, R, hrec);
	A.hrec = hrec;
	for (int a = 0; a < D; a++) A.xi_new[a] = F.xi_new[a];
	for (int e = 0; e < D * D; e++) A.J_new[e] = F.J_new[e];
	A.h = -dt;
	hipLaunchKernelGGL((kAdvectRefMap<D, Real>), gridDims(F.cell, D), gridBlock(), 0, s, A);
}

// The float kernels read the float copy of u, made here unless the
// reference-map advection of the same step made it (copiesFresh).
template <int D>
static void advectVelocityD(const Fields &F, double dt, hipStream_t s, bool copiesFresh)
{
	if constexpr (std::is_same<Real, float>::value) {
		FaceArrays<float> U{};
		char *base = static_cast<char *>(F.adv_rec);
		const RecordLayout Ly = recordLayout<D>(F, sizeof(Real));
		for (int a = 0; a < D; a++) {
			float *copy = reinterpret_cast<float *>(base + Ly.copy[a]);
			if (!copiesFresh) {
				const unsigned nb = unsigned(std::min<size_t>((F.Nf[a] + kBlock - 1) / kBlock, 8192));
				hipLaunchKernelGGL(kFaceToFloat<float>, dim3(nb), dim3(kBlock), 0, s, (const double *)F.u[a], copy, F.Nf[a]);
			}
			U.a[a] = copy;
			U.g[a] = F.face[a];
		}
		for (int a = 0; a < D; a++) {
			if constexpr (D == 3) {
				const GridDev &g = F.face[a];
				const dim3 gd(unsigned((g.n[0] + 63) / 64), unsigned((g.n[1] + 3) / 4), unsigned((g.n[2] + kVelColumn - 1) / kVelColumn));
				hipLaunchKernelGGL(kAdvectVelocityColumn<3>, gd, gridBlock(), 0, s, U, a, F.u_new[a], -dt);
			} else hipLaunchKernelGGL((kAdvectVelocityTile<D>), gridDims(F.face[a], D), gridBlock(), 0, s, U, a, F.u_new[a], -dt);
		}
	} else {
		(void)copiesFresh;
		FaceArrays<double> U{};
		for (int a = 0; a < D; a++) { U.a[a] = F.u[a]; U.g[a] = F.face[a]; }
		for (int a = 0; a < D; a++)
			hipLaunchKernelGGL((kAdvectVelocity<D, double, double>), gridDims(F.face[a], D), gridBlock(), 0, s, U, a, F.u_new[a], -dt);
	}
}

template <int D>
static void composeD(const Fields &F, hipStream_t s)
{
	ComposeArgs<D, Real> A{};
	A.cell = F.cell;
	for (int a = 0; a < D; a++) { A.xi[a] = F.xi[a]; A.phi_grad[a] = F.phi_grad[a]; }
	for (int e = 0; e < D * D; e++) A.J[e] = F.J[e];
	if constexpr (std::is_same<Real, double>::value) A.rrec = F.ref_rec64; else A.rrec = F.ref_rec32;
	A.adv = F.adv_region;
	A.cell_dist = F.cell_dist;
	A.phi = F.phi; A.phi_tmp = F.phi_tmp; A.phi_ori = F.phi_ori;
	A.band = 15.0 * F.cell.dx;
	A.tol = 3.0 * F.cell.dx;
	hipLaunchKernelGGL((kCompose<D, Real>), gridDims(F.cell, D), gridBlock(), 0, s, A);
}

#ifdef GARM_ADVECT_FP32_UNIT

void launchAdvectRefMap32(const Fields &F, double dt, hipStream_t s)
{
	if (F.D == 2) advectRefMapD<2>(F, dt, s); else advectRefMapD<3>(F, dt, s);
}

void launchAdvectVelocity32(const Fields &F, double dt, hipStream_t s, bool copiesFresh)
{
	if (F.D == 2) advectVelocityD<2>(F, dt, s, copiesFresh); else advectVelocityD<3>(F, dt, s, copiesFresh);
}

void launchCompose32(const Fields &F, hipStream_t s)
{
	if (F.D == 2) composeD<2>(F, s); else composeD<3>(F, s);
}

#else

void launchAdvectRefMap32(const Fields &F, double dt, hipStream_t s);
void launchAdvectVelocity32(const Fields &F, double dt, hipStream_t s, bool copiesFresh);
void launchCompose32(const Fields &F, hipStream_t s);

void launchAdvectRefMap(const Fields &F, double dt, hipStream_t s, bool fp32)
{
	if (fp32) launchAdvectRefMap32(F, dt, s);
	else if (F.D == 2) advectRefMapD<2>(F, dt, s);
	else advectRefMapD<3>(F, dt, s);
}

void launchAdvectVelocity(const Fields &F, double dt, hipStream_t s, bool fp32, bool copiesFresh)
{
	if (fp32) launchAdvectVelocity32(F, dt, s, copiesFresh);
	else if (F.D == 2) advectVelocityD<2>(F, dt, s, false);
	else advectVelocityD<3>(F, dt, s, false);
}

void launchCompose(const Fields &F, hipStream_t s, bool fp32)
{
	if (fp32) launchCompose32(F, s);
	else if (F.D == 2) composeD<2>(F, s);
	else composeD<3>(F, s);
}

void launchLevelSetRecords(const Fields &F, const double *phi, const double *const *grad, double *out64, float *out32, hipStream_t s)
{
	const double *g = grad[0];
	if (F.D == 2)
		hipLaunchKernelGGL(kPackRefLevelSet<2>, gridDims(F.cell, 2), gridBlock(), 0, s, F.cell, phi, g, (const double *)grad[1], g, out64, out32);
	else
		hipLaunchKernelGGL(kPackRefLevelSet<3>, gridDims(F.cell, 3), gridBlock(), 0, s, F.cell, phi, g, (const double *)grad[1],
			(const double *)grad[2], out64, out32);
}

void launchRefLevelSetRecords(const Fields &F, hipStream_t s)
{
	launchLevelSetRecords(F, F.phi_ref, F.phi_ref_grad, F.ref_rec64, F.ref_rec32, s);
}

void launchGradient4(const Fields &F, const double *src, double *const *dst, hipStream_t s)
{
	if (F.D == 2) hipLaunchKernelGGL(kGradient4<2>, gridDims(F.cell, 2), gridBlock(), 0, s, F.cell, src, dst[0], dst[1], nullptr);
	else hipLaunchKernelGGL(kGradient4<3>, gridDims(F.cell, 3), gridBlock(), 0, s, F.cell, src, dst[0], dst[1], dst[2]);
}

#endif

} // namespace garm
