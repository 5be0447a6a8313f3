// The MC-style volume tables and case formulas below follow
// MC-style-vol-eval/fraction.hpp, which carries this notice:
//
//   Copyright(c) 2021, Tetsuya Takahashi and Christopher Batty
//
//   Permission is hereby granted, free of charge, to any person obtaining a
//   copy of this software and associated documentation files (the
//   "Software"), to deal in the Software without restriction, including
//   without limitation the rights to use, copy, modify, merge, publish,
//   distribute, sublicense, and/or sell copies of the Software, and to
//   permit persons to whom the Software is furnished to do so, subject to
//   the following conditions:
//
//   The above copyright notice and this permission notice shall be included
//   in all copies or substantial portions of the Software.
//
//   THE SOFTWARE IS PROVIDED "AS IS", WITHOUT WARRANTY OF ANY KIND, EXPRESS
//   OR IMPLIED, INCLUDING BUT NOT LIMITED TO THE WARRANTIES OF
//   MERCHANTABILITY, FITNESS FOR A PARTICULAR PURPOSE AND NONINFRINGEMENT.
//   IN NO EVENT SHALL THE AUTHORS OR COPYRIGHT HOLDERS BE LIABLE FOR ANY
//   CLAIM, DAMAGES OR OTHER LIABILITY, WHETHER IN AN ACTION OF CONTRACT,
//   TORT OR OTHERWISE, ARISING FROM, OUT OF OR IN CONNECTION WITH THE
//   SOFTWARE OR THE USE OR OTHER DEALINGS IN THE SOFTWARE.
//
// Liquid volume of the level set (LevelSetLiquid::calculateVolume,
// LevelSetLiquid.cpp:1162-1190), FP64, one fused pass and a deterministic
// two-level reduction.
//
// 3D: for every cell whose +1 corner is in the grid, the MC-style volume
// fraction of the 2x2x2 stencil of phi_tmp (Fraction::get_mc_vol,
// MC-style-vol-eval/fraction.hpp:163-771; Takahashi and Batty, JCGT 2021)
// times dx^3.  The reference passes the stencil in lexicographic order where
// the fraction expects its circular corner order; the reference order keeps
// that, the corrected order permutes the stencil first.
// 2D: the area of the filled marching-squares contour (Contourer<2, true>,
// Contourer.cpp:9-96): nodes with phi <= 0 are inside, edge points are linear
// crossings, a diagonal cell is joined when the mean of its corners is <= 0.
// The contour is planar, so its triangle sum is the exact area of the inside
// polygons, computed here directly.
#include "garm_internal.hpp"
#include "../../include/garm_gpu.h"

#include <mutex>
#include <string>

namespace garm {

// Workgroups of the volume pass (one partial sum each), and the largest
// resolution per axis garm_volume_device accepts.
constexpr int kVolumeBlocks = 1024;
constexpr int kVolumeDeviceMaxRes = 4096;

// -------------------------------------------------- MC-style volume fraction
// Corner order of the fraction: 0 (0,0,0) 1 (1,0,0) 2 (1,1,0) 3 (0,1,0)
// 4 (0,0,1) 5 (1,0,1) 6 (1,1,1) 7 (0,1,1).

__device__ __forceinline__ double lenFrac(double a, double b)
{
	if (a < 0 && b < 0) return 1.0;
	if (a < 0 && 0 <= b) return a / (a - b);
	if (0 <= a && b < 0) return b / (b - a);
	return 0;
}

// Canonical configuration and orientation of each of the 256 inside/outside
// patterns (bit i set when corner i < 0): bits 0-4 the case, 5-9 the corner
// permutation, bit 10 the variant of case 9.
__constant__ uint16_t kMcDispatch[256] = {
	0x000,0x001,0x0c1,0x002,0x101,0x183,0x182,0x0e5,0x061,0x0e2,0x063,0x105,0x102,0x125,0x005,0x008,
	0x021,0x142,0x083,0x1a5,0x104,0x086,0x186,0x10b,0x0e3,0x225,0x187,0x669,0x106,0x52e,0x00c,0x2d1,
	0x0a1,0x003,0x1a2,0x025,0x1a3,0x087,0x2a5,0x5c9,0x0a4,0x0e6,0x1a6,0x10e,0x066,0x52c,0x1cb,0x0b1,
	0x022,0x245,0x145,0x1a8,0x026,0x64c,0x14e,0x211,0x0a6,0x1ab,0x54c,0x071,0x2aa,0x050,0x290,0x294,
	0x281,0x004,0x2e3,0x006,0x2a2,0x2a6,0x1c5,0x1ce,0x2a3,0x166,0x1a7,0x50c,0x1e5,0x00b,0x689,0x2f1,
	0x023,0x146,0x007,0x1ac,0x206,0x0ea,0x5cc,0x270,0x107,0x22c,0x00d,0x02f,0x1ec,0x0d0,0x0ef,0x233,
	0x1c2,0x2e6,0x185,0x14b,0x0c5,0x4cc,0x188,0x171,0x1c6,0x02a,0x18c,0x2d0,0x4ce,0x230,0x271,0x2d4,
	0x045,0x64e,0x6c9,0x291,0x18b,0x1f0,0x091,0x274,0x04c,0x130,0x2af,0x053,0x250,0x0d2,0x1f3,0x055,
	0x041,0x1e3,0x0c4,0x246,0x043,0x2a7,0x126,0x4ec,0x262,0x085,0x1e6,0x52b,0x285,0x009,0x00e,0x051,
	0x2c2,0x265,0x226,0x1ae,0x2c6,0x66c,0x1aa,0x1d0,0x165,0x088,0x16c,0x0d1,0x48b,0x191,0x2f0,0x1d4,
	0x223,0x0e7,0x0c6,0x42c,0x027,0x26d,0x6ac,0x10f,0x266,0x08c,0x00a,0x210,0x68c,0x00f,0x150,0x033,
	0x2e5,0x1a9,0x64b,0x1f1,0x2ec,0x1af,0x170,0x2b3,0x6ee,0x1d1,0x2b0,0x2b4,0x010,0x2f3,0x012,0x295,
	0x282,0x286,0x046,0x18a,0x065,0x06c,0x4cb,0x0b0,0x205,0x48e,0x20c,0x030,0x068,0x151,0x251,0x034,
	0x0a5,0x6eb,0x0ac,0x070,0x46e,0x1b0,0x0f0,0x0b2,0x089,0x2b1,0x08f,0x1b3,0x031,0x1b4,0x013,0x0b5,
	0x2c5,0x6cc,0x18e,0x110,0x189,0x18f,0x231,0x0f3,0x46b,0x190,0x090,0x112,0x1b1,0x093,0x154,0x035,
	0x2e8,0x011,0x131,0x114,0x111,0x073,0x0f4,0x075,0x0f1,0x194,0x193,0x115,0x014,0x0d5,0x015,0x016
};

// kMcRotation[r][i]: corner of the pattern that plays canonical corner i
__constant__ uint8_t kMcRotation[24][8] = {
	{0,1,2,3,4,5,6,7}, {4,5,1,0,7,6,2,3}, {7,6,5,4,3,2,1,0}, {3,2,6,7,0,1,5,4},
	{4,0,3,7,5,1,2,6}, {5,4,7,6,1,0,3,2}, {1,5,6,2,0,4,7,3}, {3,0,1,2,7,4,5,6},
	{2,3,0,1,6,7,4,5}, {1,2,3,0,5,6,7,4}, {0,4,5,1,3,7,6,2}, {0,3,7,4,1,2,6,5},
	{2,1,5,6,3,0,4,7}, {5,1,0,4,6,2,3,7}, {5,6,2,1,4,7,3,0}, {7,3,2,6,4,0,1,5},
	{2,6,7,3,1,5,4,0}, {7,4,0,3,6,5,1,2}, {1,0,4,5,2,3,7,6}, {3,7,4,0,2,6,5,1},
	{6,7,3,2,5,4,0,1}, {6,2,1,5,7,3,0,4}, {4,7,6,5,0,3,2,1}, {6,5,4,7,2,1,0,3}
};

// Edge points of the case formulas: edge e runs between corners kEA[e] and
// kEB[e]; coordinate kES[e] carries the edge fraction (or 1 - it for the
// outside point), the other two are fixed at kEX/kEY/kEZ.
__constant__ int8_t kEA[12] = {0, 1, 2, 0, 4, 5, 6, 4, 0, 1, 2, 3};
__constant__ int8_t kEB[12] = {1, 2, 3, 3, 5, 6, 7, 7, 4, 5, 6, 7};
__constant__ int8_t kES[12] = {0, 1, 0, 1, 0, 1, 0, 1, 2, 2, 2, 2};
__constant__ int8_t kEX[12] = {0, 1, 0, 0, 0, 1, 0, 0, 0, 1, 1, 0};
__constant__ int8_t kEY[12] = {0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1, 1};
__constant__ int8_t kEZ[12] = {0, 0, 0, 0, 1, 1, 1, 1, 0, 0, 0, 0};

struct P3 {
	double x, y, z;
};

__device__ __forceinline__ P3 edgePoint(const double *v, int e, bool inside)
{
	const double f = lenFrac(v[kEA[e]], v[kEB[e]]);
	const double t = inside ? f : 1.0 - f;
	P3 p{double(kEX[e]), double(kEY[e]), double(kEZ[e])};
	if (kES[e] == 0) p.x = t;
	else if (kES[e] == 1) p.y = t;
	else p.z = t;
	return p;
}

// ((q1 - q0) x (q0 - q2)) . q0
__device__ __forceinline__ double signedVol(const P3 &q0, const P3 &q1, const P3 &q2)
{
	const double ax = q1.x - q0.x, ay = q1.y - q0.y, az = q1.z - q0.z;
	const double bx = q0.x - q2.x, by = q0.y - q2.y, bz = q0.z - q2.z;
	return (ay * bz - az * by) * q0.x + (az * bx - ax * bz) * q0.y + (ax * by - ay * bx) * q0.z;
}

__device__ __forceinline__ double corner1(double a, double b, double c, double d)
{
	return lenFrac(a, b) * lenFrac(a, c) * lenFrac(a, d);
}

__device__ __forceinline__ double corner1Out(double a, double b, double c, double d)
{
	return (1.0 - lenFrac(a, b)) * (1.0 - lenFrac(a, c)) * (1.0 - lenFrac(a, d));
}

__device__ __forceinline__ double edge2(double a, double b, double c, double d, double e, double f)
{
	const double m = lenFrac(d, e);
	return (lenFrac(a, b) + m) * lenFrac(b, c) + m * lenFrac(e, f);
}

// Volume of the canonical configuration `n` (variant `m`) with corner values v.
__device__ double mcCase(int n, int m, const double *v)
{
	constexpr double s = 1.0 / 6.0;
	auto E = [&](int e, bool in) { return edgePoint(v, e, in); };
	switch (n) {
	case 0: return 0.0;
	case 1: return s * corner1(v[0], v[1], v[3], v[4]);
	case 2: return s * edge2(v[3], v[0], v[4], v[2], v[1], v[5]);
	case 3: return s * (corner1(v[0], v[1], v[3], v[4]) + corner1(v[5], v[1], v[6], v[4]));
	case 4: return s * (corner1(v[0], v[1], v[3], v[4]) + corner1(v[6], v[7], v[5], v[2]));
	case 5: {
		const double f03 = lenFrac(v[0], v[3]), f37 = lenFrac(v[3], v[7]);
		const double base = lenFrac(v[1], v[5]) * (1.0 - 0.5 * (1.0 - lenFrac(v[0], v[1])) * (1.0 - f03)) / 3.0;
		return base + s * ((lenFrac(v[2], v[6]) + f37) + f03 * f37);
	}
	case 6: return s * (edge2(v[3], v[0], v[4], v[2], v[1], v[5]) + corner1(v[6], v[7], v[5], v[2]));
	case 7: return s * (corner1(v[4], v[0], v[5], v[7]) + corner1(v[1], v[0], v[2], v[5]) + corner1(v[6], v[7], v[5], v[2]));
	case 8: return s * (2.0 * lenFrac(v[0], v[4]) + lenFrac(v[1], v[5]) + 2.0 * lenFrac(v[2], v[6]) + lenFrac(v[3], v[7]));
	case 9:
	case 15: {
		const P3 e0 = E(0, true), e1 = E(1, false), e6 = E(6, true), e7 = E(7, false), e8 = E(8, true), e10 = E(10, true);
		const double f26 = lenFrac(v[2], v[6]), f67 = lenFrac(v[6], v[7]);
		const double cap = (1.0 - 0.5 * (1.0 - f67) * (1.0 - f26)) / 3.0 + s * lenFrac(v[1], v[2]) * f26 + s * (f67 * lenFrac(v[4], v[7]));
		if (n == 15)
			return cap + s * (signedVol(e1, e6, e10) + signedVol(e1, e7, e6) + signedVol(e1, e0, e7) + signedVol(e8, e7, e0)) +
				s * corner1(v[5], v[4], v[1], v[6]);
		if (m == 0) return cap + s * (signedVol(e10, e7, e6) + signedVol(e1, e7, e10) + signedVol(e1, e8, e7) + signedVol(e1, e0, e8));
		return cap + s * (signedVol(e10, e1, e0) + signedVol(e6, e10, e0) + signedVol(e6, e0, e8) + signedVol(e6, e8, e7));
	}
	case 10: return s * (edge2(v[4], v[7], v[6], v[0], v[3], v[2]) + edge2(v[2], v[1], v[0], v[6], v[5], v[4]));
	case 11: {
		const P3 e0 = E(0, true), e1 = E(1, false), e5 = E(5, false), e6 = E(6, false), e8 = E(8, true), e11 = E(11, true);
		const double f56 = lenFrac(v[5], v[6]), f67 = lenFrac(v[6], v[7]);
		const double cap = s * (lenFrac(v[1], v[2]) + f56) + (1.0 - 0.5 * (1.0 - f67) * (1.0 - lenFrac(v[3], v[7]))) / 3.0 + s * (f56 * f67);
		return cap + s * (signedVol(e0, e8, e11) + signedVol(e0, e11, e5) + signedVol(e0, e5, e1) + signedVol(e5, e11, e6));
	}
	case 12: {
		const double f01 = lenFrac(v[0], v[1]), f15 = lenFrac(v[1], v[5]);
		return lenFrac(v[3], v[7]) * (1.0 - 0.5 * (1.0 - f01) * (1.0 - lenFrac(v[0], v[3]))) / 3.0 +
			s * ((f15 + lenFrac(v[2], v[6])) + f01 * f15 + corner1(v[4], v[0], v[5], v[7]));
	}
	case 13:
		return s * (corner1(v[4], v[0], v[5], v[7]) + corner1(v[1], v[0], v[2], v[5]) + corner1(v[6], v[7], v[5], v[2]) +
			corner1(v[3], v[0], v[2], v[7]));
	case 14: {
		const P3 e0 = E(0, false), e3 = E(3, false), e6 = E(6, true), e7 = E(7, false), e9 = E(9, true), e10 = E(10, true);
		const double f26 = lenFrac(v[2], v[6]), f67 = lenFrac(v[6], v[7]);
		const double cap = s * (lenFrac(v[1], v[5]) + f26) + s * (f67 * lenFrac(v[4], v[7])) + (1.0 - 0.5 * (1.0 - f67) * (1.0 - f26)) / 3.0;
		return cap + s * (signedVol(e0, e3, e7) + signedVol(e0, e7, e10) + signedVol(e0, e10, e9) + signedVol(e6, e10, e7));
	}
	case 16: {
		const P3 e1 = E(1, false), e3 = E(3, false), e5 = E(5, true), e6 = E(6, true), e8 = E(8, false), e9 = E(9, false), e10 = E(10, true);
		const double f26 = lenFrac(v[2], v[6]), f56 = lenFrac(v[5], v[6]), f67 = lenFrac(v[6], v[7]);
		return s * (signedVol(e1, e3, e6) + signedVol(e1, e6, e10) + signedVol(e3, e8, e6) + signedVol(e5, e6, e9) + signedVol(e8, e9, e6)) +
			s * (lenFrac(v[1], v[2]) * f26 + f56 * lenFrac(v[1], v[5])) + (1.0 - 0.5 * (1.0 - f67) * (1.0 - f26)) / 3.0 +
			(1.0 - 0.5 * (1.0 - f67) * (1.0 - f56)) / 3.0;
	}
	case 17: {
		const double f03 = lenFrac(v[0], v[3]), f37 = lenFrac(v[3], v[7]);
		return 1.0 - ((1.0 - lenFrac(v[1], v[5])) * (1.0 - 0.5 * lenFrac(v[0], v[1]) * f03) / 3.0 +
			s * ((1.0 - lenFrac(v[2], v[6])) + (1.0 - f37) + (1.0 - f03) * (1.0 - f37)));
	}
	case 18: return 1.0 - s * (corner1Out(v[0], v[1], v[3], v[4]) + corner1Out(v[6], v[5], v[2], v[7]));
	case 19: {
		const P3 e0 = E(0, false), e3 = E(3, false), e4 = E(4, true), e5 = E(5, false), e8 = E(8, false), e9 = E(9, true);
		const double f56 = lenFrac(v[5], v[6]);
		return 1.0 - (s * (-signedVol(e8, e4, e5) - signedVol(e8, e5, e3) - signedVol(e9, e0, e5) - signedVol(e0, e3, e5)) +
			(0.5 * (1.0 - lenFrac(v[1], v[5])) * (1.0 - f56)) / 3.0 + (0.5 * (1.0 - f56) * (1.0 - lenFrac(v[4], v[5]))) / 3.0);
	}
	case 20: {
		const double f12 = lenFrac(v[1], v[2]);
		return 1.0 - s * (((1.0 - lenFrac(v[0], v[3])) + (1.0 - f12)) * (1.0 - lenFrac(v[0], v[4])) + (1.0 - f12) * (1.0 - lenFrac(v[1], v[5])));
	}
	case 21: return 1.0 - s * corner1Out(v[0], v[1], v[3], v[4]);
	case 22: return 1.0;
	default: return 0.0;
	}
}

__device__ double mcVolumeFraction(const double *c)
{
	int idx = 0;
#pragma unroll
	for (int i = 0; i < 8; i++) idx |= (c[i] < 0.0) << i;
	const int d = kMcDispatch[idx];
	const int r = (d >> 5) & 31;
	double v[8];
#pragma unroll
	for (int i = 0; i < 8; i++) v[i] = c[kMcRotation[r][i]];
	return mcCase(d & 31, (d >> 10) & 1, v);
}

// --------------------------------------------- 2D filled contour, per cell
// Nodes n0 (0,0) n1 (1,0) n2 (0,1) n3 (1,1) of one cell of the node grid.
__device__ __forceinline__ double crossing(double a, double b) { return a / (a - b); }

__device__ double filledArea(double p0, double p1, double p2, double p3)
{
	const bool i0 = p0 <= 0, i1 = p1 <= 0, i2 = p2 <= 0, i3 = p3 <= 0;
	const int type = i0 | i1 << 1 | i2 << 2 | i3 << 3;
	if (type == 0) return 0.0;
	if (type == 15) return 1.0;
	// walk the boundary counter-clockwise: n0, bottom, n1, right, n3, top, n2, left
	double px[8], py[8];
	int k = 0;
	auto add = [&](double x, double y) { px[k] = x; py[k] = y; k++; };
	if (i0) add(0, 0);
	if (i0 != i1) add(crossing(p0, p1), 0);
	if (i1) add(1, 0);
	if (i1 != i3) add(1, crossing(p1, p3));
	if (i3) add(1, 1);
	if (i3 != i2) add(crossing(p2, p3), 1);
	if (i2) add(0, 1);
	if (i2 != i0) add(0, crossing(p0, p2));
	double a = 0;
	for (int q = 0; q < k; q++) {
		const int w = q + 1 == k ? 0 : q + 1;
		a += px[q] * py[w] - px[w] * py[q];
	}
	a *= 0.5;
	// a diagonal cell whose centre is outside: the two corners are separate,
	// remove the quadrilateral of the four crossings that the walk included
	if ((type == 6 || type == 9) && !((p0 + p1 + p2 + p3) / 4 <= 0)) {
		const double bx = crossing(p0, p1), ry = crossing(p1, p3), tx = crossing(p2, p3), ly = crossing(p0, p2);
		const double qx[4] = {bx, 1, tx, 0}, qy[4] = {0, ry, 1, ly};
		double q = 0;
		for (int j = 0; j < 4; j++) {
			const int w = (j + 1) & 3;
			q += qx[j] * qy[w] - qx[w] * qy[j];
		}
		a -= 0.5 * q;
	}
	return a;
}

// ----------------------------------------------------------------- kernels
template <bool CORRECTED>
__global__ __launch_bounds__(kBlock) void kVolume3(GridDev g, const double *phi, double *partials)
{
	const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
	const size_t total = size_t(nx - 1) * (ny - 1) * (nz - 1);
	double acc[1] = {0};
	for (size_t q = blockIdx.x * size_t(blockDim.x) + threadIdx.x; q < total; q += size_t(gridDim.x) * blockDim.x) {
		const int i = int(q % (nx - 1)), j = int((q / (nx - 1)) % (ny - 1)), k = int(q / (size_t(nx - 1) * (ny - 1)));
		const size_t b = (size_t(k) * ny + j) * nx + i;
		const size_t sy = nx, sz = size_t(nx) * ny;
		// lexicographic stencil (GridBasedData::stencil<1>)
		const double l[8] = {phi[b], phi[b + 1], phi[b + sy], phi[b + sy + 1], phi[b + sz], phi[b + sz + 1], phi[b + sz + sy], phi[b + sz + sy + 1]};
		if (CORRECTED) {
			const double c[8] = {l[0], l[1], l[3], l[2], l[4], l[5], l[7], l[6]};
			acc[0] += mcVolumeFraction(c);
		} else {
			acc[0] += mcVolumeFraction(l);
		}
	}
	double blk[1];
	blockSum<1>(acc, blk);
	if (threadIdx.x == 0) partials[blockIdx.x] = blk[0];
}

__global__ __launch_bounds__(kBlock) void kArea2(GridDev g, const double *phi, double *partials)
{
	const int nx = g.n[0], ny = g.n[1];
	const size_t total = size_t(nx - 1) * (ny - 1);
	double acc[1] = {0};
	for (size_t q = blockIdx.x * size_t(blockDim.x) + threadIdx.x; q < total; q += size_t(gridDim.x) * blockDim.x) {
		const int i = int(q % (nx - 1)), j = int(q / (nx - 1));
		const size_t b = size_t(j) * nx + i;
		acc[0] += filledArea(phi[b], phi[b + 1], phi[b + nx], phi[b + nx + 1]);
	}
	double blk[1];
	blockSum<1>(acc, blk);
	if (threadIdx.x == 0) partials[blockIdx.x] = blk[0];
}

__global__ __launch_bounds__(kBlock) void kVolumeFinal(const double *partials, int nb, double scale, double *out)
{
	double v[1] = {0};
	for (int i = threadIdx.x; i < nb; i += blockDim.x) v[0] += partials[i];
	double o[1];
	blockSum<1>(v, o);
	if (threadIdx.x == 0) *out = o[0] * scale;
}

// Volume of phi on grid g (dim D) into *out: one pass writing kVolumeBlocks
// partials, then one workgroup sums them in a fixed order.
void launchVolumeInto(const GridDev &g, int D, const double *phi, bool corrected, double *partials, double *out, hipStream_t s)
{
	const int nb = kVolumeBlocks;
	if (D == 3) {
		if (corrected) hipLaunchKernelGGL(kVolume3<true>, dim3(nb), dim3(kBlock), 0, s, g, phi, partials);
		else hipLaunchKernelGGL(kVolume3<false>, dim3(nb), dim3(kBlock), 0, s, g, phi, partials);
		hipLaunchKernelGGL(kVolumeFinal, dim3(1), dim3(kBlock), 0, s, (const double *)partials, nb, g.dx * g.dx * g.dx, out);
	} else {
		hipLaunchKernelGGL(kArea2, dim3(nb), dim3(kBlock), 0, s, g, phi, partials);
		hipLaunchKernelGGL(kVolumeFinal, dim3(1), dim3(kBlock), 0, s, (const double *)partials, nb, g.dx * g.dx, out);
	}
}

void launchVolume(const Fields &F, bool corrected, hipStream_t s)
{
	launchVolumeInto(F.cell, F.D, F.phi_tmp, corrected, F.partials, F.dscalars + 3, s);
}

} // namespace garm

// ------------------------------------------- allocation-free volume entry
// garm_volume_device keeps its partial sums in this file-scope array, one
// copy per device, instead of memory of a context.  In host code `scratch`
// names a host-side stand-in, never the device copy: its device address is
// taken once per device with hipGetSymbolAddress and only that address is
// passed to kernels.
static __device__ __align__(256) double scratch[garm::kVolumeBlocks];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double *g_scratch_ptr[kMaxDevices] = {};
// Device address of this device's copy of `scratch` (the device must be
// current); nullptr with the HIP error in *e when the lookup fails.
double *scratchOfDevice(int dev, hipError_t *e)
{
	std::lock_guard<std::mutex> lock(g_scratch_mutex);
	*e = hipSuccess;
	if (!g_scratch_ptr[dev]) {
		void *ptr = nullptr;
		*e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
		if (*e != hipSuccess) return nullptr;
		g_scratch_ptr[dev] = static_cast<double *>(ptr);
	}
	return g_scratch_ptr[dev];
}

int deviceFail(int code, const std::string &msg)
{
	garm::setContextFreeError(msg);
	return code;
}

} // namespace

extern "C" int garm_volume_device(int dim, const int res[3], double dx, const double *phi_dev, int order, double *out_dev, void *stream)
{
	garm::setContextFreeError("");
	if ((dim != 2 && dim != 3) || !res || !phi_dev || !out_dev || !(dx > 0))
		return deviceFail(GARM_EINVAL, "garm_volume_device: invalid arguments");
	if (order != GARM_VOLUME_REFERENCE && order != GARM_VOLUME_CORRECTED) return deviceFail(GARM_EINVAL, "garm_volume_device: unknown vertex order");
	double cells = 1;
	for (int a = 0; a < dim; a++) {
		if (res[a] < 2 || res[a] > garm::kVolumeDeviceMaxRes)
			return deviceFail(GARM_EINVAL, "garm_volume_device: resolution outside [2, " + std::to_string(garm::kVolumeDeviceMaxRes) + "]");
		cells *= res[a];
	}
	if (cells >= 2147483647.0) return deviceFail(GARM_EINVAL, "garm_volume_device: grid holds 2^31 cells or more");
	int dev = -1;
	if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return deviceFail(GARM_ENODEV, "garm_volume_device: no usable device");
	hipError_t e;
	double *partials = scratchOfDevice(dev, &e);
	if (!partials) return deviceFail(GARM_EHIP, std::string("garm_volume_device: scratch address: ") + hipGetErrorString(e));
	garm::GridDev g{};
	for (int a = 0; a < 3; a++) { g.n[a] = a < dim ? res[a] : 1; g.o[a] = 0; }
	g.dx = dx;
	g.inv = 1 / dx;
	garm::launchVolumeInto(g, dim, phi_dev, order == GARM_VOLUME_CORRECTED, partials, out_dev, static_cast<hipStream_t>(stream));
	e = hipGetLastError();
	if (e != hipSuccess) return deviceFail(GARM_EHIP, std::string("garm_volume_device: launch: ") + hipGetErrorString(e));
	return GARM_OK;
}
