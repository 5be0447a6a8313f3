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
// ORACLE — test infrastructure only (tests/, smoke() and bench.py's checker).
//
// CPU restatement of the volume measurement of LevelSetLiquid::calculateVolume
// (LevelSetLiquid.cpp:1162-1190):
//  - 3D: Fraction::get_mc_vol (MC-style-vol-eval/fraction.hpp:56-61 edge
//    fraction, :149-161 table index, :163-188 corner rotations, :190-224 edge
//    points, :226-382 the 23 case volumes, :510-771 the 256-way dispatch);
//  - 2D: the area of Contourer<2, true>'s filled contour (Contourer.cpp:9-96,
//    tables MarchingCubesTables.inc), computed as the area of the inside
//    polygon of each cell (the contour is planar, so its triangles tile it).
// The dispatch and rotation tables are data that must equal the reference's
// for the volume to match; they are written out here from fraction.hpp.
#include <array>
#include <cstdint>

namespace mcv {

using V8 = std::array<double, 8>;

// fraction.hpp:56-61
static double frac(double a, double b)
{
	if (a < 0 && b < 0) return 1.0;
	if (a < 0 && 0 <= b) return a / (a - b);
	if (0 <= a && b < 0) return b / (b - a);
	return 0;
}

// fraction.hpp:510-771 as (case, rotation, variant): case | rot << 5 | variant << 10
static const uint16_t kTable[256] = {
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

// fraction.hpp:163-188: rotated[i] = v[kRot[r][i]]
static const uint8_t kRot[24][8] = {
	{0,1,2,3,4,5,6,7}, {4,5,1,0,7,6,2,3}, {7,6,5,4,3,2,1,0}, {3,2,6,7,0,1,5,4},
	{4,0,3,7,5,1,2,6}, {5,4,7,6,1,0,3,2}, {1,5,6,2,0,4,7,3}, {3,0,1,2,7,4,5,6},
	{2,3,0,1,6,7,4,5}, {1,2,3,0,5,6,7,4}, {0,4,5,1,3,7,6,2}, {0,3,7,4,1,2,6,5},
	{2,1,5,6,3,0,4,7}, {5,1,0,4,6,2,3,7}, {5,6,2,1,4,7,3,0}, {7,3,2,6,4,0,1,5},
	{2,6,7,3,1,5,4,0}, {7,4,0,3,6,5,1,2}, {1,0,4,5,2,3,7,6}, {3,7,4,0,2,6,5,1},
	{6,7,3,2,5,4,0,1}, {6,2,1,5,7,3,0,4}, {4,7,6,5,0,3,2,1}, {6,5,4,7,2,1,0,3}
};

struct Vec {
	double x, y, z;
};

static Vec sub(const Vec &a, const Vec &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static double dot(const Vec &a, const Vec &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static Vec cross(const Vec &a, const Vec &b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

// fraction.hpp:190-224: edge point, inside or outside variant
static Vec edge(const V8 &v, int e, bool in)
{
	auto f = [&](int a, int b) { const double t = frac(v[a], v[b]); return in ? t : 1.0 - t; };
	switch (e) {
	case 0: return {f(0, 1), 0, 0};
	case 1: return {1, f(1, 2), 0};
	case 2: return {f(2, 3), 1, 0};
	case 3: return {0, f(0, 3), 0};
	case 4: return {f(4, 5), 0, 1};
	case 5: return {1, f(5, 6), 1};
	case 6: return {f(6, 7), 1, 1};
	case 7: return {0, f(4, 7), 1};
	case 8: return {0, 0, f(0, 4)};
	case 9: return {1, 0, f(1, 5)};
	case 10: return {1, 1, f(2, 6)};
	default: return {0, 1, f(3, 7)};
	}
}

// signed volume term ((b - a) x (a - c)) . a
static double sv(const Vec &a, const Vec &b, const Vec &c) { return dot(cross(sub(b, a), sub(a, c)), a); }

static double p1(double a, double b, double c, double d) { return frac(a, b) * frac(a, c) * frac(a, d); }
static double p21(double a, double b, double c, double d) { return (1.0 - frac(a, b)) * (1.0 - frac(a, c)) * (1.0 - frac(a, d)); }
static double p2(double a, double b, double c, double d, double e, double f)
{
	const double e1 = frac(d, e);
	return (frac(a, b) + e1) * frac(b, c) + e1 * frac(e, f);
}

// fraction.hpp:226-382
static double caseVolume(int n, int m, const V8 &v)
{
	const double s = 1.0 / 6.0;
	switch (n) {
	case 0: return 0.0;
	case 1: return s * p1(v[0], v[1], v[3], v[4]);
	case 2: return s * p2(v[3], v[0], v[4], v[2], v[1], v[5]);
	case 3: return s * (p1(v[0], v[1], v[3], v[4]) + p1(v[5], v[1], v[6], v[4]));
	case 4: return s * (p1(v[0], v[1], v[3], v[4]) + p1(v[6], v[7], v[5], v[2]));
	case 5: {
		const double e3 = frac(v[0], v[3]);
		const double vol0 = frac(v[1], v[5]) * (1.0 - 0.5 * (1.0 - frac(v[0], v[1])) * (1.0 - e3)) / 3.0;
		const double e11 = frac(v[3], v[7]);
		return vol0 + s * ((frac(v[2], v[6]) + e11) + e3 * e11);
	}
	case 6: return s * (p2(v[3], v[0], v[4], v[2], v[1], v[5]) + p1(v[6], v[7], v[5], v[2]));
	case 7: return s * (p1(v[4], v[0], v[5], v[7]) + p1(v[1], v[0], v[2], v[5]) + p1(v[6], v[7], v[5], v[2]));
	case 8: return s * (2.0 * frac(v[0], v[4]) + frac(v[1], v[5]) + 2.0 * frac(v[2], v[6]) + frac(v[3], v[7]));
	case 9: {
		const Vec e0 = edge(v, 0, true), e1 = edge(v, 1, false), e6 = edge(v, 6, true), e7 = edge(v, 7, false), e8 = edge(v, 8, true),
			e10 = edge(v, 10, true);
		const double _e10 = frac(v[2], v[6]), _e6 = frac(v[6], v[7]);
		const double area = (1.0 - 0.5 * (1.0 - _e6) * (1.0 - _e10)) / 3.0 + s * frac(v[1], v[2]) * _e10 + s * (_e6 * frac(v[4], v[7]));
		if (m == 0) return area + s * (sv(e10, e7, e6) + sv(e1, e7, e10) + sv(e1, e8, e7) + sv(e1, e0, e8));
		return area + s * (sv(e10, e1, e0) + sv(e6, e10, e0) + sv(e6, e0, e8) + sv(e6, e8, e7));
	}
	case 10: return s * (p2(v[4], v[7], v[6], v[0], v[3], v[2]) + p2(v[2], v[1], v[0], v[6], v[5], v[4]));
	case 11: {
		const Vec e0 = edge(v, 0, true), e1 = edge(v, 1, false), e5 = edge(v, 5, false), e6 = edge(v, 6, false), e8 = edge(v, 8, true),
			e11 = edge(v, 11, true);
		const double _e5 = frac(v[5], v[6]), _e6 = frac(v[6], v[7]);
		const double area = s * (frac(v[1], v[2]) + _e5) + (1.0 - 0.5 * (1.0 - _e6) * (1.0 - frac(v[3], v[7]))) / 3.0 + s * (_e5 * _e6);
		return area + s * (sv(e0, e8, e11) + sv(e0, e11, e5) + sv(e0, e5, e1) + sv(e5, e11, e6));
	}
	case 12: {
		const double e0 = frac(v[0], v[1]), e9 = frac(v[1], v[5]);
		return frac(v[3], v[7]) * (1.0 - 0.5 * (1.0 - e0) * (1.0 - frac(v[0], v[3]))) / 3.0 +
			s * ((e9 + frac(v[2], v[6])) + e0 * e9 + p1(v[4], v[0], v[5], v[7]));
	}
	case 13: return s * (p1(v[4], v[0], v[5], v[7]) + p1(v[1], v[0], v[2], v[5]) + p1(v[6], v[7], v[5], v[2]) + p1(v[3], v[0], v[2], v[7]));
	case 14: {
		const Vec e0 = edge(v, 0, false), e3 = edge(v, 3, false), e6 = edge(v, 6, true), e7 = edge(v, 7, false), e9 = edge(v, 9, true),
			e10 = edge(v, 10, true);
		const double _e10 = frac(v[2], v[6]), _e6 = frac(v[6], v[7]);
		const double area = s * (frac(v[1], v[5]) + _e10) + s * (_e6 * frac(v[4], v[7])) + (1.0 - 0.5 * (1.0 - _e6) * (1.0 - _e10)) / 3.0;
		return area + s * (sv(e0, e3, e7) + sv(e0, e7, e10) + sv(e0, e10, e9) + sv(e6, e10, e7));
	}
	case 15: {
		const Vec e0 = edge(v, 0, true), e1 = edge(v, 1, false), e6 = edge(v, 6, true), e7 = edge(v, 7, false), e8 = edge(v, 8, true),
			e10 = edge(v, 10, true);
		const double _e10 = frac(v[2], v[6]), _e6 = frac(v[6], v[7]);
		return (1.0 - 0.5 * (1.0 - _e6) * (1.0 - _e10)) / 3.0 + s * frac(v[1], v[2]) * _e10 + s * (_e6 * frac(v[4], v[7])) +
			s * (sv(e1, e6, e10) + sv(e1, e7, e6) + sv(e1, e0, e7) + sv(e8, e7, e0)) + s * p1(v[5], v[4], v[1], v[6]);
	}
	case 16: {
		const Vec e1 = edge(v, 1, false), e3 = edge(v, 3, false), e5 = edge(v, 5, true), e6 = edge(v, 6, true), e8 = edge(v, 8, false),
			e9 = edge(v, 9, false), e10 = edge(v, 10, true);
		const double _e10 = frac(v[2], v[6]), _e5 = frac(v[5], v[6]), _e6 = frac(v[6], v[7]);
		return s * (sv(e1, e3, e6) + sv(e1, e6, e10) + sv(e3, e8, e6) + sv(e5, e6, e9) + sv(e8, e9, e6)) +
			s * (frac(v[1], v[2]) * _e10 + _e5 * frac(v[1], v[5])) + (1.0 - 0.5 * (1.0 - _e6) * (1.0 - _e10)) / 3.0 +
			(1.0 - 0.5 * (1.0 - _e6) * (1.0 - _e5)) / 3.0;
	}
	case 17: {
		const double e3 = frac(v[0], v[3]), e11 = frac(v[3], v[7]);
		return 1.0 - ((1.0 - frac(v[1], v[5])) * (1.0 - 0.5 * frac(v[0], v[1]) * e3) / 3.0 +
			s * ((1.0 - frac(v[2], v[6])) + (1.0 - e11) + (1.0 - e3) * (1.0 - e11)));
	}
	case 18: return 1.0 - s * (p21(v[0], v[1], v[3], v[4]) + p21(v[6], v[5], v[2], v[7]));
	case 19: {
		const Vec e0 = edge(v, 0, false), e3 = edge(v, 3, false), e4 = edge(v, 4, true), e5 = edge(v, 5, false), e8 = edge(v, 8, false),
			e9 = edge(v, 9, true);
		const double _e5 = frac(v[5], v[6]);
		return 1.0 - (s * (-sv(e8, e4, e5) - sv(e8, e5, e3) - sv(e9, e0, e5) - sv(e0, e3, e5)) +
			(0.5 * (1.0 - frac(v[1], v[5])) * (1.0 - _e5)) / 3.0 + (0.5 * (1.0 - _e5) * (1.0 - frac(v[4], v[5]))) / 3.0);
	}
	case 20: {
		const double e1 = frac(v[1], v[2]);
		return 1.0 - s * (((1.0 - frac(v[0], v[3])) + (1.0 - e1)) * (1.0 - frac(v[0], v[4])) + (1.0 - e1) * (1.0 - frac(v[1], v[5])));
	}
	case 21: return 1.0 - s * p21(v[0], v[1], v[3], v[4]);
	case 22: return 1.0;
	default: return 0.0;
	}
}

} // namespace mcv

extern "C" {

// Fraction::get_mc_vol of 8 values in the fraction's circular corner order.
double ora_mc_fraction(const double *c)
{
	int idx = 0;
	for (int i = 0; i < 8; i++)
		if (c[i] < 0.0) idx |= 1 << i;
	const int t = mcv::kTable[idx];
	const int r = (t >> 5) & 31;
	mcv::V8 v;
	for (int i = 0; i < 8; i++) v[i] = c[mcv::kRot[r][i]];
	return mcv::caseVolume(t & 31, (t >> 10) & 1, v);
}

// Area of one cell of the filled contour; nodes n0 (0,0) n1 (1,0) n2 (0,1)
// n3 (1,1) (Contourer.cpp:74-91 cell type with the centre rule for 6 and 9).
double ora_filled_area(const double *p)
{
	const bool in[4] = {p[0] <= 0, p[1] <= 0, p[2] <= 0, p[3] <= 0};
	// counter-clockwise corners n0, n1, n3, n2; edge e joins ring[e] and ring[e + 1]
	const int ring[4] = {0, 1, 3, 2};
	double x[8], y[8];
	bool crossingPt[8];
	const double cx[4] = {0, 1, 0, 1}, cy[4] = {0, 0, 1, 1};
	int k = 0;
	for (int e = 0; e < 4; e++) {
		const int a = ring[e], b = ring[(e + 1) & 3];
		if (in[a]) { x[k] = cx[a]; y[k] = cy[a]; crossingPt[k] = false; k++; }
		if (in[a] != in[b]) {
			// crossing measured from the edge's lower node (Contourer.cpp:40-43)
			const int lo = a < b ? a : b, hi = a < b ? b : a;
			const double th = p[lo] / (p[lo] - p[hi]);
			x[k] = cx[lo] + th * (cx[hi] - cx[lo]);
			y[k] = cy[lo] + th * (cy[hi] - cy[lo]);
			crossingPt[k] = true;
			k++;
		}
	}
	double area = 0;
	for (int q = 0; q < k; q++) area += x[q] * y[(q + 1) % k] - x[(q + 1) % k] * y[q];
	area *= 0.5;
	const int type = in[0] | in[1] << 1 | in[2] << 2 | in[3] << 3;
	if ((type == 6 || type == 9) && !((p[0] + p[1] + p[2] + p[3]) / 4 <= 0)) {
		// separate corners: drop the middle quadrilateral of the four crossings
		double qx[4], qy[4];
		int j = 0;
		for (int q = 0; q < k; q++)
			if (crossingPt[q]) { qx[j] = x[q]; qy[j] = y[q]; j++; }
		double mid = 0;
		for (int q = 0; q < 4; q++) mid += qx[q] * qy[(q + 1) & 3] - qx[(q + 1) & 3] * qy[q];
		area -= 0.5 * mid;
	}
	return area;
}

} // extern "C"
