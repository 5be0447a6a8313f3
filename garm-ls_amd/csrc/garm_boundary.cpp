// Host-side construction of the solid boundary for a domain-box scene:
// the Boundary constructor and reset (Boundary.cpp:7-35) with the domain box
// ImplicitBox(domainOrigin, domainLengths) (ImplicitSurface.h:26-59,
// StaggeredGrid.h:40-41), then finish (Boundary.cpp:79-87) with
// getFaceFraction (Boundary.cpp:135-174).  Runs once per context on the CPU;
// the arrays are then uploaded and never rebuilt (no moving boundary).
#include "garm_boundary.hpp"

#include <cstddef>

#include <algorithm>
#include <cmath>
#include <utility>

namespace garm {

namespace {

struct Box {
	int D;
	double c[3], h[3];

	// ImplicitBox::signedDistance (ImplicitSurface.h:51-56)
	double sdf(const double *p) const
	{
		double ph[3] = {0, 0, 0};
		bool in = true;
		for (int a = 0; a < D; a++) { ph[a] = std::abs(p[a] - c[a]) - h[a]; if (!(ph[a] <= 0)) in = false; }
		if (in) { double m = ph[0]; for (int a = 1; a < D; a++) m = std::max(m, ph[a]); return m; }
		double s = 0;
		for (int a = 0; a < D; a++) { const double m = std::max(ph[a], 0.0); s += m * m; }
		return std::sqrt(s);
	}

	// ImplicitBox::closestNormal (ImplicitSurface.h:39-49); the first maximal
	// axis wins a tie and a zero vector stays zero when normalised.
	void normal(const double *p, double *n) const
	{
		double ph[3] = {0, 0, 0};
		bool in = true;
		for (int a = 0; a < D; a++) { ph[a] = std::abs(p[a] - c[a]) - h[a]; if (!(ph[a] <= 0)) in = false; }
		if (in) {
			int ax = 0;
			for (int a = 1; a < D; a++) if (ph[a] > ph[ax]) ax = a;
			for (int a = 0; a < D; a++) n[a] = a == ax ? 1.0 : 0.0;
		} else {
			for (int a = 0; a < D; a++) n[a] = std::max(ph[a], 0.0);
		}
		double s = 0;
		for (int a = 0; a < D; a++) {
			const double d = p[a] - c[a];
			n[a] *= d > 0 ? 1.0 : (d < 0 ? -1.0 : 0.0);
			s += n[a] * n[a];
		}
		if (s > 0) { const double r = std::sqrt(s); for (int a = 0; a < D; a++) n[a] /= r; }
	}
};

struct HostGrid {
	int D, n[3];
	double o[3], dx;
	size_t count() const { return size_t(n[0]) * n[1] * (D == 3 ? n[2] : 1); }
	size_t index(const int *c) const { return D == 2 ? c[0] + size_t(n[0]) * c[1] : c[0] + n[0] * (c[1] + size_t(n[1]) * c[2]); }
	void position(const int *c, double *p) const { for (int a = 0; a < D; a++) p[a] = o[a] + c[a] * dx; }
	template <typename F>
	void each(F f) const
	{
		int c[3] = {0, 0, 0};
		for (c[2] = 0; c[2] < (D == 3 ? n[2] : 1); c[2]++)
			for (c[1] = 0; c[1] < n[1]; c[1]++)
				for (c[0] = 0; c[0] < n[0]; c[0]++) f(c);
	}
};

} // namespace

void buildDomainBoundary(int D, const int *res, double dx, const double *origin, int bw,
	double *nodeDist, double *cellDist, double *frac, double *normal)
{
	Box box{D, {0, 0, 0}, {0, 0, 0}};
	for (int a = 0; a < D; a++) {
		const double dorg = origin[a] + bw * dx;
		const double dlen = double(res[a] - bw * 2) * dx;
		box.c[a] = dorg + dlen / 2;
		box.h[a] = dlen / 2;
	}
	HostGrid node{D, {1, 1, 1}, {0, 0, 0}, dx}, cell{D, {1, 1, 1}, {0, 0, 0}, dx}, face[3];
	for (int a = 0; a < 3; a++) {
		node.n[a] = a < D ? res[a] + 1 : 1;
		node.o[a] = a < D ? origin[a] : 0;
		cell.n[a] = a < D ? res[a] : 1;
		cell.o[a] = a < D ? origin[a] + dx / 2 : 0;
	}
	for (int ax = 0; ax < D; ax++) {
		face[ax] = HostGrid{D, {1, 1, 1}, {0, 0, 0}, dx};
		for (int a = 0; a < 3; a++) {
			face[ax].n[a] = a < D ? res[a] + (a == ax ? 1 : 0) : 1;
			face[ax].o[a] = a < D ? origin[a] + (a != ax ? dx / 2 : 0) : 0;
		}
	}
	node.each([&](const int *c) { double p[3]; node.position(c, p); nodeDist[node.index(c)] = -box.sdf(p); });
	cell.each([&](const int *c) { double p[3]; cell.position(c, p); cellDist[cell.index(c)] = -box.sdf(p); });

	size_t off = 0;
	for (int ax = 0; ax < D; ax++) {
		const HostGrid &g = face[ax];
		g.each([&](const int *c) {
			double p[3], n[3];
			g.position(c, p);
			box.normal(p, n);
			normal[off + g.index(c)] = -n[ax];
			// StaggeredGrid::isBoundaryFace (StaggeredGrid.h:57)
			bool inside = true;
			for (int a = 0; a < D; a++) if (c[a] < bw || c[a] >= g.n[a] - bw) inside = false;
			const bool boundary = c[ax] <= bw || c[ax] >= res[ax] - bw || !inside;
			double ff;
			if (boundary) ff = 1;
			else {
				auto theta = [](double a, double b) { return a / (a - b); };
				if (D == 2) {
					int q1[3] = {c[0], c[1], 0};
					q1[ax ^ 1] += 1;
					const double p0 = nodeDist[node.index(c)], p1 = nodeDist[node.index(q1)];
					if (p0 <= 0 && p1 <= 0) ff = 1;
					else if (p0 <= 0 && p1 > 0) ff = theta(p0, p1);
					else if (p0 > 0 && p1 <= 0) ff = theta(p1, p0);
					else ff = 0;
				} else {
					auto tri = [&](double a0, double a1, double a2) -> double {
						if (a0 > a1) std::swap(a0, a1);
						if (a1 > a2) std::swap(a1, a2);
						if (a0 > a1) std::swap(a0, a1);
						if (a2 <= 0) return 1;
						if (a1 <= 0) return 1 - theta(a2, a0) * theta(a2, a1);
						if (a0 <= 0) return theta(a0, a1) * theta(a0, a2);
						return 0;
					};
					double ph[4];
					for (int ord = 0; ord < 4; ord++) {
						int q[3] = {c[0], c[1], c[2]};
						q[(ax + 1) % 3] += ord & 1;
						q[(ax + 2) % 3] += ord >> 1 & 1;
						ph[ord] = nodeDist[node.index(q)];
					}
					const double cp = (ph[0] + ph[1] + ph[2] + ph[3]) * .25;
					ff = (tri(cp, ph[0], ph[1]) + tri(cp, ph[1], ph[3]) + tri(cp, ph[3], ph[2]) + tri(cp, ph[2], ph[0])) * .25;
				}
				ff = ff > .9 ? 1 : ff;
			}
			frac[off + g.index(c)] = ff;
		});
		off += g.count();
	}
}

} // namespace garm
