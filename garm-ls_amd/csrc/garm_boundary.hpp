// Host-side solid-boundary construction for a domain-box scene.
#pragma once

namespace garm {

// Boundary ctor + reset + finish (Boundary.cpp:7-35, 79-87, 135-174) for the
// domain box of a StaggeredGrid with `res` cells, spacing dx, lower corner
// `origin` and boundary width bw.  Face arrays are concatenated per axis.
void buildDomainBoundary(int D, const int *res, double dx, const double *origin, int bw,
	double *nodeDist, double *cellDist, double *frac, double *normal);

} // namespace garm
