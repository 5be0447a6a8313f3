// Surface tension: the implicit capillary solve on the interface faces, the
// reprojection and the little-drop fix (LevelSetLiquid::applySurfaceTension
// and trickLittleDrop, LevelSetLiquid.cpp:1026-1057).
#pragma once

#include "garm_project.hpp"

#include <string>

namespace garm {

// Device storage of the capillary system over the D face grids taken as one
// index space (axis-major, each grid x-fastest: the order of
// forEach(velocity.grids), SGridBasedData.h).
struct Capillary {
	uint8_t *flags = nullptr;     // 1 on an interface face
	int *grid2mat = nullptr, *mat2grid = nullptr;
	int *tile_count = nullptr, *tile_base = nullptr;
	int tiles = 0;
	size_t faces = 0;             // size of the face index space
	Rows<double> rows{};
	double *rhs = nullptr;
	Krylov K;
	size_t cap = 0;

	~Capillary();
	void release();
};

struct CapillaryArgs {
	double coeff;      // surface tension / density
	double band;       // the dirac band half-width in cells (surfaceTensionDelta)
	double dt;
	double tolerance;
	int max_iterations;
	int check_interval;
};

// SurfaceTension::solve (SurfaceTension.h:25-101) on F.u with the cubic view
// of F.phi: numbers the interface faces, assembles the capillary rows and
// solves them with FP64 Jacobi PCG; the solution replaces F.u on those faces.
int runCapillary(Fields &F, Capillary &C, const CapillaryArgs &a, hipStream_t s, garm_pcg_stats *st, std::string &err);

// LevelSetLiquid::trickLittleDrop (LevelSetLiquid.cpp:1037-1057)
void launchLittleDrop(const Fields &F, double dt, hipStream_t s);

} // namespace garm
