// Flexible PCG preconditioned by the U-Net, with plain PCG as the fallback.
#pragma once

#include "garm_common.hpp"
#include "garm_pcg.hpp"
#include "garm_unet.hpp"
#include "garm_multigrid.hpp"
#include "../../include/garm_gpu.h"

#include <string>

namespace garm {

// FP64 vectors and scalars of the flexible solve.
struct LearnedWork {
	double *x = nullptr, *r = nullptr, *p = nullptr, *q = nullptr, *z = nullptr, *t = nullptr, *b2 = nullptr;
	double *u = nullptr, *w = nullptr;   // the network's correction and A times it
	double *pa = nullptr, *pb = nullptr, *ls = nullptr;
	int *box = nullptr;
	double *host = nullptr;   // pinned
	hipEvent_t ready = nullptr;   // marks a readback's arrival (the host polls it)
	size_t cap = 0;
	// the flexible solves' box layout: operator (diag, idiag, wx, wy, wz) and
	// Krylov vectors (x, r, p, q, z, t) over the box's cells, box cell per unknown
	double *bpool = nullptr;
	size_t bcap = 0;              // doubles
	unsigned *boxOf = nullptr;
	size_t ocap = 0;
	~LearnedWork();
	bool reserve(size_t n, std::string &err);
	bool reserveBox(size_t cells, size_t unknowns, std::string &err);
};

// Solves A x = b; the solution is left in W.x.  grid2mat maps the cell grid
// to the compact unknowns; L is the unknowns' bounding box (origin, sizes:
// 0 and 1 along the axes past D; the cell grid's sizes; mat2grid).  What
// the flexible iteration leaves is solved by solveJacobi (mixed) with K and
// BK.
int solveLearned(int D, const Rows<double> &A64, const Rows<float> &A32, const double *b, int n, const int *grid2mat, const BoxLayout &L,
	const garm_pcg_opts &o, Krylov &K, BoxKrylov &BK, LearnedWork &W, Unet &U, Multigrid &M, hipStream_t s, garm_pcg_stats *st,
	std::string &err);

// Solves A x = b by flexible CG with the multigrid V-cycle preconditioner
// (FP32 cycle, FP64 Krylov vectors and residual); the solution is left in W.x.
int solveMultigrid(int D, const Rows<double> &A64, const Rows<float> &A32, const double *b, int n, const BoxLayout &L,
	const garm_pcg_opts &o, Krylov &K, BoxKrylov &BK, LearnedWork &W, Multigrid &M, hipStream_t s, garm_pcg_stats *st, std::string &err);

// `steps` Adam steps of U on pairs made from the pressure system of a
// projection (operator rows A64, right-hand side rhs, solution): with
// probability operatorFraction an operator pair (A e, e) for a random smooth
// e on the unknowns, otherwise (rhs, solution) with a random sign; input and
// target each scaled to unit RMS over the unknowns (the solver scales the
// correction), masked squared error.  The network
// grid is the unknowns' bounding box with two cells of context.
int trainOnSystem(int D, Unet &U, const Rows<double> &A64, const double *rhs, const double *solution, int n, const int *grid2mat,
	const int *mat2grid, const GridDev &cell, LearnedWork &W, int steps, float lr, double operatorFraction, unsigned long long seed,
	float *lossHistory, hipStream_t s, std::string &err);

// The network behind a C-ABI handle and the device it was created on.
Unet *unetOf(garm_unet *net, int *device);

} // namespace garm
