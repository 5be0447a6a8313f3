// Krylov solvers for the ghost-fluid Poisson system on gfx950.
#pragma once

#include "garm_common.hpp"
#include "../../include/garm_gpu.h"

#include <string>

namespace garm {

// The Poisson operator of Projector::buildLinearSystem (Projector.cpp:30-68)
// as rows: (A p)_r = diag_r p_r - sum_o w[o]_r p[nb[o]_r], nb = -1 where the
// neighbour is not an unknown (its ghost-fluid weight is folded into diag).
template <typename T>
struct Rows {
	int *nb[6];
	T *w[6];
	T *diag, *invdiag;
};

// device scalar slots (double) and state slots (int)
enum {
	kScRz0 = 0, kScRz1 = 1, kScAlpha = 2, kScBeta = 3, kScRn2 = 4, kScThreshold = 5, kScBnorm2 = 6, kScTol = 7,
	kScOuterRn2 = 8, kScOuterThreshold = 9
};
enum { kStDone = 0, kStIter = 1, kStMaxIter = 2, kStOuterDone = 3, kStSweep = 4, kStCounter = 5 };

// Device-resident solver workspace; vectors sized for `cap` unknowns.
struct Krylov {
	size_t cap = 0;
	double *x = nullptr, *r = nullptr, *p = nullptr, *y = nullptr;    // FP64 vectors
	float *r32 = nullptr, *p32 = nullptr, *y32 = nullptr;   // FP32 vectors of the mixed solver
	double *pa = nullptr, *pb = nullptr;                     // per-workgroup partials
	double *sc = nullptr;
	int *state = nullptr;
	int *host_state = nullptr;   // pinned
	double *host_sc = nullptr;   // pinned
	hipEvent_t ready = nullptr;  // spinWait's event for the state readback

	~Krylov();
	void release();
	bool reserve(size_t n, bool mixed, std::string &err);
};

// FP64 Jacobi-preconditioned CG, Eigen's conjugate_gradient restated
// (zero initial guess, ||r||^2 < tol^2 ||b||^2, LinearSystem.h:23-39).
int solvePcg64(int D, const Rows<double> &A, const double *b, int n, const garm_pcg_opts &o, Krylov &K,
	hipStream_t s, garm_pcg_stats *st, std::string &err);

// FP32 Jacobi-PCG with the solution accumulated in FP64 and the residual
// replaced by the FP64 residual b - A x at every 1e3 reduction; every
// reduction stays FP64 and convergence is judged on the FP64 residual.
int solveMixed(int D, const Rows<double> &A64, const Rows<float> &A32, const double *b, int n, const garm_pcg_opts &o,
	Krylov &K, hipStream_t s, garm_pcg_stats *st, std::string &err);

// The bounding box of the unknowns: origin lo and sizes n in the cell grid
// of sizes gn; mat2grid gives each unknown's cell.
struct BoxLayout {
	int lo[3], n[3], gn[3];
	const int *mat2grid;
};

// Storage of the box solvers (grown as needed).
struct BoxKrylov {
	double *pool = nullptr;
	size_t cap = 0;            // doubles
	unsigned *boxOf = nullptr; // box cell of every unknown
	size_t ocap = 0;
	double *partials = nullptr;  // per-workgroup partials of the z-walking iteration
	int *tiles = nullptr;        // active column tiles, then their count and their box cells
	unsigned char *tileMask = nullptr;
	~BoxKrylov();
};

// Jacobi PCG on the box of the unknowns with the operator held per cell
// (diagonal and -x, -y, -z face weights), one launch per iteration: FP64
// (mixed = false, the plain path) or FP32 refined in FP64 (mixed = true, the
// same contract as solveMixed).  The solution lands in K.x (compact).
// solveBoxPcg's answer when the tiles holding unknowns cover too much of
// the box (the caller solves on compact rows instead)
constexpr int kBoxTooSparse = 1;
int solveBoxPcg(int D, bool mixed, const Rows<double> &A64, const Rows<float> &A32, const double *b, int n, const BoxLayout &L,
	const garm_pcg_opts &o, Krylov &K, BoxKrylov &B, hipStream_t s, garm_pcg_stats *st, std::string &err);

// The Jacobi PCG (FP64, or mixed) in the faster of its two forms: on the
// box of the unknowns while the box holds at most kBoxSolverRatio2 (2D) /
// kBoxSolverRatio3 (3D) cells per unknown and the box solver takes it (in
// 3D it launches only the column tiles holding unknowns and declines when
// those cover too much more than the unknowns), else on the compact rows.
// *usedBox tells which; the solution lands in K.x.
constexpr double kBoxSolverRatio2 = 1.5, kBoxSolverRatio3 = 16.0;
int solveJacobi(int D, bool mixed, const Rows<double> &A64, const Rows<float> &A32, const double *b, int n, const BoxLayout &L,
	const garm_pcg_opts &o, Krylov &K, BoxKrylov &B, hipStream_t s, garm_pcg_stats *st, std::string &err, bool *usedBox);

// Launch one solver kernel `reps` times for timing (which: 0 FP64 SpMV,
// 1 FP64 x/r update, 2 FP32 SpMV, 3 FP32 update); x and r are left unchanged.
int timeSolverKernel(int D, const Rows<double> &A64, const Rows<float> &A32, int n, int which, int reps, Krylov &K,
	hipStream_t s, hipEvent_t e0, hipEvent_t e1, float *ms, double *bytes);

} // namespace garm
