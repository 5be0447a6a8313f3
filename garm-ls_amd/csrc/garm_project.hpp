// Pressure projection state (unknown maps, operator rows, PCG vectors).
#pragma once

#include "garm_internal.hpp"
#include "../../include/garm_gpu.h"

#include <string>

namespace garm {

// Three face arrays with their grids, passed by value.
struct FaceFracs {
	const double *a[3];
	GridDev g[3];
};

// The Poisson operator of Projector::buildLinearSystem (Projector.cpp:30-68)
// as rows: A p = diag p - sum_o w[o] p[nb[o]], nb = -1 where the neighbour is
// not an unknown (its ghost-fluid weight is folded into diag).
struct PoissonRows {
	int *nb[6];
	double *w[6];
	double *diag, *invdiag, *rhs;
};

// device scalar slots
enum { kScAbsNew = 0, kScAlpha = 1, kScBeta = 2, kScRn2 = 3, kScThreshold = 4, kScBnorm2 = 5, kScTol = 6 };
enum { kStDone = 0, kStIter = 1, kStMaxIter = 2 };

struct Projection {
	int *flags = nullptr, *grid2mat = nullptr, *mat2grid = nullptr;
	void *scan_tmp = nullptr;
	size_t scan_bytes = 0;
	PoissonRows rows{};
	double *x = nullptr, *r = nullptr, *p = nullptr, *y = nullptr, *partials = nullptr;
	double *sc = nullptr;
	int *state = nullptr;
	int *host_state = nullptr;
	size_t cap = 0;
	int n = 0;

	~Projection();
	void release();
	bool reserveGrid(size_t N, std::string &err);
	bool reserveRows(int D, size_t n, std::string &err);
};

int runProjection(Fields &F, Projection &P, const garm_pcg_opts &o, garm_pcg_stats *st, hipStream_t s, std::string &err);
void scatterPressure(const Fields &F, const Projection &P, double *out, hipStream_t s);
int timePcgKernel(int D, Projection &P, int which, int reps, hipStream_t s, hipEvent_t e0, hipEvent_t e1, float *ms, double *bytes);

} // namespace garm
