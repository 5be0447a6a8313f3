// Krylov solvers for the ghost-fluid Poisson system on gfx950.
//
// One conjugate-gradient iteration is three launches:
//   kSpmvAlpha   y = A p and p.y; the last workgroup to finish reduces the
//                partials in a fixed order and stores alpha
//   kUpdateBeta  x += alpha p, r -= alpha y, |r|^2 and r.M^-1 r; the last
//                workgroup tests convergence and stores beta
//   kUpdateP     p = M^-1 r + beta p
// Scalars never leave the device; the host polls a flag every few
// iterations.  Every dot product and norm is accumulated in FP64 whatever
// the vector type, and the final reduction order is fixed, so repeated
// runs give identical bits.
#include "garm_pcg.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>

namespace garm {

// ------------------------------------------------------------ reductions
__device__ __forceinline__ void storeAgent(double *p, double v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ double loadAgent(const double *p) { return __hip_atomic_load(const_cast<double *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// Block partials -> partials[NV * block]; the last block to arrive sums all
// partials in block order and returns true (totals valid in thread 0).
template <int NV>
__device__ bool lastBlockSum(double (&v)[NV], double *partials, int *counter, double (&tot)[NV])
{
	double blk[NV];
	blockSum<NV>(v, blk);
	__shared__ bool amLast;
	const int tid = threadIdx.x;
	if (tid == 0) {
#pragma unroll
		for (int q = 0; q < NV; q++) storeAgent(&partials[NV * blockIdx.x + q], blk[q]);
		__threadfence();
		const int prev = atomicAdd(counter, 1);
		amLast = prev == int(gridDim.x) - 1;
	}
	__syncthreads();
	if (!amLast) return false;
	__threadfence();
	double s[NV];
#pragma unroll
	for (int q = 0; q < NV; q++) s[q] = 0;
	for (int i = tid; i < int(gridDim.x); i += blockDim.x)
#pragma unroll
		for (int q = 0; q < NV; q++) s[q] += loadAgent(&partials[NV * i + q]);
	blockSum<NV>(s, tot);
	if (tid == 0) *counter = 0;
	return true;
}

// ---------------------------------------------------------------- kernels
// y = A p, then alpha = (r.z) / (p.y)
template <int D, typename T>
__global__ __launch_bounds__(kBlock) void kSpmvAlpha(Rows<T> A, int n, const T *p, T *y, double *partials, double *sc, int *st)
{
	if (st[kStDone]) return;
	const int r = int(blockIdx.x * blockDim.x + threadIdx.x);
	double v[1] = {0};
	if (r < n) {
		const T pr = p[r];
		T s = A.diag[r] * pr;
#pragma unroll
		for (int o = 0; o < 2 * D; o++) {
			const int c = A.nb[o][r];
			if (c >= 0) s += -A.w[o][r] * p[c];
		}
		y[r] = s;
		v[0] = double(pr) * double(s);
	}
	double tot[1];
	if (lastBlockSum<1>(v, partials, st + kStCounter, tot) && threadIdx.x == 0) sc[kScAlpha] = sc[kScAbsNew] / tot[0];
}

// x += alpha p; r -= alpha y; convergence test |r|^2 < threshold, then
// beta = (r.z)_new / (r.z)_old and the iteration count (Eigen's `i`).
template <typename T, typename TX>
__global__ __launch_bounds__(kBlock) void kUpdateBeta(int n, TX *x, T *r, const T *p, const T *y, const T *invdiag,
	double *partials, double *sc, int *st)
{
	if (st[kStDone]) return;
	const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
	double v[2] = {0, 0};
	if (i < n) {
		const T alpha = T(sc[kScAlpha]);
		x[i] += alpha * p[i];
		const T ri = r[i] - alpha * y[i];
		r[i] = ri;
		v[0] = double(ri) * double(ri);
		v[1] = double(ri) * double(invdiag[i] * ri);
	}
	double tot[2];
	if (lastBlockSum<2>(v, partials, st + kStCounter, tot) && threadIdx.x == 0) {
		sc[kScRn2] = tot[0];
		if (tot[0] < sc[kScThreshold]) { st[kStDone] = 1; return; }
		const double absOld = sc[kScAbsNew];
		sc[kScAbsNew] = tot[1];
		sc[kScBeta] = tot[1] / absOld;
		const int it = st[kStIter] + 1;
		st[kStIter] = it;
		if (it >= st[kStMaxIter]) st[kStDone] = 1;
	}
}

// p = M^-1 r + beta p
template <typename T>
__global__ void kUpdateP(int n, T *p, const T *r, const T *invdiag, const double *sc, const int *st)
{
	if (st[kStDone]) return;
	const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
	if (i < n) p[i] = invdiag[i] * r[i] + T(sc[kScBeta]) * p[i];
}

// x = 0, r = b, p = M^-1 b; start-up tests of Eigen's conjugate_gradient with
// the tolerance floor of LinearSystem::solve (LinearSystem.h:31).
__global__ __launch_bounds__(kBlock) void kStart64(int n, const double *b, const double *invdiag, double *x, double *r, double *p,
	double tol, int maxIt, double *partials, double *sc, int *st)
{
	const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
	double v[2] = {0, 0};
	if (i < n) {
		const double bi = b[i];
		x[i] = 0;
		r[i] = bi;
		const double z = invdiag[i] * bi;
		p[i] = z;
		v[0] = bi * bi;
		v[1] = bi * z;
	}
	double tot[2];
	if (lastBlockSum<2>(v, partials, st + kStCounter, tot) && threadIdx.x == 0) {
		const double b2 = tot[0];
		sc[kScBnorm2] = b2;
		sc[kScRn2] = b2;
		sc[kScAbsNew] = tot[1];
		st[kStIter] = 0;
		st[kStMaxIter] = maxIt;
		st[kStDone] = 0;
		if (b2 == 0) { st[kStDone] = 1; sc[kScThreshold] = 0; sc[kScTol] = tol; return; }
		const double t = fmax(tol, DBL_EPSILON / sqrt(b2));
		sc[kScTol] = t;
		sc[kScThreshold] = fmax(t * t * b2, DBL_MIN);
		if (b2 < sc[kScThreshold] || maxIt <= 0) st[kStDone] = 1;
	}
}

// Outer step of iterative refinement: r = b - A x in FP64; the FP32 inner
// solve restarts on r32 = r with e = 0, p = M^-1 r32.  The last block tests
// the FP64 residual against the outer threshold and sets up the inner solve.
template <int D>
__global__ __launch_bounds__(kBlock) void kResidualRefine(Rows<double> A, Rows<float> A32, int n, const double *b, const double *x,
	double *r64, float *r32, float *p32, float *e32, double tol, double innerTol, int innerMax, double *partials, double *sc, int *st)
{
	const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
	double v[3] = {0, 0, 0};
	if (i < n) {
		double ax = A.diag[i] * x[i];
#pragma unroll
		for (int o = 0; o < 2 * D; o++) {
			const int c = A.nb[o][i];
			if (c >= 0) ax += -A.w[o][i] * x[c];
		}
		const double ri = b[i] - ax;
		r64[i] = ri;
		const float rf = float(ri);
		r32[i] = rf;
		const float z = A32.invdiag[i] * rf;
		p32[i] = z;
		e32[i] = 0.0f;
		v[0] = ri * ri;
		v[1] = double(rf) * double(rf);
		v[2] = double(rf) * double(z);
	}
	double tot[3];
	if (lastBlockSum<3>(v, partials, st + kStCounter, tot) && threadIdx.x == 0) {
		const int sweep = st[kStSweep];
		if (sweep == 0) {
			const double b2 = tot[0];   // x = 0 on the first sweep
			sc[kScBnorm2] = b2;
			const double t = b2 > 0 ? fmax(tol, DBL_EPSILON / sqrt(b2)) : tol;
			sc[kScTol] = t;
			sc[kScOuterThreshold] = fmax(t * t * b2, DBL_MIN);
		}
		sc[kScOuterRn2] = tot[0];
		st[kStSweep] = sweep + 1;
		const bool done = sc[kScBnorm2] == 0 || tot[0] < sc[kScOuterThreshold];
		st[kStOuterDone] = done ? 1 : 0;
		st[kStDone] = done ? 1 : 0;
		st[kStIter] = 0;
		st[kStMaxIter] = innerMax;
		sc[kScAbsNew] = tot[2];
		sc[kScThreshold] = fmax(innerTol * innerTol * tot[1], DBL_MIN);
	}
}

__global__ void kApplyCorrection(int n, double *x, const float *e, const int *st)
{
	if (st[kStOuterDone]) return;
	const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
	if (i < n) x[i] += double(e[i]);
}

__global__ void kSetInt(int *p, int v) { *p = v; }

__global__ void kZero(double *p, int n)
{
	const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
	if (i < n) p[i] = 0;
}

// ----------------------------------------------------------------- driver
Krylov::~Krylov() { release(); }

void Krylov::release()
{
	auto fr = [](auto *&q) { if (q) { (void)hipFree(q); q = nullptr; } };
	fr(x); fr(r); fr(p); fr(y); fr(r32); fr(p32); fr(y32); fr(e32); fr(partials); fr(sc); fr(state);
	if (host_state) { (void)hipHostFree(host_state); host_state = nullptr; }
	if (host_sc) { (void)hipHostFree(host_sc); host_sc = nullptr; }
	cap = 0;
}

bool Krylov::reserve(size_t n, bool mixed, std::string &err)
{
	bool ok = true;
	if (!sc) {
		ok = hipMalloc(&sc, 32 * sizeof(double)) == hipSuccess && hipMalloc(&state, 16 * sizeof(int)) == hipSuccess
			&& hipHostMalloc(&host_state, 16 * sizeof(int)) == hipSuccess && hipHostMalloc(&host_sc, 32 * sizeof(double)) == hipSuccess
			&& hipMemset(state, 0, 16 * sizeof(int)) == hipSuccess && hipMemset(sc, 0, 32 * sizeof(double)) == hipSuccess;
		if (!ok) { err = "solver scalar allocation failed"; return false; }
	}
	if (n > cap || !x) {
		auto fr = [](auto *&q) { if (q) { (void)hipFree(q); q = nullptr; } };
		fr(x); fr(r); fr(p); fr(y); fr(r32); fr(p32); fr(y32); fr(e32); fr(partials);
		const size_t want = std::max<size_t>(n + n / 4, 1024);
		for (double **q : {&x, &r, &p, &y}) ok = ok && hipMalloc(q, want * sizeof(double)) == hipSuccess;
		ok = ok && hipMalloc(&partials, 4 * ((want + kBlock - 1) / kBlock + 1) * sizeof(double)) == hipSuccess;
		if (!ok) { err = "solver vector allocation failed"; return false; }
		cap = want;
	}
	if (mixed && !r32) {
		for (float **q : {&r32, &p32, &y32, &e32}) ok = ok && hipMalloc(q, cap * sizeof(float)) == hipSuccess;
		if (!ok) { err = "FP32 solver vector allocation failed"; return false; }
	}
	return true;
}

static int readState(Krylov &K, hipStream_t s, std::string &err)
{
	if (hipMemcpyAsync(K.host_state, K.state, 8 * sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
		hipMemcpyAsync(K.host_sc, K.sc, 16 * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
		hipStreamSynchronize(s) != hipSuccess) { err = "solver state readback failed"; return GARM_EHIP; }
	return GARM_OK;
}

// Enqueue CG iterations in chunks of `check` until the device flag is set or
// `budget` iterations were enqueued.  Returns the iterations enqueued.
template <int D, typename T, typename TX>
static int iterate(const Rows<T> &A, int n, TX *x, T *r, T *p, T *y, const T *invdiag, Krylov &K, int budget, int check,
	hipStream_t s, std::string &err, int &rc)
{
	const int nb = (n + kBlock - 1) / kBlock;
	const dim3 g1(nb), b1(kBlock);
	int launched = 0;
	rc = GARM_OK;
	for (;;) {
		if ((rc = readState(K, s, err))) return launched;
		if (K.host_state[kStDone] || launched >= budget) return launched;
		const int k = std::min(check, budget - launched);
		for (int q = 0; q < k; q++) {
			hipLaunchKernelGGL((kSpmvAlpha<D, T>), g1, b1, 0, s, A, n, (const T *)p, y, K.partials, K.sc, K.state);
			hipLaunchKernelGGL((kUpdateBeta<T, TX>), g1, b1, 0, s, n, x, r, (const T *)p, (const T *)y, invdiag, K.partials, K.sc, K.state);
			hipLaunchKernelGGL((kUpdateP<T>), g1, b1, 0, s, n, p, (const T *)r, invdiag, (const double *)K.sc, (const int *)K.state);
		}
		launched += k;
	}
}

template <int D>
static int pcg64(const Rows<double> &A, const double *b, int n, const garm_pcg_opts &o, Krylov &K, hipStream_t s,
	garm_pcg_stats *st, std::string &err)
{
	const int nb = (n + kBlock - 1) / kBlock;
	hipLaunchKernelGGL(kSetInt, dim3(1), dim3(1), 0, s, K.state + kStCounter, 0);
	hipLaunchKernelGGL(kStart64, dim3(nb), dim3(kBlock), 0, s, n, b, (const double *)A.invdiag, K.x, K.r, K.p, o.tolerance,
		o.max_iterations, K.partials, K.sc, K.state);
	int rc;
	iterate<D, double, double>(A, n, K.x, K.r, K.p, K.y, A.invdiag, K, o.max_iterations, std::max(1, o.check_interval), s, err, rc);
	if (rc) return rc;
	if ((rc = readState(K, s, err))) return rc;
	st->iterations = K.host_state[kStIter];
	st->outer_iterations = 1;
	const double b2 = K.host_sc[kScBnorm2];
	st->relative_residual = b2 > 0 ? std::sqrt(K.host_sc[kScRn2] / b2) : 0.0;
	if (st->relative_residual > K.host_sc[kScTol]) { err = "PCG did not reach the requested tolerance"; return GARM_ENOCONV; }
	return GARM_OK;
}

template <int D>
static int mixed(const Rows<double> &A64, const Rows<float> &A32, const double *b, int n, const garm_pcg_opts &o, Krylov &K,
	hipStream_t s, garm_pcg_stats *st, std::string &err)
{
	const int nb = (n + kBlock - 1) / kBlock;
	const double innerTol = o.inner_tolerance > 0 ? o.inner_tolerance : 1e-3;
	const int check = std::max(1, o.check_interval);
	const int maxSweeps = 50;
	hipLaunchKernelGGL(kSetInt, dim3(1), dim3(1), 0, s, K.state + kStCounter, 0);
	hipLaunchKernelGGL(kSetInt, dim3(1), dim3(1), 0, s, K.state + kStSweep, 0);
	hipLaunchKernelGGL(kZero, dim3(nb), dim3(kBlock), 0, s, K.x, n);
	int total = 0, sweeps = 0, rc = GARM_OK;
	for (;;) {
		hipLaunchKernelGGL(kResidualRefine<D>, dim3(nb), dim3(kBlock), 0, s, A64, A32, n, b, (const double *)K.x, K.r, K.r32, K.p32, K.e32,
			o.tolerance, innerTol, o.max_iterations, K.partials, K.sc, K.state);
		if ((rc = readState(K, s, err))) return rc;
		if (K.host_state[kStOuterDone] || sweeps >= maxSweeps || total >= o.max_iterations) break;
		iterate<D, float, float>(A32, n, K.e32, K.r32, K.p32, K.y32, A32.invdiag, K, o.max_iterations - total, check, s, err, rc);
		if (rc) return rc;
		total += K.host_state[kStIter] + 1;
		sweeps++;
		hipLaunchKernelGGL(kApplyCorrection, dim3(nb), dim3(kBlock), 0, s, n, K.x, (const float *)K.e32, (const int *)K.state);
	}
	st->iterations = total;
	st->outer_iterations = sweeps;
	const double b2 = K.host_sc[kScBnorm2];
	st->relative_residual = b2 > 0 ? std::sqrt(K.host_sc[kScOuterRn2] / b2) : 0.0;
	if (!K.host_state[kStOuterDone]) { err = "mixed-precision refinement did not reach the requested tolerance"; return GARM_ENOCONV; }
	return GARM_OK;
}

int solvePcg64(int D, const Rows<double> &A, const double *b, int n, const garm_pcg_opts &o, Krylov &K, hipStream_t s,
	garm_pcg_stats *st, std::string &err)
{
	return D == 2 ? pcg64<2>(A, b, n, o, K, s, st, err) : pcg64<3>(A, b, n, o, K, s, st, err);
}

int solveMixed(int D, const Rows<double> &A64, const Rows<float> &A32, const double *b, int n, const garm_pcg_opts &o, Krylov &K,
	hipStream_t s, garm_pcg_stats *st, std::string &err)
{
	return D == 2 ? mixed<2>(A64, A32, b, n, o, K, s, st, err) : mixed<3>(A64, A32, b, n, o, K, s, st, err);
}

int timeSolverKernel(int D, const Rows<double> &A64, const Rows<float> &A32, int n, int which, int reps, Krylov &K,
	hipStream_t s, hipEvent_t e0, hipEvent_t e1, float *ms, double *bytes)
{
	if (n <= 0 || !K.sc) return GARM_EINVAL;
	if (which >= 2 && !K.r32) return GARM_EINVAL;
	std::string err;
	if (readState(K, s, err)) return GARM_EHIP;
	double sc0[32];
	int st0[16];
	std::copy(K.host_sc, K.host_sc + 16, sc0);
	std::copy(K.host_state, K.host_state + 8, st0);
	// alpha = 0 keeps x and r as they are while the kernels do their full work
	double scz[32];
	int stz[16];
	std::copy(sc0, sc0 + 16, scz);
	std::copy(st0, st0 + 8, stz);
	scz[kScAlpha] = 0;
	scz[kScThreshold] = -1;
	stz[kStDone] = 0;
	stz[kStMaxIter] = 1 << 30;
	stz[kStCounter] = 0;
	if (hipMemcpyAsync(K.sc, scz, 16 * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess ||
		hipMemcpyAsync(K.state, stz, 8 * sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess) return GARM_EHIP;
	const int nb = (n + kBlock - 1) / kBlock;
	(void)hipEventRecord(e0, s);
	for (int q = 0; q < reps; q++) {
		if (which == 0) {
			if (D == 2) hipLaunchKernelGGL((kSpmvAlpha<2, double>), dim3(nb), dim3(kBlock), 0, s, A64, n, (const double *)K.p, K.y, K.partials, K.sc, K.state);
			else hipLaunchKernelGGL((kSpmvAlpha<3, double>), dim3(nb), dim3(kBlock), 0, s, A64, n, (const double *)K.p, K.y, K.partials, K.sc, K.state);
		} else if (which == 1) {
			hipLaunchKernelGGL((kUpdateBeta<double, double>), dim3(nb), dim3(kBlock), 0, s, n, K.x, K.r, (const double *)K.p, (const double *)K.y,
				(const double *)A64.invdiag, K.partials, K.sc, K.state);
		} else if (which == 2) {
			if (D == 2) hipLaunchKernelGGL((kSpmvAlpha<2, float>), dim3(nb), dim3(kBlock), 0, s, A32, n, (const float *)K.p32, K.y32, K.partials, K.sc, K.state);
			else hipLaunchKernelGGL((kSpmvAlpha<3, float>), dim3(nb), dim3(kBlock), 0, s, A32, n, (const float *)K.p32, K.y32, K.partials, K.sc, K.state);
		} else {
			hipLaunchKernelGGL((kUpdateBeta<float, float>), dim3(nb), dim3(kBlock), 0, s, n, K.e32, K.r32, (const float *)K.p32, (const float *)K.y32,
				(const float *)A32.invdiag, K.partials, K.sc, K.state);
		}
	}
	(void)hipEventRecord(e1, s);
	if (hipMemcpyAsync(K.sc, sc0, 16 * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess ||
		hipMemcpyAsync(K.state, st0, 8 * sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess ||
		hipStreamSynchronize(s) != hipSuccess) return GARM_EHIP;
	(void)hipEventElapsedTime(ms, e0, e1);
	const double idx = 4.0 * 2 * D;
	if (which == 0) *bytes = double(n) * (8 + idx + 8 * 2 * D + 8 + 8);
	else if (which == 1) *bytes = double(n) * (5 * 8 + 2 * 8);
	else if (which == 2) *bytes = double(n) * (4 + idx + 4 * 2 * D + 4 + 4);
	else *bytes = double(n) * (5 * 4 + 2 * 4);
	return GARM_OK;
}

} // namespace garm
