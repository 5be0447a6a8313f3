// Krylov solvers for the ghost-fluid Poisson system on gfx950.
//
// One conjugate-gradient iteration is three launches over a fixed grid of
// at most kGrid workgroups (grid-stride loops):
//   kSpmv     y = A p; per-workgroup partial of p.y
//   kUpdate   alpha = (r.z) / sum(p.y partials); x += alpha p; r -= alpha y;
//             per-workgroup partials of |r|^2 and r.M^-1 r
//   kUpdateP  convergence test on sum(|r|^2 partials), beta, p = M^-1 r + beta p
// Every workgroup of a consumer kernel sums the producer's partials itself,
// in the same order, so all of them hold bit-identical scalars without a
// device-wide fence or atomic (on gfx950 an agent-scope release fence writes
// back the whole L2 and would cost more than the kernel).  r.z is kept in
// two slots chosen by iteration parity so the workgroup that stores the new
// value never races one still reading the old.  Dot products and norms are
// accumulated in FP64 whatever the vector type.
//
// The mixed-precision solver keeps p, r, y in FP32 and the solution x in FP64,
// and replaces the recursive residual by the FP64 residual b - A x whenever
// the FP32 one has fallen by kReplaceDrop since the last replacement, and
// before convergence is accepted (residual replacement: FP64 iterative
// refinement that keeps the Krylov direction instead of restarting).
#include "garm_pcg.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>

namespace garm {

constexpr int kGrid = 1024;               // workgroups of every solver launch
constexpr double kReplaceDrop = 1e-3;     // FP32 residual reduction between replacements

static int solverBlocks(int n) { return std::max(1, std::min(kGrid, (n + kBlock - 1) / kBlock)); }

// Sum NV interleaved partials of `nb` workgroups; result valid in thread 0.
template <int NV>
__device__ __forceinline__ void sumPartials(const double *partials, int nb, double (&tot)[NV])
{
	double s[NV];
#pragma unroll
	for (int q = 0; q < NV; q++) s[q] = 0;
	for (int i = threadIdx.x; i < nb; i += blockDim.x)
#pragma unroll
		for (int q = 0; q < NV; q++) s[q] += partials[NV * i + q];
	blockSum<NV>(s, tot);
}

// Block-wide broadcast of NV doubles from thread 0.
template <int NV>
__device__ __forceinline__ void broadcast(double (&v)[NV])
{
	__shared__ double b[NV];
	__syncthreads();
	if (threadIdx.x == 0)
#pragma unroll
		for (int q = 0; q < NV; q++) b[q] = v[q];
	__syncthreads();
#pragma unroll
	for (int q = 0; q < NV; q++) v[q] = b[q];
}

template <int NV>
__device__ __forceinline__ void storePartials(double (&v)[NV], double *partials)
{
	double blk[NV];
	blockSum<NV>(v, blk);
	if (threadIdx.x == 0)
#pragma unroll
		for (int q = 0; q < NV; q++) partials[NV * blockIdx.x + q] = blk[q];
}

// ---------------------------------------------------------------- kernels
template <int D, typename T>
__device__ __forceinline__ T applyRow(const Rows<T> &A, const T *p, int r, T pr)
{
	T s = A.diag[r] * pr;
#pragma unroll
	for (int o = 0; o < 2 * D; o++) {
		const int c = A.nb[o][r];
		if (c >= 0) s += -A.w[o][r] * p[c];
	}
	return s;
}

// y = A p
template <int D, typename T>
__global__ __launch_bounds__(kBlock) void kSpmv(Rows<T> A, int n, const T *p, T *y, double *pa, const int *st)
{
	if (st[kStDone]) return;
	double v[1] = {0};
	for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
		const T pr = p[r];
		const T s = applyRow<D, T>(A, p, r, pr);
		y[r] = s;
		v[0] += double(pr) * double(s);
	}
	storePartials<1>(v, pa);
}

// x += alpha p; r -= alpha y with alpha = rz[slot] / (p.y); partials of |r|^2, r.z
template <typename T, typename TX>
__global__ __launch_bounds__(kBlock) void kUpdate(int n, TX *x, T *r, const T *p, const T *y, const T *invdiag,
	const double *pa, double *pb, double *sc, const int *st, int slot)
{
	if (st[kStDone]) return;
	double a[1];
	sumPartials<1>(pa, gridDim.x, a);
	if (threadIdx.x == 0) a[0] = sc[kScRz0 + slot] / a[0];
	broadcast<1>(a);
	if (blockIdx.x == 0 && threadIdx.x == 0) sc[kScAlpha] = a[0];
	const TX alphaX = TX(a[0]);
	const T alpha = T(a[0]);
	double v[2] = {0, 0};
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
		x[i] += alphaX * TX(p[i]);
		const T ri = r[i] - alpha * y[i];
		r[i] = ri;
		v[0] += double(ri) * double(ri);
		v[1] += double(ri) * double(invdiag[i] * ri);
	}
	storePartials<2>(v, pb);
}

// Convergence |r|^2 < threshold (Eigen's test), else beta and p update.
template <typename T>
__global__ __launch_bounds__(kBlock) void kUpdateP(int n, T *p, const T *r, const T *invdiag, const double *pb, double *sc, int *st, int slot)
{
	if (st[kStDone]) return;
	double t[2];
	sumPartials<2>(pb, gridDim.x, t);
	double b[2] = {0, 0};   // beta, converged flag
	if (threadIdx.x == 0) {
		if (t[0] < sc[kScThreshold]) b[1] = 1;
		else b[0] = t[1] / sc[kScRz0 + slot];
	}
	broadcast<2>(b);
	if (blockIdx.x == 0 && threadIdx.x == 0) {
		sc[kScRn2] = t[0];
		if (b[1] != 0) st[kStDone] = 1;
		else {
			sc[kScRz0 + (slot ^ 1)] = t[1];
			sc[kScBeta] = b[0];
			const int it = st[kStIter] + 1;
			st[kStIter] = it;
			if (it >= st[kStMaxIter]) st[kStDone] = 1;
		}
	}
	if (b[1] != 0) return;
	const T beta = T(b[0]);
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = invdiag[i] * r[i] + beta * p[i];
}

// x = 0, r = b (in T), p = M^-1 r; partials of |b|^2 (FP64) and r.z.
template <typename T>
__global__ __launch_bounds__(kBlock) void kStart(int n, const double *b, const T *invdiag, double *x, T *r, T *p, double *pb)
{
	double v[2] = {0, 0};
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
		const double bi = b[i];
		x[i] = 0;
		const T ri = T(bi);
		r[i] = ri;
		const T z = invdiag[i] * ri;
		p[i] = z;
		v[0] += bi * bi;
		v[1] += double(ri) * double(z);
	}
	storePartials<2>(v, pb);
}

// Start-up tests of Eigen's conjugate_gradient with the tolerance floor of
// LinearSystem::solve (LinearSystem.h:31).  One workgroup.
__global__ __launch_bounds__(kBlock) void kStartFinish(const double *pb, int nb, double tol, int maxIt, double *sc, int *st)
{
	double t[2];
	sumPartials<2>(pb, nb, t);
	if (threadIdx.x != 0) return;
	const double b2 = t[0];
	sc[kScBnorm2] = b2;
	sc[kScRn2] = b2;
	sc[kScOuterRn2] = b2;
	sc[kScRz0] = t[1];
	st[kStIter] = 0;
	st[kStMaxIter] = maxIt;
	st[kStDone] = 0;
	st[kStOuterDone] = 0;
	if (b2 == 0) { st[kStDone] = 1; st[kStOuterDone] = 1; sc[kScThreshold] = 0; sc[kScOuterThreshold] = 0; sc[kScTol] = tol; return; }
	const double tt = fmax(tol, DBL_EPSILON / sqrt(b2));
	sc[kScTol] = tt;
	const double thr = fmax(tt * tt * b2, DBL_MIN);
	sc[kScOuterThreshold] = thr;
	// FP64 CG tests the final threshold directly; the mixed solver first aims
	// at a reduction of kReplaceDrop, then replaces the residual
	sc[kScThreshold] = thr;
	if (b2 < thr || maxIt <= 0) { st[kStDone] = 1; st[kStOuterDone] = 1; }
}

__global__ void kMixedThreshold(double *sc, const int *st)
{
	if (st[kStOuterDone]) return;
	sc[kScThreshold] = fmax(sc[kScOuterThreshold], kReplaceDrop * kReplaceDrop * sc[kScBnorm2]);
}

// Residual replacement: r64 = b - A x in FP64, r = T(r64); partials of
// |r64|^2 and r.z.
template <int D, typename T>
__global__ __launch_bounds__(kBlock) void kReplace(Rows<double> A, const T *invdiag, int n, const double *b, const double *x, T *r, double *pb)
{
	double v[2] = {0, 0};
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
		double ax = A.diag[i] * x[i];
#pragma unroll
		for (int o = 0; o < 2 * D; o++) {
			const int c = A.nb[o][i];
			if (c >= 0) ax += -A.w[o][i] * x[c];
		}
		const double ri = b[i] - ax;
		const T rt = T(ri);
		r[i] = rt;
		v[0] += ri * ri;
		v[1] += double(rt) * double(invdiag[i] * rt);
	}
	storePartials<2>(v, pb);
}

// After a replacement: accept if the FP64 residual meets the tolerance,
// otherwise continue CG from the replaced residual with the same direction:
// beta = (r.z)_new / (r.z)_old, p = M^-1 r + beta p.
template <typename T>
__global__ __launch_bounds__(kBlock) void kReplaceFinish(int n, T *p, const T *r, const T *invdiag, const double *pb, double *sc, int *st, int slot)
{
	double t[2];
	sumPartials<2>(pb, gridDim.x, t);
	double b[2] = {0, 0};
	if (threadIdx.x == 0) {
		if (t[0] < sc[kScOuterThreshold]) b[1] = 1;
		else b[0] = t[1] / sc[kScRz0 + slot];
	}
	broadcast<2>(b);
	if (blockIdx.x == 0 && threadIdx.x == 0) {
		sc[kScOuterRn2] = t[0];
		sc[kScRn2] = t[0];
		if (b[1] != 0) { st[kStOuterDone] = 1; st[kStDone] = 1; }
		else {
			sc[kScRz0 + (slot ^ 1)] = t[1];
			sc[kScThreshold] = fmax(sc[kScOuterThreshold], kReplaceDrop * kReplaceDrop * t[0]);
			const int it = st[kStIter] + 1;
			st[kStIter] = it;
			st[kStSweep] = st[kStSweep] + 1;
			st[kStDone] = it >= st[kStMaxIter] ? 1 : 0;
		}
	}
	if (b[1] != 0) return;
	const T beta = T(b[0]);
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = invdiag[i] * r[i] + beta * p[i];
}

__global__ void kSetInt(int *p, int v) { *p = v; }

// ----------------------------------------------------------------- driver
Krylov::~Krylov() { release(); }

void Krylov::release()
{
	auto fr = [](auto *&q) { if (q) { (void)hipFree(q); q = nullptr; } };
	fr(x); fr(r); fr(p); fr(y); fr(r32); fr(p32); fr(y32); fr(pa); fr(pb); fr(sc); fr(state);
	if (host_state) { (void)hipHostFree(host_state); host_state = nullptr; }
	if (host_sc) { (void)hipHostFree(host_sc); host_sc = nullptr; }
	if (ready) { (void)hipEventDestroy(ready); ready = nullptr; }
	cap = 0;
}

bool Krylov::reserve(size_t n, bool mixed, std::string &err)
{
	bool ok = true;
	if (!sc) {
		ok = hipMalloc(&sc, 32 * sizeof(double)) == hipSuccess && hipMalloc(&state, 16 * sizeof(int)) == hipSuccess
			&& hipHostMalloc(&host_state, 16 * sizeof(int)) == hipSuccess && hipHostMalloc(&host_sc, 32 * sizeof(double)) == hipSuccess
			&& hipEventCreateWithFlags(&ready, hipEventDisableTiming) == hipSuccess
			&& hipMalloc(&pa, kGrid * sizeof(double)) == hipSuccess && hipMalloc(&pb, 2 * kGrid * sizeof(double)) == hipSuccess
			&& hipMemset(state, 0, 16 * sizeof(int)) == hipSuccess && hipMemset(sc, 0, 32 * sizeof(double)) == hipSuccess;
		if (!ok) { err = "solver scalar allocation failed"; return false; }
	}
	if (n > cap || !x) {
		auto fr = [](auto *&q) { if (q) { (void)hipFree(q); q = nullptr; } };
		fr(x); fr(r); fr(p); fr(y); fr(r32); fr(p32); fr(y32);
		const size_t want = std::max<size_t>(n + n / 4, 1024);
		for (double **q : {&x, &r, &p, &y}) ok = ok && hipMalloc(q, want * sizeof(double)) == hipSuccess;
		if (!ok) { err = "solver vector allocation failed"; return false; }
		cap = want;
	}
	if (mixed && !r32) {
		for (float **q : {&r32, &p32, &y32}) ok = ok && hipMalloc(q, cap * sizeof(float)) == hipSuccess;
		if (!ok) { err = "FP32 solver vector allocation failed"; return false; }
	}
	return true;
}

static int readState(Krylov &K, hipStream_t s, std::string &err)
{
	if (hipMemcpyAsync(K.host_state, K.state, 8 * sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
		hipMemcpyAsync(K.host_sc, K.sc, 16 * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
		spinWait(s, K.ready) != hipSuccess) { err = "solver state readback failed"; return GARM_EHIP; }
	return GARM_OK;
}

// Enqueue CG iterations in chunks of `check` until the device sets its done
// flag or `budget` more iterations were enqueued.
template <int D, typename T, typename TX>
static int iterate(const Rows<T> &A, int n, TX *x, T *r, T *p, T *y, const T *invdiag, Krylov &K, int budget, int check,
	hipStream_t s, std::string &err)
{
	const dim3 g1(solverBlocks(n)), b1(kBlock);
	int launched = 0, rc;
	for (;;) {
		if ((rc = readState(K, s, err))) return rc;
		if (K.host_state[kStDone] || launched >= budget) return GARM_OK;
		const int base = K.host_state[kStIter];
		const int k = std::min(check, budget - launched);
		for (int q = 0; q < k; q++) {
			const int slot = (base + q) & 1;
			hipLaunchKernelGGL((kSpmv<D, T>), g1, b1, 0, s, A, n, (const T *)p, y, K.pa, (const int *)K.state);
			hipLaunchKernelGGL((kUpdate<T, TX>), g1, b1, 0, s, n, x, r, (const T *)p, (const T *)y, invdiag, (const double *)K.pa, K.pb, K.sc,
				(const int *)K.state, slot);
			hipLaunchKernelGGL((kUpdateP<T>), g1, b1, 0, s, n, p, (const T *)r, invdiag, (const double *)K.pb, K.sc, K.state, slot);
		}
		launched += k;
	}
}

static void finishStats(Krylov &K, garm_pcg_stats *st)
{
	st->iterations = K.host_state[kStIter];
	const double b2 = K.host_sc[kScBnorm2];
	st->relative_residual = b2 > 0 ? std::sqrt(K.host_sc[kScOuterRn2] / b2) : 0.0;
}

template <int D>
static int pcg64(const Rows<double> &A, const double *b, int n, const garm_pcg_opts &o, Krylov &K, hipStream_t s,
	garm_pcg_stats *st, std::string &err)
{
	const int nb = solverBlocks(n);
	hipLaunchKernelGGL((kStart<double>), dim3(nb), dim3(kBlock), 0, s, n, b, (const double *)A.invdiag, K.x, K.r, K.p, K.pb);
	hipLaunchKernelGGL(kStartFinish, dim3(1), dim3(kBlock), 0, s, (const double *)K.pb, nb, o.tolerance, o.max_iterations, K.sc, K.state);
	int rc = iterate<D, double, double>(A, n, K.x, K.r, K.p, K.y, A.invdiag, K, o.max_iterations + 1, std::max(1, o.check_interval), s, err);
	if (rc || (rc = readState(K, s, err))) return rc;
	K.host_sc[kScOuterRn2] = K.host_sc[kScRn2];
	finishStats(K, st);
	st->outer_iterations = 1;
	if (st->relative_residual > K.host_sc[kScTol]) { err = "PCG did not reach the requested tolerance"; return GARM_ENOCONV; }
	return GARM_OK;
}

template <int D>
static int mixed(const Rows<double> &A64, const Rows<float> &A32, const double *b, int n, const garm_pcg_opts &o, Krylov &K,
	hipStream_t s, garm_pcg_stats *st, std::string &err)
{
	const int nb = solverBlocks(n);
	const int check = std::max(1, o.check_interval);
	hipLaunchKernelGGL((kStart<float>), dim3(nb), dim3(kBlock), 0, s, n, b, (const float *)A32.invdiag, K.x, K.r32, K.p32, K.pb);
	hipLaunchKernelGGL(kStartFinish, dim3(1), dim3(kBlock), 0, s, (const double *)K.pb, nb, o.tolerance, o.max_iterations, K.sc, K.state);
	hipLaunchKernelGGL(kMixedThreshold, dim3(1), dim3(1), 0, s, K.sc, (const int *)K.state);
	hipLaunchKernelGGL(kSetInt, dim3(1), dim3(1), 0, s, K.state + kStSweep, 0);
	int rc = GARM_OK, replacements = 0;
	for (;;) {
		if ((rc = iterate<D, float, double>(A32, n, K.x, K.r32, K.p32, K.y32, A32.invdiag, K, o.max_iterations + 1, check, s, err))) return rc;
		if ((rc = readState(K, s, err))) return rc;
		if (K.host_state[kStOuterDone] || K.host_state[kStIter] >= o.max_iterations || replacements > 200) break;
		// the FP32 residual met its target: replace it by the FP64 one
		const int slot = K.host_state[kStIter] & 1;
		hipLaunchKernelGGL((kReplace<D, float>), dim3(nb), dim3(kBlock), 0, s, A64, (const float *)A32.invdiag, n, b, (const double *)K.x, K.r32, K.pb);
		hipLaunchKernelGGL((kReplaceFinish<float>), dim3(nb), dim3(kBlock), 0, s, n, K.p32, (const float *)K.r32, (const float *)A32.invdiag,
			(const double *)K.pb, K.sc, K.state, slot);
		replacements++;
	}
	finishStats(K, st);
	st->outer_iterations = replacements;
	if (!K.host_state[kStOuterDone]) { err = "mixed-precision solve did not reach the requested tolerance"; return GARM_ENOCONV; }
	return GARM_OK;
}

// Algorithmic HBM bytes per unknown: an FP64 Jacobi-PCG iteration (product:
// p, diagonal and three face weights in, Ap out; update: x, r, p, Ap and the
// inverse diagonal in, x, r, z out; p update: z, p in, p out), its FP32
// counterpart, and an FP64 residual replacement (x, b, the rows in, r out).
constexpr double kPcg64IterBytes = 136, kPcg32IterBytes = 68, kReplaceBytes = 104;

int solvePcg64(int D, const Rows<double> &A, const double *b, int n, const garm_pcg_opts &o, Krylov &K, hipStream_t s,
	garm_pcg_stats *st, std::string &err)
{
	const int rc = D == 2 ? pcg64<2>(A, b, n, o, K, s, st, err) : pcg64<3>(A, b, n, o, K, s, st, err);
	st->hbm_bytes = kPcg64IterBytes * double(n) * st->iterations;
	st->solver_cells = n;
	return rc;
}

int solveMixed(int D, const Rows<double> &A64, const Rows<float> &A32, const double *b, int n, const garm_pcg_opts &o, Krylov &K,
	hipStream_t s, garm_pcg_stats *st, std::string &err)
{
	const int rc = D == 2 ? mixed<2>(A64, A32, b, n, o, K, s, st, err) : mixed<3>(A64, A32, b, n, o, K, s, st, err);
	st->hbm_bytes = (kPcg32IterBytes * st->iterations + kReplaceBytes * st->outer_iterations) * double(n);
	st->solver_cells = n;
	return rc;
}

int timeSolverKernel(int D, const Rows<double> &A64, const Rows<float> &A32, int n, int which, int reps, Krylov &K,
	hipStream_t s, hipEvent_t e0, hipEvent_t e1, float *ms, double *bytes)
{
	if (n <= 0 || !K.sc) return GARM_EINVAL;
	if (which >= 2 && !K.r32) return GARM_EINVAL;
	std::string err;
	if (readState(K, s, err)) return GARM_EHIP;
	double sc0[16];
	int st0[8];
	std::copy(K.host_sc, K.host_sc + 16, sc0);
	std::copy(K.host_state, K.host_state + 8, st0);
	// r.z = 0 makes alpha 0, so x and r keep their values while the update
	// kernels do their full memory traffic
	double scz[16];
	int stz[8];
	std::copy(sc0, sc0 + 16, scz);
	std::copy(st0, st0 + 8, stz);
	scz[kScRz0] = scz[kScRz1] = 0;
	stz[kStDone] = 0;
	if (hipMemcpyAsync(K.sc, scz, 16 * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess ||
		hipMemcpyAsync(K.state, stz, 8 * sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess) return GARM_EHIP;
	const int nb = solverBlocks(n);
	(void)hipEventRecord(e0, s);
	for (int q = 0; q < reps; q++) {
		if (which == 0) {
			if (D == 2) hipLaunchKernelGGL((kSpmv<2, double>), dim3(nb), dim3(kBlock), 0, s, A64, n, (const double *)K.p, K.y, K.pa, (const int *)K.state);
			else hipLaunchKernelGGL((kSpmv<3, double>), dim3(nb), dim3(kBlock), 0, s, A64, n, (const double *)K.p, K.y, K.pa, (const int *)K.state);
		} else if (which == 1) {
			hipLaunchKernelGGL((kUpdate<double, double>), dim3(nb), dim3(kBlock), 0, s, n, K.x, K.r, (const double *)K.p, (const double *)K.y,
				(const double *)A64.invdiag, (const double *)K.pa, K.pb, K.sc, (const int *)K.state, 0);
		} else if (which == 2) {
			if (D == 2) hipLaunchKernelGGL((kSpmv<2, float>), dim3(nb), dim3(kBlock), 0, s, A32, n, (const float *)K.p32, K.y32, K.pa, (const int *)K.state);
			else hipLaunchKernelGGL((kSpmv<3, float>), dim3(nb), dim3(kBlock), 0, s, A32, n, (const float *)K.p32, K.y32, K.pa, (const int *)K.state);
		} else {
			hipLaunchKernelGGL((kUpdate<float, double>), dim3(nb), dim3(kBlock), 0, s, n, K.x, K.r32, (const float *)K.p32, (const float *)K.y32,
				(const float *)A32.invdiag, (const double *)K.pa, K.pb, K.sc, (const int *)K.state, 0);
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
	else *bytes = double(n) * (8 + 4 * 4 + 8 + 4);   // x (FP64) read+write, p, y, invdiag, r read, r write
	return GARM_OK;
}

} // namespace garm
