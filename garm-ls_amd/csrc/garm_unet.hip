// Three-level 3D U-Net (16/32/64 channels) on gfx950: BF16 channels-last
// activations and weights, 3x3x3 convolutions as implicit GEMM on
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation, FP32 master weights and
// Adam.  The network has no counterpart in the reference; its layer order,
// layouts and gradients follow oracle/unet_oracle.py, which the tests hold it
// against.
//
// Forward and data-gradient convolutions: kConv3Rows for output groups of 16
// and 32 channels (an A fragment of one input row feeds every output row it
// is a tap of), kConv3 for the rest (48-channel data gradients).  Both stage
// the input tile with a one-voxel halo and the chunk's weights in LDS, 16
// input channels at a time, so every input voxel is fetched from HBM once
// per chunk and output group instead of 27 times.  Weight gradients:
// kConv3WGrad, voxels on the K dimension, per-workgroup partials summed in a
// fixed order by kReduceWGrad.
//
// MFMA 16x16x32 BF16 lane layout (wave64): lane l holds A[l%16][8(l/16)..+7],
// B[8(l/16)..+7][l%16], and D[4(l/16)+i][l%16] for i = 0..3.
#include "garm_unet.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>

namespace garm {

typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int TX = 16, TY = 4, TZ = 4, HX = TX + 2, HY = TY + 2, HZ = TZ + 2;
// forward convolution tile: CX x 4 x 4 voxels, one z-slice of CX / 16 x 4
// 16-voxel fragments per wave (CX = 32 was measured slower: twice the
// accumulators halve the occupancy)
constexpr int CX = 16, CHX = CX + 2, kFrags = CX / 16 * TY;
constexpr int kUB = 256;

static int roundUp16(int c) { return (c + 15) / 16 * 16; }

// ------------------------------------------------------------- convolution
struct ConvArgs {
	const bf16 *in;
	int inStride, Cin;
	const bf16 *w;          // [27][coutP][cinP]
	int cinP, coutP;
	const float *bias;
	bf16 *out;
	int outStride, outOff, Cout;
	int nx, ny, nz;
	int relu;
	// optional: input channels [0, upC) read from a grid of half the size
	// (nearest-neighbour upsampling folded into the staging; upC a multiple
	// of 8, of 4 for the narrow staging)
	const bf16 *up;
	int upStride, upC;
	// optional: the 2x2x2 means of the BF16 output into a grid of half the
	// size (channel co at pool[parent * poolStride + co]): the eight children
	// summed in FP32, x fastest, then y, then z, times 1/8, rounded to BF16
	bf16 *pool;
	int poolStride;
};

// K = 32 per MFMA (v_mfma_f32_16x16x32_bf16): the reduction runs over pairs
// of taps of one 16-channel chunk, lanes 0-31 of a fragment holding the
// chunk's channels at the first tap of the pair and lanes 32-63 at the
// second (the 28th tap is a zero weight).  Lane l of an A fragment holds
// A[l % 16][8 (l / 16) .. + 7], of a B fragment B[8 (l / 16) .. + 7][l % 16];
// D as for the K = 16 instruction.
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef short s8v __attribute__((ext_vector_type(8)));

template <bool WIDE>
__global__ __launch_bounds__(kUB) void kConv3(ConvArgs A)
{
	// input halo (18 x 6 x 6 voxels, 16 channels), then the chunk's weights
	// [27][NF * 16][16] (the launch sizes the LDS for NF, so a 16-channel
	// output takes 34 KB and four workgroups fit on a CU); the output tile
	// (256 voxels x 32 channels) reuses the halo once the reduction is done
	extern __shared__ __align__(16) bf16 lds[];
	bf16 *tin = lds, *tw = lds + HZ * HY * CHX * 16;
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const int kg = lane >> 4, li = lane & 15;
	const int x0 = blockIdx.x * CX, y0 = blockIdx.y * TY;
	const int nzt = (A.nz + TZ - 1) / TZ;
	const int zt = blockIdx.z % nzt, grp = blockIdx.z / nzt;
	const int z0 = zt * TZ, co0 = grp * 32;
	const int NF = min(32, A.coutP - co0) / 16;
	const int WS = NF * 16;   // weight rows per tap in LDS
	const int half = (kg & 1) * 8, second = kg >> 1;   // channel half and tap of the pair this lane holds
	f4 acc[kFrags][2];
#pragma unroll
	for (int f = 0; f < kFrags; f++)
#pragma unroll
		for (int q = 0; q < 2; q++) acc[f][q] = f4{0, 0, 0, 0};

	// 16-byte staging (all loads of a thread in flight together) when every
	// voxel's channels come in whole groups of eight (WIDE); 8-byte otherwise
	for (int ci0 = 0; ci0 < A.Cin; ci0 += 16) {
		if constexpr (WIDE) {
			constexpr int kVecs = HZ * HY * CHX * 2, kIt = (kVecs + kUB - 1) / kUB;
			uint4 hv[kIt];
#pragma unroll
			for (int r = 0; r < kIt; r++) {
				const int q = tid + r * kUB, v = q >> 1, ch = ci0 + (q & 1) * 8;
				const int hx = v % CHX, hy = (v / CHX) % HY, hz = v / (CHX * HY);
				const int gx = x0 + hx - 1, gy = y0 + hy - 1, gz = z0 + hz - 1;
				hv[r] = make_uint4(0, 0, 0, 0);
				if (q < kVecs && gx >= 0 && gx < A.nx && gy >= 0 && gy < A.ny && gz >= 0 && gz < A.nz && ch < A.Cin) {
					const bf16 *src = A.up && ch < A.upC
						? A.up + (size_t((gz >> 1) * (A.ny >> 1) + (gy >> 1)) * (A.nx >> 1) + (gx >> 1)) * A.upStride + ch
						: A.in + (size_t(gz * A.ny + gy) * A.nx + gx) * A.inStride + ch;
					hv[r] = *reinterpret_cast<const uint4 *>(src);
				}
			}
#pragma unroll
			for (int r = 0; r < kIt; r++) {
				const int q = tid + r * kUB;
				if (q < kVecs) *reinterpret_cast<uint4 *>(tin + q * 8) = hv[r];
			}
		} else {
			constexpr int kVecs = HZ * HY * CHX * 4, kIt = (kVecs + kUB - 1) / kUB;
			uint2 hv[kIt];
#pragma unroll
			for (int r = 0; r < kIt; r++) {
				const int q = tid + r * kUB, quad = q & 3, v = q >> 2;
				const int hx = v % CHX, hy = (v / CHX) % HY, hz = v / (CHX * HY);
				const int gx = x0 + hx - 1, gy = y0 + hy - 1, gz = z0 + hz - 1;
				const int ch = ci0 + quad * 4;
				hv[r] = make_uint2(0, 0);
				if (q < kVecs && gx >= 0 && gx < A.nx && gy >= 0 && gy < A.ny && gz >= 0 && gz < A.nz && ch < A.Cin) {
					const bf16 *src = A.up && ch < A.upC
						? A.up + (size_t((gz >> 1) * (A.ny >> 1) + (gy >> 1)) * (A.nx >> 1) + (gx >> 1)) * A.upStride + ch
						: A.in + (size_t(gz * A.ny + gy) * A.nx + gx) * A.inStride + ch;
					hv[r] = *reinterpret_cast<const uint2 *>(src);
				}
			}
#pragma unroll
			for (int r = 0; r < kIt; r++) {
				const int q = tid + r * kUB;
				if (q < kVecs) *reinterpret_cast<uint2 *>(tin + q * 4) = hv[r];
			}
		}
		{
			// weights [27][NF * 16][16] of this chunk (cinP and coutP are multiples of 16)
			constexpr int kMaxVecs = 27 * 32 * 2, kIt = (kMaxVecs + kUB - 1) / kUB;
			const int vecs = 27 * NF * 16 * 2;
			uint4 wv[kIt];
#pragma unroll
			for (int r = 0; r < kIt; r++) {
				const int q = tid + r * kUB, rr = q >> 1, part = q & 1;
				const int co = rr % (NF * 16), t = rr / (NF * 16);
				wv[r] = q < vecs ? *reinterpret_cast<const uint4 *>(A.w + (size_t(t) * A.coutP + co0 + co) * A.cinP + ci0 + part * 8) : make_uint4(0, 0, 0, 0);
			}
#pragma unroll
			for (int r = 0; r < kIt; r++) {
				const int q = tid + r * kUB, rr = q >> 1, part = q & 1;
				const int co = rr % (NF * 16), t = rr / (NF * 16);
				if (q < vecs) *reinterpret_cast<uint4 *>(tw + (t * WS + co) * 16 + part * 8) = wv[r];
			}
		}
		__syncthreads();
#pragma unroll 2
		for (int p = 0; p < 14; p++) {
			const int tl = 2 * p + second;                  // this lane's tap
			const bool none = tl > 26;                      // the 28th tap: zero weight
			const int t = none ? 26 : tl;                   // (a valid voxel for the A operand)
			const int dz = t / 9, dy = (t / 3) % 3, dx = t % 3;
			bf8v b[2];
#pragma unroll
			for (int q = 0; q < 2; q++) {
				const s8v raw = *reinterpret_cast<const s8v *>(tw + (t * WS + (q < NF ? q : 0) * 16 + li) * 16 + half);
				b[q] = __builtin_bit_cast(bf8v, none ? s8v{0, 0, 0, 0, 0, 0, 0, 0} : raw);
			}
#pragma unroll
			for (int f = 0; f < kFrags; f++) {
				const int fy = f & 3, fx = f >> 2;   // row of the slice, 16-voxel half of the row
				const bf8v a = __builtin_bit_cast(bf8v, *reinterpret_cast<const s8v *>(tin + (((wave + dz) * HY + fy + dy) * CHX + fx * 16 + li + dx) * 16 + half));
				acc[f][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[0], acc[f][0], 0, 0, 0);
				if (NF > 1) acc[f][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[1], acc[f][1], 0, 0, 0);
			}
		}
		__syncthreads();
	}
	// bias, activation, BF16: the tile goes to LDS as [voxel][32 channels]
	// and leaves as whole 16-byte words when the output layout allows
	bf16 *tout = tin;   // CX * 16 voxels x 32 BF16 <= the halo buffer
#pragma unroll
	for (int f = 0; f < kFrags; f++)
#pragma unroll
		for (int q = 0; q < 2; q++) {
			if (q >= NF) continue;
			const int co = co0 + q * 16 + li;
			const float bb = A.bias && co < A.Cout ? A.bias[co] : 0.0f;
#pragma unroll
			for (int i = 0; i < 4; i++) {
				float v = acc[f][q][i] + bb;
				if (A.relu) v = fmaxf(v, 0.0f);
				const int vox = (wave * TY + (f & 3)) * CX + (f >> 2) * 16 + kg * 4 + i;   // tile voxel (x fastest)
				tout[vox * 32 + q * 16 + li] = f2bf(v);
			}
		}
	__syncthreads();
	const int nch = min(NF * 16, A.Cout - co0);
	if (A.pool) {
		// the tile's origin is even along every axis, so it holds whole 2x2x2 families
		constexpr int PX = CX / 2, PY = TY / 2, PZ = TZ / 2;
		const int pnx = A.nx >> 1, pny = A.ny >> 1, pnz = A.nz >> 1;
		for (int q = tid; q < PX * PY * PZ * nch; q += kUB) {
			const int c = q % nch, pv = q / nch;
			const int px = pv % PX, py = (pv / PX) % PY, pz = pv / (PX * PY);
			const int X = (x0 >> 1) + px, Y = (y0 >> 1) + py, Z = (z0 >> 1) + pz;
			if (X >= pnx || Y >= pny || Z >= pnz) continue;
			float sum = 0;
#pragma unroll
			for (int k = 0; k < 8; k++) {
				const int vx = 2 * px + (k & 1), vy = 2 * py + (k >> 1 & 1), vz = 2 * pz + (k >> 2);
				sum += bf2f(tout[((vz * TY + vy) * CX + vx) * 32 + c]);
			}
			A.pool[(size_t(Z * pny + Y) * pnx + X) * A.poolStride + co0 + c] = f2bf(sum * 0.125f);
		}
	}
	if (nch == NF * 16 && (A.outStride & 7) == 0 && ((A.outOff + co0) & 7) == 0) {
		const int words = NF * 2;   // 16-byte words per voxel
		for (int q = tid; q < CX * TY * TZ * words; q += kUB) {
			const int vox = q / words, w8 = q % words;
			const int x = x0 + vox % CX, y = y0 + (vox / CX) % TY, z = z0 + vox / (CX * TY);
			if (x >= A.nx || y >= A.ny || z >= A.nz) continue;
			*reinterpret_cast<uint4 *>(A.out + (size_t(z * A.ny + y) * A.nx + x) * A.outStride + A.outOff + co0 + w8 * 8) =
				*reinterpret_cast<const uint4 *>(tout + vox * 32 + w8 * 8);
		}
	} else {
		for (int q = tid; q < CX * TY * TZ * nch; q += kUB) {
			const int vox = q / nch, c = q % nch;
			const int x = x0 + vox % CX, y = y0 + (vox / CX) % TY, z = z0 + vox / (CX * TY);
			if (x >= A.nx || y >= A.ny || z >= A.nz) continue;
			A.out[(size_t(z * A.ny + y) * A.nx + x) * A.outStride + A.outOff + co0 + c] = tout[vox * 32 + c];
		}
	}
}

// Row-reuse convolutions.  kConv3 reads an A fragment from LDS for every
// one or two MFMAs (one per 16 output channels).  kConv3Rows<NF, WZ> reuses
// each A fragment instead: a wave owns 4 (y) x WZ (z) rows of 16 voxels and
// NF groups of 16 output channels, and the fragment of one input row at one
// x shift feeds every output row of the wave that the row is a (dy, dz) tap
// of, for every output group -- up to 3 min(WZ, 3) NF MFMAs.  Per pair of
// the K dimension the 9 NF B fragments of the (dy, dz) taps stay in
// registers (pair 0: the taps dx = 0 in K 0-15 and dx = 1 in K 16-31; pair
// 1: the tap dx = 2 and zeros).  Per 16-channel chunk and wave, NF = 1 with
// WZ = 2: 48 A reads for 144 MFMAs; NF = 2 with WZ = 1: 36 reads for 144
// (kConv3: 70 reads for 56 or 112); the zero halves make it 144 instead of
// 108.  Workgroup tile 16 x 8 x 2 WZ voxels (the four waves 2 x 2 over y and
// z) for 16 NF output channels (output group blockIdx.z / z-tiles as in
// kConv3); its halo of the chunk and the chunk's weights in LDS (48 and
// 50 KB: three workgroups per CU).  Deeper tiles (WZ = 4 for NF = 1, 2 for
// NF = 2) read fewer A fragments per MFMA but fit two workgroups per CU and
// were measured 2-12% slower.
constexpr int RX = 16, RY = 8, RHX = RX + 2, RHY = RY + 2;
// z rows per wave and workgroups per CU of the two forms
constexpr int kRowsWZ1 = 2, kRowsOcc1 = 3, kRowsWZ2 = 1, kRowsOcc2 = 3;
template <int NF, int WZ_> struct RowsTile {
	static constexpr int WZ = WZ_, RZ = 2 * WZ, RHZ = RZ + 2, HALO = RHX * RHY * RHZ, CO = 16 * NF;
};

template <bool WIDE, int NF, int WZ_, int OCC>
__global__ __launch_bounds__(kUB, OCC) void kConv3Rows(ConvArgs A)
{
	using RT = RowsTile<NF, WZ_>;
	constexpr int WZ = RT::WZ, RZ = RT::RZ, kHalo = RT::HALO, CO = RT::CO;
	static_assert(RX * RY * RZ * CO <= kHalo * 16, "the output tile fits in the halo buffer");
	__shared__ __align__(16) bf16 tin[kHalo * 16];   // the output tile [voxel][CO] reuses it
	__shared__ __align__(16) bf16 tw[27 * CO * 16];  // [tap][co][ci] of the chunk and group
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const int kg = lane >> 4, li = lane & 15;
	const int half = (kg & 1) * 8, second = kg >> 1;   // channel half and member of the K pair this lane holds
	const int nzt = (A.nz + RZ - 1) / RZ;
	const int zt = blockIdx.z % nzt, grp = blockIdx.z / nzt;
	const int x0 = blockIdx.x * RX, y0 = blockIdx.y * RY, z0 = zt * RZ, co0 = grp * CO;
	const int wy = (wave & 1) * 4, wz = (wave >> 1) * WZ;   // the wave's rows in the tile
	f4 acc[WZ][4][NF];   // [oz][oy][output group]
#pragma unroll
	for (int a = 0; a < WZ; a++)
#pragma unroll
		for (int b = 0; b < 4; b++)
#pragma unroll
			for (int f = 0; f < NF; f++) acc[a][b][f] = f4{0, 0, 0, 0};

	for (int ci0 = 0; ci0 < A.Cin; ci0 += 16) {
		// the staging's indices are recomputed per chunk from an opaque copy of
		// the thread id: hoisted out of the loop they would hold ~50 VGPRs
		// through the MFMA phase
		int tid = threadIdx.x;
		asm volatile("" : "+v"(tid));
		// the chunk's weights [27][CO][16] of the group (cinP a multiple of 16):
		// loads issued first, stored after the halo's
		constexpr int kWVecs = 27 * CO * 2, kWIt = (kWVecs + kUB - 1) / kUB;
		uint4 wv[kWIt];
#pragma unroll
		for (int r = 0; r < kWIt; r++) {
			const int q = tid + r * kUB, row = q >> 1, part = q & 1;   // row = t * CO + co
			const int t = row / CO, co = row % CO;
			wv[r] = q < kWVecs ? *reinterpret_cast<const uint4 *>(A.w + (size_t(t) * A.coutP + co0 + co) * A.cinP + ci0 + part * 8) : make_uint4(0, 0, 0, 0);
		}
		if constexpr (WIDE) {
			constexpr int kVecs = kHalo * 2, kIt = (kVecs + kUB - 1) / kUB;
			uint4 hv[kIt];   // every load of the halo in flight together
#pragma unroll
			for (int r = 0; r < kIt; r++) {
				const int q = tid + r * kUB, v = q >> 1, ch = ci0 + (q & 1) * 8;
				const int hx = v % RHX, hy = (v / RHX) % RHY, hz = v / (RHX * RHY);
				const int gx = x0 + hx - 1, gy = y0 + hy - 1, gz = z0 + hz - 1;
				hv[r] = make_uint4(0, 0, 0, 0);
				if (q < kVecs && gx >= 0 && gx < A.nx && gy >= 0 && gy < A.ny && gz >= 0 && gz < A.nz && ch < A.Cin) {
					const bf16 *src = A.up && ch < A.upC
						? A.up + (size_t((gz >> 1) * (A.ny >> 1) + (gy >> 1)) * (A.nx >> 1) + (gx >> 1)) * A.upStride + ch
						: A.in + (size_t(gz * A.ny + gy) * A.nx + gx) * A.inStride + ch;
					hv[r] = *reinterpret_cast<const uint4 *>(src);
				}
			}
#pragma unroll
			for (int r = 0; r < kIt; r++) {
				const int q = tid + r * kUB;
				if (q < kVecs) *reinterpret_cast<uint4 *>(tin + q * 8) = hv[r];
			}
		} else {
			constexpr int kVecs = kHalo * 4, kIt = (kVecs + kUB - 1) / kUB;
			for (int r0 = 0; r0 < kIt; r0 += 8) {
				uint2 hv[8];
#pragma unroll
				for (int rr = 0; rr < 8; rr++) {
					const int q = tid + (r0 + rr) * kUB, quad = q & 3, v = q >> 2;
					const int hx = v % RHX, hy = (v / RHX) % RHY, hz = v / (RHX * RHY);
					const int gx = x0 + hx - 1, gy = y0 + hy - 1, gz = z0 + hz - 1;
					const int ch = ci0 + quad * 4;
					hv[rr] = make_uint2(0, 0);
					if (q < kVecs && gx >= 0 && gx < A.nx && gy >= 0 && gy < A.ny && gz >= 0 && gz < A.nz && ch < A.Cin) {
						const bf16 *src = A.up && ch < A.upC
							? A.up + (size_t((gz >> 1) * (A.ny >> 1) + (gy >> 1)) * (A.nx >> 1) + (gx >> 1)) * A.upStride + ch
							: A.in + (size_t(gz * A.ny + gy) * A.nx + gx) * A.inStride + ch;
						hv[rr] = *reinterpret_cast<const uint2 *>(src);
					}
				}
#pragma unroll
				for (int rr = 0; rr < 8; rr++) {
					const int q = tid + (r0 + rr) * kUB;
					if (q < kVecs) *reinterpret_cast<uint2 *>(tin + q * 4) = hv[rr];
				}
			}
		}
#pragma unroll
		for (int r = 0; r < kWIt; r++) {
			const int q = tid + r * kUB;
			if (q < kWVecs) *reinterpret_cast<uint4 *>(tw + q * 8) = wv[r];
		}
		__syncthreads();
		// per pair p: its B fragments [dz][dy][group], then every input row
		// (hy, hz) of the wave at the pair's x shifts
#pragma unroll
		for (int p = 0; p < 2; p++) {
			bf8v bw[3][3][NF];
#pragma unroll
			for (int dz = 0; dz < 3; dz++)
#pragma unroll
				for (int dy = 0; dy < 3; dy++)
#pragma unroll
					for (int f = 0; f < NF; f++) {
						const int t = dz * 9 + dy * 3 + (p == 0 ? second : 2);
						const s8v raw = *reinterpret_cast<const s8v *>(tw + (t * CO + f * 16 + li) * 16 + half);
						bw[dz][dy][f] = __builtin_bit_cast(bf8v, p == 1 && second ? s8v{0, 0, 0, 0, 0, 0, 0, 0} : raw);
					}
			const int dx = p == 0 ? second : 2;
#pragma unroll
			for (int hz = 0; hz < WZ + 2; hz++) {
#pragma unroll
				for (int hy = 0; hy < 6; hy++) {
					const bf8v a = __builtin_bit_cast(bf8v, *reinterpret_cast<const s8v *>(tin + (((wz + hz) * RHY + wy + hy) * RHX + li + dx) * 16 + half));
#pragma unroll
					for (int oz = 0; oz < WZ; oz++) {
						if (hz - oz < 0 || hz - oz > 2) continue;
#pragma unroll
						for (int oy = 0; oy < 4; oy++) {
							if (hy - oy < 0 || hy - oy > 2) continue;
#pragma unroll
							for (int f = 0; f < NF; f++)
								acc[oz][oy][f] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw[hz - oz][hy - oy][f], acc[oz][oy][f], 0, 0, 0);
						}
					}
				}
				// bounds how far the scheduler hoists the next rows' LDS reads
				__builtin_amdgcn_sched_barrier(0);
			}
		}
		__syncthreads();
	}
	// bias, activation, BF16 into the tile [voxel][CO] (x fastest)
	bf16 *tout = tin;
#pragma unroll
	for (int f = 0; f < NF; f++) {
		const int co = co0 + f * 16 + li;
		const float bb = A.bias && co < A.Cout ? A.bias[co] : 0.0f;
#pragma unroll
		for (int oz = 0; oz < WZ; oz++)
#pragma unroll
			for (int oy = 0; oy < 4; oy++)
#pragma unroll
				for (int i = 0; i < 4; i++) {
					float v = acc[oz][oy][f][i] + bb;
					if (A.relu) v = fmaxf(v, 0.0f);
					const int vox = ((wz + oz) * RY + wy + oy) * RX + kg * 4 + i;
					tout[vox * CO + f * 16 + li] = f2bf(v);
				}
	}
	__syncthreads();
	const int nch = min(CO, A.Cout - co0);
	if (A.pool) {
		// the tile's origin is even along every axis: whole 2 x 2 x 2 families
		constexpr int PX = RX / 2, PY = RY / 2, PZ = RZ / 2;
		const int pnx = A.nx >> 1, pny = A.ny >> 1, pnz = A.nz >> 1;
		for (int q = tid; q < PX * PY * PZ * nch; q += kUB) {
			const int c = q % nch, pv = q / nch;
			const int px = pv % PX, py = (pv / PX) % PY, pz = pv / (PX * PY);
			const int X = (x0 >> 1) + px, Y = (y0 >> 1) + py, Z = (z0 >> 1) + pz;
			if (X >= pnx || Y >= pny || Z >= pnz) continue;
			float sum = 0;
#pragma unroll
			for (int k = 0; k < 8; k++) {
				const int vx = 2 * px + (k & 1), vy = 2 * py + (k >> 1 & 1), vz = 2 * pz + (k >> 2);
				sum += bf2f(tout[((vz * RY + vy) * RX + vx) * CO + c]);
			}
			A.pool[(size_t(Z * pny + Y) * pnx + X) * A.poolStride + co0 + c] = f2bf(sum * 0.125f);
		}
	}
	if (nch == CO && (A.outStride & 7) == 0 && ((A.outOff + co0) & 7) == 0) {
		constexpr int words = CO / 8;   // 16-byte words per voxel
		for (int q = tid; q < RX * RY * RZ * words; q += kUB) {
			const int vox = q / words, w8 = q % words;
			const int x = x0 + vox % RX, y = y0 + (vox / RX) % RY, z = z0 + vox / (RX * RY);
			if (x >= A.nx || y >= A.ny || z >= A.nz) continue;
			*reinterpret_cast<uint4 *>(A.out + (size_t(z * A.ny + y) * A.nx + x) * A.outStride + A.outOff + co0 + w8 * 8) =
				*reinterpret_cast<const uint4 *>(tout + vox * CO + w8 * 8);
		}
	} else {
		for (int q = tid; q < RX * RY * RZ * nch; q += kUB) {
			const int vox = q / nch, c = q % nch;
			const int x = x0 + vox % RX, y = y0 + (vox / RX) % RY, z = z0 + vox / (RX * RY);
			if (x >= A.nx || y >= A.ny || z >= A.nz) continue;
			A.out[(size_t(z * A.ny + y) * A.nx + x) * A.outStride + A.outOff + co0 + c] = tout[vox * CO + c];
		}
	}
}

// Weight gradient partials: for each tap t, dW_t = X_t^T Gz over this
// workgroup's 256 voxels (16 x 4 x 4; K = voxels on the MFMA), and the bias
// partial.  On v_mfma_f32_16x16x32_bf16 a lane holds eight consecutive K
// values of one row (A: input channel, B: output channel), i.e. one channel
// at eight voxels.  So both tiles are staged in LDS channel-major: the
// output gradient as [channel][voxel], and the input as three copies
// [channel][z][y][16 x] already shifted by the tap's dx = 0, 1, 2 (copy dx
// holds halo column x + dx at x).  A K step of 32 voxels is two 16-voxel
// rows, and every fragment is one aligned 16-byte LDS read: per tap and
// chunk eight A reads for 8 NF MFMAs, the B fragments of the 256 voxels in
// registers for all taps.  (Workgroups that sum a run of tiles along z in
// registers, writing fewer partials, were measured 1.6x slower: the runs
// serialise the staging that independent workgroups overlap.)
struct WGradArgs {
	const bf16 *x;
	int xStride, Cin;
	const bf16 *gz;
	int gStride, gOff, Cout;
	int cinP, coutP;
	int nx, ny, nz;
	float *partial;    // [block][27][cinP][coutP]
	float *partialB;   // [block][coutP]
};

constexpr int kWVox = TX * TY * TZ;   // 256 voxels of a tile
constexpr int kWRow = HZ * HY * TX;   // one channel of one shifted input copy

__global__ __launch_bounds__(kUB) void kConv3WGrad(WGradArgs A)
{
	__shared__ __align__(16) bf16 tg[32 * kWVox];        // [co][voxel]
	__shared__ __align__(16) bf16 tin[3 * 16 * kWRow];   // [dx][ci][z][y][x]
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const int kg = lane >> 4, li = lane & 15;
	const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TY;
	const int nzt = (A.nz + TZ - 1) / TZ;
	const int zt = blockIdx.z % nzt, grp = blockIdx.z / nzt;
	const int z0 = zt * TZ, co0 = grp * 32;
	const int NF = min(32, A.coutP - co0) / 16;
	const int block = (zt * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
	// output-gradient tile, transposed: 8 channels of a voxel per load
	for (int q = tid; q < kWVox * 4; q += kUB) {
		const int v = q >> 2, c8 = (q & 3) * 8;
		const int vx = v % TX, vy = (v / TX) % TY, vz = v / (TX * TY);
		const int gx = x0 + vx, gy = y0 + vy, gzz = z0 + vz;
		bf16 val[8];
#pragma unroll
		for (int j = 0; j < 8; j++) val[j] = 0;
		if (gx < A.nx && gy < A.ny && gzz < A.nz) {
			const bf16 *src = A.gz + (size_t(gzz * A.ny + gy) * A.nx + gx) * A.gStride + A.gOff + co0 + c8;
			if (c8 + 8 <= NF * 16 && co0 + c8 + 8 <= A.Cout && ((A.gStride | (A.gOff + co0 + c8)) & 7) == 0) {
				const uint4 w = *reinterpret_cast<const uint4 *>(src);
				__builtin_memcpy(val, &w, 16);
			} else {
#pragma unroll
				for (int j = 0; j < 8; j++)
					if (c8 + j < NF * 16 && co0 + c8 + j < A.Cout) val[j] = src[j];
			}
		}
#pragma unroll
		for (int j = 0; j < 8; j++) tg[(c8 + j) * kWVox + v] = val[j];
	}
	__syncthreads();
	if (tid < NF * 16) {
		float sb = 0;
		for (int v = 0; v < kWVox; v++) sb += bf2f(tg[tid * kWVox + v]);
		A.partialB[size_t(block) * A.coutP + co0 + tid] = sb;
	}
	// the B fragments of the tile: K step s covers voxels 32 s .. 32 s + 31
	bf8v bq[kWVox / 32][2];
#pragma unroll
	for (int st = 0; st < kWVox / 32; st++)
#pragma unroll
		for (int f = 0; f < 2; f++)
			bq[st][f] = __builtin_bit_cast(bf8v, *reinterpret_cast<const s8v *>(tg + (f * 16 + li) * kWVox + st * 32 + kg * 8));
	for (int ci0 = 0; ci0 < A.cinP; ci0 += 16) {
		__syncthreads();
		// input halo (18 x 6 x 6 voxels, 4 channels per load) into the three
		// shifted copies
		for (int q = tid; q < HZ * HY * HX * 4; q += kUB) {
			const int quad = q & 3, v = q >> 2;
			const int hx = v % HX, hy = (v / HX) % HY, hz = v / (HX * HY);
			const int gx = x0 + hx - 1, gy = y0 + hy - 1, gzz = z0 + hz - 1;
			const int ch = ci0 + quad * 4;
			bf16 val[4] = {0, 0, 0, 0};
			if (gx >= 0 && gx < A.nx && gy >= 0 && gy < A.ny && gzz >= 0 && gzz < A.nz && ch < A.Cin) {
				const uint2 w = *reinterpret_cast<const uint2 *>(A.x + (size_t(gzz * A.ny + gy) * A.nx + gx) * A.xStride + ch);
				__builtin_memcpy(val, &w, 8);
			}
#pragma unroll
			for (int dx = 0; dx < 3; dx++) {
				const int x = hx - dx;
				if (x < 0 || x >= TX) continue;
#pragma unroll
				for (int j = 0; j < 4; j++) tin[((dx * 16 + quad * 4 + j) * HZ + hz) * HY * TX + hy * TX + x] = val[j];
			}
		}
		__syncthreads();
		for (int t = wave; t < 27; t += 4) {
			const int dz = t / 9, dy = (t / 3) % 3, dx = t % 3;
			f4 acc[2] = {f4{0, 0, 0, 0}, f4{0, 0, 0, 0}};
#pragma unroll
			for (int st = 0; st < kWVox / 32; st++) {
				// voxels 32 st + 8 kg .. + 7: one x run of row (vy, vz)
				const int v = st * 32 + kg * 8, vx = v % TX, vy = (v / TX) % TY, vz = v / (TX * TY);
				const bf8v a = __builtin_bit_cast(bf8v, *reinterpret_cast<const s8v *>(tin + ((dx * 16 + li) * HZ + vz + dz) * HY * TX + (vy + dy) * TX + vx));
				acc[0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bq[st][0], acc[0], 0, 0, 0);
				if (NF > 1) acc[1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bq[st][1], acc[1], 0, 0, 0);
			}
			float *dst = A.partial + (size_t(block) * 27 + t) * A.cinP * A.coutP;
			for (int q = 0; q < NF; q++)
#pragma unroll
				for (int i = 0; i < 4; i++) dst[size_t(ci0 + kg * 4 + i) * A.coutP + co0 + q * 16 + li] = acc[q][i];
		}
	}
}

// Deterministic sum of the partials into the FP32 gradients [27][cin][cout].
// 64 outputs per workgroup of 1024 threads: 16 groups of 64 lanes split the
// workgroups' partials (group g takes partials g, g + 16, ...), then the 16
// group sums are added in group order -- a fixed order, so the gradients
// repeat bit for bit, with 16x the parallelism of one thread per output.
constexpr int kRedGroups = 16;

__global__ __launch_bounds__(64 * kRedGroups) void kReduceWGrad(const float *partial, const float *partialB, int nblocks, int cin, int cout, int cinP, int coutP, float *gw, float *gb)
{
	__shared__ float part[kRedGroups][64];
	const int n = 27 * cin * cout;
	const int lane = int(threadIdx.x) & 63, grp = int(threadIdx.x) >> 6;
	const int e = int(blockIdx.x) * 64 + lane;
	float s = 0;
	if (e < n) {
		const int co = e % cout, ci = (e / cout) % cin, t = e / (cout * cin);
		const size_t off = (size_t(t) * cinP + ci) * coutP + co, stride = size_t(27) * cinP * coutP;
#pragma unroll 4
		for (int b = grp; b < nblocks; b += kRedGroups) s += partial[size_t(b) * stride + off];
	} else if (e < n + cout) {
		const int co = e - n;
#pragma unroll 4
		for (int b = grp; b < nblocks; b += kRedGroups) s += partialB[size_t(b) * coutP + co];
	}
	part[grp][lane] = s;
	__syncthreads();
	if (grp == 0 && e < n + cout) {
		float t = 0;
		for (int g = 0; g < kRedGroups; g++) t += part[g][lane];
		if (e < n) gw[e] = t;
		else gb[e - n] = t;
	}
}

// ------------------------------------------------------ elementwise pieces
__device__ __forceinline__ void coords(size_t v, int nx, int ny, int &x, int &y, int &z)
{
	x = int(v % nx);
	y = int((v / nx) % ny);
	z = int(v / (size_t(nx) * ny));
}

// 2x2x2 mean (child grid n = 2 * parent)
// Eight channels per thread (16 bytes): every channel count, stride and
// offset of these tensors is a multiple of eight.
__device__ __forceinline__ void unpack8(const uint4 &u, float (&f)[8])
{
	const unsigned w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
	for (int i = 0; i < 4; i++) {
		f[2 * i] = __uint_as_float(w[i] << 16);
		f[2 * i + 1] = __uint_as_float(w[i] & 0xffff0000u);
	}
}

__device__ __forceinline__ uint4 pack8(const float (&f)[8])
{
	unsigned w[4];
#pragma unroll
	for (int i = 0; i < 4; i++) w[i] = unsigned(f2bf(f[2 * i])) | unsigned(f2bf(f[2 * i + 1])) << 16;
	return make_uint4(w[0], w[1], w[2], w[3]);
}

__device__ __forceinline__ const uint4 &vec8(const bf16 *p) { return *reinterpret_cast<const uint4 *>(p); }
__device__ __forceinline__ uint4 &vec8(bf16 *p) { return *reinterpret_cast<uint4 *>(p); }

// gradient of the mean: child += parent / 8
__global__ void kPoolBack(const bf16 *gp, int pStride, int C, bf16 *gc, int cStride, int cOff, int nx, int ny, int nz)
{
	const int G = C / 8;
	const size_t total = size_t(nx) * ny * nz * G;
	for (size_t e = blockIdx.x * size_t(blockDim.x) + threadIdx.x; e < total; e += size_t(gridDim.x) * blockDim.x) {
		const int c = int(e % G) * 8;
		int x, y, z;
		coords(e / G, nx, ny, x, y, z);
		float g[8], d[8];
		unpack8(vec8(gp + (size_t((z / 2) * (ny / 2) + y / 2) * (nx / 2) + x / 2) * pStride + c), g);
		bf16 *dst = gc + (size_t(z * ny + y) * nx + x) * cStride + cOff + c;
		unpack8(vec8(dst), d);
#pragma unroll
		for (int i = 0; i < 8; i++) d[i] += g[i] * 0.125f;
		vec8(dst) = pack8(d);
	}
}

// nearest-neighbour x2 into channels [0, C) of a wider tensor (child grid n)
__global__ void kUp(const bf16 *in, int inStride, int C, bf16 *out, int outStride, int nx, int ny, int nz)
{
	const int G = C / 8;
	const size_t total = size_t(nx) * ny * nz * G;
	for (size_t e = blockIdx.x * size_t(blockDim.x) + threadIdx.x; e < total; e += size_t(gridDim.x) * blockDim.x) {
		const int c = int(e % G) * 8;
		int x, y, z;
		coords(e / G, nx, ny, x, y, z);
		vec8(out + (size_t(z * ny + y) * nx + x) * outStride + c) = vec8(in + (size_t((z / 2) * (ny / 2) + y / 2) * (nx / 2) + x / 2) * inStride + c);
	}
}

// gradient of the copy: parent = sum of its eight children
__global__ void kUpBack(const bf16 *gc, int cStride, int C, bf16 *gp, int pStride, int nx, int ny, int nz)
{
	const int px = nx / 2, py = ny / 2, pz = nz / 2, G = C / 8;
	const size_t total = size_t(px) * py * pz * G;
	for (size_t e = blockIdx.x * size_t(blockDim.x) + threadIdx.x; e < total; e += size_t(gridDim.x) * blockDim.x) {
		const int c = int(e % G) * 8;
		int x, y, z;
		coords(e / G, px, py, x, y, z);
		float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
		for (int k = 0; k < 8; k++) {
			const int cx = 2 * x + (k & 1), cy = 2 * y + (k >> 1 & 1), cz = 2 * z + (k >> 2);
			float f[8];
			unpack8(vec8(gc + (size_t(cz * ny + cy) * nx + cx) * cStride + c), f);
#pragma unroll
			for (int i = 0; i < 8; i++) s[i] += f[i];
		}
		vec8(gp + (size_t(z * py + y) * px + x) * pStride + c) = pack8(s);
	}
}

// ReLU backward in place: g *= (a > 0)
__global__ void kReluBack(bf16 *g, const bf16 *a, int stride, int off, int C, size_t vox)
{
	const size_t total = vox * C;
	for (size_t e = blockIdx.x * size_t(blockDim.x) + threadIdx.x; e < total; e += size_t(gridDim.x) * blockDim.x) {
		const size_t idx = (e / C) * stride + off + e % C;
		if (!(bf2f(a[idx]) > 0.0f)) g[idx] = 0;
	}
}

// 1x1 output convolution 16 -> 1 times the mask (input channel 1)
__global__ void kOut(const bf16 *d1, const bf16 *in, const float *w, const float *b, float *y, size_t vox)
{
	for (size_t v = blockIdx.x * size_t(blockDim.x) + threadIdx.x; v < vox; v += size_t(gridDim.x) * blockDim.x) {
		float s = b[0];
		for (int c = 0; c < 16; c++) s += bf2f(d1[v * 16 + c]) * w[c];
		y[v] = s * bf2f(in[v * 4 + 1]);
	}
}

__global__ __launch_bounds__(kUB) void kOutBack(const float *gy, const bf16 *d1, const bf16 *in, const float *w, bf16 *gd1, float *part, size_t vox)
{
	float acc[17];
	for (int c = 0; c < 17; c++) acc[c] = 0;
	for (size_t v = blockIdx.x * size_t(blockDim.x) + threadIdx.x; v < vox; v += size_t(gridDim.x) * blockDim.x) {
		const float g = gy[v] * bf2f(in[v * 4 + 1]);
		for (int c = 0; c < 16; c++) {
			acc[c] += g * bf2f(d1[v * 16 + c]);
			gd1[v * 16 + c] = f2bf(g * w[c]);
		}
		acc[16] += g;
	}
	__shared__ float red[17][kUB];
	for (int c = 0; c < 17; c++) red[c][threadIdx.x] = acc[c];
	__syncthreads();
	if (threadIdx.x < 17) {
		float s = 0;
		for (int i = 0; i < kUB; i++) s += red[threadIdx.x][i];
		part[blockIdx.x * 17 + threadIdx.x] = s;
	}
}

__global__ void kOutBackReduce(const float *part, int nb, float *gw, float *gb)
{
	const int c = threadIdx.x;
	if (c >= 17) return;
	float s = 0;
	for (int b = 0; b < nb; b++) s += part[b * 17 + c];
	if (c < 16) gw[c] = s; else gb[0] = s;
}

__global__ void kAdam(float *w, const float *g, float *m, float *v, int n, float lr, float b1, float b2, float eps, float c1, float c2)
{
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
		const float gi = g[i];
		const float mi = b1 * m[i] + (1 - b1) * gi;
		const float vi = b2 * v[i] + (1 - b2) * gi * gi;
		m[i] = mi;
		v[i] = vi;
		w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
	}
}

// forward pack [27][coutP][cinP] and data-gradient pack [27][cinP][coutP]
// (taps mirrored: the transpose of a cross-correlation)
__global__ void kPack(const float *w, int cin, int cout, int cinP, int coutP, bf16 *wf, bf16 *wd)
{
	const int n = 27 * cinP * coutP;
	for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
		const int ci = e % cinP, co = (e / cinP) % coutP, t = e / (cinP * coutP);
		const float v = (ci < cin && co < cout) ? w[(size_t(t) * cin + ci) * cout + co] : 0.0f;
		wf[(size_t(t) * coutP + co) * cinP + ci] = f2bf(v);
		const float vd = (ci < cin && co < cout) ? w[(size_t(26 - t) * cin + ci) * cout + co] : 0.0f;
		wd[(size_t(t) * cinP + ci) * coutP + co] = f2bf(vd);
	}
}

__global__ void kZeroBf(bf16 *p, size_t n)
{
	for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) p[i] = 0;
}

// input tensor (r, mask, 0, 0) from FP32 grid arrays
__global__ void kLoadInput(const float *r, const uint8_t *mask, bf16 *in, size_t vox)
{
	for (size_t v = blockIdx.x * size_t(blockDim.x) + threadIdx.x; v < vox; v += size_t(gridDim.x) * blockDim.x) {
		const bool m = mask[v] != 0;
		in[v * 4 + 0] = f2bf(m ? r[v] : 0.0f);
		in[v * 4 + 1] = f2bf(m ? 1.0f : 0.0f);
		in[v * 4 + 2] = 0;
		in[v * 4 + 3] = 0;
	}
}

// compact residual -> box input tensor, scaled to unit RMS over the n
// unknowns (rr = |r|^2 on the device)
__global__ void kGatherBox(UnetBox B, const int *grid2mat, const double *r, const double *rr, int n, bf16 *in)
{
	const double scale = rr[0] > 0 ? sqrt(double(n) / rr[0]) : 1.0;
	const size_t vox = size_t(B.n[0]) * B.n[1] * B.n[2];
	for (size_t v = blockIdx.x * size_t(blockDim.x) + threadIdx.x; v < vox; v += size_t(gridDim.x) * blockDim.x) {
		int x, y, z;
		coords(v, B.n[0], B.n[1], x, y, z);
		const int gx = B.org[0] + x, gy = B.org[1] + y, gz = B.org[2] + z;
		int m = -1;
		if (gx >= 0 && gx < B.gn[0] && gy >= 0 && gy < B.gn[1] && gz >= 0 && gz < B.gn[2])
			m = grid2mat[(size_t(gz) * B.gn[1] + gy) * B.gn[0] + gx];
		in[v * 4 + 0] = f2bf(m >= 0 ? float(r[m] * scale) : 0.0f);
		in[v * 4 + 1] = f2bf(m >= 0 ? 1.0f : 0.0f);
		in[v * 4 + 2] = 0;
		in[v * 4 + 3] = 0;
	}
}

// box output -> compact unknowns: z += y / scale
__global__ void kScatterBox(UnetBox B, const int *grid2mat, const float *y, const double *rr, int n, double *z)
{
	const double inv = rr[0] > 0 ? sqrt(rr[0] / double(n)) : 1.0;
	const size_t vox = size_t(B.n[0]) * B.n[1] * B.n[2];
	for (size_t v = blockIdx.x * size_t(blockDim.x) + threadIdx.x; v < vox; v += size_t(gridDim.x) * blockDim.x) {
		int x, yy, zz;
		coords(v, B.n[0], B.n[1], x, yy, zz);
		const int gx = B.org[0] + x, gy = B.org[1] + yy, gz = B.org[2] + zz;
		if (gx < 0 || gx >= B.gn[0] || gy < 0 || gy >= B.gn[1] || gz < 0 || gz >= B.gn[2]) continue;
		const int m = grid2mat[(size_t(gz) * B.gn[1] + gy) * B.gn[0] + gx];
		if (m >= 0) z[m] += double(y[v]) * inv;
	}
}

// The same two for r and z laid out on the flexible solve's box V (origin,
// sizes, cell order (k ny + j) nx + i; it holds every unknown); grid2mat
// still tells the unknowns.
__device__ __forceinline__ size_t vecBoxIndex(const VecBox &V, int gx, int gy, int gz)
{
	return (size_t(gz - V.org[2]) * V.n[1] + size_t(gy - V.org[1])) * V.n[0] + size_t(gx - V.org[0]);
}

__global__ void kGatherVecBox(UnetBox B, VecBox V, const int *grid2mat, const double *r, const double *rr, int n, bf16 *in)
{
	const double scale = rr[0] > 0 ? sqrt(double(n) / rr[0]) : 1.0;
	const size_t vox = size_t(B.n[0]) * B.n[1] * B.n[2];
	for (size_t v = blockIdx.x * size_t(blockDim.x) + threadIdx.x; v < vox; v += size_t(gridDim.x) * blockDim.x) {
		int x, y, z;
		coords(v, B.n[0], B.n[1], x, y, z);
		const int gx = B.org[0] + x, gy = B.org[1] + y, gz = B.org[2] + z;
		int m = -1;
		if (gx >= 0 && gx < B.gn[0] && gy >= 0 && gy < B.gn[1] && gz >= 0 && gz < B.gn[2])
			m = grid2mat[(size_t(gz) * B.gn[1] + gy) * B.gn[0] + gx];
		in[v * 4 + 0] = f2bf(m >= 0 ? float(r[vecBoxIndex(V, gx, gy, gz)] * scale) : 0.0f);
		in[v * 4 + 1] = f2bf(m >= 0 ? 1.0f : 0.0f);
		in[v * 4 + 2] = 0;
		in[v * 4 + 3] = 0;
	}
}

__global__ void kScatterVecBox(UnetBox B, VecBox V, const int *grid2mat, const float *y, const double *rr, int n, double *z)
{
	const double inv = rr[0] > 0 ? sqrt(rr[0] / double(n)) : 1.0;
	const size_t vox = size_t(B.n[0]) * B.n[1] * B.n[2];
	for (size_t v = blockIdx.x * size_t(blockDim.x) + threadIdx.x; v < vox; v += size_t(gridDim.x) * blockDim.x) {
		int x, yy, zz;
		coords(v, B.n[0], B.n[1], x, yy, zz);
		const int gx = B.org[0] + x, gy = B.org[1] + yy, gz = B.org[2] + zz;
		if (gx < 0 || gx >= B.gn[0] || gy < 0 || gy >= B.gn[1] || gz < 0 || gz >= B.gn[2]) continue;
		if (grid2mat[(size_t(gz) * B.gn[1] + gy) * B.gn[0] + gx] >= 0) z[vecBoxIndex(V, gx, gy, gz)] += double(y[v]) * inv;
	}
}

// ------------------------------------------------------------------ host
static unsigned blocksFor(size_t n) { return unsigned(std::min<size_t>((n + kUB - 1) / kUB, 8192)); }

struct LayerIO {
	const bf16 *in;
	int inStride;
	bf16 *out;
	int outStride, outOff;
	int level;
};

static const int kCin[kUnetLayers] = {4, 16, 16, 32, 32, 64, 96, 32, 48, 16};
static const int kCout[kUnetLayers] = {16, 16, 32, 32, 64, 64, 32, 32, 16, 16};
static const int kLevel[kUnetLayers] = {0, 0, 1, 1, 2, 2, 1, 1, 0, 0};

UnetBuffers::~UnetBuffers()
{
	for (void *p : owned) (void)hipFree(p);
}

bool UnetBuffers::reserve(size_t maxVox, bool train, std::string &err)
{
	if (maxVox <= cap_vox && (!train || has_grad)) return true;
	for (void *p : owned) (void)hipFree(p);
	owned.clear();
	maxVox = std::max(maxVox, cap_vox);
	const size_t v0 = maxVox, v1 = maxVox / 8 + 1, v2 = maxVox / 64 + 1;
	bool ok = true;
	auto A = [&](auto *&p, size_t count) {
		void *q = nullptr;
		if (ok && hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(*p)) == hipSuccess) { owned.push_back(q); p = static_cast<std::remove_reference_t<decltype(p)>>(q); }
		else ok = false;
	};
	A(in, v0 * 4); A(a1, v0 * 16); A(c1, v0 * 48); A(a5, v0 * 16); A(d1, v0 * 16);
	A(p1, v1 * 16); A(a2, v1 * 32); A(c2, v1 * 96); A(a4, v1 * 32); A(d2, v1 * 32);
	A(p2, v2 * 32); A(a3, v2 * 64); A(b3, v2 * 64);
	A(y, v0);
	if (train) {
		A(gy, v0);
		A(g_a1, v0 * 16); A(g_c1, v0 * 48); A(g_a5, v0 * 16); A(g_d1, v0 * 16);
		A(g_p1, v1 * 16); A(g_a2, v1 * 32); A(g_c2, v1 * 96); A(g_a4, v1 * 32); A(g_d2, v1 * 32);
		A(g_p2, v2 * 32); A(g_a3, v2 * 64); A(g_b3, v2 * 64);
		// weight-gradient partials: one per 256-voxel tile (tiles of a level
		// hold at most vox / 256 + faces of partial tiles; bound generously)
		size_t cap = 0;
		for (int i = 0; i < kUnetLayers; i++) {
			const size_t vl = kLevel[i] == 0 ? v0 : kLevel[i] == 1 ? v1 : v2;
			const size_t blocks = vl / 64 + 64;
			cap = std::max(cap, blocks * (27 * roundUp16(kCin[i]) * roundUp16(kCout[i]) + roundUp16(kCout[i])));
		}
		A(partial, cap);
		partial_cap = cap;
		A(red, 17 * 1024);
	} else {
		gy = nullptr; partial = nullptr; red = nullptr; partial_cap = 0;
		g_a1 = g_c1 = g_a5 = g_d1 = g_p1 = g_a2 = g_c2 = g_a4 = g_d2 = g_p2 = g_a3 = g_b3 = nullptr;
	}
	if (!ok) { err = "network buffer allocation failed"; cap_vox = 0; has_grad = false; return false; }
	cap_vox = maxVox;
	has_grad = train;
	return true;
}

bool UnetBuffers::setDims(int nx, int ny, int nz, std::string &err)
{
	if (nx % 4 || ny % 4 || nz % 4 || nx < 4 || ny < 4 || nz < 4) { err = "network grid sizes must be positive multiples of 4"; return false; }
	if (size_t(nx) * ny * nz > cap_vox) { err = "network grid exceeds the reserved capacity"; return false; }
	for (int l = 0; l < 3; l++) {
		n[l][0] = nx >> l; n[l][1] = ny >> l; n[l][2] = nz >> l;
		vox[l] = size_t(n[l][0]) * n[l][1] * n[l][2];
	}
	return true;
}

Unet::~Unet()
{
	for (void *p : owned) (void)hipFree(p);
}

size_t unetParamCount()
{
	size_t n = 0;
	for (int i = 0; i < kUnetLayers; i++) n += size_t(27) * kCin[i] * kCout[i] + kCout[i];
	return n + 16 + 1;
}

bool unetCreate(Unet &U, unsigned long long seed, hipStream_t s, std::string &err)
{
	U.s = s;
	bool ok = true;
	auto A = [&](float *&p, size_t count) {
		void *q = nullptr;
		if (ok && hipMalloc(&q, count * sizeof(float)) == hipSuccess) { U.owned.push_back(q); p = static_cast<float *>(q); (void)hipMemset(q, 0, count * sizeof(float)); }
		else ok = false;
	};
	auto Ab = [&](bf16 *&p, size_t count) {
		void *q = nullptr;
		if (ok && hipMalloc(&q, count * sizeof(bf16)) == hipSuccess) { U.owned.push_back(q); p = static_cast<bf16 *>(q); }
		else ok = false;
	};
	for (int i = 0; i < kUnetLayers; i++) {
		UnetLayer &L = U.L[i];
		L.cin = kCin[i]; L.cout = kCout[i]; L.level = kLevel[i];
		const size_t nw = size_t(27) * L.cin * L.cout;
		A(L.w, nw); A(L.gw, nw); A(L.mw, nw); A(L.vw, nw);
		A(L.b, L.cout); A(L.gb, L.cout); A(L.mb, L.cout); A(L.vb, L.cout);
		const size_t np = size_t(27) * roundUp16(L.cin) * roundUp16(L.cout);
		Ab(L.wf, np); Ab(L.wd, np);
	}
	A(U.ow, 16); A(U.gow, 16); A(U.mow, 16); A(U.vow, 16);
	A(U.ob, 1); A(U.gob, 1); A(U.mob, 1); A(U.vob, 1);
	if (!ok) { err = "network parameter allocation failed"; return false; }
	// He-normal initialisation (std sqrt(2 / fan_in)), zero biases, seeded
	std::mt19937_64 rng(seed);
	std::vector<float> flat(unetParamCount(), 0.0f);
	size_t off = 0;
	for (int i = 0; i < kUnetLayers; i++) {
		const int fanIn = 27 * (i == 0 ? 2 : kCin[i]);
		std::normal_distribution<double> nd(0.0, std::sqrt(2.0 / fanIn));
		for (int t = 0; t < 27; t++)
			for (int ci = 0; ci < kCin[i]; ci++)
				for (int co = 0; co < kCout[i]; co++) {
					const double v = nd(rng);
					flat[off + (size_t(t) * kCin[i] + ci) * kCout[i] + co] = (i == 0 && ci >= 2) ? 0.0f : float(v);
				}
		off += size_t(27) * kCin[i] * kCout[i] + kCout[i];
	}
	std::normal_distribution<double> nd(0.0, std::sqrt(2.0 / 16));
	for (int c = 0; c < 16; c++) flat[off + c] = float(nd(rng));
	return unetSetParams(U, flat.data(), err);
}

// flat order: per layer w [27][cin][cout] then b [cout]; then out.w [16], out.b [1]
bool unetSetParams(Unet &U, const float *flat, std::string &err)
{
	size_t off = 0;
	for (int i = 0; i < kUnetLayers; i++) {
		UnetLayer &L = U.L[i];
		const size_t nw = size_t(27) * L.cin * L.cout;
		if (hipMemcpy(L.w, flat + off, nw * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
			hipMemcpy(L.b, flat + off + nw, L.cout * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) { err = "parameter upload failed"; return false; }
		off += nw + L.cout;
	}
	if (hipMemcpy(U.ow, flat + off, 16 * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
		hipMemcpy(U.ob, flat + off + 16, sizeof(float), hipMemcpyHostToDevice) != hipSuccess) { err = "parameter upload failed"; return false; }
	U.adam_t = 0;
	unetPackWeights(U, U.s);
	return hipStreamSynchronize(U.s) == hipSuccess;
}

bool unetGetParams(Unet &U, float *flat, bool grads, std::string &err)
{
	if (hipStreamSynchronize(U.s) != hipSuccess) { err = "synchronize failed"; return false; }
	size_t off = 0;
	for (int i = 0; i < kUnetLayers; i++) {
		UnetLayer &L = U.L[i];
		const size_t nw = size_t(27) * L.cin * L.cout;
		if (hipMemcpy(flat + off, grads ? L.gw : L.w, nw * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess ||
			hipMemcpy(flat + off + nw, grads ? L.gb : L.b, L.cout * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) { err = "parameter download failed"; return false; }
		off += nw + L.cout;
	}
	if (hipMemcpy(flat + off, grads ? U.gow : U.ow, 16 * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess ||
		hipMemcpy(flat + off + 16, grads ? U.gob : U.ob, sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) { err = "parameter download failed"; return false; }
	return true;
}

void unetPackWeights(Unet &U, hipStream_t s)
{
	for (int i = 0; i < kUnetLayers; i++) {
		UnetLayer &L = U.L[i];
		hipLaunchKernelGGL(kPack, dim3(blocksFor(size_t(27) * roundUp16(L.cin) * roundUp16(L.cout))), dim3(kUB), 0, s,
			(const float *)L.w, L.cin, L.cout, roundUp16(L.cin), roundUp16(L.cout), L.wf, L.wd);
	}
}

static void conv(const UnetBuffers &B, const UnetLayer &L, int level, const bf16 *in, int inStride, int cin, const bf16 *w, int cinP, int coutP,
	const float *bias, bf16 *out, int outStride, int outOff, int cout, bool relu, hipStream_t s, const bf16 *up = nullptr, int upStride = 0, int upC = 0,
	bf16 *pool = nullptr, int poolStride = 0)
{
	ConvArgs A{};
	A.in = in; A.inStride = inStride; A.Cin = cin;
	A.up = up; A.upStride = upStride; A.upC = upC;
	A.pool = pool; A.poolStride = poolStride;
	A.w = w; A.cinP = cinP; A.coutP = coutP;
	A.bias = bias;
	A.out = out; A.outStride = outStride; A.outOff = outOff; A.Cout = cout;
	A.nx = B.n[level][0]; A.ny = B.n[level][1]; A.nz = B.n[level][2];
	A.relu = relu ? 1 : 0;
	const bool wide = (inStride & 7) == 0 && (cin & 7) == 0 && (!up || ((upStride & 7) == 0 && (upC & 7) == 0));
	if (coutP == 16) {
		constexpr int RZ1 = RowsTile<1, kRowsWZ1>::RZ;
		const dim3 g((A.nx + RX - 1) / RX, (A.ny + RY - 1) / RY, (A.nz + RZ1 - 1) / RZ1);
		if (wide) hipLaunchKernelGGL((kConv3Rows<true, 1, kRowsWZ1, kRowsOcc1>), g, dim3(kUB), 0, s, A);
		else hipLaunchKernelGGL((kConv3Rows<false, 1, kRowsWZ1, kRowsOcc1>), g, dim3(kUB), 0, s, A);
		return;
	}
	const int groups = (coutP + 31) / 32;
	if (coutP % 32 == 0) {
		constexpr int RZ2 = RowsTile<2, kRowsWZ2>::RZ;
		const dim3 g((A.nx + RX - 1) / RX, (A.ny + RY - 1) / RY, ((A.nz + RZ2 - 1) / RZ2) * groups);
		if (wide) hipLaunchKernelGGL((kConv3Rows<true, 2, kRowsWZ2, kRowsOcc2>), g, dim3(kUB), 0, s, A);
		else hipLaunchKernelGGL((kConv3Rows<false, 2, kRowsWZ2, kRowsOcc2>), g, dim3(kUB), 0, s, A);
		return;
	}
	const dim3 g((A.nx + CX - 1) / CX, (A.ny + TY - 1) / TY, ((A.nz + TZ - 1) / TZ) * groups);
	const size_t lds = size_t(HZ * HY * CHX * 16 + 27 * std::min(32, coutP) * 16) * sizeof(bf16);
	if (wide) hipLaunchKernelGGL(kConv3<true>, g, dim3(kUB), lds, s, A);
	else hipLaunchKernelGGL(kConv3<false>, g, dim3(kUB), lds, s, A);
	(void)L;
}

// forward layer table: input tensor / stride, output tensor / stride / channel offset
static void layerIO(UnetBuffers &B, int i, LayerIO &io)
{
	switch (i) {
	case 0: io = {B.in, 4, B.a1, 16, 0, 0}; break;
	case 1: io = {B.a1, 16, B.c1, 48, 32, 0}; break;
	case 2: io = {B.p1, 16, B.a2, 32, 0, 1}; break;
	case 3: io = {B.a2, 32, B.c2, 96, 64, 1}; break;
	case 4: io = {B.p2, 32, B.a3, 64, 0, 2}; break;
	case 5: io = {B.a3, 64, B.b3, 64, 0, 2}; break;
	case 6: io = {B.c2, 96, B.a4, 32, 0, 1}; break;
	case 7: io = {B.a4, 32, B.d2, 32, 0, 1}; break;
	case 8: io = {B.c1, 48, B.a5, 16, 0, 0}; break;
	default: io = {B.a5, 16, B.d1, 16, 0, 0}; break;
	}
}

// gradient tensors mirroring layerIO: gradient of the layer output and of its input
static void layerGrad(UnetBuffers &B, int i, bf16 *&gOut, bf16 *&gIn)
{
	switch (i) {
	case 0: gOut = B.g_a1; gIn = nullptr; break;
	case 1: gOut = B.g_c1; gIn = B.g_a1; break;
	case 2: gOut = B.g_a2; gIn = B.g_p1; break;
	case 3: gOut = B.g_c2; gIn = B.g_a2; break;
	case 4: gOut = B.g_a3; gIn = B.g_p2; break;
	case 5: gOut = B.g_b3; gIn = B.g_a3; break;
	case 6: gOut = B.g_a4; gIn = B.g_c2; break;
	case 7: gOut = B.g_d2; gIn = B.g_a4; break;
	case 8: gOut = B.g_a5; gIn = B.g_c1; break;
	default: gOut = B.g_d1; gIn = B.g_a5; break;
	}
}

// The decoder's inputs c2 = [up(b3) | skip] and c1 = [up(d2) | skip]: the
// forward pass reads the upsampled channels straight from b3 and d2 in the
// convolutions' staging (the same values, so the same results), and
// unetBackward writes them into c2 and c1 for the weight gradients.
static void materializeUp(UnetBuffers &B, hipStream_t s)
{
	const int *n0 = B.n[0], *n1 = B.n[1];
	hipLaunchKernelGGL(kUp, dim3(blocksFor(B.vox[1] * 8)), dim3(kUB), 0, s, (const bf16 *)B.b3, 64, 64, B.c2, 96, n1[0], n1[1], n1[2]);
	hipLaunchKernelGGL(kUp, dim3(blocksFor(B.vox[0] * 4)), dim3(kUB), 0, s, (const bf16 *)B.d2, 32, 32, B.c1, 48, n0[0], n0[1], n0[2]);
}

void unetForward(Unet &U, hipStream_t s)
{
	UnetBuffers &B = U.buf;
	auto run = [&](int i, const bf16 *up = nullptr, int upStride = 0, int upC = 0, bf16 *pool = nullptr, int poolStride = 0) {
		LayerIO io;
		layerIO(B, i, io);
		const UnetLayer &L = U.L[i];
		conv(B, L, io.level, io.in, io.inStride, L.cin, L.wf, roundUp16(L.cin), roundUp16(L.cout), L.b, io.out, io.outStride, io.outOff, L.cout, true, s,
			up, upStride, upC, pool, poolStride);
	};
	// the encoder's pooling happens in the epilogues of layers 1 and 3
	run(0);
	run(1, nullptr, 0, 0, B.p1, 16);
	run(2);
	run(3, nullptr, 0, 0, B.p2, 32);
	run(4);
	run(5);
	run(6, B.b3, 64, 64);
	run(7);
	run(8, B.d2, 32, 32);
	run(9);
	hipLaunchKernelGGL(kOut, dim3(blocksFor(B.vox[0])), dim3(kUB), 0, s, (const bf16 *)B.d1, (const bf16 *)B.in, (const float *)U.ow,
		(const float *)U.ob, B.y, B.vox[0]);
}

// exact size of the weight-gradient partials for the current dimensions
static size_t partialNeed(const UnetBuffers &B)
{
	size_t need = 0;
	for (int i = 0; i < kUnetLayers; i++) {
		const int *n = B.n[kLevel[i]];
		const size_t tiles = size_t((n[0] + TX - 1) / TX) * ((n[1] + TY - 1) / TY) * ((n[2] + TZ - 1) / TZ);
		need = std::max(need, tiles * (27 * roundUp16(kCin[i]) * roundUp16(kCout[i]) + roundUp16(kCout[i])));
	}
	return need;
}

bool unetBackward(Unet &U, hipStream_t s, std::string &err)
{
	UnetBuffers &B = U.buf;
	if (!B.has_grad) { err = "network buffers were reserved without gradients"; return false; }
	const size_t need = partialNeed(B);
	if (need > B.partial_cap) {
		(void)hipStreamSynchronize(s);
		auto it = std::find(B.owned.begin(), B.owned.end(), static_cast<void *>(B.partial));
		if (it != B.owned.end()) { (void)hipFree(*it); B.owned.erase(it); }
		void *q = nullptr;
		if (hipMalloc(&q, need * sizeof(float)) != hipSuccess) { err = "gradient partial allocation failed"; B.partial = nullptr; B.partial_cap = 0; return false; }
		B.owned.push_back(q);
		B.partial = static_cast<float *>(q);
		B.partial_cap = need;
	}
	const int *n0 = B.n[0], *n1 = B.n[1];
	materializeUp(B, s);
	// output layer
	const int nbOut = int(std::min<size_t>((B.vox[0] + kUB - 1) / kUB, 1024));
	hipLaunchKernelGGL(kOutBack, dim3(nbOut), dim3(kUB), 0, s, (const float *)B.gy, (const bf16 *)B.d1, (const bf16 *)B.in,
		(const float *)U.ow, B.g_d1, B.red, B.vox[0]);
	hipLaunchKernelGGL(kOutBackReduce, dim3(1), dim3(32), 0, s, (const float *)B.red, nbOut, U.gow, U.gob);

	auto back = [&](int i) {
		const UnetLayer &L = U.L[i];
		LayerIO io;
		layerIO(B, i, io);
		bf16 *gOut, *gIn;
		layerGrad(B, i, gOut, gIn);
		const int lv = io.level;
		// dL/dz = dL/da * (a > 0), in place
		hipLaunchKernelGGL(kReluBack, dim3(blocksFor(B.vox[lv] * L.cout)), dim3(kUB), 0, s, gOut, (const bf16 *)io.out, io.outStride, io.outOff,
			L.cout, B.vox[lv]);
		// weight and bias gradients
		WGradArgs W{};
		W.x = io.in; W.xStride = io.inStride; W.Cin = L.cin;
		W.gz = gOut; W.gStride = io.outStride; W.gOff = io.outOff; W.Cout = L.cout;
		W.cinP = roundUp16(L.cin); W.coutP = roundUp16(L.cout);
		W.nx = B.n[lv][0]; W.ny = B.n[lv][1]; W.nz = B.n[lv][2];
		const int tx = (W.nx + TX - 1) / TX, ty = (W.ny + TY - 1) / TY, tz = (W.nz + TZ - 1) / TZ;
		const int nblocks = tx * ty * tz;
		W.partial = B.partial;
		W.partialB = B.partial + size_t(nblocks) * 27 * W.cinP * W.coutP;
		const int groups = (W.coutP + 31) / 32;
		hipLaunchKernelGGL(kConv3WGrad, dim3(tx, ty, tz * groups), dim3(kUB), 0, s, W);
		hipLaunchKernelGGL(kReduceWGrad, dim3(unsigned((size_t(27) * L.cin * L.cout + L.cout + 63) / 64)), dim3(64 * kRedGroups), 0, s, (const float *)W.partial,
			(const float *)W.partialB, nblocks, L.cin, L.cout, W.cinP, W.coutP, L.gw, L.gb);
		// input gradient: the transposed convolution (mirrored taps, swapped channels)
		if (gIn) {
			const bf16 *gz = gOut + io.outOff;
			conv(B, L, lv, gz, io.outStride, L.cout, L.wd, roundUp16(L.cout), roundUp16(L.cin), nullptr, gIn, io.inStride, 0, L.cin, false, s);
		}
	};
	back(9);
	back(8);
	hipLaunchKernelGGL(kUpBack, dim3(blocksFor(B.vox[1] * 4)), dim3(kUB), 0, s, (const bf16 *)B.g_c1, 48, 32, B.g_d2, 32, n0[0], n0[1], n0[2]);
	back(7);
	back(6);
	hipLaunchKernelGGL(kUpBack, dim3(blocksFor(B.vox[2] * 8)), dim3(kUB), 0, s, (const bf16 *)B.g_c2, 96, 64, B.g_b3, 64, n1[0], n1[1], n1[2]);
	back(5);
	back(4);
	hipLaunchKernelGGL(kPoolBack, dim3(blocksFor(B.vox[1] * 4)), dim3(kUB), 0, s, (const bf16 *)B.g_p2, 32, 32, B.g_c2, 96, 64, n1[0], n1[1], n1[2]);
	back(3);
	back(2);
	hipLaunchKernelGGL(kPoolBack, dim3(blocksFor(B.vox[0] * 2)), dim3(kUB), 0, s, (const bf16 *)B.g_p1, 16, 16, B.g_c1, 48, 32, n0[0], n0[1], n0[2]);
	back(1);
	back(0);
	return true;
}

void unetAdam(Unet &U, float lr, float b1, float b2, float eps, hipStream_t s)
{
	U.adam_t++;
	const float c1 = 1 - std::pow(b1, float(U.adam_t)), c2 = 1 - std::pow(b2, float(U.adam_t));
	for (int i = 0; i < kUnetLayers; i++) {
		UnetLayer &L = U.L[i];
		const int nw = 27 * L.cin * L.cout;
		hipLaunchKernelGGL(kAdam, dim3(blocksFor(nw)), dim3(kUB), 0, s, L.w, (const float *)L.gw, L.mw, L.vw, nw, lr, b1, b2, eps, c1, c2);
		hipLaunchKernelGGL(kAdam, dim3(1), dim3(kUB), 0, s, L.b, (const float *)L.gb, L.mb, L.vb, L.cout, lr, b1, b2, eps, c1, c2);
	}
	hipLaunchKernelGGL(kAdam, dim3(1), dim3(kUB), 0, s, U.ow, (const float *)U.gow, U.mow, U.vow, 16, lr, b1, b2, eps, c1, c2);
	hipLaunchKernelGGL(kAdam, dim3(1), dim3(kUB), 0, s, U.ob, (const float *)U.gob, U.mob, U.vob, 1, lr, b1, b2, eps, c1, c2);
	unetPackWeights(U, s);
}

void unetLoadInput(Unet &U, const float *r, const uint8_t *mask, hipStream_t s)
{
	hipLaunchKernelGGL(kLoadInput, dim3(blocksFor(U.buf.vox[0])), dim3(kUB), 0, s, r, mask, U.buf.in, U.buf.vox[0]);
}

void unetApplyCompact(Unet &U, const UnetBox &B, const int *grid2mat, const double *r, const double *rr, int n, double *z, hipStream_t s)
{
	const size_t vox = size_t(B.n[0]) * B.n[1] * B.n[2];
	hipLaunchKernelGGL(kGatherBox, dim3(blocksFor(vox)), dim3(kUB), 0, s, B, grid2mat, r, rr, n, U.buf.in);
	unetForward(U, s);
	hipLaunchKernelGGL(kScatterBox, dim3(blocksFor(vox)), dim3(kUB), 0, s, B, grid2mat, (const float *)U.buf.y, rr, n, z);
}

void unetGatherCompact(Unet &U, const UnetBox &B, const int *grid2mat, const double *r, const double *rr, int n, hipStream_t s)
{
	const size_t vox = size_t(B.n[0]) * B.n[1] * B.n[2];
	hipLaunchKernelGGL(kGatherBox, dim3(blocksFor(vox)), dim3(kUB), 0, s, B, grid2mat, r, rr, n, U.buf.in);
}

void unetApplyVecBox(Unet &U, const UnetBox &B, const VecBox &V, const int *grid2mat, const double *r, const double *rr, int n, double *z,
	hipStream_t s)
{
	const size_t vox = size_t(B.n[0]) * B.n[1] * B.n[2];
	hipLaunchKernelGGL(kGatherVecBox, dim3(blocksFor(vox)), dim3(kUB), 0, s, B, V, grid2mat, r, rr, n, U.buf.in);
	unetForward(U, s);
	hipLaunchKernelGGL(kScatterVecBox, dim3(blocksFor(vox)), dim3(kUB), 0, s, B, V, grid2mat, (const float *)U.buf.y, rr, n, z);
}

} // namespace garm
