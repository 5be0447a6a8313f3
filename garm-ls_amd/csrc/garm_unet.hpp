// Learned pressure preconditioner: a three-level 3D U-Net (16/32/64
// channels, 3x3x3 convolutions) with BF16 activations and weights on MFMA,
// FP32 accumulation, FP32 master weights and Adam, on gfx950.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

namespace garm {

using bf16 = uint16_t;

__host__ __device__ __forceinline__ float bf2f(bf16 h)
{
	uint32_t u = uint32_t(h) << 16;
	float f;
	std::memcpy(&f, &u, 4);
	return f;
}

// FP32 -> BF16, round to nearest even (NaN stays NaN)
__host__ __device__ __forceinline__ bf16 f2bf(float f)
{
	uint32_t u;
	std::memcpy(&u, &f, 4);
	if ((u & 0x7fffffffu) > 0x7f800000u) return bf16(0x7fc0);
	u += 0x7fffu + ((u >> 16) & 1u);
	return bf16(u >> 16);
}


constexpr int kUnetLayers = 10;   // L1a L1b L2a L2b L3a L3b L4a L4b L5a L5b

struct UnetLayer {
	int cin, cout, level;
	float *w = nullptr;      // FP32 master [27][cin][cout]
	float *b = nullptr;      // [cout]
	float *gw = nullptr, *gb = nullptr;
	float *mw = nullptr, *vw = nullptr, *mb = nullptr, *vb = nullptr;   // Adam moments
	bf16 *wf = nullptr;      // forward pack  [27][coutP][cinP]
	bf16 *wd = nullptr;      // data-gradient pack (flipped taps, swapped channels) [27][cinP][coutP]
};

// Activation buffers for one grid size (nx, ny, nz multiples of 4).
struct UnetBuffers {
	int n[3][3] = {};        // per level: nx, ny, nz
	size_t vox[3] = {};
	// forward (channels-last BF16)
	bf16 *in = nullptr, *a1 = nullptr, *c1 = nullptr, *a5 = nullptr, *d1 = nullptr;   // level 1: 4, 16, 48, 16, 16
	bf16 *p1 = nullptr, *a2 = nullptr, *c2 = nullptr, *a4 = nullptr, *d2 = nullptr;   // level 2: 16, 32, 96, 32, 32
	bf16 *p2 = nullptr, *a3 = nullptr, *b3 = nullptr;                                  // level 3: 32, 64, 64
	float *y = nullptr;       // network output, level 1
	// backward (gradients of the same tensors)
	bf16 *g_a1 = nullptr, *g_c1 = nullptr, *g_a5 = nullptr, *g_d1 = nullptr;
	bf16 *g_p1 = nullptr, *g_a2 = nullptr, *g_c2 = nullptr, *g_a4 = nullptr, *g_d2 = nullptr;
	bf16 *g_p2 = nullptr, *g_a3 = nullptr, *g_b3 = nullptr, *g_in = nullptr;
	float *gy = nullptr;
	float *partial = nullptr; // weight-gradient partials
	size_t partial_cap = 0;
	float *red = nullptr;     // small reduction scratch
	size_t cap_vox = 0;       // level-1 voxels the buffers hold
	bool has_grad = false;
	std::vector<void *> owned;
	~UnetBuffers();
	// capacity for `maxVox` level-1 voxels (gradient buffers only if `train`)
	bool reserve(size_t maxVox, bool train, std::string &err);
	// dimensions of the next passes (multiples of 4, within capacity)
	bool setDims(int nx, int ny, int nz, std::string &err);
};

struct Unet {
	UnetLayer L[kUnetLayers];
	float *ow = nullptr, *ob = nullptr, *gow = nullptr, *gob = nullptr;   // 1x1 output conv (16 -> 1)
	float *mow = nullptr, *vow = nullptr, *mob = nullptr, *vob = nullptr;
	int adam_t = 0;
	hipStream_t s = nullptr;
	UnetBuffers buf;
	std::vector<void *> owned;
	~Unet();
};

bool unetCreate(Unet &U, unsigned long long seed, hipStream_t s, std::string &err);
// Upload FP32 parameters in the oracle's order and layout; refreshes the BF16 packs.
bool unetSetParams(Unet &U, const float *flat, std::string &err);
size_t unetParamCount();
bool unetGetParams(Unet &U, float *flat, bool gradients, std::string &err);

// in: FP32 grid channels (r, mask) already placed in U.buf.in (BF16); runs
// the forward pass into U.buf.y.
void unetForward(Unet &U, hipStream_t s);
// Backward from U.buf.gy (gradient of the output); fills the layer gradients.
bool unetBackward(Unet &U, hipStream_t s, std::string &err);
void unetAdam(Unet &U, float lr, float beta1, float beta2, float eps, hipStream_t s);
void unetPackWeights(Unet &U, hipStream_t s);

// Place r (FP32 or FP64 grid, scaled by `scale`) and the mask into buf.in;
// `box` selects the sub-block of a larger grid (offset, size).
void unetLoadInput(Unet &U, const float *r, const uint8_t *mask, hipStream_t s);

// Preconditioner application over a sub-box B of a cell grid: gathers the
// residual r (n unknowns, |r|^2 in device memory at rr) into B scaled to unit
// RMS, runs the network and adds the output, scaled back, to z.  r and z are
// compact (unknowns in grid2mat's numbering), or, with unetApplyVecBox, laid
// out on a box V of the cell grid that holds every unknown (cell order
// (k ny + j) nx + i over V's sizes; grid2mat tells the unknowns).
struct UnetBox {
	int org[3];              // box origin in cell coordinates
	int n[3];                // box size (multiples of 4)
	int gn[3];               // full cell-grid size
};
struct VecBox {
	int org[3];
	int n[3];
};
void unetApplyCompact(Unet &U, const UnetBox &B, const int *grid2mat, const double *r, const double *rr, int n, double *z, hipStream_t s);
// The gather of unetApplyCompact alone: compact r scaled to unit RMS into U.buf.in.
void unetGatherCompact(Unet &U, const UnetBox &B, const int *grid2mat, const double *r, const double *rr, int n, hipStream_t s);
void unetApplyVecBox(Unet &U, const UnetBox &B, const VecBox &V, const int *grid2mat, const double *r, const double *rr, int n, double *z,
	hipStream_t s);

} // namespace garm
