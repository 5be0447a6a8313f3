// Internal declarations shared by the HIP translation units of libgarm_gpu.
#pragma once

#include "garm_common.hpp"
#include "garm_boundary.hpp"

#include <cstdint>
#include <string>
#include <vector>

namespace garm {

// Face-centred velocity: one array per axis on its own face grid
// (SGridBasedData<Dim,double>, StaggeredGrid.h:104-120).
struct FaceSet {
	const double *a[3];
	GridDev g[3];
};

constexpr int kColliderLayers = 6;   // _boundaryExtrapolationRadius (LevelSetLiquid.h:92)

// Every device-resident field of one LevelSetLiquid<Dim>
// (LevelSetLiquid.h:27-121), structure-of-arrays.
struct Fields {
	int D = 3;
	int bw = 2;
	int res[3] = {1, 1, 1};
	GridDev cell{}, node{}, face[3]{};
	size_t N = 0, Nn = 0, Nf[3] = {0, 0, 0};

	double *u[3] = {}, *u_new[3] = {};               // _velocity (+ double buffer)
	double *phi = nullptr, *phi_grad[3] = {};         // _levelSet, _levelSetGrad
	double *xi[3] = {}, *J[9] = {};                   // _refMap, _refMapGrad: J[c*D + r] = d xi_r / d x_c
	double *xi_new[3] = {}, *J_new[9] = {};
	double *phi_tmp = nullptr, *phi_ori = nullptr;    // _tmpLevelSet, _oriLevelSet
	double *phi_ref = nullptr, *phi_ref_grad[3] = {}; // _refLevelSet, _refLevelSetGrad
	double *ref_rec64 = nullptr;                      // _refLevelSetView's Hermite data, 2^D per cell (FP64)
	float *ref_rec32 = nullptr;                       // ... rounded to FP32
	void *adv_rec = nullptr;                          // advection records (velocity per face, reference map per cell)
	uint8_t *adv_region = nullptr;                    // _advRegion
	double *cell_dist = nullptr, *node_dist = nullptr;// _boundary.cellDist / nodeDist
	double *frac[3] = {}, *bvel[3] = {}, *bnormal[3] = {};   // _boundary.fraction / velocity / normal
	double *scratch[2] = {};                          // two cell arrays (reinit stages)
	uint8_t *layer = nullptr, *layer2 = nullptr;      // extrapolation layer index per cell (two fronts)
	uint8_t *layer_f[3] = {};                         // ... per face
	uint8_t *open_f[3] = {};                          // 1 where fraction < 1 (set with the boundary)
	int *front_list[6] = {};                          // extrapolation fronts (element ids): two lists per front
	int *front_cnt = nullptr;                         // three counters per front (9), one for the boundary faces, eight for the collider layers
	uint8_t *coll_layer = nullptr;                    // collider extrapolation: layer per cell (fixed by cellDist)
	int *coll_list = nullptr;                         // ... cells of layers 1..6, one run per layer
	int coll_off[kColliderLayers + 2] = {}, coll_cnt[kColliderLayers + 1] = {};   // host: run offsets and lengths
	int *bface = nullptr;                             // faces Boundary::enforce acts on (global face ids)
	double *bface_rec = nullptr;                      // per listed face: unit normal, solid velocity, result
	int nbface = 0;                                   // listed faces
	double *partials = nullptr;                       // reduction partials
	size_t partials_cap = 0;
	double *dscalars = nullptr;                       // small device scalars
	int *host_ints = nullptr;                         // pinned, 16 ints: small readbacks
	hipEvent_t wait = nullptr;                        // spinWait's event on the context's stream
};

// ---- launchers (one per kernel family; all enqueue on `s`) ---------------
void launchAdvectRefMap(const Fields &F, double dt, hipStream_t s, bool fp32 = false);
// copiesFresh: the float copy of u made by launchAdvectRefMap(fp32) of the same u is current
void launchAdvectVelocity(const Fields &F, double dt, hipStream_t s, bool fp32 = false, bool copiesFresh = false);
void launchCompose(const Fields &F, hipStream_t s, bool fp32 = false);
// Hermite records of the reference level set (after phi_ref or its gradient changes)
void launchRefLevelSetRecords(const Fields &F, hipStream_t s);
// the same records of any scalar field and its gradient (D arrays)
void launchLevelSetRecords(const Fields &F, const double *phi, const double *const *grad, double *out64, float *out32, hipStream_t s);
// LevelSetLiquid::restart of the default build: Hybrid Newton from
// (phi_ori, phi_grad) with closest-point guesses from phi, characteristic
// steps elsewhere; result in phi_ref / phi_ref_grad, reference map reset
void launchRestartNewton(const Fields &F, hipStream_t s);
void launchGradient4(const Fields &F, const double *src, double *const *dst, hipStream_t s);
void launchReinit(const Fields &F, int iterations, double cfl, hipStream_t s, bool fp32 = false);
// one Runge-Kutta stage L(phi) -> scratch[0], for timing (the FP32 stage
// reads the float copy of phi that launchReinitPrepare makes)
void launchReinitStage(const Fields &F, double dtau, hipStream_t s, bool fp32 = false);
void launchReinitPrepare(const Fields &F, hipStream_t s, bool fp32);

void launchUnionBox(const Fields &F, const double *center, const double *half, hipStream_t s);
void launchUnionSphere(const Fields &F, const double *center, double r, hipStream_t s);
void launchUnionPlane(const Fields &F, const double *pos, const double *nrm, hipStream_t s);
void launchResetRefMap(const Fields &F, hipStream_t s);
// volume of phi_tmp into F.dscalars[3] (corrected: circular corner order)
void launchVolume(const Fields &F, bool corrected, hipStream_t s);
// message of the last failed call that has no context (garm_create_error)
void setContextFreeError(const std::string &msg);
void launchFill(double *p, double v, size_t n, hipStream_t s);
void launchFillU8(uint8_t *p, uint8_t v, size_t n, hipStream_t s);
// open_f from the face fractions (after the boundary fields change)
void launchOpenFaces(const Fields &F, hipStream_t s);
void launchCopy(double *dst, const double *src, size_t n, hipStream_t s);
void launchEnforceLevelSet(const Fields &F, double *sdf, double tol, hipStream_t s);
void launchGravity(const Fields &F, double dt, hipStream_t s);
// Enright deformation times coeff on the face arrays (3D)
void launchEnrightVelocity(const Fields &F, double coeff, hipStream_t s);
void launchExtrapolateCollider(const Fields &F, hipStream_t s);
// finds the collider extrapolation's layers (after cellDist is set); synchronises
bool buildColliderLayers(Fields &F, hipStream_t s);
// Contourer<3>::contour of `field` (a cell field) at 0 with SurfaceMesh
// normals: FP64 positions and float normals (3 per vertex), uint indices
// (3 per triangle) in the reference's vertex order.
int contourMesh(const Fields &F, const double *field, std::vector<double> &pos, std::vector<float> &nrm, std::vector<unsigned> &idx,
	hipStream_t s, std::string &err);
// velocityOnly: extrapolateFieldsToAir(true), the velocity alone
void launchExtrapolateAir(const Fields &F, hipStream_t s, bool velocityOnly = false);
// Boundary::enforce on the listed faces, in place on F.u
void launchEnforceBoundary(const Fields &F, hipStream_t s);
// lists the faces with 0.1 < fraction <= 1 into ids (count in *cnt, zeroed by the caller)
void launchBoundaryFaces(const Fields &F, int *ids, int *cnt, hipStream_t s);
// unit normal and solid velocity of the F.nbface listed faces into F.bface_rec
void launchBoundaryRecords(const Fields &F, hipStream_t s);
// results land in F.dscalars: [0] = max |u|; [1] = tot, [2] = cnt (as doubles)
void launchMaxSpeed(const Fields &F, hipStream_t s);
void launchCheckRefMap(const Fields &F, double threshold, double band, hipStream_t s);

} // namespace garm
