/*
 * garm_gpu.h — C-ABI of the MI355X (gfx950) implementation of the GARM-LS
 * per-timestep grid path.
 *
 * The reference composes its step from compile-time C++ templates called by
 * LevelSetLiquid<Dim> (Source/Core/LevelSetLiquid.cpp:608-698) behind the
 * virtual Simulation interface (Common/Util/Simulation.h:21-33).  It has no
 * FFI.  This header is the seam a host program binds instead: one opaque
 * context owns every field on the device, and each entry point below replaces
 * one reference operator (cited per entry).  Host buffers cross the boundary
 * in the reference's own array-of-structures layout (the layout of its .sav
 * files, GridBasedData.h:102-103, SGridBasedData.h:54-55); the device keeps
 * structure-of-arrays.
 *
 * Conventions
 *  - Every call returns GARM_OK (0) or a negative garm_status; the reference
 *    calls std::exit(-1) instead (LinearSystem.h:26-37).  The message of the
 *    last failure is available from garm_last_error().
 *  - A context is bound to one HIP device and one HIP stream and is not
 *    thread-safe (the reference is single-threaded at this level too).
 *  - Cell arrays are x-fastest (Grid.h:40-44).  Face arrays of all axes are
 *    concatenated axis 0..Dim-1 (SGridBasedData.h:54-55).
 */
#ifndef GARM_GPU_H
#define GARM_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct garm_ctx garm_ctx;
typedef struct garm_unet garm_unet;

typedef enum garm_status {
	GARM_OK = 0,
	GARM_EINVAL = -1,    /* bad argument, size or field id */
	GARM_EHIP = -2,      /* a HIP runtime call failed */
	GARM_ENOCONV = -3,   /* the pressure solve did not reach its tolerance */
	GARM_ENOMEM = -4,    /* device allocation failed */
	GARM_ENODEV = -5     /* no usable gfx950 device */
} garm_status;

/* Field identifiers for garm_upload / garm_download.  Element type is double
 * unless stated; "D" is the dimension. */
typedef enum garm_field {
	GARM_PHI = 0,            /* _levelSet, 1/cell                              */
	GARM_PHI_GRAD = 1,       /* _levelSetGrad, D/cell                          */
	GARM_REFMAP = 2,         /* _refMap (xi), D/cell                           */
	GARM_REFMAP_GRAD = 3,    /* _refMapGrad (J = d xi / dx), D*D/cell, col-major */
	GARM_PHI_TMP = 4,        /* _tmpLevelSet, 1/cell                           */
	GARM_PHI_ORI = 5,        /* _oriLevelSet, 1/cell                           */
	GARM_PHI_REF = 6,        /* _refLevelSet, 1/cell                           */
	GARM_PHI_REF_GRAD = 7,   /* _refLevelSetGrad, D/cell                       */
	GARM_ADV_REGION = 8,     /* _advRegion, uint8 1/cell                       */
	GARM_VELOCITY = 9,       /* _velocity, faces                               */
	GARM_CELL_DIST = 10,     /* _boundary.cellDist, 1/cell                     */
	GARM_NODE_DIST = 11,     /* _boundary.nodeDist, 1/node                     */
	GARM_FRACTION = 12,      /* _boundary.fraction, faces                      */
	GARM_SOLID_VELOCITY = 13,/* _boundary.velocity, faces                      */
	GARM_NORMAL = 14,        /* _boundary.normal, faces                        */
	GARM_PRESSURE = 15,      /* Projector::_rdP scattered to cells (0 elsewhere), 1/cell */
	GARM_RHS = 16,           /* Projector::_rhs of the last solve scattered to cells
	                            (0 elsewhere), 1/cell, download only */
	GARM_UNKNOWNS = 17,      /* 1 where the last solve had an unknown, else 0, 1/cell,
	                            download only (Projector::setUnknowns) */
	GARM_NUM_FIELDS = 18
} garm_field;

/* Pressure solver selection for garm_project. */
typedef enum garm_pcg_mode {
	GARM_PCG_FP64 = 0,       /* FP64 Jacobi-preconditioned CG                         */
	GARM_PCG_MIXED = 1,      /* FP32 Jacobi-PCG refined in FP64: solution accumulated in
	                            FP64, residual replaced by the FP64 b - Ax at every 1e3
	                            reduction, FP64 reductions, FP64 convergence test */
	GARM_PCG_LEARNED = 2,    /* flexible CG preconditioned by the U-Net attached with
	                            garm_attach_preconditioner (Jacobi pre- and post-smoothing
	                            around it, its correction scaled to minimise the A-norm
	                            error); once the A-norm error stops halving per iteration
	                            (an iteration's energy decrease (r.p)^2/(p.Ap) above a
	                            quarter of the one before), fallback_mode takes over;
	                            judged by the FP64 b - Ax */
	GARM_PCG_MULTIGRID = 3   /* flexible CG preconditioned by an FP32 multigrid V-cycle
	                            (Galerkin coarsening of the assembled operator on the
	                            bounding box of the unknowns, red-black Gauss-Seidel),
	                            FP64 Krylov vectors; falls back to GARM_PCG_MIXED if
	                            the residual stops halving; judged by the FP64 b - Ax */
} garm_pcg_mode;

typedef struct garm_pcg_opts {
	int mode;                /* garm_pcg_mode */
	double tolerance;        /* relative residual ||b-Ax||/||b|| (reference: 1e-5, Projector.cpp:161) */
	int max_iterations;      /* reference: 2000 */
	int check_interval;      /* iterations between host convergence checks (>=1) */
	int fallback_mode;       /* GARM_PCG_LEARNED only: what takes over when the network
	                            stalls -- 0 or GARM_PCG_MIXED: the mixed Jacobi solver on
	                            the remaining residual; GARM_PCG_MULTIGRID: the same
	                            flexible CG continues with the multigrid V-cycle */
} garm_pcg_opts;

typedef struct garm_pcg_stats {
	int unknowns;            /* number of pressure unknowns (Projector::setUnknowns) */
	int iterations;          /* CG iterations, counted as the reference's solver does */
	int outer_iterations;    /* FP64 residual replacements (mixed) */
	int fell_back;           /* 1 if the learned solve fell back to plain PCG */
	double relative_residual;/* final ||b-Ax||/||b|| computed in FP64 */
	int preconditioner_applications; /* learned: flexible-CG iterations with the network
	                            (included in iterations) */
	double hbm_bytes;        /* algorithmic HBM bytes of the solver's iterations (each array a
	                            kernel must touch read or written once per launch; the
	                            network's own activations are excluded: it is bound by MFMA) */
	int solver_cells;        /* cells an iteration covers: the unknowns (compact rows), the
	                            unknowns' bounding box, or (3D Jacobi PCG on the box) the
	                            box cells of the column tiles that hold unknowns */
} garm_pcg_stats;

/* Precision of advection and reinitialisation. */
typedef enum garm_precision { GARM_FP64 = 0, GARM_FP32 = 1 } garm_precision;

/* LevelSetLiquid's surface-tension settings (LevelSetLiquid.h:86-91). */
typedef struct garm_surface_tension_opts {
	double coeff;            /* _surfaceTensionCoeff / _density (reference 7.28e-2 / 1e3;
	                            LevelSetLiquidBuilder::buildCase1: 30 / 1e3) */
	double delta_min;        /* band half-width in cells: delta_min + (delta_max - delta_min)
	                            * 2 / (1 + exp(elapsed / delta_crit_time)) (reference 2, 8, 1;
	                            buildCase1 5.5, 5.5) */
	double delta_max;
	double delta_crit_time;
	double tolerance;        /* capillary solve, relative residual (reference 1e-5,
	                            SurfaceTension.h:93) */
	int max_iterations;      /* reference 2000 */
	int little_drop;         /* 1: trickLittleDrop after the reprojection (the reference
	                            always runs it with surface tension) */
} garm_surface_tension_opts;

typedef struct garm_step_opts {
	int reinit_interval;     /* WENO5 reinitialisation every k steps (reference: every step, k=1);
	                            a step that restarts the reference map reinitialises first */
	int reinit_iterations;   /* reference: 30 (LevelSetLiquid.cpp:1156) */
	double reinit_cfl;       /* reference: 0.9 */
	int advect_precision;    /* garm_precision of advection, composition and reinit
	                            (computed in FP32, stored in FP64; garm_initialize's
	                            reinitialisation is always FP64) */
	garm_pcg_opts pcg;
	int restart_mode;        /* garm_restart_mode of a step whose reference map is too
	                            distorted (0, the plain branch, unless set) */
	const garm_surface_tension_opts *surface_tension;
	                         /* NULL: off (the reference's _enableSurfaceTension = false);
	                            otherwise applySurfaceTension + trickLittleDrop after the
	                            projection (LevelSetLiquid.cpp:671-675) */
} garm_step_opts;

typedef struct garm_surface_tension_stats {
	double band;             /* the surfaceTensionDelta used (cells) */
	garm_pcg_stats capillary;/* unknowns = interface faces; FP64 Jacobi PCG */
	garm_pcg_stats reproject;/* the second pressure solve */
} garm_surface_tension_stats;

/* How a step restarts the reference map (LevelSetLiquid::restart,
 * LevelSetLiquid.cpp:1285-1352). */
typedef enum garm_restart_mode {
	GARM_RESTART_PLAIN = 0,      /* the branch without Hybrid Newton (:1331-1352): phi_ref =
	                                phi, grad phi_ref by computeFirst<4> */
	GARM_RESTART_NEWTON = 1      /* the default build (RM, GA, Newton) with enable-fmm=n
	                                (:1290-1303): Hybrid Newton closest points and 20
	                                characteristic steps */
} garm_restart_mode;

typedef struct garm_step_stats {
	double t_advect_ms;      /* reference bucket tadvect  (LevelSetLiquid.h:103-109) */
	double t_restart_ms;     /* trestart: reinit + restart check + restart */
	double t_project_ms;     /* tproject */
	double t_extrap_ms;      /* textrap */
	double t_total_ms;
	int restarted;           /* 1 if the reference map was restarted this step */
	int reinitialized;
	garm_pcg_stats pcg;
	garm_surface_tension_stats surface_tension;   /* zero when off */
	double max_speed;        /* garm_max_speed of the velocity the step leaves: the next
	                            step's CFL input (LevelSetLiquid.h:129), formed on the
	                            device at the step's end so no separate call is needed */
} garm_step_stats;

/* ---- context ---------------------------------------------------------- */

/* Replaces the StaggeredGrid + LevelSetLiquid constructors
 * (StaggeredGrid.h:27-36, LevelSetLiquid.cpp:47-83) and the domain-box
 * Boundary constructor + finish (Boundary.cpp:7-35, 79-87).  res[] holds the
 * cell counts (res[2] ignored for dim 2), origin[] the grid's lower corner. */
int garm_create(garm_ctx **out, int dim, const int res[3], double dx, const double origin[3], int boundary_width);
void garm_destroy(garm_ctx *ctx);
const char *garm_last_error(const garm_ctx *ctx);
/* Last error of a failed garm_create or garm_volume_device (calls without a
 * context). */
const char *garm_create_error(void);

/* Number of elements (not bytes) of a field. */
size_t garm_field_count(const garm_ctx *ctx, int field);
/* Host -> device / device -> host in the reference's AoS layout.  n is the
 * element count and must equal garm_field_count. */
int garm_upload(garm_ctx *ctx, int field, const void *host, size_t n);
int garm_download(garm_ctx *ctx, int field, void *host, size_t n);

/* LevelSetLiquid::unionLevelSet with ImplicitBox / ImplicitSphere /
 * ImplicitPlane (LevelSetLiquid.cpp:700-707, ImplicitSurface.h:7-59,117-133). */
int garm_union_box(garm_ctx *ctx, const double min_corner[3], const double lengths[3]);
int garm_union_sphere(garm_ctx *ctx, const double center[3], double radius);
int garm_union_plane(garm_ctx *ctx, const double position[3], const double normal[3]);

/* ---- operators (each replaces the cited reference function) ------------ */

/* LevelSetLiquid::initialize (LevelSetLiquid.cpp:528-598) with WENO5
 * reinitialisation. */
int garm_initialize(garm_ctx *ctx, const garm_step_opts *opts);
/* advectFields part 1: dxi, J by RK2 + Hermite (LevelSetLiquid.cpp:739-748,
 * Advection.h:24-31,59-82,147-160, HermiteIntrpl.h). */
int garm_advect_refmap(garm_ctx *ctx, double dt);
/* advectFields part 2: cubic semi-Lagrangian velocity (LevelSetLiquid.cpp:825,
 * Advection.h:15-22,33-43,137-145). */
int garm_advect_velocity(garm_ctx *ctx, double dt);
/* generateCurrentFields (LevelSetLiquid.cpp:1218-1244). */
int garm_compose_levelset(garm_ctx *ctx);
/* Reinitialization::solve<5> on _levelSet (Reinitialization.h:20-80). */
int garm_reinit_weno5(garm_ctx *ctx, int iterations, double cfl);
/* Compute precision (garm_precision) of the four calls above and of
 * garm_time_kernel's advection / reinit kernels; FP64 by default.  Fields
 * stay FP64 in memory either way; garm_step takes its own setting from
 * garm_step_opts.advect_precision.  No counterpart in the reference. */
int garm_set_precision(garm_ctx *ctx, int precision);
/* extrapolateFieldsToCollider (LevelSetLiquid.cpp:1086-1137). */
int garm_extrapolate_collider(garm_ctx *ctx);
/* extrapolateFieldsToAir (LevelSetLiquid.cpp:1059-1084). */
int garm_extrapolate_air(garm_ctx *ctx);
/* Boundary::enforce(velocity) (Boundary.cpp:89-120). */
int garm_enforce_boundary(garm_ctx *ctx);
/* applyBodyForces, gravity only (LevelSetLiquid.cpp:834-841). */
int garm_apply_gravity(garm_ctx *ctx, double dt);
/* checkReferenceMap (LevelSetLiquid.cpp:1246-1283); *ok = its return value. */
int garm_check_refmap(garm_ctx *ctx, double threshold, double band_steps, double ratio, int *ok);
/* restart without Hybrid Newton (LevelSetLiquid.cpp:1331-1352). */
int garm_restart(garm_ctx *ctx);
/* restart of the default build with enable-fmm=n (LevelSetLiquid.cpp:1290-1303):
 * from x - phi(x) grad phi(x) (the cubic view of _levelSet) the modified Newton
 * iteration finds the closest point on the Hermite interpolant of
 * (_oriLevelSet, _levelSetGrad) for the cells next to a sign change and one
 * layer more (HybridNewton.h:119-136, 147-467; HybridNewton.cpp:64-79,
 * 103-123); the other cells take 20 characteristic RK3 steps of 0.9 dx
 * (HybridNewton.cpp:30-49).  The result becomes _refLevelSet and its
 * gradient; the reference map is reset. */
int garm_restart_newton(garm_ctx *ctx);
/* Projector::project (Projector.cpp:13-20) with matrix-free PCG in place of
 * Eigen's IC-preconditioned CG (LinearSystem.h:23-39). */
int garm_project(garm_ctx *ctx, const garm_pcg_opts *opts, garm_pcg_stats *stats);
/* LevelSetLiquid::applySurfaceTension then trickLittleDrop
 * (LevelSetLiquid.cpp:1026-1057), on the state a projection left:
 * extrapolateFieldsToAir(true) (the velocity alone) and the boundary
 * enforcement; SurfaceTension::solve (SurfaceTension.h:12-101) on the faces
 * with |phi| < band dx under the cubic view of phi and fraction < 1, its
 * implicit capillary system solved with FP64 Jacobi PCG in place of Eigen's
 * IC-preconditioned CG; Projector::reproject (Projector.cpp:22-28) with
 * `reproject` -- the same unknowns and operator as the projection, the
 * right-hand side of the new velocity, solved from zero where the
 * reference starts from the projection's pressure.  The band uses the
 * context's elapsed time (advanced by garm_step as the reference's
 * _elapsedTime, LevelSetLiquid.cpp:619). */
int garm_apply_surface_tension(garm_ctx *ctx, double dt, const garm_surface_tension_opts *opts, const garm_pcg_opts *reproject,
	garm_surface_tension_stats *stats);
/* LevelSetLiquid::_elapsedTime (LevelSetLiquid.h:83): garm_step adds
 * dt * pi / 2 after the advection (LevelSetLiquid.cpp:619); 0 at creation. */
int garm_set_elapsed_time(garm_ctx *ctx, double t);
int garm_get_elapsed_time(garm_ctx *ctx, double *t);
/* LevelSetLiquid::advance (LevelSetLiquid.cpp:608-698); step_index selects
 * the reinitialisation cadence (reinit when step_index % reinit_interval == 0). */
int garm_step(garm_ctx *ctx, double dt, int step_index, const garm_step_opts *opts, garm_step_stats *stats);
/* b = A x with the pressure operator of the last garm_project / garm_step
 * (Projector::buildLinearSystem's matrix, Projector.cpp:30-68) applied to a
 * cell field x (N doubles, values at non-unknown cells ignored); b is 0 at
 * non-unknown cells.  Used to make (A x, x) training pairs. */
int garm_apply_pressure_system(garm_ctx *ctx, const double *x, double *b);
/* SGridBasedData::normMax of the velocity (SGridBasedData.h:48-52), used by
 * getTimeStep (LevelSetLiquid.h:129). */
int garm_max_speed(garm_ctx *ctx, double *out);
/* Analytic velocity fields of the reference's advection tests. */
typedef enum garm_velocity_field {
	GARM_VELOCITY_ENRIGHT = 0    /* Enright deformation on [0,1]^3 (TestAdvect3D.cpp:152-190):
	                                u = 2 sin^2(pi x) sin(2 pi y) sin(2 pi z) cos(pi t / T),
	                                v = -sin^2(pi y) sin(2 pi x) sin(2 pi z) cos(pi t / T),
	                                w = -sin^2(pi z) sin(2 pi x) sin(2 pi y) cos(pi t / T) */
} garm_velocity_field;

/* LevelSetLiquid::saveFrame / loadFrame (LevelSetLiquid.cpp:425-526) of the
 * default build: writes / reads velocity.sav, refMap.sav, refLevelSet.sav,
 * advRegion.sav (uint8), refMapGrad.sav and refLevelSetGrad.sav in
 * frame_dir, headerless raw arrays in the reference's layout (IO::writeArray,
 * IO.h:26; GridBasedData.h:102-103, SGridBasedData.h:54-55), so files
 * written by either program load in the other.  These fields determine the
 * next step: every other field is recomputed from them within garm_step.
 * Loading checks each file's size against the grid (GARM_EINVAL on a
 * mismatch); the boundary is the context's own, as the reference keeps its
 * Boundary and only calls finish (LevelSetLiquid.cpp:477). */
int garm_save_frame(garm_ctx *ctx, const char *frame_dir);
int garm_load_frame(garm_ctx *ctx, const char *frame_dir);

/* Contourer<3>::contour of _tmpLevelSet at 0 (Contourer.cpp:9-97) with
 * SurfaceMesh::computeNormals (SurfaceMesh.cpp:191-216), on the device: the
 * reference's vertex numbering (first cube holding an edge, cube order then
 * edge order) and positions, each cube's triangles from this library's
 * table (crossing edges paired per cube face, an ambiguous face cutting off
 * each inside corner; watertight), oriented from the liquid to the air.
 * 2D contexts: Contourer<2, true>::contour, the filled contour of the
 * reference's 2D frames (inside nodes first, then edge crossings; cell types
 * 6 and 9 joined when the cell mean is <= 0), on the host from a copy of
 * _tmpLevelSet, its triangles (counter-clockwise) from this library's table;
 * no normals.  positions: D floats per vertex; normals: 3 per vertex (3D
 * only, may be NULL in 2D); indices: 3 per triangle.  The counts are always
 * set; the arrays are filled when the capacities (in vertices and in
 * indices) suffice, else GARM_EINVAL. */
int garm_contour(garm_ctx *ctx, float *positions, float *normals, size_t vertex_capacity, unsigned *indices, size_t index_capacity,
	size_t *n_vertices, size_t *n_indices);
/* The same mesh written as the reference's frame files (LevelSetLiquid.cpp:
 * 306-347 without refinement; either path may be NULL): mesh_path in
 * SurfaceMesh::write's binary layout (SurfaceMesh.cpp:103-124, liquid.mesh:
 * uint32 vertex count, float positions (D per vertex), float normals (3D
 * only), uint32 index count, uint32 indices), obj_path as
 * SurfaceMesh::writeOBJ (SurfaceMesh.cpp:126-144, render/N.obj: "v" lines of
 * the FP64 positions printed %.15e (z = 0 in 2D), no "vn" lines -- the
 * contoured mesh's normals member is empty -- and 1-based "f" lines). */
int garm_write_mesh(garm_ctx *ctx, const char *mesh_path, const char *obj_path);

/* Sets the face velocity of a 3D context to an analytic field at time t
 * with period T, on the device (the reference's tests evaluate it per
 * sample, TestAdvect3D.cpp:152-190, 390-397). */
int garm_set_velocity_field(garm_ctx *ctx, int field, double time, double period);

/* Vertex order of the 3D volume evaluator. */
typedef enum garm_volume_order {
	GARM_VOLUME_REFERENCE = 0,   /* lexicographic stencil passed to the circular-order
	                                fraction, as LevelSetLiquid.cpp:1184-1185 does */
	GARM_VOLUME_CORRECTED = 1    /* stencil permuted to the fraction's corner order */
} garm_volume_order;

/* LevelSetLiquid::calculateVolume (LevelSetLiquid.cpp:1162-1190) on phi_tmp,
 * in FP64: 3D MC-style volume fractions (Fraction::get_mc_vol,
 * MC-style-vol-eval/fraction.hpp:510-771), 2D filled-contour area
 * (Contourer<2, true>, Contourer.cpp:9-96); `order` applies in 3D only. */
int garm_volume(garm_ctx *ctx, int order, double *out);
/* Allocation-free form of garm_volume for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  It has no counterpart in the reference
 * and needs no context: phi_dev is a device array of res[0]*res[1](*res[2])
 * FP64 cell values (x fastest), *out_dev (device memory) receives the volume
 * with cell size dx, computed exactly as garm_volume computes it (same
 * kernels, same reduction order, same bits).  The work is enqueued on
 * `stream` (a hipStream_t; NULL is the null stream) and the call returns
 * without waiting.  Partial sums live in a file-scope device array of 1024
 * doubles (8 KiB per device, held until the process ends); no call allocates
 * or frees memory.  Every call on one device shares that array, so calls
 * whose work overlaps on one device (two streams, two threads) race.
 * Limits: dim 2 or 3, 2 <= res[a] <= 4096, fewer than 2^31 cells; anything
 * else returns GARM_EINVAL with the reason in garm_create_error(). */
int garm_volume_device(int dim, const int res[3], double dx, const double *phi_dev, int order, double *out_dev, void *stream);
/* Blocks until all work queued on the context's stream is done. */
int garm_synchronize(garm_ctx *ctx);

/* ---- instrumentation (no counterpart in the reference) ----------------- */

/* Kernels that garm_time_kernel can time in isolation. */
typedef enum garm_kernel {
	GARM_K_ADVECT_REFMAP = 0,   /* GARM xi/J advection (one launch)            */
	GARM_K_ADVECT_VELOCITY = 1, /* cubic velocity advection (Dim launches)     */
	GARM_K_COMPOSE = 2,         /* phi / grad-phi composition                  */
	GARM_K_REINIT_STAGE = 3,    /* one WENO5 Runge-Kutta stage                 */
	GARM_K_PCG_SPMV = 4,        /* matrix-free Poisson operator + p.Ap partials */
	GARM_K_PCG_UPDATE = 5,      /* x, r update + |r|^2, r.z partials           */
	GARM_K_PCG32_SPMV = 6,      /* the same two in FP32 (mixed-precision inner) */
	GARM_K_PCG32_UPDATE = 7,
	GARM_K_COPY = 8,            /* plain FP64 copy of one cell field (bandwidth reference) */
	GARM_NUM_KERNELS = 9
} garm_kernel;

/* Launches kernel `which` `reps` times on the context's stream on the current
 * state, bracketed by HIP events on that stream, and returns the mean launch
 * time and the kernel's algorithmic HBM bytes per launch (each array it must
 * touch read or written once).  The state is left as a normal step would
 * leave it.  The PCG kernels need a preceding garm_project / garm_step. */
int garm_time_kernel(garm_ctx *ctx, int which, int reps, double *avg_ms, double *bytes_per_launch);

/* ---- learned preconditioner (no counterpart in the reference) ----------
 *
 * A three-level 3D U-Net (16/32/64 channels, 3x3x3 convolutions, BF16
 * weights and activations on MFMA, FP32 accumulation, FP32 master weights)
 * mapping a residual r, scaled to unit RMS over its unknowns, to an
 * approximate solution of A z = r.  Network grids are nx*ny*nz voxels,
 * x-fastest, each size a multiple of 4 in [4, 1024].  The network is bound to
 * the device current at creation and owns one HIP stream. */
int garm_unet_create(garm_unet **out, unsigned long long seed);
void garm_unet_destroy(garm_unet *net);
const char *garm_unet_last_error(const garm_unet *net);
const char *garm_unet_create_error(void);
/* Parameters as FP32, per layer w[27][cin][cout] then b[cout] (layers
 * L1a L1b L2a L2b L3a L3b L4a L4b L5a L5b), then the 1x1 output w[16], b[1]. */
size_t garm_unet_param_count(void);
int garm_unet_get_params(garm_unet *net, float *out, size_t n);
int garm_unet_set_params(garm_unet *net, const float *in, size_t n);
/* y = net(r, mask) on one grid (inputs used as given, no scaling). */
int garm_unet_forward(garm_unet *net, const int dims[3], const float *r, const uint8_t *mask, float *y);
/* Mean squared error over the mask of net(r, mask) against target and its
 * gradient with respect to every parameter (layout of garm_unet_get_params). */
int garm_unet_loss_grad(garm_unet *net, const int dims[3], const float *r, const uint8_t *mask, const float *target,
	double *loss, float *grad, size_t n);
/* Adam training on n_pairs (rhs, pressure, mask) grids stored one after the
 * other: each step draws a pair and a sign (seeded), scales the pair by the
 * sign and by the factor that gives rhs unit RMS over its mask, and takes one
 * Adam step (beta 0.9 / 0.999) on the masked squared error.  loss_history
 * (optional) receives `steps` losses. */
int garm_unet_train(garm_unet *net, const int dims[3], int n_pairs, const float *rhs, const float *pressure, const uint8_t *mask,
	int steps, float lr, unsigned long long seed, float *loss_history);
/* Average time of `reps` forward passes on a dims grid of synthetic input,
 * after one untimed pass, measured with HIP events on the network's stream,
 * and the multiply-add FLOPs of one pass (2 * 27 * cin * cout per voxel of
 * every layer's level, plus the 1x1 output). */
int garm_unet_time_forward(garm_unet *net, const int dims[3], int reps, double *avg_ms, double *flops);
/* `steps` Adam steps (beta 0.9 / 0.999) of the network on pairs made on the
 * device from the pressure system of the context's last projection: with
 * probability operator_fraction an operator pair (A e, e), A that
 * projection's matrix (Projector::buildLinearSystem, Projector.cpp:30-68) and
 * e a random smooth field on its unknowns (12 plane waves, wavelengths
 * log-uniform from 4 cells to the unknowns' extent or 32 cells, whichever is
 * smaller -- about the network's receptive field --, a random power-law
 * spectrum) -- the error a Jacobi-smoothed flexible-CG residual leaves --,
 * otherwise the projection's own (divergence, pressure) pair with a random
 * sign.  Inputs are scaled to unit RMS over the unknowns, the loss is the
 * masked squared error, the network grid is the unknowns' bounding box with
 * two cells of context.  Nothing leaves the device but loss_history
 * (optional, `steps` floats).  Seeded; no counterpart in the reference. */
int garm_unet_train_on_system(garm_unet *net, garm_ctx *ctx, int steps, float lr, double operator_fraction,
	unsigned long long seed, float *loss_history);
/* Uses `net` for GARM_PCG_LEARNED solves of this context (NULL detaches).
 * The network must outlive the attachment and live on the same device. */
int garm_attach_preconditioner(garm_ctx *ctx, garm_unet *net);

#ifdef __cplusplus
}
#endif

#endif
