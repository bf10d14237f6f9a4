/*
 * mgx.h -- C ABI of the MI355X-native projection solver (libmgx.so).
 *
 * This is the drop-in boundary for incflo's projection hot path.  Every entry
 * point returns an int status (MGX_OK == 0); nothing throws across the ABI.
 * A non-zero status leaves a message in mgx_ctx_last_error().  The C++ facade
 * (include/mgx_amrex.H) maps a non-zero status to an abort with that message,
 * which is what amrex::Abort does at the reference call sites
 * (src/projection/incflo_projection_bc.cpp:36,74).
 *
 * Data layout ("box-offset table" layout).  A level is a list of boxes (cell
 * index ranges lo..hi inclusive).  A field on a level is ONE device
 * allocation holding the boxes back to back, in box order; each box is an
 * AMReX-style Array4: x fastest, then y, z, then component, over the box
 * grown by the field's ghost width ng.  Cell-centred boxes span
 * (hi-lo+1+2ng) per direction, nodal boxes (hi-lo+2+2ng).  Box b starts at
 * ncomp * sum_{b'<b} volume(b').  The caller owns all field memory; the
 * handle owns the multigrid workspace.
 *
 * All device work is enqueued on the context's HIP stream; project/solve
 * synchronise only to read residual norms and to return the info struct.
 * A handle is used from one host thread at a time.
 */
#ifndef MGX_H
#define MGX_H

#ifdef __cplusplus
extern "C" {
#endif

#define MGX_API __attribute__((visibility("default")))

/* status codes */
enum {
    MGX_OK = 0,
    MGX_ERR_ARG = 1,           /* invalid argument (amrex::Abort at the reference site) */
    MGX_ERR_HIP = 2,           /* HIP runtime error                                     */
    MGX_ERR_NOT_CONVERGED = 3, /* MLMG did not reach the tolerance within maxiter       */
    MGX_ERR_STATE = 4,         /* call out of order (e.g. project before sigma)          */
    MGX_ERR_UNSUPPORTED = 5    /* valid input this build does not implement              */
};

/* boundary types: mirrors amrex::LinOpBCType as produced by
 * incflo::get_projection_bc / get_mac_projection_bc
 * (src/projection/incflo_projection_bc.cpp:6-78) */
enum {
    MGX_BC_PERIODIC = 0,
    MGX_BC_DIRICHLET = 1,
    MGX_BC_NEUMANN = 2,
    MGX_BC_INFLOW = 3,
    MGX_BC_ROBIN = 4,
    /* incflo's direction_dependent ("dd") faces
     * (src/boundary_conditions/boundary_conditions.cpp:82-87): the reference
     * maps them to nodal inflow / MAC Neumann and, because such a face lets
     * fluid both in and out, rescales the outflowing velocities with
     * HydroUtils::enforceInOutSolvability before each projection
     * (src/projection/incflo_apply_nodal_projection.cpp:166-178,
     * src/convection/incflo_compute_MAC_projected_velocities.cpp:266-270).
     * That rescaling is not implemented: set_domain_bc refuses this value
     * with MGX_ERR_UNSUPPORTED rather than return a different velocity. */
    MGX_BC_INOUT = 5
};

/* solve modes */
enum {
    MGX_MODE_FP64 = 0,   /* plain FP64 MLMG                                            */
    MGX_MODE_MIXED = 1,  /* FP32 V-cycle as the correction inside FP64 refinement      */
    MGX_MODE_LEARNED = 2 /* mixed + learned initial guess / coarse correction          */
};

typedef struct mgx_ctx_s* mgx_ctx;
typedef struct mgx_nodal_s* mgx_nodal;

/* Geometry + box list of one AMR level (amrex::Geometry + BoxArray). */
typedef struct {
    int n_cell[3];        /* domain cells                                  */
    double prob_lo[3];
    double prob_hi[3];
    int is_periodic[3];
    int nboxes;
    const int* boxes;     /* nboxes*6 ints: lo[0..2], hi[0..2] (cell index, inclusive) */
} mgx_level_desc;

/* bottom_solver (MLMG::setBottomSolver through the nodal_proj.bottom_solver /
 * mac_proj.bottom_solver / diffusion bottom_solver ParmParse strings,
 * Docs/sphinx_documentation/source/InputsMultigrid.rst:34-35,
 * src/diffusion/DiffusionScalarOp.cpp:306-314).  A Krylov bottom solve that
 * does not reach bottom_rtol within bottom_maxiter is followed by 8
 * smoothing sweeps (from zero when it broke down); "hypre" is refused with
 * MGX_ERR_UNSUPPORTED. */
enum {
    MGX_BOTTOM_DEFAULT = 0,   /* bicgcg                                        */
    MGX_BOTTOM_BICGSTAB = 1,
    MGX_BOTTOM_CG = 2,
    MGX_BOTTOM_BICGCG = 3,    /* BiCGStab, then CG from zero if it fails       */
    MGX_BOTTOM_CGBICG = 4,    /* CG, then BiCGStab from zero if it fails       */
    MGX_BOTTOM_SMOOTHER = 5,  /* 8 smoothing sweeps of the bottom level        */
    MGX_BOTTOM_HYPRE = 6      /* refused                                       */
};

/* Smoother of the nodal V-cycle's per-level sweeps.  MLNodeLaplacian's GPU
 * smoother is weighted Jacobi; the Chebyshev form runs the same sweeps with
 * the factors of the degree-nu Chebyshev polynomial for the high-frequency
 * interval of the Jacobi-scaled operator (a smaller smoothing factor per
 * pair of sweeps: 0.14 against 0.31 at constant sigma).  Default: Jacobi for
 * FP64 hierarchies (MGX_MODE_FP64, as MLMG), Chebyshev for the FP32 V-cycles
 * of the mixed and learned modes. */
enum { MGX_SMOOTHER_DEFAULT = 0, MGX_SMOOTHER_JACOBI = 1, MGX_SMOOTHER_CHEBYSHEV = 2 };

/* MLMG knobs: the nodal_proj.* / mac_proj.* ParmParse names
 * (src/incflo.H:409-445, src/setup/init.cpp:165-177,
 *  Docs/sphinx_documentation/source/InputsMultigrid.rst:10-36). */
typedef struct {
    int verbose;
    int bottom_verbose;
    int maxiter;               /* default 100 (nodal) / 200 (MAC)             */
    int bottom_maxiter;        /* default 100 (nodal) / 200 (MAC)             */
    double bottom_rtol;        /* default 1e-4                                */
    double bottom_atol;        /* default -1 (unused)                         */
    int max_coarsening_level;  /* LPInfo::setMaxCoarseningLevel, default 100  */
    int nu1;                   /* pre-smoothing sweeps, default 2             */
    int nu2;                   /* post-smoothing sweeps, default 2            */
    double omega;              /* weighted-Jacobi factor, default 8/9         */
    int stall_iters;           /* learned mode: fall back to plain FP64 MLMG
                                  after this many consecutive outer
                                  iterations that each reduced the residual
                                  by less than 10% (default 3; 0 disables)   */
    int cycles_per_update;     /* mixed / learned modes (no reference
                                  counterpart): FP32 V-cycles per FP64
                                  refinement step, 0..8; each after the first
                                  starts from the correction so far; the last
                                  step runs only the cycles the measured
                                  contraction says are left.  Honoured by the
                                  single-level nodal and MAC solvers (the
                                  AMR solvers run one composite cycle per
                                  step).  0 means the default, 3.            */
    int bottom_solver;         /* MGX_BOTTOM_*, default bicgcg                 */
    int smoother;              /* MGX_SMOOTHER_* (no reference counterpart):
                                  the per-level sweeps of the nodal V-cycle    */
    int learned_steps;         /* MGX_MODE_LEARNED (no reference counterpart):
                                  the refinement steps whose FP32 V-cycles
                                  start from the network's correction of that
                                  step's residual (kept only when it lowers
                                  the FP32 residual): 1 the first step only
                                  (the initial guess; also 0, the default),
                                  k the first k steps, -1 every step        */
} mgx_mlmg_opts;

typedef struct {
    int iters;          /* outer (V-cycle) iterations                        */
    int converged;
    int nlevels;
    int fallback_used;  /* learned mode: 1 if the plain MLMG fallback ran    */
    double bnorm;       /* ||b||_inf (after the singular-problem mean fix)   */
    double resnorm0;    /* initial ||b - L phi||_inf                         */
    double resnorm;     /* final ||b - L phi||_inf (FP64)                    */
    double rel_resid;   /* resnorm / bnorm                                   */
    int vcycles;        /* V-cycles run (iters, or more in the mixed forms
                           with cycles_per_update > 1; a learned step
                           counts as one)                                   */
} mgx_solve_info;

/* ---- library / context ------------------------------------------------ */
MGX_API const char* mgx_version(void);
/* stream: a hipStream_t (NULL = the device's null stream). */
MGX_API int mgx_ctx_create(int device, void* stream, mgx_ctx* out);
MGX_API int mgx_ctx_destroy(mgx_ctx ctx);
MGX_API const char* mgx_ctx_last_error(mgx_ctx ctx);
MGX_API int mgx_ctx_synchronize(mgx_ctx ctx);
/* Device-memory plumbing for hosts without their own arena. */
MGX_API int mgx_dev_malloc(mgx_ctx ctx, unsigned long long bytes, void** out);
MGX_API int mgx_dev_free(mgx_ctx ctx, void* p);
MGX_API int mgx_memcpy_h2d(mgx_ctx ctx, void* dst, const void* src, unsigned long long bytes);
MGX_API int mgx_memcpy_d2h(mgx_ctx ctx, void* dst, const void* src, unsigned long long bytes);
MGX_API int mgx_memset_d(mgx_ctx ctx, void* dst, int value, unsigned long long bytes);
/* stream-ordered device-to-device copy (returns without synchronising) */
MGX_API int mgx_memcpy_d2d(mgx_ctx ctx, void* dst, const void* src, unsigned long long bytes);
/* Instrumentation (no reference counterpart): a plain streaming copy kernel
 * (16-byte loads and stores; bytes a multiple of 16, 16-byte aligned
 * pointers) whose known traffic calibrates hardware byte counters. */
MGX_API int mgx_stream_copy(mgx_ctx ctx, void* dst, const void* src, unsigned long long bytes);
/* Stream-ordered timing: record event `slot` (0..15) on the context's
 * stream; elapsed milliseconds between two recorded slots. */
MGX_API int mgx_ctx_record(mgx_ctx ctx, int slot);
MGX_API int mgx_ctx_elapsed_ms(mgx_ctx ctx, int slot_a, int slot_b, double* ms);
MGX_API void mgx_default_mlmg_opts(mgx_mlmg_opts* o);

/* ---- nodal projection (Hydro::NodalProjector + MLNodeLaplacian + MLMG) -- */
/* Hydro::NodalProjector constructor, geometry part
 * (src/projection/incflo_apply_nodal_projection.cpp:187-192). */
MGX_API int mgx_nodal_create(mgx_ctx ctx, const mgx_level_desc* level, mgx_nodal* out);
MGX_API int mgx_nodal_destroy(mgx_nodal h);
/* NodalProjector::setDomainBC(lo, hi) (:194); values MGX_BC_*. */
MGX_API int mgx_nodal_set_domain_bc(mgx_nodal h, const int lo[3], const int hi[3]);
/* LPInfo / nodal_proj.* knobs (:181-182). */
MGX_API int mgx_nodal_set_options(mgx_nodal h, const mgx_mlmg_opts* opts);
/* Instrumentation (no reference counterpart; the reference's BL_PROFILE regions,
 * e.g. src/incflo.cpp:220, play this role): when on, every finest-level
 * smoother launch is bracketed by HIP events on the context's stream;
 * get_timing(which=0) returns the summed milliseconds and the launch count. */
MGX_API int mgx_nodal_set_timing(mgx_nodal h, int on);
MGX_API int mgx_nodal_get_timing(mgx_nodal h, int which, double* total_ms, long long* count);
/* Launch mode (no reference counterpart; for callers that must not use HIP
 * graphs, and for tests): on (default), each V-cycle and the learned
 * predictor's network forward are captured once as HIP graphs and replayed;
 * off, their kernels are launched one by one.  The kernels and results are
 * the same either way. */
MGX_API int mgx_nodal_set_graphs(mgx_nodal h, int on);
/* Storage of the finest level of the mixed-mode V-cycle (no reference
 * counterpart; MLMG runs in FP64 throughout): 32 (default) stores FP32; 16
 * stores x, b and sigma in FP16 with the operator and the residual scaled
 * into FP16's range and computes in FP32.  The outer residual, the update
 * and the convergence test are FP64 either way. */
MGX_API int mgx_nodal_set_vcycle_bits(mgx_nodal h, int bits);
/* Launch-shape autotuner (no reference counterpart; AMReX picks fixed GPU
 * launch sizes).  Times candidate launch shapes of the nodal stencil kernel
 * (nodes per thread 1/2/4, z-run length through a target workgroup count,
 * and for the 4-byte Jacobi sweep the variant whose plane loads go straight
 * into LDS) on every level of this handle's hierarchy for `mode` (FP64 or the FP32
 * V-cycle of the mixed forms), for the Jacobi, first-sweep and residual
 * passes and the level-0 FP64 residual / update passes, then the number of
 * coarsest levels run inside the single-workgroup bottom launch (whole
 * V-cycles timed), and keeps the fastest of each in a process-wide table
 * keyed by level shape, storage and pass.  Needs sigma (one projection or
 * set_sigma first); results of later solves are unchanged (launch shapes
 * only).  report (may be NULL) receives one text line per tuned pass with
 * the chosen shape and its time against the default's. */
MGX_API int mgx_nodal_autotune(mgx_nodal h, int mode, char* report, int report_len);
/* The tuning table as text ("stencil nn0 nn1 nn2 word cword pass fine npt
 * target kzmin dma" lines, "coarse_nodes N", "jacobi_pairs 0|1" and the AMR
 * fine levels' stencil tiles "amr_tile_width 16|64" (FP64),
 * "amr_tile_width32 16|64" (FP32) and "amr_tile_kz K" planes, and
 * "mac_coarse_cells N": cell-centred multigrid levels below the first with
 * at most N cells run in one workgroup per V-cycle half, 0: none), for caching
 * per device model; mgx_tuning_set loads such a text (replacing the table;
 * lines left out take their defaults). */
MGX_API int mgx_tuning_get(char* buf, int len);
MGX_API int mgx_tuning_set(const char* text);
/* Variable sigma, a boxed cell field with ng ghosts (:191, sigma = s/rho at :117). */
MGX_API int mgx_nodal_set_sigma(mgx_nodal h, const double* sigma, int sigma_ng);
/* Scalar sigma = s/rho_0 (:186-188). */
MGX_API int mgx_nodal_set_const_sigma(mgx_nodal h, double sigma);
/* NodalProjector::project(rtol, atol) (:215): rhs = D.vel (vel: boxed cell
 * field, 3 comps, vel_ng >= 1 ghosts; ghosts outside non-periodic domain faces
 * are read as given, e.g. 0 or the inflow velocity), solve L phi = rhs, then
 * vel -= sigma G(phi) on valid cells, in place. */
MGX_API int mgx_nodal_project(mgx_nodal h, double* vel, int vel_ng, double rtol, double atol,
                              int mode, mgx_solve_info* info);
/* NodalProjector::getPhi() / getGradPhi() (:218-219): copy out the last
 * projection's phi (boxed nodal, 1 comp) and G(phi) (boxed cell, 3 comps). */
MGX_API int mgx_nodal_get_phi(mgx_nodal h, double* phi, int phi_ng);
MGX_API int mgx_nodal_get_gradphi(mgx_nodal h, double* gphi, int gphi_ng);
/* MLLinOp::apply on level 0: out = L phi with the current sigma (boxed nodal
 * fields; used to check a solution's residual with the device operator). */
MGX_API int mgx_nodal_apply(mgx_nodal h, const double* phi, int phi_ng, double* out, int out_ng);
/* MLMG::solve(sol, rhs, rtol, atol) on the nodal operator: boxed nodal fields
 * (src/diffusion/DiffusionTensorOp.cpp:188-215 shows the call shape).
 * phi holds the initial guess on entry. */
MGX_API int mgx_nodal_solve(mgx_nodal h, double* phi, int phi_ng, const double* rhs, int rhs_ng,
                            double rtol, double atol, int mode, mgx_solve_info* info);
/* Whole incflo::ApplyNodalProjection (wrapper + core,
 * src/projection/incflo_apply_nodal_projection.cpp:29-267) in one call:
 *   if !incremental: u += (s/rho) gp                          (:38-62)
 *   if incremental or vel_old given for the small-dt form: u -= vel_old (:64-71)
 *   sigma = s/rho (or s/const_rho when const_rho > 0)          (:101-121, :184-188)
 *   ghost cells: 0, inflow faces = inflow_vel[face][comp]      (:137-163)
 *   project(rtol, atol)                                        (:215)
 *   gp = G(phi) (+= if incremental), p_nd = phi (+= if incr.)  (:221-256)
 *   if vel_old given: u += vel_old                             (:85-92)
 * density: boxed cell field (rho_ng ghosts) or NULL with const_rho > 0;
 * vel: 3-comp cell field (vel_ng >= 1); gp: 3-comp cell field; p_nd: nodal.
 * inflow_vel: 6x3 doubles indexed [2*dir + (0 lo | 1 hi)][comp] or NULL. */
MGX_API int mgx_apply_nodal_projection(mgx_nodal h, const double* density, int rho_ng,
                                       double const_rho, double* vel, int vel_ng,
                                       const double* vel_old, int vel_old_ng, double* gp,
                                       int gp_ng, double* p_nd, int p_ng,
                                       double scaling_factor, int incremental,
                                       const double* inflow_vel, double rtol, double atol,
                                       int mode, mgx_solve_info* info);

/* ---- embedded boundaries on the nodal projection ----------------------------
 * The reference builds EB data with EB2::CylinderIF + EB2::Build
 * (src/embedded_boundaries/eb_cylinder.cpp:53-63) and MLNodeLaplacian
 * assembles its cut-cell operator; here a cut cell's Q1 element is assembled
 * from the fluid moments of the cell, coarse levels use Galerkin cut elements.
 * With EB the solve runs in FP64 (MGX_MODE_FP64).
 * A cylinder along z with fluid outside (|x - (cx, cy)| >= r): */
MGX_API int mgx_nodal_set_eb_cylinder(mgx_nodal h, double cx, double cy, double r);
/* Any geometry: per canonical cell (x fastest, n0*n1*n2 cells) a type (0
 * covered, 1 regular, 2 cut) and 27 fluid moments int xi^p eta^q zeta^r over
 * the fluid part of the unit cell (index p + 3q + 9r); type NULL removes the EB. */
MGX_API int mgx_nodal_set_eb_moments(mgx_nodal h, const int* type, const double* moments);

/* getLinOp().setEBInflowVelocity(lev, eb_vel) (src/projection/incflo_apply_nodal_projection.cpp:196-200):
 * a velocity on the body surface (boxed cell field, 3 comps, ng ghosts; the
 * values of cut cells are used) adds int over the surface of N (u_eb . n) to
 * the divergence, n out of the fluid.  Cylinder geometry only; NULL removes it. */
MGX_API int mgx_nodal_set_eb_inflow_velocity(mgx_nodal h, const double* eb_vel, int ng);
/* getLinOp().setOversetMask(lev, mask) (src/projection/incflo_apply_nodal_projection.cpp:203-211),
 * which incflo uses for mixed boundaries (src/boundary_conditions/incflo_set_bcs.cpp:11-53):
 * a boxed nodal int field (ng ghosts), 1 = solve, 0 = node fixed at phi = 0;
 * NULL removes it.  Not with embedded boundaries or the learned mode. */
MGX_API int mgx_nodal_set_overset_mask(mgx_nodal h, const int* mask, int ng);
/* Host-only: body-surface weights of the cylinder, 24 per cell ([corner][d],
 * x bit 0, y bit 1, z bit 2), (1/V) int over the surface in the cell of
 * N_p n_d dA with n out of the fluid; zero outside cut cells (no GPU). */
MGX_API int mgx_eb_cylinder_surface_weights(const mgx_level_desc* level, double cx, double cy,
                                            double r, double* weights);
/* Host-only: the cylinder's face apertures on a level (canonical face arrays,
 * n_d + 1 faces along d, x fastest) (no GPU). */
MGX_API int mgx_eb_cylinder_apertures(const mgx_level_desc* level, double cx, double cy, double r,
                                      double* ax, double* ay, double* az);
/* Host-only: the cylinder's per-cell types and moments on a level (no GPU). */
MGX_API int mgx_eb_cylinder_geometry(const mgx_level_desc* level, double cx, double cy, double r,
                                     int* type, double* moments);

/* ---- AMR: composite nodal projection over the levels ---------------------
 * incflo projects every AMR level with one NodalProjector
 * (src/projection/incflo_apply_nodal_projection.cpp:124-215: vel and sigma
 * per level, project(rtol, atol), getPhi / getGradPhi per level) and then
 * averages gp down (:258-266).  levels[0] tiles the domain; each further
 * level (up to MGX_AMR_MAX_LEVELS in all) has n_cell = 2 x the level below,
 * the same domain and periodicity, boxes aligned with the cells of the level
 * below (even lo, odd hi) and properly nested in it (every node of the level
 * below that touches a cell under the finer level lies inside the level
 * below's own region: one cell of buffer away from domain faces).  Fine
 * boxes may touch Dirichlet and Neumann domain faces, whose ghost velocity
 * the caller sets to 0, but not inflow faces (MGX_ERR_UNSUPPORTED).  The
 * composite problem is the conforming Q1 discretisation on the composite
 * mesh (each level's nodes inside its region, nodes on a level's boundary
 * following the level below); cells under a finer level receive the average
 * of their 8 children for the velocity and G phi (applied from the finest
 * level down), nodes under a finer level its value. */
#define MGX_AMR_MAX_LEVELS 4
typedef struct mgx_amr_s* mgx_amr_nodal;
MGX_API int mgx_amr_nodal_create(mgx_ctx ctx, const mgx_level_desc* levels, int nlevels,
                                 mgx_amr_nodal* out);
MGX_API int mgx_amr_nodal_destroy(mgx_amr_nodal h);
/* setDomainBC (:194), as mgx_nodal_set_domain_bc (sigma handed over before
 * is kept). */
MGX_API int mgx_amr_nodal_set_domain_bc(mgx_amr_nodal h, const int lo[3], const int hi[3]);
MGX_API int mgx_amr_nodal_set_options(mgx_amr_nodal h, const mgx_mlmg_opts* opts);
/* sigma per level (:101-121): sigma[lev] is a boxed cell field of level lev (ng ghosts). */
MGX_API int mgx_amr_nodal_set_sigma(mgx_amr_nodal h, const double* const* sigma, int sigma_ng);
/* project(rtol, atol) over the levels (:215): vel[lev] (boxed cell, 3 comps,
 * vel_ng ghosts) is overwritten in place; phi[lev] (boxed nodal) and
 * gphi[lev] (boxed cell, 3 comps) receive getPhi() / getGradPhi() (:218-219);
 * phi / gphi or their entries may be NULL.  MGX_MODE_FP64 or MGX_MODE_MIXED. */
MGX_API int mgx_amr_nodal_project(mgx_amr_nodal h, double* const* vel, int vel_ng,
                                  double* const* phi, int phi_ng, double* const* gphi,
                                  int gphi_ng, double rtol, double atol, int mode,
                                  mgx_solve_info* info);

/* ---- AMR: composite MAC projection over the levels ----------------------
 * incflo builds one MacProjector over all AMR levels
 * (src/convection/incflo_compute_MAC_projected_velocities.cpp:96-133:
 * initProjector with beta per level, setDomainBC; :272-300: setUMAC with the
 * face velocities of every level, project(rtol, atol) or project(phi, rtol,
 * atol)), and the projector averages the fine faces down (:301-302).  Level
 * rules as for mgx_amr_nodal_create (up to MGX_AMR_MAX_LEVELS, properly
 * nested: every cell around a cell under a finer level belongs to its own
 * level), except that fine boxes may touch every kind of non-periodic face.
 * The composite problem: each level's cells not under a finer level are the
 * unknowns; every face between two of them carries beta (phi_j - phi_i) /
 * (distance of the cell centres), 1.5 h across a face on a level's boundary
 * to the cell of the level below; that cell takes the 4 finer fluxes on each
 * such face (reflux).  Boundary types: Dirichlet, Neumann or inflow (no
 * Robin). */
typedef struct mgx_amr_mac_s* mgx_amr_mac;
MGX_API int mgx_amr_mac_create(mgx_ctx ctx, const mgx_level_desc* levels, int nlevels,
                               mgx_amr_mac* out);
MGX_API int mgx_amr_mac_destroy(mgx_amr_mac h);
/* setDomainBC (:114), as mgx_mac_set_domain_bc without ROBIN (beta handed
 * over before is kept). */
MGX_API int mgx_amr_mac_set_domain_bc(mgx_amr_mac h, const int lo[3], const int hi[3]);
/* mac_proj.* knobs (maxiter / bottom_maxiter default 200). */
MGX_API int mgx_amr_mac_set_options(mgx_amr_mac h, const mgx_mlmg_opts* opts);
/* initProjector(info, beta) / updateCoeffs(beta) (:112, :131): beta[3 lev + d]
 * is the boxed face field of level lev, direction d (ng ghosts). */
MGX_API int mgx_amr_mac_set_beta(mgx_amr_mac h, const double* const* beta, int ng);
/* setUMAC(umac) + project(phi, rtol, atol) (:272-299): umac[3 lev + d] (boxed
 * faces, umac_ng ghosts) is overwritten with the projected velocity, faces
 * under a finer level with the mean of its faces on them (from the finest
 * level down); phi[lev] (boxed cells, may be NULL) receives phi, cells under
 * a finer level the mean of its 8 cells.  MGX_MODE_FP64 or MGX_MODE_MIXED. */
MGX_API int mgx_amr_mac_project(mgx_amr_mac h, double* const* umac, int umac_ng,
                                double* const* phi, int phi_ng, double rtol, double atol,
                                int mode, mgx_solve_info* info);

/* AMR synchronisation, refinement ratio 2 (fine boxes aligned with coarse cells):
 * average_down_faces (src/convection/incflo_compute_advection_term.cpp:559-569):
 * every coarse face under the fine level (its closure included) becomes the
 * mean of the 4 fine faces on it, in every coarse box holding it; boxed face
 * fields with ncomp components (fine_faces[d] / crse_faces[d]). */
MGX_API int mgx_amr_average_down_faces(mgx_ctx ctx, const mgx_level_desc* crse,
                                       const mgx_level_desc* fine, const double* const* fine_faces,
                                       int fine_ng, double* const* crse_faces, int crse_ng,
                                       int ncomp);
/* average_down of cell fields (src/incflo_update_density.cpp:47-51,
 * src/incflo_update_tracer.cpp:61-65): coarse cell = mean of its 8 fine cells. */
MGX_API int mgx_amr_average_down(mgx_ctx ctx, const mgx_level_desc* crse,
                                 const mgx_level_desc* fine, const double* fine_cells, int fine_ng,
                                 double* crse_cells, int crse_ng, int ncomp);

/* ---- MAC projection (Hydro::MacProjector + MLABecLaplacian/MLPoisson + MLMG) */
typedef struct mgx_mac_s* mgx_mac;
/* MacProjector(Vector<Geometry>) (src/incflo.cpp:291); the level's box list
 * lays out the face fields (x-faces of box b span lo..hi+1 in x). */
MGX_API int mgx_mac_create(mgx_ctx ctx, const mgx_level_desc* level, mgx_mac* out);
MGX_API int mgx_mac_destroy(mgx_mac h);
/* setDomainBC (src/convection/incflo_compute_MAC_projected_velocities.cpp:114) with the
 * get_mac_projection_bc mapping (src/projection/incflo_projection_bc.cpp:44-78):
 * DIRICHLET for pressure in/outflow, NEUMANN (or INFLOW) for walls and mass inflow,
 * ROBIN for mixed faces (coefficients from mgx_mac_set_level_bc). */
MGX_API int mgx_mac_set_domain_bc(mgx_mac h, const int lo[3], const int hi[3]);
/* setLevelBC(lev, nullptr, &robin_a, &robin_b, &robin_f)
 * (src/convection/incflo_compute_MAC_projected_velocities.cpp:115-122, data from
 * src/boundary_conditions/incflo_set_bcs.cpp:122-190): coefficients of
 * a phi + b dphi/dn = f on the MGX_BC_ROBIN sides, read from the first ghost
 * cell outside each such face of three boxed cell fields (ng >= 1 ghosts).
 * Needs a >= 0, b >= 0, a h/2 + b > 0 there.  Call after setDomainBC (which
 * clears them); the solve refuses Robin sides without them. */
MGX_API int mgx_mac_set_level_bc(mgx_mac h, const double* robin_a, const double* robin_b,
                                 const double* robin_f, int ng);
/* mac_proj.* knobs (maxiter / bottom_maxiter default 200). */
MGX_API int mgx_mac_set_options(mgx_mac h, const mgx_mlmg_opts* opts);
/* Launch mode (no reference counterpart; for callers that must not use HIP
 * graphs, and for tests): on (default), each V-cycle of the cell-centred
 * multigrid is captured once as a HIP graph and replayed; off, its kernels
 * are launched one by one.  The kernels and results are the same either way. */
MGX_API int mgx_mac_set_graphs(mgx_mac h, int on);
/* initProjector(lp_info, beta) / updateCoeffs(beta) (:112, :131): face coefficient
 * fields (boxed face layout, ng ghosts). */
MGX_API int mgx_mac_set_beta(mgx_mac h, const double* bx, const double* by, const double* bz,
                             int ng);
/* initProjector(ba, dm, lp_info, beta) / updateBeta(beta) (:108, :127): constant beta. */
MGX_API int mgx_mac_set_const_beta(mgx_mac h, double beta);
/* incflo's beta_f = dt / (face average of rho) (:70-91), from a boxed cell density. */
MGX_API int mgx_mac_set_beta_from_density(mgx_mac h, const double* rho, int rho_ng, double dt);
/* setUMAC + project(rtol, atol) / project(phi, rtol, atol) (:272, :294-299):
 * solve div(beta grad phi) = div(u_mac), then u_mac -= beta grad phi in place.
 * phi (boxed cell field, may be NULL) is the initial guess and receives phi. */
MGX_API int mgx_mac_project(mgx_mac h, double* umac, double* vmac, double* wmac, int umac_ng,
                            double* phi, int phi_ng, double rtol, double atol, int mode,
                            mgx_solve_info* info);
/* getFluxes (src/projection/incflo_apply_cc_projection.cpp:387-389): -beta grad phi on faces. */
MGX_API int mgx_mac_get_fluxes(mgx_mac h, double* fx, double* fy, double* fz, int ng);
/* incflo::ApplyCCProjection (src/projection/incflo_apply_cc_projection.cpp:199-487),
 * the projection incflo runs instead of the nodal one when use_cc_proj is set
 * (dispatch in incflo::ApplyProjection, src/incflo.cpp:213-231), on the MAC
 * handle's geometry, boundary types and options (single level, no EB):
 *   if !incremental: u += (s/rho) gp;  if incremental: u -= vel_old
 *   ghost cells 0, faces marked MGX_BC_INFLOW the inflow_vel table (6x3,
 *     [2*dir + side][comp], not when incremental)
 *   u_mac = MOL::ExtrapVelToFaces(u) (limited-slope upwind face velocity)
 *   beta = s / face-average(rho) (or s / const_rho), MAC projection of u_mac
 *     from phi = 0, fluxes F = -beta grad phi
 *   u += mean of the two fluxes per direction (average_mac_to_ccvel),
 *   p_cc = phi, gp = -(that mean) rho / s (both += / -= when incremental);
 *   if incremental: u += vel_old.
 * density: boxed cells or NULL with const_rho > 0; vel: 3 comps, vel_ng >= 1;
 * gp: 3 comps; p_cc: 1 comp; umac/vmac/wmac (may be NULL): the projected face
 * velocities. */
MGX_API int mgx_apply_cc_projection(mgx_mac h, const double* density, int rho_ng,
                                    double const_rho, double* vel, int vel_ng,
                                    const double* vel_old, int vel_old_ng, double* gp, int gp_ng,
                                    double* p_cc, int p_ng, double scaling_factor, int incremental,
                                    const double* inflow_vel, double* umac, double* vmac,
                                    double* wmac, int umac_ng, double rtol, double atol, int mode,
                                    mgx_solve_info* info);
/* EB form (MacProjector with the EB factory, src/incflo.cpp:286-289; MLEBABecLap):
 * a cylinder along z (fluid outside).  Face apertures a_f weight the
 * coefficients (operator div(a beta grad)), the right-hand side is div(a u),
 * covered cells drop out and fully covered faces get u = 0.  Call after
 * setDomainBC and before the coefficients; needs a Dirichlet face. */
MGX_API int mgx_mac_set_eb_cylinder(mgx_mac h, double cx, double cy, double r);
/* setEBInflowVelocity(lev, eb_vel) (src/convection/incflo_compute_MAC_projected_velocities.cpp:277-279):
 * velocity on the body surface (boxed cell field, 3 comps): the surface flux
 * -sum_d (a_d+ - a_d-) u_eb,d / h_d joins the divergence of each cut cell.
 * NULL removes it. */
MGX_API int mgx_mac_set_eb_inflow_velocity(mgx_mac h, const double* eb_vel, int ng);
/* MLMG::solve on the MAC operator (boxed cell fields; phi holds the initial guess). */
MGX_API int mgx_mac_solve(mgx_mac h, double* phi, int phi_ng, const double* rhs, int rhs_ng,
                          double rtol, double atol, int mode, mgx_solve_info* info);

/* ---- Godunov advection on the predictor path (single level, no EB) --------
 * The reference calls AMReX-Hydro's Godunov here; the method (PPM with the
 * Colella-Woodward limiter, corner-transport-upwind transverse terms,
 * use_forces_in_trans = false, src/incflo.H:516,525) is restated in
 * incflo_amd/csrc/godunov_kernels.h.  Boxed layouts as above; faces shared by
 * two boxes must agree.  Domain faces of non-periodic directions follow
 * amrex::BCType records (6 ints per component: lo x, y, z, hi x, y, z; ext_dir
 * 3, foextrap 2, hoextrap 4, reflect_even 1, reflect_odd -1; periodic
 * directions int_dir 0), as incflo's get_velocity_bcrec / get_density_bcrec /
 * get_tracer_bcrec; cell states then carry 2 filled ghost cells (incflo's
 * FillPatch; an ext_dir ghost holds the boundary value). */
typedef struct mgx_adv_s* mgx_adv;
MGX_API int mgx_adv_create(mgx_ctx ctx, const mgx_level_desc* level, mgx_adv* out);
MGX_API int mgx_adv_destroy(mgx_adv h);
/* get_velocity_bcrec() (src/incflo.H): 3 x 6 BCType values; required before
 * the velocity calls on a bounded domain. */
MGX_API int mgx_adv_set_velocity_bcrec(mgx_adv h, const int* bcrec);
/* HydroUtils::ExtrapVelToFaces (src/convection/incflo_compute_MAC_projected_velocities.cpp:168-178):
 * cell velocity (3 comps, vel_ng ghosts) and forces (3 comps, f_ng ghosts,
 * NULL = none) -> face-normal velocities at t^{n+1/2} (valid faces of the
 * boxed face fields with umac_ng ghosts are written). */
MGX_API int mgx_adv_extrap_vel_to_faces(mgx_adv h, const double* vel, int vel_ng,
                                        const double* force, int f_ng, double dt, double* umac,
                                        double* vmac, double* wmac, int umac_ng);
/* ComputeFluxesOnBoxFromState + ComputeDivergence(mult = -1) + ComputeConvectiveTerm
 * for the velocity in convective form (src/convection/incflo_compute_advection_term.cpp:361-428,
 * 575-661): edge states of u, v, w by Godunov with the MAC velocities,
 * fluxes u_mac * state (not area weighted), conv = -div(flux) + u div(u_mac)
 * (3 comps, conv_ng ghosts).  face_* / flux_* (boxed faces, 3 comps, no
 * ghosts) may be NULL. */
MGX_API int mgx_adv_convective_term(mgx_adv h, const double* vel, int vel_ng, const double* force,
                                    int f_ng, const double* umac, const double* vmac,
                                    const double* wmac, int umac_ng, double dt, double* conv,
                                    int conv_ng, double* face_x, double* face_y, double* face_z,
                                    double* flux_x, double* flux_y, double* flux_z);

/* ComputeFluxesOnBoxFromState (is_velocity = false) + ComputeDivergence(mult = -1)
 * + ComputeConvectiveTerm for density and tracers
 * (src/convection/incflo_compute_advection_term.cpp:446-561, 666-760): edge states
 * of ncomp cell scalars (state_ng ghosts) with the MAC velocities, fluxes
 * u_mac * state, conv = -div(flux), plus state * div(u_mac) for components
 * with iconserv[n] == 0 (iconserv NULL: all convective); bcrec: ncomp x 6
 * BCType values (NULL on periodic domains).  Conservative
 * components (density; rho * tracer for conservative tracers, which the caller
 * forms as incflo does at :498-511) take conservative transverse terms.
 * face_* / flux_* (boxed faces, ncomp comps, no ghosts) may be NULL. */
MGX_API int mgx_adv_scalar_term(mgx_adv h, const double* state, int state_ng, int ncomp,
                                const int* iconserv, const int* bcrec, const double* force,
                                int f_ng,
                                const double* umac, const double* vmac, const double* wmac,
                                int umac_ng, double dt, double* conv, int conv_ng, double* face_x,
                                double* face_y, double* face_z, double* flux_x, double* flux_y,
                                double* flux_z);

/* ---- the predictor step around the projections (single level) -----------
 * incflo::ApplyPredictor (src/incflo_apply_predictor.cpp) runs, between the
 * MAC and the nodal projection, the forces, the convective term, the density
 * and velocity updates and the viscous solve.  These entry points keep that
 * work on the device, on the caller's boxed fields (layout as above). */
typedef struct mgx_step_s* mgx_step;
MGX_API int mgx_step_create(mgx_ctx ctx, const mgx_level_desc* level, mgx_step* out);
MGX_API int mgx_step_destroy(mgx_step h);
/* MultiFab::FillBoundary(geom.periodicity()) of a boxed cell field with ng
 * ghosts and ncomp components: every ghost cell that lies inside the domain
 * or across a periodic face takes the valid value there; ghosts outside
 * non-periodic faces are left as they are (the caller's physical BC fill,
 * as incflo's fillpatch / fillphysbc). */
MGX_API int mgx_fill_boundary(mgx_step h, double* field, int ng, int ncomp);
/* mgx_fill_boundary without any allocation (no counterpart in the reference):
 * for callers that cannot allocate device memory -- code under HIP graph
 * capture, tight stepping loops, harnesses that forbid hipMalloc.  The
 * canonical staging copy lives in a fixed 64 MiB device array that each GPU
 * holds until the process ends, so ncomp x cells must be at most 8,388,608
 * (MGX_ERR_ARG otherwise).  Every call on one GPU shares that array: calls
 * that overlap on one GPU (several streams or host threads) race. */
MGX_API int mgx_fill_boundary_noalloc(mgx_step h, double* field, int ng, int ncomp);
/* incflo::compute_vel_forces (src/incflo_compute_forces.cpp:50-151, the
 * non-Boussinesq branch :134-148): forces = -(gp + gp0)/rho + gravity
 * (without gp when include_pressure_gradient == 0), plus divtau when given
 * (the Godunov forcing, src/convection/incflo_compute_MAC_projected_velocities.cpp:24-57).
 * gravity / gp0: 3 doubles or NULL (0). */
MGX_API int mgx_compute_vel_forces(mgx_step h, const double* gp, int gp_ng, const double* rho,
                                   int rho_ng, const double* divtau, int divtau_ng,
                                   const double gravity[3], const double gp0[3],
                                   int include_pressure_gradient, double* forces, int f_ng);
/* incflo::update_velocity, predictor (src/incflo_update_velocity.cpp:14-143):
 *   velocity form (rho_* NULL):  vel = vel_old + dt (conv + forces + c divtau)
 *   momentum form:  vel = (rho_old vel_old + dt (conv + rho_nph forces + c divtau)) / rho_new
 * with c = divtau_coef: 1 explicit, 0.5 Crank-Nicolson, 0 (or divtau NULL)
 * implicit; the implicit and Crank-Nicolson forms are followed by
 * mgx_diff_solve (:322-334). */
MGX_API int mgx_update_velocity(mgx_step h, double* vel, int vel_ng, const double* vel_old,
                                int vo_ng, const double* conv, int conv_ng, const double* forces,
                                int f_ng, const double* divtau, int divtau_ng, double divtau_coef,
                                const double* rho_old, const double* rho_new,
                                const double* rho_nph, int rho_ng, double dt);
/* incflo::update_density, predictor (src/incflo_update_density.cpp:14-64):
 * rho_new = rho_old + dt drdt, rho_nph = (rho_old + rho_new)/2 (rho_nph may be NULL). */
MGX_API int mgx_update_density(mgx_step h, const double* rho_old, int ro_ng, const double* drdt,
                               int c_ng, double dt, double* rho_new, int rn_ng, double* rho_nph,
                               int rh_ng);

/* ---- cell-centred viscous operator and Helmholtz solve ---------------------
 * DiffusionScalarOp (src/diffusion/DiffusionScalarOp.cpp): per component,
 * MLABecLaplacian with alpha = 1, a = rho, beta = dt and b = eta averaged to
 * faces (average_scalar_eta_to_faces, src/diffusion/incflo_diffusion.cpp:289-380).
 * The same red-black Gauss-Seidel multigrid as the MAC projection, with the
 * cell coefficient on every level (coarse a = mean of 8 fine cells).
 * Dirichlet faces take their value from the field's first ghost cell
 * (setLevelBC(lev, &vel), two-point face flux as mg_maxorder = 2). */
typedef struct mgx_diff_s* mgx_diff;
MGX_API int mgx_diff_create(mgx_ctx ctx, const mgx_level_desc* level, mgx_diff* out);
MGX_API int mgx_diff_destroy(mgx_diff h);
/* get_diffuse_velocity_bc / get_diffuse_scalar_bc (src/diffusion/incflo_diffusion.cpp:156-260):
 * LinOpBCType per side for component comp (-1: all); periodic, Dirichlet or Neumann. */
MGX_API int mgx_diff_set_domain_bc(mgx_diff h, int comp, const int lo[3], const int hi[3]);
/* scalar_diffusion.* knobs (DiffusionScalarOp.cpp:130-146; maxiter 100). */
MGX_API int mgx_diff_set_options(mgx_diff h, const mgx_mlmg_opts* opts);
/* viscosity eta: boxed cell field (ng ghosts) or a positive constant when eta is NULL. */
MGX_API int mgx_diff_set_eta(mgx_diff h, const double* eta, int ng, double const_eta);
/* compute_divtau (scalar form, :584-760) / compute_laps (:482-580):
 * out_c = div(eta grad field_c) [/ rho when rho is given], c < ncomp. */
MGX_API int mgx_diff_apply(mgx_diff h, const double* field, int ng, int ncomp, const double* rho,
                           int rho_ng, double* out, int out_ng);
/* diffuse_vel_components / diffuse_scalar (:332-480, MLMG::solve to rtol, atol):
 * every component c < ncomp solves (a - dt div(eta grad)) x = a field_c with
 * a = acoef (boxed cell field) or 1 when NULL; field is the initial guess and
 * receives x.  info sums iterations and V-cycles over the components. */
MGX_API int mgx_diff_solve(mgx_diff h, double* field, int ng, int ncomp, const double* acoef,
                           int a_ng, double dt, double rtol, double atol, int mode,
                           mgx_solve_info* info);

/* ---- AMReX-format output and restart ---------------------------------------
 * incflo writes plotfiles and checkpoints through AMReX
 * (src/utilities/io.cpp:16-111 and WritePlotFile); these write the same
 * formats from device fields (copied to the host first).
 * VisMF::Write of one boxed field: <prefix>_H and <prefix>_D_00000, fabs over
 * the boxes grown by ng, cell or nodal (nodal != 0) boxes, ncomp components. */
MGX_API int mgx_write_multifab(mgx_ctx ctx, const char* prefix, const mgx_level_desc* level,
                               const double* field, int ng, int ncomp, int nodal);
/* VisMF::Read into a boxed device field of the same box list (valid cells and
 * the ghosts both sides have); MGX_ERR_ARG when the boxes differ. */
MGX_API int mgx_read_multifab(mgx_ctx ctx, const char* prefix, const mgx_level_desc* level,
                              double* field, int ng, int ncomp, int nodal);
/* incflo::WriteHeader for a checkpoint (src/utilities/io.cpp:16-68):
 * "Checkpoint version: 1", finest level, step, time, dt, prev_dt,
 * prev_prev_dt, prob_lo, prob_hi, one BoxArray per level; creates dir. */
MGX_API int mgx_write_checkpoint_header(const char* dir, int nlevels, const mgx_level_desc* levels,
                                        int nstep, double cur_time, double dt, double prev_dt,
                                        double prev_prev_dt);
/* incflo::ReadCheckpointFile's header part (:113-205): times = {cur_time,
 * dt, prev_dt, prev_prev_dt}; per level its box count and the boxes (6 ints
 * each, levels back to back, at most max_boxes in all). */
MGX_API int mgx_read_checkpoint_header(const char* dir, int* finest_level, int* nstep,
                                       double* times, double* prob_lo, double* prob_hi,
                                       int* nboxes_per_level, int max_levels, int* boxes,
                                       int max_boxes);
/* amrex::WriteMultiLevelPlotfile ("HyperCLaw-V1.1" Header, Level_<l>/Cell
 * multifabs): data[l] is level l's boxed cell field with nvars components
 * and ng ghosts; level_steps / ref_ratio may be NULL (0 / 2). */
MGX_API int mgx_write_plotfile(mgx_ctx ctx, const char* dir, int nlevels,
                               const mgx_level_desc* levels, const double* const* data, int ng,
                               int nvars, const char* const* varnames, double time,
                               const int* level_steps, const int* ref_ratio);

/* ---- dynamic refinement (ratio 2 between neighbouring levels) ------------
 * incflo regrids through AmrCore: ErrorEst tags the cells of each level,
 * AMReX's grid generation clusters them into the boxes of the next level, and
 * MakeNewLevelFromCoarse / RemakeLevel fill the fields of every new level
 * (src/incflo_regrid.cpp:9-113).
 *
 * incflo::ErrorEst (src/incflo_tagging.cpp:11-198) on one level: tags[c] = 1
 * (canonical cell index, x fastest; device bytes, other entries left as they
 * are) where rho > rhoerr, where the largest one-sided density difference
 * reaches gradrhoerr (rho then needs rho_ng >= 1 filled ghost cells), or
 * where the cell centre lies in [region_lo, region_hi] (both NULL: no
 * region).  A criterion whose threshold is not finite (INFINITY, NAN) is off,
 * as an absent incflo.rhoerr / incflo.gradrhoerr is. */
MGX_API int mgx_tag_cells(mgx_ctx ctx, const mgx_level_desc* level, const double* rho, int rho_ng,
                          double rhoerr, double gradrhoerr, const double* region_lo,
                          const double* region_hi, unsigned char* tags);
/* AmrMesh::MakeNewGrids for level l + 1 from the tags of level l over its
 * domain of n_cell cells (host memory): tags buffered by n_error_buf cells (periodic directions
 * wrap), coarsened to blocks of blocking_factor/2 coarse cells, clustered by
 * Berger-Rigoutsos to grid_eff, chopped to max_grid_size and refined: fine
 * boxes (6 ints each, fine index space) aligned to blocking_factor; at most
 * max_boxes (MGX_ERR_ARG beyond).  amr.n_error_buf, amr.blocking_factor,
 * amr.max_grid_size and amr.grid_eff of the inputs. */
MGX_API int mgx_cluster_tags(const unsigned char* tags, const int n_cell[3],
                             const int is_periodic[3], int n_error_buf, int blocking_factor,
                             int max_grid_size, double grid_eff, int* boxes, int max_boxes,
                             int* nboxes);
/* FillPatchTwoLevels / InterpFromCoarseLevel with cell_cons_interp
 * (src/boundary_conditions/incflo_fillpatch.cpp; ng = 0 as
 * MakeNewLevelFromCoarse / RemakeLevel call them): the valid cells of the
 * fine boxed field (f_ng ghosts, ncomp <= 8 components) take the old fine
 * level's values where one of its boxes covers them (old_fine / old_field
 * NULL: none, a level made from the coarse one) and the conservative linear
 * interpolation of the coarse field elsewhere.  The components share each
 * direction's slope limiter (as one FillPatch of a multi-component
 * field does); bcrec: ncomp x 6 amrex::BCType values (NULL: all int_dir)
 * select the one-sided slope next to ext_dir / hoextrap faces.  The coarse
 * field needs c_ng >= 1 filled ghost cells on a bounded domain.  The coarse
 * level may itself be a fine level (any level below the one filled): its
 * boxes must cover the parents of the fine cells and their neighbours
 * (MGX_ERR_ARG otherwise). */
MGX_API int mgx_fill_fine_level(mgx_ctx ctx, const mgx_level_desc* crse, const double* crse_field,
                                int c_ng, int ncomp, const int* bcrec,
                                const mgx_level_desc* fine, double* fine_field, int f_ng,
                                const mgx_level_desc* old_fine, const double* old_field,
                                int of_ng);

/* ---- learned coarse-level predictor --------------------------------------
 * No counterpart in the reference (incflo has no learned component); these
 * entry points serve MGX_MODE_LEARNED of the nodal solve.  A 3-D conv U-net
 * (4 levels, base_channels -> 8*base_channels, 3x3x3 kernels, BF16 weights,
 * FP32 accumulation on the matrix cores) maps the normalised coarse residual
 * and coarse log(sigma) to the normalised coarse correction.
 * Tensors are device FP32, channels last: in[batch][nz][ny][nx][8]
 * (channels 0,1 used), out[batch][nz][ny][nx]; sizes multiples of 8. */
typedef struct mgx_unet_s* mgx_unet;
MGX_API int mgx_unet_create(mgx_ctx ctx, int base_channels, int levels, unsigned long long seed,
                            mgx_unet* out);
MGX_API int mgx_unet_destroy(mgx_unet u);
MGX_API int mgx_unet_num_params(mgx_unet u, long long* n);
/* FP32 master parameters (host arrays of num_params floats). */
MGX_API int mgx_unet_get_params(mgx_unet u, float* host, long long n);
MGX_API int mgx_unet_set_params(mgx_unet u, const float* host, long long n);
MGX_API int mgx_unet_forward(mgx_unet u, const float* in, float* out, int nx, int ny, int nz,
                             int batch, const int periodic[3]);
/* forward + backward with the given upstream gradient dL/d(out); the
 * parameter gradients are copied to host_grads (n = num_params). */
MGX_API int mgx_unet_gradients(mgx_unet u, const float* in, const float* dout, int nx, int ny,
                               int nz, int batch, const int periodic[3], float* host_grads,
                               long long n);
/* one Adam step on the mean-squared error against target[batch][nz][ny][nx]. */
MGX_API int mgx_unet_train_step(mgx_unet u, const float* in, const float* target, int nx, int ny,
                                int nz, int batch, const int periodic[3], float lr, float* loss);
/* The BF16-MFMA 3x3x3 convolution the U-net is built from (operator level):
 * out = conv(in, w) + bias (ReLU if relu), w: FP32 [cout][27*cin] with
 * k = tap*cin + ci, tap = (dz+1)*9 + (dy+1)*3 + (dx+1); periodic or zero padding. */
MGX_API int mgx_conv3d_forward(mgx_ctx ctx, const float* in, const float* w, const float* bias,
                               float* out, int nx, int ny, int nz, int batch, int cin, int cout,
                               const int periodic[3], int relu);
/* Gradients of that convolution (no ReLU) for upstream dout: din (may be
 * NULL), dw [cout][27*cin] and db [cout] (either may be NULL), overwritten. */
MGX_API int mgx_conv3d_backward(mgx_ctx ctx, const float* in, const float* w, const float* dout,
                                float* din, float* dw, float* db, int nx, int ny, int nz,
                                int batch, int cin, int cout, const int periodic[3]);
/* Attach a U-net to a nodal projector for MGX_MODE_LEARNED: it acts on the
 * multigrid level `level` (coarsened `level` times; sizes there must be
 * multiples of 8, periodic domains only).  u may be NULL to detach. */
MGX_API int mgx_nodal_set_unet(mgx_nodal h, mgx_unet u, int level);
/* After a solve: the network input (normalised coarse rhs and log sigma,
 * [nz][ny][nx][8] at `level`) and the normalised target (the solution
 * injected at that level's nodes) -- one training pair for mgx_unet_train_step. */
MGX_API int mgx_nodal_training_pair(mgx_nodal h, int level, float* in, float* target);

#ifdef __cplusplus
}
#endif
#endif /* MGX_H */
