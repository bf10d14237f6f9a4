/*
 * orc.h -- FP64 CPU oracle for the projection hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path
 * is compared against; it is never linked into, called by or shipped with the
 * product library (incflo_amd/csrc).  Only tests/, the smoke entry point and
 * the CPU-baseline leg of bench.py load it.
 *
 * It restates, in plain C + OpenMP, the discretisation that incflo's nodal
 * projection (src/projection/incflo_apply_nodal_projection.cpp:95-267) hands
 * to AMReX's MLNodeLaplacian/MLMG and its MAC projection
 * (src/convection/incflo_compute_MAC_projected_velocities.cpp:96-300) hands to
 * MLABecLaplacian/MLMG.  AMReX/AMReX-Hydro are not available, so the operator
 * definitions are the ones written out in the repository's design notes
 * (Q1-FEM nodal operator, 7-point MAC operator) and are pinned by closed
 * forms and independent assemblies in tests/ rather than by an import.
 *
 * Array conventions: x fastest, idx = i + nx*(j + ny*k).
 *   cells : n[0]*n[1]*n[2]
 *   nodes : nn[d] = periodic[d] ? n[d] : n[d]+1
 *   cell vectors with a ghost layer: (n[d]+2) per dim, component-major.
 */
#ifndef ORC_H
#define ORC_H

#ifdef __cplusplus
extern "C" {
#endif

/* Boundary types; same numbering as include/mgx.h (mirrors amrex::LinOpBCType). */
enum { ORC_BC_PERIODIC = 0, ORC_BC_DIRICHLET = 1, ORC_BC_NEUMANN = 2,
       ORC_BC_INFLOW = 3, ORC_BC_ROBIN = 4 };

typedef struct {
    int n[3];        /* cells per direction            */
    int periodic[3]; /* 1 if periodic                  */
    int bclo[3];     /* ORC_BC_* for non-periodic dims */
    int bchi[3];
    double dx[3];    /* cell size                      */
} orc_geom;

typedef struct {
    int max_iter;          /* outer V-cycles (nodal_proj.maxiter, default 100) */
    int bottom_max_iter;   /* BiCGStab iterations (default 100)               */
    double bottom_rtol;    /* BiCGStab relative tolerance (default 1e-4)      */
    int nu1, nu2;          /* pre/post smoothing sweeps                       */
    double omega;          /* weighted-Jacobi factor                          */
    int max_coarsening_level;
    int verbose;
} orc_mg_opts;

typedef struct {
    int iters;
    int converged;
    double bnorm;      /* ||b||_inf after the singular mean fix */
    double resnorm0;   /* ||b - L x0||_inf                       */
    double resnorm;    /* final ||b - L x||_inf                  */
    int nlevels;
    int bottom_iters;  /* BiCGStab iterations summed over the solve */
} orc_solve_info;

void orc_default_opts(orc_mg_opts* o);

/* ---- nodal operator pieces ------------------------------------------- */
void orc_nodal_apply(const orc_geom* g, const double* sigma, const double* phi, double* out);
void orc_nodal_diag(const orc_geom* g, const double* sigma, double* diag);
int  orc_is_dirichlet_node(const orc_geom* g, int i, int j, int k);
int  orc_nodal_is_singular(const orc_geom* g);
void orc_nodal_residual(const orc_geom* g, const double* sigma, const double* b,
                        const double* x, double* r);
void orc_nodal_jacobi(const orc_geom* g, const double* sigma, const double* b,
                      const double* xin, double* xout, double omega);
void orc_nodal_restrict(const orc_geom* gf, const double* rf, double* rc);
void orc_nodal_prolong_add(const orc_geom* gf, const double* xc, double* xf);
void orc_sigma_coarsen(const orc_geom* gf, const double* sf, double* sc);
int  orc_nodal_nlevels(const orc_geom* g, int max_coarsening_level);
double orc_nodal_omega(const orc_geom* g, double omega);

/* D.u (nodal, from cell velocity with one ghost layer) and G.phi (cell). */
void orc_fill_vel_ghosts(const orc_geom* g, double* ug, const double* inflow_vel /*6x3 or NULL*/);
void orc_nodal_divergence(const orc_geom* g, const double* ug, double* rhs);
void orc_nodal_gradient(const orc_geom* g, const double* phi, double* gphi);

/* ---- nodal MLMG ------------------------------------------------------- */
int orc_nodal_solve(const orc_geom* g, const double* sigma, const double* b_in,
                    double* x, double rtol, double atol, const orc_mg_opts* o,
                    orc_solve_info* info);
int orc_nodal_bicgstab(const orc_geom* g, const double* sigma, const double* b,
                       double* x, double rtol, int maxiter, int* iters_out);

/* Whole nodal projection (core of incflo::ApplyNodalProjection):
 * vel_g: cell velocity with a ghost layer (ghosts already filled);
 * sigma: cell coefficient; on return vel_g valid cells hold u - sigma*G(phi). */
int orc_nodal_project(const orc_geom* g, const double* sigma, double* vel_g,
                      double* phi, double* gphi, double rtol, double atol,
                      const orc_mg_opts* o, orc_solve_info* info);

/* ---- MAC projection (cell-centred variable-beta Poisson on faces) ----- */
long orc_mac_face_count(const orc_geom* g, int d);
int  orc_mac_is_singular(const orc_geom* g);
void orc_mac_apply(const orc_geom* g, const double* bx, const double* by, const double* bz,
                   const double* phi, double* out);
void orc_mac_residual(const orc_geom* g, const double* bx, const double* by, const double* bz,
                      const double* b, const double* x, double* r);
void orc_mac_gsrb(const orc_geom* g, const double* bx, const double* by, const double* bz,
                  const double* b, double* x);
void orc_mac_restrict(const orc_geom* gf, const double* rf, double* rc);
void orc_mac_prolong_add(const orc_geom* gf, const double* xc, double* xf);
void orc_mac_beta_coarsen(const orc_geom* gf, const double* bf, int d, double* bc);
void orc_mac_beta_from_density(const orc_geom* g, const double* rho, double dt, double* bx,
                               double* by, double* bz);
void orc_mac_divergence(const orc_geom* g, const double* u, const double* v, const double* w,
                        double* div);
void orc_mac_fluxes(const orc_geom* g, const double* bx, const double* by, const double* bz,
                    const double* phi, double* fx, double* fy, double* fz);
int  orc_mac_bicgstab(const orc_geom* g, const double* bx, const double* by, const double* bz,
                      const double* b, double* x, double rtol, int maxiter, int* iters_out);
int  orc_mac_solve(const orc_geom* g, const double* bx, const double* by, const double* bz,
                   const double* b_in, double* x, double rtol, double atol,
                   const orc_mg_opts* o, orc_solve_info* info);
int  orc_mac_project(const orc_geom* g, const double* bx, const double* by, const double* bz,
                     double* u, double* v, double* w, double* phi, double rtol, double atol,
                     const orc_mg_opts* o, orc_solve_info* info);

/* ---- learned predictor: 3x3x3 convolution and U-net (orc_unet.c) ------- */
void orc_conv3d_fwd(const double* in, const double* w, const double* bias, double* out, int nx,
                    int ny, int nz, int batch, int cin, int cout, const int per[3], int relu,
                    int bf16);
void orc_conv3d_bwd(const double* in, const double* w, const double* dout, double* din,
                    double* dw, double* db, int nx, int ny, int nz, int batch, int cin, int cout,
                    const int per[3], int bf16);
long orc_unet_nparam(int base, int levels, int cin0);
void orc_unet_forward(const double* p, int base, int levels, int cin0, const double* in,
                      double* out, int nx, int ny, int nz, int batch, const int per[3], int bf16);
void orc_unet_grad(const double* p, int base, int levels, int cin0, const double* in,
                   const double* dout, double* grad, int nx, int ny, int nz, int batch,
                   const int per[3], int bf16);

/* ---- Godunov PPM advection (orc_godunov.c), periodic canonical arrays --- */
void orc_ppm_predict(const orc_geom* g, const double* s, int d, const double* ucell,
                     const double* uface, double dt, double* Im, double* Ip);
void orc_godunov_extrap_vel(const orc_geom* g, const double* u, const double* v, const double* w,
                            const double* fu, const double* fv, const double* fw, double dt,
                            double* umac, double* vmac, double* wmac);
void orc_godunov_edge_states(const orc_geom* g, const double* u, const double* v, const double* w,
                             const double* fu, const double* fv, const double* fw,
                             const double* umac, const double* vmac, const double* wmac,
                             double dt, double* states);
void orc_godunov_scalar_states(const orc_geom* g, const double* s, int ncomp, const int* iconserv,
                               const double* force, const double* umac, const double* vmac,
                               const double* wmac, double dt, double* states);
void orc_scalar_advection(const orc_geom* g, const double* s, int ncomp, const int* iconserv,
                          const double* umac, const double* vmac, const double* wmac,
                          const double* states, double* conv);
/* general form: ghosted cell states (2 layers), amrex::BCType per component
 * and side (bc[6 n + 3 side + d]), faces with n_d + 1 entries along d */
void orc_gd_extrap_vel(const orc_geom* g, const double* vel_g, const int* bc, const double* force,
                       double dt, double* umac, double* vmac, double* wmac);
void orc_gd_edge_states(const orc_geom* g, int ncomp, const double* s_g, const int* bc,
                        const int* iconserv, const double* force, const double* umac,
                        const double* vmac, const double* wmac, double dt, double* sx, double* sy,
                        double* sz);
void orc_gd_conv(const orc_geom* g, int ncomp, const double* s, const int* iconserv,
                 const double* umac, const double* vmac, const double* wmac, const double* sx,
                 const double* sy, const double* sz, double* conv);
void orc_convective_term(const orc_geom* g, const double* u, const double* v, const double* w,
                         const double* umac, const double* vmac, const double* wmac,
                         const double* states, double* conv);

/* ---- embedded boundary, nodal (orc_eb.c) ------------------------------- */
void orc_eb_cylinder_boundary(const orc_geom* g, const double prob_lo[3], double cx, double cy,
                              double r, const int* ctype, const double* mom, double* B);
void orc_eb_inflow_rhs(const orc_geom* g, const int* ctype, const double* B, const double* ueb,
                       double* rhs);
/* ctype per cell: 0 covered, 1 regular, 2 cut; mom: 27 fluid moments per cell */
void orc_eb_cylinder(const orc_geom* g, const double prob_lo[3], double cx, double cy, double r,
                     int* ctype, double* mom);
void orc_eb_element(const double* mom, const double dx[3], double E[64]);
void orc_eb_node_mask(const orc_geom* g, const int* ctype, int* mask);
void orc_eb_nodal_apply(const orc_geom* g, const int* ctype, const double* mom,
                        const double* sigma, const double* phi, double* out);
void orc_eb_nodal_divergence(const orc_geom* g, const int* ctype, const double* mom,
                             const double* ug, double* rhs);
void orc_eb_nodal_gradient(const orc_geom* g, const int* ctype, const double* mom,
                           const double* phi, double* gphi);
int  orc_eb_nodal_solve(const orc_geom* g, const int* ctype, const double* mom,
                        const double* sigma, const double* b_in, double* x, double rtol,
                        int maxiter, int* iters);

/* ---- helpers ---------------------------------------------------------- */
double orc_norm_inf(const double* a, long n);
double orc_dot(const double* a, const double* b, long n);
double orc_mean(const double* a, long n);

#ifdef __cplusplus
}
#endif
#endif
