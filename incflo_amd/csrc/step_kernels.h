// step_kernels.h -- device kernels of the predictor step around the two
// projections: forces, the velocity and density updates, periodic ghost
// filling, and the cell-centred viscous operator / Helmholtz right-hand side.
// Boxed fields are the caller's (box-offset table, one tile of 256 cells of
// one box per workgroup); the viscous kernels work on canonical level arrays
// (x fastest, no ghosts) with face coefficient arrays of n_d + 1 faces along d.
#pragma once

#include <hip/hip_runtime.h>

#include "box_kernels.h"

namespace mgx {

// ---- boxed <-> canonical, one component of an ncomp field -----------------
static __global__ void __launch_bounds__(256) comp_copy_kernel(KGeom g, const BoxInfo* boxes,
                                                              const Tile* tiles, double* boxed,
                                                              int ng, int ncomp, int comp,
                                                              double* canon, int to_box) {
    int bi, i, j, k;
    if (!tile_point(boxes, tiles, REG_CELL, bi, i, j, k)) return;
    const BoxInfo& b = boxes[bi];
    const long o = field_offset(b, ng, ncomp, false) + box_index(b, ng, false, i, j, k, comp);
    const long c = i + (long)g.n[0] * (j + (long)g.n[1] * k);
    if (to_box) boxed[o] = canon[c];
    else canon[c] = boxed[o];
}

// The ghost cell just outside domain face (d, side) of component comp of a
// boxed cell field, one value per boundary face (plane element t, x fastest
// over the two remaining directions); NaN where no box holds the adjacent cell.
static __global__ void __launch_bounds__(256) comp_face_ghost_kernel(KGeom g, const BoxInfo* boxes,
                                                                    int nboxes,
                                                                    const double* __restrict__ in,
                                                                    int ng, int ncomp, int comp,
                                                                    int d, int side,
                                                                    double* __restrict__ out) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int a0 = d == 0 ? 1 : 0, a1 = d == 2 ? 1 : 2;
    const long m0 = g.n[a0];
    if (t >= m0 * g.n[a1]) return;
    int p[3];
    p[d] = side ? g.n[d] - 1 : 0;
    p[a0] = (int)(t % m0);
    p[a1] = (int)(t / m0);
    double v = __builtin_nan("");
    for (int bi = 0; bi < nboxes; ++bi) {
        const BoxInfo& b = boxes[bi];
        if (p[0] < b.lo[0] || p[0] >= b.lo[0] + b.len[0] || p[1] < b.lo[1] ||
            p[1] >= b.lo[1] + b.len[1] || p[2] < b.lo[2] || p[2] >= b.lo[2] + b.len[2])
            continue;
        int q[3] = {p[0], p[1], p[2]};
        q[d] += side ? 1 : -1;
        v = in[field_offset(b, ng, ncomp, false) + box_index(b, ng, false, q[0], q[1], q[2], comp)];
        break;
    }
    out[t] = v;
}

// Periodic ghost filling (FillBoundary with the domain's periodicity) in two
// passes through a canonical copy of all components: valid cells -> canon,
// then every grown cell of every box whose (wrapped) index lies inside the
// domain <- canon.  Ghosts outside non-periodic faces are left as they are.
static __global__ void __launch_bounds__(256) canon_all_kernel(KGeom g, const BoxInfo* boxes,
                                                              const Tile* tiles,
                                                              const double* __restrict__ boxed,
                                                              int ng, int ncomp,
                                                              double* __restrict__ canon) {
    int bi, i, j, k;
    if (!tile_point(boxes, tiles, REG_CELL, bi, i, j, k)) return;
    const BoxInfo& b = boxes[bi];
    const long base = field_offset(b, ng, ncomp, false);
    const long nc = (long)g.n[0] * g.n[1] * g.n[2];
    const long c = i + (long)g.n[0] * (j + (long)g.n[1] * k);
    for (int q = 0; q < ncomp; ++q) canon[q * nc + c] = boxed[base + box_index(b, ng, false, i, j, k, q)];
}

// one thread per cell of a box grown by ng (tiles over the grown boxes come
// from the host, one table per ghost width)
static __global__ void __launch_bounds__(256) fill_ghosts_kernel(KGeom g, const BoxInfo* boxes,
                                                                const Tile* tiles,
                                                                double* __restrict__ boxed, int ng,
                                                                int ncomp,
                                                                const double* __restrict__ canon) {
    const Tile t = tiles[blockIdx.x];
    const BoxInfo& b = boxes[t.box];
    int len[3], lo[3];
    for (int d = 0; d < 3; ++d) { lo[d] = b.lo[d] - ng; len[d] = b.len[d] + 2 * ng; }
    const int lin = t.start + (int)threadIdx.x;
    if (lin >= len[0] * len[1] * len[2]) return;
    const int qq = lin / len[0];
    int p[3] = {lo[0] + (lin - qq * len[0]), lo[1] + qq % len[1], lo[2] + qq / len[1]};
    bool valid = true;
    for (int d = 0; d < 3; ++d)
        if (p[d] < b.lo[d] || p[d] >= b.lo[d] + b.len[d]) valid = false;
    if (valid) return;
    int w[3];
    for (int d = 0; d < 3; ++d) {
        w[d] = p[d];
        if (w[d] < 0 || w[d] >= g.n[d]) {
            if (!g.per[d]) return;  // physical boundary ghost: not ours
            w[d] = ((w[d] % g.n[d]) + g.n[d]) % g.n[d];
        }
    }
    const long nc = (long)g.n[0] * g.n[1] * g.n[2];
    const long c = w[0] + (long)g.n[0] * (w[1] + (long)g.n[1] * w[2]);
    const long base = field_offset(b, ng, ncomp, false);
    for (int q = 0; q < ncomp; ++q) boxed[base + box_index(b, ng, false, p[0], p[1], p[2], q)] = canon[q * nc + c];
}

// the same fill with a (chunk, box) grid: blockIdx.y is the box, blockIdx.x
// a 256-cell chunk of its grown region (chunks past a smaller box return)
static __global__ void __launch_bounds__(256) fill_ghosts_boxgrid_kernel(
    KGeom g, const BoxInfo* boxes, double* __restrict__ boxed, int ng, int ncomp,
    const double* __restrict__ canon) {
    const BoxInfo& b = boxes[blockIdx.y];
    int len[3], lo[3];
    for (int d = 0; d < 3; ++d) { lo[d] = b.lo[d] - ng; len[d] = b.len[d] + 2 * ng; }
    const long lin = (long)blockIdx.x * 256 + threadIdx.x;
    if (lin >= (long)len[0] * len[1] * len[2]) return;
    const int qq = (int)(lin / len[0]);
    int p[3] = {lo[0] + (int)(lin - (long)qq * len[0]), lo[1] + qq % len[1], lo[2] + qq / len[1]};
    bool valid = true;
    for (int d = 0; d < 3; ++d)
        if (p[d] < b.lo[d] || p[d] >= b.lo[d] + b.len[d]) valid = false;
    if (valid) return;
    int w[3];
    for (int d = 0; d < 3; ++d) {
        w[d] = p[d];
        if (w[d] < 0 || w[d] >= g.n[d]) {
            if (!g.per[d]) return;  // physical boundary ghost: not ours
            w[d] = ((w[d] % g.n[d]) + g.n[d]) % g.n[d];
        }
    }
    const long nc = (long)g.n[0] * g.n[1] * g.n[2];
    const long c = w[0] + (long)g.n[0] * (w[1] + (long)g.n[1] * w[2]);
    const long base = field_offset(b, ng, ncomp, false);
    for (int q = 0; q < ncomp; ++q) boxed[base + box_index(b, ng, false, p[0], p[1], p[2], q)] = canon[q * nc + c];
}

// ---- predictor step pieces (boxed, valid cells, 3 components) -------------
struct ForceArgs {
    const BoxInfo* boxes;
    const Tile* tiles;
    const double* gp; int gp_ng;
    const double* rho; int rho_ng;
    const double* divtau; int dt_ng;  // may be null
    double* out; int out_ng;
    double gravity[3];
    double gp0[3];
    int include_gp;
};

// compute_vel_forces (src/incflo_compute_forces.cpp:134-148):
//   f = -(gp + gp0)/rho + g   (or g - gp0/rho without the pressure gradient),
// plus divtau when given (the Godunov forcing of
// src/convection/incflo_compute_MAC_projected_velocities.cpp:24-57)
static __global__ void __launch_bounds__(256) vel_forces_kernel(ForceArgs a) {
    int bi, i, j, k;
    if (!tile_point(a.boxes, a.tiles, REG_CELL, bi, i, j, k)) return;
    const BoxInfo& b = a.boxes[bi];
    const double rinv = 1.0 / a.rho[field_offset(b, a.rho_ng, 1, false) +
                                     box_index(b, a.rho_ng, false, i, j, k, 0)];
    const long og = field_offset(b, a.gp_ng, 3, false);
    const long oo = field_offset(b, a.out_ng, 3, false);
    const long od = a.divtau ? field_offset(b, a.dt_ng, 3, false) : 0;
    for (int c = 0; c < 3; ++c) {
        const double gpc = a.include_gp ? a.gp[og + box_index(b, a.gp_ng, false, i, j, k, c)] : 0.0;
        double f = -(gpc + a.gp0[c]) * rinv + a.gravity[c];
        if (a.divtau) f += a.divtau[od + box_index(b, a.dt_ng, false, i, j, k, c)];
        a.out[oo + box_index(b, a.out_ng, false, i, j, k, c)] = f;
    }
}

struct UpdVelArgs {
    const BoxInfo* boxes;
    const Tile* tiles;
    double* vel; int vel_ng;
    const double* vel_old; int vo_ng;
    const double* conv; int conv_ng;
    const double* force; int f_ng;
    const double* divtau; int dt_ng;     // may be null
    const double* rho_old; int ro_ng;    // momentum form only
    const double* rho_new; int rn_ng;
    const double* rho_nph; int rh_ng;
    double dt;
    double divtau_coef;                  // 1 explicit, 0.5 Crank-Nicolson, 0 implicit
    int momentum;
};

// update_velocity, predictor (src/incflo_update_velocity.cpp:14-143):
//   velocity form: u = u_old + dt (conv + f + c divtau_o)
//   momentum form: u = (rho_old u_old + dt (conv + rho_nph f + c divtau_o)) / rho_new
static __global__ void __launch_bounds__(256) update_velocity_kernel(UpdVelArgs a) {
    int bi, i, j, k;
    if (!tile_point(a.boxes, a.tiles, REG_CELL, bi, i, j, k)) return;
    const BoxInfo& b = a.boxes[bi];
    double ro = 1.0, rn = 1.0, rh = 1.0;
    if (a.momentum) {
        ro = a.rho_old[field_offset(b, a.ro_ng, 1, false) + box_index(b, a.ro_ng, false, i, j, k, 0)];
        rn = a.rho_new[field_offset(b, a.rn_ng, 1, false) + box_index(b, a.rn_ng, false, i, j, k, 0)];
        rh = a.rho_nph[field_offset(b, a.rh_ng, 1, false) + box_index(b, a.rh_ng, false, i, j, k, 0)];
    }
    const long ov = field_offset(b, a.vel_ng, 3, false), oo = field_offset(b, a.vo_ng, 3, false);
    const long oc = field_offset(b, a.conv_ng, 3, false), of = field_offset(b, a.f_ng, 3, false);
    const long od = a.divtau ? field_offset(b, a.dt_ng, 3, false) : 0;
    for (int c = 0; c < 3; ++c) {
        const double uo = a.vel_old[oo + box_index(b, a.vo_ng, false, i, j, k, c)];
        const double cv = a.conv[oc + box_index(b, a.conv_ng, false, i, j, k, c)];
        const double f = a.force[of + box_index(b, a.f_ng, false, i, j, k, c)];
        const double dv = a.divtau ? a.divtau_coef * a.divtau[od + box_index(b, a.dt_ng, false, i, j, k, c)]
                                   : 0.0;
        double u;
        if (a.momentum) u = (ro * uo + a.dt * (cv + rh * f + dv)) / rn;
        else u = uo + a.dt * (cv + f + dv);
        a.vel[ov + box_index(b, a.vel_ng, false, i, j, k, c)] = u;
    }
}

// update_density, predictor (src/incflo_update_density.cpp:14-64):
//   rho_new = rho_old + dt drdt;  rho_nph = (rho_old + rho_new) / 2
static __global__ void __launch_bounds__(256) update_density_kernel(
    const BoxInfo* boxes, const Tile* tiles, const double* __restrict__ rho_old, int ro_ng,
    const double* __restrict__ drdt, int c_ng, double dt, double* __restrict__ rho_new, int rn_ng,
    double* __restrict__ rho_nph, int rh_ng) {
    int bi, i, j, k;
    if (!tile_point(boxes, tiles, REG_CELL, bi, i, j, k)) return;
    const BoxInfo& b = boxes[bi];
    const double ro = rho_old[field_offset(b, ro_ng, 1, false) + box_index(b, ro_ng, false, i, j, k, 0)];
    const double rn = ro + dt * drdt[field_offset(b, c_ng, 1, false) + box_index(b, c_ng, false, i, j, k, 0)];
    rho_new[field_offset(b, rn_ng, 1, false) + box_index(b, rn_ng, false, i, j, k, 0)] = rn;
    if (rho_nph)
        rho_nph[field_offset(b, rh_ng, 1, false) + box_index(b, rh_ng, false, i, j, k, 0)] =
            0.5 * (ro + rn);
}

// ---- cell-centred viscous operator on canonical arrays ---------------------
struct DiffGeomK {
    int n[3];
    int per[3];
    int dlo[3], dhi[3];  // Dirichlet sides (boundary value from the planes)
    double ih2[3];
};

__device__ __forceinline__ long dcell(const DiffGeomK& g, int i, int j, int k) {
    return i + (long)g.n[0] * (j + (long)g.n[1] * k);
}

__device__ __forceinline__ long dface(const DiffGeomK& g, int d, int i, int j, int k) {
    const long f0 = g.n[0] + (d == 0), f1 = g.n[1] + (d == 1);
    return i + f0 * (j + f1 * (long)k);
}

__device__ __forceinline__ long dplane(const DiffGeomK& g, int d, int i, int j, int k) {
    const int p[3] = {i, j, k};
    const int a0 = d == 0 ? 1 : 0, a1 = d == 2 ? 1 : 2;
    return p[a0] + (long)g.n[a0] * p[a1];
}

// eta on the faces of direction d: the mean of the two cells; a non-periodic
// domain face takes the adjacent cell's value (average_scalar_eta_to_faces,
// src/diffusion/incflo_diffusion.cpp:289-380).  eta null: the constant ceta.
static __global__ void __launch_bounds__(256) eta_faces_kernel(DiffGeomK g, int d,
                                                              const double* __restrict__ eta,
                                                              double ceta, double scale,
                                                              double* __restrict__ out) {
    const int fx = g.n[0] + (d == 0), fy = g.n[1] + (d == 1), fz = g.n[2] + (d == 2);
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)fx * fy * fz) return;
    int p[3] = {(int)(t % fx), (int)((t / fx) % fy), (int)(t / ((long)fx * fy))};
    double v = ceta;
    if (eta) {
        int L[3] = {p[0], p[1], p[2]}, R[3] = {p[0], p[1], p[2]};
        L[d] = p[d] - 1;
        if (g.per[d]) {
            if (L[d] < 0) L[d] += g.n[d];
            if (R[d] >= g.n[d]) R[d] -= g.n[d];
        } else {
            if (L[d] < 0) L[d] = R[d];
            if (R[d] >= g.n[d]) R[d] = L[d];
        }
        v = 0.5 * (eta[dcell(g, L[0], L[1], L[2])] + eta[dcell(g, R[0], R[1], R[2])]);
    }
    out[t] = scale * v;
}

// div(B grad u) at cell (i,j,k) with homogeneous boundary data; bnd receives
// the inhomogeneous part sum over Dirichlet sides of 2 B_f g / h^2
__device__ __forceinline__ double diff_apply_cell(const DiffGeomK& g, const double* const B[3],
                                                  const double* u, const double* const pl[6],
                                                  int i, int j, int k, double& bnd) {
    const int p[3] = {i, j, k};
    const double uc = u[dcell(g, i, j, k)];
    double acc = 0.0;
    bnd = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        for (int side = 0; side < 2; ++side) {
            int f[3] = {p[0], p[1], p[2]};
            if (side) f[d] += 1;
            const double bf = B[d][dface(g, d, f[0], f[1], f[2])] * g.ih2[d];
            const bool edge = side ? (p[d] == g.n[d] - 1) : (p[d] == 0);
            if (!edge || g.per[d]) {
                int q[3] = {p[0], p[1], p[2]};
                q[d] += side ? 1 : -1;
                if (q[d] < 0) q[d] += g.n[d];
                if (q[d] >= g.n[d]) q[d] -= g.n[d];
                acc += bf * (u[dcell(g, q[0], q[1], q[2])] - uc);
            } else if (side ? g.dhi[d] : g.dlo[d]) {
                acc -= 2.0 * bf * uc;
                const double* P = pl[2 * d + side];
                if (P) bnd += 2.0 * bf * P[dplane(g, d, i, j, k)];
            }
        }
    }
    return acc;
}

// out = (div(B grad u) incl. boundary values) * (rho ? 1/rho : 1)
static __global__ void __launch_bounds__(256) diff_apply_kernel(DiffGeomK g, const double* bx,
                                                               const double* by, const double* bz,
                                                               const double* __restrict__ u,
                                                               const double* p0, const double* p1,
                                                               const double* p2, const double* p3,
                                                               const double* p4, const double* p5,
                                                               const double* __restrict__ rho,
                                                               double* __restrict__ out) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nc = (long)g.n[0] * g.n[1] * g.n[2];
    if (t >= nc) return;
    const int i = (int)(t % g.n[0]), j = (int)((t / g.n[0]) % g.n[1]),
              k = (int)(t / ((long)g.n[0] * g.n[1]));
    const double* B[3] = {bx, by, bz};
    const double* P[6] = {p0, p1, p2, p3, p4, p5};
    double bnd;
    const double v = diff_apply_cell(g, B, u, P, i, j, k, bnd) + bnd;
    out[t] = rho ? v / rho[t] : v;
}

// Helmholtz right-hand side of the solver's form div(B grad u) - A u = rhs:
//   rhs = -(a u_old) - sum over Dirichlet sides of 2 B_f g / h^2,
// with A = a (rho, or 1 when null) and B = dt eta_f (already scaled)
static __global__ void __launch_bounds__(256) diff_rhs_kernel(DiffGeomK g, const double* bx,
                                                             const double* by, const double* bz,
                                                             const double* __restrict__ u,
                                                             const double* p0, const double* p1,
                                                             const double* p2, const double* p3,
                                                             const double* p4, const double* p5,
                                                             const double* __restrict__ a,
                                                             double* __restrict__ rhs) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nc = (long)g.n[0] * g.n[1] * g.n[2];
    if (t >= nc) return;
    const int i = (int)(t % g.n[0]), j = (int)((t / g.n[0]) % g.n[1]),
              k = (int)(t / ((long)g.n[0] * g.n[1]));
    const int p[3] = {i, j, k};
    const double* B[3] = {bx, by, bz};
    const double* P[6] = {p0, p1, p2, p3, p4, p5};
    double bnd = 0.0;
    for (int d = 0; d < 3; ++d) {
        if (g.per[d]) continue;
        for (int side = 0; side < 2; ++side) {
            const bool edge = side ? (p[d] == g.n[d] - 1) : (p[d] == 0);
            const double* Pl = P[2 * d + side];
            if (!edge || !(side ? g.dhi[d] : g.dlo[d]) || !Pl) continue;
            int f[3] = {p[0], p[1], p[2]};
            if (side) f[d] += 1;
            bnd += 2.0 * B[d][dface(g, d, f[0], f[1], f[2])] * g.ih2[d] * Pl[dplane(g, d, i, j, k)];
        }
    }
    rhs[t] = -(a ? a[t] : 1.0) * u[t] - bnd;
}

static __global__ void __launch_bounds__(256) scale_copy_kernel(const double* __restrict__ in,
                                                               double s, double* __restrict__ out,
                                                               long n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = s * in[t];
}

static __global__ void __launch_bounds__(256) const_fill_kernel(double* a, long n, double v) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) a[t] = v;
}

}  // namespace mgx
