"""MAC operator with Robin ("mixed") domain faces: FP64 CPU oracle.

TEST INFRASTRUCTURE ONLY (the checker for the GPU Robin MAC solve); the
product library never loads it.

incflo maps a `mixed` domain face to LinOpBCType::Robin for the MAC
projection (src/projection/incflo_projection_bc.cpp:69-73) and hands MLMG the
coefficients of

    a phi + b dphi/dn = f        (n the outward normal)

stored in the ghost cells just outside the face
(src/boundary_conditions/incflo_set_bcs.cpp:122-190, setLevelBC at
src/convection/incflo_compute_MAC_projected_velocities.cpp:115-122).  For the
MAC projection the data are a = 1, b = 0 on the outflow part and a = 0, b = 1
on the inflow part, f = 0 (src/prob/prob_bc.cpp:97-197).

The cell-centred discretisation restated here: the face value of phi is the
mean of the boundary cell and its ghost, the outward derivative their
difference over h, so the Robin relation fixes the ghost and

    dphi/dn |face = (f - a phi_0) / (a h / 2 + b)

(a = 1, b = 0 gives the Dirichlet face of the MAC operator, 2 phi_0 / h; a = 0
gives a Neumann face with flux f / b).  The face flux is beta times the
derivative along +x_d, so F_hi = beta dphi/dn and F_lo = -beta dphi/dn, and
(L phi)_c = sum_d (F_hi - F_lo) / h_d as for the other faces (orc_mac.c).

Robin data are per face plane: robin[(d, side)] = (a, b, f), each array
shaped like the face plane of direction d in numpy order (for d = 0:
(nz, ny); d = 1: (nz, nx); d = 2: (ny, nx)).
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import BC_DIRICHLET, BC_ROBIN


def plane_shape(g, d):
    n = [g.n[2], g.n[1], g.n[0]]  # numpy axes z, y, x
    ax = 2 - d
    return tuple(n[a] for a in range(3) if a != ax)


def _operator(g, beta, robin):
    """Sparse linear part A and constant part k of L: L phi = A phi + k."""
    nz, ny, nx = g.cell_shape()
    N = nx * ny * nz
    idx = np.arange(N).reshape(nz, ny, nx)
    rows, cols, vals = [], [], []
    k = np.zeros((nz, ny, nx))
    diag = np.zeros((nz, ny, nx))
    for d in range(3):
        ax = 2 - d
        h = g.dx[d]
        ih2 = 1.0 / (h * h)
        b = beta[d]
        n = g.n[d]
        # interior faces 1..n-1 along the axis
        sl_l = [slice(None)] * 3
        sl_r = [slice(None)] * 3
        sl_f = [slice(None)] * 3
        sl_l[ax] = slice(0, n - 1)
        sl_r[ax] = slice(1, n)
        sl_f[ax] = slice(1, n)
        bf = b[tuple(sl_f)] * ih2
        L_id, R_id = idx[tuple(sl_l)], idx[tuple(sl_r)]
        rows += [L_id.ravel(), R_id.ravel()]
        cols += [R_id.ravel(), L_id.ravel()]
        vals += [bf.ravel(), bf.ravel()]
        diag[tuple(sl_l)] -= bf
        diag[tuple(sl_r)] -= bf
        for side in (0, 1):
            fi = 0 if side == 0 else n
            ci = 0 if side == 0 else n - 1
            sl_b = [slice(None)] * 3
            sl_b[ax] = fi
            sl_c = [slice(None)] * 3
            sl_c[ax] = ci
            bb = b[tuple(sl_b)]
            if g.periodic[d]:
                if side == 1:
                    oc = [slice(None)] * 3
                    oc[ax] = 0
                    bw = bb * ih2
                    A_id, B_id = idx[tuple(sl_c)], idx[tuple(oc)]
                    rows += [A_id.ravel(), B_id.ravel()]
                    cols += [B_id.ravel(), A_id.ravel()]
                    vals += [bw.ravel(), bw.ravel()]
                    diag[tuple(sl_c)] -= bw
                    diag[tuple(oc)] -= bw
                continue
            bc = g.bchi[d] if side else g.bclo[d]
            if bc == BC_DIRICHLET:
                diag[tuple(sl_c)] -= 2.0 * bb * ih2
            elif bc == BC_ROBIN:
                a, bb_, f = robin[(d, side)]
                den = a * h / 2.0 + bb_
                # (F_hi - F_lo)/h picks +beta dphi/dn / h on either side
                diag[tuple(sl_c)] -= bb * a / den / h
                k[tuple(sl_c)] += bb * f / den / h
    rows.append(idx.ravel())
    cols.append(idx.ravel())
    vals.append(diag.ravel())
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(N, N))
    return A, k


def robin_apply(g, beta, robin, phi):
    """L phi (including the constant Robin term)."""
    A, k = _operator(g, beta, robin)
    return (A @ phi.ravel()).reshape(g.cell_shape()) + k


def robin_solve(g, beta, robin, rhs):
    """phi with L phi = rhs (direct sparse solve; the problem must be
    non-singular: a Dirichlet face or a Robin face with some a > 0)."""
    A, k = _operator(g, beta, robin)
    x = spla.spsolve(A.tocsc(), (rhs - k).ravel())
    return x.reshape(g.cell_shape())


def robin_fluxes(g, beta, robin, phi):
    """F_d = beta dphi/dx_d on every face (Robin faces by the relation above)."""
    out = []
    for d in range(3):
        ax = 2 - d
        h = g.dx[d]
        n = g.n[d]
        F = np.zeros(beta[d].shape)
        sl_f = [slice(None)] * 3
        sl_f[ax] = slice(1, n)
        sl_l = [slice(None)] * 3
        sl_l[ax] = slice(0, n - 1)
        sl_r = [slice(None)] * 3
        sl_r[ax] = slice(1, n)
        F[tuple(sl_f)] = beta[d][tuple(sl_f)] * (phi[tuple(sl_r)] - phi[tuple(sl_l)]) / h
        for side in (0, 1):
            fi = 0 if side == 0 else n
            ci = 0 if side == 0 else n - 1
            sl_b = [slice(None)] * 3
            sl_b[ax] = fi
            sl_c = [slice(None)] * 3
            sl_c[ax] = ci
            bb = beta[d][tuple(sl_b)]
            p0 = phi[tuple(sl_c)]
            if g.periodic[d]:
                o0, o1 = [slice(None)] * 3, [slice(None)] * 3
                o0[ax], o1[ax] = 0, n - 1
                F[tuple(sl_b)] = bb * (phi[tuple(o0)] - phi[tuple(o1)]) / h
                continue
            bc = g.bchi[d] if side else g.bclo[d]
            if bc == BC_DIRICHLET:
                dn = -2.0 * p0 / h
            elif bc == BC_ROBIN:
                a, b_, f = robin[(d, side)]
                dn = (f - a * p0) / (a * h / 2.0 + b_)
            else:
                dn = np.zeros_like(p0)
            F[tuple(sl_b)] = bb * dn if side else -bb * dn
        out.append(F)
    return out


def robin_project(g, beta, robin, u):
    """MacProjector::project with Robin faces: solve L phi = div(u), then
    u - F on every face (Neumann faces keep u).  Returns (u_new, phi)."""
    nz, ny, nx = g.cell_shape()
    div = ((u[0][:, :, 1:] - u[0][:, :, :-1]) / g.dx[0] + (u[1][:, 1:, :] - u[1][:, :-1, :]) / g.dx[1] +
           (u[2][1:, :, :] - u[2][:-1, :, :]) / g.dx[2])
    phi = robin_solve(g, beta, robin, div)
    F = robin_fluxes(g, beta, robin, phi)
    return [u[d] - F[d] for d in range(3)], phi
