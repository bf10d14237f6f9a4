"""Cell-centred viscous operator and Helmholtz solve: FP64 CPU oracle.

TEST INFRASTRUCTURE ONLY (the checker for the GPU diffusion solver); the
product library never loads it.

incflo's implicit viscous step with separate solves per velocity component
(DiffusionScalarOp::diffuse_vel_components,
src/diffusion/DiffusionScalarOp.cpp:332-480, called from
src/incflo_update_velocity.cpp:322-334 through
src/diffusion/incflo_diffusion.cpp:76-89) solves, per component,

    alpha a u - beta div(b grad u) = rhs,   alpha = 1, a = rho, beta = dt,
    b = eta on faces, rhs = rho u*

with MLMG (MLABecLaplacian), and the explicit form (compute_divtau, scalar
branch, :584-760) is divtau = div(eta grad u) / rho.  Face coefficients come
from average_scalar_eta_to_faces (src/diffusion/incflo_diffusion.cpp:289-309):
the arithmetic mean of the two cells, and on a non-periodic domain face the
adjacent interior cell's value (fixup_eta_on_domain_faces, :311-380).

The discretisation restated here (this build's own, second order):
    F_f = b_f (u_R - u_L) / h            interior and periodic faces
    F_lo = b_f (u_0 - g) / (h/2)         Dirichlet face, g the boundary value
    F_hi = b_f (g - u_{n-1}) / (h/2)
    F = 0                                Neumann face (homogeneous)
    (L u)_c = sum_d (F_hi - F_lo) / h_d
incflo runs these solves with mg_maxorder = 2 (src/diffusion/DiffusionScalarOp.H:65),
i.e. the ghost value is extrapolated linearly from the face value and the
first interior cell, 2 g - u_0, which is the two-point Dirichlet flux above.

Boundary values follow incflo's convention that a Dirichlet face value lives
in the ghost cell outside it (setLevelBC(lev, &phi), DiffusionScalarOp.cpp:446):
bvals[(d, side)] is an array shaped like the face plane of direction d (numpy
order, as oracle/mac_robin.py's planes), absent for homogeneous data.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import BC_DIRICHLET, BC_NEUMANN, BC_PERIODIC


def eta_faces(eta, periodic):
    """average_cellcenter_to_face + fixup_eta_on_domain_faces: cell eta
    (nz, ny, nx) -> 3 face arrays with n_d + 1 entries along d."""
    out = []
    for d in range(3):
        ax = 2 - d
        n = eta.shape[ax]
        if periodic[d]:
            lo = np.roll(eta, 1, axis=ax)
            mid = 0.5 * (lo + eta)  # face f between cells f-1 and f (f = 0 wraps)
            f = np.concatenate([mid, np.take(mid, [0], axis=ax)], axis=ax)
        else:
            inner = 0.5 * (np.take(eta, range(0, n - 1), axis=ax) +
                           np.take(eta, range(1, n), axis=ax))
            f = np.concatenate([np.take(eta, [0], axis=ax), inner,
                                np.take(eta, [n - 1], axis=ax)], axis=ax)
        out.append(f)
    return out


def _side_bc(g, d, side):
    if g.periodic[d]:
        return BC_PERIODIC
    return g.bchi[d] if side else g.bclo[d]


def operator(g, bf):
    """Sparse L (homogeneous boundary data) and the per-side coefficient
    planes k[(d, side)] with L u = A u + sum k g on the boundary cells."""
    nz, ny, nx = g.cell_shape()
    N = nx * ny * nz
    idx = np.arange(N).reshape(nz, ny, nx)
    rows, cols, vals = [], [], []
    diag = np.zeros((nz, ny, nx))
    kside = {}
    for d in range(3):
        ax = 2 - d
        n = g.n[d]
        ih2 = 1.0 / (g.dx[d] * g.dx[d])
        b = bf[d]

        def sl(a, b_):
            s = [slice(None)] * 3
            s[ax] = slice(a, b_)
            return tuple(s)

        # interior faces 1..n-1 between cells f-1 and f
        c = np.take(b, range(1, n), axis=ax) * ih2
        iL = idx[sl(0, n - 1)]
        iR = idx[sl(1, n)]
        rows += [iL.ravel(), iR.ravel()]
        cols += [iR.ravel(), iL.ravel()]
        vals += [c.ravel(), c.ravel()]
        diag[sl(0, n - 1)] -= c
        diag[sl(1, n)] -= c
        if g.periodic[d]:
            c0 = np.take(b, [0], axis=ax) * ih2  # face 0 joins cell n-1 and cell 0
            i0 = idx[sl(0, 1)]
            i1 = idx[sl(n - 1, n)]
            rows += [i0.ravel(), i1.ravel()]
            cols += [i1.ravel(), i0.ravel()]
            vals += [c0.ravel(), c0.ravel()]
            diag[sl(0, 1)] -= c0
            diag[sl(n - 1, n)] -= c0
            continue
        for side in (0, 1):
            bc = _side_bc(g, d, side)
            f = n if side else 0
            cell = sl(n - 1, n) if side else sl(0, 1)
            cb = 2.0 * np.take(b, [f], axis=ax) * ih2
            if bc == BC_DIRICHLET:
                diag[cell] -= cb
                kside[(d, side)] = np.squeeze(cb, axis=ax)
            elif bc != BC_NEUMANN:
                raise ValueError(f"diffusion oracle: unsupported BC {bc}")
    rows.append(idx.ravel())
    cols.append(idx.ravel())
    vals.append(diag.ravel())
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(N, N))
    return A, kside


def _bterm(g, kside, bvals):
    """sum over Dirichlet sides of k g on the boundary cells."""
    t = np.zeros(g.cell_shape())
    for (d, side), k in kside.items():
        v = None if bvals is None else bvals.get((d, side))
        if v is None:
            continue
        ax = 2 - d
        s = [slice(None)] * 3
        s[ax] = slice(g.n[d] - 1, g.n[d]) if side else slice(0, 1)
        t[tuple(s)] += np.expand_dims(k * v, axis=ax)
    return t


def apply(g, bf, u, bvals=None):
    """(L u)_c = div(b grad u) with the boundary values bvals."""
    A, k = operator(g, bf)
    return (A @ u.ravel()).reshape(g.cell_shape()) + _bterm(g, k, bvals)


def helmholtz_solve(g, acoef, bf, beta, rhs, bvals=None, rtol=1e-14, x0=None):
    """Solve acoef u - beta L u = rhs (acoef > 0 per cell) by conjugate
    gradients on the symmetric positive definite matrix."""
    A, k = operator(g, bf)
    M = sp.diags(acoef.ravel()) - beta * A
    r = (rhs + beta * _bterm(g, k, bvals)).ravel()
    dinv = 1.0 / M.diagonal()
    pre = spla.LinearOperator(M.shape, matvec=lambda v: dinv * v)
    x, info = spla.cg(M, r, x0=None if x0 is None else x0.ravel(), rtol=rtol, atol=0.0,
                      maxiter=20000, M=pre)
    if info != 0:
        raise RuntimeError(f"diffusion oracle: CG did not converge ({info})")
    return x.reshape(g.cell_shape())


def diffuse_velocity(g, vel, rho, eta, dt, bvals=None, rtol=1e-14):
    """DiffusionScalarOp::diffuse_vel_components with constant or cell eta:
    every component solves (rho - dt div(eta grad)) u_new = rho u_old.
    bvals: per component dict of boundary planes (or None)."""
    e = eta if np.ndim(eta) else np.full(g.cell_shape(), float(eta))
    bf = eta_faces(e, g.periodic)
    out = np.empty_like(vel)
    for c in range(vel.shape[0]):
        bv = None if bvals is None else bvals[c]
        out[c] = helmholtz_solve(g, rho, bf, dt, rho * vel[c], bv, rtol=rtol, x0=vel[c])
    return out


def divtau(g, vel, rho, eta, bvals=None):
    """compute_divtau, scalar form: div(eta grad u_c) / rho per component."""
    e = eta if np.ndim(eta) else np.full(g.cell_shape(), float(eta))
    bf = eta_faces(e, g.periodic)
    return np.stack([apply(g, bf, vel[c], None if bvals is None else bvals[c]) / rho
                     for c in range(vel.shape[0])])
