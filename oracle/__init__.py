"""ctypes wrapper around the FP64 CPU oracle (oracle/_build/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke entry point and the
CPU-baseline leg of bench.py, where it is the checker.  The product library
never loads it.

Arrays are numpy float64, x fastest (numpy shape (nz, ny, nx), C order).
Cell vectors with one ghost layer have shape (3, nz+2, ny+2, nx+2).
"""
import ctypes as C
import os
import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liborc.so")

BC_PERIODIC, BC_DIRICHLET, BC_NEUMANN, BC_INFLOW, BC_ROBIN = 0, 1, 2, 3, 4


class Geom(C.Structure):
    _fields_ = [("n", C.c_int * 3), ("periodic", C.c_int * 3), ("bclo", C.c_int * 3),
                ("bchi", C.c_int * 3), ("dx", C.c_double * 3)]

    @staticmethod
    def make(n, periodic=(1, 1, 1), bclo=(0, 0, 0), bchi=(0, 0, 0), prob_lo=(0, 0, 0),
             prob_hi=(1, 1, 1)):
        g = Geom()
        for d in range(3):
            g.n[d] = int(n[d])
            g.periodic[d] = int(periodic[d])
            g.bclo[d] = int(bclo[d]) if not periodic[d] else BC_PERIODIC
            g.bchi[d] = int(bchi[d]) if not periodic[d] else BC_PERIODIC
            g.dx[d] = (prob_hi[d] - prob_lo[d]) / n[d]
        return g

    def nodes(self):
        return tuple(self.n[d] if self.periodic[d] else self.n[d] + 1 for d in range(3))

    def node_shape(self):
        nx, ny, nz = self.nodes()
        return (nz, ny, nx)

    def cell_shape(self):
        return (self.n[2], self.n[1], self.n[0])

    def ghost_vec_shape(self):
        return (3, self.n[2] + 2, self.n[1] + 2, self.n[0] + 2)

    def coarsen(self):
        g = Geom()
        for d in range(3):
            g.n[d] = self.n[d] // 2
            g.periodic[d] = self.periodic[d]
            g.bclo[d] = self.bclo[d]
            g.bchi[d] = self.bchi[d]
            g.dx[d] = 2 * self.dx[d]
        return g


class Opts(C.Structure):
    _fields_ = [("max_iter", C.c_int), ("bottom_max_iter", C.c_int),
                ("bottom_rtol", C.c_double), ("nu1", C.c_int), ("nu2", C.c_int),
                ("omega", C.c_double), ("max_coarsening_level", C.c_int),
                ("verbose", C.c_int)]


class Info(C.Structure):
    _fields_ = [("iters", C.c_int), ("converged", C.c_int), ("bnorm", C.c_double),
                ("resnorm0", C.c_double), ("resnorm", C.c_double), ("nlevels", C.c_int),
                ("bottom_iters", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        _lib = C.CDLL(_LIB_PATH)
    return _lib


_dp = C.POINTER(C.c_double)


def _p(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(_dp)


def default_opts(**kw):
    o = Opts()
    lib().orc_default_opts(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def nodal_apply(g, sigma, phi):
    out = np.zeros(g.node_shape())
    lib().orc_nodal_apply(C.byref(g), _p(sigma), _p(phi), _p(out))
    return out


def nodal_diag(g, sigma):
    out = np.zeros(g.node_shape())
    lib().orc_nodal_diag(C.byref(g), _p(sigma), _p(out))
    return out


def nodal_residual(g, sigma, b, x):
    out = np.zeros(g.node_shape())
    lib().orc_nodal_residual(C.byref(g), _p(sigma), _p(b), _p(x), _p(out))
    return out


def nodal_jacobi(g, sigma, b, x, omega):
    out = np.zeros(g.node_shape())
    lib().orc_nodal_jacobi(C.byref(g), _p(sigma), _p(b), _p(x), _p(out), C.c_double(omega))
    return out


def nodal_restrict(gf, rf):
    out = np.zeros(gf.coarsen().node_shape())
    lib().orc_nodal_restrict(C.byref(gf), _p(rf), _p(out))
    return out


def nodal_prolong_add(gf, xc, xf):
    xf = np.ascontiguousarray(xf, dtype=np.float64).copy()
    lib().orc_nodal_prolong_add(C.byref(gf), _p(xc), _p(xf))
    return xf


def sigma_coarsen(gf, sf):
    out = np.zeros(gf.coarsen().cell_shape())
    lib().orc_sigma_coarsen(C.byref(gf), _p(sf), _p(out))
    return out


def nodal_nlevels(g, max_coarsening_level=100):
    return lib().orc_nodal_nlevels(C.byref(g), C.c_int(max_coarsening_level))


def fill_vel_ghosts(g, ug, inflow_vel=None):
    iv = None
    if inflow_vel is not None:
        iv = np.ascontiguousarray(inflow_vel, dtype=np.float64).reshape(6, 3)
    lib().orc_fill_vel_ghosts(C.byref(g), _p(ug), _p(iv) if iv is not None else None)
    return ug


def nodal_divergence(g, ug):
    out = np.zeros(g.node_shape())
    lib().orc_nodal_divergence(C.byref(g), _p(ug), _p(out))
    return out


def nodal_gradient(g, phi):
    out = np.zeros((3,) + g.cell_shape())
    lib().orc_nodal_gradient(C.byref(g), _p(phi), _p(out))
    return out


def nodal_bicgstab(g, sigma, b, x0=None, rtol=1e-4, maxiter=100):
    x = np.zeros(g.node_shape()) if x0 is None else np.array(x0, dtype=np.float64)
    its = C.c_int(0)
    ret = lib().orc_nodal_bicgstab(C.byref(g), _p(sigma), _p(b), _p(x), C.c_double(rtol),
                                   C.c_int(maxiter), C.byref(its))
    return x, its.value, ret


def nodal_solve(g, sigma, b, x0=None, rtol=1e-10, atol=0.0, opts=None):
    x = np.zeros(g.node_shape()) if x0 is None else np.array(x0, dtype=np.float64)
    info = Info()
    o = opts if opts is not None else default_opts()
    ret = lib().orc_nodal_solve(C.byref(g), _p(sigma), _p(np.ascontiguousarray(b)), _p(x),
                                C.c_double(rtol), C.c_double(atol), C.byref(o), C.byref(info))
    return x, info, ret


def nodal_project(g, sigma, vel_g, rtol=1e-10, atol=0.0, opts=None):
    """vel_g is updated in place (valid cells); returns (phi, gphi, info)."""
    phi = np.zeros(g.node_shape())
    gphi = np.zeros((3,) + g.cell_shape())
    info = Info()
    o = opts if opts is not None else default_opts()
    ret = lib().orc_nodal_project(C.byref(g), _p(sigma), _p(vel_g), _p(phi), _p(gphi),
                                  C.c_double(rtol), C.c_double(atol), C.byref(o),
                                  C.byref(info))
    return phi, gphi, info, ret


def _setup():
    L = lib()
    for name in ["orc_norm_inf", "orc_dot", "orc_mean"]:
        getattr(L, name).restype = C.c_double
    L.orc_nodal_nlevels.restype = C.c_int
    L.orc_unet_nparam.restype = C.c_long


def load():
    lib()
    _setup()
    return _lib


# ---- MAC (cell-centred) ---------------------------------------------------
def face_shape(g, d):
    n = [g.n[0], g.n[1], g.n[2]]
    n[d] += 1
    return (n[2], n[1], n[0])


def mac_apply(g, beta, phi):
    out = np.zeros(g.cell_shape())
    lib().orc_mac_apply(C.byref(g), _p(beta[0]), _p(beta[1]), _p(beta[2]), _p(phi), _p(out))
    return out


def mac_residual(g, beta, b, x):
    out = np.zeros(g.cell_shape())
    lib().orc_mac_residual(C.byref(g), _p(beta[0]), _p(beta[1]), _p(beta[2]), _p(b), _p(x),
                           _p(out))
    return out


def mac_gsrb(g, beta, b, x):
    x = np.array(x, dtype=np.float64)
    lib().orc_mac_gsrb(C.byref(g), _p(beta[0]), _p(beta[1]), _p(beta[2]), _p(b), _p(x))
    return x


def mac_restrict(gf, rf):
    out = np.zeros(gf.coarsen().cell_shape())
    lib().orc_mac_restrict(C.byref(gf), _p(rf), _p(out))
    return out


def mac_prolong_add(gf, xc, xf):
    xf = np.array(xf, dtype=np.float64)
    lib().orc_mac_prolong_add(C.byref(gf), _p(xc), _p(xf))
    return xf


def mac_beta_coarsen(gf, bf, d):
    out = np.zeros(face_shape(gf.coarsen(), d))
    lib().orc_mac_beta_coarsen(C.byref(gf), _p(bf), C.c_int(d), _p(out))
    return out


def mac_beta_from_density(g, rho, dt):
    b = [np.zeros(face_shape(g, d)) for d in range(3)]
    lib().orc_mac_beta_from_density(C.byref(g), _p(rho), C.c_double(dt), _p(b[0]), _p(b[1]),
                                    _p(b[2]))
    return b


def mac_divergence(g, u):
    out = np.zeros(g.cell_shape())
    lib().orc_mac_divergence(C.byref(g), _p(u[0]), _p(u[1]), _p(u[2]), _p(out))
    return out


def mac_fluxes(g, beta, phi):
    f = [np.zeros(face_shape(g, d)) for d in range(3)]
    lib().orc_mac_fluxes(C.byref(g), _p(beta[0]), _p(beta[1]), _p(beta[2]), _p(phi), _p(f[0]),
                         _p(f[1]), _p(f[2]))
    return f


def mac_bicgstab(g, beta, b, rtol=1e-4, maxiter=200):
    x = np.zeros(g.cell_shape())
    its = C.c_int(0)
    ret = lib().orc_mac_bicgstab(C.byref(g), _p(beta[0]), _p(beta[1]), _p(beta[2]), _p(b), _p(x),
                                 C.c_double(rtol), C.c_int(maxiter), C.byref(its))
    return x, its.value, ret


def mac_solve(g, beta, b, x0=None, rtol=1e-10, atol=0.0, opts=None):
    x = np.zeros(g.cell_shape()) if x0 is None else np.array(x0, dtype=np.float64)
    info = Info()
    o = opts if opts is not None else default_opts(max_iter=200, bottom_max_iter=200)
    ret = lib().orc_mac_solve(C.byref(g), _p(beta[0]), _p(beta[1]), _p(beta[2]),
                              _p(np.ascontiguousarray(b)), _p(x), C.c_double(rtol),
                              C.c_double(atol), C.byref(o), C.byref(info))
    return x, info, ret


def mac_project(g, beta, u, rtol=1e-10, atol=0.0, opts=None, phi0=None):
    """u: list of 3 face arrays, updated in place; returns (phi, info, ret)."""
    phi = np.zeros(g.cell_shape()) if phi0 is None else np.array(phi0, dtype=np.float64)
    info = Info()
    o = opts if opts is not None else default_opts(max_iter=200, bottom_max_iter=200)
    ret = lib().orc_mac_project(C.byref(g), _p(beta[0]), _p(beta[1]), _p(beta[2]), _p(u[0]),
                                _p(u[1]), _p(u[2]), _p(phi), C.c_double(rtol), C.c_double(atol),
                                C.byref(o), C.byref(info))
    return phi, info, ret


# ---- learned predictor: 3x3x3 convolution and U-net ------------------------
# Tensors are channels last: shape (batch, nz, ny, nx, c), float64.
def _per(periodic):
    return (C.c_int * 3)(*[1 if p else 0 for p in periodic])


def conv3d_forward(x, w, bias, cout, periodic=(1, 1, 1), relu=False, bf16=False):
    """x: (batch, nz, ny, nx, cin); w: (cout, 27*cin); bias: (cout,) or None."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    batch, nz, ny, nx, cin = x.shape
    w = np.ascontiguousarray(w, dtype=np.float64)
    assert w.shape == (cout, 27 * cin)
    out = np.zeros((batch, nz, ny, nx, cout))
    bp = _p(np.ascontiguousarray(bias, dtype=np.float64)) if bias is not None else None
    lib().orc_conv3d_fwd(_p(x), _p(w), bp, _p(out), nx, ny, nz, batch, cin, cout,
                         _per(periodic), int(relu), int(bf16))
    return out


def conv3d_backward(x, w, dout, periodic=(1, 1, 1), bf16=False):
    """Returns (din, dw, db) for out = conv(x, w) + b and upstream dout."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    dout = np.ascontiguousarray(dout, dtype=np.float64)
    batch, nz, ny, nx, cin = x.shape
    cout = dout.shape[-1]
    w = np.ascontiguousarray(w, dtype=np.float64)
    din = np.zeros_like(x)
    dw = np.zeros((cout, 27 * cin))
    db = np.zeros(cout)
    lib().orc_conv3d_bwd(_p(x), _p(w), _p(dout), _p(din), _p(dw), _p(db), nx, ny, nz, batch, cin,
                         cout, _per(periodic), int(bf16))
    return din, dw, db


def unet_nparam(base=16, levels=4, cin0=8):
    return int(lib().orc_unet_nparam(base, levels, cin0))


def unet_forward(params, x, base=16, levels=4, periodic=(1, 1, 1), bf16=False):
    """x: (batch, nz, ny, nx, cin0) -> (batch, nz, ny, nx)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    batch, nz, ny, nx, cin0 = x.shape
    p = np.ascontiguousarray(params, dtype=np.float64)
    assert p.size == unet_nparam(base, levels, cin0)
    out = np.zeros((batch, nz, ny, nx))
    lib().orc_unet_forward(_p(p), base, levels, cin0, _p(x), _p(out), nx, ny, nz, batch,
                           _per(periodic), int(bf16))
    return out


def unet_grad(params, x, dout, base=16, levels=4, periodic=(1, 1, 1), bf16=False):
    """Gradient of sum(dout * unet(x)) with respect to the flat parameters."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    batch, nz, ny, nx, cin0 = x.shape
    p = np.ascontiguousarray(params, dtype=np.float64)
    g = np.zeros_like(p)
    lib().orc_unet_grad(_p(p), base, levels, cin0, _p(x),
                        _p(np.ascontiguousarray(dout, dtype=np.float64)), _p(g), nx, ny, nz,
                        batch, _per(periodic), int(bf16))
    return g


# ---- Godunov PPM advection (periodic, canonical arrays) -----------------------
# cells (nz, ny, nx); d-faces have n entries along d (face i = low face of cell i)
def _opt(a):
    return _p(np.ascontiguousarray(a, dtype=np.float64)) if a is not None else None


def ppm_predict(g, s, d, ucell=None, uface=None, dt=0.0):
    Im = np.zeros(g.cell_shape())
    Ip = np.zeros(g.cell_shape())
    lib().orc_ppm_predict(C.byref(g), _p(np.ascontiguousarray(s)), int(d), _opt(ucell),
                          _opt(uface), C.c_double(dt), _p(Im), _p(Ip))
    return Im, Ip


def godunov_extrap_vel(g, vel, dt, force=None):
    """vel: (3, nz, ny, nx) -> [umac, vmac, wmac] at t^{n+1/2}."""
    vel = np.ascontiguousarray(vel, dtype=np.float64)
    f = [None] * 3 if force is None else [np.ascontiguousarray(force[n]) for n in range(3)]
    out = [np.zeros(g.cell_shape()) for _ in range(3)]
    lib().orc_godunov_extrap_vel(C.byref(g), _p(vel[0]), _p(vel[1]), _p(vel[2]), _opt(f[0]),
                                 _opt(f[1]), _opt(f[2]), C.c_double(dt), _p(out[0]), _p(out[1]),
                                 _p(out[2]))
    return out


def godunov_edge_states(g, vel, umac, dt, force=None):
    """Edge states of the 3 components on all faces: shape (3 faces, 3 comps, nz, ny, nx)."""
    vel = np.ascontiguousarray(vel, dtype=np.float64)
    f = [None] * 3 if force is None else [np.ascontiguousarray(force[n]) for n in range(3)]
    st = np.zeros((3, 3) + g.cell_shape())
    um = [np.ascontiguousarray(a, dtype=np.float64) for a in umac]
    lib().orc_godunov_edge_states(C.byref(g), _p(vel[0]), _p(vel[1]), _p(vel[2]), _opt(f[0]),
                                  _opt(f[1]), _opt(f[2]), _p(um[0]), _p(um[1]), _p(um[2]),
                                  C.c_double(dt), _p(st))
    return st


def convective_term(g, vel, umac, states):
    vel = np.ascontiguousarray(vel, dtype=np.float64)
    um = [np.ascontiguousarray(a, dtype=np.float64) for a in umac]
    conv = np.zeros((3,) + g.cell_shape())
    lib().orc_convective_term(C.byref(g), _p(vel[0]), _p(vel[1]), _p(vel[2]), _p(um[0]),
                              _p(um[1]), _p(um[2]), _p(np.ascontiguousarray(states)), _p(conv))
    return conv


def godunov_scalar_states(g, s, umac, dt, iconserv=None, force=None):
    """Edge states of ncomp scalars (s: (ncomp, nz, ny, nx)) with the MAC
    velocities: shape (3 faces, ncomp, nz, ny, nx)."""
    s = np.ascontiguousarray(s, dtype=np.float64)
    nc = s.shape[0]
    ic = (C.c_int * nc)(*([0] * nc if iconserv is None else [int(v) for v in iconserv]))
    st = np.zeros((3, nc) + g.cell_shape())
    um = [np.ascontiguousarray(a, dtype=np.float64) for a in umac]
    lib().orc_godunov_scalar_states(C.byref(g), _p(s), nc, ic, _opt(force), _p(um[0]), _p(um[1]),
                                    _p(um[2]), C.c_double(dt), _p(st))
    return st


def scalar_advection(g, s, umac, states, iconserv=None):
    """conv = -div(u_mac s_face) (+ s div(u_mac) for convective components)."""
    s = np.ascontiguousarray(s, dtype=np.float64)
    nc = s.shape[0]
    ic = (C.c_int * nc)(*([0] * nc if iconserv is None else [int(v) for v in iconserv]))
    um = [np.ascontiguousarray(a, dtype=np.float64) for a in umac]
    conv = np.zeros((nc,) + g.cell_shape())
    lib().orc_scalar_advection(C.byref(g), _p(s), nc, ic, _p(um[0]), _p(um[1]), _p(um[2]),
                               _p(np.ascontiguousarray(states)), _p(conv))
    return conv


# General form with domain boundaries: cell states carry 2 ghost layers
# (shape (ncomp, nz+4, ny+4, nx+4), filled by the caller), bc is an int array
# (ncomp, 6) of amrex::BCType values (lo x, y, z, hi x, y, z), faces have
# n_d + 1 entries along d (face_shape).
BCT_RODD, BCT_INT, BCT_REVEN, BCT_FOE, BCT_EXT, BCT_HOE = -1, 0, 1, 2, 3, 4


def _bc_arr(bc, ncomp):
    a = np.ascontiguousarray(np.asarray(bc, dtype=np.int32).reshape(ncomp, 6))
    return a, a.ctypes.data_as(C.POINTER(C.c_int))


def gd_extrap_vel(g, vel_g, bc, dt, force=None):
    vel_g = np.ascontiguousarray(vel_g, dtype=np.float64)
    b, bp = _bc_arr(bc, 3)
    out = [np.zeros(face_shape(g, d)) for d in range(3)]
    lib().orc_gd_extrap_vel(C.byref(g), _p(vel_g), bp, _opt(force), C.c_double(dt), _p(out[0]),
                            _p(out[1]), _p(out[2]))
    return out


def gd_edge_states(g, s_g, bc, umac, dt, iconserv=None, force=None):
    """-> [states on x-faces (ncomp, ...), y-faces, z-faces]"""
    s_g = np.ascontiguousarray(s_g, dtype=np.float64)
    nc = s_g.shape[0]
    b, bp = _bc_arr(bc, nc)
    ic = (C.c_int * nc)(*([0] * nc if iconserv is None else [int(v) for v in iconserv]))
    um = [np.ascontiguousarray(a, dtype=np.float64) for a in umac]
    out = [np.zeros((nc,) + face_shape(g, d)) for d in range(3)]
    lib().orc_gd_edge_states(C.byref(g), nc, _p(s_g), bp, ic, _opt(force), _p(um[0]), _p(um[1]),
                             _p(um[2]), C.c_double(dt), _p(out[0]), _p(out[1]), _p(out[2]))
    return out


def gd_conv(g, s, umac, states, iconserv=None):
    s = np.ascontiguousarray(s, dtype=np.float64)
    nc = s.shape[0]
    ic = (C.c_int * nc)(*([0] * nc if iconserv is None else [int(v) for v in iconserv]))
    um = [np.ascontiguousarray(a, dtype=np.float64) for a in umac]
    st = [np.ascontiguousarray(a, dtype=np.float64) for a in states]
    conv = np.zeros((nc,) + g.cell_shape())
    lib().orc_gd_conv(C.byref(g), nc, _p(s), ic, _p(um[0]), _p(um[1]), _p(um[2]), _p(st[0]),
                      _p(st[1]), _p(st[2]), _p(conv))
    return conv


# ---- embedded boundary, nodal ------------------------------------------------
_ip = C.POINTER(C.c_int)


def eb_cylinder(g, cx, cy, r, prob_lo=(0.0, 0.0, 0.0)):
    """-> (ctype (nz,ny,nx) int32: 0 covered / 1 regular / 2 cut, moments (nz,ny,nx,27))"""
    ct = np.zeros(g.cell_shape(), dtype=np.int32)
    mom = np.zeros(g.cell_shape() + (27,))
    lib().orc_eb_cylinder(C.byref(g), (C.c_double * 3)(*prob_lo), C.c_double(cx),
                          C.c_double(cy), C.c_double(r), ct.ctypes.data_as(_ip), _p(mom))
    return ct, mom


def eb_element(mom27, dx):
    E = np.zeros(64)
    lib().orc_eb_element(_p(np.ascontiguousarray(mom27, dtype=np.float64)),
                         (C.c_double * 3)(*dx), _p(E))
    return E.reshape(8, 8)


def eb_node_mask(g, ctype):
    m = np.zeros(g.node_shape(), dtype=np.int32)
    lib().orc_eb_node_mask(C.byref(g), np.ascontiguousarray(ctype, np.int32).ctypes.data_as(_ip),
                           m.ctypes.data_as(_ip))
    return m


def eb_nodal_apply(g, ctype, mom, sigma, phi):
    out = np.zeros(g.node_shape())
    lib().orc_eb_nodal_apply(C.byref(g), np.ascontiguousarray(ctype, np.int32).ctypes.data_as(_ip),
                             _p(np.ascontiguousarray(mom)), _p(np.ascontiguousarray(sigma)),
                             _p(np.ascontiguousarray(phi)), _p(out))
    return out


def eb_nodal_divergence(g, ctype, mom, ug):
    out = np.zeros(g.node_shape())
    lib().orc_eb_nodal_divergence(C.byref(g),
                                  np.ascontiguousarray(ctype, np.int32).ctypes.data_as(_ip),
                                  _p(np.ascontiguousarray(mom)), _p(np.ascontiguousarray(ug)),
                                  _p(out))
    return out


def eb_cylinder_boundary(g, ctype, mom, cx, cy, r, prob_lo=(0.0, 0.0, 0.0)):
    """Body-surface weights B (nz,ny,nx,8,3): (1/V) int N_p n_d dA, n out of the fluid."""
    B = np.zeros(g.cell_shape() + (8, 3))
    lib().orc_eb_cylinder_boundary(C.byref(g), (C.c_double * 3)(*prob_lo), C.c_double(cx),
                                   C.c_double(cy), C.c_double(r),
                                   np.ascontiguousarray(ctype, np.int32).ctypes.data_as(_ip),
                                   _p(np.ascontiguousarray(mom)), _p(B))
    return B


def eb_inflow_rhs(g, ctype, B, ueb, rhs):
    """rhs + the body-surface term of the EB inflow velocity ueb (3, nz, ny, nx)."""
    out = np.array(rhs, dtype=np.float64)
    lib().orc_eb_inflow_rhs(C.byref(g), np.ascontiguousarray(ctype, np.int32).ctypes.data_as(_ip),
                            _p(np.ascontiguousarray(B)), _p(np.ascontiguousarray(ueb)), _p(out))
    return out


def eb_nodal_gradient(g, ctype, mom, phi):
    out = np.zeros((3,) + g.cell_shape())
    lib().orc_eb_nodal_gradient(C.byref(g),
                                np.ascontiguousarray(ctype, np.int32).ctypes.data_as(_ip),
                                _p(np.ascontiguousarray(mom)), _p(np.ascontiguousarray(phi)),
                                _p(out))
    return out


def eb_nodal_solve(g, ctype, mom, sigma, b, x0=None, rtol=1e-12, maxiter=20000):
    x = np.zeros(g.node_shape()) if x0 is None else np.array(x0, dtype=np.float64)
    its = C.c_int(0)
    ret = lib().orc_eb_nodal_solve(C.byref(g),
                                   np.ascontiguousarray(ctype, np.int32).ctypes.data_as(_ip),
                                   _p(np.ascontiguousarray(mom)), _p(np.ascontiguousarray(sigma)),
                                   _p(np.ascontiguousarray(b)), _p(x), C.c_double(rtol),
                                   C.c_int(maxiter), C.byref(its))
    return x, its.value, ret


def dirichlet_node_mask(g):
    """1 on nodes the domain boundary conditions leave free, 0 on Dirichlet nodes"""
    m = np.ones(g.node_shape(), bool)
    for d, ax in ((0, 2), (1, 1), (2, 0)):
        if g.periodic[d]:
            continue
        sl = [slice(None)] * 3
        if g.bclo[d] == BC_DIRICHLET:
            sl[ax] = 0
            m[tuple(sl)] = False
        if g.bchi[d] == BC_DIRICHLET:
            sl[ax] = -1
            m[tuple(sl)] = False
    return m


def nodal_masked_solve(g, sigma, b, mask, rtol=1e-13, maxiter=20000):
    """Nodal solve with an overset mask (MLNodeLaplacian::setOversetMask, used by
    incflo for mixed boundaries, src/projection/incflo_apply_nodal_projection.cpp:203-211):
    nodes with mask 0 are fixed at phi = 0 and carry no equation.  Jacobi-
    preconditioned conjugate gradients on the remaining nodes (the operator
    restricted to them is symmetric negative definite once a node is fixed).
    Returns (phi, iterations, 0 on success)."""
    m = np.asarray(mask, bool) & dirichlet_node_mask(g)
    mf = m.astype(np.float64)

    def A(x):
        return -mf * nodal_apply(g, sigma, mf * x)

    d = -mf * nodal_diag(g, sigma)
    d = np.where(m & (d > 0), d, 1.0)
    r = -mf * np.asarray(b, np.float64)
    x = np.zeros(g.node_shape())
    bn = np.abs(r).max()
    if bn == 0.0:
        return x, 0, 0
    z = r / d
    p = z.copy()
    rz = np.sum(r * z)
    for it in range(1, maxiter + 1):
        q = A(p)
        alpha = rz / np.sum(p * q)
        x += alpha * p
        r -= alpha * q
        if np.abs(r).max() <= rtol * bn:
            return x, it, 0
        z = r / d
        rz2 = np.sum(r * z)
        p = z + (rz2 / rz) * p
        rz = rz2
    return x, maxiter, 1
