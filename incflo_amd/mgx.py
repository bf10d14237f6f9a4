"""ctypes binding of libmgx.so (include/mgx.h) used by the tests and bench.

This is plumbing over the C ABI: device memory comes from the library's own
mgx_dev_malloc / memcpy helpers, so no second HIP runtime is loaded.  The
library is required -- there is no CPU fallback: loading raises when
libmgx.so is missing and every call raises MgxError on a non-zero status.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmgx.so")

MGX_OK, MGX_ERR_ARG, MGX_ERR_HIP, MGX_ERR_NOT_CONVERGED, MGX_ERR_STATE, MGX_ERR_UNSUPPORTED = range(6)
BC_PERIODIC, BC_DIRICHLET, BC_NEUMANN, BC_INFLOW, BC_ROBIN, BC_INOUT = range(6)
MODE_FP64, MODE_MIXED, MODE_LEARNED = 0, 1, 2
(BOTTOM_DEFAULT, BOTTOM_BICGSTAB, BOTTOM_CG, BOTTOM_BICGCG, BOTTOM_CGBICG, BOTTOM_SMOOTHER,
 BOTTOM_HYPRE) = range(7)
# the bottom_solver strings of incflo's inputs (nodal_proj.bottom_solver, ...)
SMOOTHER_DEFAULT, SMOOTHER_JACOBI, SMOOTHER_CHEBYSHEV = 0, 1, 2
BOTTOM_SOLVERS = {"bicgstab": BOTTOM_BICGSTAB, "cg": BOTTOM_CG, "bicgcg": BOTTOM_BICGCG,
                  "cgbicg": BOTTOM_CGBICG, "smoother": BOTTOM_SMOOTHER, "hypre": BOTTOM_HYPRE}


class MgxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mgx status {code}: {msg}")
        self.code = code


class LevelDesc(C.Structure):
    _fields_ = [("n_cell", C.c_int * 3), ("prob_lo", C.c_double * 3),
                ("prob_hi", C.c_double * 3), ("is_periodic", C.c_int * 3),
                ("nboxes", C.c_int), ("boxes", C.POINTER(C.c_int))]


class MLMGOpts(C.Structure):
    _fields_ = [("verbose", C.c_int), ("bottom_verbose", C.c_int), ("maxiter", C.c_int),
                ("bottom_maxiter", C.c_int), ("bottom_rtol", C.c_double),
                ("bottom_atol", C.c_double), ("max_coarsening_level", C.c_int),
                ("nu1", C.c_int), ("nu2", C.c_int), ("omega", C.c_double),
                ("stall_iters", C.c_int), ("cycles_per_update", C.c_int),
                ("bottom_solver", C.c_int), ("smoother", C.c_int), ("learned_steps", C.c_int)]


class SolveInfo(C.Structure):
    _fields_ = [("iters", C.c_int), ("converged", C.c_int), ("nlevels", C.c_int),
                ("fallback_used", C.c_int), ("bnorm", C.c_double), ("resnorm0", C.c_double),
                ("resnorm", C.c_double), ("rel_resid", C.c_double), ("vcycles", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m incflo_amd.build`")
        L = C.CDLL(LIB_PATH)
        L.mgx_version.restype = C.c_char_p
        L.mgx_ctx_last_error.restype = C.c_char_p
        L.mgx_ctx_last_error.argtypes = [C.c_void_p]
        vp = C.c_void_p
        for name, args in {
            "mgx_ctx_create": [C.c_int, vp, C.POINTER(vp)],
            "mgx_ctx_destroy": [vp],
            "mgx_ctx_synchronize": [vp],
            "mgx_dev_malloc": [vp, C.c_ulonglong, C.POINTER(vp)],
            "mgx_dev_free": [vp, vp],
            "mgx_memcpy_h2d": [vp, vp, vp, C.c_ulonglong],
            "mgx_memcpy_d2h": [vp, vp, vp, C.c_ulonglong],
            "mgx_memset_d": [vp, vp, C.c_int, C.c_ulonglong],
            "mgx_memcpy_d2d": [vp, vp, vp, C.c_ulonglong],
            "mgx_stream_copy": [vp, vp, vp, C.c_ulonglong],
            "mgx_nodal_set_timing": [vp, C.c_int],
            "mgx_nodal_set_graphs": [vp, C.c_int],
            "mgx_nodal_set_overset_mask": [vp, vp, C.c_int],
            "mgx_nodal_set_vcycle_bits": [vp, C.c_int],
            "mgx_nodal_get_timing": [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_longlong)],
            "mgx_ctx_record": [vp, C.c_int],
            "mgx_ctx_elapsed_ms": [vp, C.c_int, C.c_int, C.POINTER(C.c_double)],
            "mgx_nodal_create": [vp, C.POINTER(LevelDesc), C.POINTER(vp)],
            "mgx_nodal_destroy": [vp],
            "mgx_nodal_set_domain_bc": [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)],
            "mgx_nodal_set_options": [vp, C.POINTER(MLMGOpts)],
            "mgx_nodal_set_sigma": [vp, vp, C.c_int],
            "mgx_nodal_set_const_sigma": [vp, C.c_double],
            "mgx_nodal_project": [vp, vp, C.c_int, C.c_double, C.c_double, C.c_int,
                                  C.POINTER(SolveInfo)],
            "mgx_nodal_get_phi": [vp, vp, C.c_int],
            "mgx_nodal_apply": [vp, vp, C.c_int, vp, C.c_int],
            "mgx_nodal_get_gradphi": [vp, vp, C.c_int],
            "mgx_nodal_solve": [vp, vp, C.c_int, vp, C.c_int, C.c_double, C.c_double, C.c_int,
                                C.POINTER(SolveInfo)],
            "mgx_mac_create": [vp, C.POINTER(LevelDesc), C.POINTER(vp)],
            "mgx_mac_destroy": [vp],
            "mgx_mac_set_domain_bc": [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)],
            "mgx_mac_set_options": [vp, C.POINTER(MLMGOpts)],
            "mgx_mac_set_graphs": [vp, C.c_int],
            "mgx_mac_set_beta": [vp, vp, vp, vp, C.c_int],
            "mgx_mac_set_const_beta": [vp, C.c_double],
            "mgx_mac_set_beta_from_density": [vp, vp, C.c_int, C.c_double],
            "mgx_mac_project": [vp, vp, vp, vp, C.c_int, vp, C.c_int, C.c_double, C.c_double,
                                C.c_int, C.POINTER(SolveInfo)],
            "mgx_mac_get_fluxes": [vp, vp, vp, vp, C.c_int],
            "mgx_mac_solve": [vp, vp, C.c_int, vp, C.c_int, C.c_double, C.c_double, C.c_int,
                              C.POINTER(SolveInfo)],
            "mgx_unet_create": [vp, C.c_int, C.c_int, C.c_ulonglong, C.POINTER(vp)],
            "mgx_unet_destroy": [vp],
            "mgx_unet_num_params": [vp, C.POINTER(C.c_longlong)],
            "mgx_unet_get_params": [vp, vp, C.c_longlong],
            "mgx_unet_set_params": [vp, vp, C.c_longlong],
            "mgx_unet_forward": [vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)],
            "mgx_unet_gradients": [vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int,
                                   C.POINTER(C.c_int), vp, C.c_longlong],
            "mgx_unet_train_step": [vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.POINTER(C.c_int), C.c_float, C.POINTER(C.c_float)],
            "mgx_conv3d_forward": [vp, vp, vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                   C.c_int, C.POINTER(C.c_int), C.c_int],
            "mgx_conv3d_backward": [vp, vp, vp, vp, vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.c_int, C.c_int, C.POINTER(C.c_int)],
            "mgx_adv_create": [vp, C.POINTER(LevelDesc), C.POINTER(vp)],
            "mgx_adv_destroy": [vp],
            "mgx_adv_extrap_vel_to_faces": [vp, vp, C.c_int, vp, C.c_int, C.c_double, vp, vp, vp,
                                            C.c_int],
            "mgx_adv_scalar_term": [vp, vp, C.c_int, C.c_int, vp, vp, vp, C.c_int, vp, vp, vp,
                                    C.c_int, C.c_double, vp, C.c_int, vp, vp, vp, vp, vp, vp],
            "mgx_adv_set_velocity_bcrec": [vp, vp],
            "mgx_adv_convective_term": [vp, vp, C.c_int, vp, C.c_int, vp, vp, vp, C.c_int,
                                        C.c_double, vp, C.c_int, vp, vp, vp, vp, vp, vp],
            "mgx_nodal_set_eb_cylinder": [vp, C.c_double, C.c_double, C.c_double],
            "mgx_nodal_set_eb_moments": [vp, vp, vp],
            "mgx_eb_cylinder_geometry": [C.POINTER(LevelDesc), C.c_double, C.c_double, C.c_double,
                                         vp, vp],
            "mgx_mac_set_eb_cylinder": [vp, C.c_double, C.c_double, C.c_double],
            "mgx_mac_set_level_bc": [vp, vp, vp, vp, C.c_int],
            "mgx_amr_average_down_faces": [vp, C.POINTER(LevelDesc), C.POINTER(LevelDesc), vp,
                                           C.c_int, vp, C.c_int, C.c_int],
            "mgx_amr_average_down": [vp, C.POINTER(LevelDesc), C.POINTER(LevelDesc), vp, C.c_int,
                                     vp, C.c_int, C.c_int],
            "mgx_mac_set_eb_inflow_velocity": [vp, vp, C.c_int],
            "mgx_nodal_set_eb_inflow_velocity": [vp, vp, C.c_int],
            "mgx_eb_cylinder_surface_weights": [C.POINTER(LevelDesc), C.c_double, C.c_double,
                                                C.c_double, vp],
            "mgx_eb_cylinder_apertures": [C.POINTER(LevelDesc), C.c_double, C.c_double, C.c_double,
                                          vp, vp, vp],
            "mgx_nodal_set_unet": [vp, vp, C.c_int],
            "mgx_nodal_training_pair": [vp, C.c_int, vp, vp],
            "mgx_amr_nodal_create": [vp, C.POINTER(LevelDesc), C.c_int, C.POINTER(vp)],
            "mgx_amr_nodal_destroy": [vp],
            "mgx_amr_nodal_set_domain_bc": [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)],
            "mgx_amr_nodal_set_options": [vp, C.POINTER(MLMGOpts)],
            "mgx_amr_nodal_set_sigma": [vp, C.POINTER(vp), C.c_int],
            "mgx_amr_nodal_project": [vp, C.POINTER(vp), C.c_int, C.POINTER(vp), C.c_int,
                                      C.POINTER(vp), C.c_int, C.c_double, C.c_double, C.c_int,
                                      C.POINTER(SolveInfo)],
            "mgx_amr_mac_create": [vp, C.POINTER(LevelDesc), C.c_int, C.POINTER(vp)],
            "mgx_amr_mac_destroy": [vp],
            "mgx_amr_mac_set_domain_bc": [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)],
            "mgx_amr_mac_set_options": [vp, C.POINTER(MLMGOpts)],
            "mgx_amr_mac_set_beta": [vp, C.POINTER(vp), C.c_int],
            "mgx_amr_mac_project": [vp, C.POINTER(vp), C.c_int, C.POINTER(vp), C.c_int, C.c_double,
                                    C.c_double, C.c_int, C.POINTER(SolveInfo)],
            "mgx_apply_cc_projection": [vp, vp, C.c_int, C.c_double, vp, C.c_int, vp, C.c_int, vp,
                                        C.c_int, vp, C.c_int, C.c_double, C.c_int, vp, vp, vp,
                                        vp, C.c_int, C.c_double, C.c_double, C.c_int,
                                        C.POINTER(SolveInfo)],
            "mgx_nodal_autotune": [vp, C.c_int, C.c_char_p, C.c_int],
            "mgx_tuning_get": [C.c_char_p, C.c_int],
            "mgx_tuning_set": [C.c_char_p],
            "mgx_write_multifab": [vp, C.c_char_p, C.POINTER(LevelDesc), vp, C.c_int, C.c_int,
                                   C.c_int],
            "mgx_read_multifab": [vp, C.c_char_p, C.POINTER(LevelDesc), vp, C.c_int, C.c_int,
                                  C.c_int],
            "mgx_write_checkpoint_header": [C.c_char_p, C.c_int, C.POINTER(LevelDesc), C.c_int,
                                            C.c_double, C.c_double, C.c_double, C.c_double],
            "mgx_read_checkpoint_header": [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                           C.POINTER(C.c_double), C.POINTER(C.c_double),
                                           C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_int,
                                           C.POINTER(C.c_int), C.c_int],
            "mgx_write_plotfile": [vp, C.c_char_p, C.c_int, C.POINTER(LevelDesc), C.POINTER(vp),
                                   C.c_int, C.c_int, C.POINTER(C.c_char_p), C.c_double,
                                   C.POINTER(C.c_int), C.POINTER(C.c_int)],
            "mgx_tag_cells": [vp, C.POINTER(LevelDesc), vp, C.c_int, C.c_double, C.c_double,
                              vp, vp, vp],
            "mgx_cluster_tags": [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int, C.c_int,
                                 C.c_int, C.c_double, C.POINTER(C.c_int), C.c_int,
                                 C.POINTER(C.c_int)],
            "mgx_fill_fine_level": [vp, C.POINTER(LevelDesc), vp, C.c_int, C.c_int,
                                    C.POINTER(C.c_int), C.POINTER(LevelDesc), vp, C.c_int,
                                    C.POINTER(LevelDesc), vp, C.c_int],
            "mgx_step_create": [vp, C.POINTER(LevelDesc), C.POINTER(vp)],
            "mgx_step_destroy": [vp],
            "mgx_fill_boundary": [vp, vp, C.c_int, C.c_int],
            "mgx_fill_boundary_noalloc": [vp, vp, C.c_int, C.c_int],
            "mgx_compute_vel_forces": [vp, vp, C.c_int, vp, C.c_int, vp, C.c_int, vp, vp, C.c_int,
                                       vp, C.c_int],
            "mgx_update_velocity": [vp, vp, C.c_int, vp, C.c_int, vp, C.c_int, vp, C.c_int, vp,
                                    C.c_int, C.c_double, vp, vp, vp, C.c_int, C.c_double],
            "mgx_update_density": [vp, vp, C.c_int, vp, C.c_int, C.c_double, vp, C.c_int, vp,
                                   C.c_int],
            "mgx_diff_create": [vp, C.POINTER(LevelDesc), C.POINTER(vp)],
            "mgx_diff_destroy": [vp],
            "mgx_diff_set_domain_bc": [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)],
            "mgx_diff_set_options": [vp, C.POINTER(MLMGOpts)],
            "mgx_diff_set_eta": [vp, vp, C.c_int, C.c_double],
            "mgx_diff_apply": [vp, vp, C.c_int, C.c_int, vp, C.c_int, vp, C.c_int],
            "mgx_diff_solve": [vp, vp, C.c_int, C.c_int, vp, C.c_int, C.c_double, C.c_double,
                               C.c_double, C.c_int, C.POINTER(SolveInfo)],
            "mgx_apply_nodal_projection": [vp, vp, C.c_int, C.c_double, vp, C.c_int, vp, C.c_int,
                                           vp, C.c_int, vp, C.c_int, C.c_double, C.c_int, vp,
                                           C.c_double, C.c_double, C.c_int,
                                           C.POINTER(SolveInfo)],
        }.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        L.mgx_default_mlmg_opts.argtypes = [C.POINTER(MLMGOpts)]
        L.mgx_default_mlmg_opts.restype = None
        _lib = L
    return _lib


def tuning_get():
    """the process-wide launch-shape table as text"""
    buf = C.create_string_buffer(1 << 16)
    st = lib().mgx_tuning_get(buf, len(buf))
    if st != MGX_OK:
        raise MgxError(st, "mgx_tuning_get failed")
    return buf.value.decode()


def tuning_set(text):
    st = lib().mgx_tuning_set(text.encode())
    if st != MGX_OK:
        raise MgxError(st, "mgx_tuning_set: bad table")


def default_opts(**kw):
    o = MLMGOpts()
    lib().mgx_default_mlmg_opts(C.byref(o))
    for k, v in kw.items():
        if k == "bottom_solver" and isinstance(v, str):
            v = BOTTOM_SOLVERS[v]
        setattr(o, k, v)
    return o


class Context:
    def __init__(self, device=0, stream=None):
        self.h = C.c_void_p()
        st = lib().mgx_ctx_create(device, stream, C.byref(self.h))
        if st != MGX_OK:
            raise MgxError(st, f"mgx_ctx_create(device={device}) failed (no GPU?)")

    def check(self, st):
        if st != MGX_OK:
            raise MgxError(st, lib().mgx_ctx_last_error(self.h).decode())

    def sync(self):
        self.check(lib().mgx_ctx_synchronize(self.h))

    def record(self, slot):
        self.check(lib().mgx_ctx_record(self.h, slot))

    def elapsed_ms(self, a, b):
        ms = C.c_double()
        self.check(lib().mgx_ctx_elapsed_ms(self.h, a, b, C.byref(ms)))
        return ms.value

    def close(self):
        if self.h:
            lib().mgx_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def array(self, host_or_n, dtype=np.float64):
        return DeviceArray(self, host_or_n, dtype)


class DeviceArray:
    """A flat device buffer with numpy upload/download."""

    def __init__(self, ctx, host_or_n, dtype=np.float64):
        self.ctx = ctx
        if isinstance(host_or_n, np.ndarray):
            host = np.ascontiguousarray(host_or_n, dtype=dtype)
            self.n = host.size
        else:
            host = None
            self.n = int(host_or_n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        self.ptr = C.c_void_p()
        ctx.check(lib().mgx_dev_malloc(ctx.h, max(self.nbytes, 1), C.byref(self.ptr)))
        if host is not None:
            self.upload(host)
        else:
            ctx.check(lib().mgx_memset_d(ctx.h, self.ptr, 0, self.nbytes))

    def upload(self, host):
        host = np.ascontiguousarray(host, dtype=self.dtype)
        assert host.size == self.n
        self.ctx.check(lib().mgx_memcpy_h2d(self.ctx.h, self.ptr, host.ctypes.data, self.nbytes))

    def copy_from(self, other):
        """stream-ordered device-to-device copy"""
        assert other.nbytes == self.nbytes
        self.ctx.check(lib().mgx_memcpy_d2d(self.ctx.h, self.ptr, other.ptr, self.nbytes))

    def download(self):
        out = np.empty(self.n, dtype=self.dtype)
        self.ctx.check(lib().mgx_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr, self.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().mgx_dev_free(self.ctx.h, self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def make_level(n_cell, boxes, periodic=(1, 1, 1), prob_lo=(0, 0, 0), prob_hi=(1, 1, 1)):
    d = LevelDesc()
    for k in range(3):
        d.n_cell[k] = int(n_cell[k])
        d.prob_lo[k] = float(prob_lo[k])
        d.prob_hi[k] = float(prob_hi[k])
        d.is_periodic[k] = int(periodic[k])
    flat = np.ascontiguousarray(np.array([list(lo) + list(hi) for lo, hi in boxes], dtype=np.int32))
    d.nboxes = len(boxes)
    d._keep = flat
    d.boxes = flat.ctypes.data_as(C.POINTER(C.c_int))
    return d


class NodalProjector:
    """Thin mirror of Hydro::NodalProjector over the C ABI (test plumbing)."""

    def __init__(self, ctx, level):
        self.ctx = ctx
        self.level = level
        self.h = C.c_void_p()
        ctx.check(lib().mgx_nodal_create(ctx.h, C.byref(level), C.byref(self.h)))

    def setDomainBC(self, lo, hi):
        a = (C.c_int * 3)(*lo)
        b = (C.c_int * 3)(*hi)
        self.ctx.check(lib().mgx_nodal_set_domain_bc(self.h, a, b))

    def set_options(self, opts):
        self.ctx.check(lib().mgx_nodal_set_options(self.h, C.byref(opts)))

    def set_timing(self, on):
        self.ctx.check(lib().mgx_nodal_set_timing(self.h, int(on)))

    def set_graphs(self, on):
        self.ctx.check(lib().mgx_nodal_set_graphs(self.h, int(on)))

    def setEBInflowVelocity(self, eb_vel, ng=0):
        """Body-surface velocity (device boxed cell field, 3 comps; None removes it)."""
        self.ctx.check(lib().mgx_nodal_set_eb_inflow_velocity(
            self.h, eb_vel.ptr if eb_vel is not None else None, ng))

    def setOversetMask(self, mask_dev, ng=0):
        """mask_dev: device int32 boxed nodal field (or None to remove)"""
        self.ctx.check(lib().mgx_nodal_set_overset_mask(self.h, mask_dev.ptr if mask_dev else None, ng))

    def set_vcycle_bits(self, bits):
        self.ctx.check(lib().mgx_nodal_set_vcycle_bits(self.h, int(bits)))

    def get_timing(self):
        ms, n = C.c_double(), C.c_longlong()
        self.ctx.check(lib().mgx_nodal_get_timing(self.h, 0, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def autotune(self, mode=MODE_FP64):
        """mgx_nodal_autotune: returns the text report."""
        buf = C.create_string_buffer(1 << 16)
        self.ctx.check(lib().mgx_nodal_autotune(self.h, mode, buf, len(buf)))
        return buf.value.decode()

    def set_sigma(self, sigma_dev, ng=0):
        self.ctx.check(lib().mgx_nodal_set_sigma(self.h, sigma_dev.ptr, ng))

    def set_const_sigma(self, v):
        self.ctx.check(lib().mgx_nodal_set_const_sigma(self.h, v))

    def project(self, vel_dev, vel_ng, rtol, atol=0.0, mode=MODE_FP64):
        info = SolveInfo()
        self.ctx.check(lib().mgx_nodal_project(self.h, vel_dev.ptr, vel_ng, rtol, atol, mode,
                                               C.byref(info)))
        return info

    def getPhi(self, phi_dev, ng=0):
        self.ctx.check(lib().mgx_nodal_get_phi(self.h, phi_dev.ptr, ng))

    def getGradPhi(self, g_dev, ng=0):
        self.ctx.check(lib().mgx_nodal_get_gradphi(self.h, g_dev.ptr, ng))

    def apply(self, phi_dev, phi_ng, out_dev, out_ng):
        self.ctx.check(lib().mgx_nodal_apply(self.h, phi_dev.ptr, phi_ng, out_dev.ptr, out_ng))

    def solve(self, phi_dev, phi_ng, rhs_dev, rhs_ng, rtol, atol=0.0, mode=MODE_FP64):
        info = SolveInfo()
        self.ctx.check(lib().mgx_nodal_solve(self.h, phi_dev.ptr, phi_ng, rhs_dev.ptr, rhs_ng,
                                             rtol, atol, mode, C.byref(info)))
        return info

    def apply_nodal_projection(self, vel, vel_ng, gp, gp_ng, p_nd, p_ng, scaling_factor,
                               density=None, rho_ng=0, const_rho=0.0, vel_old=None,
                               vel_old_ng=0, incremental=False, inflow_vel=None, rtol=1e-11,
                               atol=1e-14, mode=MODE_FP64):
        info = SolveInfo()
        iv = None
        if inflow_vel is not None:
            iv = np.ascontiguousarray(inflow_vel, dtype=np.float64).reshape(18)
        st = lib().mgx_apply_nodal_projection(
            self.h, density.ptr if density is not None else None, rho_ng, const_rho, vel.ptr,
            vel_ng, vel_old.ptr if vel_old is not None else None, vel_old_ng, gp.ptr, gp_ng,
            p_nd.ptr, p_ng, scaling_factor, int(incremental),
            iv.ctypes.data if iv is not None else None, rtol, atol, mode, C.byref(info))
        self.ctx.check(st)
        return info

    def set_eb_cylinder(self, cx, cy, r):
        """Embedded cylinder along z (fluid outside)."""
        self.ctx.check(lib().mgx_nodal_set_eb_cylinder(self.h, cx, cy, r))

    def set_eb_moments(self, ctype, moments):
        """ctype: (nz, ny, nx) int 0/1/2; moments: (nz, ny, nx, 27); None removes the EB."""
        if ctype is None:
            self.ctx.check(lib().mgx_nodal_set_eb_moments(self.h, None, None))
            return
        self._eb_keep = (np.ascontiguousarray(ctype, dtype=np.int32),
                         np.ascontiguousarray(moments, dtype=np.float64))
        self.ctx.check(lib().mgx_nodal_set_eb_moments(self.h, self._eb_keep[0].ctypes.data,
                                                      self._eb_keep[1].ctypes.data))

    def set_unet(self, net, level):
        """Attach a UNet for MODE_LEARNED (None detaches)."""
        self.ctx.check(lib().mgx_nodal_set_unet(self.h, net.h if net is not None else None, level))

    def training_pair(self, level, in_dev, target_dev):
        self.ctx.check(lib().mgx_nodal_training_pair(self.h, level, in_dev.ptr, target_dev.ptr))

    def close(self):
        if self.h:
            lib().mgx_nodal_destroy(self.h)
            self.h = C.c_void_p()


class MacProjector:
    """Thin mirror of Hydro::MacProjector over the C ABI (test plumbing)."""

    def __init__(self, ctx, level):
        self.ctx = ctx
        self.level = level
        self.h = C.c_void_p()
        ctx.check(lib().mgx_mac_create(ctx.h, C.byref(level), C.byref(self.h)))

    def setDomainBC(self, lo, hi):
        self.ctx.check(lib().mgx_mac_set_domain_bc(self.h, (C.c_int * 3)(*lo), (C.c_int * 3)(*hi)))

    def set_options(self, opts):
        self.ctx.check(lib().mgx_mac_set_options(self.h, C.byref(opts)))

    def set_graphs(self, on):
        self.ctx.check(lib().mgx_mac_set_graphs(self.h, int(on)))

    def set_beta(self, bx, by, bz, ng=0):
        self.ctx.check(lib().mgx_mac_set_beta(self.h, bx.ptr, by.ptr, bz.ptr, ng))

    def updateBeta(self, beta):
        self.ctx.check(lib().mgx_mac_set_const_beta(self.h, beta))

    def set_beta_from_density(self, rho, rho_ng, dt):
        self.ctx.check(lib().mgx_mac_set_beta_from_density(self.h, rho.ptr, rho_ng, dt))

    def project(self, u, v, w, umac_ng, rtol, atol=0.0, phi=None, phi_ng=0, mode=MODE_FP64):
        info = SolveInfo()
        self.ctx.check(lib().mgx_mac_project(self.h, u.ptr, v.ptr, w.ptr, umac_ng,
                                             phi.ptr if phi is not None else None, phi_ng, rtol,
                                             atol, mode, C.byref(info)))
        return info

    def set_eb_cylinder(self, cx, cy, r):
        """Embedded cylinder along z (fluid outside); before the coefficients."""
        self.ctx.check(lib().mgx_mac_set_eb_cylinder(self.h, cx, cy, r))

    def setEBInflowVelocity(self, eb_vel, ng=0):
        self.ctx.check(lib().mgx_mac_set_eb_inflow_velocity(
            self.h, eb_vel.ptr if eb_vel is not None else None, ng))

    def setLevelBC(self, robin_a, robin_b, robin_f, ng=1):
        """Robin coefficients a phi + b dphi/dn = f, boxed cell fields whose
        first ghost cell outside each Robin face holds the values."""
        self.ctx.check(lib().mgx_mac_set_level_bc(self.h, robin_a.ptr, robin_b.ptr, robin_f.ptr,
                                                  ng))

    def apply_cc_projection(self, vel, vel_ng, gp, gp_ng, p_cc, p_ng, scaling_factor,
                            density=None, rho_ng=0, const_rho=0.0, vel_old=None, vel_old_ng=0,
                            incremental=False, inflow_vel=None, umac=(None, None, None),
                            umac_ng=0, rtol=1e-11, atol=1e-14, mode=MODE_FP64):
        """incflo::ApplyCCProjection on this projector (mgx_apply_cc_projection)."""
        info = SolveInfo()
        iv = None
        if inflow_vel is not None:
            iv = np.ascontiguousarray(inflow_vel, dtype=np.float64).reshape(18)
        p = lambda a: a.ptr if a is not None else None
        self.ctx.check(lib().mgx_apply_cc_projection(
            self.h, p(density), rho_ng, const_rho, vel.ptr, vel_ng, p(vel_old), vel_old_ng,
            gp.ptr, gp_ng, p_cc.ptr, p_ng, scaling_factor, int(incremental),
            iv.ctypes.data if iv is not None else None, p(umac[0]), p(umac[1]), p(umac[2]),
            umac_ng, rtol, atol, mode, C.byref(info)))
        return info

    def getFluxes(self, fx, fy, fz, ng=0):
        self.ctx.check(lib().mgx_mac_get_fluxes(self.h, fx.ptr, fy.ptr, fz.ptr, ng))

    def solve(self, phi, phi_ng, rhs, rhs_ng, rtol, atol=0.0, mode=MODE_FP64):
        info = SolveInfo()
        self.ctx.check(lib().mgx_mac_solve(self.h, phi.ptr, phi_ng, rhs.ptr, rhs_ng, rtol, atol,
                                           mode, C.byref(info)))
        return info

    def close(self):
        if self.h:
            lib().mgx_mac_destroy(self.h)
            self.h = C.c_void_p()


def _per(periodic):
    return (C.c_int * 3)(*[1 if p else 0 for p in periodic])


class UNet:
    """The learned coarse-level predictor (BF16-MFMA 3-D conv U-net)."""

    def __init__(self, ctx, base=16, levels=4, seed=0):
        self.ctx = ctx
        self.h = C.c_void_p()
        ctx.check(lib().mgx_unet_create(ctx.h, base, levels, seed, C.byref(self.h)))
        n = C.c_longlong()
        ctx.check(lib().mgx_unet_num_params(self.h, C.byref(n)))
        self.nparam = n.value

    def get_params(self):
        out = np.empty(self.nparam, dtype=np.float32)
        self.ctx.check(lib().mgx_unet_get_params(self.h, out.ctypes.data, self.nparam))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparam
        self.ctx.check(lib().mgx_unet_set_params(self.h, p.ctypes.data, self.nparam))

    def forward(self, x_dev, out_dev, shape, periodic=(1, 1, 1)):
        """shape = (batch, nz, ny, nx); x: [batch][nz][ny][nx][8], out: [batch][nz][ny][nx]."""
        b, nz, ny, nx = shape
        self.ctx.check(lib().mgx_unet_forward(self.h, x_dev.ptr, out_dev.ptr, nx, ny, nz, b,
                                              _per(periodic)))

    def gradients(self, x_dev, dout_dev, shape, periodic=(1, 1, 1)):
        b, nz, ny, nx = shape
        g = np.empty(self.nparam, dtype=np.float32)
        self.ctx.check(lib().mgx_unet_gradients(self.h, x_dev.ptr, dout_dev.ptr, nx, ny, nz, b,
                                                _per(periodic), g.ctypes.data, self.nparam))
        return g

    def train_step(self, x_dev, t_dev, shape, lr, periodic=(1, 1, 1)):
        b, nz, ny, nx = shape
        loss = C.c_float()
        self.ctx.check(lib().mgx_unet_train_step(self.h, x_dev.ptr, t_dev.ptr, nx, ny, nz, b,
                                                 _per(periodic), lr, C.byref(loss)))
        return loss.value

    def close(self):
        if self.h:
            lib().mgx_unet_destroy(self.h)
            self.h = C.c_void_p()


def conv3d_forward(ctx, x_dev, w_dev, b_dev, out_dev, shape, cin, cout, periodic=(1, 1, 1),
                   relu=False):
    b, nz, ny, nx = shape
    ctx.check(lib().mgx_conv3d_forward(ctx.h, x_dev.ptr, w_dev.ptr,
                                       b_dev.ptr if b_dev is not None else None, out_dev.ptr,
                                       nx, ny, nz, b, cin, cout, _per(periodic), int(relu)))


def conv3d_backward(ctx, x_dev, w_dev, dout_dev, din_dev, dw_dev, db_dev, shape, cin, cout,
                    periodic=(1, 1, 1)):
    b, nz, ny, nx = shape
    opt = lambda a: a.ptr if a is not None else None
    ctx.check(lib().mgx_conv3d_backward(ctx.h, x_dev.ptr, w_dev.ptr, dout_dev.ptr, opt(din_dev),
                                        opt(dw_dev), opt(db_dev), nx, ny, nz, b, cin, cout,
                                        _per(periodic)))


class Advection:
    """Godunov PPM advection pieces of the predictor (periodic, single level)."""

    def __init__(self, ctx, level):
        self.ctx = ctx
        self.h = C.c_void_p()
        ctx.check(lib().mgx_adv_create(ctx.h, C.byref(level), C.byref(self.h)))

    def extrap_vel_to_faces(self, vel, vel_ng, dt, umac, vmac, wmac, umac_ng=0, force=None,
                            f_ng=0):
        self.ctx.check(lib().mgx_adv_extrap_vel_to_faces(
            self.h, vel.ptr, vel_ng, force.ptr if force is not None else None, f_ng, dt,
            umac.ptr, vmac.ptr, wmac.ptr, umac_ng))

    def convective_term(self, vel, vel_ng, umac, vmac, wmac, umac_ng, dt, conv, conv_ng=0,
                        force=None, f_ng=0, faces=(None, None, None), fluxes=(None, None, None)):
        p = lambda a: a.ptr if a is not None else None
        self.ctx.check(lib().mgx_adv_convective_term(
            self.h, vel.ptr, vel_ng, p(force), f_ng, umac.ptr, vmac.ptr, wmac.ptr, umac_ng, dt,
            conv.ptr, conv_ng, p(faces[0]), p(faces[1]), p(faces[2]), p(fluxes[0]),
            p(fluxes[1]), p(fluxes[2])))

    def set_velocity_bcrec(self, bcrec):
        """amrex::BCType values, shape (3, 6): lo x, y, z, hi x, y, z per component."""
        b = (C.c_int * 18)(*[int(v) for v in np.asarray(bcrec).ravel()])
        self.ctx.check(lib().mgx_adv_set_velocity_bcrec(self.h, b))

    def scalar_term(self, state, state_ng, ncomp, umac, vmac, wmac, umac_ng, dt, conv,
                    conv_ng=0, iconserv=None, force=None, f_ng=0, faces=(None, None, None),
                    fluxes=(None, None, None), bcrec=None):
        """Density / tracer advection: edge states, fluxes and conv (see mgx.h)."""
        p = lambda a: a.ptr if a is not None else None
        ic = None if iconserv is None else (C.c_int * ncomp)(*[int(v) for v in iconserv])
        bc = None if bcrec is None else (C.c_int * (6 * ncomp))(
            *[int(v) for v in np.asarray(bcrec).ravel()])
        self.ctx.check(lib().mgx_adv_scalar_term(
            self.h, state.ptr, state_ng, ncomp, ic, bc, p(force), f_ng, umac.ptr, vmac.ptr,
            wmac.ptr, umac_ng, dt, conv.ptr, conv_ng, p(faces[0]), p(faces[1]), p(faces[2]),
            p(fluxes[0]), p(fluxes[1]), p(fluxes[2])))

    def close(self):
        if self.h:
            lib().mgx_adv_destroy(self.h)
            self.h = C.c_void_p()


def _ptr(a):
    return a.ptr if a is not None else None


def _d3(v):
    return None if v is None else (C.c_double * 3)(*[float(x) for x in v])


class Step:
    """Predictor-step pieces on one level (mgx_step_*): periodic ghost
    filling, forces, velocity and density updates."""

    def __init__(self, ctx, level):
        self.ctx = ctx
        self.h = C.c_void_p()
        ctx.check(lib().mgx_step_create(ctx.h, C.byref(level), C.byref(self.h)))

    def fill_boundary(self, field, ng, ncomp):
        self.ctx.check(lib().mgx_fill_boundary(self.h, field.ptr, ng, ncomp))

    def fill_boundary_noalloc(self, field, ng, ncomp):
        self.ctx.check(lib().mgx_fill_boundary_noalloc(self.h, field.ptr, ng, ncomp))

    def compute_vel_forces(self, gp, gp_ng, rho, rho_ng, forces, f_ng, divtau=None, divtau_ng=0,
                           gravity=None, gp0=None, include_pressure_gradient=True):
        self.ctx.check(lib().mgx_compute_vel_forces(
            self.h, _ptr(gp), gp_ng, rho.ptr, rho_ng, _ptr(divtau), divtau_ng, _d3(gravity),
            _d3(gp0), 1 if include_pressure_gradient else 0, forces.ptr, f_ng))

    def update_velocity(self, vel, vel_ng, vel_old, vo_ng, conv, conv_ng, forces, f_ng, dt,
                        divtau=None, divtau_ng=0, divtau_coef=0.0, rho_old=None, rho_new=None,
                        rho_nph=None, rho_ng=0):
        self.ctx.check(lib().mgx_update_velocity(
            self.h, vel.ptr, vel_ng, vel_old.ptr, vo_ng, conv.ptr, conv_ng, forces.ptr, f_ng,
            _ptr(divtau), divtau_ng, divtau_coef, _ptr(rho_old), _ptr(rho_new), _ptr(rho_nph),
            rho_ng, dt))

    def update_density(self, rho_old, ro_ng, drdt, c_ng, dt, rho_new, rn_ng, rho_nph=None,
                       rh_ng=0):
        self.ctx.check(lib().mgx_update_density(self.h, rho_old.ptr, ro_ng, drdt.ptr, c_ng, dt,
                                                rho_new.ptr, rn_ng, _ptr(rho_nph), rh_ng))

    def close(self):
        if self.h:
            lib().mgx_step_destroy(self.h)
            self.h = C.c_void_p()


class Diffusion:
    """Cell-centred viscous operator and Helmholtz solve (mgx_diff_*),
    DiffusionScalarOp's component solves."""

    def __init__(self, ctx, level):
        self.ctx = ctx
        self.h = C.c_void_p()
        ctx.check(lib().mgx_diff_create(ctx.h, C.byref(level), C.byref(self.h)))

    def setDomainBC(self, lo, hi, comp=-1):
        l = (C.c_int * 3)(*lo)
        u = (C.c_int * 3)(*hi)
        self.ctx.check(lib().mgx_diff_set_domain_bc(self.h, comp, l, u))

    def set_options(self, opts):
        self.ctx.check(lib().mgx_diff_set_options(self.h, C.byref(opts)))

    def set_eta(self, eta=None, ng=0, const_eta=0.0):
        self.ctx.check(lib().mgx_diff_set_eta(self.h, _ptr(eta), ng, const_eta))

    def apply(self, field, ng, ncomp, out, out_ng=0, rho=None, rho_ng=0):
        self.ctx.check(lib().mgx_diff_apply(self.h, field.ptr, ng, ncomp, _ptr(rho), rho_ng,
                                            out.ptr, out_ng))

    def solve(self, field, ng, ncomp, dt, acoef=None, a_ng=0, rtol=1e-11, atol=1e-14,
              mode=MODE_FP64):
        info = SolveInfo()
        self.ctx.check(lib().mgx_diff_solve(self.h, field.ptr, ng, ncomp, _ptr(acoef), a_ng, dt,
                                            rtol, atol, mode, C.byref(info)))
        return info

    def close(self):
        if self.h:
            lib().mgx_diff_destroy(self.h)
            self.h = C.c_void_p()


def write_multifab(ctx, prefix, level, field, ng, ncomp, nodal=False):
    ctx.check(lib().mgx_write_multifab(ctx.h, prefix.encode(), C.byref(level), field.ptr, ng,
                                       ncomp, int(nodal)))


def read_multifab(ctx, prefix, level, field, ng, ncomp, nodal=False):
    ctx.check(lib().mgx_read_multifab(ctx.h, prefix.encode(), C.byref(level), field.ptr, ng,
                                      ncomp, int(nodal)))


def write_checkpoint_header(path, levels, nstep, cur_time, dt, prev_dt, prev_prev_dt):
    arr = (LevelDesc * len(levels))(*levels)
    st = lib().mgx_write_checkpoint_header(path.encode(), len(levels), arr, nstep, cur_time, dt,
                                           prev_dt, prev_prev_dt)
    if st != MGX_OK:
        raise MgxError(st, f"mgx_write_checkpoint_header({path}) failed")


def read_checkpoint_header(path, max_levels=8, max_boxes=1 << 16):
    fl, ns = C.c_int(), C.c_int()
    times = (C.c_double * 4)()
    lo, hi = (C.c_double * 3)(), (C.c_double * 3)()
    nb = (C.c_int * max_levels)()
    bx = (C.c_int * (6 * max_boxes))()
    st = lib().mgx_read_checkpoint_header(path.encode(), C.byref(fl), C.byref(ns), times, lo, hi,
                                          nb, max_levels, bx, max_boxes)
    if st != MGX_OK:
        raise MgxError(st, f"mgx_read_checkpoint_header({path}) failed")
    boxes, used = [], 0
    for l in range(fl.value + 1):
        lv = []
        for b in range(nb[l]):
            v = bx[6 * (used + b): 6 * (used + b) + 6]
            lv.append((tuple(v[0:3]), tuple(v[3:6])))
        used += nb[l]
        boxes.append(lv)
    return {"finest_level": fl.value, "nstep": ns.value, "cur_time": times[0], "dt": times[1],
            "prev_dt": times[2], "prev_prev_dt": times[3], "prob_lo": list(lo),
            "prob_hi": list(hi), "boxes": boxes}


def write_plotfile(ctx, path, levels, data, ng, varnames, time, level_steps=None,
                   ref_ratio=None):
    nl = len(levels)
    arr = (LevelDesc * nl)(*levels)
    ptrs = (C.c_void_p * nl)(*[d.ptr for d in data])
    names = (C.c_char_p * len(varnames))(*[v.encode() for v in varnames])
    ls = None if level_steps is None else (C.c_int * nl)(*level_steps)
    rr = None if ref_ratio is None else (C.c_int * max(nl - 1, 1))(*ref_ratio)
    ctx.check(lib().mgx_write_plotfile(ctx.h, path.encode(), nl, arr, ptrs, ng, len(varnames),
                                       names, time, ls, rr))


# ---- dynamic refinement (ErrorEst, grid generation, FillPatch) ---------------
def tag_cells(ctx, level, rho=None, rho_ng=0, rhoerr=None, gradrhoerr=None, region=None,
              tags=None):
    """incflo::ErrorEst on one level -> host uint8 tags (nz, ny, nx); None
    disables a criterion; tags (a device uint8 array) accumulates when given."""
    n = [level.n_cell[d] for d in range(3)]
    t = tags if tags is not None else ctx.array(np.zeros(n[0] * n[1] * n[2], np.uint8), np.uint8)
    lo = hi = None
    if region is not None:
        lo = (C.c_double * 3)(*region[0])
        hi = (C.c_double * 3)(*region[1])
    inf = float("inf")
    ctx.check(lib().mgx_tag_cells(ctx.h, C.byref(level), _ptr(rho), rho_ng,
                                  inf if rhoerr is None else rhoerr,
                                  inf if gradrhoerr is None else gradrhoerr,
                                  C.cast(lo, C.c_void_p) if lo else None,
                                  C.cast(hi, C.c_void_p) if hi else None, t.ptr))
    out = t.download().reshape(n[2], n[1], n[0])
    if tags is None:
        t.free()
    return out


def cluster_tags(tags, periodic=(1, 1, 1), n_error_buf=1, blocking_factor=8, max_grid_size=32,
                 grid_eff=0.7, max_boxes=1 << 16):
    """AmrMesh::MakeNewGrids for level l+1 from level-l tags (nz, ny, nx) ->
    boxes [(lo, hi)] in the index space of level l+1."""
    t = np.ascontiguousarray(tags, dtype=np.uint8)
    n = (C.c_int * 3)(t.shape[2], t.shape[1], t.shape[0])
    per = (C.c_int * 3)(*periodic)
    out = (C.c_int * (6 * max_boxes))()
    nb = C.c_int()
    st = lib().mgx_cluster_tags(t.ctypes.data, n, per, n_error_buf, blocking_factor,
                                max_grid_size, grid_eff, out, max_boxes, C.byref(nb))
    if st != MGX_OK:
        raise MgxError(st, "mgx_cluster_tags failed")
    return [(tuple(out[6 * b:6 * b + 3]), tuple(out[6 * b + 3:6 * b + 6]))
            for b in range(nb.value)]


def fill_fine_level(ctx, crse, crse_field, c_ng, ncomp, fine, fine_field, f_ng, bcrec=None,
                    old_fine=None, old_field=None, of_ng=0):
    """FillPatchTwoLevels / InterpFromCoarseLevel with cell_cons_interp."""
    bc = None
    if bcrec is not None:
        flat = np.ascontiguousarray(np.asarray(bcrec, dtype=np.int32).ravel())
        bc = (C.c_int * flat.size)(*flat.tolist())
    ctx.check(lib().mgx_fill_fine_level(
        ctx.h, C.byref(crse), crse_field.ptr, c_ng, ncomp, bc, C.byref(fine), fine_field.ptr, f_ng,
        C.byref(old_fine) if old_fine is not None else None,
        old_field.ptr if old_field is not None else None, of_ng))


def eb_cylinder_geometry(level, cx, cy, r):
    """Host-side cylinder geometry of a level: (ctype (nz,ny,nx), moments (nz,ny,nx,27))."""
    n = [level.n_cell[d] for d in range(3)]
    ct = np.zeros((n[2], n[1], n[0]), dtype=np.int32)
    mom = np.zeros((n[2], n[1], n[0], 27))
    st = lib().mgx_eb_cylinder_geometry(C.byref(level), cx, cy, r, ct.ctypes.data, mom.ctypes.data)
    if st != MGX_OK:
        raise MgxError(st, "mgx_eb_cylinder_geometry failed")
    return ct, mom


def eb_cylinder_surface_weights(level, cx, cy, r):
    """Host-side body-surface weights (nz, ny, nx, 8, 3) of the cylinder."""
    n = [level.n_cell[d] for d in range(3)]
    out = np.zeros((n[2], n[1], n[0], 8, 3))
    st = lib().mgx_eb_cylinder_surface_weights(C.byref(level), cx, cy, r, out.ctypes.data)
    if st != MGX_OK:
        raise MgxError(st, "mgx_eb_cylinder_surface_weights failed")
    return out


def eb_cylinder_apertures(level, cx, cy, r):
    """Host-side face apertures [ax, ay, az] (canonical faces, (nz, ny, nx) + 1 along d)."""
    n = [level.n_cell[d] for d in range(3)]
    out = []
    for d in range(3):
        shape = [n[2], n[1], n[0]]
        shape[2 - d] += 1
        out.append(np.zeros(shape))
    st = lib().mgx_eb_cylinder_apertures(C.byref(level), cx, cy, r, out[0].ctypes.data,
                                         out[1].ctypes.data, out[2].ctypes.data)
    if st != MGX_OK:
        raise MgxError(st, "mgx_eb_cylinder_apertures failed")
    return out


def _ptrs(arrs):
    """array of device pointers (None -> NULL) for the per-level arguments"""
    return (C.c_void_p * len(arrs))(*[a.ptr if a is not None else None for a in arrs])


def amr_average_down_faces(ctx, crse, fine, fine_faces, fine_ng, crse_faces, crse_ng, ncomp=1):
    """average_down_faces over two levels (ratio 2): device face arrays per direction."""
    ff = (C.c_void_p * 3)(*[a.ptr for a in fine_faces])
    cf = (C.c_void_p * 3)(*[a.ptr for a in crse_faces])
    ctx.check(lib().mgx_amr_average_down_faces(ctx.h, C.byref(crse), C.byref(fine), ff, fine_ng,
                                               cf, crse_ng, ncomp))


def amr_average_down(ctx, crse, fine, fine_cells, fine_ng, crse_cells, crse_ng, ncomp=1):
    """average_down of a cell field over two levels (ratio 2)."""
    ctx.check(lib().mgx_amr_average_down(ctx.h, C.byref(crse), C.byref(fine), fine_cells.ptr,
                                         fine_ng, crse_cells.ptr, crse_ng, ncomp))


class AmrNodalProjector:
    """Composite nodal projection over AMR levels (mgx_amr_nodal_*): levels[0]
    tiles the domain, each further level holds the boxes of the next finer
    level (refinement ratio 2, properly nested; at most MGX_AMR_MAX_LEVELS)."""

    def __init__(self, ctx, levels):
        self.ctx = ctx
        self.levels = levels
        arr = (LevelDesc * len(levels))(*levels)
        self._arr = arr
        self.h = C.c_void_p()
        ctx.check(lib().mgx_amr_nodal_create(ctx.h, arr, len(levels), C.byref(self.h)))

    def setDomainBC(self, lo, hi):
        a = (C.c_int * 3)(*lo)
        b = (C.c_int * 3)(*hi)
        self.ctx.check(lib().mgx_amr_nodal_set_domain_bc(self.h, a, b))

    def set_options(self, opts):
        self.ctx.check(lib().mgx_amr_nodal_set_options(self.h, C.byref(opts)))

    def set_sigma(self, sigmas, ng=0):
        self.ctx.check(lib().mgx_amr_nodal_set_sigma(self.h, _ptrs(sigmas), ng))

    def project(self, vels, vel_ng, rtol, atol=0.0, phis=None, phi_ng=0, gphis=None, gphi_ng=0,
                mode=MODE_FP64):
        info = SolveInfo()
        self.ctx.check(lib().mgx_amr_nodal_project(
            self.h, _ptrs(vels), vel_ng, _ptrs(phis) if phis else None, phi_ng,
            _ptrs(gphis) if gphis else None, gphi_ng, rtol, atol, mode, C.byref(info)))
        return info

    def close(self):
        if self.h:
            lib().mgx_amr_nodal_destroy(self.h)
            self.h = C.c_void_p()


class AmrMacProjector:
    """Composite MAC projection over AMR levels (mgx_amr_mac_*): levels[0]
    tiles the domain, each further level holds the boxes of the next finer
    level (refinement ratio 2, properly nested).  Face fields are passed per
    level as [x, y, z]."""

    def __init__(self, ctx, levels):
        self.ctx = ctx
        self.levels = levels
        arr = (LevelDesc * len(levels))(*levels)
        self._arr = arr
        self.h = C.c_void_p()
        ctx.check(lib().mgx_amr_mac_create(ctx.h, arr, len(levels), C.byref(self.h)))

    def setDomainBC(self, lo, hi):
        a = (C.c_int * 3)(*lo)
        b = (C.c_int * 3)(*hi)
        self.ctx.check(lib().mgx_amr_mac_set_domain_bc(self.h, a, b))

    def set_options(self, opts):
        self.ctx.check(lib().mgx_amr_mac_set_options(self.h, C.byref(opts)))

    def initProjector(self, betas, ng=0):
        """betas: per level a list of 3 boxed face arrays"""
        self.ctx.check(lib().mgx_amr_mac_set_beta(self.h, _ptrs([b for lev in betas for b in lev]), ng))

    updateCoeffs = initProjector

    def project(self, umacs, umac_ng, rtol, atol=0.0, phis=None, phi_ng=0, mode=MODE_FP64):
        info = SolveInfo()
        self.ctx.check(lib().mgx_amr_mac_project(
            self.h, _ptrs([u for lev in umacs for u in lev]), umac_ng,
            _ptrs(phis) if phis else None, phi_ng, rtol, atol, mode, C.byref(info)))
        return info

    def close(self):
        if self.h:
            lib().mgx_amr_mac_destroy(self.h)
            self.h = C.c_void_p()
