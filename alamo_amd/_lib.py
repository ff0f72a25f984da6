"""ctypes binding of libalamo_amd.so (the C ABI declared in include/alamo_amd.h).

The library is built in-tree by alamo_amd/build.py.  Loading fails loudly when it
is missing: there is no CPU fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# ALAMO_AMD_LIB selects another build of the library (comparisons of kernel variants)
LIB_PATH = os.environ.get("ALAMO_AMD_LIB") or os.path.join(HERE, "lib", "libalamo_amd.so")

c_int, c_double, c_void_p = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
P = ctypes.POINTER


class MGDesc(ctypes.Structure):
    _fields_ = [("n_cell", c_int * 3), ("dx", c_double * 3), ("sym", c_int),
                ("bc_type", (c_int * 3) * 26), ("bc_val", (c_double * 3) * 26),
                ("max_coarsening_level", c_int), ("precision", c_int), ("cycle", c_int),
                ("pre_smooth", c_int), ("post_smooth", c_int), ("sweeps_per_smooth", c_int),
                ("omega", c_double), ("max_iter", c_int), ("bottom_max_iter", c_int),
                ("bottom_tol_rel", c_double), ("verbose", c_int), ("smoother", c_int), ("coarse_op", c_int),
                ("fixed_iter", c_int),
                ("bottom_solver", c_int), ("max_fmg_iter", c_int), ("sweeps_finest", c_int)]


class SolveStats(ctypes.Structure):
    _fields_ = [("iterations", c_int), ("converged", c_int), ("initial_res_inf", c_double),
                ("final_res_inf", c_double), ("final_res_l2", c_double), ("rhs_inf", c_double),
                ("target", c_double), ("seconds", c_double), ("fallback_events", c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class KrylovOpts(ctypes.Structure):
    _fields_ = [("restart", c_int), ("precond", c_int), ("max_iter", c_int), ("stagnation_ratio", c_double),
                ("verbose", c_int)]


# name -> (restype, argtypes)
class TrainOpts(ctypes.Structure):
    _fields_ = [("npairs", c_int), ("steps", c_int), ("lr", c_double), ("seed", ctypes.c_uint64),
                ("solve_max_iter", c_int), ("solve_tol", c_double), ("verbose", c_int), ("loss", c_int)]


class TrainStats(ctypes.Structure):
    _fields_ = [("pairs", c_int), ("pairs_converged", c_int), ("steps", c_int),
                ("first_loss", c_double), ("last_loss", c_double),
                ("seconds_data", c_double), ("seconds_train", c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


SIGNATURES = {
    "alamo_last_error": (ctypes.c_char_p, []),
    "alamo_version": (c_int, []),
    "alamo_mg_desc_default": (c_int, [P(MGDesc)]),
    "alamo_mg_create": (c_int, [P(MGDesc), P(c_void_p)]),
    "alamo_mg_destroy": (c_int, [c_void_p]),
    "alamo_mg_set_stream": (c_int, [c_void_p, c_void_p]),
    "alamo_mg_num_levels": (c_int, [c_void_p, P(c_int)]),
    "alamo_mg_level_ncell": (c_int, [c_void_p, c_int, P(c_int)]),
    "alamo_mg_level_layout": (c_int, [c_void_p, c_int, ctypes.POINTER(ctypes.c_int64)]),
    "alamo_mg_update_opts": (c_int, [c_void_p, P(MGDesc)]),
    "alamo_mg_strain": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p]),
    "alamo_mg_stress": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p]),
    "alamo_mg_energy": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "alamo_newton_dw": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "alamo_newton_w": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "alamo_mg_set_uniform": (c_int, [c_void_p, c_int]),
    "alamo_mg_norm_inf": (c_int, [c_void_p, c_int, c_void_p, P(c_double), c_void_p]),
    "alamo_mg_norm_l2": (c_int, [c_void_p, c_int, c_void_p, P(c_double), c_void_p]),
    "alamo_mg_reflux": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "alamo_learned_solve": (c_int, [c_void_p, c_void_p, c_void_p, c_double, c_double, P(KrylovOpts), P(SolveStats)]),
    "alamo_newton_solve": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_double, c_double, c_double,
                                   P(SolveStats), P(c_double)]),
    "alamo_mg_trac_xhi": (c_int, [c_void_p, c_void_p, c_void_p, P(c_double)]),
    "alamo_blend_models": (c_int, [ctypes.c_int64, c_int, c_int, c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    "alamo_mg_set_psi": (c_int, [c_void_p, c_int, c_void_p, c_double]),
    "alamo_mg_set_psi_host": (c_int, [c_void_p, c_int, c_void_p, c_double]),
    "alamo_mg_set_model": (c_int, [c_void_p, c_int, c_void_p, c_int]),
    "alamo_mg_set_model_host": (c_int, [c_void_p, c_int, c_void_p, c_int]),
    "alamo_mg_set_model_grains_host": (c_int, [c_void_p, c_void_p, c_void_p, c_int]),
    "alamo_mg_prepare": (c_int, [c_void_p]),
    "alamo_mg_get_coeff": (c_int, [c_void_p, c_int, c_void_p]),
    "alamo_mg_fapply": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "alamo_mg_diagonal": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p]),
    "alamo_mg_fsmooth": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "alamo_mg_fsmooth_noalloc": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "alamo_mg_error0x": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "alamo_mg_normalize": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p]),
    "alamo_mg_residual": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "alamo_mg_restrict": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "alamo_mg_interpolate": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "alamo_mg_norms": (c_int, [c_void_p, c_int, c_void_p, P(c_double), P(c_double), c_void_p]),
    "alamo_bc_init": (c_int, [c_void_p, c_void_p, c_int]),
    "alamo_newton_rhs": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "alamo_linear_solve": (c_int, [c_void_p, c_void_p, c_void_p, c_double, c_double, P(SolveStats)]),
    "alamo_krylov_opts_default": (c_int, [P(KrylovOpts)]),
    "alamo_krylov_solve": (c_int, [c_void_p, c_void_p, c_void_p, c_double, c_double, P(KrylovOpts),
                                   P(SolveStats)]),
    "alamo_net_create": (c_int, [P(c_void_p), ctypes.c_uint64]),
    "alamo_net_destroy": (c_int, [c_void_p]),
    "alamo_net_num_params": (c_int, [c_void_p, P(ctypes.c_int64)]),
    "alamo_net_get_params": (c_int, [c_void_p, c_void_p]),
    "alamo_net_set_params": (c_int, [c_void_p, c_void_p]),
    "alamo_net_get_grads": (c_int, [c_void_p, c_void_p]),
    "alamo_net_zero_grads": (c_int, [c_void_p]),
    "alamo_net_forward_host": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "alamo_net_backward_host": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "alamo_net_adam_step": (c_int, [c_void_p, c_double]),
    "alamo_net_train": (c_int, [c_void_p, c_void_p, P(TrainOpts), P(TrainStats)]),
    "alamo_net_save": (c_int, [c_void_p, ctypes.c_char_p]),
    "alamo_net_load": (c_int, [c_void_p, ctypes.c_char_p]),
    "alamo_mg_set_net": (c_int, [c_void_p, c_void_p]),
    "alamo_mg_precond": (c_int, [c_void_p, c_int, c_void_p, c_void_p]),
    "alamo_mg_apply_transpose": (c_int, [c_void_p, c_void_p, c_void_p]),
    "alamo_amr_fine_mask": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "alamo_amr_reflux": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "alamo_amr_average_down": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "alamo_amr_fixup_residual_mask": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "alamo_net_mfma_probe": (c_int, [c_void_p, c_void_p, c_void_p]),
    "alamo_net_mfma_probe32": (c_int, [c_void_p, c_void_p, c_void_p]),
    "alamo_write_plotfile": (c_int, [ctypes.c_char_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                                     c_void_p, c_void_p, c_double, c_void_p]),
    "alamo_write_plotfile_host": (c_int, [ctypes.c_char_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                          c_void_p, c_int, c_void_p, c_void_p, c_double, c_void_p]),
    "alamo_net_wgrad_host": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "alamo_stream_probe": (c_int, [c_void_p, c_void_p, ctypes.c_int64, c_void_p]),
    "alamo_amr_create": (c_int, [P(MGDesc), c_int, c_void_p, c_void_p, c_void_p, P(c_void_p)]),
    "alamo_amr_destroy": (c_int, [c_void_p]),
    "alamo_amr_level": (c_int, [c_void_p, c_int, P(c_void_p)]),
    "alamo_amr_masks": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "alamo_amr_solve": (c_int, [c_void_p, c_void_p, c_void_p, c_double, c_double, P(SolveStats)]),
    "alamo_net_layer_times": (c_int, [c_void_p, c_void_p, c_int, c_void_p]),
    "alamo_net_layer_times_bwd": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libalamo_amd.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class AlamoError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("alamo error %d: %s" % (code, msg))
        self.code = code


def dev_ptr(t, n, what, dtype="float64"):
    """Pointer of a device field, checked before any launch: a contiguous torch tensor on the GPU
    of the given dtype holding at least n elements.  The kernels index n elements from the
    pointer, so a host tensor, another dtype or a short buffer is an error here rather than an
    out-of-bounds access on the device."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise AlamoError(1, "%s: expected a torch tensor on the GPU, got %s" % (what, type(t).__name__))
    if not t.is_cuda:
        raise AlamoError(1, "%s: tensor is on %s, not on the GPU" % (what, t.device))
    if t.dtype != getattr(torch, dtype):
        raise AlamoError(1, "%s: dtype %s, expected %s" % (what, t.dtype, dtype))
    if not t.is_contiguous():
        raise AlamoError(1, "%s: tensor is not contiguous" % what)
    if t.numel() < n:
        raise AlamoError(1, "%s: %d elements, need at least %d" % (what, t.numel(), n))
    return ctypes.c_void_p(t.data_ptr())


def check(rc):
    if rc != 0:
        raise AlamoError(rc, lib().alamo_last_error().decode())
    return rc
