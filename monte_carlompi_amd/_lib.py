"""ctypes bindings of include/mcsnow.h (libmcsnow.so, built in-tree).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libmcsnow.so')

MCS_OK = 0
ERRORS = {-1: 'MCS_ERR_ARG', -2: 'MCS_ERR_HIP', -3: 'MCS_ERR_NOMEM', -4: 'MCS_ERR_STEP_CAP',
          -5: 'MCS_ERR_REPLAY', -6: 'MCS_ERR_STATE', -7: 'MCS_ERR_PHI_UNBOUND'}
MCS_NATIVE = 0
MCS_REPLAY = 1
THETA_BINS = 18
PHI_BINS = 36
BRF_BINS = THETA_BINS * PHI_BINS
TALLY_F64 = 656
TALLY_U64 = 16
MAX_BANDS = 256

# names declared in include/mcsnow.h (tests check every one is exported)
EXPORTS = ['mcs_create', 'mcs_destroy', 'mcs_last_error', 'mcs_device_cus', 'mcs_set_cases',
           'mcs_set_launch', 'mcs_get_launch', 'mcs_walk', 'mcs_sync', 'mcs_read_records',
           'mcs_read_tallies', 'mcs_read_counts', 'mcs_last_stats', 'mcs_autotune', 'mcs_philox',
           'mcs_set_polar', 'mcs_read_stokes', 'mcs_walk_noalloc',
           'mcs_mlp_create', 'mcs_mlp_destroy', 'mcs_mlp_last_error', 'mcs_mlp_param_count',
           'mcs_mlp_set_params', 'mcs_mlp_get_params', 'mcs_mlp_forward', 'mcs_mlp_last_ms',
           'mcs_mlp_grads', 'mcs_mlp_set_dataset', 'mcs_mlp_train_steps', 'mcs_mlp_set_serving',
           'mcs_mlp_serve', 'mcs_write_records_text', 'mcs_repr_doubles']
MCS_FEATURES_IDENTITY = 0
MCS_FEATURES_SNOW_BRF = 1


class McsCase(ctypes.Structure):
    _fields_ = [('tau_tot', ctypes.c_double), ('mux0', ctypes.c_double), ('muz0', ctypes.c_double),
                ('r_lambertian', ctypes.c_double), ('lambertian_bottom', ctypes.c_int32),
                ('band_begin', ctypes.c_int32), ('n_bands', ctypes.c_int32),
                ('lambertian_surface', ctypes.c_int32), ('n_photons', ctypes.c_int64)]


class McsBand(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in
                ('wvn', 'g', 'ssa_ice', 'ssa_imp', 'ext_cff', 'p_ext_imp', 'snow_depth', 'cdf')]


class McsWalkArgs(ctypes.Structure):
    _fields_ = [('mode', ctypes.c_int32), ('want_records', ctypes.c_int32),
                ('seed', ctypes.c_uint64), ('max_steps', ctypes.c_int64),
                ('photon_band', ctypes.c_void_p), ('replay_dev', ctypes.c_void_p),
                ('replay_off', ctypes.c_void_p), ('async_', ctypes.c_int32),
                ('use_graph', ctypes.c_int32)]


DEFAULT_COOP_LIVE = 16


class McsLaunch(ctypes.Structure):
    _fields_ = [('block', ctypes.c_int32), ('blocks_per_cu', ctypes.c_int32),
                ('chunk', ctypes.c_int32), ('refill_min', ctypes.c_int32),
                ('coop_live', ctypes.c_int32)]

    @classmethod
    def of(cls, shape):
        """(block, blocks_per_cu, chunk, refill_min[, coop_live])."""
        shape = tuple(int(v) for v in shape)
        if len(shape) == 4:
            shape += (DEFAULT_COOP_LIVE,)
        return cls(*shape)

    def as_tuple(self):
        return (self.block, self.blocks_per_cu, self.chunk, self.refill_min, self.coop_live)


class McsPolarCase(ctypes.Structure):
    _fields_ = [('table', ctypes.c_int32), ('cutoff', ctypes.c_int32),
                ('stokes0', ctypes.c_double * 4)]


class McsRecords(ctypes.Structure):
    _fields_ = [(k, ctypes.c_void_p) for k in
                ('condition', 'band', 'n_scat', 'theta_n', 'phi_n', 'path_length', 'wvn', 'snow_depth')]


class McsTextRecords(ctypes.Structure):
    """Host record columns for mcs_write_records_text (mcs_text_records)."""
    _fields_ = [(k, ctypes.c_void_p) for k in
                ('condition', 'wvn', 'theta_n', 'phi_n', 'n_scat', 'path_length', 'snow_depth', 'q', 'u', 'v')]


class McsDevOut(ctypes.Structure):
    """Device pointers of mcs_walk_noalloc's outputs (mcs_dev_out)."""
    _fields_ = [(k, ctypes.c_void_p) for k in
                ('tallies', 'counts', 'n_scat', 'cond_band', 'theta_n', 'phi_n', 'path_length', 'stokes')]


NOALLOC_MAX_CASES = 65536


class McsStats(ctypes.Structure):
    _fields_ = [('photons', ctypes.c_uint64), ('steps', ctypes.c_uint64),
                ('capped', ctypes.c_uint64), ('phi_unbound', ctypes.c_uint64),
                ('kernel_ms', ctypes.c_double)]


class McsMlpDims(ctypes.Structure):
    _fields_ = [('n_in', ctypes.c_int32), ('hidden', ctypes.c_int32),
                ('n_hidden_layers', ctypes.c_int32), ('n_out', ctypes.c_int32),
                ('max_batch', ctypes.c_int32), ('reserved', ctypes.c_int32)]


class McsAdam(ctypes.Structure):
    _fields_ = [('lr', ctypes.c_float), ('beta1', ctypes.c_float), ('beta2', ctypes.c_float),
                ('eps', ctypes.c_float), ('weight_decay', ctypes.c_float), ('step', ctypes.c_int32)]


_LIB = None


class McsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__('%s (%d): %s' % (ERRORS.get(code, 'MCS_ERR'), code, msg))
        self.code = code


def lib():
    """Load libmcsnow.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError('%s is not built: run `python -m monte_carlompi_amd.build` '
                           '(there is no CPU fallback)' % LIB_PATH)
    # One HIP runtime per process.  torch carries its own libamdhip64 /
    # libhsa-runtime64, which its libraries load by file name: if this library
    # bound /opt/rocm's copies first, importing torch later would load a second
    # runtime that cannot open the GPU.  Loading torch first makes this
    # library bind to torch's copies (same sonames).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    I = ctypes.c_int
    sig = {
        'mcs_create': ([I, ctypes.POINTER(P)], I),
        'mcs_destroy': ([P], I),
        'mcs_last_error': ([P], ctypes.c_char_p),
        'mcs_device_cus': ([P, ctypes.POINTER(ctypes.c_int32)], I),
        'mcs_set_cases': ([P, P, ctypes.c_int32, P, ctypes.c_int32], I),
        'mcs_set_launch': ([P, ctypes.POINTER(McsLaunch)], I),
        'mcs_get_launch': ([P, ctypes.POINTER(McsLaunch)], I),
        'mcs_walk': ([P, ctypes.POINTER(McsWalkArgs)], I),
        'mcs_sync': ([P], I),
        'mcs_read_records': ([P, ctypes.POINTER(McsRecords), ctypes.c_int64], I),
        'mcs_read_tallies': ([P, P, ctypes.c_int32], I),
        'mcs_read_counts': ([P, P, ctypes.c_int32], I),
        'mcs_last_stats': ([P, ctypes.POINTER(McsStats)], I),
        'mcs_autotune': ([P, P, ctypes.c_int32, ctypes.c_int64, ctypes.c_int32, ctypes.c_uint64,
                          P, ctypes.POINTER(McsLaunch)], I),
        'mcs_philox': ([P, P, ctypes.c_uint32, ctypes.c_uint32, P, ctypes.c_int64], I),
        'mcs_set_polar': ([P, P, ctypes.c_int32, P, ctypes.c_int32, ctypes.c_int32], I),
        'mcs_read_stokes': ([P, P, P, P, ctypes.c_int64], I),
        'mcs_walk_noalloc': ([P, ctypes.POINTER(McsWalkArgs), ctypes.POINTER(McsDevOut)], I),
        'mcs_mlp_create': ([I, ctypes.POINTER(McsMlpDims), ctypes.POINTER(P)], I),
        'mcs_mlp_destroy': ([P], I),
        'mcs_mlp_last_error': ([P], ctypes.c_char_p),
        'mcs_mlp_set_params': ([P, P, ctypes.c_int64], I),
        'mcs_mlp_get_params': ([P, P, ctypes.c_int64], I),
        'mcs_mlp_param_count': ([P, ctypes.POINTER(ctypes.c_int64)], I),
        'mcs_mlp_forward': ([P, P, ctypes.c_int64, P], I),
        'mcs_mlp_last_ms': ([P, ctypes.POINTER(ctypes.c_double)], I),
        'mcs_mlp_grads': ([P, P, P, P, ctypes.c_int32, P, ctypes.POINTER(ctypes.c_float)], I),
        'mcs_mlp_set_dataset': ([P, P, P, ctypes.c_int64], I),
        'mcs_mlp_train_steps': ([P, P, ctypes.c_int32, ctypes.c_int32, P, ctypes.POINTER(McsAdam), P], I),
        'mcs_mlp_set_serving': ([P, ctypes.c_int32, P, P, P, P, P], I),
        'mcs_mlp_serve': ([P, P, ctypes.c_int64, P], I),
        'mcs_write_records_text': ([ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(McsTextRecords),
                                    ctypes.c_int64], I),
        'mcs_repr_doubles': ([P, ctypes.c_int64, P, ctypes.c_int64], ctypes.c_int64),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _LIB = L
    return L


def check(code, ctx=None, errfn='mcs_last_error'):
    if code != MCS_OK:
        msg = ''
        if ctx:
            raw = getattr(lib(), errfn)(ctx)
            msg = raw.decode() if raw else ''
        raise McsError(code, msg)
