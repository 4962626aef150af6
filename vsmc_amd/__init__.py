"""vsmc_amd -- MI355X (gfx950) backend of vSMC's per-step particle path.

The compute lives in the in-tree shared library ``libvsmch.so`` (HIP
kernels behind the C ABI of ``include/vsmch.h``); this module loads it with
ctypes and offers thin Python wrappers over device tensors (torch is used
only to allocate device memory). There is no CPU fallback: every call goes
to the HIP library, and loading fails loudly if it is missing.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvsmch.so")

WEIGHT_SET_LOG, WEIGHT_ADD_LOG, WEIGHT_SET, WEIGHT_MUL = 0, 1, 3, 4

PHILOX4X32 = 0
THREEFRY4X64 = 1

# vsmc::MatrixLayout (include/vsmc/internal/defines.hpp:49)
ROW_MAJOR = 101
COL_MAJOR = 102

# vsmc::ResampleScheme order (include/vsmc/internal/defines.hpp:49-56)
MULTINOMIAL = 0
STRATIFIED = 1
SYSTEMATIC = 2
RESIDUAL = 3
RESIDUAL_STRATIFIED = 4
RESIDUAL_SYSTEMATIC = 5
SCHEMES = {
    "multinomial": MULTINOMIAL,
    "stratified": STRATIFIED,
    "systematic": SYSTEMATIC,
    "residual": RESIDUAL,
    "residual_stratified": RESIDUAL_STRATIFIED,
    "residual_systematic": RESIDUAL_SYSTEMATIC,
}

# Every C symbol of include/vsmch.h with (restype, argtypes).
_c = ctypes
_P = _c.c_void_p
SIGNATURES = {
    "vsmch_last_error": (_c.c_char_p, []),
    "vsmch_version": (_c.c_int, []),
    "vsmch_set_stream": (_c.c_int, [_P]),
    "vsmch_synchronize": (_c.c_int, []),
    "vsmch_rng_words": (_c.c_int, [_c.c_int, _c.c_uint64, _c.c_uint64, _c.c_size_t, _P]),
    "vsmch_u01_fill": (_c.c_int, [_c.c_int, _c.c_uint64, _c.c_uint64, _c.c_size_t, _P]),
    "vsmch_normal_fill": (_c.c_int, [_c.c_int, _c.c_uint64, _c.c_uint64, _c.c_size_t,
                                     _c.c_double, _c.c_double, _c.c_int, _P]),
    "vsmch_resample_trans_u01_rep": (_c.c_int, [_c.c_int, _c.c_int, _P, _P, _P]),
    "vsmch_resample_trans_u01_rep_noalloc": (_c.c_int, [_c.c_int, _c.c_int, _P, _P, _P]),
    "vsmch_resample_trans_u01_index": (_c.c_int, [_c.c_int, _c.c_int, _P, _P, _P]),
    "vsmch_resample_trans_rep_index": (_c.c_int, [_c.c_int, _c.c_int, _P, _P]),
    "vsmch_resample_trans_index_rep": (_c.c_int, [_c.c_int, _c.c_int, _P, _P]),
    "vsmch_resample_trans_residual": (_c.c_int, [_c.c_int, _c.c_int, _P, _P, _P,
                                                 _c.POINTER(_c.c_int)]),
    "vsmch_resample": (_c.c_int, [_c.c_int, _c.c_int, _c.c_uint32,
                                  _c.POINTER(_c.c_uint64), _P, _P, _c.c_int]),
    "vsmch_exact_cumsum": (_c.c_int, [_c.c_size_t, _P, _P]),
    "vsmch_sorted_uniforms": (_c.c_int, [_c.c_size_t, _c.c_uint32, _c.c_uint64, _P]),
    "vsmch_particle_create": (_c.c_int, [_c.POINTER(_P), _P]),
    "vsmch_particle_destroy": (_c.c_int, [_P]),
    "vsmch_particle_state": (_c.c_int, [_P, _c.POINTER(_P)]),
    "vsmch_particle_read_state": (_c.c_int, [_P, _P]),
    "vsmch_particle_write_state": (_c.c_int, [_P, _P]),
    "vsmch_particle_log_weight": (_c.c_int, [_P, _c.POINTER(_P)]),
    "vsmch_particle_keys": (_c.c_int, [_P, _c.POINTER(_c.c_uint32),
                                       _c.POINTER(_c.c_uint32), _c.POINTER(_c.c_uint64)]),
    "vsmch_weight_set_equal": (_c.c_int, [_P]),
    "vsmch_weight_set_log": (_c.c_int, [_P, _P]),
    "vsmch_weight_add_log": (_c.c_int, [_P, _P]),
    "vsmch_weight_update": (_c.c_int, [_P, _c.c_int, _P, _c.c_int64]),
    "vsmch_particle_copy": (_c.c_int, [_P, _c.c_int64, _P]),
    "vsmch_weight_ess": (_c.c_int, [_P, _c.POINTER(_c.c_double)]),
    "vsmch_weight_log_zconst": (_c.c_int, [_P, _c.POINTER(_c.c_double)]),
    "vsmch_weight_read": (_c.c_int, [_P, _P]),
    "vsmch_particle_resample": (_c.c_int, [_P, _c.c_int, _c.c_double,
                                           _c.POINTER(_c.c_int), _P, _P]),
    "vsmch_particle_resample_count": (_c.c_int, [_P, _c.POINTER(_c.c_int64)]),
    "vsmch_particle_rng_words": (_c.c_int, [_P, _c.POINTER(_P)]),
    "vsmch_particle_scratch": (_c.c_int, [_P, _c.POINTER(_P)]),
    "vsmch_weighted_sum": (_c.c_int, [_P, _c.c_int, _P, _c.POINTER(_c.c_double)]),
    "vsmch_monitor_state_moments": (_c.c_int, [_P, _c.c_int, _c.c_int,
                                               _c.POINTER(_c.c_double)]),
    "vsmch_resample_multinomial": (_c.c_int, [_c.c_int, _c.c_int, _c.c_uint32,
                                              _c.POINTER(_c.c_uint64), _P, _P]),
    "vsmch_resample_stratified": (_c.c_int, [_c.c_int, _c.c_int, _c.c_uint32,
                                             _c.POINTER(_c.c_uint64), _P, _P]),
    "vsmch_resample_systematic": (_c.c_int, [_c.c_int, _c.c_int, _c.c_uint32,
                                             _c.POINTER(_c.c_uint64), _P, _P]),
    "vsmch_resample_residual": (_c.c_int, [_c.c_int, _c.c_int, _c.c_uint32,
                                           _c.POINTER(_c.c_uint64), _P, _P]),
    "vsmch_resample_residual_stratified": (_c.c_int, [_c.c_int, _c.c_int, _c.c_uint32,
                                                      _c.POINTER(_c.c_uint64), _P, _P]),
    "vsmch_resample_residual_systematic": (_c.c_int, [_c.c_int, _c.c_int, _c.c_uint32,
                                                      _c.POINTER(_c.c_uint64), _P, _P]),
    "vsmch_u01_sorted": (_c.c_int, [_c.c_int, _P, _P]),
    "vsmch_u01_stratified": (_c.c_int, [_c.c_int, _P, _P]),
    "vsmch_u01_systematic": (_c.c_int, [_c.c_int, _c.c_double, _P]),
    "vsmch_particle_cess": (_c.c_int, [_P, _P, _c.c_double, _c.POINTER(_c.c_double)]),
    "vsmch_particle_cess_solve": (_c.c_int, [_P, _P, _c.c_double, _c.c_double, _c.c_int,
                                             _c.POINTER(_c.c_double)]),
    "vsmch_cv_init": (_c.c_int, [_P, _P, _c.c_double, _c.c_double]),
    "vsmch_cv_move": (_c.c_int, [_P, _P, _c.c_int64, _c.c_double, _c.c_double]),
    "vsmch_cv_run": (_c.c_int, [_P, _P, _c.c_int64, _P, _P, _c.c_int, _c.c_double,
                                _P, _P, _P, _P]),
    "vsmch_cv_iterate": (_c.c_int, [_P, _P, _c.c_int64, _c.c_int64, _P, _P, _c.c_int,
                                    _c.c_double, _P, _P, _P, _P]),
    "vsmch_rindex_create": (_c.c_int, [_c.POINTER(_P), _c.c_int64]),
    "vsmch_rindex_destroy": (_c.c_int, [_P]),
    "vsmch_rindex_iter_size": (_c.c_int, [_P, _c.POINTER(_c.c_int64)]),
    "vsmch_rindex_reset": (_c.c_int, [_P]),
    "vsmch_rindex_clear": (_c.c_int, [_P]),
    "vsmch_rindex_push_back": (_c.c_int, [_P, _P]),
    "vsmch_rindex_insert": (_c.c_int, [_P, _c.c_int64, _P]),
    "vsmch_rindex_index": (_c.c_int, [_P, _c.c_int64, _c.c_int64, _c.POINTER(_c.c_int64)]),
    "vsmch_rindex_index_matrix": (_c.c_int, [_P, _c.c_int, _P]),
    "vsmch_mlp_create": (_c.c_int, [_c.POINTER(_P)]),
    "vsmch_mlp_destroy": (_c.c_int, [_P]),
    "vsmch_mlp_set_params": (_c.c_int, [_P, _P, _P, _P]),
    "vsmch_mlp_get_params": (_c.c_int, [_P, _P]),
    "vsmch_mlp_forward": (_c.c_int, [_P, _c.c_int64, _P, _P]),
    "vsmch_mlp_loss_grad": (_c.c_int, [_P, _c.c_int64, _P, _P, _P, _P]),
    "vsmch_mlp_train": (_c.c_int, [_P, _c.c_int64, _P, _P, _c.c_int64, _c.c_int64,
                                   _c.c_float, _P]),
    "vsmch_cv_transitions": (_c.c_int, [_P, _P, _c.c_int64, _P, _P, _c.c_int, _c.c_double,
                                        _c.c_uint32, _P, _P]),
    "vsmch_cv_move_learned": (_c.c_int, [_P, _P, _P, _c.c_int64, _c.c_double, _c.c_double]),
    "vsmch_batch_create": (_c.c_int, [_c.POINTER(_P), _c.c_int, _P, _P, _c.c_int,
                                      _c.c_double]),
    "vsmch_batch_destroy": (_c.c_int, [_P]),
    "vsmch_batch_set_obs": (_c.c_int, [_P, _P]),
    "vsmch_batch_run": (_c.c_int, [_P, _P, _c.c_int]),
    "vsmch_batch_histories": (_c.c_int, [_P, _P, _P, _P]),
    "vsmch_batch_read_filter": (_c.c_int, [_P, _c.c_int, _P, _P]),
    "vsmch_batch_info": (_c.c_int, [_P, _c.POINTER(_c.c_int64), _c.POINTER(_c.c_int64),
                                    _c.POINTER(_c.c_int), _c.POINTER(_c.c_double)]),
    "vsmch_cv_run_learned": (_c.c_int, [_P, _P, _P, _c.c_int64, _P, _P, _c.c_int,
                                        _c.c_double, _P, _P, _P, _P]),
    "vsmch_gmm_create": (_c.c_int, [_c.POINTER(_P), _P, _P, _c.c_int64]),
    "vsmch_gmm_destroy": (_c.c_int, [_P]),
    "vsmch_gmm_run": (_c.c_int, [_P, _c.c_int64, _c.POINTER(_c.c_int64), _P, _P, _P, _P,
                                 _P, _P]),
    "vsmch_gmm_particle": (_c.c_int, [_P, _c.POINTER(_P)]),
    "vsmch_gmm_move": (_c.c_int, [_P, _c.c_int, _c.c_double, _c.c_int64, _c.c_int]),
}


class GmmConfig(ctypes.Structure):
    """vsmch_gmm_config."""
    _fields_ = [("n", ctypes.c_int64), ("ncomp", ctypes.c_int), ("seed", ctypes.c_uint32),
                ("cess", ctypes.c_double), ("moves", ctypes.c_int),
                ("search_rounds", ctypes.c_int), ("scheme", ctypes.c_int),
                ("threshold", ctypes.c_double)]

MLP_PARAMS = 5128


class ParticleConfig(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("dim", ctypes.c_int),
                ("state_bytes", ctypes.c_int), ("seed", ctypes.c_uint32)]


class CvModel(ctypes.Structure):
    """vsmch_cv_model: constant-velocity PF constants (example/pf/include/pf.hpp)."""
    _fields_ = [("dt", ctypes.c_double), ("sd_pos", ctypes.c_double),
                ("sd_vel", ctypes.c_double), ("sd_pos0", ctypes.c_double),
                ("sd_vel0", ctypes.c_double), ("obs_model", ctypes.c_int),
                ("obs_scale", ctypes.c_double), ("obs_df", ctypes.c_double)]


class BatchFilter(ctypes.Structure):
    """vsmch_batch_filter."""
    _fields_ = [("n", ctypes.c_int64), ("T", ctypes.c_int64), ("seed", ctypes.c_uint32)]


def cv_model(dt=0.1, sd_pos=0.02, sd_vel=0.02, sd_pos0=2.0, sd_vel0=1.0,
             obs_model=0, obs_scale=0.1, obs_df=10.0):
    return CvModel(dt, sd_pos, sd_vel, sd_pos0, sd_vel0, obs_model, obs_scale, obs_df)


class VsmchError(RuntimeError):
    pass


_lib = None


def hip_runtimes():
    """Paths of the HIP runtime libraries mapped into this process."""
    try:
        with open("/proc/self/maps") as f:
            return sorted({ln.split()[-1] for ln in f if "libamdhip64" in ln})
    except OSError:
        return []


def lib():
    """Loads libvsmch.so (built by vsmc_amd.build); raises if it is absent.

    torch is imported first so that the library's libamdhip64.so.7 dependency
    binds to the HIP runtime torch already carries (same soname); device
    pointers then come from one runtime. Two runtimes in one process would
    hand each other unknown allocations, so that case raises."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise VsmchError(
                f"{LIB_PATH} is missing: build it with `python -m vsmc_amd.build`")
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        if len(hip_runtimes()) > 1:
            raise VsmchError(f"two HIP runtimes are loaded: {hip_runtimes()}")
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().vsmch_last_error()
        raise VsmchError(f"vsmch status {rc}: {msg.decode() if msg else ''}")


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream_sync():
    check(lib().vsmch_synchronize())


# ------------------------------------------------------------ streams

def rng_words(kind, key, offset, n, device="cuda"):
    import torch
    dt = torch.int32 if kind == PHILOX4X32 else torch.int64
    out = torch.empty(n, dtype=dt, device=device)
    check(lib().vsmch_rng_words(kind, key, offset, n, _ptr(out)))
    _stream_sync()
    return out


def u01_fill(kind, key, offset, n, out=None, device="cuda"):
    import torch
    if out is None:
        out = torch.empty(n, dtype=torch.float64, device=device)
    check(lib().vsmch_u01_fill(kind, key, offset, n, _ptr(out)))
    return out


def normal_fill(kind, key, offset, n, mean=0.0, sd=1.0, form=0, out=None, device="cuda"):
    import torch
    if out is None:
        out = torch.empty(n, dtype=torch.float64, device=device)
    check(lib().vsmch_normal_fill(kind, key, offset, n, mean, sd, form, _ptr(out)))
    return out


# ------------------------------------------------- resample transforms

def resample_trans_u01_rep(weight, u01, m=None, n=None):
    """resample_trans_u01_rep (include/vsmc/resample/transform.hpp:42-71)."""
    import torch
    m = weight.numel() if m is None else m
    n = u01.numel() if n is None else n
    rep = torch.empty(max(m, 1), dtype=torch.int32, device=weight.device)
    check(lib().vsmch_resample_trans_u01_rep(m, n, _ptr(weight), _ptr(u01), _ptr(rep)))
    _stream_sync()
    return rep[:m]


NOALLOC_MAX = 1 << 22


def resample_trans_u01_rep_noalloc(weight, u01, m=None, n=None, rep=None):
    """The counts of resample_trans_u01_rep from the library's fixed device
    scratch (vsmch_resample_trans_u01_rep_noalloc: no device allocation in
    the library; m, n <= NOALLOC_MAX). `rep` may be passed in."""
    import torch
    m = weight.numel() if m is None else m
    n = u01.numel() if n is None else n
    if rep is None:
        rep = torch.empty(max(m, 1), dtype=torch.int32, device=weight.device)
    check(lib().vsmch_resample_trans_u01_rep_noalloc(m, n, _ptr(weight), _ptr(u01),
                                                     _ptr(rep)))
    _stream_sync()
    return rep[:m]


def resample_trans_u01_index(weight, u01):
    import torch
    m, n = weight.numel(), u01.numel()
    idx = torch.empty(max(n, 1), dtype=torch.int32, device=weight.device)
    check(lib().vsmch_resample_trans_u01_index(m, n, _ptr(weight), _ptr(u01), _ptr(idx)))
    _stream_sync()
    return idx[:n]


def resample_trans_rep_index(rep, n=None):
    import torch
    m = rep.numel()
    n = m if n is None else n
    idx = torch.empty(max(n, 1), dtype=torch.int32, device=rep.device)
    check(lib().vsmch_resample_trans_rep_index(m, n, _ptr(rep), _ptr(idx)))
    _stream_sync()
    return idx[:n]


def resample_trans_index_rep(idx, m):
    import torch
    rep = torch.empty(max(m, 1), dtype=torch.int32, device=idx.device)
    check(lib().vsmch_resample_trans_index_rep(m, idx.numel(), _ptr(idx), _ptr(rep)))
    _stream_sync()
    return rep[:m]


def resample_trans_residual(weight, n):
    import torch
    m = weight.numel()
    resid = torch.empty(max(m, 1), dtype=torch.float64, device=weight.device)
    integ = torch.empty(max(m, 1), dtype=torch.int32, device=weight.device)
    R = ctypes.c_int(0)
    check(lib().vsmch_resample_trans_residual(m, n, _ptr(weight), _ptr(resid),
                                              _ptr(integ), ctypes.byref(R)))
    return resid[:m], integ[:m], R.value


def resample(scheme, weight, n, key, word_offset=0, m=None):
    """Resample<Scheme>()(M, N, rng, W, rep) with the Philox stream (key,
    word_offset); returns (rep, new word offset)."""
    import torch
    m = weight.numel() if m is None else m
    rep = torch.empty(max(m, 1), dtype=torch.int32, device=weight.device)
    off = ctypes.c_uint64(word_offset)
    check(lib().vsmch_resample(m, n, key, ctypes.byref(off), _ptr(weight), _ptr(rep),
                               scheme))
    return rep[:m], off.value


def exact_cumsum(w):
    import torch
    out = torch.empty_like(w)
    check(lib().vsmch_exact_cumsum(w.numel(), _ptr(w), _ptr(out)))
    _stream_sync()
    return out


def sorted_uniforms(n, key, word_offset=0, device="cuda"):
    import torch
    out = torch.empty(max(n, 1), dtype=torch.float64, device=device)
    check(lib().vsmch_sorted_uniforms(n, key, word_offset, _ptr(out)))
    _stream_sync()
    return out[:n]


_SCHEME_NAMES = ("multinomial", "stratified", "systematic", "residual",
                 "residual_stratified", "residual_systematic")


def resample_scheme(name, weight, n, key, word_offset=0, m=None):
    """vsmc_resample_<name>(m, n, rng, W, rep): resample with the scheme in
    the entry point's name; returns (rep, new word offset)."""
    import torch
    if name not in _SCHEME_NAMES:
        raise ValueError(f"unknown resampling scheme {name!r}")
    m = weight.numel() if m is None else m
    rep = torch.empty(max(m, 1), dtype=torch.int32, device=weight.device)
    off = ctypes.c_uint64(word_offset)
    check(getattr(lib(), "vsmch_resample_" + name)(m, n, key, ctypes.byref(off),
                                                    _ptr(weight), _ptr(rep)))
    return rep[:m], off.value


def u01_sorted(u01):
    """u01_sorted of n uniforms (include/vsmc/rng/u01_sequence.hpp:54-62)."""
    import torch
    out = torch.empty(max(u01.numel(), 1), dtype=torch.float64, device=u01.device)
    check(lib().vsmch_u01_sorted(u01.numel(), _ptr(u01), _ptr(out)))
    _stream_sync()
    return out[:u01.numel()]


def u01_stratified(u01):
    """u01_stratified of n uniforms (include/vsmc/rng/u01_sequence.hpp:67-74)."""
    import torch
    out = torch.empty(max(u01.numel(), 1), dtype=torch.float64, device=u01.device)
    check(lib().vsmch_u01_stratified(u01.numel(), _ptr(u01), _ptr(out)))
    _stream_sync()
    return out[:u01.numel()]


def u01_systematic(n, u01, device="cuda"):
    """u01_systematic of one uniform (include/vsmc/rng/u01_sequence.hpp:80-89)."""
    import torch
    out = torch.empty(max(n, 1), dtype=torch.float64, device=device)
    check(lib().vsmch_u01_systematic(n, float(u01), _ptr(out)))
    _stream_sync()
    return out[:n]


from .sampler import (Particle, CvFilter, ProposalNet, CvBatch, GmmSampler,  # noqa: E402,F401
                      ResampleIndex)
