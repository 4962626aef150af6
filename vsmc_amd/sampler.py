"""Particle / Weight / Monitor over the C ABI (include/vsmch.h).

Mirrors the reference's Particle<T> (include/vsmc/core/particle.hpp),
Weight (include/vsmc/core/weight.hpp) and Monitor::eval
(include/vsmc/core/monitor.hpp) for an SoA state living on the device.
"""
import ctypes

from . import lib, check, ParticleConfig, cv_model, VsmchError  # noqa: F401


class Particle:
    """Particle<T> with a device state of `dim` components (float32/float64)."""

    def __init__(self, n, dim, dtype="float32", seed=0):
        import torch
        self.n, self.dim = n, dim
        self.dtype = torch.float32 if dtype in ("float32", torch.float32) else torch.float64
        cfg = ParticleConfig(n, dim, 4 if self.dtype == torch.float32 else 8, seed)
        h = ctypes.c_void_p()
        check(lib().vsmch_particle_create(ctypes.byref(h), ctypes.byref(cfg)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().vsmch_particle_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # raw device pointers, wrapped as tensors that alias the library buffers
    def _state_ptr(self):
        p = ctypes.c_void_p()
        check(lib().vsmch_particle_state(self._h, ctypes.byref(p)))
        return p.value

    def _logw_ptr(self):
        p = ctypes.c_void_p()
        check(lib().vsmch_particle_log_weight(self._h, ctypes.byref(p)))
        return p.value

    def state(self):
        """Copy of the state as a (dim, n) tensor (ColMajor StateMatrix)."""
        import torch
        out = torch.empty((self.dim, self.n), dtype=self.dtype, device="cuda")
        torch.cuda.synchronize()
        check(lib().vsmch_particle_read_state(self._h, ctypes.c_void_p(out.data_ptr())))
        return out

    def set_state(self, x):
        import torch
        x = x.to(device="cuda", dtype=self.dtype).contiguous()
        assert x.shape == (self.dim, self.n)
        torch.cuda.synchronize()
        check(lib().vsmch_particle_write_state(self._h, ctypes.c_void_p(x.data_ptr())))

    def keys(self):
        k0, kr = ctypes.c_uint32(), ctypes.c_uint32()
        w = ctypes.c_uint64()
        check(lib().vsmch_particle_keys(self._h, ctypes.byref(k0), ctypes.byref(kr),
                                        ctypes.byref(w)))
        return k0.value, kr.value, w.value

    # Weight
    def set_equal(self):
        check(lib().vsmch_weight_set_equal(self._h))

    def set_log(self, logw):
        logw = logw.to(device="cuda", dtype=_f64()).contiguous()
        check(lib().vsmch_weight_set_log(self._h, ctypes.c_void_p(logw.data_ptr())))
        self._keep = logw

    def add_log(self, inc):
        inc = inc.to(device="cuda", dtype=_f64()).contiguous()
        check(lib().vsmch_weight_add_log(self._h, ctypes.c_void_p(inc.data_ptr())))
        self._keep = inc

    def ess(self):
        v = ctypes.c_double()
        check(lib().vsmch_weight_ess(self._h, ctypes.byref(v)))
        return v.value

    def log_zconst(self):
        v = ctypes.c_double()
        check(lib().vsmch_weight_log_zconst(self._h, ctypes.byref(v)))
        return v.value

    def weights(self):
        import torch
        W = torch.empty(self.n, dtype=torch.float64, device="cuda")
        check(lib().vsmch_weight_read(self._h, ctypes.c_void_p(W.data_ptr())))
        check(lib().vsmch_synchronize())
        return W

    def resample(self, scheme, threshold, want_rep=False):
        """Particle::resample(op, threshold): returns (resampled, rep, idx)."""
        import torch
        rep = idx = None
        if want_rep:
            rep = torch.empty(self.n, dtype=torch.int32, device="cuda")
            idx = torch.full((self.n,), -1, dtype=torch.int32, device="cuda")
        r = ctypes.c_int()
        check(lib().vsmch_particle_resample(
            self._h, scheme, threshold, ctypes.byref(r),
            ctypes.c_void_p(rep.data_ptr()) if rep is not None else None,
            ctypes.c_void_p(idx.data_ptr()) if idx is not None else None))
        return bool(r.value), rep, idx

    def resample_async(self, scheme, threshold):
        """Particle::resample with the decision left on the device (no host
        synchronisation; the step can be queued back to back)."""
        check(lib().vsmch_particle_resample(self._h, scheme, threshold, None, None, None))

    def resample_count(self):
        v = ctypes.c_int64()
        check(lib().vsmch_particle_resample_count(self._h, ctypes.byref(v)))
        return v.value

    def moments(self, c0, dim):
        """Monitor::eval: sum_i W_i x_d(i) for components c0..c0+dim-1."""
        out = (ctypes.c_double * dim)()
        check(lib().vsmch_monitor_state_moments(self._h, c0, dim, out))
        return list(out)


def _f64():
    import torch
    return torch.float64


class CvFilter(Particle):
    """The constant-velocity bootstrap PF of user_guide/cpp/pf_seq.cpp on
    the device: Sampler(N, scheme, threshold) + init/move/monitor("pos")."""

    def __init__(self, n, dtype="float32", seed=0, **model):
        super().__init__(n, 4, dtype=dtype, seed=seed)
        self.model = cv_model(**model)

    def init(self, ox, oy):
        check(lib().vsmch_cv_init(self._h, ctypes.byref(self.model), ox, oy))

    def move(self, t, ox, oy):
        check(lib().vsmch_cv_move(self._h, ctypes.byref(self.model), t, ox, oy))

    def run(self, obs_x, obs_y, scheme, threshold):
        """initialize + iterate(T - 1); returns dict of per-step histories."""
        import numpy as np
        T = len(obs_x)
        ox = np.ascontiguousarray(obs_x, dtype=np.float64)
        oy = np.ascontiguousarray(obs_y, dtype=np.float64)
        ess = np.zeros(T)
        rs = np.zeros(T, dtype=np.int32)
        pos = np.zeros(2 * T)
        logz = np.zeros(T)
        P = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
        check(lib().vsmch_cv_run(self._h, ctypes.byref(self.model), T, P(ox), P(oy),
                                 scheme, threshold, P(ess), P(rs), P(pos), P(logz)))
        return {"ess": ess, "resampled": rs, "pos": pos.reshape(T, 2), "logz": logz}
