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
        # a borrowed view (GmmSampler.particle) belongs to its owner
        if getattr(self, "_h", None) and not getattr(self, "_borrowed", False):
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

    def weight_update(self, op, src, stride=1):
        """Weight::set / mul / set_log / add_log with a stride (op one of
        WEIGHT_SET, WEIGHT_MUL, WEIGHT_SET_LOG, WEIGHT_ADD_LOG): particle i
        takes src[i * stride] of the flat device array src."""
        src = src.to(device="cuda", dtype=_f64()).contiguous()
        check(lib().vsmch_weight_update(self._h, op, ctypes.c_void_p(src.data_ptr()), stride))
        self._keep = src

    def copy(self, idx):
        """StateMatrix::copy(N, idx): particle i takes the state of idx[i]."""
        import torch
        idx = idx.to(device="cuda", dtype=torch.int32).contiguous()
        check(lib().vsmch_particle_copy(self._h, idx.numel(), ctypes.c_void_p(idx.data_ptr())))

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

    def cess(self, ll, d):
        """Conditional ESS of the increment d of the log-likelihoods ll (one
        per particle) under the current weights."""
        ll = ll.to(device="cuda", dtype=_f64()).contiguous()
        v = ctypes.c_double()
        check(lib().vsmch_particle_cess(self._h, ctypes.c_void_p(ll.data_ptr()), float(d),
                                        ctypes.byref(v)))
        return v.value

    def cess_solve(self, ll, target, dmax, rounds=10):
        """The largest increment d in (0, dmax] with cess(ll, d) >= target."""
        ll = ll.to(device="cuda", dtype=_f64()).contiguous()
        v = ctypes.c_double()
        check(lib().vsmch_particle_cess_solve(self._h, ctypes.c_void_p(ll.data_ptr()),
                                              float(target), float(dmax), int(rounds),
                                              ctypes.byref(v)))
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

    @staticmethod
    def history_buffers(k):
        import torch
        return {"ess": torch.empty(k, dtype=torch.float64, device="cuda"),
                "resampled": torch.empty(k, dtype=torch.int32, device="cuda"),
                "pos": torch.empty((k, 2), dtype=torch.float64, device="cuda"),
                "logz": torch.empty(k, dtype=torch.float64, device="cuda")}

    def iterate(self, t0, obs_x, obs_y, scheme, threshold, histories=True):
        """Steps t0 .. t0 + len(obs_x) - 1 (Sampler::initialize when t0 == 0,
        then iterate), queued without synchronising: each move gathers the
        previous event's ancestors itself. histories: True (new device
        buffers), False (none) or a dict from history_buffers(). Returns the
        device histories (ess, resampled, pos, logz) or None."""
        import numpy as np
        import torch
        k = len(obs_x)
        ox = np.ascontiguousarray(obs_x, dtype=np.float64)
        oy = np.ascontiguousarray(obs_y, dtype=np.float64)
        h = None
        ptr = [None] * 4
        if isinstance(histories, dict):
            h = histories   # preallocated device tensors of at least k records
        elif histories:
            h = self.history_buffers(k)
            torch.cuda.synchronize()
        if h is not None:
            ptr = [ctypes.c_void_p(h[n].data_ptr()) for n in ("ess", "resampled", "pos", "logz")]
        P = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
        check(lib().vsmch_cv_iterate(self._h, ctypes.byref(self.model), t0, k, P(ox), P(oy),
                                     scheme, threshold, *ptr))
        self._keep_obs = (ox, oy)
        return h

    def move_learned(self, net, t, ox, oy):
        check(lib().vsmch_cv_move_learned(self._h, ctypes.byref(self.model), net._h, t,
                                          ox, oy))

    def run_learned(self, net, obs_x, obs_y, scheme, threshold):
        """run() with the learned proposal for every move t >= 1."""
        import numpy as np
        T = len(obs_x)
        ox = np.ascontiguousarray(obs_x, dtype=np.float64)
        oy = np.ascontiguousarray(obs_y, dtype=np.float64)
        ess = np.zeros(T)
        rs = np.zeros(T, dtype=np.int32)
        pos = np.zeros(2 * T)
        logz = np.zeros(T)
        P = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
        check(lib().vsmch_cv_run_learned(self._h, ctypes.byref(self.model), net._h, T,
                                         P(ox), P(oy), scheme, threshold, P(ess), P(rs),
                                         P(pos), P(logz)))
        return {"ess": ess, "resampled": rs, "pos": pos.reshape(T, 2), "logz": logz}

    def transitions(self, obs_x, obs_y, scheme, threshold, train_seed):
        """Training transitions from this bootstrap filter's own clouds over
        the observations: (feat[(T-1) n, 8], target[(T-1) n, 4]) on the
        device."""
        import numpy as np
        import torch
        T = len(obs_x)
        ox = np.ascontiguousarray(obs_x, dtype=np.float64)
        oy = np.ascontiguousarray(obs_y, dtype=np.float64)
        rows = (T - 1) * self.n
        feat = torch.empty((rows, 8), dtype=torch.float32, device="cuda")
        target = torch.empty((rows, 4), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        P = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
        check(lib().vsmch_cv_transitions(self._h, ctypes.byref(self.model), T, P(ox), P(oy),
                                         scheme, threshold, train_seed,
                                         ctypes.c_void_p(feat.data_ptr()),
                                         ctypes.c_void_p(target.data_ptr())))
        return feat, target


class ProposalNet:
    """The learned proposal's network (6 -> 64 -> 64 -> 8, bfloat16 matrix
    cores): parameters, forward, loss gradient and Adam training, all on the
    device through the C ABI."""

    NPARAMS = 5128

    def __init__(self):
        h = ctypes.c_void_p()
        check(lib().vsmch_mlp_create(ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            check(lib().vsmch_mlp_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params=None, shift=None, scale=None):
        import numpy as np
        P = lambda a: None if a is None else np.ascontiguousarray(a, np.float32).ctypes.data_as(ctypes.c_void_p)  # noqa: E731,E501
        keep = [None if a is None else np.ascontiguousarray(a, np.float32)
                for a in (params, shift, scale)]
        check(lib().vsmch_mlp_set_params(self._h, *[P(a) for a in keep]))

    def params(self):
        import numpy as np
        out = np.zeros(self.NPARAMS, dtype=np.float32)
        check(lib().vsmch_mlp_get_params(self._h, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def forward(self, feat):
        import torch
        feat = feat.to(device="cuda", dtype=torch.float32).contiguous()
        out = torch.empty((feat.shape[0], 8), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        check(lib().vsmch_mlp_forward(self._h, feat.shape[0], ctypes.c_void_p(feat.data_ptr()),
                                      ctypes.c_void_p(out.data_ptr())))
        torch.cuda.synchronize()
        return out

    def loss_grad(self, feat, target):
        import torch
        feat = feat.to(device="cuda", dtype=torch.float32).contiguous()
        target = target.to(device="cuda", dtype=torch.float32).contiguous()
        g = torch.empty(self.NPARAMS + 1, dtype=torch.float32, device="cuda")
        loss = ctypes.c_float()
        torch.cuda.synchronize()
        check(lib().vsmch_mlp_loss_grad(self._h, feat.shape[0],
                                        ctypes.c_void_p(feat.data_ptr()),
                                        ctypes.c_void_p(target.data_ptr()),
                                        ctypes.c_void_p(g.data_ptr()), ctypes.byref(loss)))
        return loss.value, g[:self.NPARAMS].cpu().numpy()

    def init_random(self, seed=0, gain=1.0):
        """Hidden layers N(0, gain^2 / fan_in), biases and the last layer 0:
        the proposal starts as the transition (bootstrap) and, unlike an
        all-zero network, the hidden layers receive gradient."""
        import numpy as np
        rs = np.random.default_rng(seed)
        p = np.zeros(self.NPARAMS, dtype=np.float32)
        p[0:384] = rs.normal(0, gain / np.sqrt(6), 384)
        p[448:4544] = rs.normal(0, gain / np.sqrt(64), 4096)
        self.set_params(p)

    def fit_normalisation(self, feat):
        """Feature shift / scale from the training set (mean, 1 / std)."""
        f = feat[:, :6].double()
        shift = f.mean(dim=0).cpu().numpy()
        sd = f.std(dim=0).cpu().numpy()
        sd[sd <= 0] = 1.0
        self.set_params(None, shift, 1.0 / sd)

    def train(self, feat, target, batch=4096, steps=1000, lr=1e-3):
        import numpy as np
        import torch
        loss = np.zeros(max(steps, 1), dtype=np.float32)
        torch.cuda.synchronize()
        check(lib().vsmch_mlp_train(self._h, feat.shape[0], ctypes.c_void_p(feat.data_ptr()),
                                    ctypes.c_void_p(target.data_ptr()), batch, steps, lr,
                                    loss.ctypes.data_as(ctypes.c_void_p)))
        return loss[:steps]

    # Loss of the bootstrap proposal (mean 0, log-scale 0) on standardised
    # transitions: E[sum_d z_d^2 / 2] over 4 components.
    BOOTSTRAP_LOSS = 2.0

    @classmethod
    def gain(cls, losses, last=20):
        """Nats per transition the trained network gains over the bootstrap
        proposal (mean of the last training losses)."""
        import numpy as np
        return float(cls.BOOTSTRAP_LOSS - np.mean(np.asarray(losses)[-last:]))

    def select(self, losses, min_gain=0.1):
        """The proposal a filter should run with: this network, or None (the
        bootstrap proposal) when its gain over the bootstrap is below
        min_gain nats per transition -- too little to pay for the network
        (for a gain g the ESS improves by at most about e^g)."""
        return self if self.gain(losses) >= min_gain else None


class GmmSampler:
    """Tempered SMC for a c-component Gaussian mixture on the device (the
    sampler of example/gmm/include/gmm.hpp with adaptive tempering to a
    conditional ESS target); particles are 3c + 2 FP32 components."""

    def __init__(self, n, data, ncomp=4, seed=101, cess=0.9, moves=5, search_rounds=8,
                 scheme=None, threshold=0.5):
        import numpy as np
        from . import GmmConfig, STRATIFIED
        self.n, self.c = int(n), int(ncomp)
        cfg = GmmConfig(self.n, self.c, seed, cess, moves, search_rounds,
                        STRATIFIED if scheme is None else scheme, threshold)
        self.data = np.ascontiguousarray(data, dtype=np.float64)
        h = ctypes.c_void_p()
        check(lib().vsmch_gmm_create(ctypes.byref(h), ctypes.byref(cfg),
                                     self.data.ctypes.data_as(ctypes.c_void_p), len(self.data)))
        self._h = h
        ph = ctypes.c_void_p()
        check(lib().vsmch_gmm_particle(self._h, ctypes.byref(ph)))
        self.particle = Particle.__new__(Particle)
        self.particle._h, self.particle.n, self.particle.dim = ph, self.n, 3 * self.c + 2
        import torch
        self.particle.dtype = torch.float32
        self.particle._borrowed = True

    def close(self):
        if getattr(self, "_h", None):
            check(lib().vsmch_gmm_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, max_steps=200):
        import numpy as np
        T = int(max_steps)
        steps = ctypes.c_int64()
        a, e, z, g = (np.zeros(T) for _ in range(4))
        rs = np.zeros(T, dtype=np.int32)
        acc = np.zeros((T, 3))
        P = lambda x: x.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
        check(lib().vsmch_gmm_run(self._h, T, ctypes.byref(steps), P(a), P(e), P(rs), P(z),
                                  P(g), P(acc)))
        k = steps.value
        return {"steps": k, "alpha": a[:k], "ess": e[:k], "resampled": rs[:k],
                "logz": z[:k], "integrand": g[:k], "accept": acc[:k],
                "path_sampling": float(np.sum(0.5 * (g[1:k] + g[:k - 1]) * np.diff(a[:k])))}

    def move(self, kind, alpha, iteration, m):
        check(lib().vsmch_gmm_move(self._h, kind, alpha, iteration, m))

    def state(self):
        return self.particle.state()

    def weights(self):
        return self.particle.weights()


class CvBatch:
    """Many independent constant-velocity filters (FP32 state) advanced
    together on the device: filter f has ns[f] particles, horizon Ts[f] and
    seed seeds[f], and gives the results of CvFilter(ns[f], "float32",
    seed=seeds[f]).run(...) on its own observations."""

    def __init__(self, ns, Ts, seeds, scheme, threshold=0.5, **model):
        from . import BatchFilter
        self.ns = [int(n) for n in ns]
        self.Ts = [int(t) for t in Ts]
        self.nf = len(self.ns)
        arr = (BatchFilter * self.nf)(*[BatchFilter(n, t, s) for n, t, s
                                        in zip(self.ns, self.Ts, seeds)])
        self.model = cv_model(**model)
        h = ctypes.c_void_p()
        check(lib().vsmch_batch_create(ctypes.byref(h), self.nf, ctypes.cast(arr, ctypes.c_void_p),
                                       ctypes.byref(self.model), scheme, threshold))
        self._h = h
        self.tmax = max(self.Ts)

    def close(self):
        if getattr(self, "_h", None):
            check(lib().vsmch_batch_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_obs(self, tracks):
        """tracks[f] = (obs_x, obs_y), each of length >= Ts[f]."""
        import numpy as np
        obs = np.zeros((self.nf, self.tmax, 2))
        for f, (ox, oy) in enumerate(tracks):
            T = self.Ts[f]
            obs[f, :T, 0] = np.asarray(ox, dtype=np.float64)[:T]
            obs[f, :T, 1] = np.asarray(oy, dtype=np.float64)[:T]
        check(lib().vsmch_batch_set_obs(self._h, obs.ctypes.data_as(ctypes.c_void_p)))

    def run(self, net=None, graph=True):
        check(lib().vsmch_batch_run(self._h, net._h if net is not None else None, int(graph)))

    def histories(self):
        import numpy as np
        logz = np.zeros((self.nf, self.tmax))
        ess = np.zeros((self.nf, self.tmax))
        rs = np.zeros((self.nf, self.tmax), dtype=np.int32)
        P = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
        check(lib().vsmch_batch_histories(self._h, P(logz), P(ess), P(rs)))
        return {"logz": logz, "ess": ess, "resampled": rs}

    def filter_state(self, f):
        import torch
        n = self.ns[f]
        st = torch.empty((4, n), dtype=torch.float32, device="cuda")
        W = torch.empty(n, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        check(lib().vsmch_batch_read_filter(self._h, f, ctypes.c_void_p(st.data_ptr()),
                                            ctypes.c_void_p(W.data_ptr())))
        return st, W

    def info(self):
        tm, npart = ctypes.c_int64(), ctypes.c_int64()
        mb, ms = ctypes.c_int(), ctypes.c_double()
        check(lib().vsmch_batch_info(self._h, ctypes.byref(tm), ctypes.byref(npart),
                                     ctypes.byref(mb), ctypes.byref(ms)))
        return {"tmax": tm.value, "particles": npart.value, "move_blocks": mb.value,
                "run_ms": ms.value}


class ResampleIndex:
    """ResampleIndex (include/vsmc/resample/index.hpp:46-206) with the
    history on the device: row t holds the ancestor index of every particle
    at iteration t. push_back() with no argument records a step without
    resampling (the identity); push_back(idx) records the ancestors of a
    resampling step, e.g. the idx that Particle.resample(..., want_rep=True)
    returns."""

    def __init__(self, n):
        self.n = n
        h = ctypes.c_void_p()
        check(lib().vsmch_rindex_create(ctypes.byref(h), n))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().vsmch_rindex_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _ptr(idx):
        if idx is None:
            return None
        import torch
        if idx.dtype != torch.int32 or not idx.is_cuda or not idx.is_contiguous():
            raise VsmchError("ResampleIndex: idx must be a contiguous int32 CUDA tensor")
        return ctypes.c_void_p(idx.data_ptr())

    def _check_len(self, idx):
        if idx is not None and idx.numel() != self.n:
            raise VsmchError(f"ResampleIndex: idx has {idx.numel()} entries, expected {self.n}")

    def iter_size(self):
        v = ctypes.c_int64()
        check(lib().vsmch_rindex_iter_size(self._h, ctypes.byref(v)))
        return v.value

    def reset(self):
        check(lib().vsmch_rindex_reset(self._h))

    def clear(self):
        check(lib().vsmch_rindex_clear(self._h))

    def push_back(self, idx=None):
        self._check_len(idx)
        check(lib().vsmch_rindex_push_back(self._h, self._ptr(idx)))

    def insert(self, idx=None, iter=-1):
        """insert() / insert(first) / insert(iter, first): iter -1 is the
        last iteration."""
        self._check_len(idx)
        check(lib().vsmch_rindex_insert(self._h, iter, self._ptr(idx)))

    def index(self, id, iter):
        v = ctypes.c_int64()
        check(lib().vsmch_rindex_index(self._h, id, iter, ctypes.byref(v)))
        return v.value

    def index_matrix(self, layout=None):
        """index_matrix<RowMajor> (n x iter_size, row i = particle i's
        lineage) or <ColMajor> (iter_size x n) as an int32 CUDA tensor."""
        import torch
        from . import ROW_MAJOR, COL_MAJOR
        layout = ROW_MAJOR if layout is None else layout
        T = self.iter_size()
        shape = (self.n, T) if layout == ROW_MAJOR else (T, self.n)
        out = torch.empty(shape, dtype=torch.int32, device="cuda")
        if T > 0:
            check(lib().vsmch_rindex_index_matrix(self._h, layout,
                                                  ctypes.c_void_p(out.data_ptr())))
        return out
