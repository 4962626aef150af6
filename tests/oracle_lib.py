"""ctypes access to the CPU oracle (oracle/liboracle.so) and to the reference's
own C RNG headers compiled unchanged (oracle/_ref/libvsmc_rngc.so).

Test infrastructure only: the checker, never the thing measured.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")
REF = os.path.join(ROOT, "oracle", "_ref", "libvsmc_rngc.so")

P = np.ctypeslib.ndpointer
c = ctypes

MULTINOMIAL, STRATIFIED, SYSTEMATIC, RESIDUAL, RESIDUAL_STRATIFIED, RESIDUAL_SYSTEMATIC = range(6)


class Philox(ctypes.Structure):
    _fields_ = [("buf", c.c_uint32 * 4), ("ctr", c.c_uint32 * 4),
                ("key", c.c_uint32 * 2), ("index", c.c_uint32)]


class Threefry(ctypes.Structure):
    _fields_ = [("buf", c.c_uint64 * 4), ("ctr", c.c_uint64 * 4),
                ("key", c.c_uint64 * 4), ("index", c.c_uint32)]


class PFConfig(ctypes.Structure):
    _fields_ = [("N", c.c_int64), ("T", c.c_int64), ("seed", c.c_uint32),
                ("scheme", c.c_int), ("threshold", c.c_double), ("dt", c.c_double),
                ("sd_pos", c.c_double), ("sd_vel", c.c_double),
                ("sd_pos0", c.c_double), ("sd_vel0", c.c_double),
                ("obs_model", c.c_int), ("obs_scale", c.c_double),
                ("obs_df", c.c_double), ("threads", c.c_int)]


_o = None
_r = None


def ensure_built():
    if not os.path.exists(ORACLE):
        import subprocess
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True,
                       capture_output=True)


def oracle():
    global _o
    if _o is None:
        ensure_built()
        L = ctypes.CDLL(ORACLE)
        sig = {
            "vo_philox4x32_10": (None, [P(np.uint32), P(np.uint32), P(np.uint32)]),
            "vo_threefry4x64_20": (None, [P(np.uint64), P(np.uint64), P(np.uint64)]),
            "vo_philox_seed": (None, [c.POINTER(Philox), c.c_uint32]),
            "vo_philox_next": (c.c_uint32, [c.POINTER(Philox)]),
            "vo_philox_fill": (None, [c.POINTER(Philox), c.c_size_t, P(np.uint32)]),
            "vo_philox_discard": (None, [c.POINTER(Philox), c.c_uint32]),
            "vo_philox_set_ctr": (None, [c.POINTER(Philox), P(np.uint32)]),
            "vo_threefry_seed": (None, [c.POINTER(Threefry), c.c_uint64]),
            "vo_threefry_next": (c.c_uint64, [c.POINTER(Threefry)]),
            "vo_threefry_fill": (None, [c.POINTER(Threefry), c.c_size_t, P(np.uint64)]),
            "vo_threefry_discard": (None, [c.POINTER(Threefry), c.c_uint64]),
            "vo_philox_words": (None, [c.c_uint32, c.c_uint64, c.c_size_t, P(np.uint32)]),
            "vo_threefry_words": (None, [c.c_uint64, c.c_uint64, c.c_size_t, P(np.uint64)]),
            "vo_u01_co_u32d": (c.c_double, [c.c_uint32]),
            "vo_u01_co_u64d": (c.c_double, [c.c_uint64]),
            "vo_u01_co_u32f": (c.c_float, [c.c_uint32]),
            "vo_u01_fill": (None, [c.c_int, c.c_uint64, c.c_uint64, c.c_size_t, P(np.float64)]),
            "vo_normal_fill": (None, [c.c_int, c.c_uint64, c.c_uint64, c.c_size_t,
                                      c.c_double, c.c_double, P(np.float64)]),
            "vo_normal_fill_scalar": (None, [c.c_int, c.c_uint64, c.c_uint64, c.c_size_t,
                                             c.c_double, c.c_double, P(np.float64)]),
            "vo_weight_set_equal": (c.c_double, [c.c_size_t, P(np.float64)]),
            "vo_weight_set_log": (c.c_double, [c.c_size_t, P(np.float64), P(np.float64),
                                               c.POINTER(c.c_double)]),
            "vo_weight_add_log": (c.c_double, [c.c_size_t, P(np.float64), P(np.float64),
                                               c.POINTER(c.c_double)]),
            "vo_weight_ess": (c.c_double, [c.c_size_t, P(np.float64)]),
            "vo_u01_sorted": (None, [c.c_size_t, P(np.float64), P(np.float64)]),
            "vo_u01_stratified": (None, [c.c_size_t, P(np.float64), P(np.float64)]),
            "vo_u01_systematic": (None, [c.c_size_t, c.c_double, P(np.float64)]),
            "vo_trans_u01_rep": (None, [c.c_size_t, c.c_size_t, P(np.float64),
                                        P(np.float64), P(np.int64)]),
            "vo_trans_u01_index": (None, [c.c_size_t, c.c_size_t, P(np.float64),
                                          P(np.float64), P(np.int64)]),
            "vo_trans_rep_index": (None, [c.c_size_t, c.c_size_t, P(np.int64), P(np.int64)]),
            "vo_trans_index_rep": (None, [c.c_size_t, c.c_size_t, P(np.int64), P(np.int64)]),
            "vo_rindex_index": (c.c_int64, [c.c_size_t, c.c_size_t, P(np.int64), c.c_size_t,
                                            c.c_size_t]),
            "vo_rindex_matrix": (None, [c.c_size_t, c.c_size_t, P(np.int64), c.c_int,
                                        P(np.int64)]),
            "vo_rw_q": (c.c_double, [c.c_uint, c.c_double, c.c_double, c.c_double,
                                     c.c_double, c.POINTER(c.c_double)]),
            "vo_rw_flag": (c.c_uint, [c.c_double, c.c_double]),
            "vo_rw_dim": (c.c_int, [c.c_int]),
            "vo_rw_init": (None, [c.c_int, c.c_size_t, P(np.float64)]),
            "vo_rw_run": (c.c_int, [c.c_int, c.c_uint32, c.c_size_t, c.c_size_t, c.c_size_t,
                                    P(np.float64)]),
            "vo_trans_residual": (c.c_size_t, [c.c_size_t, c.c_size_t, P(np.float64),
                                               P(np.float64), P(np.int64)]),
            "vo_resample": (c.c_uint64, [c.c_int, c.c_size_t, c.c_size_t,
                                         c.POINTER(Philox), P(np.float64), P(np.int64)]),
            "vo_scheme_u01": (None, [c.c_int, c.c_size_t, c.POINTER(Philox), c.c_size_t,
                                     P(np.float64)]),
            "vo_copy_rowmajor": (None, [c.c_size_t, c.c_size_t, P(np.float64), P(np.int64)]),
            "vo_copy_colmajor": (None, [c.c_size_t, c.c_size_t, P(np.float64), P(np.int64)]),
            "vo_monitor_moments": (None, [c.c_size_t, c.c_size_t, P(np.float64),
                                          P(np.float64), P(np.float64)]),
            "vo_pf_run": (c.c_double, [c.POINTER(PFConfig), P(np.float64), P(np.float64),
                                       c.c_void_p, c.c_void_p, c.c_void_p, c.c_void_p,
                                       c.c_void_p, c.c_void_p]),
            "vo_kalman_cv": (c.c_double, [c.c_int64, c.c_double, c.c_double, c.c_double,
                                          c.c_double, c.c_double, c.c_double,
                                          P(np.float64), P(np.float64), c.c_void_p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _o = L
    return _o


def ref_available():
    return os.path.exists(REF)


def ref():
    global _r
    if _r is None:
        L = ctypes.CDLL(REF)
        L.ref_philox4x32_words.argtypes = [c.c_uint32, c.c_uint64, c.c_size_t, P(np.uint32)]
        L.ref_threefry4x64_words.argtypes = [c.c_uint64, c.c_uint64, c.c_size_t, P(np.uint64)]
        L.ref_philox4x32_block.argtypes = [P(np.uint32), P(np.uint32), P(np.uint32)]
        L.ref_threefry4x64_block.argtypes = [P(np.uint64), P(np.uint64), P(np.uint64)]
        L.ref_u01_co_u32d.argtypes = [c.c_size_t, P(np.uint32), P(np.float64)]
        L.ref_u01_co_u64d.argtypes = [c.c_size_t, P(np.uint64), P(np.float64)]
        L.ref_u01_co_u32f.argtypes = [c.c_size_t, P(np.uint32), P(np.float32)]
        _r = L
    return _r


# ---------------------------------------------------------- conveniences

def philox_words(key, offset, n):
    out = np.zeros(max(n, 1), np.uint32)
    oracle().vo_philox_words(key, offset, n, out)
    return out[:n]


def threefry_words(key, offset, n):
    out = np.zeros(max(n, 1), np.uint64)
    oracle().vo_threefry_words(key, offset, n, out)
    return out[:n]


def u01_fill(kind, key, offset, n):
    out = np.zeros(max(n, 1))
    oracle().vo_u01_fill(kind, key, offset, n, out)
    return out[:n]


def normal_fill(kind, key, offset, n, mean=0.0, sd=1.0, form=0):
    out = np.zeros(max(n, 1))
    f = oracle().vo_normal_fill if form == 0 else oracle().vo_normal_fill_scalar
    f(kind, key, offset, n, mean, sd, out)
    return out[:n]


def trans_u01_rep(W, u, M=None, N=None):
    W = np.ascontiguousarray(W, np.float64)
    u = np.ascontiguousarray(u, np.float64)
    M = len(W) if M is None else M
    N = len(u) if N is None else N
    rep = np.zeros(max(M, 1), np.int64)
    oracle().vo_trans_u01_rep(M, N, W if len(W) else np.zeros(1),
                              u if len(u) else np.zeros(1), rep)
    return rep[:M]


def trans_rep_index(rep, N=None):
    rep = np.ascontiguousarray(rep, np.int64)
    M = len(rep)
    N = M if N is None else N
    idx = np.zeros(max(N, 1), np.int64)
    oracle().vo_trans_rep_index(M, N, rep if M else np.zeros(1, np.int64), idx)
    return idx[:N]


def rindex_index(hist, id, iter):
    """ResampleIndex::index over hist[T][N] (int64)."""
    hist = np.ascontiguousarray(hist, np.int64)
    T, N = hist.shape
    return int(oracle().vo_rindex_index(N, T, hist, id, iter))


def rindex_matrix(hist, layout=101):
    """ResampleIndex::index_matrix<RowMajor (101) / ColMajor (102)>."""
    hist = np.ascontiguousarray(hist, np.int64)
    T, N = hist.shape
    out = np.zeros(max(N * T, 1), np.int64)
    oracle().vo_rindex_matrix(N, T, hist, layout, out)
    return out[:N * T].reshape((N, T) if layout == 101 else (T, N))


# random walk: scenario ids and the per-particle output record
RW_NORMAL, RW_GAMMA, RW_BETA, RW_NEG_GAMMA, RW_MV3, RW_BLOCK, RW_FLAT, \
    RW_FLAT_MV, RW_G, RW_BLOCK1 = range(10)
RW_NSCEN = 10
RW_OUT = 8   # x[3], log gamma(x), accepted, g[2], words used


def rw_q(flag, x, z, a=-np.inf, b=np.inf):
    y = c.c_double(0.0)
    q = oracle().vo_rw_q(flag, x, z, a, b, c.byref(y))
    return y.value, q


def rw_flag(a, b):
    return int(oracle().vo_rw_flag(a, b))


def rw_init(scenario, i):
    x = np.zeros(4)
    oracle().vo_rw_init(scenario, i, x)
    return x[:oracle().vo_rw_dim(scenario)]


def rw_run(scenario, seed, n, iters, steps):
    out = np.zeros((max(n, 1), RW_OUT))
    rc = oracle().vo_rw_run(scenario, seed, n, iters, steps, out)
    if rc != 0:
        raise ValueError(f"unknown random-walk scenario {scenario}")
    return out[:n]


def trans_u01_index(W, u):
    W = np.ascontiguousarray(W, np.float64)
    u = np.ascontiguousarray(u, np.float64)
    idx = np.zeros(max(len(u), 1), np.int64)
    oracle().vo_trans_u01_index(len(W), len(u), W, u, idx)
    return idx[:len(u)]


def trans_residual(W, N):
    W = np.ascontiguousarray(W, np.float64)
    resid = np.zeros(max(len(W), 1))
    integ = np.zeros(max(len(W), 1), np.int64)
    R = oracle().vo_trans_residual(len(W), N, W, resid, integ)
    return resid[:len(W)], integ[:len(W)], R


def resample(scheme, W, N, key, skip=0, M=None):
    """Resample<Scheme>()(M, N, rng, W, rep) with a Philox engine seeded with
    key that has already handed out `skip` words; returns (rep, words used)."""
    W = np.ascontiguousarray(W, np.float64)
    M = len(W) if M is None else M
    e = Philox()
    oracle().vo_philox_seed(c.byref(e), key)
    for _ in range(skip):
        oracle().vo_philox_next(c.byref(e))
    rep = np.zeros(max(M, 1), np.int64)
    used = oracle().vo_resample(scheme, M, N, c.byref(e), W if len(W) else np.zeros(1), rep)
    return rep[:M], used


def scheme_u01(scheme, N, key, nu, skip=0):
    e = Philox()
    oracle().vo_philox_seed(c.byref(e), key)
    for _ in range(skip):
        oracle().vo_philox_next(c.byref(e))
    out = np.zeros(max(nu, 1))
    oracle().vo_scheme_u01(scheme, N, c.byref(e), nu, out)
    return out[:nu]


def weight_set_log(logw):
    logw = np.ascontiguousarray(logw, np.float64)
    W = np.zeros_like(logw)
    dz = c.c_double()
    ess = oracle().vo_weight_set_log(len(logw), W, logw, c.byref(dz))
    return W, ess, dz.value


def weight_add_log(W, inc):
    W = np.ascontiguousarray(W, np.float64).copy()
    inc = np.ascontiguousarray(inc, np.float64)
    dz = c.c_double()
    ess = oracle().vo_weight_add_log(len(W), W, inc, c.byref(dz))
    return W, ess, dz.value


def pf_run(N, T, seed, scheme, threshold, obs_x, obs_y, obs_model=0, sd_pos=0.02,
           sd_vel=0.02, sd_pos0=2.0, sd_vel0=1.0, dt=0.1, obs_scale=0.1, obs_df=10.0,
           threads=0, want_state=False):
    cfg = PFConfig(N, T, seed, scheme, threshold, dt, sd_pos, sd_vel, sd_pos0, sd_vel0,
                   obs_model, obs_scale, obs_df, threads)
    ess = np.zeros(T)
    rs = np.zeros(T, np.int64)
    pos = np.zeros(2 * T)
    logz = np.zeros(T)
    state = np.zeros(4 * N) if want_state else None
    W = np.zeros(N) if want_state else None
    vp = lambda a: a.ctypes.data_as(c.c_void_p) if a is not None else None  # noqa: E731
    secs = oracle().vo_pf_run(c.byref(cfg), np.ascontiguousarray(obs_x, np.float64),
                              np.ascontiguousarray(obs_y, np.float64), vp(ess), vp(rs),
                              vp(pos), vp(logz), vp(state), vp(W))
    ph = np.zeros(5)
    L = oracle()
    L.vo_pf_phase_seconds.restype = None
    L.vo_pf_phase_seconds.argtypes = [P(np.float64)]
    L.vo_pf_phase_seconds(ph)
    out = {"ess": ess, "resampled": rs, "pos": pos.reshape(T, 2), "logz": logz,
           "seconds": secs,
           "phase_seconds": dict(zip(("move", "weights", "resample", "copy", "monitor"),
                                     ph.tolist()))}
    if want_state:
        out["state"] = state.reshape(N, 4)
        out["W"] = W
    return out


def kalman_cv(obs_x, obs_y, dt=0.1, sd_pos=0.02, sd_vel=0.02, sd_pos0=2.0, sd_vel0=1.0,
              sd_obs=0.01):
    T = len(obs_x)
    cum = np.zeros(T)
    ll = oracle().vo_kalman_cv(T, dt, sd_pos, sd_vel, sd_pos0, sd_vel0, sd_obs,
                               np.ascontiguousarray(obs_x, np.float64),
                               np.ascontiguousarray(obs_y, np.float64),
                               cum.ctypes.data_as(c.c_void_p))
    return ll, cum


def cumsum(w):
    w = np.ascontiguousarray(w, np.float64)
    L = oracle()
    L.vo_cumsum.restype = None
    L.vo_cumsum.argtypes = [c.c_size_t, P(np.float64), P(np.float64)]
    s = np.zeros(max(len(w), 1))
    L.vo_cumsum(len(w), w if len(w) else np.zeros(1), s)
    return s[:len(w)]


# ------------------------------------------------------ learned proposal

MLP_NPARAMS = 5128
MLP_OFF = {"W1": (0, (64, 6)), "b1": (384, (64,)), "W2": (448, (64, 64)),
           "b2": (4544, (64,)), "W3": (4608, (8, 64)), "b3": (5120, (8,))}


def _mlp_sig():
    L = oracle()
    if getattr(L, "_mlp_ready", False):
        return L
    L.vo_mlp_nparams.restype = c.c_int
    L.vo_bf16.restype = c.c_double
    L.vo_bf16.argtypes = [c.c_double]
    L.vo_cv_features.restype = None
    L.vo_cv_features.argtypes = [c.c_double] * 7 + [P(np.float64)]
    L.vo_mlp_forward.restype = None
    L.vo_mlp_forward.argtypes = [P(np.float32), P(np.float64), P(np.float64), c.c_int64,
                                 P(np.float64), c.c_int, P(np.float64)]
    L.vo_mlp_loss_grad.restype = c.c_double
    L.vo_mlp_loss_grad.argtypes = [P(np.float32), P(np.float64), P(np.float64), c.c_int64,
                                   P(np.float64), P(np.float64), c.c_int, P(np.float64)]
    L.vo_cv_learned_particle.restype = c.c_double
    L.vo_cv_learned_particle.argtypes = [P(np.float64), c.c_double, c.c_double,
                                         P(np.float64), P(np.float64), c.c_double,
                                         c.c_double, c.c_double, c.c_int, c.c_double,
                                         c.c_double, P(np.float64)]
    L.vo_cv_loglik.restype = c.c_double
    L.vo_cv_loglik.argtypes = [c.c_double] * 4 + [c.c_int, c.c_double, c.c_double]
    L._mlp_ready = True
    return L


def bf16(x):
    return _mlp_sig().vo_bf16(float(x))


def cv_features(x, y, vx, vy, ox, oy, dt=0.1):
    f = np.zeros(6)
    _mlp_sig().vo_cv_features(x, y, vx, vy, ox, oy, dt, f)
    return f


def mlp_forward(params, shift, scale, feat, round_bf16=True):
    feat = np.ascontiguousarray(feat, np.float64).reshape(-1, 8)
    out = np.zeros((len(feat), 8))
    _mlp_sig().vo_mlp_forward(np.ascontiguousarray(params, np.float32),
                              np.ascontiguousarray(shift, np.float64),
                              np.ascontiguousarray(scale, np.float64), len(feat), feat,
                              int(round_bf16), out)
    return out


def mlp_loss_grad(params, shift, scale, feat, target, round_bf16=True):
    feat = np.ascontiguousarray(feat, np.float64).reshape(-1, 8)
    target = np.ascontiguousarray(target, np.float64).reshape(-1, 4)
    g = np.zeros(MLP_NPARAMS)
    loss = _mlp_sig().vo_mlp_loss_grad(np.ascontiguousarray(params, np.float32),
                                       np.ascontiguousarray(shift, np.float64),
                                       np.ascontiguousarray(scale, np.float64), len(feat),
                                       feat, target, int(round_bf16), g)
    return loss, g


def cv_learned_particle(s0, ox, oy, o, z, dt=0.1, sd_pos=0.02, sd_vel=0.02, obs_model=0,
                        obs_scale=0.1, obs_df=10.0):
    s1 = np.zeros(4)
    inc = _mlp_sig().vo_cv_learned_particle(np.ascontiguousarray(s0, np.float64), ox, oy,
                                            np.ascontiguousarray(o, np.float64),
                                            np.ascontiguousarray(z, np.float64), dt, sd_pos,
                                            sd_vel, obs_model, obs_scale, obs_df, s1)
    return s1, inc


def cv_loglik(x, y, ox, oy, obs_model=0, obs_scale=0.1, obs_df=10.0):
    return _mlp_sig().vo_cv_loglik(x, y, ox, oy, obs_model, obs_scale, obs_df)


# ------------------------------------------------- tempered mixture SMC

def _gmm_sig():
    L = oracle()
    if getattr(L, "_gmm_ready", False):
        return L
    d = P(np.float64)
    L.vo_gmm_hyper.restype = None
    L.vo_gmm_hyper.argtypes = [d, c.c_int64, d]
    L.vo_gmm_log_prior.restype = c.c_double
    L.vo_gmm_log_prior.argtypes = [c.c_int, d, d, d]
    L.vo_gmm_log_lik.restype = c.c_double
    L.vo_gmm_log_lik.argtypes = [c.c_int, d, d, d, d, c.c_int64]
    L.vo_gmm_scales.restype = None
    L.vo_gmm_scales.argtypes = [c.c_double, d]
    L.vo_gmm_move.restype = c.c_int
    L.vo_gmm_move.argtypes = [c.c_int, c.c_int, d, d, c.c_double, c.c_double, d, d,
                              c.c_int64]
    L.vo_cess.restype = c.c_double
    L.vo_cess.argtypes = [c.c_int64, d, d, c.c_double]
    L.vo_cess_solve.restype = c.c_double
    L.vo_cess_solve.argtypes = [c.c_int64, d, d, c.c_double, c.c_double, c.c_double]
    L._gmm_ready = True
    return L


def _a(x):
    return np.ascontiguousarray(x, np.float64)


def gmm_hyper(x):
    h = np.zeros(4)
    _gmm_sig().vo_gmm_hyper(_a(x), len(x), h)
    return h


def gmm_log_prior(mu, lam, h):
    return _gmm_sig().vo_gmm_log_prior(len(mu), _a(mu), _a(lam), _a(h))


def gmm_log_lik(mu, lam, w, x):
    return _gmm_sig().vo_gmm_log_lik(len(mu), _a(mu), _a(lam), _a(w), _a(x), len(x))


def gmm_scales(alpha):
    sd = np.zeros(3)
    _gmm_sig().vo_gmm_scales(alpha, sd)
    return sd


def gmm_move(kind, p, z, log_u, alpha, h, x):
    """p = (mu, lambda, weight, log prior, log likelihood); returns
    (accepted, new p)."""
    q = _a(p).copy()
    c_ = (len(q) - 2) // 3
    acc = _gmm_sig().vo_gmm_move(kind, c_, q, _a(z), log_u, alpha, _a(h), _a(x), len(x))
    return acc, q


def cess(W, ll, d):
    return _gmm_sig().vo_cess(len(W), _a(W), _a(ll), d)


def cess_solve(W, ll, dmax, target, tol=1e-12):
    return _gmm_sig().vo_cess_solve(len(W), _a(W), _a(ll), dmax, target, tol)
