"""The free resampling entry points called from several host threads at once,
each on its own stream (the reference's vsmc_resample_* keep no shared
state, lib/src/vsmc_resample.cpp): every thread's counts equal the ones the
same inputs give on one thread, and the calls never share scratch."""
import threading

import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu


def _inputs(k):
    rs = np.random.default_rng(100 + k)
    # sizes differ per thread so a shared workspace would be regrown under
    # another thread's queued kernels
    m = [9000, 300001, 65539, 1 << 20][k % 4]
    lw = rs.normal(0.0, [0.5, 2.0, 8.0, 2.0][k % 4], m)
    W = np.exp(lw - lw.max())
    W /= W.sum()
    u = np.sort(rs.random(m))
    return W, u


def test_threads_trans_u01_rep(dev):
    import torch
    import vsmc_amd as va
    nthreads, reps = 4, 3
    data = [_inputs(k) for k in range(nthreads)]
    expect = [ol.trans_u01_rep(W, u) for W, u in data]
    dW = [torch.as_tensor(W, device="cuda") for W, _ in data]
    du = [torch.as_tensor(u, device="cuda") for _, u in data]
    torch.cuda.synchronize()
    results = [[None] * reps for _ in range(nthreads)]
    errors = []

    def work(k):
        try:
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                va.check(va.lib().vsmch_set_stream(va.ctypes.c_void_p(s.cuda_stream)))
                for r in range(reps):
                    results[k][r] = va.resample_trans_u01_rep(dW[k], du[k]).cpu().numpy()
                va.check(va.lib().vsmch_set_stream(None))
        except Exception as e:   # reported on the main thread
            errors.append((k, repr(e)))

    print(f"threads={nthreads} reps={reps}", flush=True)
    ts = [threading.Thread(target=work, args=(k,)) for k in range(nthreads)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors
    for k in range(nthreads):
        for r in range(reps):
            got = results[k][r]
            assert (got == expect[k]).all(), (k, r, np.nonzero(got != expect[k])[0][:5])


def test_thread_errors_are_per_thread(dev):
    """A failure on one thread does not change another thread's last error."""
    import vsmc_amd as va
    L = va.lib()
    assert L.vsmch_resample_trans_u01_rep(-1, 1, None, None, None) != 0
    mine = L.vsmch_last_error()
    seen = []

    def other():
        seen.append(L.vsmch_last_error())

    t = threading.Thread(target=other)
    t.start()
    t.join()
    assert b"trans_u01_rep" in mine
    assert seen == [b""]
