"""Random-walk MH moves on the device (include/vsmch/random_walk.hpp, run
by tests/cpp/rw_hip.cpp through the Sampler's MCMC queue) against the
oracle (oracle/rw.c), particle by particle: the engine words each particle
consumed and its accept count exactly, the state, log gamma(x) and g(x) to
1e-9 (device and host exp / log / sin / cos may differ in the last ulp).
Reference: include/vsmc/rng/random_walk.hpp:49-555."""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "rw_hip")


def _run(scenario, n, iters, steps, seed):
    print(f"rw_hip scenario={scenario} N={n} iters={iters} steps={steps} seed={seed}",
          flush=True)
    r = subprocess.run([EXE, str(scenario), str(n), str(iters), str(steps), str(seed)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().split("\n")
    assert lines[-1].startswith("ACCEPT ")
    total = int(lines[-1].split()[1])
    out = np.array([[float(v) for v in ln.split()] for ln in lines[:-1]])
    assert out.shape == (n, ol.RW_OUT)
    return out, total


@pytest.mark.parametrize("scenario", range(ol.RW_NSCEN))
def test_scenario_vs_oracle(scenario):
    n, iters, steps, seed = 3000, 3, 7, 1000 + scenario
    dev, total = _run(scenario, n, iters, steps, seed)
    ref = ol.rw_run(scenario, seed, n, iters, steps)
    np.testing.assert_array_equal(dev[:, 7], ref[:, 7])   # words consumed
    np.testing.assert_array_equal(dev[:, 4], ref[:, 4])   # accepts
    assert total == int(ref[:, 4].sum())
    np.testing.assert_allclose(dev[:, :3], ref[:, :3], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(dev[:, 3], ref[:, 3], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(dev[:, 5:7], ref[:, 5:7], rtol=1e-9, atol=1e-9)


def test_long_chains_one_particle_per_lane():
    """Many steps in one call (the cached Box-Muller half carried across
    steps) and a particle count that is not a multiple of the block."""
    n, iters, steps, seed = 1001, 1, 64, 7
    for scenario in (ol.RW_GAMMA, ol.RW_MV3, ol.RW_BLOCK):
        dev, total = _run(scenario, n, iters, steps, seed)
        ref = ol.rw_run(scenario, seed, n, iters, steps)
        np.testing.assert_array_equal(dev[:, 7], ref[:, 7])
        np.testing.assert_array_equal(dev[:, 4], ref[:, 4])
        np.testing.assert_allclose(dev[:, :3], ref[:, :3], rtol=1e-9, atol=1e-9)
