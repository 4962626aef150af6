"""The drop-in C++ layer (include/vsmch/sampler.hpp): the reference's PF
example written with device functors (tests/cpp/pf_hip.cpp) against the
oracle's Sampler restatement -- same resampled flags, ESS, "pos" monitor and
log-evidence per step."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "cpp", "pf_hip")


def c0_obs():
    with open(os.path.join(HERE, "golden", "cv_obs_c0.json")) as f:
        g = json.load(f)
    return (np.array([float.fromhex(x) for x in g["obs_x"]]),
            np.array([float.fromhex(x) for x in g["obs_y"]]))


def run_exe(N, T, scheme, thr, seed, obs_path):
    r = subprocess.run([EXE, str(N), str(T), str(scheme), repr(thr), str(seed), obs_path],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().split("\n")
    head = lines[0].rstrip("\t").split("\t")
    rows = [ln.rstrip("\t").split("\t") for ln in lines[1:] if not ln.startswith("LOGZ")]
    logz = [float(x) for x in lines[-1].split("\t")[1:]]
    tab = {h: np.array([float(r_[i]) for r_ in rows]) for i, h in enumerate(head)}
    return tab, np.array(logz)


@pytest.mark.parametrize("scheme", [ol.STRATIFIED, ol.SYSTEMATIC, ol.RESIDUAL_STRATIFIED])
def test_cpp_sampler_vs_oracle(dev, tmp_path, scheme):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} not built (run __graft_entry__.build())")
    ox, oy = c0_obs()
    obs = tmp_path / "pf.data"
    with open(obs, "w") as f:
        for a, b in zip(ox, oy):
            # repr of a Python float round-trips exactly through operator>>
            f.write(f"{float(a)!r} {float(b)!r}\n")
    N, T, seed = 10000, 100, 101
    tab, logz = run_exe(N, T, scheme, 0.5, seed, str(obs))
    e = ol.pf_run(N, T, seed, scheme, 0.5, ox, oy)
    assert list(tab) == ["Size", "Resampled", "Accept.0", "ESS", "pos.0", "pos.1"]
    assert (tab["Size"] == N).all()
    assert (tab["Resampled"] == e["resampled"]).all()
    np.testing.assert_allclose(tab["ESS"], e["ess"], rtol=1e-9)
    np.testing.assert_allclose(np.column_stack([tab["pos.0"], tab["pos.1"]]), e["pos"],
                               rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(logz, e["logz"], rtol=1e-9, atol=1e-9)
