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


def test_cpp_summary_and_state(dev, tmp_path):
    """Sampler::summary_header(_int) / summary_data(_int) in both layouts
    (include/vsmc/core/sampler.hpp:551-648) against the printed summary, and
    the final state's text output (StateMatrix operator<<,
    state_matrix.hpp:172-205) against the oracle's final particles."""
    ox, oy = c0_obs()
    obs = tmp_path / "pf.data"
    with open(obs, "w") as f:
        for a, b in zip(ox, oy):
            f.write(f"{float(a)!r} {float(b)!r}\n")
    N, T, seed = 3000, 40, 101
    r = subprocess.run([EXE, str(N), str(T), str(ol.SYSTEMATIC), "0.5", str(seed), str(obs),
                        "full"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.rstrip("\n").split("\n")
    iz = next(i for i, ln in enumerate(lines) if ln.startswith("LOGZ"))
    head = lines[0].rstrip("\t").split("\t")
    printed = np.array([[float(x) for x in ln.rstrip("\t").split("\t")] for ln in lines[1:iz]])
    assert printed.shape == (T, len(head))
    tagged = {}
    for ln in lines[iz + 1:]:
        if ln == "STATE":
            break
        tag, *vals = ln.split("\t")
        tagged[tag] = vals
    assert tagged["SHI"] == ["Size", "Resampled", "Accept.0"]
    assert tagged["SH"] == ["ESS", "pos.0", "pos.1"]
    assert head == tagged["SHI"] + tagged["SH"]
    di_row = np.array(tagged["SDIR"], dtype=np.int64).reshape(T, 3)
    di_col = np.array(tagged["SDIC"], dtype=np.int64).reshape(3, T)
    np.testing.assert_array_equal(di_row, di_col.T)
    np.testing.assert_array_equal(di_row, printed[:, :3].astype(np.int64))
    df_row = np.array([float(x) for x in tagged["SDR"]]).reshape(T, 3)
    df_col = np.array([float(x) for x in tagged["SDC"]]).reshape(3, T)
    np.testing.assert_array_equal(df_row, df_col.T)
    np.testing.assert_array_equal(df_row, printed[:, 3:])
    ist = lines.index("STATE")
    state = np.array([[float(x) for x in ln.split("\t")] for ln in lines[ist + 1:]])
    assert state.shape == (N, 4)
    e = ol.pf_run(N, T, seed, ol.SYSTEMATIC, 0.5, ox, oy, want_state=True)
    assert (di_row[:, 1] == e["resampled"]).all()
    np.testing.assert_allclose(state, e["state"], rtol=1e-9, atol=1e-9)
