"""The C-ABI library loads and exports every symbol include/*.h declares (no
compute calls: this runs without a GPU)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"\b(vsmch_[a-z0-9_]+)\s*\(", src):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "vsmch_resample_trans_u01_rep" in names
    assert "vsmch_resample" in names
    assert len(names) >= 30


def test_library_exports_all_symbols():
    import vsmc_amd
    if not os.path.exists(vsmc_amd.LIB_PATH):
        pytest.skip("libvsmch.so not built (run python -m vsmc_amd.build)")
    L = vsmc_amd.lib()   # loads torch's HIP runtime first, then the library
    missing = [n for n in declared_symbols() if not hasattr(L, n)]
    assert not missing, missing
    # the Python wrapper binds exactly the declared set
    assert set(vsmc_amd.SIGNATURES) == set(declared_symbols())


def test_version_without_gpu():
    import vsmc_amd
    if not os.path.exists(vsmc_amd.LIB_PATH):
        pytest.skip("libvsmch.so not built")
    assert vsmc_amd.lib().vsmch_version() == 10000


def test_single_hip_runtime():
    """Device pointers cross between torch and the library: both must use one
    HIP runtime (the library binds to torch's libamdhip64.so.7)."""
    import vsmc_amd
    if not os.path.exists(vsmc_amd.LIB_PATH):
        pytest.skip("libvsmch.so not built")
    vsmc_amd.lib()
    assert len(vsmc_amd.hip_runtimes()) == 1, vsmc_amd.hip_runtimes()


def test_workspace_pointers_are_assigned():
    """Every device buffer of the resampling workspace is assigned through
    the member it belongs to (a local that shadows a member once left one
    null): the allocation list and the release list name the same members."""
    src = open(os.path.join(ROOT, "vsmc_amd", "csrc", "resample.hip")).read()
    ens = src[src.index("bool Workspace::ensure"):src.index("bool Workspace::all_allocated")]
    alloc = set(re.findall(r"A\(&(\w+),", ens))
    rel = src[src.index("void Workspace::release()"):]
    rel = rel[:rel.index("}")]
    freed = set(re.findall(r"\b(\w+)\b", rel[rel.index("{", rel.index("ps[]")):]))
    hdr = open(os.path.join(ROOT, "vsmc_amd", "csrc", "internal.hpp")).read()
    members = set(re.findall(r"^\s+[\w:]+(?:\s+\w+)?\s*\*(\w+) = nullptr;", hdr, re.M))
    assert alloc <= members, alloc - members
    assert "groups" in alloc and alloc <= freed
