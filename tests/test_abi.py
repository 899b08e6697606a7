"""CPU checks of the C ABI: the library loads, exports every entry point that
include/lpmdx.h declares, and fails loudly (no CPU fallback) without a GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lpmdx.h")


def declared():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"LPMDX_API\s+[\w\s\*]*?\b(lpmdx_\w+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared()
    assert "lpmdx_update_forces" in names and "lpmdx_velocityverlet" in names
    assert len(names) >= 15


def test_library_exports_every_declared_symbol():
    import lpmd2_amd

    lib = lpmd2_amd.lib()
    for name in declared():
        assert hasattr(lib, name), name
    assert lib.lpmdx_abi_version() == 2


def test_no_device_fails_loudly():
    import lpmd2_amd

    if lpmd2_amd.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(lpmd2_amd.LpmdxError) as e:
        lpmd2_amd.System(np.zeros((4, 3)), [10.0] * 3, 18)
    assert e.value.code == lpmd2_amd.EDEVICE
    from lpmd2_amd.batch import Batch
    from lpmd2_amd.server import Server

    with pytest.raises(lpmd2_amd.LpmdxError) as e:
        Batch(1024)
    assert e.value.code == lpmd2_amd.EDEVICE
    with pytest.raises(lpmd2_amd.LpmdxError) as e:
        Server(1 << 30)
    assert e.value.code == lpmd2_amd.EDEVICE


def test_null_arguments_are_rejected():
    import lpmd2_amd

    lib = lpmd2_amd.lib()
    assert lib.lpmdx_system_create(None, 4, None, None, None) == lpmd2_amd.EINVAL
    assert lib.lpmdx_update_forces(None, None, None, None) == lpmd2_amd.EINVAL
    assert b"NULL" in lib.lpmdx_last_error()
    from lpmd2_amd.server import _bind

    L = _bind()
    assert L.lpmdx_server_submit(None, 4, None, None, None, None, None, None, 1, None, None, None, None, 1.0, 1,
                                 None) == lpmd2_amd.EINVAL
    assert L.lpmdx_server_stats(None, None) == lpmd2_amd.EINVAL
    assert L.lpmdx_server_create(None, 1 << 30, 16) == lpmd2_amd.EINVAL


def test_lattice_generator_matches_crystal3d(golden_dir):
    import json

    import lpmd2_amd

    meta = json.load(open(os.path.join(golden_dir, "golden.json")))["crystal"]
    for key, m in meta.items():
        d = np.load(os.path.join(golden_dir, f"crystal_{key}.npz"))
        x, L = lpmd2_amd.fcc(m["nx"], m["a"])
        assert np.array_equal(x, d["x"]) and np.array_equal(L, d["L"])


def test_maxwell_boltzmann_exact_temperature():
    import lpmd2_amd
    from lpmd2_amd.lattice import temperature

    m = np.full(500, 63.546)
    v = lpmd2_amd.maxwell_boltzmann(m, 300.0, seed=4)
    assert abs(temperature(v, m) - 300.0) < 1e-9
    assert np.abs((v * m[:, None]).sum(axis=0)).max() < 1e-12
