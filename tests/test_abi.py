"""The C ABI library loads on a CPU-only host and exports every entry point of include/dsh.h;
the ctypes mirror has the C struct layouts.  No compute call needs a GPU here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from dustyshock3d_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "dsh.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dsh_[a-z0-9_]+)\s*\(", text)) - {"dsh_clock_fn"})


def test_every_declared_symbol_is_exported():
    lib = abi.lib()
    names = declared_functions()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts_match():
    out = (C.c_size_t * 4)()
    assert abi.lib().dsh_struct_sizes(out) == abi.OK
    assert list(out) == [C.sizeof(abi.Params), C.sizeof(abi.SoaView), C.sizeof(abi.TubeSpec),
                         C.sizeof(abi.ServerOpts)]


def test_gpu_entry_points_fail_cleanly_without_a_device():
    import torch  # noqa: F401  (only to know whether a device exists)
    if torch.cuda.is_available():
        return
    h = C.c_void_p()
    p = abi.Params()
    abi.lib().dsh_default_params(C.byref(p))
    assert abi.lib().dsh_create(C.byref(p), C.byref(h)) == abi.ERR_HIP


def test_errors_on_bad_arguments():
    lib = abi.lib()
    assert lib.dsh_tube_count(None, None) == abi.ERR_ARG
    spec = abi.TubeSpec()
    n = (C.c_int * 3)()
    assert lib.dsh_tube_count(C.byref(spec), n) == abi.ERR_ARG   # zero spacing
    assert lib.dsh_read_init(b"/nonexistent/Init.txt", C.byref(abi.Params())) == abi.ERR_IO
    assert np.isfinite(0.0)


def test_field_length_mismatch_is_rejected(oracle_mod, tmp_path):
    """Every field handed to the C ABI must hold n entries (n = len(x)); a shorter one would be
    read past its end, so the host layer raises ValueError before any call."""
    from dustyshock3d_amd import abi, configs
    from dustyshock3d_amd.sim import write_data, reference_nn, tube_params

    spec = configs.tube(0.03, 0.01, 1.0, 1.0, max_steps=2)
    ic = oracle_mod.tube_ic(spec, 0)
    n = len(ic["x"])
    ok = abi.field_arrays(ic, abi.ALL_FIELDS)
    assert set(ok) <= set(abi.ALL_FIELDS) and all(len(a) == n for a in ok.values())
    for field in ("vx", "h", "ind"):
        bad = dict(ic)
        bad[field] = ic[field][:-1]
        with pytest.raises(ValueError, match=field):
            abi.field_arrays(bad, abi.ALL_FIELDS)
        with pytest.raises(ValueError, match=field):
            oracle_mod.OracleSim(oracle_mod.tube_params(spec)).upload(0, bad)
        with pytest.raises(ValueError, match=field):
            write_data(str(tmp_path), 0, 0.0, 1e-3, 0.01, 0, bad)
    two_d = dict(ic)
    two_d["m"] = np.stack([ic["m"], ic["m"]], 1)
    with pytest.raises(ValueError, match="m"):
        abi.field_arrays(two_d, abi.ALL_FIELDS)
    with pytest.raises(ValueError):
        reference_nn(tube_params(spec), ic["x"], ic["y"][:-1], ic["z"])
    with pytest.raises(TypeError):
        abi.make_view({"x": np.zeros(3, np.float32)})
