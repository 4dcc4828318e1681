"""The C ABI library loads on a CPU-only host and exports every entry point of include/dsh.h;
the ctypes mirror has the C struct layouts.  No compute call needs a GPU here."""
import ctypes as C
import os
import re

import numpy as np

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
