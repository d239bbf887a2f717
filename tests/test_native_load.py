"""CPU checks of the C-ABI library: it loads, exports every function that
include/rocketcombustor.h declares, and its host-side ordering utility
returns a valid colour-major permutation."""
import ctypes
import os
import re

import numpy as np
import ctypes as C

import pytest

from rocketcombustor_amd import native, nozzle

from helpers import face_adjacency, sweep_run_levels

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "rocketcombustor.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rc_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_header():
    lib = native.lib()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(native.SIGNATURES), "ctypes signatures out of sync with the header"


def test_return_codes_match_header():
    text = open(HEADER).read()
    codes = dict((k, int(v)) for k, v in re.findall(r"\b(RC_(?:OK|ERR_[A-Z]+))\s*=\s*(\d+)", text))
    assert len(codes) >= 9
    for k, v in codes.items():
        assert getattr(native, k) == v, k


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(native, "_LIB", None)
    monkeypatch.setattr(native, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(ImportError):
        native.lib()


def test_error_without_device_is_reported():
    """Creating a case without a usable device returns an error code, never
    falls back to a CPU path."""
    case = nozzle.nozzle_case(6, 3)
    from rocketcombustor_amd.solver import SteadyCase
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    with pytest.raises(native.RcError) as ei:
        SteadyCase(case)
    assert ei.value.code == 6


def test_argument_errors_need_no_device():
    """The entry points reject missing handles and bad arguments with
    RC_ERR_ARGUMENT before touching a device (batch launch shape, network,
    learned source)."""
    L = native.lib()
    arg = native.RC_ERR_ARGUMENT
    assert L.rc_batch_set_launch(None, 256, 64, 1) == arg
    assert L.rc_set_surrogate(None, None, None, 0.0, 0.0) == arg
    out = (C.c_longlong * 2)()
    assert L.rc_get_surrogate_stats(None, out) == arg
    h = C.c_void_p()
    assert L.rc_mlp_create(0, 0, 128, 3, 3, C.byref(h)) == arg          # no inputs
    assert L.rc_mlp_create(0, 6, 128, 0, 3, C.byref(h)) == arg          # no hidden layer
    assert L.rc_mlp_train(None, 10, 64, 1e-3, 1, 1, None) == arg
    assert L.rc_mlp_forward(None, None, 1, None) == arg
    assert L.rc_mlp_num_params(None) == -1
    assert L.rc_mlp_last_error(None)


@pytest.mark.parametrize("nx,ny,block", [(24, 6, 1), (24, 6, 4), (150, 40, 1), (150, 40, 4)])
def test_multicolor_order(nx, ny, block):
    case = nozzle.nozzle_case(nx, ny)
    perm, nc = native.order_multicolor(case.mesh.cells, block=block)
    ne = case.mesh.ne
    assert sorted(perm.tolist()) == list(range(ne))
    assert 2 <= nc <= 6
    cells = case.mesh.cells[perm]
    adj = face_adjacency(cells)
    # cell wavefront levels of the new order: no two adjacent cells share a
    # level, and a colour of blocks of up to `block` cells needs at most
    # `block` levels
    level = np.zeros(ne, dtype=int)
    for i in range(ne):
        lower = [level[j] for j in adj[i] if j < i]
        level[i] = 1 + max(lower) if lower else 0
    assert level.max() + 1 <= block * nc
    for i in range(ne):
        for j in adj[i]:
            assert level[i] != level[j]
    if block == 1:
        assert level.max() + 1 <= nc
    # each block is consecutive and face-connected in the order, so the
    # library's sweep runs recover the blocks (block <= 4): one level per colour
    if block <= 4:
        first, lev = sweep_run_levels(adj, 4)
        assert lev.max() + 1 == nc
        assert len(first) <= ne // block + ne // 10


def test_ctypes_structs_match_header_layout():
    # rc_mechanism / rc_params sizes follow from the header's field list
    assert ctypes.sizeof(native.RcParams) == 8 + 4 + 4 + 8 + 8 * 8 + 16 + 8 * 8 + 16
    assert ctypes.sizeof(native.RcMechanism) % 8 == 0
