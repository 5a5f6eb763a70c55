"""CPU-side checks of the product library: it builds for gfx950, loads, exports
every symbol the header declares, and refuses to run without a GPU."""

import ctypes
import os
import re

import pytest

from .conftest import ROOT


def _declared():
    import glob
    names = set()
    for path in glob.glob(os.path.join(ROOT, "include", "*.h")):
        names |= set(re.findall(r"\b(gx_[a-z_0-9]+)\s*\(", open(path).read()))
    return sorted(names)


def test_header_declares_entry_points():
    names = _declared()
    assert "gx_build_create" in names and "gx_build_finish" in names and "gx_build_emit_part" in names
    assert "gx_mlp_train_step" in names and "gx_build_set_sizer" in names


def test_library_exports_all_declared_symbols():
    from genomix_amd import gx
    if not os.path.exists(gx.LIB_PATH):
        from genomix_amd import build
        build.build()
    L = ctypes.CDLL(gx.LIB_PATH)
    for name in _declared():
        assert hasattr(L, name), name
    assert set(_declared()) <= set(gx.EXPORTED) | {"gx_unitig_run", "gx_unitig_info", "gx_unitig_emit"}


def test_no_cpu_fallback_without_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from genomix_amd import gx
    with pytest.raises(gx.GxError) as e:
        gx.GraphBuilder(21)
    assert e.value.code in (gx.GX_EDEVICE,)


def test_create_validates_config():
    from genomix_amd import gx
    for kw in ({"k": 2}, {"k": 64}, {"k": 21, "max_readids_per_edge": -1}):
        with pytest.raises(gx.GxError) as e:
            gx.GraphBuilder(**kw)
        assert e.value.code == gx.GX_EINVAL


def test_unitig_limits():
    """The unitig pass refuses graphs its 30-bit edge entries cannot address
    and heads beyond ReadHeadInfo's 23-bit offset (the oracle's makeUUID
    throws there, as the reference's does)."""
    from genomix_amd import gx
    from oracle import oracle
    lim = gx.unitig_limits()
    assert lim == {"max_nodes": 1 << 30, "max_head_offset": (1 << 23) - 1}
    oracle.make_uuid(0, 0, 1, lim["max_head_offset"])
    with pytest.raises(oracle.OracleError):
        oracle.make_uuid(0, 0, 1, lim["max_head_offset"] + 1)
