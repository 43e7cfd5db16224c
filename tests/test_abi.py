"""The C ABI: libnphys_hip.so loads on a machine without a GPU, exports every entry point that
include/nphys.h declares, and fails loudly (never falls back to a CPU path) without a device."""
import ctypes as C
import os
import re

import pytest

from nphysics_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "nphys.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nphys_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "nphys_step" in syms and "nphys_world_create" in syms
    assert len(syms) >= 20


def test_library_exports_every_declared_symbol(nphys):
    L = nphys.load_library()
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    assert set(nphys.EXPORTED) <= set(declared_symbols())


def test_params_struct_layout(nphys):
    L = nphys.load_library()
    p = abi.Params()
    L.nphys_params_default(C.byref(p))
    q = abi.default_params()
    for name, _ in abi.Params._fields_:
        if name == "gravity":
            assert list(p.gravity) == list(q.gravity)
        else:
            assert getattr(p, name) == pytest.approx(getattr(q, name)), name
    assert C.sizeof(abi.Params) == 8 * 10 + 4 * 8 + 8 * 3
    assert C.sizeof(abi.Counters) == 8 * 8 + 8 * 7 + 8 + 8 + 8 + 8 + 8 * 11 + 8 * 3


def test_body_desc_fields_follow_the_header():
    """abi.BodyDesc lists nphys_body_desc's members in the header's order."""
    import re

    text = open(os.path.join(os.path.dirname(__file__), "..", "include", "nphys.h")).read()
    body = text[text.index("typedef struct nphys_body_desc {"):text.index("} nphys_body_desc;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = re.findall(r"(\w+)\s*;", body)
    assert names == [f for f, _ in abi.BodyDesc._fields_]


def test_no_device_fails_loudly(nphys):
    import torch

    if torch.cuda.is_available() or os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    with pytest.raises(nphys.NphysError, match="no HIP device"):
        nphys.MechanicalWorld()


def test_invalid_arguments_rejected(nphys):
    L = nphys.load_library()
    h = C.c_void_p()
    p = abi.default_params(dt=-0.5)
    assert L.nphys_world_create(C.byref(p), C.byref(h)) == abi.NPHYS_ERR_INVALID
    assert L.nphys_world_create(None, C.byref(h)) == abi.NPHYS_ERR_INVALID
    assert L.nphys_step(None, 1) == abi.NPHYS_ERR_INVALID
    assert L.nphys_world_destroy(None) == abi.NPHYS_ERR_INVALID


def test_oracle_library_builds_and_loads(oracle_mod):
    L = oracle_mod.lib()
    for s in ("orc_world_create", "orc_step", "orc_collide", "orc_read_pairs", "orc_read_contacts",
              "orc_read_impulses", "orc_residual", "orc_colour_manifolds"):
        assert hasattr(L, s)


def test_product_does_not_link_the_oracle():
    """The shipped library and package never reference the checker."""
    import subprocess

    so = os.path.join(ROOT, "nphysics_amd", "libnphys_hip.so")
    out = subprocess.run(["nm", "-D", so], capture_output=True, text=True).stdout
    assert "orc_" not in out
    for f in os.listdir(os.path.join(ROOT, "nphysics_amd")):
        if f.endswith(".py"):
            src = open(os.path.join(ROOT, "nphysics_amd", f)).read()
            assert not re.search(r"^\s*(from|import)\s+oracle", src, flags=re.M), f
