"""The C-ABI library loads on a host without a GPU, exports every entry point the header declares,
and refuses bad configurations before touching the device (no compute calls here)."""
import ctypes
import os
import re

import pytest

from orrery_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "orrery_lbm.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:lbm|lbm2d|lbm3d|lbm_batch)_\w+)\s*\(", text)))


def test_header_and_binding_list_agree():
    assert set(declared_symbols()) == {name for name, _, _ in abi.EXPORTS} | set(
        n for n in declared_symbols() if n.startswith("lbm_batch_") or n.startswith("lbm_serve_"))


def test_library_exports_every_declared_symbol():
    lib = abi.load()
    missing = [name for name in declared_symbols() if not hasattr(lib, name)]
    assert not missing, missing


def test_validation_refuses_before_the_device():
    from cases import periodic_box
    from orrery_amd import Lbm3DSolver, LbmError, fluid_mask_3d

    c = periodic_box(8, 8, 2, 0.6)
    with pytest.raises(LbmError, match="at least 3 nodes"):
        Lbm3DSolver(c, fluid_mask_3d(8, 8, 2))
    c = periodic_box(8, 8, 8, 0.5)
    with pytest.raises(LbmError, match="tau must be finite and above 1/2"):
        Lbm3DSolver(c, fluid_mask_3d(8, 8, 8))
    c = periodic_box(8, 8, 8, 0.6)
    c.body_force_x = 1e39
    with pytest.raises(LbmError, match="outside the binary32 range"):
        Lbm3DSolver(c, fluid_mask_3d(8, 8, 8))


def test_perturbation_matches_the_oracle():
    from oracle import load
    lib = abi.load()
    orc = load()
    for seed in (0, 1, 2**63 + 5):
        for cell in (0, 1, 12345, 2**26):
            for comp in range(3):
                v = lib.lbm_perturbation(seed, cell, comp)
                assert v == orc.orc_perturbation(seed, cell, comp)
                assert -1.0 <= v < 1.0
