"""The C-ABI library loads and exports every symbol include/smx.h declares (CPU)."""
import ctypes
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, 'include', 'smx.h')).read()
    return sorted(set(re.findall(r'\b(smx_[a-z0-9_]+)\s*\(', txt)))


def test_library_exports_every_declared_symbol():
    from schwarmax_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared()
    assert len(names) >= 15
    for nm in names:
        assert hasattr(lib, nm), nm
    assert set(names) == set(_lib.EXPORTED)


def test_struct_layouts_match_header():
    """ctypes structs have the header's field order and sizes."""
    from schwarmax_amd import _lib
    assert ctypes.sizeof(_lib.SmxPotential) == 8 * (5 + 3 + 9 + 4 + 3 + 9 + 5 + 5 + 1)
    assert ctypes.sizeof(_lib.SmxGrids) == 4 * 8 + 8 * (3 + 3 + 2 + 2 + 9 + 1)
    assert ctypes.sizeof(_lib.SmxSchedule) == 8 + 8 * 5


def test_workspace_queries_without_gpu():
    from schwarmax_amd import _lib, model as M
    L = _lib.load_library()
    g = M.make_grids(100, (8, 4, 6), [[0, 10], [-3, 3], [-np.pi, np.pi]], [[-5, 5], [-5, 5]],
                     (20, 20), np.eye(3), True)
    P = 8 * 4 * 6
    assert L.smx_library_workspace_bytes(g, 10) == 10 * 8 * (((P + 7) // 8) * 8 + 8 * 100)
    assert L.smx_nnls_workspace_bytes(1000, 700, 20) > 4 * 1000 * 1000
    assert L.smx_nnls_workspace_bytes(0, 1, 1) == -1


def test_product_path_refuses_without_gpu():
    import pytest
    import torch
    from schwarmax_amd import _lib
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(RuntimeError, match="needs a ROCm GPU"):
        _lib.lib()


def test_noncanonical_potential_is_rejected():
    import pytest
    from schwarmax_amd import model as M
    with pytest.raises(NotImplementedError):
        M._canonical(lambda *a: 0, lambda *a: 0)
