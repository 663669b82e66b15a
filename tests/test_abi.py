"""CPU checks of the C ABI: the library loads, exports every entry point
declared in include/nitorch_amd.h, and the Python boundary refuses CPU
tensors (no silent CPU fallback)."""
import ctypes
import os
import re

import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER_DIR = os.path.join(ROOT, "include")


def declared_symbols():
    names = []
    for f in sorted(os.listdir(HEADER_DIR)):
        if f.endswith(".h"):
            text = open(os.path.join(HEADER_DIR, f)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names += re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(ni_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_symbols()
    for n in ("ni_grid_pull", "ni_grid_push", "ni_grid_count", "ni_grid_grad",
              "ni_grid_pull_backward", "ni_grid_push_backward", "ni_grid_count_backward",
              "ni_grid_grad_backward", "ni_last_error", "ni_pushpull"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from nitorch_amd import _lib
    assert os.path.exists(_lib.LIB_PATH), "build with python -m nitorch_amd.build"
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_argument_errors_are_reported():
    from nitorch_amd import _lib as L
    lib = L.lib()
    # invalid order: reported before any device work
    b = L.i32arr([2, 2, 2])
    o = L.i32arr([9, 9, 9])
    rc = lib.ni_grid_pull(0, 3, b, o, 0, 1, 1, L.i64arr([2, 2, 2]), L.i64arr([2, 2, 2]),
                          None, L.i64arr([8, 8, 4, 2, 1]), ctypes.c_void_p(1), L.i64arr([24, 12, 6, 3, 1]),
                          None, L.i64arr([8, 8, 4, 2, 1]), None)
    assert rc != 0
    assert "order" in L.last_error()


def test_cpu_tensors_raise():
    from nitorch_amd import spatial
    x = torch.zeros(1, 1, 4, 4, 4)
    g = torch.zeros(1, 4, 4, 4, 3)
    with pytest.raises(RuntimeError, match="ROCm GPU"):
        spatial.grid_pull(x, g)


def test_name_normalisation():
    from nitorch_amd.grid import bound_to_nitorch, inter_to_nitorch, BoundType, InterpolationType
    assert bound_to_nitorch(["reflect", "wrap", "nearest", "zeros"]) == ["dct2", "dft", "replicate", "zero"]
    assert bound_to_nitorch("dirichlet", "enum") == BoundType.dst2
    assert bound_to_nitorch(["mirror", "antimirror"], "int") == [1, 3]
    assert inter_to_nitorch(["cubic", 5, "nearest"], "int") == [3, 5, 0]
    assert inter_to_nitorch(3, "enum") == InterpolationType.cubic
    with pytest.raises(ValueError):
        bound_to_nitorch("sliding")
