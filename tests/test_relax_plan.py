"""The library's Gauss-Seidel colour classes (ni_relax_colouring, host only,
no GPU needed) against the oracle's: both pick the same plan -- linear
colouring and coefficients, or per-axis period classes -- on every shape,
bound and stencil band here, so the GPU relax and FMG tests compare the same
sweep order.  The oracle's plans are pinned in test_solver_oracle.py (each
class an independent set of the operator's coupling graph)."""
import itertools

import pytest

from oracle import solver as OS

PEN = {3: (1e-4, 1e-3, 0.2, 0.05, 0.2), 2: (0.01, 0.1, 0.0, 0.1, 0.3), 0: (0.01, 0.2, 0, 0, 0), 1: (0.1, 0, 0, 0, 0)}
SHAPES = [(160, 192, 160), (80, 96, 80), (40, 48, 40), (20, 24, 20), (10, 12, 10), (5, 6, 5),
          (16, 12, 14), (8, 6, 10), (7, 5, 9), (5, 6, 4), (3, 4, 5), (1, 8, 8), (2, 2, 9), (1, 1, 1), (33, 17, 64)]


@pytest.mark.parametrize("bound", ["dft", "dct2", "dct1", "dst1", "replicate", "zero"])
def test_library_plan_matches_oracle(bound):
    from nitorch_amd import solve as SV
    try:
        SV.L.lib()
    except Exception as e:   # the library is built by __graft_entry__.build()
        pytest.skip(f"library not built: {e}")
    for shape, band in itertools.product(SHAPES, PEN):
        assert SV.relax_colouring(shape, bound, PEN[band]) == OS.relax_plan(shape, bound, band), (shape, band)


def test_mixed_bounds_plan_matches_oracle():
    from nitorch_amd import solve as SV
    try:
        SV.L.lib()
    except Exception as e:
        pytest.skip(f"library not built: {e}")
    for bounds in (["dft", "dct2", "dft"], ["dct2", "dft", "dft"], ["dft", "dft", "zero"]):
        for shape, band in itertools.product(SHAPES, PEN):
            assert SV.relax_colouring(shape, bounds, PEN[band]) == OS.relax_plan(shape, bounds, band), \
                (shape, bounds, band)
