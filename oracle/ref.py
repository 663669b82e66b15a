"""ORACLE — test infrastructure only.

Loads the reference's own C++ CPU backend (nitorch/_C, built from the
reference sources by `oracle/Makefile.ref` into `oracle/_ref/`) and exposes
its entry points with the argument conventions of `nitorch/_C/grid.py` and
`nitorch/_C/solve.py`.  Returns None from `load()` when it has not been built
(the GPU box only has it if it was built in-tree beforehand).
"""
import importlib.machinery
import importlib.util
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
SO_PATH = os.path.join(_HERE, "_ref", "nitorch_ref_spatial.so")
_mod = None


def available():
    return os.path.exists(SO_PATH)


def load():
    global _mod
    if _mod is None:
        if not available():
            return None
        import torch  # noqa: F401  (libtorch must be loaded first)
        loader = importlib.machinery.ExtensionFileLoader("nitorch_ref_spatial", SO_PATH)
        spec = importlib.util.spec_from_file_location("nitorch_ref_spatial", SO_PATH, loader=loader)
        mod = importlib.util.module_from_spec(spec)
        loader.exec_module(mod)
        _mod = mod
    return _mod


_NAMES = {0: "replicate", 1: "dct1", 2: "dct2", 3: "dst1", 4: "dst2", 5: "dft", 7: "zero"}


def bounds(bound):
    m = load()
    out = []
    for b in (bound if isinstance(bound, (list, tuple)) else [bound]):
        name = _NAMES[b] if isinstance(b, int) else b
        out.append(getattr(m.BoundType, name))
    return out


def orders(order):
    m = load()
    names = ["nearest", "linear", "quadratic", "cubic", "fourth", "fifth", "sixth", "seventh"]
    return [getattr(m.InterpolationType, names[int(o)]) for o in
            (order if isinstance(order, (list, tuple)) else [order])]
