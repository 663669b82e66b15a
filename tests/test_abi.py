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


def test_device_code_has_no_calls_needing_a_dynamic_stack(tmp_path):
    """Every gfx950 kernel of the library has a static private segment: no
    recursive call (e.g. a constexpr helper left unfolded) and no dynamic
    stack, whose size the runtime cannot know (an overflow faults the GPU).
    Reads the AMDGPU metadata notes of the library's own code objects."""
    import shutil
    import subprocess
    from nitorch_amd import _lib
    bundler = "/opt/rocm/llvm/bin/clang-offload-bundler"
    readelf = "/opt/rocm/llvm/bin/llvm-readelf"
    objcopy = shutil.which("objcopy")
    if not (objcopy and os.path.exists(bundler) and os.path.exists(readelf)):
        pytest.skip("binutils / ROCm llvm tools not available")
    fat = tmp_path / "fat.bin"
    # (-O binary with an output file: objcopy never rewrites the library)
    subprocess.run([objcopy, "-O", "binary", "--only-section=.hip_fatbin", _lib.LIB_PATH, str(fat)], check=True)
    data = fat.read_bytes()
    magic = b"__CLANG_OFFLOAD_BUNDLE__"
    starts = [m.start() for m in re.finditer(re.escape(magic), data)]
    assert starts
    kernels = 0
    for i, s in enumerate(starts):
        part, co = tmp_path / f"b{i}.bin", tmp_path / f"b{i}.co"
        part.write_bytes(data[s:starts[i + 1] if i + 1 < len(starts) else len(data)])
        subprocess.run([bundler, "--unbundle", "--type=o", f"--input={part}",
                        "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", f"--output={co}"], check=True)
        notes = subprocess.run([readelf, "--notes", str(co)], check=True, capture_output=True, text=True).stdout
        kernels += notes.count(".name:")
        assert ".uses_dynamic_stack: true" not in notes, f"code object {i}: a kernel uses a dynamic stack"
    assert kernels > 100
