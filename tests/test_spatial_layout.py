"""CPU checks of the host-side spatial glue: how caller shapes fold into the
backend's (B, C, *spatial) / (B, *spatial, D) layout and back
(nitorch/spatial/_grid.py:90-143), the resize lattice maps of each anchor
(nitorch/spatial/_grid.py:618-754, _affine.py:1712-1790) and the alignment
names (nitorch/_C/resize.py:16-53).  No kernel is called."""

import numpy as np
import pytest
import torch

from nitorch_amd import spatial as SP
from nitorch_amd.affine import LatticeAxis, affine_resize, affine_matvec
from nitorch_amd.resize import GridAlignType, align_to_nitorch

# (grid shape, image shape, push?) -> (backend grid, backend image, caller output of a pull/push)
# Written out by hand from the documented contract: leading dimensions of
# grid and image broadcast; the output keeps the image's channel dimension
# when it has one, gets a singleton channel when there are leading
# dimensions, and none otherwise.
CASES = [
    ((5, 6, 7, 3), (5, 6, 7), False, (1, 5, 6, 7, 3), (1, 1, 5, 6, 7), (5, 6, 7)),
    ((5, 6, 7, 3), (4, 5, 6, 7), False, (1, 5, 6, 7, 3), (1, 4, 5, 6, 7), (4, 5, 6, 7)),
    ((2, 5, 6, 7, 3), (5, 6, 7), False, (2, 5, 6, 7, 3), (2, 1, 5, 6, 7), (2, 1, 5, 6, 7)),
    ((2, 5, 6, 7, 3), (4, 8, 9, 10), False, (2, 5, 6, 7, 3), (2, 4, 8, 9, 10), (2, 4, 5, 6, 7)),
    ((2, 5, 6, 7, 3), (2, 4, 8, 9, 10), False, (2, 5, 6, 7, 3), (2, 4, 8, 9, 10), (2, 4, 5, 6, 7)),
    ((3, 1, 5, 6, 7, 3), (2, 4, 8, 9, 10), False, (6, 5, 6, 7, 3), (6, 4, 8, 9, 10), (3, 2, 4, 5, 6, 7)),
    ((1, 5, 6, 7, 3), (3, 2, 8, 9, 10), False, (3, 5, 6, 7, 3), (3, 2, 8, 9, 10), (3, 2, 5, 6, 7)),
    ((5, 6, 2), (2, 5, 6), False, (1, 5, 6, 2), (1, 2, 5, 6), (2, 5, 6)),
    ((7, 1), (4, 9), False, (1, 7, 1), (1, 4, 9), (4, 7)),
    # push: the spatial shapes broadcast too
    ((5, 1, 7, 3), (4, 5, 6, 1), True, (1, 5, 6, 7, 3), (1, 4, 5, 6, 7), (4, 5, 6, 7)),
    ((2, 5, 6, 7, 3), (5, 6, 7), True, (2, 5, 6, 7, 3), (2, 1, 5, 6, 7), (2, 1, 5, 6, 7)),
]


@pytest.mark.parametrize("gshape,ishape,push,gb,ib,out", CASES)
def test_fold_and_restore_shapes(gshape, ishape, push, gb, ib, out):
    grid = torch.zeros(gshape)
    image = torch.zeros(ishape)
    g, x, layout = SP._fold_pair(grid, image, shared_space=push)
    assert tuple(g.shape) == gb and tuple(x.shape) == ib
    # a backend output on the grid's lattice (pull) or the image's (push)
    space = tuple(x.shape[2:]) if push else tuple(g.shape[1:-1])
    res = torch.zeros((x.shape[0], x.shape[1]) + space)
    assert tuple(layout.restore(res).shape) == out
    # the gradient layout appends the vector dimension
    resg = torch.zeros((x.shape[0], x.shape[1]) + space + (len(space),))
    assert tuple(layout.restore(resg, vector=True).shape) == out + (len(space),)


def test_fold_keeps_values_in_order():
    """Folding broadcast leading dimensions keeps (batch, channel) pairs in
    row-major order, so the restored result of a per-batch operation lines
    up with the caller's indices."""
    grid = torch.arange(3 * 2 * 4 * 2, dtype=torch.float64).reshape(3, 1, 2, 4, 2)
    image = torch.arange(2 * 5 * 2 * 4, dtype=torch.float64).reshape(2, 5, 2, 4)
    g, x, layout = SP._fold_pair(grid, image)
    assert g.shape == (6, 2, 4, 2) and x.shape == (6, 5, 2, 4)
    for a in range(3):
        for b in range(2):
            assert torch.equal(g[a * 2 + b], grid[a, 0])
            assert torch.equal(x[a * 2 + b], image[b])
    back = layout.restore(x * 2)
    assert back.shape == (3, 2, 5, 2, 4)
    assert torch.equal(back[2, 1], image[1] * 2)


def test_fold_grid_alone():
    g, layout = SP._fold_grid(torch.zeros(2, 3, 4, 5, 6, 3))
    assert g.shape == (6, 4, 5, 6, 3)
    assert layout.restore(torch.zeros(6, 1, 7, 8, 9)).shape == (2, 3, 1, 7, 8, 9)
    g, layout = SP._fold_grid(torch.zeros(4, 5, 6, 3))
    assert layout.restore(torch.zeros(1, 1, 4, 5, 6)).shape == (4, 5, 6)


@pytest.mark.parametrize("n_in,n_out,f", [(12, 18, 1.5), (10, 8, 0.8), (9, 18, 2.0), (7, 3, 0.45), (5, 5, 1.0)])
def test_lattice_axis_anchors(n_in, n_out, f):
    """Each anchor pins its defining points: 'c' the first and last centres,
    'e' the outer edges (-0.5 and n - 0.5), 'f' the first centre and 'l' the
    last one, with spacing 1/f for the last two; coordinates() is
    scale * o + shift."""
    o = np.arange(n_out, dtype=np.float64)
    for anchor in "cefl":
        ax = LatticeAxis(anchor, f, n_in, n_out)
        c = ax.coordinates(dtype=torch.float64).numpy()
        assert np.allclose(c, ax.scale * o + ax.shift, atol=1e-12)
        if anchor == "c":
            assert abs(c[0]) < 1e-12 and abs(c[-1] - (n_in - 1)) < 1e-12
        elif anchor == "e":
            assert abs(ax.scale * -0.5 + ax.shift + 0.5) < 1e-12
            assert abs(ax.scale * (n_out - 0.5) + ax.shift - (n_in - 0.5)) < 1e-12
        elif anchor == "f":
            assert c[0] == 0 and abs(ax.scale - 1 / f) < 1e-15
        else:
            assert abs(c[-1] - (n_in - 1)) < 1e-12 and abs(ax.scale - 1 / f) < 1e-15


def test_lattice_axis_singletons_and_errors():
    # a length-1 axis on either side is edge-anchored
    ax = LatticeAxis("c", 0.25, 4, 1)
    assert ax.anchor == "e" and ax.scale == 4.0 and ax.shift == 1.5
    ax = LatticeAxis("f", 3.0, 1, 3)
    assert ax.anchor == "e" and abs(ax.scale - 1 / 3) < 1e-15
    # both of length 1: scale 1 / factor, no shift
    ax = LatticeAxis("l", 2.0, 1, 1)
    assert ax.scale == 0.5 and ax.shift == 0.0
    with pytest.raises(ValueError):
        LatticeAxis("x", 1.0, 4, 4)


def test_affine_resize_maps_the_same_world_points():
    """World position of output voxel o under the resized affine equals the
    world position of input coordinate scale * o + shift under the original."""
    rng = np.random.default_rng(0)
    aff = torch.eye(4, dtype=torch.float64)
    aff[:3, :] = torch.from_numpy(rng.standard_normal((3, 4)))
    shape, factor = (12, 10, 9), (1.5, 0.8, 2.0)
    for anchor in ("c", "e", "f", "l", ["c", "e", "l"]):
        new, shape_out = affine_resize(aff, shape, factor, anchor)
        assert shape_out == tuple(max(1, int(round(n * f))) for n, f in zip(shape, factor))
        anchors = anchor if isinstance(anchor, list) else [anchor] * 3
        axes = [LatticeAxis(a, f, n, m) for a, f, n, m in zip(anchors, factor, shape, shape_out)]
        o = torch.from_numpy(rng.uniform(0, 8, (5, 3)))
        src = torch.stack([o[:, d] * axes[d].scale + axes[d].shift for d in range(3)], -1)
        assert torch.allclose(affine_matvec(new, o), affine_matvec(aff, src), atol=1e-12)


def test_resize_lattice_shapes():
    x = torch.zeros(2, 3, 12, 10, 9)
    n_in, n_out, factor, anchor, axes = SP._resize_lattice(x, [1.5, 0.8, 2.0], None, "c", None)
    assert n_out == [18, 8, 18] and factor == [1.5, 0.8, 2.0]
    n_in, n_out, factor, anchor, axes = SP._resize_lattice(x, None, (6, 5, 3), "e", None)
    assert n_out == [6, 5, 3] and factor == [0.5, 0.5, 3 / 9]
    n_in, n_out, factor, anchor, axes = SP._resize_lattice(x, 0.01, None, "c", None)
    assert n_out == [1, 1, 1]
    # an axis with neither a factor nor a length collapses to length 1
    n_in, n_out, factor, anchor, axes = SP._resize_lattice(x, None, [4, None, 3], "c", None)
    assert n_out == [4, 1, 3]


def test_align_names():
    assert align_to_nitorch("centers") == "center"
    assert align_to_nitorch(["e", "First", "LAST"]) == ["edge", "first", "last"]
    assert align_to_nitorch(("c", "e"), "int") == (GridAlignType.center, GridAlignType.edge)
    assert align_to_nitorch(GridAlignType.last) == "last"
    assert int(align_to_nitorch("f", int)) == 2
    with pytest.raises(ValueError):
        align_to_nitorch("middle")
