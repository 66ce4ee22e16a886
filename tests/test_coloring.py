"""Greedy element colouring of the coloured assembly mode (host-only entry of the C-ABI library)."""
import ctypes as C

import numpy as np

from specfem3d_amd import abi
from specfem3d_amd.mesh.box import build_box, homogeneous


def color(db, elems):
    lib = abi.load()
    ib = np.ascontiguousarray(db.ibool.reshape(db.nspec, 125).astype(np.int32))
    el = np.ascontiguousarray(np.asarray(elems, dtype=np.int32))
    col = np.zeros(el.size, np.int32)
    n = lib.sf_color_elements(C.c_int(db.nglob), C.c_int(el.size), abi.ptr(ib), abi.ptr(el), abi.ptr(col))
    return n, col


def check_proper(db, elems, col):
    """No two elements of one colour share a global point (brute force over point incidences)."""
    ib = db.ibool.reshape(db.nspec, 125)
    owner = {}
    for e, c in zip(elems, col):
        for g in np.unique(ib[e]):
            assert (int(g), int(c)) not in owner, (g, c, owner.get((int(g), int(c))), e)
            owner[(int(g), int(c))] = e


def test_structured_box_takes_eight_balanced_colours():
    db = build_box(6, 4, 4, (0, 6000, 0, 4000, -4000, 0), material=homogeneous(2300, 2800, 1500))
    elems = np.arange(db.nspec)
    n, col = color(db, elems)
    assert n == 8
    check_proper(db, elems, col)
    counts = np.bincount(col, minlength=8)
    assert counts.min() == counts.max() == db.nspec // 8


def test_subset_and_warped_mesh():
    from conftest import smooth_warp
    db = build_box(5, 3, 4, (0, 2000, 0, 2000, -2000, 0), material=homogeneous(2300, 2800, 1500),
                   warp=smooth_warp(2000.0))
    rng = np.random.default_rng(3)
    elems = rng.permutation(db.nspec)[: db.nspec // 2 + 3]
    n, col = color(db, elems)
    assert 1 <= n <= 27
    check_proper(db, elems, col)


def test_morton_launch_order_is_a_local_permutation():
    """The host's Morton launch order lists every element once and keeps face neighbours close in
    launch position: far more neighbour pairs lie within 32 launch slots than in the x-fastest
    processing order, where every z neighbour is a whole plane (64 elements) away."""
    from specfem3d_amd.solver import morton_order
    db = build_box(8, 8, 8, (0, 8000, 0, 8000, -8000, 0), material=homogeneous(2300, 2800, 1500))
    ref = np.arange(db.nspec)
    mo = morton_order(db, ref)
    assert sorted(mo.tolist()) == ref.tolist()
    pos = np.empty(db.nspec, np.int64)
    pos[mo] = np.arange(db.nspec)
    e = np.arange(db.nspec)
    ex, ey, ez = e % 8, (e // 8) % 8, e // 64
    pairs = [(e[ex < 7], e[ex < 7] + 1), (e[ey < 7], e[ey < 7] + 8), (e[ez < 7], e[ez < 7] + 64)]
    near_ref = np.mean(np.concatenate([np.abs(b - a) for a, b in pairs]) < 32)
    near_mo = np.mean(np.concatenate([np.abs(pos[b] - pos[a]) for a, b in pairs]) < 32)
    assert near_ref < 0.7 and near_mo > 0.8
    n, col = color(db, mo)
    assert n == 8
    check_proper(db, mo, col)
