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
