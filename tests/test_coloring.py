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


def test_brick_launch_order_interleaves_few_sequential_rows():
    """Slab ("brick") launch order (z // bz, y // by, x, y % by, z % bz): a permutation in which every
    y and z face neighbour within a slab is at most by * bz * nx launches away, and the element ids
    visited form by * bz increasing sequences (sequential streams of the per-element arrays)."""
    from specfem3d_amd.solver import brick_order
    nx = 6
    db = build_box(nx, 4, 4, (0, 6000, 0, 4000, -4000, 0), material=homogeneous(2300, 2800, 1500))
    ref = np.arange(db.nspec)
    for by, bz in ((1, 2), (2, 2), (2, 4)):
        o = brick_order(db, ref, by, bz)
        assert sorted(o.tolist()) == ref.tolist()
        ex, ey, ez = o % nx, (o // nx) % 4, o // (nx * 4)
        streams = {}
        for e, y, z in zip(o, ey, ez):
            streams.setdefault((y % by, z % bz), []).append(e)
        assert len(streams) == by * bz
        assert all(np.all(np.diff(v) > 0) for v in streams.values())
        pos = np.empty(db.nspec, np.int64)
        pos[o] = np.arange(db.nspec)
        e = ref
        x, y, z = e % nx, (e // nx) % 4, e // (nx * 4)
        inner_z = (z % bz) < bz - 1
        assert np.all(np.abs(pos[e[inner_z] + nx * 4] - pos[e[inner_z]]) <= by * bz)


def point_flags(db, elems, col, forced):
    lib = abi.load()
    ib = np.ascontiguousarray(db.ibool.reshape(db.nspec, 125).astype(np.int32))
    order = np.argsort(col, kind="stable")
    el = np.ascontiguousarray(np.asarray(elems, np.int32)[order])
    cl = np.ascontiguousarray(np.asarray(col, np.int32)[order])
    fl = np.zeros(db.nspec * 125, np.uint8)
    rc = lib.sf_batch_point_flags(C.c_int(db.nglob), C.c_int(el.size), abi.ptr(ib), abi.ptr(el), abi.ptr(cl),
                                  abi.ptr(np.ascontiguousarray(forced, np.uint8)), abi.ptr(fl))
    return rc, el, cl, fl.reshape(db.nspec, 125)


def fused_assembly(db, el, cl, fl, forced, ext, contrib):
    """The fused batched mode's accumulation on the CPU: `a` starts as garbage (NaN) except where the
    Stacey/source pass added `ext`; per colour, first writers store, others add; the finisher reads
    the total. Returns the totals seen by the finishers and how many finishers each point had."""
    ib = db.ibool.reshape(db.nspec, 125)
    a = np.full(db.nglob, np.nan, np.float32)
    a[forced] = np.float32(0) + ext[forced]
    total = np.full(db.nglob, np.nan, np.float32)
    nfin = np.zeros(db.nglob, int)
    for c in np.unique(cl):
        for e in el[cl == c]:
            for p in range(125):
                g, f = ib[e, p], fl[e, p]
                a[g] = (np.float32(0) if f & 1 else a[g]) + contrib[e, p]
                if f & 2:
                    total[g] = a[g]
                    nfin[g] += 1
                    if f & 4:
                        a[g] = 0
    return total, nfin


def test_fused_mode_point_flags_reproduce_the_plain_sum():
    """Flags of the fused batched assembly (sf_batch_point_flags): with `a` never cleared except at
    points Stacey/sources add to, first-writer/finisher flags must give every point exactly one
    finisher and the same float32 total as zero-initialised accumulation in colour order."""
    db = build_box(5, 4, 3, (0, 5000, 0, 4000, -3000, 0), material=homogeneous(2300, 2800, 1500))
    rng = np.random.default_rng(7)
    elems = rng.permutation(db.nspec).astype(np.int32)
    n, col = color(db, elems)
    forced = np.zeros(db.nglob, bool)
    forced[np.unique(db.ibool[db.abs_faces.ispec].ravel())[::3]] = True     # some Stacey points
    forced[np.unique(db.ibool[elems[5]])] = True                               # a source element
    rc, el, cl, fl = point_flags(db, elems, col, forced)
    assert rc == 0
    ext = rng.normal(size=db.nglob).astype(np.float32)
    contrib = rng.normal(size=(db.nspec, 125)).astype(np.float32)
    total, nfin = fused_assembly(db, el, cl, fl, forced, ext, contrib)
    ref = np.where(forced, ext, np.float32(0)).astype(np.float32)
    ib = db.ibool.reshape(db.nspec, 125)
    for c in np.unique(cl):
        for e in el[cl == c]:
            for p in range(125):
                ref[ib[e, p]] = ref[ib[e, p]] + contrib[e, p]
    assert (nfin == 1).all()
    assert np.array_equal(total, ref)
    # test of the test: a first-writer flag on a forced point drops its Stacey/source term
    bad = fl.copy()
    e, p = next((e, p) for e in el for p in range(125) if forced[ib[e, p]] and not fl[e, p] & 1)
    bad[ib == ib[e, p]] &= ~np.uint8(1)
    bad[e, p] |= 1
    t2, _ = fused_assembly(db, el, cl, bad, forced, ext, contrib)
    assert not np.array_equal(t2, ref)


def test_point_flags_refuse_a_colour_that_shares_points():
    db = build_box(2, 1, 1, (0, 2000, 0, 1000, -1000, 0), material=homogeneous(2300, 2800, 1500))
    rc, *_ = point_flags(db, np.arange(2), np.zeros(2, np.int32), np.zeros(db.nglob, bool))
    assert rc == -1
