"""The workloads at their full sizes, a few steps each, against the CPU oracle bit for bit.

Every other parity test runs shrunk meshes. Here the exact build with ordered assembly (the CPU's
element order) starts from a smooth non-zero state, a Gaussian displacement bump (and a potential
bump in the fluid) centred in the mesh, so every kernel of a step works on non-zero values at every
point within a few element widths of the centre, and the whole state after the steps (displacement,
velocity, acceleration, potentials, all 28.6 M / 7.2 M points) must equal the FP32 oracle's:

* the 96 x 96 x 48 fluid-over-viscoelastic workload (6 fluid layers, 3-SLS solid, free surface,
  coupling, Stacey on 5 sides, its CMT source), 3 steps;
* the 48^3 elastic adjoint workload, 4 steps;
* the 64 x 64 x 32 batched workload: 8 of its 32 shots in one batch (exact, coloured, fused update);
  three of them equal their single-shot runs after 6 steps, state and seismograms;
* the fluid-over-viscoelastic builder of the 160 x 160 x 80 workload at scale 0.6 (96 x 96 x 48, 6
  fluid layers, its memory variables just under 4 GiB per shot): 8 shots in one batch; two of them
  equal their single-shot runs after 4 steps.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def bump_state(db):
    """Gaussian displacement bump (x: 1, y: -0.5, z: 0.25 times the bump; zero on fluid-only points)
    and a potential bump, float32."""
    x, y, z = (np.asarray(c, np.float64) for c in (db.xstore, db.ystore, db.zstore))
    c = [(x.min() + x.max()) / 2, (y.min() + y.max()) / 2, (z.min() + z.max()) / 2]
    s = min(x.max() - x.min(), y.max() - y.min(), z.max() - z.min()) / 8
    g = np.exp(-((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) / (2 * s * s))
    u = np.zeros((db.nglob, 3), np.float32)
    u[:, 0], u[:, 1], u[:, 2] = 1e-3 * g, -5e-4 * g, 2.5e-4 * g
    ib = db.ibool.reshape(db.nspec, 125)
    el = np.zeros(db.nglob, bool)
    el[ib[db.ispec_is_elastic.astype(bool)].ravel()] = True
    u[~el] = 0.0
    ac = np.zeros(db.nglob, bool)
    ac[ib[db.ispec_is_acoustic.astype(bool)].ravel()] = True
    pot = np.where(ac, 1e-2 * g, 0.0).astype(np.float32)
    return u.ravel(), pot


def gpu_state(db, src, rec, nstep, att, u0, p0):
    from specfem3d_amd import abi
    from specfem3d_amd.solver import GPUSolver
    g = GPUSolver(db, src, rec, nstep=nstep, attenuation=att, exact=True, assembly=1)
    z = np.zeros_like(u0)
    u = u0.copy()
    g.L.transfer_fields_el_to_device_(abi.Args().i(u.size), abi.ptr(u), abi.ptr(z), abi.ptr(z), g.MP)
    if db.nspec_acoustic:
        p, zp = p0.copy(), np.zeros_like(p0)
        g.L.transfer_fields_ac_to_device_(abi.Args().i(p.size), abi.ptr(p), abi.ptr(zp), abi.ptr(zp), g.MP)
    seis = g.run().copy()
    out = g.fields()
    if db.nspec_acoustic:
        out = out + g.potentials()
    g.close()
    return seis, out


def oracle_state(oracle_mod, db, src, rec, nstep, att, u0, p0):
    o = oracle_mod.OracleSolver(db, src, rec, "f32", attenuation=att)
    o.displ[:] = u0
    if db.nspec_acoustic:
        o.pot[:] = p0
    seis = o.run(nstep)
    out = (o.displ, o.veloc, o.accel)
    if db.nspec_acoustic:
        out = out + (o.pot, o.potd, o.potdd)
    return seis, out


def check_equal(gpu, ref, names):
    for name, a, b in zip(names, gpu, ref):
        assert np.abs(b).max() > 0, name
        assert np.array_equal(a, b), (name, rel_l2(a, b), int((a != b).sum()))


def test_sea_over_visco_full_size_bitwise(oracle_mod):
    from specfem3d_amd import workloads as W
    w = W.sea_over_visco_box(nstep=3)
    db = w.db
    u0, p0 = bump_state(db)
    print(f"[gpu] full-size {w.name}: NGLOB {db.nglob}, 3 steps from a bump, exact + ordered", flush=True)
    sg, gs = gpu_state(db, w.sources, w.receivers, 3, w.attenuation, u0, p0)
    so, os_ = oracle_state(oracle_mod, db, w.sources, w.receivers, 3, w.attenuation, u0, p0)
    check_equal(gs, os_, ("displ", "veloc", "accel", "potential", "potential_dot", "potential_dot_dot"))
    assert np.array_equal(sg, so)


def test_gradient_adjoint_box_full_size_bitwise(oracle_mod):
    from specfem3d_amd import workloads as W
    w = W.gradient_adjoint_box(nstep=4)
    db = w.db
    u0, p0 = bump_state(db)
    print(f"[gpu] full-size {w.name}: NGLOB {db.nglob}, 4 steps from a bump, exact + ordered", flush=True)
    sg, gs = gpu_state(db, w.sources, w.receivers, 4, None, u0, p0)
    so, os_ = oracle_state(oracle_mod, db, w.sources, w.receivers, 4, None, u0, p0)
    check_equal(gs, os_, ("displ", "veloc", "accel"))
    assert np.array_equal(sg, so)


def _batch_vs_single(w, shots, nstep, att, sc, assembly, check):
    from specfem3d_amd.batch import BatchSolver
    from specfem3d_amd.solver import GPUSolver
    shots = [(s, r, nstep) for s, r, _ in shots]
    b = BatchSolver(w.db, shots, exact=True, shots_per_group=sc, assembly=assembly, attenuation=att)
    b.run(graph_steps=1)
    for i in check:
        src, rec, n = shots[i]
        g = GPUSolver(w.db, src, rec, nstep=n, attenuation=att, exact=True, assembly=2, element_order="morton")
        ref = g.run().copy()
        u, _, _ = g.fields()
        p = g.potentials()[0] if w.db.nspec_acoustic else None
        g.close()
        assert np.abs(u).max() > 0, i
        assert np.array_equal(b.seismograms(i), ref), i
        assert np.array_equal(b.displ(i), u), (i, rel_l2(b.displ(i), u))
        if p is not None:
            assert np.array_equal(b.potential(i), p), (i, rel_l2(b.potential(i), p))
    b.close()


def test_gradient_shots_full_size_batch_bitwise():
    from specfem3d_amd import workloads as W
    w = W.gradient_shots_box()
    print(f"[gpu] full-size {w.name}: NGLOB {w.db.nglob}, 8 shots x 6 steps, batch vs single shots", flush=True)
    _batch_vs_single(w, w.shots[:8], 6, None, 8, 3, check=(0, 3, 7))


def test_sea_over_visco_batched_scale_06_bitwise():
    from specfem3d_amd import workloads as W
    w = W.sea_over_visco_large(scale=0.6, nshots=8, nstep=4)
    print(f"[gpu] {w.name}: NGLOB {w.db.nglob}, 8 coupled viscoelastic shots x 4 steps, batch vs single shots",
          flush=True)
    _batch_vs_single(w, w.shots, 4, w.attenuation, 4, 3, check=(0, 5))
