"""C-PML layers on the GPU (elastic elements, forward runs) through the C-ABI.

* exact build + ordered assembly equals the FP32 oracle bit for bit: seismograms and the whole
  displacement field, on regular and on warped (irregular) elements, with 3-SLS attenuation in the
  elements inside the layers, and with a layer in place of the free surface;
* the FMA build with atomic and with coloured assembly stays within 1e-5 of the oracle;
* the reference example CPML_examples/homogeneous_halfspace_HEX8_elastic_absorbing_CPML_5sides (2400
  steps, irregular CUBIT mesh) against its REF_SEIS and against the maximum displacement norms its log
  prints every 100 steps;
* transfer_pml_displ_from_device_ returns the oracle's PML_displ_old, and writing it back changes
  nothing.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import rel_l2, small_pml_box, smooth_warp

pytestmark = pytest.mark.gpu


def gpu_run(db, src, rec, nstep, exact, assembly, att=None):
    from specfem3d_amd.solver import GPUSolver
    print(f"[gpu] C-PML nspec={db.nspec} ncpml={db.pml.n} nstep={nstep} exact={exact} assembly={assembly} "
          f"att={att is not None}", flush=True)
    g = GPUSolver(db, src, rec, nstep=nstep, exact=exact, assembly=assembly, attenuation=att)
    seis = g.run().copy()
    u, v, _ = g.fields()
    g.close()
    return seis, u, v


def oracle_run(oracle_mod, db, src, rec, nstep, att=None):
    o = oracle_mod.OracleSolver(db, src, rec, "f32", attenuation=att)
    seis = o.run(nstep)
    return seis, o.displ.copy(), o.veloc.copy(), o


@pytest.mark.parametrize("case", ["regular", "warped", "attenuation", "six_sides"])
def test_pml_exact_ordered_bitwise(oracle_mod, case):
    warp = smooth_warp(3000.0) if case == "warped" else None
    db, src, rec, att = small_pml_box(n=6, layers=2, warp=warp, attenuation=(case == "attenuation"),
                                      top=(case == "six_sides"))
    nstep = 250
    rs, ru, rv, _ = oracle_run(oracle_mod, db, src, rec, nstep, att)
    gs, gu, gv = gpu_run(db, src, rec, nstep, True, 1, att)
    assert np.abs(rs).max() > 0
    assert np.array_equal(gs, rs), rel_l2(gs, rs)
    assert np.array_equal(gu, ru), rel_l2(gu, ru)
    assert np.array_equal(gv, rv), rel_l2(gv, rv)


@pytest.mark.parametrize("assembly", [0, 2])
def test_pml_fast_within_tolerance(oracle_mod, assembly):
    db, src, rec, _ = small_pml_box(n=6, layers=2, warp=smooth_warp(3000.0))
    nstep = 400
    rs, ru, _, _ = oracle_run(oracle_mod, db, src, rec, nstep)
    gs, gu, _ = gpu_run(db, src, rec, nstep, False, assembly)
    e, eu = rel_l2(gs, rs), rel_l2(gu, ru)
    print(f"[gpu] C-PML fast assembly={assembly}: rel-L2 seismograms {e:.2e}, displacement {eu:.2e}", flush=True)
    assert e < 1e-5 and eu < 1e-5


def test_pml_transfer_roundtrip(oracle_mod):
    from specfem3d_amd.solver import GPUSolver
    db, src, rec, _ = small_pml_box(n=6, layers=2)
    n = 120
    _, _, _, o = oracle_run(oracle_mod, db, src, rec, n)
    g = GPUSolver(db, src, rec, nstep=2 * n, exact=True, assembly=1)
    for _ in range(n):
        g.step()
    size = C.c_int(3 * 125 * db.pml.n)
    old = np.zeros(size.value, np.float32)
    new = np.zeros(size.value, np.float32)
    g.L.transfer_pml_displ_from_device_(C.byref(size), old.ctypes.data_as(C.c_void_p),
                                        new.ctypes.data_as(C.c_void_p), g.MP)
    # the oracle stores PML_displ_old per local point (ispec_CPML, point, component)
    assert np.array_equal(old, o.pml_old)
    g.L.transfer_pml_displ_to_device_(C.byref(size), old.ctypes.data_as(C.c_void_p),
                                      new.ctypes.data_as(C.c_void_p), g.MP)
    for _ in range(n):
        g.step()
    seis = g.seismograms("d").copy()
    g.close()
    ref = oracle_mod.OracleSolver(db, src, rec, "f32").run(2 * n)
    assert np.array_equal(seis, ref)


def test_cpml5_reference_example():
    """REF_SEIS of the reference's own (NPROC=4, CPU) run of the example, and the maximum displacement
    norms its solver log prints every 100 steps (taken at the start of those steps), from the production
    build (FMA, automatic assembly)."""
    from specfem3d_amd.examples import cpml5_example
    from specfem3d_amd.solver import GPUSolver
    db, src, rec, ref, dt, z = cpml5_example()
    nstep = ref.shape[0]
    steps = [int(s) for s in z["log_norm_steps"]]
    print(f"[gpu] CPML 5-sides example: nglob={db.nglob} ncpml={db.pml.n} nstep={nstep}", flush=True)
    g = GPUSolver(db, src, rec, nstep=nstep, exact=False, assembly=-1)
    norms = {}
    for it in range(1, nstep + 1):
        if it in steps:
            norms[it] = g.max_norm()  # check_stability runs at the start of step it (iterate_time.F90:222)
        g.step()
    out = g.seismograms("d").copy()
    g.close()
    e = rel_l2(out, ref)
    cc = min(np.corrcoef(out[:, r, c], ref[:, r, c])[0, 1] for r in range(ref.shape[1]) for c in range(3))
    lg = dict(zip(steps, z["log_norms"]))
    ratio = np.array([norms[s] / lg[s] for s in steps])
    print(f"[gpu] CPML 5-sides example: rel-L2 vs REF_SEIS {e:.3e}, min correlation {cc:.6f}, "
          f"norm / log norm in [{ratio.min():.5f}, {ratio.max():.5f}]", flush=True)
    print("[gpu] norms", {s: (f"{norms[s]:.4e}", f"{lg[s]:.4e}") for s in steps}, flush=True)
    assert e < 2e-3 and cc > 0.9999
    assert np.all(np.abs(ratio - 1) < 5e-3)


def test_pml_launch_timing_scope():
    """sf_set_tuning("timing_scope", 1) times the C-PML launch alone: one event pair per step, each
    shorter than the whole elastic stiffness call that contains it (timing_scope 0)."""
    from specfem3d_amd.solver import GPUSolver
    db, src, rec, _ = small_pml_box(n=6, layers=2)
    g = GPUSolver(db, src, rec, nstep=60, exact=False, assembly=0)
    out = {}
    for scope in (0, 1):
        g.L.sf_set_tuning(g.MP, b"timing_scope", C.c_int(scope))
        g.L.sf_timing_start(g.MP, C.c_int(20))
        for _ in range(20):
            g.step()
        tot, cnt = C.c_double(0), C.c_int(0)
        g.L.sf_timing_read(g.MP, C.byref(tot), C.byref(cnt))
        out[scope] = (tot.value, cnt.value)
    g.close()
    print("[gpu] stiffness call / C-PML launch (ms, launches):", out, flush=True)
    assert out[0][1] == 20 and out[1][1] == 20
    assert 0 < out[1][0] < out[0][0]
