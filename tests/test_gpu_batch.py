"""Batched many-shot engine (BatchSolver / sf_batch_*) against single-shot runs and the oracle.

* exact kernels + coloured assembly: every shot of the batch is bit-identical to its single-shot
  GPUSolver run with the same options (same arithmetic, same colour order), also with the Newmark
  update fused into the coloured stiffness launches (assembly 3);
* hipGraph replay: steps replayed from a captured group of 10 steps equal eager steps, for every
  step of every shot, with eager steps and host synchronisation between the replays;
* fast kernels: each shot within 1e-5 rel-L2 of the FP32 oracle, for every autotuner candidate;
* uneven record lengths: a shot stops at its own length (its final displacement equals the
  single-shot run's);
* 3-SLS viscoelastic batches: every shot bit-identical to its single-shot viscoelastic run;
* compaction: once at least 8 shots have finished, the live shots get narrower rows; with 12
  shots (rows of 16) every shot still equals its single-shot run, elastic and viscoelastic,
  eager and graph-replayed, and finished shots keep their final displacement;
* fluid over solid (the 160x160x80 coupled workload's builder at scale 0.1: one acoustic layer,
  free surface, fluid and solid Stacey faces, coupling): with 10 shots of uneven lengths every
  shot's seismograms, final displacement and final potential equal its single-shot run, elastic
  and 3-SLS viscoelastic, separate and fused update, eager and graph-replayed, and through
  compaction; fast kernels with atomic assembly stay within 1e-5 of the FP32 oracle.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_batch():
    from specfem3d_amd import workloads as W
    return W.gradient_shots_box(scale=0.1, nshots=5, nrec=5, nstep_range=(60, 200), seed=11)


def single_shot(w, i, exact, assembly):
    from specfem3d_amd.solver import GPUSolver
    src, rec, n = w.shots[i]
    g = GPUSolver(w.db, src, rec, nstep=n, exact=exact, assembly=assembly, element_order="morton")
    s = g.run().copy()
    u, _, _ = g.fields()
    g.close()
    return s, u


@pytest.fixture(scope="module")
def single_runs(small_batch):
    return [single_shot(small_batch, i, True, 2) for i in range(len(small_batch.shots))]


@pytest.mark.parametrize("assembly,sc,overlay", [(2, 4, 0), (3, 4, 0), (3, 8, 0), (3, 4, 1), (3, 8, 1), (2, 8, 1)])
def test_batch_shots_bitwise_equal_single_shot_runs(small_batch, single_runs, assembly, sc, overlay):
    from specfem3d_amd.batch import BatchSolver
    w = small_batch
    print(f"[gpu] batch of {len(w.shots)} shots, nstep {[s[2] for s in w.shots]}, assembly {assembly}, "
          f"shots per group {sc}, LDS overlay {overlay}", flush=True)
    b = BatchSolver(w.db, w.shots, exact=True, shots_per_group=sc, assembly=assembly)
    b.set_tuning("lds_overlay", overlay)
    b.run(graph_steps=1)
    for i in range(len(w.shots)):
        ref, uref = single_runs[i]
        out = b.seismograms(i)
        assert np.abs(ref).max() > 0
        assert np.array_equal(out, ref), (i, rel_l2(out, ref))
        assert np.array_equal(b.displ(i), uref), i
    b.close()


@pytest.mark.parametrize("assembly,overlay", [(2, 1), (3, 0), (3, 1)])
def test_graph_replay_equals_eager_steps(small_batch, assembly, overlay):
    from specfem3d_amd.batch import BatchSolver
    w = small_batch
    print(f"[gpu] graph replay, assembly {assembly}, LDS overlay {overlay}", flush=True)
    # coloured assembly: deterministic, so eager and replayed steps must agree bit for bit
    b = BatchSolver(w.db, w.shots, exact=False, shots_per_group=8, assembly=assembly)
    b.set_tuning("lds_overlay", overlay)
    b.run(graph_steps=1)
    eager = [b.seismograms(i) for i in range(len(w.shots))]
    b.reset()
    b.run(7, graph_steps=1)          # eager steps, then replays of a 10-step graph with syncs between
    while b.current_step() <= max(b.nstep):
        b.run(30, graph_steps=10)
        b.synchronize()
    for i in range(len(w.shots)):
        out = b.seismograms(i)
        assert np.array_equal(out, eager[i]), (i, rel_l2(out, eager[i]))
    b.close()


def test_autotuned_batch_matches_oracle(small_batch, oracle_mod):
    from specfem3d_amd.batch import BatchSolver
    w = small_batch
    b = BatchSolver(w.db, w.shots, exact=False)
    choice = b.autotune(trial_steps=3)
    print(f"[gpu] autotuner choice {choice}", flush=True)
    assert choice["shots_per_group"] in (4, 8) and choice["assembly"] in (0, 2, 3) and choice["lds_overlay"] in (0, 1)
    b.run(graph_steps=10)
    for i, (src, rec, n) in enumerate(w.shots):
        ref = oracle_mod.OracleSolver(w.db, src, rec, "f32").run(n)
        e = rel_l2(b.seismograms(i), ref)
        print(f"[gpu] shot {i} ({n} steps): rel-L2 vs FP32 oracle {e:.2e}", flush=True)
        assert e < 1e-5
    b.close()


@pytest.fixture(scope="module")
def visco_batch():
    from specfem3d_amd import workloads as W
    from specfem3d_amd.attenuation import setup_attenuation
    w = W.gradient_shots_box(scale=0.1, nshots=5, nrec=5, nstep_range=(60, 160), seed=13)
    att = setup_attenuation(w.db, qmu=60.0, qkappa=9000.0)
    refs = []
    from specfem3d_amd.solver import GPUSolver
    for src, rec, n in w.shots:
        g = GPUSolver(w.db, src, rec, nstep=n, attenuation=att, exact=True, assembly=2, element_order="morton")
        s = g.run().copy()
        u, _, _ = g.fields()
        g.close()
        refs.append((s, u))
    return w, att, refs


@pytest.mark.parametrize("assembly,sc", [(2, 4), (3, 8)])
def test_visco_batch_bitwise_equal_single_shot_runs(visco_batch, assembly, sc):
    from specfem3d_amd.batch import BatchSolver
    w, att, refs = visco_batch
    print(f"[gpu] viscoelastic batch of {len(w.shots)} shots, assembly {assembly}, shots per group {sc}", flush=True)
    b = BatchSolver(w.db, w.shots, exact=True, shots_per_group=sc, assembly=assembly, attenuation=att)
    b.run(graph_steps=10)
    for i, (ref, uref) in enumerate(refs):
        out = b.seismograms(i)
        assert np.abs(ref).max() > 0
        assert np.array_equal(out, ref), (i, rel_l2(out, ref))
        assert np.array_equal(b.displ(i), uref), i
    b.close()


@pytest.fixture(scope="module")
def twelve_shots():
    from specfem3d_amd import workloads as W
    from specfem3d_amd.attenuation import setup_attenuation
    from specfem3d_amd.solver import GPUSolver
    w = W.gradient_shots_box(scale=0.1, nshots=12, nrec=3, nstep_range=(30, 150), seed=21)
    att = setup_attenuation(w.db, qmu=60.0, qkappa=9000.0)
    refs = {}
    for key, a in (("elastic", None), ("visco", att)):
        out = []
        for src, rec, n in w.shots:
            g = GPUSolver(w.db, src, rec, nstep=n, attenuation=a, exact=True, assembly=2, element_order="morton")
            s = g.run().copy()
            u, _, _ = g.fields()
            g.close()
            out.append((s, u))
        refs[key] = out
    return w, att, refs


@pytest.mark.parametrize("kind,assembly,graph_steps", [("elastic", 3, 1), ("elastic", 3, 10), ("elastic", 2, 10),
                                                       ("visco", 3, 10)])
def test_compaction_bitwise_equal_single_shot_runs(twelve_shots, kind, assembly, graph_steps):
    from specfem3d_amd.batch import BatchSolver
    w, att, refs = twelve_shots
    ns = sorted((s[2] for s in w.shots), reverse=True)
    assert ns[8] < ns[0]  # the live shots drop to 8 or fewer before the longest record ends
    print(f"[gpu] 12-shot {kind} batch, assembly {assembly}, graph steps {graph_steps}, nstep {ns}", flush=True)
    b = BatchSolver(w.db, w.shots, exact=True, shots_per_group=8, assembly=assembly,
                    attenuation=att if kind == "visco" else None)
    b.run(graph_steps=graph_steps)
    for i, (ref, uref) in enumerate(refs[kind]):
        out = b.seismograms(i)
        assert np.abs(ref).max() > 0
        assert np.array_equal(out, ref), (i, rel_l2(out, ref))
        assert np.array_equal(b.displ(i), uref), i
    # a rewind restores the full rows: the second run reproduces the first
    b.reset()
    b.run(graph_steps=graph_steps)
    for i, (ref, _) in enumerate(refs[kind]):
        assert np.array_equal(b.seismograms(i), ref), i
    b.close()


@pytest.fixture(scope="module")
def coupled_shots():
    from specfem3d_amd import workloads as W
    from specfem3d_amd.solver import GPUSolver
    w = W.sea_over_visco_large(scale=0.1, nshots=10, nstep=150, nrec_side=3)
    rng = np.random.default_rng(31)
    w.shots = [(s, r, int(rng.integers(40, 151))) for s, r, _ in w.shots]
    assert w.db.nspec_acoustic > 0 and w.db.coupling_faces.n > 0 and w.db.free_faces.n > 0
    refs = {}
    for key, a in (("elastic", None), ("visco", w.attenuation)):
        out = []
        for src, rec, n in w.shots:
            g = GPUSolver(w.db, src, rec, nstep=n, attenuation=a, exact=True, assembly=2, element_order="morton")
            s = g.run().copy()
            u, _, _ = g.fields()
            p, _, _ = g.potentials()
            g.close()
            out.append((s, u, p))
        refs[key] = out
    return w, refs


@pytest.mark.parametrize("kind,assembly,sc,graph_steps", [("elastic", 2, 4, 1), ("elastic", 3, 8, 10),
                                                          ("visco", 3, 4, 10), ("visco", 2, 8, 1)])
def test_coupled_batch_bitwise_equal_single_shot_runs(coupled_shots, kind, assembly, sc, graph_steps):
    from specfem3d_amd.batch import BatchSolver
    w, refs = coupled_shots
    ns = sorted((s[2] for s in w.shots), reverse=True)
    print(f"[gpu] coupled {kind} batch of {len(w.shots)} shots, assembly {assembly}, shots per group {sc}, "
          f"graph steps {graph_steps}, nstep {ns}", flush=True)
    b = BatchSolver(w.db, w.shots, exact=True, shots_per_group=sc, assembly=assembly,
                    attenuation=w.attenuation if kind == "visco" else None)
    b.run(graph_steps=graph_steps)
    for i, (ref, uref, pref) in enumerate(refs[kind]):
        out = b.seismograms(i)
        assert np.abs(ref).max() > 0 and np.abs(pref).max() > 0
        assert np.array_equal(out, ref), (i, rel_l2(out, ref))
        assert np.array_equal(b.displ(i), uref), i
        assert np.array_equal(b.potential(i), pref), (i, rel_l2(b.potential(i), pref))
    b.close()


def test_coupled_batch_fast_atomic_matches_oracle(coupled_shots, oracle_mod):
    """FMA kernels with atomic assembly over 500 steps (the records span the sources' main pulses;
    over the first ~150 steps they hold only the low early tail of the Ricker wavelets)."""
    from specfem3d_amd.batch import BatchSolver
    from specfem3d_amd.solver import GPUSolver
    w, _ = coupled_shots
    shots = [(s, r, 500) for s, r, _ in w.shots[:2]]
    b = BatchSolver(w.db, shots, exact=False, shots_per_group=4, assembly=0)
    b.run(graph_steps=10)
    for i, (src, rec, n) in enumerate(shots):
        ref = oracle_mod.OracleSolver(w.db, src, rec, "f32").run(n)
        e = rel_l2(b.seismograms(i), ref)
        g = GPUSolver(w.db, src, rec, nstep=n, exact=False, assembly=0)
        e1 = rel_l2(g.run(), ref)
        g.close()
        print(f"[gpu] coupled shot {i} ({n} steps, FMA + atomics): rel-L2 vs FP32 oracle {e:.2e} "
              f"(single-shot run {e1:.2e})", flush=True)
        assert e < 1e-5
    b.close()
