#!/usr/bin/env python3
"""Benchmark of the single-partition solver loop on one MI355X.

Workload (default): the single-shot GPU configuration -- 96 x 96 x 48 elements (NGLOB =
28,607,425 GLL points), top 6 element layers acoustic coupled to a viscoelastic (Q_mu = 100,
3 SLS) elastic half-space with a depth gradient x (1 + 5% Gaussian random field), one seeded
moment-tensor source, 100 sea-floor receivers. One "step" = one Newmark time step driven
through the reference's C-ABI call sequence (update_displacement, acoustic forces + Stacey +
coupling, elastic forces + Stacey + coupling + source, kernel_3, seismograms).

Metric: GLL-point updates/s = NGLOB x steps / wall time of the timed steps (inputs resident
in HBM). Prints one JSON line with the roofline of the dominant kernel (viscoelastic element
stiffness, HBM bound, timed with HIP events on its stream) and the CPU oracle timed on a
bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0


def copy_peak_gbs() -> float:
    """Achievable HBM bandwidth on this GPU: a 1 GiB device-to-device copy (torch ``copy_``), read +
    written bytes over the average time of 20 copies (HIP events on torch's stream)."""
    import torch
    n = 256 * 1024 * 1024
    x = torch.ones(n, device="cuda:0", dtype=torch.float32)
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(20):
        y.copy_(x)
    t1.record()
    t1.synchronize()
    ms = t0.elapsed_time(t1) / 20
    del x, y
    torch.cuda.empty_cache()
    return 2 * 4 * n / (ms / 1e3) / 1e9


def measured_traffic(assembly: int, exact: int):
    """HBM bytes per stiffness launch from the committed rocprofv3 PMC summary of this mode.

    The summary holds FETCH_SIZE and WRITE_SIZE of the kernel, each scaled by what the same
    counter reads for a plain 1 GiB device copy; returns (bytes, summary path) or (None, None).
    """
    if exact:
        return None, None
    name = os.path.join("profiles", f"k_elastic_pt_sea_over_visco_96x96x48_{['atomic', 'ordered', 'coloured'][assembly]}.md")
    try:
        txt = open(os.path.join(HERE, name)).read()
    except OSError:
        return None, None
    m = re.search(r"traffic \(read \+ write\) \| ([0-9.]+) GB", txt)
    return (float(m.group(1)) * 1e9, name) if m else (None, None)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def stiffness_bytes_per_launch(db, attenuation=None) -> float:
    """Algorithmic HBM bytes of one element-stiffness launch (unique-traffic model).

    Per global point of the elastic domain: displ read (12 B), accel read + write (24 B),
    plus veloc read (12 B) with attenuation. Per local point of the elastic elements:
    ibool + kappa + mu (12 B); with attenuation the 18 memory variables read + written
    (144 B), the 6 stored strains read + written (48 B) and the 6 SLS factors, read as a 2-byte
    row index into a table of distinct rows when there are at most 65535 (else 24 B).
    Regular elements carry no Jacobian arrays; irregular ones add 40 B per local point.
    """
    import numpy as np
    el = db.ispec_is_elastic
    nloc_el = int(el.sum()) * 125
    nglob_el = int(np.unique(db.ibool[el]).size)
    per_glob = 12 + 24 + (12 if attenuation is not None else 0)
    per_loc = 12
    if attenuation is not None:
        rows = np.unique(np.concatenate([attenuation.factor_common.reshape(-1, 3),
                                         attenuation.factor_common_kappa.reshape(-1, 3)], 1), axis=0).shape[0]
        per_loc += 144 + 48 + (2 if rows <= 65535 else 24)
    nirr_el = int((db.irregular_element_number[el] > 0).sum()) * 125
    return float(per_glob * nglob_el + per_loc * nloc_el + 40 * nirr_el)


def batch_stiffness_bytes(db, nshots: int, fused: bool = False, attenuation=None) -> float:
    """Algorithmic HBM bytes of one batched stiffness pass with every shot live: per shot and global
    point u read (12 B) and accel read + write (24 B); per local point, shared by the shots, ibool +
    kappa + mu (12 B) and 40 B of Jacobian terms in irregular elements. With the Newmark update fused
    into the pass (assembly 3) the pass is the whole step: u, v, accel each read and written
    (72 B per shot and point) plus the inverse mass (12 B per point). With attenuation each shot also
    reads and writes its 18 memory variables and 6 stored strains per local point (192 B), and the
    SLS factors are shared (a 2-byte row index, or 24 B of planes beyond 65535 distinct rows); a
    separate (unfused) pass also reads v (12 B per shot and point)."""
    import numpy as np
    el = db.ispec_is_elastic
    nloc = int(el.sum()) * 125
    nglob_el = int(np.unique(db.ibool[el]).size)
    nirr = int((db.irregular_element_number[el] > 0).sum()) * 125
    per_shot = 72 if fused else 36
    per_shot_loc = 0
    shared_loc = 12
    if attenuation is not None:
        per_shot += 0 if fused else 12
        per_shot_loc = 192
        rows = np.unique(np.concatenate([attenuation.factor_common.reshape(-1, 3),
                                         attenuation.factor_common_kappa.reshape(-1, 3)], 1), axis=0).shape[0]
        shared_loc += 2 if rows <= 65535 else 24
    return float(per_shot * nglob_el * nshots + per_shot_loc * nloc * nshots + shared_loc * nloc + 40 * nirr
                 + (12 * nglob_el if fused else 0))


def batch_step_bytes(db, shot_steps: int, nsteps: int) -> float:
    """Algorithmic bytes of a whole batched run (the unique-traffic model: every array read and written
    once per step at its natural size): 72 B per live shot and point per step for u, v, accel, plus per
    step the shared ibool + kappa + mu (12 B per local point), irregular Jacobian terms (40 B) and the
    inverse mass (12 B per point)."""
    import numpy as np
    el = db.ispec_is_elastic
    nloc = int(el.sum()) * 125
    nglob_el = int(np.unique(db.ibool[el]).size)
    nirr = int((db.irregular_element_number[el] > 0).sum()) * 125
    return float(72 * nglob_el * shot_steps + (12 * nloc + 40 * nirr + 12 * nglob_el) * nsteps)


def batched_measurement(args) -> dict:
    """Many shots on one mesh: 64x64x32 elastic box (gradient + GRF model, no attenuation), 32 shots with
    seeded random CMT sources (Ricker 0.5-2 Hz), 50 receivers each and record lengths drawn from
    1000-4000 steps, run to the end of the longest record with autotuned shots-per-workgroup/assembly and
    hipGraph-captured step groups. Also times the plain single-shot path on the same mesh."""
    import numpy as np
    from specfem3d_amd import workloads as W
    from specfem3d_amd.batch import BatchSolver
    from specfem3d_amd.solver import GPUSolver
    w = W.gradient_shots_box(scale=args.batch_scale, nshots=args.batch_shots)
    if args.batch_max_steps > 0:
        w.shots = [(s, r, min(n, args.batch_max_steps)) for s, r, n in w.shots]
    db = w.db
    nsteps = [n for _, _, n in w.shots]
    log(f"[bench] batched workload {w.name}: NGLOB {db.nglob}, record lengths {min(nsteps)}-{max(nsteps)}")
    b = BatchSolver(db, w.shots, exact=False)
    choice = b.autotune(trial_steps=5)
    log(f"[bench] autotuner: {choice}")
    kms, sms = b.time(20)            # all shots live (first records are >= 1000 steps)
    b.reset()
    b.synchronize()
    t = time.perf_counter()
    b.run(graph_steps=10)
    b.synchronize()
    wall = time.perf_counter() - t
    finite = all(bool(np.isfinite(b.seismograms(i)).all()) for i in range(len(w.shots)))
    b.close()
    value = db.nglob * float(sum(nsteps)) / wall
    kbytes = batch_stiffness_bytes(db, len(w.shots), fused=choice["assembly"] == 3)
    ach = kbytes / (kms / 1e3) / 1e9
    run_gbs = batch_step_bytes(db, sum(nsteps), max(nsteps)) / wall / 1e9
    # plain single-shot path on the same mesh (first shot), timed the same way
    src, rec, n = w.shots[0]
    ps = min(args.plain_steps, n)
    pw = min(10, ps // 2)
    # the fastest single-shot configuration on this elastic mesh: coloured assembly (plain
    # read-modify-write per colour) runs the stiffness ~3x faster than float atomics here
    g = GPUSolver(db, src, rec, nstep=ps, exact=False, assembly=2)
    for _ in range(pw):
        g.step()
    g.synchronize()
    tp = time.perf_counter()
    for _ in range(ps - pw):
        g.step()
    g.synchronize()
    plain = db.nglob * (ps - pw) / (time.perf_counter() - tp)
    g.close()
    log(f"[bench] batched {value:.3e} GLL-pt updates/s in {wall:.2f}s (stiffness {kms:.3f} ms/step, step {sms:.3f} ms "
        f"with all shots live), plain single shot {plain:.3e}, ratio {value / plain:.2f}, finite={finite}")
    return {"workload": w.name, "value": value, "unit": "GLL-point updates/s", "wall_s": round(wall, 3),
            "shot_steps": int(sum(nsteps)), "plain_single_shot": plain, "plain_assembly": "coloured",
            "ratio_vs_plain": value / plain,
            "all_live_value": db.nglob * len(w.shots) / (sms / 1e3),
            "all_live_ratio_vs_plain": db.nglob * len(w.shots) / (sms / 1e3) / plain,
            "autotuned": choice, "seismograms_finite": finite,
            "run_roofline": {"bound": "hbm", "achieved": round(run_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": round(run_gbs / HBM_PEAK_GBS, 4),
                             "bytes_per_update": round(batch_step_bytes(db, sum(nsteps), max(nsteps))
                                                       / (db.nglob * float(sum(nsteps))), 2)},
            "roofline": {"bound": "hbm", "kernel": "k_elastic_batch (all shots live"
                         + (", Newmark update fused)" if choice["assembly"] == 3 else ")"), "achieved": round(ach, 1),
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                         "bytes_per_launch": kbytes, "avg_pass_ms": round(kms, 4), "step_ms": round(sms, 4)}}


def batched_visco_measurement(args) -> dict:
    """Many viscoelastic shots on one mesh: the 64x64x32 gradient + GRF box with 3-SLS attenuation
    (Q_mu = 100), ``--visco-shots`` seeded random CMT shots of ``--visco-steps`` steps each (every shot
    live for the whole run), against the single-shot viscoelastic path on the same mesh."""
    import numpy as np
    from specfem3d_amd import workloads as W
    from specfem3d_amd.attenuation import setup_attenuation
    from specfem3d_amd.batch import BatchSolver
    from specfem3d_amd.solver import GPUSolver
    w = W.gradient_shots_box(scale=args.batch_scale, nshots=args.visco_shots)
    att = setup_attenuation(w.db, qmu=100.0, qkappa=9000.0)
    n = args.visco_steps
    shots = [(s, r, n) for s, r, _ in w.shots]
    db = w.db
    log(f"[bench] batched viscoelastic workload: NGLOB {db.nglob}, {len(shots)} shots x {n} steps")
    b = BatchSolver(db, shots, exact=False, attenuation=att)
    choice = b.autotune(trial_steps=3)
    log(f"[bench] viscoelastic autotuner: {choice}")
    kms, sms = b.time(10)
    b.reset()
    b.synchronize()
    t = time.perf_counter()
    b.run(graph_steps=10)
    b.synchronize()
    wall = time.perf_counter() - t
    finite = all(bool(np.isfinite(b.seismograms(i)).all()) for i in range(len(shots)))
    b.close()
    value = db.nglob * float(n * len(shots)) / wall
    kbytes = batch_stiffness_bytes(db, len(shots), fused=choice["assembly"] == 3, attenuation=att)
    ach = kbytes / (kms / 1e3) / 1e9
    src, rec, _ = shots[0]
    ps = min(args.plain_steps, n)
    pw = min(10, ps // 2)
    g = GPUSolver(db, src, rec, nstep=ps, attenuation=att, exact=False, assembly=0)
    for _ in range(pw):
        g.step()
    g.synchronize()
    tp = time.perf_counter()
    for _ in range(ps - pw):
        g.step()
    g.synchronize()
    plain = db.nglob * (ps - pw) / (time.perf_counter() - tp)
    g.close()
    log(f"[bench] batched viscoelastic {value:.3e} GLL-pt updates/s in {wall:.2f}s (pass {kms:.3f} ms, step "
        f"{sms:.3f} ms), plain viscoelastic single shot {plain:.3e}, ratio {value / plain:.2f}, finite={finite}")
    return {"workload": f"{w.name.split(',')[0]} viscoelastic Q_mu=100, {len(shots)} shots x {n} steps",
            "value": value, "unit": "GLL-point updates/s", "wall_s": round(wall, 3), "plain_single_shot": plain,
            "ratio_vs_plain": value / plain, "autotuned": choice, "seismograms_finite": finite,
            "roofline": {"bound": "hbm", "kernel": "k_visco_batch (all shots live"
                         + (", Newmark update fused)" if choice["assembly"] == 3 else ")"), "achieved": round(ach, 1),
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                         "bytes_per_launch": kbytes, "avg_pass_ms": round(kms, 4), "step_ms": round(sms, 4)}}


def homogeneous_box_measurement(args) -> dict:
    """The 20^3 homogeneous elastic box (531,441 points, 2000 steps, the CPU-runnable case) through the
    ABI call sequence with the library's defaults (FMA kernels, coloured assembly without attenuation),
    the FP32 oracle on one host core over a sample of its steps, and the reference's own Fortran
    solver figure for the same case (measured on another host)."""
    import numpy as np
    from specfem3d_amd import workloads as W
    from specfem3d_amd.solver import GPUSolver
    w = W.homogeneous_box()
    db = w.db
    n = w.nstep
    g = GPUSolver(db, w.sources, w.receivers, nstep=n, exact=False, assembly=-1)
    for _ in range(20):
        g.step()
    g.synchronize()
    t = time.perf_counter()
    for _ in range(n - 20):
        g.step()
    g.synchronize()
    wall = time.perf_counter() - t
    finite = bool(np.isfinite(g.seis_d[: g.seismo_current]).all())
    g.close()
    value = db.nglob * (n - 20) / wall
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(HERE, "oracle"))
        import oracle
        ora = oracle.OracleSolver(db, w.sources, w.receivers, "f32")
        ns = 100
        tc = time.perf_counter()
        for _ in range(ns):
            ora.step()
        cpu = {"value": db.nglob * ns / (time.perf_counter() - tc), "unit": "GLL-point updates/s", "cores": 1,
               "kind": "port", "sample": f"first {ns} of the {n} steps (FP32 C oracle, one thread)"}
    log(f"[bench] 20^3 homogeneous box: {value:.3e} GLL-pt updates/s ({wall / (n - 20) * 1e3:.3f} ms/step), "
        f"finite={finite}")
    return {"workload": w.name, "nglob": db.nglob, "value": value, "unit": "GLL-point updates/s",
            "ms_per_step": wall / (n - 20) * 1e3, "steps": n - 20, "assembly": "library default (coloured)",
            "seismograms_finite": finite, "cpu_baseline": cpu,
            "reference_fortran_cpu": {"value": 7.16e6, "unit": "GLL-point updates/s", "cores": 1,
                                      "kind": "reference",
                                      "sample": "the reference's own FP32 solver (flang -O3, serial), whole "
                                                "2000-step run, measured on another host (8-core Xeon @ 2.70 GHz)"}}


def cpml_example_measurement(args) -> dict:
    """The reference example CPML_examples/homogeneous_halfspace_HEX8_elastic_absorbing_CPML_5sides
    (36 x 36 x 16 irregular elements, 1,366,625 points, 9036 C-PML elements, 2400 steps) with the
    library defaults (FMA kernels, coloured assembly), and its seismograms against the example's
    REF_SEIS."""
    import ctypes as C
    import numpy as np
    from specfem3d_amd.examples import cpml5_example
    from specfem3d_amd.solver import GPUSolver
    db, src, rec, ref, dt, _ = cpml5_example()
    n = ref.shape[0]
    g = GPUSolver(db, src, rec, nstep=n, exact=False, assembly=args.cpml_assembly)
    warm = 20
    half = (n - warm) // 2
    for _ in range(warm):
        g.step()
    g.synchronize()
    tot, cnt = C.c_double(0), C.c_int(0)
    ptot, pcnt = C.c_double(0), C.c_int(0)
    # HIP events on the mesh stream: the whole elastic stiffness call over the first half of the timed
    # steps, its C-PML launch alone over the second half (timing_scope 1)
    g.L.sf_timing_start(g.MP, C.c_int(half))
    t = time.perf_counter()
    for it in range(n - warm):
        if it == half:
            g.L.sf_timing_read(g.MP, C.byref(tot), C.byref(cnt))
            g.L.sf_set_tuning(g.MP, b"timing_scope", C.c_int(1))
            g.L.sf_timing_start(g.MP, C.c_int(n - warm - half))
        g.step()
    g.synchronize()
    wall = time.perf_counter() - t
    g.L.sf_timing_read(g.MP, C.byref(ptot), C.byref(pcnt))
    out = g.seismograms("d").copy()
    g.close()
    e = float(np.sqrt(((out.astype(np.float64) - ref) ** 2).sum() / (ref.astype(np.float64) ** 2).sum()))
    value = db.nglob * (n - warm) / wall
    stiff_ms = tot.value / max(cnt.value, 1)
    pml_ms = ptot.value / max(pcnt.value, 1)
    nbytes = pml_bytes_per_launch(db)
    achieved = nbytes / (pml_ms * 1e-3) / 1e9
    log(f"[bench] C-PML example: {value:.3e} GLL-pt updates/s ({wall / (n - warm) * 1e3:.3f} ms/step, stiffness "
        f"incl. C-PML {stiff_ms:.3f} ms, C-PML launch {pml_ms * 1e3:.1f} us = {achieved:.0f} GB/s algorithmic), "
        f"rel-L2 vs REF_SEIS {e:.2e}")
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(HERE, "oracle"))
        import oracle
        ora = oracle.OracleSolver(db, src, rec, "f32")
        ns = 12
        tc = time.perf_counter()
        for _ in range(ns):
            ora.step()
        cpu = {"value": db.nglob * ns / (time.perf_counter() - tc), "unit": "GLL-point updates/s", "cores": 1,
               "kind": "port", "sample": f"first {ns} of the {n} steps (FP32 C oracle with C-PML, one thread)"}
        log(f"[bench] C-PML example oracle: {cpu['value']:.3e} GLL-pt updates/s on one core")
    return {"workload": "CPML_examples/homogeneous_halfspace_HEX8_elastic_absorbing_CPML_5sides", "nglob": db.nglob,
            "nspec": db.nspec, "nspec_cpml": db.pml.n, "value": value, "unit": "GLL-point updates/s",
            "ms_per_step": wall / (n - warm) * 1e3, "steps": n - warm,
            "stiffness_ms": stiff_ms, "rel_l2_vs_ref_seis": e,
            "roofline": {"kernel": "k_elastic_pml (C-PML elements)", "bound": "hbm", "achieved": round(achieved, 1),
                         "peak": 8000.0, "unit": "GB/s", "frac": round(achieved / 8000.0, 4),
                         "bytes_per_launch": nbytes, "avg_launch_ms": pml_ms,
                         "traffic": 400.6e6, "traffic_source": "profiles/k_elastic_pml_cpml5_example.md"},
            "cpu_baseline": cpu}


def pml_bytes_per_launch(db) -> float:
    """Algorithmic HBM bytes of one C-PML stiffness launch: every value moved once. Per C-PML element,
    the live memory-variable planes read and written (a plane is live when the coefficient plane that
    multiplies it is not zero over the element) and the coefficient planes read (125 values, or one
    for a plane uniform over the element); per local point ibool, kappa/mu and, on irregular
    elements, the ten Jacobian terms; per global point of the C-PML elements u, v and PML_displ_old
    read and accel read and written."""
    import numpy as np
    from specfem3d_amd.gll import get_gll
    p = db.pml
    n = p.n
    w = get_gll().wgll
    w3 = (w[None, None, :] * w[None, :, None] * w[:, None, None]).ravel().astype(np.float32)
    rho = db.rhostore.reshape(db.nspec, 125)[p.cpml_to_spec]
    cf = np.concatenate([p.coef_alpha.reshape(n, 125, 9), p.coef_beta.reshape(n, 125, 9),
                         p.coef_abar.reshape(n, 125, 5), p.coef_strain.reshape(n, 125, 18),
                         (w3[None, :] * rho)[:, :, None]], axis=2)
    uni = np.all(cf == cf[:, :1, :], axis=1)
    nz = np.any(cf != 0, axis=1)
    live = 0
    for a in range(9):
        for s in range(3):
            live += int(nz[:, 23 + 4 * (a // 3) + 1 + s].sum())
    for b in range(12):
        live += int(nz[:, 23 + 13 + 2 * (b // 4)].sum())
    live += 3 * int(nz[:, 20:23].sum())
    ib = db.ibool.reshape(db.nspec, 125)[p.cpml_to_spec]
    nirr = int((db.irregular_element_number[p.cpml_to_spec] != 0).sum())
    return float(live * 125 * 8 + int((~uni).sum()) * 125 * 4 + int(uni.sum()) * 4
                 + n * 125 * 12 + nirr * 125 * 40 + np.unique(ib).size * 60)


def coupled_batch_measurement(args) -> dict:
    """The fluid-over-viscoelastic many-shot workload (the 160 x 160 x 80 builder: 10 % of the element
    layers fluid, 3-SLS solid with Q_mu = 100, seeded random CMT shots below the sea floor, one
    sea-floor receiver grid) at ``--coupled-scale``: ``--coupled-shots`` shots of ``--coupled-steps``
    steps batched (autotuned, graph-replayed), the single-shot path on the same mesh, and the
    checkpointed adjoint (forward with snapshots + adjoint with recomputation) of the first shot
    against synthetics of a 3 % Gaussian mu perturbation."""
    import numpy as np
    from specfem3d_amd import adjoint as A
    from specfem3d_amd import workloads as W
    from specfem3d_amd.autograd import perturbed_database
    from specfem3d_amd.batch import BatchSolver
    from specfem3d_amd.solver import GPUSolver
    n = args.coupled_steps
    t0 = time.time()
    w = W.sea_over_visco_large(scale=args.coupled_scale, nshots=args.coupled_shots, nstep=n)
    db = w.db
    shots = w.shots
    log(f"[bench] coupled batch workload {w.name}: NGLOB {db.nglob}, built in {time.time() - t0:.1f}s")
    b = BatchSolver(db, shots, exact=False, attenuation=w.attenuation)
    choice = b.autotune(trial_steps=3)
    log(f"[bench] coupled autotuner: {choice}")
    kms, sms = b.time(10)
    b.reset()
    b.synchronize()
    t = time.perf_counter()
    b.run(graph_steps=10)
    b.synchronize()
    wall = time.perf_counter() - t
    finite = all(bool(np.isfinite(b.seismograms(i)).all()) for i in range(len(shots)))
    b.close()
    value = db.nglob * float(n * len(shots)) / wall
    src, rec, _ = shots[0]
    ps = min(args.plain_steps, n)
    pw = min(10, ps // 2)
    g = GPUSolver(db, src, rec, nstep=ps, attenuation=w.attenuation, exact=False)
    for _ in range(pw):
        g.step()
    g.synchronize()
    tp = time.perf_counter()
    for _ in range(ps - pw):
        g.step()
    g.synchronize()
    plain = db.nglob * (ps - pw) / (time.perf_counter() - tp)
    g.close()
    log(f"[bench] coupled batch {value:.3e} GLL-pt updates/s in {wall:.2f}s (step {sms:.3f} ms), single shot "
        f"{plain:.3e}, ratio {value / plain:.2f}, finite={finite}")
    adj = None
    if args.coupled_adjoint:
        na = args.coupled_adjoint_steps
        from specfem3d_amd.attenuation import estimate_min_resolved_period, setup_attenuation
        dbp = perturbed_database(db, dlnmu=0.03 * W.gaussian_bump(db))
        # the attenuated moduli of the observed model come from its own (perturbed) mu; same SLS band
        att_obs = setup_attenuation(dbp, qmu=100.0, qkappa=9000.0,
                                    min_resolved_period=estimate_min_resolved_period(db))
        g = GPUSolver(dbp, src, rec, nstep=na, attenuation=att_obs, exact=False)
        obs = g.run().copy()
        g.close()
        ks = A.KernelSolver(db, src, rec, na, nt_dump=64, attenuation=w.attenuation)
        ks.synchronize()
        t = time.perf_counter()
        syn = ks.forward()
        tf = time.perf_counter() - t
        rk, mk, kk = ks.adjoint(A.waveform_adjoint_source(syn, obs))
        aw = time.perf_counter() - t
        ka = ks.acoustic_kernels
        ks.close()
        chi = A.waveform_misfit(syn, obs, db.dt)
        kfin = bool(np.isfinite(mk).all() and np.abs(mk).max() > 0)
        log(f"[bench] coupled kernel run {aw:.2f}s (forward with snapshots {tf:.2f}s), misfit {chi:.4e}, "
            f"max |mu_kl| {float(np.abs(mk).max()):.3e}, non-finite mu/rho/kappa "
            f"{int((~np.isfinite(mk)).sum())}/{int((~np.isfinite(rk)).sum())}/{int((~np.isfinite(kk)).sum())}, "
            f"fluid kernels finite={bool(np.isfinite(ka[0]).all() and np.isfinite(ka[1]).all())}, kernels finite={kfin}")
        adj = {"wall_s": round(aw, 3), "forward_s": round(tf, 3), "steps": na, "nt_dump": 64, "misfit": chi,
               "kernels_finite": kfin, "unit": "s (forward with checkpoints + adjoint with recompute, one shot)"}
    return {"workload": w.name, "scale": args.coupled_scale, "nglob": db.nglob, "value": value,
            "unit": "GLL-point updates/s", "wall_s": round(wall, 3), "step_ms_all_live": round(sms, 4),
            "plain_single_shot": plain, "ratio_vs_plain": value / plain, "autotuned": choice,
            "seismograms_finite": finite, "adjoint": adj}


def adjoint_measurement(args) -> dict:
    """Wall time of a checkpointed sensitivity-kernel run: forward with a snapshot every 64 steps +
    adjoint with recomputation (kernels copied back to the host included), on the 48^3 elastic
    workload with an L2 waveform misfit against synthetics of a 3% Gaussian mu perturbation."""
    import numpy as np
    from specfem3d_amd import adjoint as A
    from specfem3d_amd import workloads as W
    from specfem3d_amd.autograd import perturbed_database
    from specfem3d_amd.solver import GPUSolver
    w = W.gradient_adjoint_box(scale=args.adjoint_scale, nstep=args.adjoint_steps)
    db = w.db
    n = args.adjoint_steps
    g = GPUSolver(perturbed_database(db, dlnmu=0.03 * W.gaussian_bump(db)), w.sources, w.receivers, nstep=n,
                  exact=False, assembly=0)
    obs = g.run().copy()
    g.close()
    log(f"[bench] adjoint workload {w.name}: NGLOB {db.nglob}, observed data computed")
    ks = A.KernelSolver(db, w.sources, w.receivers, n, nt_dump=64, assembly=args.adjoint_assembly)
    ks.synchronize()
    t = time.perf_counter()
    syn = ks.forward()
    tf = time.perf_counter() - t
    rk, mk, kk = ks.adjoint(A.waveform_adjoint_source(syn, obs))
    wall = time.perf_counter() - t
    ks.close()
    chi = A.waveform_misfit(syn, obs, db.dt)
    finite = bool(np.isfinite(mk).all() and np.abs(mk).max() > 0)
    log(f"[bench] kernel run {wall:.2f}s (forward with snapshots {tf:.2f}s), misfit {chi:.4e}, kernels finite={finite}")
    return {"workload": w.name, "wall_s": round(wall, 3), "forward_s": round(tf, 3), "steps": n, "nt_dump": 64,
            "assembly": ["atomic", "ordered", "coloured"][args.adjoint_assembly],
            "nglob": db.nglob, "misfit": chi, "kernels_finite": finite,
            "unit": "s (forward with checkpoints + adjoint with recompute)"}


def large_measurement(args) -> dict:
    """Largest single-GPU workload: 160x160x80 elements (131.9 M GLL points), 10 acoustic layers over a
    3-SLS viscoelastic solid, one of its shots run for a few steps through the ABI call sequence:
    throughput, stiffness roofline and device memory in use."""
    import ctypes as C
    import numpy as np
    from specfem3d_amd import workloads as W
    from specfem3d_amd.solver import GPUSolver
    nx = max(2, int(round(160 * args.large_scale)))
    nz = max(8, int(round(80 * args.large_scale)))
    r_bytes = 18 * 125 * 4 * nx * nx * (nz - max(1, nz // 8))
    if r_bytes >= 4 * 2 ** 30:
        # At 160x160x80 (131.9 M points) the stepping has twice produced no output for minutes after a
        # complete 11 s set-up; the one size-related difference to the runs that work is per-point state
        # above 4 GiB (memory variables 18.4 GB). Until the cause is found, any size with a
        # memory-variable array of 4 GiB or more is refused.
        raise SystemExit(f"bench.py: --large-steps with a {r_bytes / 2 ** 30:.1f} GiB memory-variable array is "
                         "disabled: the 131.9 M-point single-shot run stalled after set-up for a reason not yet "
                         "found (runs below 4 GiB of per-point state only)")
    t0 = time.time()
    n = args.large_steps + 3
    w = W.sea_over_visco_large(scale=args.large_scale, nshots=1, nstep=n)
    db = w.db
    log(f"[bench] built {w.name}: NGLOB {db.nglob} in {time.time() - t0:.1f}s")
    os.environ["SF_VERBOSE_PREPARE"] = "1"     # the set-up of 131.9 M points takes minutes: report its stages
    g = GPUSolver(db, w.sources, w.receivers, nstep=n, attenuation=w.attenuation, exact=False, assembly=0)
    os.environ.pop("SF_VERBOSE_PREPARE", None)
    log("[bench] large: device arrays prepared")
    for i in range(3):
        ts = time.perf_counter()
        g.step()
        g.synchronize()
        log(f"[bench] large: warm-up step {i + 1} took {time.perf_counter() - ts:.3f}s")
    free, used, total = C.c_float(0), C.c_float(0), C.c_float(0)
    g.L.get_free_device_memory_(C.byref(free), C.byref(used), C.byref(total))
    g.L.sf_timing_start(g.MP, C.c_int(args.large_steps))
    t = time.perf_counter()
    for _ in range(args.large_steps):
        g.step()
    g.synchronize()
    wall = time.perf_counter() - t
    tot, cnt = C.c_double(0.0), C.c_int(0)
    g.L.sf_timing_read(g.MP, C.byref(tot), C.byref(cnt))
    finite = bool(np.isfinite(g.seis_d[: g.seismo_current]).all())
    g.close()
    kavg = tot.value / max(cnt.value, 1)
    kbytes = stiffness_bytes_per_launch(db, w.attenuation)
    ach = kbytes / (kavg / 1e3) / 1e9
    log(f"[bench] large: {args.large_steps} steps in {wall:.2f}s, stiffness {kavg:.2f} ms, device memory used "
        f"{used.value:.0f} of {total.value:.0f} MB, finite={finite}")
    return {"workload": w.name, "nglob": db.nglob, "value": db.nglob * args.large_steps / wall,
            "unit": "GLL-point updates/s", "steps": args.large_steps, "ms_per_step": wall / args.large_steps * 1e3,
            "device_mb_used": round(used.value), "seismograms_finite": finite,
            "roofline": {"bound": "hbm", "kernel": "k_elastic_pt (viscoelastic 3-SLS stiffness)",
                         "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(ach / HBM_PEAK_GBS, 4), "bytes_per_launch": kbytes,
                         "avg_launch_ms": round(kavg, 4)}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--scale", type=float, default=1.0, help="shrink the mesh (tests only)")
    ap.add_argument("--exact", type=int, default=0, help="1: kernels without FMA (CPU rounding)")
    ap.add_argument("--assembly", type=int, default=0,
                    help="0 atomic scatter, 1 ordered (CPU element order), 2 coloured (no atomics)")
    ap.add_argument("--element-order", default="none",
                    choices=["morton", "none", "brick1x2", "brick2x1", "brick2x2", "brick1x4", "brick4x4", "brick2x4"],
                    help="launch order of the elastic elements of the single-shot run")
    ap.add_argument("--att-variant", type=int, default=-1,
                    help="attenuation schedule of the single-shot stiffness kernel (-1: library default)")
    ap.add_argument("--skip-single", action="store_true",
                    help="skip the single-shot 96x96x48 measurement (profiling runs of the other parts)")
    ap.add_argument("--cpu-steps", type=int, default=4, help="steps of the CPU oracle sample (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-shots", type=int, default=32, help="shots of the batched 64x64x32 workload (0: skip)")
    ap.add_argument("--batch-scale", type=float, default=1.0, help="shrink the batched mesh (tests only)")
    ap.add_argument("--visco-shots", type=int, default=16,
                    help="viscoelastic shots batched on the 64x64x32 mesh with attenuation (0: skip)")
    ap.add_argument("--visco-steps", type=int, default=300, help="steps of each batched viscoelastic shot")
    ap.add_argument("--plain-steps", type=int, default=300, help="steps of the single-shot run on the batched mesh")
    ap.add_argument("--batch-max-steps", type=int, default=0,
                    help="clip every shot's record length to this many steps (profiling runs; 0: full records)")
    ap.add_argument("--adjoint-steps", type=int, default=4000,
                    help="steps of the checkpointed kernel run on the 48^3 adjoint workload (0: skip)")
    ap.add_argument("--adjoint-assembly", type=int, default=2,
                    help="assembly of the kernel run: 0 atomic, 1 ordered, 2 coloured (default: the fastest here)")
    ap.add_argument("--adjoint-scale", type=float, default=1.0, help="shrink the adjoint mesh (tests only)")
    ap.add_argument("--homogeneous-box", type=int, default=1,
                    help="1: also run the 20^3 homogeneous box single-shot figures")
    ap.add_argument("--cpml-example", type=int, default=1,
                    help="1: also run the reference's C-PML example (2400 steps) and compare with its REF_SEIS")
    ap.add_argument("--cpml-assembly", type=int, default=-1,
                    help="assembly of the C-PML example run (-1: library default, coloured without attenuation)")
    ap.add_argument("--coupled-shots", type=int, default=8,
                    help="shots of the batched fluid-over-viscoelastic workload (0: skip)")
    ap.add_argument("--coupled-steps", type=int, default=300, help="steps of each batched coupled shot")
    ap.add_argument("--coupled-scale", type=float, default=0.6,
                    help="scale of the 160x160x80 coupled builder for the batched run (0.6: 96x96x48)")
    ap.add_argument("--coupled-adjoint", type=int, default=1,
                    help="1: also time the checkpointed adjoint of the first coupled shot")
    ap.add_argument("--coupled-adjoint-steps", type=int, default=1500,
                    help="steps of the coupled checkpointed adjoint run")
    ap.add_argument("--large-steps", type=int, default=0,
                    help="steps of one shot of the 160x160x80 viscoelastic workload (0: skip; the mesh build "
                         "takes minutes)")
    ap.add_argument("--large-scale", type=float, default=1.0, help="shrink the large mesh (tests only)")
    ap.add_argument("--kernel-probe", action="store_true",
                    help="only run a short fixed sequence of batched and single-shot steps (PMC counter runs)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB device copy (calibrates rocprof FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()

    if args.kernel_probe:
        # short fixed sequence for PMC counter runs: 4 batched steps of the 32-shot workload with
        # each batched kernel variant, then 6 single-shot steps with coloured assembly
        from specfem3d_amd import workloads as W
        from specfem3d_amd.batch import BatchSolver
        from specfem3d_amd.solver import GPUSolver
        w = W.gradient_shots_box(scale=args.batch_scale, nshots=32)
        w.shots = [(s_, r_, min(n_, 40)) for s_, r_, n_ in w.shots]
        b = BatchSolver(w.db, w.shots, exact=False, shots_per_group=8, assembly=3)
        for ov in (0, 1):
            b.set_tuning("lds_overlay", ov)
            b.reset()
            b.run(4, graph_steps=1)
            log(f"[bench] kernel probe: 4 batched steps, lds_overlay {ov}")
        b.close()
        src, rec, _ = w.shots[0]
        g = GPUSolver(w.db, src, rec, nstep=6, exact=False, assembly=2)
        g.run()
        g.close()
        log("[bench] kernel probe: 6 single-shot steps, coloured assembly")
        return

    if args.copy_probe:
        import torch
        n = 256 * 1024 * 1024
        x = torch.ones(n, device="cuda:0", dtype=torch.float32)
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        log(f"[bench] copy probe: 3 copies of {4 * n} bytes")
        return

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); this build has no multi-GPU path. "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from specfem3d_amd import workloads as W
    from specfem3d_amd.solver import GPUSolver

    if args.skip_single:
        out = {"metric": "GLL-point updates/s/GPU (forward, single shot)", "value": None,
               "batched": batched_measurement(args) if args.batch_shots > 0 else None,
               "batched_viscoelastic": batched_visco_measurement(args) if args.visco_shots > 0 else None,
               "adjoint": adjoint_measurement(args) if args.adjoint_steps > 0 else None,
               "coupled_batch": coupled_batch_measurement(args) if args.coupled_shots > 0 else None,
               "cpml_example": cpml_example_measurement(args) if args.cpml_example else None,
               "large": large_measurement(args) if args.large_steps > 0 else None}
        print(json.dumps(out), flush=True)
        return

    t0 = time.time()
    nstep_total = args.warmup + args.steps
    w = W.sea_over_visco_box(scale=args.scale, nstep=max(nstep_total, 1))
    db = w.db
    log(f"[bench] built {w.name}: {db.summary()} in {time.time() - t0:.1f}s")

    tuning = {}
    if args.att_variant >= 0:
        tuning["att_variant"] = args.att_variant
    g = GPUSolver(db, w.sources, w.receivers, nstep=nstep_total, attenuation=w.attenuation,
                  exact=bool(args.exact), assembly=args.assembly,
                  element_order=None if args.element_order == "none" else args.element_order, tuning=tuning)
    log("[bench] prepared device arrays")
    for _ in range(args.warmup):
        g.step()
    g.synchronize()
    log(f"[bench] warmup {args.warmup} steps done")
    import ctypes as C
    g.L.sf_timing_start(g.MP, C.c_int(args.steps))
    g.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        g.step()
    g.synchronize()
    wall = time.perf_counter() - t
    tot = C.c_double(0.0)
    cnt = C.c_int(0)
    g.L.sf_timing_read(g.MP, C.byref(tot), C.byref(cnt))
    seis = g.seis_d[: g.seismo_current]
    finite = bool(np.isfinite(seis).all())
    g.close()
    log(f"[bench] timed {args.steps} steps: {wall:.3f}s, stiffness kernel {tot.value / max(cnt.value, 1):.3f} ms avg "
        f"over {cnt.value} launches, seismograms finite={finite}")

    nglob = db.nglob
    value = nglob * args.steps / wall
    kbytes = stiffness_bytes_per_launch(db, w.attenuation)
    kavg_s = tot.value / max(cnt.value, 1) / 1e3
    achieved = kbytes / kavg_s / 1e9
    traffic, tsrc = measured_traffic(args.assembly, args.exact) if args.scale == 1.0 else (None, None)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "k_elastic_pt (viscoelastic 3-SLS stiffness)", "bytes_per_launch": kbytes,
                "avg_launch_ms": round(kavg_s * 1e3, 4), "traffic_source": tsrc}

    copy_gbs = copy_peak_gbs()
    roofline["measured_copy_peak"] = round(copy_gbs, 1)
    roofline["frac_of_copy_peak"] = round(achieved / copy_gbs, 4)
    log(f"[bench] 1 GiB device copy: {copy_gbs:.0f} GB/s")

    cpu = None
    if not args.no_cpu_baseline and args.cpu_steps > 0:
        sys.path.insert(0, os.path.join(HERE, "oracle"))
        import oracle
        # the host's cores (16 on the GPU box, OMP_NUM_THREADS there), capped at 16
        threads = max(1, min(16, int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)))
        ora = oracle.OracleSolver(db, w.sources, w.receivers, "f32", attenuation=w.attenuation, threads=threads)
        nthr = args.cpu_steps * 4 if threads > 1 else args.cpu_steps
        tc = time.perf_counter()
        for _ in range(nthr):
            ora.step()
        ct = time.perf_counter() - tc
        cpu = {"value": nglob * nthr / ct, "unit": "GLL-point updates/s", "cores": threads, "kind": "port",
               "sample": f"{nthr} time steps of the same {w.name} mesh (FP32 C oracle, {threads} OpenMP threads)"}
        log(f"[bench] CPU oracle {nthr} steps on {threads} threads in {ct:.1f}s")
        if threads > 1:
            ora1 = oracle.OracleSolver(db, w.sources, w.receivers, "f32", attenuation=w.attenuation)
            tc = time.perf_counter()
            for _ in range(args.cpu_steps // 2 or 1):
                ora1.step()
            c1 = time.perf_counter() - tc
            cpu["single_thread_value"] = nglob * (args.cpu_steps // 2 or 1) / c1
            log(f"[bench] CPU oracle {args.cpu_steps // 2 or 1} steps on 1 thread in {c1:.1f}s")

    adj = adjoint_measurement(args) if args.adjoint_steps > 0 else None
    batched = batched_measurement(args) if args.batch_shots > 0 else None
    batched_visco = batched_visco_measurement(args) if args.visco_shots > 0 else None
    coupled = coupled_batch_measurement(args) if args.coupled_shots > 0 else None
    box = homogeneous_box_measurement(args) if args.homogeneous_box else None
    cpml = cpml_example_measurement(args) if args.cpml_example else None
    large = large_measurement(args) if args.large_steps > 0 else None

    out = {
        "metric": "GLL-point updates/s/GPU (forward, single shot)",
        "value": value,
        "unit": "GLL-point updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded gradient + Gaussian random field model, random CMT source)",
        "config": {"workload": w.name, "nglob": nglob, "nspec": db.nspec, "nspec_acoustic": db.nspec_acoustic,
                   "attenuation": (f"{len(w.attenuation.tau_sigma)} SLS, Q_mu={w.meta.get('qmu'):g}"
                                   if w.attenuation is not None else "none"),
                   "receivers": w.receivers.n, "exact": args.exact,
                   "assembly": ["atomic", "ordered", "coloured"][args.assembly],
                   "element_order": args.element_order,
                   "att_variant": "default" if args.att_variant < 0 else args.att_variant},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batched": batched,
        "batched_viscoelastic": batched_visco,
        "coupled_batch": coupled,
        "homogeneous_box": box,
        "cpml_example": cpml,
        "adjoint": adj,
    }
    if large is not None:
        out["large"] = large
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
