"""Graph mode (mbpol_context_set_graph; HIP extension with no counterpart in the
reference): evaluations captured in a HIP graph and replayed, against a second
context that evaluates the same sequence of configurations eagerly.

The graph runs the eager path's kernels with the same arguments, and every sum
is in fixed point or in a fixed order, so energies, forces, dipoles and the
iteration counts must be equal bit for bit.  Between replays the program does
eager work of its own on the same stream (the allocation-free electrostatic
potential, the system moments, a dipole read-back) and synchronises; every
replay is checked.

* the 256-water PME box, mixed precision, conjugate gradients started from the
  previous step's dipoles, on eight perturbed configurations;
* the same with the captured solve cut to two field evaluations
  (MBPOL_HIP_GRAPH_STEPS, a host-side setting): every replay stops before the
  stopping rule, and each evaluation must be run again eagerly -- the fallback
  path, with the same results;
* a parameter change between replays (a tighter dipole tolerance): the graph
  is captured anew;
* 64 independent clusters of the batched-cluster workload (direct-sum
  electrostatics, per-cluster energies).
"""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def mb():
    import mbpol_openmm_plugin_amd as m
    if m.device_count() < 1:
        pytest.fail('no GPU visible to libmbpol_hip.so')
    return m


def box_system(mb):
    waters, L = mb.systems.config1()
    sysm = mb.forcefield.create_system(256, nonbonded_method='PME', box=L, ewald_error_tolerance=1e-5,
                                       dipole_epsilon=1e-8)
    return sysm, mb.forcefield.water_positions_with_m(waters), L


def configurations(pos, count, seed, scale=0.002):
    rng = np.random.default_rng(seed)
    return [pos + rng.normal(scale=scale, size=pos.shape) * (k > 0) for k in range(count)]


def other_eager_work(mb, ctx, L):
    """Eager work of the program between replays, on the context's stream."""
    lib = mb.api.library()
    dp = ctypes.POINTER(ctypes.c_double)
    pts = np.ascontiguousarray(np.random.default_rng(3).uniform(0, L, (16, 3)))
    out = np.zeros(16)
    assert lib.mbpol_context_get_electrostatic_potential_noalloc(ctx._h, 16, pts.ctypes.data_as(dp),
                                                                 out.ctypes.data_as(dp)) == 0
    assert np.all(np.isfinite(out))
    masses = np.tile([15.999, 1.008, 1.008, 0.0], ctx.system.getNumParticles() // 4)
    mom = np.zeros(13)
    assert lib.mbpol_context_get_system_electrostatics_moments(ctx._h, masses.ctypes.data_as(dp),
                                                               mom.ctypes.data_as(dp)) == 0
    assert np.all(np.isfinite(mom))
    ctx.getInducedDipoles()


def evaluate(ctx, p):
    ctx.uploadPositions(p)
    e = ctx.evaluateResident()
    return e, ctx.downloadForces(), ctx.getInducedDipoles(), ctx.getStats()['iterations']


def compare(step, a, b):
    (ea, fa, da, ia), (eb, fb, db, ib) = a, b
    print('step %d: energy %.9f vs %.9f, iterations %d vs %d' % (step, ea[5], eb[5], ia, ib), flush=True)
    assert np.array_equal(ea, eb), (step, ea - eb)
    assert np.array_equal(fa, fb), (step, np.abs(fa - fb).max())
    assert np.array_equal(da, db), (step, np.abs(da - db).max())
    assert ia == ib


@pytest.mark.parametrize('steps', [None, 2])
def test_graph_replays_equal_eager(mb, steps, monkeypatch):
    if steps is not None:
        monkeypatch.setenv('MBPOL_HIP_GRAPH_STEPS', str(steps))
    sysm, pos, L = box_system(mb)
    eager = mb.Context(sysm, precision='mixed')
    graph = mb.Context(sysm, precision='mixed')
    for c in (eager, graph):
        c.setDipoleSolver('pcg')
        c.setDipoleGuess('previous')
    graph.setGraph(True)
    confs = configurations(pos, 8, seed=21)
    for k, p in enumerate(confs):
        a = evaluate(eager, p)
        b = evaluate(graph, p)
        compare(k, a, b)
        other_eager_work(mb, graph, L)
    caps, reps, falls = (graph.getTuning(n) for n in ('graph_captures', 'graph_replays', 'graph_fallbacks'))
    print('captures %d, replays %d, fallbacks %d, captured solve %d field evaluations' %
          (caps, reps, falls, graph.getTuning('graph_steps')), flush=True)
    # the first two evaluations are eager (sizing, then the first forked one); the third captures
    assert reps == len(confs) - 2
    if steps is None:
        # a solve longer than the captured one falls back and is captured anew
        assert caps <= 1 + falls and falls <= 2
    else:
        assert caps == 1 and falls == reps  # two field evaluations never reach the stopping rule
    assert eager.getTuning('graph_replays') == 0


def test_graph_recaptures_after_a_parameter_change(mb):
    sysm, pos, L = box_system(mb)
    eager = mb.Context(sysm, precision='mixed')
    graph = mb.Context(sysm, precision='mixed')
    graph.setGraph(True)
    confs = configurations(pos, 6, seed=22)
    for k, p in enumerate(confs):
        if k == 4:
            for c in (eager, graph):
                c.setDipoleConvergence(1e-10, 200)
        compare(k, evaluate(eager, p), evaluate(graph, p))
        other_eager_work(mb, graph, L)
    caps, reps, falls = (graph.getTuning(n) for n in ('graph_captures', 'graph_replays', 'graph_fallbacks'))
    print('captures %d, replays %d, fallbacks %d' % (caps, reps, falls), flush=True)
    # captured anew after the change; the tighter solve is longer than the captured one
    # (the last solve's + 1), so its first replay falls back and the next call captures again
    assert reps == len(confs) - 2 and 2 <= caps <= 2 + falls


def test_graph_batched_clusters(mb):
    clusters = mb.systems.config4(count=64)
    sysb, pos, _ = mb.forcefield.create_cluster_batch(clusters)
    eager = mb.Context(sysb, precision='mixed')
    graph = mb.Context(sysb, precision='mixed')
    graph.setGraph(True)
    for k, p in enumerate(configurations(pos, 5, seed=23)):
        a = evaluate(eager, p)
        b = evaluate(graph, p)
        compare(k, a, b)
        assert np.array_equal(eager.getGroupEnergies(), graph.getGroupEnergies())
        graph.getInducedDipoles()
    reps = graph.getTuning('graph_replays')
    print('replays', reps, flush=True)
    assert reps == 3
