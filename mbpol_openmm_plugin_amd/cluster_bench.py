#!/usr/bin/env python3
"""Batched-cluster throughput: 2048 independent water clusters of 3..64
waters (seeded, sphere radius 2.6 n^(1/3) A) evaluated together as groups of
one system, mixed precision, energies and forces of every cluster.

Prints one JSON document: batch evaluations per second and clusters per
second over `--steps` timed evaluations (positions resident on the GPU), the
same replayed from one captured HIP graph per evaluation (graph mode; its
energies must equal the eager ones bit for bit), the per-phase split from HIP
events in a further pass, and the counts of the batch.
Run:  python -m mbpol_openmm_plugin_amd.cluster_bench
"""
import argparse
import json
import time

import numpy as np

from . import api, forcefield, systems

PHASES = ('evaluate', 'lists', 'onebody', 'twobody', 'threebody', 'dispersion', 'elec_setup', 'elec_iteration',
          'elec_energy')


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--steps', type=int, default=10)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--count', type=int, default=2048)
    args = ap.parse_args(argv)
    clusters = systems.config4(count=args.count)
    sysb, pos, start = forcefield.create_cluster_batch(clusters)
    ctx = api.Context(sysb, precision='mixed')
    ctx.uploadPositions(pos)
    for _ in range(args.warmup):
        ctx.evaluateResident()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.evaluateResident()
    ms = (time.perf_counter() - t0) * 1000.0 / args.steps
    stats = ctx.getStats()
    e_eager = ctx.evaluateResident()
    ctx.setGraph(True)
    for _ in range(2):
        ctx.evaluateResident()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        e_graph = ctx.evaluateResident()
    ms_graph = (time.perf_counter() - t0) * 1000.0 / args.steps
    graph = {'batch_evals_per_s': 1000.0 / ms_graph, 'clusters_per_s': args.count * 1000.0 / ms_graph,
             'ms_per_batch': ms_graph, 'energies_equal_eager': bool(np.array_equal(e_graph, e_eager)),
             'captured_field_evaluations': ctx.getTuning('graph_steps'),
             'captures': ctx.getTuning('graph_captures'), 'replays': ctx.getTuning('graph_replays'),
             'fallbacks': ctx.getTuning('graph_fallbacks')}
    ctx.setGraph(False)
    ctx.setTiming(True)
    for _ in range(args.steps):
        ctx.evaluateResident()
    split = {}
    for name in PHASES:
        t, c = ctx.getTiming(name)
        split[name] = t / args.steps
    ctx.setTiming(False)
    sizes = np.array([len(c) for c in clusters])
    out = {
        'workload': 'config4: %d independent clusters, %d..%d waters (%d in all), mixed precision' % (
            len(clusters), sizes.min(), sizes.max(), int(sizes.sum())),
        'batch_evals_per_s': 1000.0 / ms,
        'clusters_per_s': len(clusters) * 1000.0 / ms,
        'ms_per_batch': ms,
        'steps': args.steps,
        'graph_mode': graph,
        'autotuned': {n: ctx.getTuning(n) for n in ('fork_early', 'pip_waves', 'pip3_waves',
                                                     'disp_waves', 'tune_evaluations')},
        'counts': stats,
        'phase_ms_timed_serially': split,
    }
    print(json.dumps(out, indent=1), flush=True)
    return out


if __name__ == '__main__':
    main()
