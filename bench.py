#!/usr/bin/env python3
"""MB-pol energy+force evaluations per second, 4096 waters, periodic, on one MI355X.

One "step" is one full MB-pol evaluation (1-body, 2-body, 3-body, dispersion
and PME electrostatics with the induced-dipole solve to 1e-8) of the
4096-water box (cubic 49.66 A, synthetic seeded geometry), positions already
resident in HBM.  Prints one JSON line.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# algorithmic work per unit (mul + add counted from the reference's polynomial source text,
# poly-3b-v2x.cpp / poly-2b-v6x.cpp, plus the variable/switch chain rule)
FLOPS_PER_TRIMER = 112.5e3
FLOPS_PER_DIMER = 31.0e3
PEAK_FP32_TFLOPS = 157.3


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--precision', default='mixed', choices=['mixed', 'double', 'single-pip'])
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--cpu-sample-seconds', type=float, default=20.0)
    return ap.parse_args()


def workload():
    import mbpol_openmm_plugin_amd as m
    waters, L = m.systems.config2()
    sysm = m.forcefield.create_system(4096, nonbonded_method='PME', box=L, ewald_error_tolerance=1e-5,
                                      dipole_epsilon=1e-8)
    return m, sysm, m.forcefield.water_positions_with_m(waters), L


def cpu_baseline(budget_s):
    """Single-threaded FP64 oracle on a bounded sample of the same 4096-water workload:
    every component except the electrostatics on all molecules would take minutes,
    so the sample times each component on a subset and scales it to one full
    evaluation: 1-body on all waters, 2-body on a random subset of the dimer list,
    3-body on a random subset of the trimer list, dispersion and PME
    electrostatics on the 256-water box scaled by their pair counts."""
    code = r'''
import sys, time, json, numpy as np
sys.path.insert(0, %(root)r); sys.path.insert(0, %(oracle)r)
import pyoracle as po, mbpol_openmm_plugin_amd as m
budget = %(budget)f
waters, L = m.systems.config2()
pos = m.forcefield.water_positions_with_m(waters)
idx = np.array([[4*w, 4*w+1, 4*w+2] for w in range(4096)])
box = [L]*3
res = {}
t = time.perf_counter(); po.onebody(pos, idx, True, box); res['onebody'] = time.perf_counter() - t
opos = pos[0::4]
pairs = po.dimer_list(opos, 0.65, True, box)
trip = po.trimer_list(opos, 0.45, True, box)
rng = np.random.default_rng(0)
def sample(items, fn, share):
    k = max(1, len(items))
    n = 64
    while True:
        sel = items[rng.choice(len(items), size=min(n, len(items)), replace=False)]
        t = time.perf_counter(); fn(sel); dt = time.perf_counter() - t
        if dt > share or n >= len(items):
            return dt * len(items) / len(sel), len(sel)
        n *= 4
res['twobody'], n2 = sample(pairs, lambda s: po.twobody(pos, idx, s, True, box), budget * 0.15)
res['threebody'], n3 = sample(trip, lambda s: po.threebody(pos, idx, s, True, box), budget * 0.35)
w1, L1 = m.systems.config1()
s1 = m.forcefield.create_system(256, nonbonded_method='PME', box=L1, ewald_error_tolerance=1e-5, dipole_epsilon=1e-8,
                                components=8 | 16)
p1 = m.forcefield.water_positions_with_m(w1)
t = time.perf_counter(); E, F = po.evaluate_system(s1, p1, components=16); te = time.perf_counter() - t
t = time.perf_counter(); po.evaluate_system(s1, p1, components=8); td = time.perf_counter() - t
# electrostatics real-space loops are all-pairs in the reference (N^2); dispersion pairs scale with N
res['electrostatics'] = te * (4096 / 256) ** 2
res['dispersion'] = td * (4096 / 256) ** 2
total = sum(res.values())
print(json.dumps(dict(seconds_per_eval=total, parts=res, n2_sample=n2, n3_sample=n3, ndimers=len(pairs),
                      ntrimers=len(trip))))
''' % dict(root=ROOT, oracle=os.path.join(ROOT, 'oracle'), budget=budget_s)
    env = dict(os.environ, OMP_NUM_THREADS='1')
    cmd = [sys.executable, '-c', code]
    if os.path.exists('/usr/bin/taskset'):
        cmd = ['taskset', '-c', '0'] + cmd
    out = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=600)
    if out.returncode != 0:
        return {'value': None, 'error': out.stderr[-500:]}
    r = json.loads(out.stdout.strip().splitlines()[-1])
    return {'value': 1.0 / r['seconds_per_eval'], 'unit': 'evals/s', 'cores': 1, 'kind': 'port',
            'sample': ('FP64 oracle, one core: 1-body on all 4096 waters, 2-body on %d of %d dimers, 3-body on '
                       '%d of %d trimers (scaled to the full lists), dispersion and all-pairs PME electrostatics '
                       'timed on the 256-water box and scaled by (4096/256)^2; seconds per full evaluation %.1f'
                       % (r['n2_sample'], r['ndimers'], r['n3_sample'], r['ntrimers'], r['seconds_per_eval'])),
            'parts_s': r['parts']}


def main():
    args = parse()
    if args.gpus != 1:
        print('bench.py: this build runs on one GPU only (--gpus %d requested)' % args.gpus, file=sys.stderr)
        sys.exit(2)
    import ctypes
    m, sysm, pos, L = workload()
    ctx = m.Context(sysm, precision=args.precision)
    Lib = m.library()
    h = ctx._h
    posd = np.ascontiguousarray(sysm.computeVirtualSites(pos))
    m.api._check(Lib.mbpol_context_upload_positions(h, posd.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    energies = np.zeros(6)
    ep = energies.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    for _ in range(args.warmup):
        m.api._check(Lib.mbpol_context_evaluate_resident(h, ep, ctypes.c_uint32(31)))
    m.api._check(Lib.mbpol_context_set_timing(h, 1))
    t0 = time.perf_counter()
    for _ in range(args.steps):
        m.api._check(Lib.mbpol_context_evaluate_resident(h, ep, ctypes.c_uint32(31)))
    t1 = time.perf_counter()
    stats = ctx.getStats()
    timing = {}
    for name in ('evaluate', 'lists', 'onebody', 'twobody', 'threebody', 'dispersion', 'elec_setup', 'elec_iteration',
                 'elec_energy'):
        ms, cnt = ctypes.c_double(0), ctypes.c_int64(0)
        m.api._check(Lib.mbpol_context_get_timing(h, name.encode(), ctypes.byref(ms), ctypes.byref(cnt)))
        timing[name] = (ms.value, cnt.value)
    ms_per_step = (t1 - t0) * 1000.0 / args.steps
    # roofline of the 3-body PIP kernel (the largest single kernel)
    t3_ms = timing['threebody'][0] / max(timing['threebody'][1], 1)
    achieved = FLOPS_PER_TRIMER * stats['num_trimers'] / (t3_ms * 1e-3) / 1e12
    line = {
        'metric': 'MB-pol E+F evals/s @4096 waters PBC',
        'value': 1000.0 / ms_per_step,
        'unit': 'evals/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': ms_per_step,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32+f64' if args.precision == 'mixed' else 'f64',
        'data': 'synthetic (seeded jittered 16x16x16 lattice, 4096 waters, 49.66 A box)',
        'config': {'workload': 'MB-pol 4096 waters PBC, PME tol 1e-5 (alpha %.6f, grid 121^3), rc 0.9 nm, '
                               'dipole eps 1e-8, %s precision' % (3.654826, args.precision),
                   'waters': 4096, 'box_nm': L, 'precision': args.precision},
        'roofline': {'bound': 'mfma', 'kernel': 'threebody PIP (FP32 VALU)', 'achieved': achieved,
                     'peak': PEAK_FP32_TFLOPS, 'unit': 'TFLOP/s', 'frac': achieved / PEAK_FP32_TFLOPS,
                     'traffic': None, 'flops_per_trimer': FLOPS_PER_TRIMER, 'trimers': stats['num_trimers'],
                     'avg_kernel_ms': t3_ms},
        'breakdown_ms_per_step': {k: v[0] / max(args.steps, 1) for k, v in timing.items()},
        'iterations': stats['iterations'],
        'energies_kj_mol': energies.tolist(),
    }
    if not args.no_cpu_baseline:
        line['cpu_baseline'] = cpu_baseline(args.cpu_sample_seconds)
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
