#!/usr/bin/env python3
"""MB-pol energy+force evaluations per second, 4096 waters, periodic, one MI355X.

One step is one full MB-pol evaluation -- 1-body, 2-body and 3-body PIPs with
their O-O lists, damped dispersion, and PME electrostatics with the Thole
induced-dipole solve to epsilon 1e-8 -- of the 4096-water box (cubic 49.66 A,
synthetic seeded geometry), positions resident in HBM before the timed region.

Prints one JSON line:
  value / ms_per_step  wall-clock throughput of the default (mixed) path over
                       `--steps` evaluations, timing switched off;
  fp64_path            the same for the FP64 path, and the error of the mixed
                       path against it per component;
  roofline             the kernel with the largest share of the evaluation,
                       from HIP events recorded on the library's stream during
                       a second pass of `--steps` evaluations; its algorithmic
                       bytes or flops per launch (models below) / average
                       launch time, against the MI355X peak;
  cpu_baseline         the FP64 oracle on one host core over a bounded sample
                       of the same evaluation (see cpu_baseline()).
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

# MI355X peaks: HBM3E bandwidth, FP64 vector/matrix rate
PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6
# algorithmic work per unit: mul + add of the reference polynomials counted from
# their source (poly-3b-v2x.cpp: 111,817; poly-2b-v6x.cpp: 30,600) plus the
# variables, switches and chain rule around them
FLOPS_PER_TRIMER = 112.5e3
FLOPS_PER_DIMER = 31.0e3
# HBM bytes per launch from rocprofv3 FETCH_SIZE + WRITE_SIZE counters on this
# workload, calibrated against the runtime's plain copy kernel in the same runs
# (FETCH_SIZE x 1.885, WRITE_SIZE x 1.000; the runs are summarised under
# profiles/); kernels not listed report null
MEASURED_TRAFFIC = {
    'elec_iterate': 85816165,
    'spread_fft_z': 84248140,
    'fft_z_r2c': 28177550,
    'fft_y': 25651329,
    'fft_x_conv': 38354591,
    'fft_z_c2r': 25337943,
    'threebody': 114310400,
    'twobody': 58527178,
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=100)
    ap.add_argument('--warmup', type=int, default=20)
    ap.add_argument('--precision', default='mixed', choices=['mixed', 'double', 'single-pip'])
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--no-fp64-path', action='store_true')
    ap.add_argument('--no-dipole-iterations', action='store_true')
    ap.add_argument('--no-graph', action='store_true', help='skip the graph-mode leg')
    ap.add_argument('--cpu-config1-evals', type=int, default=1,
                    help='complete oracle evaluations of the 256-water box timed beside the sampled estimate')
    return ap.parse_args()


def workload(m):
    waters, L = m.systems.config2()
    sysm = m.forcefield.create_system(4096, nonbonded_method='PME', box=L, ewald_error_tolerance=1e-5,
                                      dipole_epsilon=1e-8)
    return sysm, m.forcefield.water_positions_with_m(waters), L


def kernel_models(stats, grid, nsites, npol):
    """Algorithmic work per launch of each individually timed kernel:
    (kind, amount, bound, peak, unit)."""
    K0, K1, K2 = grid
    ngrid, nhalf = K0 * K1 * K2, K0 * K1 * (K2 // 2 + 1)
    return {
        'threebody': ('flops', FLOPS_PER_TRIMER * stats['num_trimers'], 'valu', PEAK_FP64_TFLOPS, 'TFLOP/s'),
        'twobody': ('flops', FLOPS_PER_DIMER * stats['num_dimers'], 'valu', PEAK_FP64_TFLOPS, 'TFLOP/s'),
        # the induced-field kernel (the conjugate-gradient matrix-vector product; the SOR
        # step in the FP64 path): rows of the polarisable sites (int index + 2 FP64
        # prefactors per entry), the potential grid once, positions and dipoles of every
        # site once, the site's own vector / polarisability / row length and its result
        'elec_iterate': ('bytes', stats['num_pol_pairs'] * 20 + ngrid * 8 + nsites * 48 + npol * (24 + 24 + 8 + 4 + 4),
                         'hbm', PEAK_HBM_GBS, 'GB/s'),
        # row starts, entries (site + 3 FP64 coefficients), dipoles, grid write
        'pme_spread': ('bytes', (ngrid + 1) * 4 + npol * 125 * 28 + npol * 24 + ngrid * 8, 'hbm', PEAK_HBM_GBS, 'GB/s'),
        'fft_z_r2c': ('bytes', ngrid * 8 + nhalf * 16, 'hbm', PEAK_HBM_GBS, 'GB/s'),
        # the dipole spreading fused with the z pass: row starts, entries, dipoles, half spectrum write
        'spread_fft_z': ('bytes', (ngrid + 1) * 4 + npol * 125 * 28 + npol * 24 + nhalf * 16, 'hbm', PEAK_HBM_GBS,
                         'GB/s'),
        'fft_y': ('bytes', nhalf * 32, 'hbm', PEAK_HBM_GBS, 'GB/s'),
        'fft_x_conv': ('bytes', nhalf * 40, 'hbm', PEAK_HBM_GBS, 'GB/s'),
        'fft_z_c2r': ('bytes', nhalf * 16 + ngrid * 8, 'hbm', PEAK_HBM_GBS, 'GB/s'),
        # z and y passes fused per x plane (when a plane fits in LDS): the real grid and
        # the half spectrum once each
        'fft_plane_fwd': ('bytes', ngrid * 8 + nhalf * 16, 'hbm', PEAK_HBM_GBS, 'GB/s'),
        'fft_plane_inv': ('bytes', nhalf * 16 + ngrid * 8, 'hbm', PEAK_HBM_GBS, 'GB/s'),
    }


PHASES = ('evaluate', 'lists', 'onebody', 'twobody', 'threebody', 'dispersion', 'elec_setup', 'elec_iteration',
          'elec_energy')


def run_steps(ctx, steps, per_eval=None):
    """Mean wall time per evaluation (ms) over `steps` evaluations and the last energies;
    with a list in `per_eval`, each evaluation's own wall time (ms) is appended to it."""
    t0 = time.perf_counter()
    for _ in range(steps):
        t = time.perf_counter()
        e = ctx.evaluateResident()  # returns after the energies are on the host
        if per_eval is not None:
            per_eval.append((time.perf_counter() - t) * 1000.0)
    return (time.perf_counter() - t0) * 1000.0 / steps, e


LADDER = [3e-1, 1e-1, 3e-2, 1e-2, 3e-3, 1e-3, 3e-4, 1e-4, 3e-5, 1e-5, 3e-6, 1e-6, 1e-7, 1e-8]


def dipole_iterations(ctx):
    """Induced-field evaluations of the dipole solve on the benchmark box, from
    alpha * E_fixed (the reference's start): to the convergence target 1e-8,
    and the fewest after which the total energy is within 1e-6 (relative) of
    the converged one (a ladder of looser targets), for the reference's SOR
    ("plain") and for this library's default conjugate-gradient solve."""
    ctx.setDipoleGuess('alpha_e')
    ctx.setDipoleSolver('pcg')
    ctx.setDipoleConvergence(1e-11, 500)
    e_ref = ctx.evaluateResident()[5]
    out = {}
    for solver in ('sor', 'pcg'):
        ctx.setDipoleSolver(solver)
        ctx.setDipoleConvergence(1e-8, 500)
        ctx.evaluateResident()
        full = ctx.getStats()['iterations']
        e6 = None
        for tol in LADDER:
            ctx.setDipoleConvergence(tol, 500)
            e = ctx.evaluateResident()
            if abs(e[5] - e_ref) <= 1e-6 * abs(e_ref):
                e6 = ctx.getStats()['iterations']
                break
        out[('plain_sor' if solver == 'sor' else 'pcg') + '_from_alpha_e'] = {
            'evaluations_to_eps_1e-8': full, 'evaluations_to_1e-6_rel_E': e6}
    ctx.setDipoleConvergence(1e-8, 500)
    ctx.setDipoleSolver('pcg')
    ctx.evaluateResident()  # leave the converged state of the default path behind
    out['learned_guess_study'] = 'profiles/dipole_guess_study_4096w.md (300 K trajectory of this box)'
    return out


def cpu_baseline(gpu_iterations):
    """FP64 oracle on one core, on a bounded sample of the same evaluation:
    1-body, the O-O lists and dispersion in full; 2-body and 3-body on the first
    2000 entries of their lists, scaled to the full lists; electrostatics run
    once with a single dipole iteration, its setup + energy time plus
    `gpu_iterations` (the SOR iteration count of this configuration, as the
    reference iterates) times the measured iteration time.  About 25 s."""
    code = r'''
import sys, time, json, numpy as np
sys.path.insert(0, %(root)r); sys.path.insert(0, %(oracle)r)
import pyoracle as po, mbpol_openmm_plugin_amd as m
waters, L = m.systems.config2()
pos = m.forcefield.water_positions_with_m(waters)
idx = np.array([[4*w, 4*w+1, 4*w+2] for w in range(4096)]); box = [L]*3
res = {}
t = time.perf_counter(); po.onebody(pos, idx, True, box); res['onebody'] = time.perf_counter() - t
t = time.perf_counter(); opos = pos[0::4]; pairs = po.dimer_list(opos, 0.65, True, box)
trip = po.trimer_list(opos, 0.45, True, box); res['lists'] = time.perf_counter() - t
k2 = min(2000, len(pairs)); k3 = min(2000, len(trip))
t = time.perf_counter(); po.twobody(pos, idx, pairs[:k2], True, box); res['twobody'] = (time.perf_counter() - t) * len(pairs) / k2
t = time.perf_counter(); po.threebody(pos, idx, trip[:k3], True, box); res['threebody'] = (time.perf_counter() - t) * len(trip) / k3
sd = m.forcefield.create_system(4096, nonbonded_method='PME', box=L, components=8)
t = time.perf_counter(); po.evaluate_system(sd, pos, components=8); res['dispersion'] = time.perf_counter() - t
se = m.forcefield.create_system(4096, nonbonded_method='PME', box=L, ewald_error_tolerance=1e-5, dipole_epsilon=1e9,
                                components=16)
E, F, Ls = po.evaluate_system(se, pos, components=16, return_lists=True)
r = Ls['elec']
res['electrostatics'] = r['setup_s'] + %(iters)d * r['iteration_s'] + r['energy_s']
print(json.dumps(dict(parts=res, n2=k2, n3=k3, ndimers=len(pairs), ntrimers=len(trip),
                      elec=dict(setup_s=r['setup_s'], iteration_s=r['iteration_s'], energy_s=r['energy_s']))))
''' % dict(root=ROOT, oracle=os.path.join(ROOT, 'oracle'), iters=gpu_iterations)
    env = dict(os.environ, OMP_NUM_THREADS='1')
    cmd = [sys.executable, '-c', code]
    if os.path.exists('/usr/bin/taskset'):
        cmd = ['/usr/bin/taskset', '-c', '0'] + cmd
    out = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=900)
    if out.returncode != 0:
        return {'value': None, 'unit': 'evals/s', 'cores': 1, 'kind': 'port', 'error': out.stderr[-400:]}
    r = json.loads(out.stdout.strip().splitlines()[-1])
    total = sum(r['parts'].values())
    return {'value': 1.0 / total, 'unit': 'evals/s', 'cores': 1, 'kind': 'port',
            'sample': ('FP64 oracle (oracle/, single thread, taskset core 0) on the same 4096-water evaluation: '
                       '1-body, O-O lists and dispersion in full; 2-body on %d of %d dimers and 3-body on %d of %d '
                       'trimers, scaled to the full lists; electrostatics run with one dipole iteration, '
                       'setup + energy + %d x the measured iteration time. Estimated seconds per evaluation: %.1f'
                       % (r['n2'], r['ndimers'], r['n3'], r['ntrimers'], gpu_iterations, total)),
            'seconds_per_eval_by_component': r['parts'], 'electrostatics_phases_s': r['elec']}


def cpu_full_config1(evals):
    """The FP64 oracle's complete evaluation of the 256-water box (all five
    components, PME, SOR from alpha E_fixed as the reference iterates, its O-O
    lists built in full), `evals` times on one core (taskset core 0): median
    seconds per evaluation, with the CPU model."""
    code = r'''
import sys, time, json, statistics
sys.path.insert(0, %(root)r); sys.path.insert(0, %(oracle)r)
import pyoracle as po, mbpol_openmm_plugin_amd as m
waters, L = m.systems.config1()
s = m.forcefield.create_system(256, nonbonded_method='PME', box=L, ewald_error_tolerance=1e-5, dipole_epsilon=1e-8)
pos = m.forcefield.water_positions_with_m(waters)
ts = []
for _ in range(%(evals)d):
    t = time.perf_counter(); E, F, Ls = po.evaluate_system(s, pos, return_lists=True); ts.append(time.perf_counter() - t)
model = ''
try:
    for ln in open('/proc/cpuinfo'):
        if ln.startswith('model name'):
            model = ln.split(':', 1)[1].strip(); break
except OSError:
    pass
print(json.dumps(dict(seconds=ts, iterations=Ls['elec']['iterations'], cpu_model=model)))
''' % dict(root=ROOT, oracle=os.path.join(ROOT, 'oracle'), evals=evals)
    cmd = [sys.executable, '-c', code]
    if os.path.exists('/usr/bin/taskset'):
        cmd = ['/usr/bin/taskset', '-c', '0'] + cmd
    out = subprocess.run(cmd, capture_output=True, text=True, env=dict(os.environ, OMP_NUM_THREADS='1'),
                         timeout=120 * evals + 60)
    if out.returncode != 0:
        return {'error': out.stderr[-400:]}
    r = json.loads(out.stdout.strip().splitlines()[-1])
    med = float(np.median(r['seconds']))
    return {'workload': '256 waters, PBC 19.71 A, PME tol 1e-5, all five components, SOR from alpha E_fixed',
            'evaluations': len(r['seconds']), 'seconds_per_eval_median': med, 'evals_per_s': 1.0 / med,
            'seconds': r['seconds'], 'dipole_iterations': r['iterations'], 'cores': 1, 'cpu_model': r['cpu_model'],
            'oracle_mode': 'oracle/: the reference loop structure (all-pairs electrostatics with the cutoff '
                           'test inside, PME reciprocal part, SOR omega 0.55), polynomials term by term'}


def main():
    args = parse()
    if args.gpus != 1:
        print('bench.py: this build runs on one GPU only (--gpus %d requested)' % args.gpus, file=sys.stderr)
        sys.exit(2)
    import mbpol_openmm_plugin_amd as m
    sysm, pos, L = workload(m)
    ctx = m.Context(sysm, precision=args.precision)
    ctx.uploadPositions(pos)
    for _ in range(args.warmup):
        ctx.evaluateResident()
    # timed region: throughput, no events
    per_eval = []
    ms_per_step, energies = run_steps(ctx, args.steps, per_eval)
    stats = ctx.getStats()
    # second timed region with HIP events on the library's stream
    ctx.setTiming(True)
    run_steps(ctx, args.steps)
    timing = {}
    for name in PHASES + tuple(kernel_models(stats, (2, 2, 2), 1, 1)):
        ms, cnt = ctx.getTiming(name)
        timing[name] = {'ms_per_step': ms / args.steps, 'launches_per_step': cnt / args.steps,
                        'avg_us': 1000.0 * ms / cnt if cnt else None}
    ctx.setTiming(False)
    el = [f for f in sysm._forces if type(f).__name__ == 'MBPolElectrostaticsForce'][0]
    alpha, grid = el.getAEwald(), el.getPmeGridDimensions()
    if grid[0] == 0:  # automatic choice, as the library makes it
        import math
        tol, rc = el.getEwaldErrorTolerance(), el.getCutoffDistance()
        alpha = math.sqrt(-math.log(2 * tol)) / rc
        grid = [max(int(math.ceil(2 * alpha * L / (3 * tol ** 0.2))), 6)] * 3
    nsites = sysm.getNumParticles()
    npol = sum(1 for i in range(el.getNumElectrostatics()) if el.getElectrostaticsParameters(i)[4] != 0.0)
    models = kernel_models(stats, grid, nsites, npol)
    kernels = {}
    for name, (kind, amount, bound, peak, unit) in models.items():
        t = timing[name]
        if not t['avg_us']:
            continue
        per_launch = amount  # every model is per launch
        rate = per_launch / (t['avg_us'] * 1e-6) / (1e12 if kind == 'flops' else 1e9)
        kernels[name] = {'ms_per_step': t['ms_per_step'], 'launches_per_step': t['launches_per_step'],
                         'avg_us': t['avg_us'], kind + '_per_launch': per_launch, 'achieved': rate,
                         'unit': unit, 'peak': peak, 'frac': rate / peak, 'bound': bound,
                         'traffic': MEASURED_TRAFFIC.get(name)}
    dom = max(kernels, key=lambda k: kernels[k]['ms_per_step'])
    d = kernels[dom]
    roofline = {'bound': d['bound'], 'achieved': d['achieved'], 'peak': d['peak'], 'unit': d['unit'],
                'frac': d['frac'], 'traffic': MEASURED_TRAFFIC.get(dom), 'kernel': dom,
                'avg_kernel_us': d['avg_us'], 'launches_per_step': d['launches_per_step'],
                'share_of_step': d['ms_per_step'] / ms_per_step,
                'algorithmic_' + ('flops' if 'flops_per_launch' in d else 'bytes') + '_per_launch':
                    d.get('flops_per_launch', d.get('bytes_per_launch')),
                'all_kernels': kernels}
    if dom == 'threebody':
        # The evaluation runs the 3-body grid capped (waves per SIMD, tuned) so that it
        # leaves the electrostatics room; the same kernel at full occupancy (one block per
        # 64 trimers, two waves per SIMD), timed alone on a second context
        roofline['production_waves_per_simd'] = ctx.getTuning('pip3_waves')
        saved = {k: os.environ.get(k) for k in ('MBPOL_HIP_3B_WAVES', 'MBPOL_HIP_TUNE')}
        os.environ['MBPOL_HIP_3B_WAVES'] = '0'
        os.environ['MBPOL_HIP_TUNE'] = '0'
        try:
            full = m.Context(sysm, precision=args.precision)
            full.uploadPositions(pos)
            for _ in range(2):
                full.evaluateResident()
            full.setTiming(True)
            run_steps(full, 5)
            ms3, cnt3 = full.getTiming('threebody')
            full.setTiming(False)
            del full
        finally:
            for k, v in saved.items():
                if v is None:
                    os.environ.pop(k, None)
                else:
                    os.environ[k] = v
        us3 = 1000.0 * ms3 / cnt3
        rate3 = d['flops_per_launch'] / (us3 * 1e-6) / 1e12
        roofline['full_occupancy'] = {'avg_kernel_us': us3, 'achieved': rate3, 'frac': rate3 / d['peak'],
                                      'waves_per_simd': 'one block per 64 trimers (2 resident per SIMD)'}
    line = {
        'metric': 'MB-pol E+F evals/s @4096 waters PBC; dipole iters vs plain at 1e-6 rel E',
        'value': 1000.0 / ms_per_step,
        'unit': 'evals/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': ms_per_step,
        'per_eval_ms': {'median': float(np.median(per_eval)), 'p10': float(np.percentile(per_eval, 10)),
                        'p90': float(np.percentile(per_eval, 90))},
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': {'mixed': 'f64 (f32 dispersion pair terms)', 'double': 'f64',
                  'single-pip': 'f32 polynomials, f64 elsewhere'}[args.precision],
        'data': 'synthetic: seeded jittered 16x16x16 lattice, 4096 waters, cubic 49.66 A box',
        'config': {'workload': 'MB-pol, 4096 waters, PBC 49.66 A, PME tol 1e-5 (alpha %.6f nm^-1, grid %dx%dx%d), '
                               'rc 0.9 nm, 2B rc 0.65 nm, 3B rc 0.45 nm, dipole epsilon 1e-8, %s precision'
                               % (alpha, grid[0], grid[1], grid[2], args.precision),
                   'waters': 4096, 'box_nm': L, 'precision': args.precision},
        'iterations': stats['iterations'],
        'autotuned': {n: ctx.getTuning(n) for n in ('site_lanes', 'fork_early', 'pip_waves',
                                                     'pip3_waves', 'disp_waves', 'tune_evaluations')},
        'dipole_solver': 'sor' if args.precision == 'double' else 'pcg',
        'counts': {k: stats[k] for k in ('num_dimers', 'num_trimers', 'num_elec_pairs', 'num_pol_pairs')},
        'energies_kj_mol': dict(zip(['onebody', 'twobody', 'threebody', 'dispersion', 'electrostatics', 'total'],
                                    energies.tolist())),
        'breakdown_ms_per_step': {k: timing[k]['ms_per_step'] for k in PHASES},
        'roofline': roofline,
    }
    # the same evaluations replayed from a HIP graph (mbpol_context_set_graph): captured
    # once, checked after every replay; the energies must equal the eager ones bit for bit
    if not args.no_graph:
        ctx.setGraph(True)
        for _ in range(max(args.warmup, 2)):
            ctx.evaluateResident()
        ms_graph, e_graph = run_steps(ctx, args.steps)
        line['graph_path'] = {'value': 1000.0 / ms_graph, 'ms_per_step': ms_graph,
                              'energies_equal_eager': bool(np.array_equal(e_graph, energies)),
                              'captured_field_evaluations': ctx.getTuning('graph_steps'),
                              'captures': ctx.getTuning('graph_captures'),
                              'replays': ctx.getTuning('graph_replays'),
                              'fallbacks': ctx.getTuning('graph_fallbacks')}
        ctx.setGraph(False)
    # the same evaluation through the host-buffer entry point mbpol_context_evaluate
    # (positions copied in, virtual sites placed on the device, forces copied out)
    import ctypes
    lib = m.library()
    dp = ctypes.POINTER(ctypes.c_double)
    hpos = np.ascontiguousarray(pos, dtype=np.float64)
    hfrc, hen = np.zeros_like(hpos), np.zeros(6)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        m.api._check(lib.mbpol_context_evaluate(ctx._h, hpos.ctypes.data_as(dp), hfrc.ctypes.data_as(dp),
                                                hen.ctypes.data_as(dp), ctypes.c_uint32(31)))
    ms_host = (time.perf_counter() - t0) * 1000.0 / args.steps
    line['host_buffer_path'] = {'value': 1000.0 / ms_host, 'ms_per_step': ms_host,
                                'note': 'positions copied in and forces copied out on every evaluation'}
    if args.precision != 'double' and not args.no_dipole_iterations:
        line['dipole_iterations'] = dipole_iterations(ctx)
    sor_iterations = stats['iterations']
    if not args.no_fp64_path and args.precision != 'double':
        f_mixed = ctx.downloadForces()
        ctx64 = m.Context(sysm, precision='double')
        ctx64.uploadPositions(pos)
        for _ in range(args.warmup):
            ctx64.evaluateResident()
        ms64, e64 = run_steps(ctx64, args.steps)
        f64 = ctx64.downloadForces()
        sor_iterations = ctx64.getStats()['iterations']
        tot = abs(e64[5])
        line['fp64_path'] = {
            'value': 1000.0 / ms64, 'ms_per_step': ms64, 'dipole_solver': 'sor', 'iterations': sor_iterations,
            'speedup_of_default_path': ms64 / ms_per_step,
            'mixed_vs_fp64_rel_error_of_total_by_component': dict(
                zip(['onebody', 'twobody', 'threebody', 'dispersion', 'electrostatics', 'total'],
                    (np.abs(energies - e64) / tot).tolist())),
            'mixed_vs_fp64_rel_rms_force': float(np.sqrt(((f_mixed - f64) ** 2).mean() / (f64 ** 2).mean()))}
        # FP64 everywhere with the conjugate-gradient solve: separates the solver's gain from precision's
        ctx64.setDipoleSolver('pcg')
        for _ in range(args.warmup):
            ctx64.evaluateResident()
        ms64p, e64p = run_steps(ctx64, args.steps)
        line['fp64_pcg_path'] = {
            'value': 1000.0 / ms64p, 'ms_per_step': ms64p, 'dipole_solver': 'pcg',
            'iterations': ctx64.getStats()['iterations'],
            'speedup_over_fp64_sor': ms64 / ms64p, 'speedup_of_default_path_over_it': ms64p / ms_per_step,
            'rel_error_of_total_vs_fp64_sor': float(abs(e64p[5] - e64[5]) / tot)}
        del ctx64
    if not args.no_cpu_baseline:
        if args.precision != 'double' and args.no_fp64_path:  # iterations of the reference's SOR solve
            if 'dipole_iterations' in line:
                sor_iterations = line['dipole_iterations']['plain_sor_from_alpha_e']['evaluations_to_eps_1e-8']
            else:
                ctx.setDipoleSolver('sor')
                ctx.evaluateResident()
                sor_iterations = ctx.getStats()['iterations']
                ctx.setDipoleSolver('pcg')
        line['cpu_baseline'] = cpu_baseline(sor_iterations)
        if args.cpu_config1_evals > 0:
            line['cpu_baseline']['full_config1'] = cpu_full_config1(args.cpu_config1_evals)
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
