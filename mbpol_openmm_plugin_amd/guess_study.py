#!/usr/bin/env python3
"""Learned dipole guess study on the 4096-water box.

Protocol:
  1. Langevin dynamics at 300 K (0.5 fs steps, friction 5/ps) from the
     seeded 4096-water lattice, the solver started from the previous step's
     dipoles; an equilibration stretch, then `--snapshots` snapshots
     `--stride` steps apart, each storing the network's inputs and the
     converged dipoles (relative to alpha * E_fixed) on the GPU.
  2. Training of the 48 -> 64 -> 64 -> 3 network on them (Adam, BF16 MFMA).
  3. On `--test` later snapshots of the same trajectory, for each starting
     guess -- alpha * E_fixed (the reference's), zero, the previous MD step's
     dipoles, the trained network -- and each solver (SOR as the reference,
     PCG): the induced-field evaluations needed for the reference's
     convergence measure to fall below 1e-8, and the fewest evaluations after
     which the total energy is within 1e-6 (relative) of the converged one
     (found by solving to a ladder of looser targets).

Prints one JSON document.  Run:  python -m mbpol_openmm_plugin_amd.guess_study
"""
import argparse
import json
import time

import numpy as np

from . import api, forcefield, systems


LADDER = [3e-1, 1e-1, 3e-2, 1e-2, 3e-3, 1e-3, 3e-4, 1e-4, 3e-5, 1e-5, 3e-6, 1e-6, 1e-7, 1e-8]


def solve_counts(ctx, gm, solver, mode, e_ref, prev):
    """(field evaluations to epsilon 1e-8, fewest evaluations within 1e-6 of e_ref) from one guess."""
    ctx.setDipoleSolver(solver)

    def set_mode():
        if mode == 'learned':
            gm.attach(True)
        elif mode == 'previous':
            ctx.setDipoleGuess('user', prev)
        else:
            ctx.setDipoleGuess(mode)

    set_mode()
    ctx.setDipoleConvergence(1e-8, 500)
    ctx.evaluateResident()
    it_full = ctx.getStats()['iterations']
    it_e6 = None
    for tol in LADDER:
        set_mode()
        ctx.setDipoleConvergence(tol, 500)
        e = ctx.evaluateResident()
        if abs(e[5] - e_ref) <= 1e-6 * abs(e_ref):
            it_e6 = ctx.getStats()['iterations']
            break
    ctx.setDipoleConvergence(1e-8, 500)
    return it_full, it_e6


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--snapshots', type=int, default=2000)
    ap.add_argument('--stride', type=int, default=5)
    ap.add_argument('--equilibrate', type=int, default=2000)
    ap.add_argument('--epochs', type=int, default=20)
    ap.add_argument('--batch', type=int, default=4096)
    ap.add_argument('--lr', type=float, default=1e-3)
    ap.add_argument('--test', type=int, default=8)
    ap.add_argument('--waters', type=int, default=4096, choices=[256, 4096])
    args = ap.parse_args(argv)
    t0 = time.perf_counter()
    if args.waters == 4096:
        waters, L = systems.config2()
    else:
        waters, L = systems.config1()
    sysm = forcefield.create_system(args.waters, nonbonded_method='PME', box=L, ewald_error_tolerance=1e-5,
                                    dipole_epsilon=1e-8)
    ctx = api.Context(sysm, precision='mixed')
    ctx.setPositions(forcefield.water_positions_with_m(waters))
    ctx.setDipoleGuess('previous')
    md = api.LangevinIntegrator(ctx, 300.0, 5.0, 0.0005, seed=11)
    e, k = md.step(args.equilibrate)
    nat = 3 * args.waters
    log = {'equilibration': {'steps': args.equilibrate, 'potential_kj_mol': e[5],
                             'temperature_K': k / (1.5 * nat * 0.0083144626)}}
    gm = api.GuessModel(ctx, max_snapshots=args.snapshots, seed=3)
    md_iters = []
    t1 = time.perf_counter()
    for _ in range(args.snapshots):
        md.step(args.stride)
        md_iters.append(ctx.getStats()['iterations'])
        gm.addSnapshot()
    log['trajectory'] = {'snapshots': args.snapshots, 'stride_steps': args.stride, 'seconds': time.perf_counter() - t1,
                         'pcg_evaluations_per_step_from_previous': float(np.mean(md_iters))}
    t1 = time.perf_counter()
    losses = gm.train(args.epochs, batch=args.batch, learning_rate=args.lr)
    log['training'] = {'epochs': args.epochs, 'samples': args.snapshots * nat, 'batch': args.batch,
                       'seconds': time.perf_counter() - t1, 'loss_first_epoch': float(losses[0]),
                       'loss_last_epoch': float(losses[-1])}
    modes = ('alpha_e', 'zero', 'previous', 'learned')
    res = {(s, m): [] for s in ('sor', 'pcg') for m in modes}
    for _ in range(args.test):
        md.step(args.stride - 1)
        prev = ctx.getInducedDipoles()  # converged dipoles one step before the test configuration
        md.step(1)
        ctx.setDipoleSolver('pcg')
        ctx.setDipoleGuess('alpha_e')
        ctx.setDipoleConvergence(1e-11, 500)
        e_ref = ctx.evaluateResident()[5]
        ctx.setDipoleConvergence(1e-8, 500)
        for solver in ('sor', 'pcg'):
            for m in modes:
                res[(solver, m)].append(solve_counts(ctx, gm, solver, m, e_ref, prev))
        ctx.setDipoleSolver('pcg')
        ctx.setDipoleGuess('previous')
    out = {}
    for (s, m), v in res.items():
        full = [a for a, _ in v]
        e6 = [b for _, b in v if b is not None]
        out['%s/%s' % (s, m)] = {'evaluations_to_eps_1e-8': float(np.mean(full)),
                                 'evaluations_to_1e-6_rel_E': float(np.mean(e6)) if e6 else None,
                                 'samples': len(v)}
    log['iterations'] = out
    log['seconds_total'] = time.perf_counter() - t0
    print(json.dumps(log, indent=1), flush=True)
    return log


if __name__ == '__main__':
    main()
