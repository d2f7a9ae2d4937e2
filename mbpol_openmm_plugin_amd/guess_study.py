#!/usr/bin/env python3
"""Learned dipole guess study on the 4096-water box.

Protocol:
  1. Langevin dynamics at 300 K (0.5 fs steps, friction 5/ps) from the
     seeded 4096-water lattice, the solver started from the previous step's
     dipoles; an equilibration stretch, then `--snapshots` snapshots
     `--stride` steps apart, each storing the network's inputs and the
     converged dipoles (relative to alpha * E_fixed) on the GPU.
  2. Two networks (48 -> 64 -> 64 -> 3, Adam, BF16 MFMA) trained on the
     first `--snapshots` - `--holdout` of them: one correcting alpha * E_fixed
     (kind 'alpha_e'), one correcting the previous MD step's converged dipoles
     (kind 'previous', anchored on the change of alpha * E_fixed since that
     step).  Their mean squared error on the `--holdout` later snapshots,
     which no training step saw, is reported beside the error of a zero
     correction.
  3. On `--test` later configurations of the same trajectory, for each
     starting guess -- alpha * E_fixed (the reference's), zero, the previous
     MD step's dipoles, the two trained networks -- the start's error
     |mu_0 - mu*| / |mu*| against the dipoles mu* converged to 1e-11, the
     epsilon of the start (the reference's convergence measure after one SOR
     step from it, a scaled residual rather than an error), and
     for each solver (SOR as the reference, PCG) the induced-field
     evaluations needed for epsilon to fall below 1e-8 and the fewest
     evaluations after which the total energy is within 1e-6 (relative) of
     the converged one (found by solving to a ladder of looser targets).

Prints one JSON document.  Run:  python -m mbpol_openmm_plugin_amd.guess_study
"""
import argparse
import json
import time

import numpy as np

from . import api, forcefield, systems


LADDER = [3e-1, 1e-1, 3e-2, 1e-2, 3e-3, 1e-3, 3e-4, 1e-4, 3e-5, 1e-5, 3e-6, 1e-6, 1e-7, 1e-8]


def set_start(ctx, mode, guesses):
    if mode in guesses:
        ctx.setDipoleGuess('user', guesses[mode])
    else:
        ctx.setDipoleGuess(mode)


def start_epsilon(ctx, mode, guesses):
    """The reference's convergence measure after one SOR step from the start."""
    ctx.setDipoleSolver('sor')
    set_start(ctx, mode, guesses)
    ctx.setDipoleConvergence(1e3, 1)
    ctx.evaluateResident()
    eps = ctx.getStats()['epsilon']
    ctx.setDipoleConvergence(1e-8, 500)
    return eps


def solve_counts(ctx, solver, mode, guesses, e_ref):
    """(field evaluations to epsilon 1e-8, fewest evaluations within 1e-6 of e_ref, converged energy)."""
    ctx.setDipoleSolver(solver)
    set_start(ctx, mode, guesses)
    ctx.setDipoleConvergence(1e-8, 500)
    e_full = ctx.evaluateResident()[5]
    it_full = ctx.getStats()['iterations']
    it_e6 = None
    for tol in LADDER:
        set_start(ctx, mode, guesses)
        ctx.setDipoleConvergence(tol, 500)
        e = ctx.evaluateResident()
        if abs(e[5] - e_ref) <= 1e-6 * abs(e_ref):
            it_e6 = ctx.getStats()['iterations']
            break
    ctx.setDipoleConvergence(1e-8, 500)
    return it_full, it_e6, e_full


def zero_network(ctx):
    """A network with every parameter zero: its correction is zero, so its guess
    is alpha * E_fixed itself, through the same apply path as the trained ones."""
    gz = api.GuessModel(ctx, max_snapshots=1, seed=1)
    norm = np.concatenate([np.zeros(api.GUESS_FEATURES), np.ones(api.GUESS_FEATURES), [1.0]]).astype(np.float32)
    gz.setParameters(np.zeros(api.GUESS_PARAMS, dtype=np.float32), norm)
    return gz


def rel_error(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def zero_loss(gm, first, count):
    """Mean squared error of a zero correction on stored snapshots [first, first + count)."""
    scale = float(gm.getParameters()[1][-1])
    tot, n = 0.0, 0
    for k in range(first, first + count):
        _, t = gm.getSnapshot(k)
        tot += float(np.sum(t.astype(np.float64) ** 2))
        n += len(t)
    return tot / n / scale ** 2


MODES = ('alpha_e', 'zero', 'previous', 'learned', 'learned_previous')


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--snapshots', type=int, default=2000)
    ap.add_argument('--holdout', type=int, default=200)
    ap.add_argument('--stride', type=int, default=5)
    ap.add_argument('--equilibrate', type=int, default=2000)
    ap.add_argument('--epochs', type=int, default=20)
    ap.add_argument('--batch', type=int, default=4096)
    ap.add_argument('--lr', type=float, default=1e-3)
    ap.add_argument('--test', type=int, default=8)
    ap.add_argument('--waters', type=int, default=4096, choices=[256, 4096])
    ap.add_argument('--quiet', action='store_true')
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
    nets = {'learned': api.GuessModel(ctx, max_snapshots=args.snapshots, seed=3, kind='alpha_e'),
            'learned_previous': api.GuessModel(ctx, max_snapshots=args.snapshots, seed=4, kind='previous')}
    md_iters = []
    t1 = time.perf_counter()
    for _ in range(args.snapshots):
        md.step(args.stride)
        md_iters.append(ctx.getStats()['iterations'])
        for gm in nets.values():
            gm.addSnapshot()
    log['trajectory'] = {'snapshots': args.snapshots, 'stride_steps': args.stride, 'seconds': time.perf_counter() - t1,
                         'pcg_evaluations_per_step_from_previous': float(np.mean(md_iters))}
    ntrain = args.snapshots - args.holdout
    log['training'] = {}
    for name, gm in nets.items():
        t1 = time.perf_counter()
        losses = gm.train(args.epochs, batch=args.batch, learning_rate=args.lr, ntrain=ntrain)
        secs = time.perf_counter() - t1
        log['training'][name] = {
            'kind': gm.kind, 'epochs': args.epochs, 'train_snapshots': ntrain, 'samples': ntrain * nat,
            'batch': args.batch, 'seconds': secs, 'loss_first_epoch': float(losses[0]),
            'loss_last_epoch': float(losses[-1]),
            'holdout_snapshots': args.holdout,
            'holdout_loss': gm.validationLoss(ntrain, args.holdout, batch=args.batch) if args.holdout else None,
            'holdout_loss_zero_output': zero_loss(gm, ntrain, args.holdout) if args.holdout else None}
    res = {(s, m): [] for s in ('sor', 'pcg') for m in MODES}
    eps0 = {m: [] for m in MODES}
    err0 = {m: [] for m in MODES}
    gz = zero_network(ctx)
    spread = []
    for _ in range(args.test):
        md.step(args.stride - 1)
        prev = ctx.getInducedDipoles()  # converged dipoles one step before the test configuration
        md.step(1)
        # the networks' guesses for this configuration: from its alpha E_fixed, and from the step before
        guesses = {'previous': prev}
        for name, gm in nets.items():
            guesses[name] = gm.predictLast()
        alpha_e = gz.predictLast()
        ctx.setDipoleSolver('pcg')
        ctx.setDipoleGuess('alpha_e')
        ctx.setDipoleConvergence(1e-11, 500)
        e_ref = ctx.evaluateResident()[5]
        mu_star = ctx.getInducedDipoles()
        ctx.setDipoleConvergence(1e-8, 500)
        for m in MODES:
            start = alpha_e if m == 'alpha_e' else np.zeros_like(mu_star) if m == 'zero' else guesses[m]
            err0[m].append(rel_error(start, mu_star))
        energies = []
        for m in MODES:
            eps0[m].append(start_epsilon(ctx, m, guesses))
        for solver in ('sor', 'pcg'):
            for m in MODES:
                a, b, en = solve_counts(ctx, solver, m, guesses, e_ref)
                res[(solver, m)].append((a, b))
                energies.append(en)
        spread.append(float((max(energies) - min(energies)) / abs(e_ref)))
        ctx.setDipoleSolver('pcg')
        ctx.setDipoleGuess('previous')
    out = {}
    for (s, m), v in res.items():
        full = [a for a, _ in v]
        e6 = [b for _, b in v if b is not None]
        out['%s/%s' % (s, m)] = {'evaluations_to_eps_1e-8': float(np.mean(full)),
                                 'evaluations_to_1e-6_rel_E': float(np.mean(e6)) if e6 else None,
                                 'samples': len(v)}
    log['start_error'] = {m: float(np.mean(v)) for m, v in err0.items()}
    log['start_epsilon'] = {m: float(np.mean(v)) for m, v in eps0.items()}
    log['iterations'] = out
    log['converged_energy_spread_rel'] = max(spread) if spread else 0.0
    log['seconds_total'] = time.perf_counter() - t0
    if not args.quiet:
        print(json.dumps(log, indent=1), flush=True)
    return log


if __name__ == '__main__':
    main()
