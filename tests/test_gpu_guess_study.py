"""The learned dipole guess on the 4096-water box at full size (HIP extension
with no counterpart in the reference; the reference always starts the
induced-dipole solve from alpha * E_fixed, MBPolReferenceElectrostaticsForce.cpp
:628-639).

Runs mbpol_openmm_plugin_amd.guess_study: a 300 K Langevin trajectory of the
4096 waters, 2000 snapshots 5 steps apart; two networks trained on the first
1800 of them -- one correcting alpha * E_fixed, one correcting the previous
MD step's converged dipoles -- and scored on the last 200, which no training
step saw; then, on three later configurations, every start's error against
the dipoles converged to 1e-11, its epsilon (the reference's convergence
measure after one SOR step: a scaled residual), and the SOR and PCG
evaluations to epsilon 1e-8 and to 1e-6 of the converged energy.

Checked: each network's held-out error is below that of a zero correction
(0.6 of it for the alpha * E_fixed network, which explains about half of its
targets' variance; 0.95 for the previous-step one, whose targets -- the
change of the dipoles over one MD step beyond alpha times the change of the
field -- it explains about 15% of) and within 1.15 of its last training
epoch's (it generalises and does not overfit); each network's start is closer
to the converged dipoles than the start it corrects (alpha * E_fixed, the
previous step's dipoles); every start reaches the same converged energy (1e-7
relative).  The epsilons and
iteration counts of each start and solver are printed, not asserted: the
one-step residual weighs the parts of the error otherwise than the per-atom
error the networks are trained on, and a start closer to the solution can
have a larger one.
"""
import json

import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def mb():
    import mbpol_openmm_plugin_amd as m
    if m.device_count() < 1:
        pytest.fail('no GPU visible to libmbpol_hip.so')
    return m


def test_learned_guess_4096_waters(mb):
    from mbpol_openmm_plugin_amd import guess_study
    log = guess_study.main(['--snapshots', '2000', '--holdout', '200', '--stride', '5', '--equilibrate', '1000',
                            '--epochs', '10', '--test', '3', '--quiet'])
    print(json.dumps(log, indent=1), flush=True)
    for name, t in log['training'].items():
        print('%s: train loss %.4f -> %.4f, held-out %.4f, zero correction %.4f' %
              (name, t['loss_first_epoch'], t['loss_last_epoch'], t['holdout_loss'], t['holdout_loss_zero_output']),
              flush=True)
        bound = {'learned': 0.6, 'learned_previous': 0.95}[name]
        assert t['holdout_loss'] < bound * t['holdout_loss_zero_output'], name
        assert t['holdout_loss'] <= 1.15 * t['loss_last_epoch'], name
    err = log['start_error']
    eps = log['start_epsilon']
    its = log['iterations']
    for m in guess_study.MODES:
        print('%-17s start error %.3e eps %.3e  SOR %.2f (%s to 1e-6 E)  PCG %.2f (%s to 1e-6 E)' %
              (m, err[m], eps[m], its['sor/' + m]['evaluations_to_eps_1e-8'],
               its['sor/' + m]['evaluations_to_1e-6_rel_E'], its['pcg/' + m]['evaluations_to_eps_1e-8'],
               its['pcg/' + m]['evaluations_to_1e-6_rel_E']), flush=True)
    assert err['learned'] < err['alpha_e']
    assert err['learned_previous'] < err['previous']
    assert log['converged_energy_spread_rel'] <= 1e-7
