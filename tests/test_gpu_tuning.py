"""Launch-shape tuning (HIP extension, no counterpart in the reference).

The grid caps of the 2-body, 3-body and dispersion kernels are tuned on a
context's first forked evaluation by timing trial evaluations; the tuning
relies on every cap giving the same results bit for bit (fixed-point force
and energy sums, per-unit energies summed in list order).  Checked on the
256-water PME box over four configurations, with the solver started from the
previous step's dipoles (so that the trial solves must leave that state as
they found it):

* contexts with caps and fork points fixed to other values by their
  environment variables equal the untuned context bit for bit (energies,
  forces, dipoles, iterations), the short-range work enqueued by the side
  enqueuer's host thread among them, and the short-range terms enqueued in
  another order (MBPOL_HIP_SR_ORDER);
* the tuned context equals the untuned one bit for bit, ran trial evaluations,
  and reports caps among its candidates.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def mb():
    import mbpol_openmm_plugin_amd as m
    if m.device_count() < 1:
        pytest.fail('no GPU visible to libmbpol_hip.so')
    return m


def run(mb, precision='mixed'):
    waters, L = mb.systems.config1()
    sysm = mb.forcefield.create_system(256, nonbonded_method='PME', box=L, ewald_error_tolerance=1e-5,
                                       dipole_epsilon=1e-8)
    pos = mb.forcefield.water_positions_with_m(waters)
    ctx = mb.Context(sysm, precision=precision)
    ctx.setDipoleGuess('previous')
    rng = np.random.default_rng(31)
    out = []
    for k in range(4):
        ctx.uploadPositions(pos + rng.normal(scale=0.002, size=pos.shape) * (k > 0))
        e = ctx.evaluateResident()
        out.append((e, ctx.downloadForces(), ctx.getInducedDipoles(), ctx.getStats()['iterations']))
    caps = {n: ctx.getTuning(n) for n in ('fork_early', 'pip_waves', 'pip3_waves', 'disp_waves', 'tune_evaluations')}
    return out, caps


def same(a, b):
    for k, ((ea, fa, da, ia), (eb, fb, db, ib)) in enumerate(zip(a, b)):
        assert np.array_equal(ea, eb), (k, ea - eb)
        assert np.array_equal(fa, fb), (k, np.abs(fa - fb).max())
        assert np.array_equal(da, db), k
        assert ia == ib, k


def test_caps_change_no_result_and_tuning_is_invisible(mb, monkeypatch):
    monkeypatch.setenv('MBPOL_HIP_TUNE', '0')
    ref, caps0 = run(mb)
    print('untuned', caps0, 'energies', ref[-1][0], flush=True)
    assert caps0['tune_evaluations'] == 0
    for pip, pip3, disp, fork in ((2, 0, 2, 'late'), (0, 2, 1, 'early'), (1, 1, 0, 'thread')):
        monkeypatch.setenv('MBPOL_HIP_PIP_WAVES', str(pip))
        monkeypatch.setenv('MBPOL_HIP_3B_WAVES', str(pip3))
        monkeypatch.setenv('MBPOL_HIP_DISP_WAVES', str(disp))
        monkeypatch.setenv('MBPOL_HIP_FORK', fork)
        got, caps = run(mb)
        print('fixed caps', caps, flush=True)
        assert (caps['pip_waves'], caps['pip3_waves'], caps['disp_waves']) == (pip, pip3, disp)
        assert caps['fork_early'] == {'late': 0, 'early': 1, 'thread': 2}[fork]
        same(ref, got)
    # the short-range terms in another order on their stream
    monkeypatch.setenv('MBPOL_HIP_SR_ORDER', 'd32')
    got, caps = run(mb)
    print('dispersion first', caps, flush=True)
    same(ref, got)
    for name in ('MBPOL_HIP_TUNE', 'MBPOL_HIP_PIP_WAVES', 'MBPOL_HIP_3B_WAVES', 'MBPOL_HIP_DISP_WAVES',
                 'MBPOL_HIP_FORK', 'MBPOL_HIP_SR_ORDER'):
        monkeypatch.delenv(name)
    got, caps = run(mb)
    print('tuned', caps, flush=True)
    assert caps['tune_evaluations'] >= 4 * 5
    assert caps['pip_waves'] in (0, 1, 2) and caps['pip3_waves'] in (0, 1, 2) and caps['disp_waves'] in (0, 1, 2)
    same(ref, got)
