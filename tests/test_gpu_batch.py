"""Independent water clusters evaluated together as groups of one system
(HIP extension for the batched-cluster workload; no counterpart in the
reference, which evaluates one cluster per Context).

* five clusters of 3..40 waters in one batch against the FP64 oracle run on
  each cluster alone: per-cluster energies of every component within 1e-6 of
  the cluster's total energy, per-cluster forces within 1e-4 relative RMS
  (the mixed-precision bar);
* every one of the 2048 clusters of 3..64 waters against the oracle fixture
  tests/golden/config4_oracle.npz (energies and force checksums);
* the 2048-cluster batch of 3..64 waters: group energies add up to the
  system's energies, and a cluster's energies do not depend on the other
  clusters in the batch (same cluster alone and at another place in a
  reordered batch, within the dipole tolerance's effect, 1e-7).
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


def cluster_system(mb, n):
    return mb.forcefield.create_system(n, nonbonded_method='CutoffNonPeriodic', twobody_cutoff=1e10,
                                       threebody_cutoff=1e10, nonbonded_cutoff=1e3, dipole_epsilon=1e-8)


def test_batch_matches_oracle_per_cluster(mb, oracle):
    sizes = [3, 6, 14, 23, 40]
    clusters = [mb.systems.cluster(n, seed=100 + k) for k, n in enumerate(sizes)]
    sysb, pos, start = mb.forcefield.create_cluster_batch(clusters)
    ctx = mb.Context(sysb, precision='mixed')
    ctx.setPositions(pos)
    st = ctx.getState()
    ge = ctx.getGroupEnergies()
    F = st.getForces()
    names = ('onebody', 'twobody', 'threebody', 'dispersion', 'electrostatics')
    for g, (w, n) in enumerate(zip(clusters, sizes)):
        E, Fo = oracle.evaluate_system(cluster_system(mb, n), mb.forcefield.water_positions_with_m(w))
        tot = abs(E['total'])
        for k, name in enumerate(names):
            assert abs(ge[g, k] - E[name]) <= 1e-6 * tot, (n, name, ge[g, k], E[name])
        assert abs(ge[g, 5] - E['total']) <= 1e-6 * tot
        Fg = F[4 * start[g]:4 * start[g + 1]]
        err = float(np.sqrt(np.mean((Fg - Fo) ** 2)) / np.sqrt(np.mean(Fo ** 2)))
        print('cluster of %d waters: E %.6f vs %.6f, force rel rms %.2e' % (n, ge[g, 5], E['total'], err), flush=True)
        assert err <= 1e-4
    assert np.allclose(ge.sum(0), st._e, rtol=0, atol=1e-6 * np.abs(st._e).max())


def test_full_batch_is_independent_per_cluster(mb):
    clusters = mb.systems.config4()
    assert len(clusters) == 2048
    sysb, pos, start = mb.forcefield.create_cluster_batch(clusters)
    ctx = mb.Context(sysb, precision='mixed')
    ctx.setPositions(pos)
    st = ctx.getState()
    ge = ctx.getGroupEnergies()
    print('2048 clusters, %d waters: E %.4f kJ/mol, %s' % (start[-1], st.getPotentialEnergy(), st.stats), flush=True)
    assert np.all(np.isfinite(ge)) and np.all(np.isfinite(st.getForces()))
    assert np.abs(ge.sum(0) - st._e).max() <= 1e-6 * np.abs(st._e).max()
    # reversed order: the same cluster energies
    rev = clusters[::-1]
    sysr, posr, startr = mb.forcefield.create_cluster_batch(rev)
    ctxr = mb.Context(sysr, precision='mixed')
    ctxr.setPositions(posr)
    ctxr.getState()
    ger = ctxr.getGroupEnergies()[::-1]
    scale = np.abs(ge[:, 5])[:, None] + 1.0
    assert np.abs(ger - ge).max(axis=1).max() <= 1e-7 * scale.max()
    assert np.all(np.abs(ger - ge) <= 1e-7 * scale)
    # the largest cluster alone
    k = int(np.argmax([len(c) for c in clusters]))
    sys1, pos1, _ = mb.forcefield.create_cluster_batch([clusters[k]])
    ctx1 = mb.Context(sys1, precision='mixed')
    ctx1.setPositions(pos1)
    ctx1.getState()
    e1 = ctx1.getGroupEnergies()[0]
    print('largest cluster (%d waters) alone %s, in batch %s' % (len(clusters[k]), e1, ge[k]), flush=True)
    assert np.all(np.abs(e1 - ge[k]) <= 1e-7 * (abs(ge[k, 5]) + 1.0))


def test_full_batch_matches_oracle_per_cluster(mb):
    """All 2048 clusters of the batch against the FP64 oracle run on each
    cluster alone (tests/golden/config4_oracle.npz, written by
    tests/golden/make_config4_fixture.py): every component's energy and the
    total within 1e-6 of the cluster's energy scale, and two force checksums
    per cluster -- the RMS force within 1e-4 relative, and the projection on a
    seeded normal field within 1e-4 of |F| |u| (the mixed-precision bar).  The
    energy scale is the sum of the absolute component energies: the components
    of a cluster cancel, down to totals of 0.02 kJ/mol among these clusters, so
    the FP32 rounding of the dispersion pair terms (about 5e-9 of the
    dispersion energy) is measured against the energies it acts on.  The error
    relative to the total is printed too."""
    import hashlib
    import os
    ref = np.load(os.path.join(os.path.dirname(__file__), 'golden', 'config4_oracle.npz'))
    clusters = mb.systems.config4()
    h = hashlib.sha256()
    for c in clusters:
        h.update(np.ascontiguousarray(c, dtype=np.float64).tobytes())
    assert h.hexdigest() == str(ref['positions_sha256'])
    sysb, pos, start = mb.forcefield.create_cluster_batch(clusters)
    ctx = mb.Context(sysb, precision='mixed')
    ctx.setPositions(pos)
    F = ctx.getState().getForces()
    ge = ctx.getGroupEnergies()
    E = ref['energies']
    scale = np.abs(E[:, :5]).sum(axis=1)
    err_e = np.abs(ge - E).max(axis=1) / scale
    err_tot = np.abs(ge[:, 5] - E[:, 5]) / np.abs(E[:, 5])
    frms = np.zeros(len(clusters))
    fdot = np.zeros(len(clusters))
    for k in range(len(clusters)):
        Fk = F[4 * start[k]:4 * start[k + 1]]
        u = np.random.default_rng(1_000_003 + k).standard_normal(Fk.shape)
        frms[k] = np.sqrt(np.mean(Fk ** 2))
        fdot[k] = np.sum(Fk * u)
        scale = np.sqrt(np.sum(Fk ** 2)) * np.sqrt(np.sum(u ** 2))
        fdot[k] = abs(fdot[k] - ref['force_dot'][k]) / scale
    err_rms = np.abs(frms - ref['force_rms']) / ref['force_rms']
    worst = int(np.argmax(err_e))
    print('2048 clusters: worst energy error %.2e of the energy scale (cluster %d, %d waters); relative to the '
          'total: median %.1e, worst %.1e; worst force rms error %.2e, worst projection error %.2e' %
          (err_e.max(), worst, len(clusters[worst]), np.median(err_tot), err_tot.max(), err_rms.max(), fdot.max()),
          flush=True)
    assert err_e.max() <= 1e-6
    assert err_rms.max() <= 1e-4
    assert fdot.max() <= 1e-4
