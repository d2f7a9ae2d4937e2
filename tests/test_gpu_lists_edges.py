"""O-O neighbour lists at their edge cases, HIP path against the oracle, as sets.

* Ties: O-O distances exactly equal to the cut-off in FP64 (pairs displaced
  along one axis by exactly rc, at coordinates where (x + rc) - x == rc and the
  minimum-image difference is rc too).  The reference compares d^2 > rc^2 to
  skip a pair (ReferenceThreeNeighborList.cpp:158-160), so a tie is inside.
* The chain list's voxel edge: with a periodic box, a molecule whose wrapped
  coordinate lands on voxel index nv (an O at x = -1e-17 wraps to x = L
  exactly, ReferenceThreeNeighborList.cpp:86-97,193-200) is stored under a key
  that the wrapped lookups of later molecules never produce, so its pairs with
  later molecules are missing from the chain list.  The dimer list (OpenMM's
  voxel hash, distance <= rc) keeps them.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

L = 1.8  # nm; 3-body voxels of 0.45 nm, nv = 4 per axis
RC2, RC3 = 0.65, 0.45


@pytest.fixture(scope='module')
def mb():
    import mbpol_openmm_plugin_amd as m
    if m.device_count() < 1:
        pytest.fail('no GPU visible to libmbpol_hip.so')
    return m


def place(mb, oxy_nm, seed):
    """Distorted waters whose O atoms sit exactly at oxy_nm (nm)."""
    rng = np.random.default_rng(seed)
    w = np.array([mb.systems.monomer(rng, o * 10.0) for o in oxy_nm]) * 0.1
    return w - w[:, :1, :] + np.asarray(oxy_nm)[:, None, :]


def periodic_diff(a, b, period):
    d = a - b
    return d - np.floor(d / period + 0.5) * period


def tie_x(rc, start):
    """The first x >= start (on a 1/64 nm grid, x + rc < 1 nm) at which x + rc lies exactly rc
    away, directly and as a minimum-image difference."""
    for k in range(64):
        x = (np.floor(start * 64.0) + k) / 64.0
        y = x + rc
        if y < 1.0 and y - x == rc and periodic_diff(y, x, L) == rc:
            return x
    raise AssertionError('no exact tie found')


def random_oxygens(rng, n, taken, box):
    out = []
    while len(out) < n:
        p = rng.uniform(0.0, box, 3)
        allp = np.array(taken + out) if (taken or out) else np.zeros((0, 3))
        d = allp - p
        d -= box * np.round(d / box)
        if len(allp) and np.min(np.linalg.norm(d, axis=1)) < 0.26:
            continue
        out.append(p)
    return out


def tie_oxygens():
    """Pairs displaced by exactly rc along x (3-body 0.45 nm and 2-body 0.65 nm)."""
    out = []
    for k, (x0, rc) in enumerate([(0.0, RC3), (0.015625, RC3), (0.03125, RC3), (0.0625, RC2), (0.125, RC2)]):
        x = tie_x(rc, x0)
        y, z = 0.2 + 0.35 * k, 0.9 + 0.15 * (k % 2)
        out += [np.array([x, y, z]), np.array([x + rc, y, z])]
    return out


def lists(mb, waters, periodic):
    n = len(waters)
    kw = dict(twobody_cutoff=RC2, threebody_cutoff=RC3, components=2 | 4)
    if periodic:
        sysm = mb.forcefield.create_system(n, nonbonded_method='PME', box=L, **kw)
    else:
        sysm = mb.forcefield.create_system(n, nonbonded_method='CutoffNonPeriodic', **kw)
    pos = mb.forcefield.water_positions_with_m(waters)
    ctx = mb.Context(sysm, precision='double')
    ctx.setListMode(True)
    ctx.setPositions(pos)
    ctx.getState(components=2 | 4)
    return pos[0::4], sorted(map(tuple, ctx.getDimerList().tolist())), sorted(map(tuple, ctx.getTrimerList().tolist()))


@pytest.mark.parametrize('periodic', [False, True])
def test_lists_with_ties_at_the_cutoff(mb, oracle, periodic):
    rng = np.random.default_rng(21)
    ties = tie_oxygens()
    oxy = ties + random_oxygens(rng, 26, ties, L)
    opos, got2, got3 = lists(mb, place(mb, np.array(oxy), 3), periodic)
    box = [L] * 3 if periodic else None
    want2 = sorted(map(tuple, oracle.dimer_list(opos, RC2, periodic, box).tolist()))
    want3 = sorted(map(tuple, oracle.trimer_list(opos, RC3, periodic, box).tolist()))
    # the ties are in both lists
    assert (0, 1) in want2 and (6, 7) in want2
    assert any(t[0] == 1 and t[1] == 0 for t in want3) or any(t[1] == 1 and t[2] == 0 for t in want3)
    print('dimers', len(got2), len(want2), 'trimers', len(got3), len(want3), flush=True)
    assert got2 == want2
    assert got3 == want3


def test_chain_list_voxel_edge(mb, oracle):
    rng = np.random.default_rng(22)
    # molecule 0 sits at x = -1e-17: wrapped, x = L exactly, voxel index nv along x
    first = [np.array([-1e-17, 0.8, 0.8]), np.array([0.9, 1.7, 1.0])]
    # its neighbours, later in index order, just across the periodic boundary
    near = [np.array([0.3, 0.8, 0.8]), np.array([1.5, 0.75, 0.9]), np.array([0.05, 1.05, 0.7])]
    oxy = first + near
    oxy += random_oxygens(rng, 25, oxy, L)
    opos, got2, got3 = lists(mb, place(mb, np.array(oxy), 4), True)
    box = [L] * 3
    want2 = sorted(map(tuple, oracle.dimer_list(opos, RC2, True, box).tolist()))
    want3 = sorted(map(tuple, oracle.trimer_list(opos, RC3, True, box).tolist()))
    # the dimer list has molecule 0's neighbours; the chain list lacks the pairs (later, 0), so
    # no chain ends at molecule 0: (4, 2, 0) is missing although d(4,2) and d(2,0) are within rc
    assert (0, 2) in want2 and (0, 3) in want2 and (0, 4) in want2
    near3 = set(map(tuple, oracle.dimer_list(opos, RC3, True, box).tolist()))
    assert (2, 4) in near3 and (0, 2) in near3
    assert not any(t[2] == 0 for t in want3) and (4, 2, 0) not in want3
    print('dimers', len(got2), len(want2), 'trimers', len(got3), len(want3), flush=True)
    assert got2 == want2
    assert got3 == want3
