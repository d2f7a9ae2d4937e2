#!/usr/bin/env python3
"""Write tests/golden/config2_oracle.npz: the FP64 oracle's evaluation of the
4096-water periodic box (systems.config2(): 49.66 A cubic box, PME tolerance
1e-5 with alpha 3.654826 nm^-1 and a 121^3 grid, rc 0.9 nm, 2-body 0.65 nm,
3-body 0.45 nm, dipole epsilon 1e-8).

Stored: the per-component energies (kJ/mol), the forces on every particle
(kJ/mol/nm, float64), the oracle's dipole iteration count, and the O-O lists
as sets -- the number of dimers / chain trimers, the SHA-256 of the sorted
int32 rows, and the number of rows per first molecule (to locate a mismatch).
The positions' SHA-256 is stored too, so a test can check that it evaluates
the same coordinates.  The oracle runs single-threaded for a few minutes.

Run from the repository root after `make -C oracle`:
    python3 tests/golden/make_config2_fixture.py
"""
import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, '..', '..'))
for p in (ROOT, os.path.join(ROOT, 'oracle')):
    sys.path.insert(0, p)
import pyoracle  # noqa: E402
import mbpol_openmm_plugin_amd as m  # noqa: E402

OUT = os.path.join(HERE, 'config2_oracle.npz')


def config2_system():
    waters, L = m.systems.config2()
    sysm = m.forcefield.create_system(4096, nonbonded_method='PME', box=L, ewald_error_tolerance=1e-5,
                                      dipole_epsilon=1e-8, ewald_alpha=3.654826, pme_grid=(121, 121, 121))
    return sysm, m.forcefield.water_positions_with_m(waters)


def list_digest(rows, nmol):
    """(count, sha256 of the lexicographically sorted int32 rows, rows per first molecule)."""
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    order = np.lexsort(rows.T[::-1]) if len(rows) else np.zeros(0, np.int64)
    srt = np.ascontiguousarray(rows[order])
    per_first = np.bincount(rows[:, 0], minlength=nmol).astype(np.int32) if len(rows) else np.zeros(nmol, np.int32)
    return len(rows), hashlib.sha256(srt.tobytes()).hexdigest(), per_first


def positions_digest(pos):
    return hashlib.sha256(np.ascontiguousarray(pos, dtype=np.float64).tobytes()).hexdigest()


def main():
    sysm, pos = config2_system()
    t0 = time.time()
    E, F, lists = pyoracle.evaluate_system(sysm, pos, return_lists=True)
    dt = time.time() - t0
    nd, hd, pd = list_digest(lists['dimers'], 4096)
    nt, ht, pt = list_digest(lists['trimers'], 4096)
    comps = np.array([E[k] for k in pyoracle.COMPONENTS + ('total',)])
    np.savez_compressed(OUT, components=np.array(pyoracle.COMPONENTS + ('total',)), energies=comps,
                        forces=F, iterations=lists['elec']['iterations'], positions_sha256=positions_digest(pos),
                        dimer_count=nd, dimer_sha256=hd, dimers_per_first=pd,
                        trimer_count=nt, trimer_sha256=ht, trimers_per_first=pt, oracle_seconds=dt)
    print('oracle %.1f s; energies %s; dimers %d trimers %d; iterations %d' %
          (dt, dict(zip(pyoracle.COMPONENTS + ('total',), comps)), nd, nt, lists['elec']['iterations']))


if __name__ == '__main__':
    main()
