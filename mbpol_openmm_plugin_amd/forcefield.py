"""Assemble an MB-pol System the way the reference's force field does.

The reference builds its System from ``python/mbpol.xml`` through the
generators in ``python/mbpol.py``; this module does the same job for water
(and optional Cl-) systems given as coordinates:

* particle order per water: O, H1, H2, M (python/mbpol.xml:9-16);
* M is a ThreeParticleAverageSite with weights 0.573293118 / 0.213353441 x2
  (python/mbpol.xml:14);
* 1-body, 2-body and 3-body Forces take {O, H1, H2} per water
  (python/mbpol.py:65-76), cutoffs 0.65 and 0.45 nm (python/mbpol.xml:31,34);
* electrostatics: one entry per particle, moleculeIndex = water index,
  atomType 0/1/2 for O/H/M, charge/damping/polarisability from
  python/mbpol.xml:24-26, Thole vector python/mbpol.py:266-269;
* dispersion: O/H/M/Cl types with the C6/d6 tables of python/mbpol.xml:52-83,
  same-residue and M pairs excluded (:100-104);
* nonbonded-method mapping of python/mbpol.py:46-49,118-121,191-194,291-293
  and the dispersion script (:89-93).
"""
import json
import os

import numpy as np

from . import api

_TABLES = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'params', 'mbpol_tables.json')

NoCutoff, CutoffNonPeriodic, CutoffPeriodic, PME = 'NoCutoff', 'CutoffNonPeriodic', 'CutoffPeriodic', 'PME'


def tables():
    with open(_TABLES) as fh:
        return json.load(fh)


def msite_weights():
    return tuple(tables()['forcefield']['msite_weights'])


def water_positions_with_m(waters):
    """waters: (W, 3, 3) O,H1,H2 in nm -> (4W, 3) O,H1,H2,M."""
    w = np.asarray(waters, dtype=np.float64).reshape(-1, 3, 3)
    wo, wh1, wh2 = msite_weights()
    m = wo * w[:, 0] + wh1 * w[:, 1] + wh2 * w[:, 2]
    return np.concatenate([w, m[:, None, :]], axis=1).reshape(-1, 3)


def create_system(num_waters, nonbonded_method=PME, nonbonded_cutoff=0.9, box=None, ions=0,
                  twobody_cutoff=None, threebody_cutoff=None, ewald_error_tolerance=None, ewald_alpha=None,
                  pme_grid=None, dipole_epsilon=None, max_iterations=None, components=api.MBPOL_ALL):
    """System for ``num_waters`` waters (4 particles each, O,H1,H2,M) followed by
    ``ions`` Cl- particles.  ``box`` (nm, cubic or 3-vector) is needed for the
    periodic methods."""
    t = tables()
    ff = t['forcefield']
    sys = api.System()
    W = int(num_waters)
    wo, wh1, wh2 = ff['msite_weights']
    for w in range(W):
        for mass in (15.99943, 1.007947, 1.007947, 0.0):
            sys.addParticle(mass)
        sys.setVirtualSite(4 * w + 3, api.ThreeParticleAverageSite(4 * w, 4 * w + 1, 4 * w + 2, wo, wh1, wh2))
    for _ in range(ions):
        sys.addParticle(35.4532)
    periodic = nonbonded_method in (PME, CutoffPeriodic)
    if box is not None:
        b = np.broadcast_to(np.asarray(box, dtype=np.float64), (3,))
        sys.setDefaultPeriodicBoxVectors([b[0], 0, 0], [0, b[1], 0], [0, 0, b[2]])
    elif periodic:
        raise ValueError('periodic method needs a box')
    mols = [[4 * w, 4 * w + 1, 4 * w + 2] for w in range(W)]

    if ions and components & api.MBPOL_ELECTROSTATICS:
        raise ValueError('MB-pol electrostatics covers water only; build ion systems without it')
    if components & api.MBPOL_ELECTROSTATICS and W > 0:
        # CutoffNonPeriodic maps to NoCutoff for the electrostatics (python/mbpol.py:287-293)
        el = api.MBPolElectrostaticsForce()
        el.setCutoffDistance(nonbonded_cutoff)
        el.setNonbondedMethod(el.PME if nonbonded_method == PME else el.NoCutoff)
        el.setTholeParameters(ff['thole'])
        atoms = ff['electrostatics_atoms']
        for w in range(W):
            for name, typ in (('O', 0), ('H', 1), ('H', 1), ('M', 2)):
                a = atoms[name]
                el.addElectrostatics(a['charge'], w, typ, a['damping'], a['polarizability'])
        if ewald_error_tolerance is not None:
            el.setEwaldErrorTolerance(ewald_error_tolerance)
        if ewald_alpha is not None:
            el.setAEwald(ewald_alpha)
        if pme_grid is not None:
            el.setPmeGridDimensions(list(pme_grid))
        if dipole_epsilon is not None:
            el.setMutualInducedTargetEpsilon(dipole_epsilon)
        if max_iterations is not None:
            el.setMutualInducedMaxIterations(max_iterations)
        sys.addForce(el)
    if components & api.MBPOL_ONEBODY and W > 0:
        ob = api.MBPolOneBodyForce()
        ob.setNonbondedMethod(ob.Periodic if periodic else ob.NonPeriodic)
        for m in mols:
            ob.addOneBody(m)
        sys.addForce(ob)
    for cls, cut, key, flag in ((api.MBPolTwoBodyForce, twobody_cutoff, 'twobody_cutoff_nm', api.MBPOL_TWOBODY),
                                (api.MBPolThreeBodyForce, threebody_cutoff, 'threebody_cutoff_nm',
                                 api.MBPOL_THREEBODY)):
        if not (components & flag) or W < 2:
            continue
        f = cls()
        f.setCutoff(cut if cut is not None else ff[key])
        f.setNonbondedMethod({NoCutoff: f.NoCutoff, CutoffNonPeriodic: f.CutoffNonPeriodic,
                              CutoffPeriodic: f.CutoffPeriodic, PME: f.CutoffPeriodic}[nonbonded_method])
        for m in mols:
            f.addParticle(m)
        sys.addForce(f)
    if components & api.MBPOL_DISPERSION:
        d = api.MBPolDispersionForce()
        d.setTypeTables(4, ff['dispersion_C6'], ff['dispersion_d6'])
        d.setNonbondedMethod({NoCutoff: d.NoCutoff, CutoffNonPeriodic: d.CutoffNonPeriodic,
                              CutoffPeriodic: d.CutoffPeriodic, PME: d.CutoffPeriodic}[nonbonded_method])
        d.setCutoff(nonbonded_cutoff)
        for w in range(W):
            for typ in (0, 1, 1, 2):
                d.addParticle(typ, w)
        for k in range(ions):
            d.addParticle(3, W + k)
        sys.addForce(d)
    return sys


def read_pdb(path):
    """Minimal PDB reader for water / ion fixtures: returns (positions nm, atom
    names, residue names, box nm or None)."""
    pos, names, res, box = [], [], [], None
    with open(path) as fh:
        for line in fh:
            if line.startswith('CRYST1'):
                box = [float(line[6:15]) / 10, float(line[15:24]) / 10, float(line[24:33]) / 10]
            if line.startswith(('ATOM', 'HETATM')):
                names.append(line[12:16].strip())
                res.append(line[17:20].strip())
                pos.append([float(line[30:38]) / 10, float(line[38:46]) / 10, float(line[46:54]) / 10])
    return np.array(pos), names, res, box


def create_cluster_batch(clusters, dipole_epsilon=1e-8, components=api.MBPOL_ALL):
    """One System holding independent water clusters (lists of (n, 3, 3) water
    coordinates in nm) as groups: the settings of a single cluster (no cutoffs
    on the 2- and 3-body lists, direct-sum electrostatics, dispersion without a
    cutoff) with no interaction between clusters.  Returns (system, positions
    with M sites, water offsets of the clusters)."""
    sizes = [len(c) for c in clusters]
    start = np.concatenate([[0], np.cumsum(sizes)]).astype(int)
    sys_ = create_system(int(start[-1]), nonbonded_method='CutoffNonPeriodic', twobody_cutoff=1e10,
                         threebody_cutoff=1e10, nonbonded_cutoff=1e3, dipole_epsilon=dipole_epsilon,
                         components=components)
    sys_.setWaterGroups(start)
    pos = water_positions_with_m(np.concatenate(clusters))
    return sys_, pos, start
