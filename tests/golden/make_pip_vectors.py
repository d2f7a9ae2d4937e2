#!/usr/bin/env python3
"""Write tests/golden/pip_vectors.json: energies and gradients of the two
reference polynomials (poly_2b_v6x_eval, poly_3b_v2x::eval) at seeded random
variable vectors, computed by the reference's own sources compiled into
oracle/_ref/libref_poly.so (oracle/ref/Makefile).  Also records reference
gammq(3/4, x) values.  Run in a container that holds the reference tree."""
import ctypes
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, '..', '..'))
sys.path.insert(0, os.path.join(ROOT, 'oracle'))
import pyoracle  # noqa: E402


def main():
    L = pyoracle.ref_lib()
    if L is None:
        raise SystemExit('oracle/_ref/libref_poly.so not built (make -C oracle ref)')
    with open(os.path.join(ROOT, 'mbpol_openmm_plugin_amd', 'params', 'mbpol_tables.json')) as fh:
        t = json.load(fh)
    a2 = np.array(t['twobody']['thefit'])
    a3 = np.array(t['threebody']['thefit'])
    dp = ctypes.POINTER(ctypes.c_double)
    rng = np.random.Generator(np.random.PCG64(2024))
    out = {'source': 'reference poly-2b-v6x.cpp / poly-3b-v2x.cpp / gammq.cpp compiled in oracle/_ref',
           'poly2b': [], 'poly3b': [], 'gammq34': []}
    for key, fn, a, nv in (('poly2b', L.ref_poly_2b, a2, 31), ('poly3b', L.ref_poly_3b, a3, 36)):
        for k in range(12):
            # variables are exponentials of distances: keep them in a realistic (0, 1.5) range
            x = rng.uniform(0.0, 1.5, nv) if k < 10 else rng.uniform(0.0, 0.3, nv)
            g = np.zeros(nv)
            e = fn(a.ctypes.data_as(dp), x.ctypes.data_as(dp), g.ctypes.data_as(dp))
            out[key].append({'x': x.tolist(), 'e': e, 'g': g.tolist()})
    for x in [0.0, 1e-6, 0.01, 0.3, 0.75, 1.0, 1.7499, 1.75, 2.0, 5.0, 17.0, 40.0, 120.0, 700.0]:
        out['gammq34'].append([x, L.ref_gammq(0.75, x)])
    with open(os.path.join(HERE, 'pip_vectors.json'), 'w') as fh:
        json.dump(out, fh)


if __name__ == '__main__':
    main()
