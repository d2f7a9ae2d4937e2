"""Seeded synthetic water systems for the benchmark and parity workloads.

Geometry recipes (all coordinates returned in nm, O,H1,H2 per water):
* ``cluster``: O uniform in a sphere, O-O >= 2.6 A by rejection, uniformly
  random orientation, OH = 0.9572 A + N(0, 0.02 A), HOH = 104.52 deg + N(0, 3 deg);
* ``lattice_box``: O on a jittered lattice (sigma 0.3 A, O-O >= 2.6 A) in a
  cubic periodic box, monomers as above.
"""
import numpy as np

OH_BOND_A = 0.9572
HOH_DEG = 104.52


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def _random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def monomer(rng, oxygen_A):
    """One distorted water around an O position (Angstrom in, Angstrom out)."""
    r1 = OH_BOND_A + rng.normal(0.0, 0.02)
    r2 = OH_BOND_A + rng.normal(0.0, 0.02)
    th = np.deg2rad(HOH_DEG + rng.normal(0.0, 3.0))
    local = np.array([[0.0, 0.0, 0.0],
                      [r1 * np.sin(th / 2), r1 * np.cos(th / 2), 0.0],
                      [-r2 * np.sin(th / 2), r2 * np.cos(th / 2), 0.0]])
    R = _random_rotation(rng)
    return oxygen_A + local @ R.T


def cluster(n, seed=0, radius_A=None):
    """n waters in a sphere (default radius 6 A for 14 waters, 2.6 n^(1/3) A otherwise)."""
    rng = _rng(seed)
    if radius_A is None:
        radius_A = 6.0 if n == 14 else 2.6 * n ** (1.0 / 3.0)
    oxy = []
    tries = 0
    while len(oxy) < n:
        tries += 1
        if tries > 200000:
            raise RuntimeError('could not place %d waters in a %.2f A sphere' % (n, radius_A))
        p = rng.uniform(-radius_A, radius_A, 3)
        if p @ p > radius_A * radius_A:
            continue
        if oxy and np.min(np.linalg.norm(np.array(oxy) - p, axis=1)) < 2.6:
            continue
        oxy.append(p)
    waters = np.array([monomer(rng, o) for o in oxy])
    return waters * 0.1


def lattice_box(dims, box_A, n, seed=0, sigma_A=0.3):
    """n waters on a jittered dims[0] x dims[1] x dims[2] lattice in a cubic box of side box_A."""
    rng = _rng(seed)
    nx, ny, nz = dims
    sites = np.array([[(i + 0.5) * box_A / nx, (j + 0.5) * box_A / ny, (k + 0.5) * box_A / nz]
                      for i in range(nx) for j in range(ny) for k in range(nz)])
    chosen = np.sort(rng.choice(len(sites), size=n, replace=False))
    base = sites[chosen]
    L = box_A
    # Start from the bare lattice (valid: spacing > 2.6 A) and jitter the sites one
    # at a time; a draw is kept only when every O-O distance stays >= 2.6 A, else
    # redrawn (up to 200 times, after which the site keeps its lattice point).
    oxy = base.copy()
    for s_ in rng.permutation(n):
        for _ in range(200):
            p = base[s_] + rng.normal(0.0, sigma_A, 3)
            d = oxy - p
            d -= L * np.round(d / L)
            r2 = np.einsum('ij,ij->i', d, d)
            r2[s_] = np.inf
            if r2.min() >= 2.6 * 2.6:
                oxy[s_] = p
                break
    waters = np.array([monomer(rng, o) for o in oxy])
    return waters * 0.1, box_A * 0.1


# The five workloads (cluster / 256-water PBC / 4096-water PBC / batched clusters)
def config0(seed=0):
    """14-water cluster (no cutoff)."""
    return cluster(14, seed=seed, radius_A=6.0)


def config1(seed=1):
    """256 waters, cubic 19.71 A box, 7x7x6 lattice."""
    return lattice_box((7, 7, 6), 19.71, 256, seed=seed)


def config2(seed=2):
    """4096 waters, cubic 49.66 A box, 16x16x16 lattice."""
    return lattice_box((16, 16, 16), 49.66, 4096, seed=seed)


def config4(seed=4, count=2048, nmin=3, nmax=64):
    """Batched clusters: sizes uniform in [nmin, nmax], radius 2.6 n^(1/3) A."""
    rng = _rng(seed)
    sizes = rng.integers(nmin, nmax + 1, size=count)
    return [cluster(int(n), seed=int(seed * 100003 + i)) for i, n in enumerate(sizes)]
