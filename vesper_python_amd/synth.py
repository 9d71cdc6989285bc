"""Seeded synthetic density maps for the benchmark and parity workloads.

Pseudo-atom chains are random walks (3.8 A steps, uniform directions, a step
is rejected if it lands < 3 A from an earlier atom or outside the box minus a
6 A margin; a walk that cannot continue backtracks).  Each atom is an
isotropic Gaussian (sigma 1.5 A, amplitude U(0.8, 1.2)); maps get N(0, 0.02)
noise.  The query map B is a contiguous sub-chain of A, moved by a uniform
SO(3) rotation about its centroid plus a U(-10, 10) A translation, rendered in
its own box aligned to A's grid around the sub-chain's original centroid.
Maps are returned indexed [x][y][z] with their MRC origin and voxel size.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
from scipy.ndimage import gaussian_filter
from scipy.spatial.transform import Rotation

from .mrc import write_mrc


@dataclass
class SynthMap:
    data: np.ndarray            # [x][y][z] float32
    origin: np.ndarray          # float32 (x, y, z), Angstrom
    width: float                # voxel size, Angstrom
    threshold: float            # 90th percentile
    prob: np.ndarray | None = None   # (nz, ny, nx, 4) float32, MRC order


@dataclass
class SynthPair:
    a: SynthMap
    b: SynthMap
    rotation: np.ndarray        # 3x3 applied to the sub-chain
    translation: np.ndarray     # Angstrom
    centroid: np.ndarray        # rotation centre (sub-chain centroid in A's frame)
    meta: dict = field(default_factory=dict)


def random_walk_chain(rng, natoms, box, step=3.8, min_dist=3.0, margin=6.0, tries=64):
    lo, hi = margin, box - margin
    cell = min_dist
    grid: dict = {}

    def key(p):
        return tuple((p // cell).astype(np.int64))

    def ok(p):
        if np.any(p < lo) or np.any(p > hi):
            return False
        k = key(p)
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for q in grid.get((k[0] + dx, k[1] + dy, k[2] + dz), ()):
                        if np.sum((q - p) ** 2) < min_dist * min_dist:
                            return False
        return True

    def add(p):
        grid.setdefault(key(p), []).append(p)

    def remove(p):
        lst = grid[key(p)]
        for i, q in enumerate(lst):
            if q is p:
                del lst[i]
                return

    centre = np.full(3, box / 2.0)
    chain = [centre + rng.uniform(-box / 8, box / 8, 3)]
    add(chain[0])
    stuck = 0
    while len(chain) < natoms:
        head = chain[-1]
        placed = False
        for _ in range(tries):
            d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            p = head + step * d
            if ok(p):
                chain.append(p)
                add(p)
                placed = True
                break
        if not placed:
            stuck += 1
            for _ in range(min(len(chain) - 1, 1 + stuck % 8)):
                remove(chain.pop())
    return np.array(chain)


def render(atoms, amps, shape, origin, width, sigma=1.5, radius=4.0):
    """Sum of isotropic Gaussians on the voxel centres origin + index * width."""
    shape = tuple(int(s) for s in shape)
    r = int(np.ceil(radius * sigma / width))
    off = np.arange(-r, r + 1)
    ox, oy, oz = np.meshgrid(off, off, off, indexing="ij")
    off3 = np.stack([ox.ravel(), oy.ravel(), oz.ravel()], axis=1)
    out = np.zeros(int(np.prod(shape)), dtype=np.float64)
    rel = (np.asarray(atoms) - np.asarray(origin, dtype=np.float64)) / width
    base = np.rint(rel).astype(np.int64)
    chunk = 2048
    inv2s2 = 1.0 / (2.0 * sigma * sigma)
    for s in range(0, len(rel), chunk):
        b = base[s:s + chunk]
        idx = b[:, None, :] + off3[None, :, :]
        pos = idx * width
        d2 = np.sum((pos - rel[s:s + chunk, None, :] * width) ** 2, axis=2)
        w = np.asarray(amps[s:s + chunk])[:, None] * np.exp(-d2 * inv2s2)
        inside = np.all((idx >= 0) & (idx < np.array(shape)), axis=2)
        flat = (idx[..., 0] * shape[1] + idx[..., 1]) * shape[2] + idx[..., 2]
        out += np.bincount(flat[inside], weights=w[inside], minlength=out.size)
    return out.reshape(shape)


def _finish(rng, dens, origin, width):
    dens = dens + rng.normal(0.0, 0.02, dens.shape)
    dens = dens.astype(np.float32)
    return SynthMap(data=dens, origin=np.asarray(origin, dtype=np.float32), width=float(width),
                    threshold=float(np.percentile(dens, 90)))


def make_pair(seed, a_edge=100, natoms=500, b_edge=60, nsub=150, width=1.0, rotation=None, translation=None):
    """One (A, B) pair: A = a_edge^3 voxels with `natoms` pseudo-atoms; B =
    b_edge^3 voxels around a moved `nsub`-atom sub-chain.  `rotation` (3x3)
    and `translation` (Angstrom) replace the drawn ones when given (the draws
    still happen, so every other value of the pair is the same)."""
    rng = np.random.default_rng(seed)
    box = a_edge * width
    chain = random_walk_chain(rng, natoms, box)
    amps = rng.uniform(0.8, 1.2, natoms)
    a_orig = np.zeros(3)
    a = _finish(rng, render(chain, amps, (a_edge,) * 3, a_orig, width), a_orig, width)
    start = int(rng.integers(0, natoms - nsub + 1))
    sub = chain[start:start + nsub]
    rot = Rotation.random(random_state=rng)
    t = rng.uniform(-10.0, 10.0, 3)
    if rotation is not None:
        rot = Rotation.from_matrix(np.asarray(rotation, dtype=np.float64))
    if translation is not None:
        t = np.asarray(translation, dtype=np.float64)
    c = sub.mean(axis=0)
    moved = rot.apply(sub - c) + c + t
    b_orig = np.floor(c / width) * width - 0.5 * b_edge * width
    b = _finish(rng, render(moved, amps[start:start + nsub], (b_edge,) * 3, b_orig, width), b_orig, width)
    return SynthPair(a=a, b=b, rotation=rot.as_matrix(), translation=t, centroid=c,
                     meta=dict(seed=seed, a_edge=a_edge, natoms=natoms, b_edge=b_edge, nsub=nsub,
                               width=width))


def add_prob_channels(rng, m: SynthMap, sigma=2.0):
    """4 probability channels: Dirichlet(0.5) inside the threshold mask,
    Gaussian-smoothed (sigma in Angstrom), renormalised; 0 outside.  Stored in
    MRC order (nz, ny, nx, 4)."""
    mask = m.data >= m.threshold
    p = rng.dirichlet(np.full(4, 0.5), size=m.data.shape).astype(np.float64)
    p *= mask[..., None]
    for k in range(4):
        p[..., k] = gaussian_filter(p[..., k], sigma / m.width)
    s = p.sum(axis=-1, keepdims=True)
    p = np.where(mask[..., None] & (s > 0), p / np.where(s > 0, s, 1.0), 0.0)
    m.prob = np.ascontiguousarray(p.astype(np.float32).transpose(2, 1, 0, 3))
    return m


# Workload presets: (a_edge, natoms, b_edge, nsub)
CONFIGS = {
    "pair100": dict(a_edge=100, natoms=500, b_edge=60, nsub=150),      # -g 16 -s 7 -A 30
    "pair200": dict(a_edge=200, natoms=3000, b_edge=100, nsub=800),    # -g 16 -s 2 -A 10
}


def make_config_pair(name, seed):
    return make_pair(seed, **CONFIGS[name])


def make_uneven_pair(seed):
    """A pair from the uneven batch generator: A edge U{96..320}, 1000-12000
    atoms (scaled to the box), B a 40-60 % sub-chain in a box that holds it."""
    rng = np.random.default_rng(seed)
    a_edge = int(rng.integers(96, 321))
    natoms = int(min(12000, max(1000, round(12000 * (a_edge / 320.0) ** 3))))
    frac = float(rng.uniform(0.4, 0.6))
    nsub = int(natoms * frac)
    b_edge = int(min(a_edge, max(48, round(a_edge * 0.6))))
    return make_pair(int(rng.integers(0, 2**31 - 1)), a_edge=a_edge, natoms=natoms, b_edge=b_edge,
                     nsub=nsub)


def write_pair(pair: SynthPair, prefix: str):
    """Write A/B as MRC files (and .npy probability channels when present)."""
    paths = {}
    for tag, m in (("A", pair.a), ("B", pair.b)):
        p = f"{prefix}_{tag}.mrc"
        write_mrc(p, m.data.transpose(2, 1, 0), m.width, m.origin)
        paths[tag] = p
        if m.prob is not None:
            q = f"{prefix}_{tag}_prob.npy"
            np.save(q, m.prob)
            paths[tag + "_prob"] = q
    return paths
