"""Synthetic inputs of the benchmark configurations (seeded, deterministic).

Procedural definitions fixed here (the configurations name them but leave
their functional forms open):

* ``pcg_hash``: the PCG output hash of the reference shader (common.glsl:13-18)
  applied to a uint32 key; the value-noise lattice hash.
* ``value_noise3``: trilinear interpolation, with smoothstep fade, of lattice
  values in [-1, 1) drawn from ``pcg_hash`` of the lattice coordinates, the
  octave and the seed.
* ``fbm3``: 4 octaves, base frequency 2, lacunarity 2, gain 0.5, normalised
  by the sum of the octave amplitudes so that it lies in [-1, 1].
* ``icosphere_fbm``: an icosahedron subdivided ``subdiv`` times (each edge
  midpoint projected to the unit sphere; 20*4^subdiv triangles), each vertex
  then scaled radially by ``radius * (1 + amplitude * fbm3(v))``.  Vertices
  are shared through the subdivision, so the displaced surface stays closed.
* ``sphere_rays``: origins ``3 * u`` with ``u`` uniform on the unit sphere
  (normalised Gaussian triple), targets uniform in the unit ball
  (radius U^(1/3) along a uniform direction), direction = normalise(target -
  origin) computed in float64, both stored as float32.
"""
from __future__ import annotations

import numpy as np

_M32 = np.uint64(0xFFFFFFFF)


def pcg_hash(x: np.ndarray) -> np.ndarray:
    """PCG hash of uint32 keys (state*747796405 + 2891336453, RXS-M-XS output)."""
    s = np.asarray(x, dtype=np.uint64) & _M32
    state = (s * np.uint64(747796405) + np.uint64(2891336453)) & _M32
    word = (((state >> ((state >> np.uint64(28)) + np.uint64(4))) ^ state) * np.uint64(277803737)) & _M32
    return ((word >> np.uint64(22)) ^ word) & _M32


def _lattice(ix, iy, iz, octave: int, seed: int) -> np.ndarray:
    m = np.uint64(0xFFFFFFFF)
    k = pcg_hash((ix.astype(np.int64).astype(np.uint64) & m) ^ np.uint64((seed * 0x9E3779B9) & 0xFFFFFFFF))
    k = pcg_hash(k ^ (iy.astype(np.int64).astype(np.uint64) & m))
    k = pcg_hash(k ^ (iz.astype(np.int64).astype(np.uint64) & m))
    k = pcg_hash(k ^ np.uint64(octave))
    return k.astype(np.float64) / 2147483648.0 - 1.0


def value_noise3(p: np.ndarray, octave: int, seed: int) -> np.ndarray:
    f = np.floor(p)
    i = f.astype(np.int64)
    t = p - f
    w = t * t * (3.0 - 2.0 * t)
    out = np.zeros(p.shape[0])
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                val = _lattice(i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz, octave, seed)
                wx = w[:, 0] if dx else 1.0 - w[:, 0]
                wy = w[:, 1] if dy else 1.0 - w[:, 1]
                wz = w[:, 2] if dz else 1.0 - w[:, 2]
                out += val * wx * wy * wz
    return out


def fbm3(p: np.ndarray, seed: int, octaves: int = 4, base_freq: float = 2.0, lacunarity: float = 2.0,
         gain: float = 0.5) -> np.ndarray:
    total = np.zeros(p.shape[0])
    amp, freq, norm = 1.0, base_freq, 0.0
    for o in range(octaves):
        total += amp * value_noise3(p * freq, o, seed)
        norm += amp
        amp *= gain
        freq *= lacunarity
    return total / norm


def icosphere(subdiv: int) -> tuple[np.ndarray, np.ndarray]:
    """Unit icosphere: (vertices float64 [V,3], faces int64 [F,3]) with F = 20*4^subdiv."""
    g = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, g, 0], [1, g, 0], [-1, -g, 0], [1, -g, 0], [0, -1, g], [0, 1, g], [0, -1, -g], [0, 1, -g],
                  [g, 0, -1], [g, 0, 1], [-g, 0, -1], [-g, 0, 1]], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    for _ in range(subdiv):
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        edges = np.concatenate([np.stack([a, b], 1), np.stack([b, c], 1), np.stack([c, a], 1)])
        edges.sort(axis=1)
        uniq, inv = np.unique(edges, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        mid = v[uniq[:, 0]] + v[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = v.shape[0]
        v = np.concatenate([v, mid])
        m = f.shape[0]
        ab, bc, ca = inv[:m] + base, inv[m:2 * m] + base, inv[2 * m:] + base
        f = np.concatenate([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1), np.stack([c, ca, bc], 1),
                            np.stack([ab, bc, ca], 1)])
    return v, f


def icosphere_fbm(subdiv: int = 6, radius: float = 1.0, amplitude: float = 0.1, seed: int = 1) -> np.ndarray:
    """Displaced icosphere as a float32 triangle soup (9 floats per triangle)."""
    v, f = icosphere(subdiv)
    r = radius * (1.0 + amplitude * fbm3(v, seed))
    v = v * r[:, None]
    return v[f].reshape(-1).astype(np.float32)


def _unit_vectors(rng: np.random.Generator, n: int) -> np.ndarray:
    g = rng.standard_normal((n, 3))
    return g / np.linalg.norm(g, axis=1, keepdims=True)


def sphere_rays(n: int, seed: int = 1, origin_radius: float = 3.0) -> tuple[np.ndarray, np.ndarray]:
    """Rays from a sphere of radius 3 aimed at uniform points of the unit ball (float32 xyz)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    org = origin_radius * _unit_vectors(rng, n)
    tgt = _unit_vectors(rng, n) * (rng.random(n) ** (1.0 / 3.0))[:, None]
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return org.astype(np.float32), d.astype(np.float32)


CFG1_SUBDIV = 6
CFG1_RAYS = 1 << 20


def config1(n_rays: int = CFG1_RAYS, seed: int = 1) -> dict:
    """Plain traversal workload: 81,920-triangle displaced icosphere and 2^20 sphere rays."""
    pos = icosphere_fbm(CFG1_SUBDIV, 1.0, 0.1, seed)
    org, d = sphere_rays(n_rays, seed + 1)
    return {"positions": pos, "origins": org, "dirs": d}
