"""Seeded synthetic node fields for the benchmark and test workloads.

A value belongs to a physical node, not to a node copy: every copy of a
duplicated interface node gets the value of the node, so fields are
consistent across subdomains.  The value of a node is a function of
SplitMix64(seed, component) and a canonical node id that does not depend on
the subdomain split:

    id = ((d * (E+1) + i) * (E+1) + j) * (Lr+1) + g

with d the lowest diamond holding the node, (i, j) the node's global lateral
index in that diamond (E = 2^lateral_level) and g its global shell index
(Lr = 2^radial_level).

uniform: 53-bit mantissa mapped to [-1, 1).
normal : Box-Muller on two uniforms.
"""
from __future__ import annotations

import numpy as np

_M = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x):
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def subdomain_info(lateral_sdr, radial_sdr):
    """(diamond, sx, sy, sr) per local subdomain, enumeration order diamond, x, y, r."""
    S, Sr = 1 << lateral_sdr, 1 << radial_sdr
    out = [(d, x, y, r) for d in range(10) for x in range(S) for y in range(S) for r in range(Sr)]
    return np.array(out, dtype=np.int64)


def node_ids(coords, lateral_level, radial_level, lateral_sdr, radial_sdr, nr):
    """Canonical node id for every node copy, shape (nsd, nx, nx, nr)."""
    nsd, nx = coords.shape[0], coords.shape[1]
    E, Lr = 1 << lateral_level, 1 << radial_level
    info = subdomain_info(lateral_sdr, radial_sdr)
    assert info.shape[0] == nsd
    base = E // (1 << lateral_sdr)
    d = np.broadcast_to(info[:, 0, None, None], (nsd, nx, nx))
    i = info[:, 1, None, None] * base + np.arange(nx)[None, :, None]
    j = info[:, 2, None, None] * base + np.arange(nx)[None, None, :]
    i = np.broadcast_to(i, (nsd, nx, nx))
    j = np.broadcast_to(j, (nsd, nx, nx))
    key = np.ascontiguousarray(coords.reshape(-1, 3)).view(np.uint64)
    _, inv = np.unique(key, axis=0, return_inverse=True)
    inv = inv.ravel()
    nlat = inv.max() + 1
    lat_id = (d.ravel() * (E + 1) + i.ravel()) * (E + 1) + j.ravel()
    # owner lateral id: copy with the lowest diamond (ties share (i, j))
    order = np.lexsort((d.ravel(), inv))
    first = np.ones(order.size, dtype=bool)
    first[1:] = inv[order][1:] != inv[order][:-1]
    owner_lat = np.empty(nlat, dtype=np.int64)
    owner_lat[inv[order][first]] = lat_id[order][first]
    lat_canon = owner_lat[inv].reshape(nsd, nx, nx)
    sr = info[:, 3]
    g = sr[:, None] * (nr - 1) + np.arange(nr)[None, :]  # (nsd, nr)
    return (lat_canon[..., None] * (Lr + 1) + g[:, None, None, :]).astype(np.uint64)


def _stream(seed, comp, ids):
    s = splitmix64(np.uint64(seed) * np.uint64(256) + np.uint64(comp))
    return splitmix64(s ^ ids)


def uniform(seed, comp, ids):
    h = _stream(seed, comp, ids)
    return (h >> np.uint64(11)).astype(np.float64) * (2.0 ** -53) * 2.0 - 1.0


def normal(seed, comp, ids):
    h1 = _stream(seed, comp, ids)
    h2 = splitmix64(h1)
    u1 = ((h1 >> np.uint64(11)).astype(np.float64) + 1.0) * (2.0 ** -53)
    u2 = (h2 >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


class Workload:
    """Node ids and radii of one mesh, for generating fields."""

    def __init__(self, coords, radii, lateral_level, radial_level, lateral_sdr=0, radial_sdr=0):
        nr = radii.shape[1]
        self.ids = node_ids(coords, lateral_level, radial_level, lateral_sdr, radial_sdr, nr)
        self.shape = self.ids.shape
        self.r = np.broadcast_to(radii[:, None, None, :], self.shape)
        self.coords = coords
        self.r_min, self.r_max = float(radii.min()), float(radii.max())

    def velocity(self, seed=101):
        return np.stack([uniform(seed, c, self.ids) for c in range(3)])

    def eta_lognormal(self, seed=102, sigma=1.0):
        return np.exp(sigma * normal(seed, 0, self.ids))

    def eta_layered(self, seed=102, jump=1e3, r_jump=0.85, sigma=0.5):
        """Layered viscosity: jump below r_jump times log-normal(sigma) noise."""
        return np.where(self.r < r_jump, jump, 1.0) * np.exp(sigma * normal(seed, 0, self.ids))

    def eta_radial_contrast(self, seed=103, contrast=1e4, sigma=1.0):
        """Radial profile falling by `contrast` from the inner to the outer
        surface (exponential in r) times log-normal(sigma) noise."""
        t = (self.r_max - self.r) / (self.r_max - self.r_min)
        return np.exp(np.log(contrast) * t) * np.exp(sigma * normal(seed, 0, self.ids))

    def eta_plumes(self, seed=105, jump=1e2, r_jump=0.85, n_plumes=6, anomaly=1e2, width=0.2, sigma=1.0):
        """Viscosity of the largest single-GPU case: `jump` below r_jump,
        divided by up to `anomaly` in n_plumes radial columns around seeded
        axes (Gaussian in the angle from the nearest axis, `width` radians),
        times log-normal(sigma) noise."""
        ax = np.stack([normal(seed, 1 + c, np.arange(n_plumes, dtype=np.uint64)) for c in range(3)], axis=1)
        ax /= np.linalg.norm(ax, axis=1, keepdims=True)
        x = self.coords / np.linalg.norm(self.coords, axis=-1, keepdims=True)
        shape = np.zeros(self.coords.shape[:3])
        for a in ax:
            ang = np.arccos(np.clip(x @ a, -1.0, 1.0))
            shape = np.maximum(shape, np.exp(-0.5 * (ang / width) ** 2))
        return (np.where(self.r < r_jump, jump, 1.0) * np.exp(-np.log(anomaly) * shape)[..., None]
                * np.exp(sigma * normal(seed, 0, self.ids)))

    def temperature_harmonics(self, seed=104, max_degree=8):
        """sum_{l <= max_degree, |m| <= l} a_lm Y_lm(lat, lon) * sin(pi (r - r_min) / (r_max - r_min)),
        real orthonormal spherical harmonics, a_lm ~ N(0, 1) from the seed (component l*l + l + m)."""
        from scipy.special import sph_harm_y

        x = self.coords.reshape(-1, 3)
        theta = np.arccos(np.clip(x[:, 2] / np.linalg.norm(x, axis=1), -1.0, 1.0))
        phi = np.arctan2(x[:, 1], x[:, 0])
        lat = np.zeros(x.shape[0])
        for l in range(max_degree + 1):
            for m in range(-l, l + 1):
                a = float(normal(seed, l * l + l + m, np.zeros(1, dtype=np.uint64))[0])
                y = sph_harm_y(l, abs(m), theta, phi)
                if m > 0:
                    y = np.sqrt(2.0) * (-1) ** m * y.real
                elif m < 0:
                    y = np.sqrt(2.0) * (-1) ** m * y.imag
                else:
                    y = y.real
                lat += a * y
        lat = lat.reshape(self.coords.shape[:3])
        radial = np.sin(np.pi * (self.r[0, 0, 0] - self.r_min) / (self.r_max - self.r_min))
        return lat[..., None] * radial[None, None, None, :]
