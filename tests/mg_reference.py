"""numpy restatement of the multigrid V-cycle of the Stokes velocity block on
top of the oracle's operators (test infrastructure).

linalg/solvers/multigrid.hpp:288-432 (pre-smooth, residual, restriction,
zero coarse error, recursion, additive prolongation, post-smooth; PCG on the
coarsest level), linalg/solvers/chebyshev.hpp:102-165 (rho carried across the
outer iterations), linalg/solvers/pcg.hpp:81-170 (identity preconditioner),
linalg/solvers/power_iteration.hpp:13-58 (spectral radius of D^{-1} A from a
random start vector; the start vector is the library's SplitMix64 stream, so
the estimates agree with the device to rounding), with the coarse-level
viscosity injected from the finest level.
"""
import numpy as np

import oracle as O


M64 = (1 << 64) - 1


def splitmix_field(shape, ncomp, seed):
    """Per-component SplitMix64 uniforms in [0, 1) over the flat node copies."""
    n = int(np.prod(shape))
    i = np.arange(1, n + 1, dtype=np.uint64)
    out = np.empty((ncomp, n))
    with np.errstate(over="ignore"):
        for c in range(ncomp):
            base = np.uint64((seed * 0xD1B54A32D192ED03 + c * 0x8CB92BA72F3D8DD7) & M64)
            z = base + i * np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z ^= z >> np.uint64(31)
            out[c] = (z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return out.reshape((ncomp,) + tuple(shape))


class VCycle:
    def __init__(self, lmin, lmax, k_fine, bcs, order=2, steps=2, power_iterations=10, seed=12345,
                 coarse=(50, 1e-6, 1e-16)):
        self.meshes = [O.Mesh(l) for l in range(lmin, lmax + 1)]
        self.bcs = bcs
        self.ks = [None] * len(self.meshes)
        self.ks[-1] = np.asarray(k_fine, dtype=np.float64)
        for i in range(len(self.meshes) - 2, -1, -1):
            self.ks[i] = np.ascontiguousarray(self.ks[i + 1][:, ::2, ::2, ::2])
        self.invd = [1.0 / self.A(i, np.ones((3,) + m.shape), diagonal=True) for i, m in enumerate(self.meshes)]
        self.order, self.steps = order, steps
        self.coarse = coarse
        # (pre, post) estimate per level; level 0 is the coarsest
        self.est = [tuple(self.power_iteration(i, power_iterations, seed + 1000 * i + post) for post in (0, 1))
                    for i in range(len(self.meshes))]

    def power_iteration(self, i, iterations, seed):
        it = splitmix_field(self.meshes[i].shape, 3, seed)
        it = it / np.sqrt(self.dot(i, it, it))
        aux = self.invd[i] * self.A(i, it)
        radius = 0.0
        for _ in range(iterations):
            it = aux / np.sqrt(self.dot(i, aux, aux))
            aux = self.invd[i] * self.A(i, it)
            radius = self.dot(i, it, aux)
        return radius if radius > 0.0 else 1.0

    def A(self, i, x, diagonal=False):
        return self.meshes[i].epsdivdiv(self.ks[i], x, *self.bcs, diagonal=diagonal, path="fast")

    def dot(self, i, a, b):
        return self.meshes[i].masked_dot(a, b)

    def cheb(self, i, x, b, post):
        est = self.est[i][post]
        lmax, lmin = 1.5 * est, 0.1 * est
        th, de = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
        sig = th / de
        rho = 1.0 / sig
        for _ in range(self.steps):
            z = self.invd[i] * (b - self.A(i, x))
            d = z / th
            x = x + d
            for _ in range(2, self.order + 1):
                rn = 1.0 / (2.0 * sig - rho)
                z = self.invd[i] * (b - self.A(i, x))
                d = (2.0 * rn / de) * z + (rn * rho) * d
                x = x + d
                rho = rn
        return x

    def pcg(self, x, b):
        maxit, rtol, atol = self.coarse
        r = b - self.A(0, x)
        z = r.copy()
        p = z.copy()
        r0 = np.sqrt(self.dot(0, r, r))
        if r0 < atol:
            return x
        for _ in range(maxit):
            an = self.dot(0, z, r)
            ap = self.A(0, p)
            a = an / self.dot(0, ap, p)
            x = x + a * p
            r = r - a * ap
            res = np.sqrt(self.dot(0, r, r))
            if res / r0 <= rtol or res < atol:
                return x
            z = r.copy()
            p = z + (self.dot(0, z, r) / an) * p
        return x

    def cycle(self, i, x, b):
        if i == 0:
            return self.pcg(x, b)
        x = self.cheb(i, x, b, 0)
        r = b - self.A(i, x)
        rc = O.restrict(self.meshes[i], self.meshes[i - 1], r)
        ec = self.cycle(i - 1, np.zeros_like(rc), rc)
        x = x + O.prolongate(self.meshes[i - 1], self.meshes[i], ec)
        return self.cheb(i, x, b, 1)

    def __call__(self, x, b):
        return self.cycle(len(self.meshes) - 1, x, b)
