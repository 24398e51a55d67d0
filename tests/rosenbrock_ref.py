"""Dense NumPy Rosenbrock replay and tangent-linear model (test infrastructure).

Independent of the device code: the mechanism enters only through the plan's
flux groups (reactant powers and stoichiometry), the step sizes are a given
(frozen) sequence, and every linear algebra operation is a dense
numpy.linalg.solve.  Used as the reference tangent-linear run for the device
adjoint (dot-product identity) and as a second forward replay.

All arrays carry the cells on the first axis.  y: (cells, N) physical order
(fixed species included, they stay constant); k: (cells, NG).
"""

from __future__ import annotations

import json
import os

import numpy as np

from mkpp_amd.gen.tableau import TABLEAUX

PLANS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mkpp_amd", "plans")


class DenseMechanism:
    def __init__(self, name: str):
        with open(os.path.join(PLANS, name + ".json")) as f:
            p = json.load(f)
        self.n = len(p["species"])
        self.fixed = np.array(p["fixed"], dtype=bool)
        self.var = np.flatnonzero(~self.fixed)
        groups = p["groups"]
        self.ng = len(groups)
        # reactant entries padded to 3 with a slot that reads 1.0 (index n)
        self.s = np.full((self.ng, 3), self.n, dtype=int)
        self.q = np.zeros((self.ng, 3), dtype=int)
        self.nu = np.zeros((self.n, self.ng))
        for g, grp in enumerate(groups):
            for e, (sp, q) in enumerate(grp["powers"]):
                self.s[g, e] = sp
                self.q[g, e] = q
            for sp, c in grp["nu"]:
                if not self.fixed[sp]:
                    self.nu[sp, g] += c
        self.unit = np.array(p["unit_value"], dtype=bool)

    def _yext(self, y):
        ye = np.concatenate([y, np.ones((y.shape[0], 1))], axis=1)
        ye[:, :self.n][:, self.unit] = 1.0
        return ye

    def monomial(self, y):
        """pw, d1, d2 of every group: (cells, NG, 3), (cells, NG, 3), (cells, NG, 3, 3)."""
        ye = self._yext(y)
        v = ye[:, self.s]                                   # (cells, NG, 3)
        q = self.q[None]
        pw = np.where(q == 0, 1.0, v ** q)
        pm1 = np.where(q == 0, 0.0, q * v ** np.maximum(q - 1, 0))
        pm2 = np.where(q <= 1, 0.0, q * (q - 1) * v ** np.maximum(q - 2, 0))
        d1 = np.empty_like(pw)
        d2 = np.empty(pw.shape + (3,))
        for e in range(3):
            others = [o for o in range(3) if o != e]
            d1[..., e] = pm1[..., e] * pw[..., others[0]] * pw[..., others[1]]
            for f in range(3):
                if f == e:
                    d2[..., e, f] = pm2[..., e] * pw[..., others[0]] * pw[..., others[1]]
                else:
                    (o,) = [o for o in range(3) if o not in (e, f)]
                    d2[..., e, f] = pm1[..., e] * pm1[..., f] * pw[..., o]
        return pw, d1, d2

    def rates(self, y, k):
        pw, _, _ = self.monomial(y)
        flux = k * pw.prod(axis=2)
        return flux @ self.nu.T                              # (cells, N), fixed rows 0

    def jacobian(self, y, k):
        """dF/dy, (cells, N, N) (fixed columns included, fixed rows 0)."""
        _, d1, _ = self.monomial(y)
        cells = y.shape[0]
        D = np.zeros((cells, self.ng, self.n + 1))
        for e in range(3):
            np.add.at(D, (slice(None), np.arange(self.ng), self.s[:, e]), k * d1[..., e])
        return np.einsum("ig,cgj->cij", self.nu, D[..., :self.n])

    def _dy_flux(self, dy):
        """dy as the fluxes see it: a unit-value species reads as the constant
        1.0, so its direction does not enter any monomial."""
        d = np.array(dy, dtype=np.float64, copy=True)
        d[:, self.unit] = 0.0
        return d

    def djac_dir(self, y, k, K, dy, dk):
        """(d/dy J . dy + d/dk J . dk) K for every cell: (cells, N).  J's
        columns of unit-value species (evaluated at 1.0) multiply their own K."""
        pw, d1, d2 = self.monomial(y)
        z = np.zeros((y.shape[0], 1))
        Ke = np.concatenate([K, z], axis=1)[:, self.s]      # (cells, NG, 3)
        de = np.concatenate([self._dy_flux(dy), z], axis=1)[:, self.s]
        second = np.einsum("cgef,cge,cgf->cg", d2, Ke, de)
        first = np.einsum("cge,cge->cg", d1, Ke)
        return (k * second + dk * first) @ self.nu.T

    def drates(self, y, k, dy, dk):
        """dF = J dy + dF/dk dk."""
        pw, d1, _ = self.monomial(y)
        z = np.zeros((y.shape[0], 1))
        de = np.concatenate([self._dy_flux(dy), z], axis=1)[:, self.s]
        return (k * np.einsum("cge,cge->cg", d1, de) + dk * pw.prod(axis=2)) @ self.nu.T


def replay_tlm(mech: DenseMechanism, y0, k, hs, nsteps, method="rodas3", dy0=None, dk=None, states=None):
    """Integrate the frozen step sequences (hs[cell, :nsteps[cell]]) from y0 and,
    if dy0/dk are given, the tangent-linear model along it.  Returns (y, dy).

    states (cells, max_steps, N), optional: the state at the start of every
    accepted step of another run of the same map (a device checkpoint).  Each
    step then starts from that state, so the tangent-linear model linearises
    around that run's trajectory instead of this replay's own, which for an
    ill-conditioned mechanism drifts away from it over many steps although
    each step agrees to rounding."""
    tab = TABLEAUX[method]
    S = tab.stages
    g = tab.gamma[0]
    var = mech.var
    y = y0.copy()
    dy = None if dy0 is None else dy0.copy()
    dk = np.zeros_like(k) if (dy0 is not None and dk is None) else dk
    cells = y.shape[0]
    eye = np.eye(len(var))
    for n in range(int(nsteps.max())):
        act = n < nsteps
        if not act.any():
            break
        h = np.where(act, hs[:, n] if n < hs.shape[1] else 1.0, 1.0)
        if states is not None:
            y = np.where(act[:, None], states[:, min(n, states.shape[1] - 1)], y)
        J = mech.jacobian(y, k)[:, var][:, :, var]
        W = eye[None] / (g * h)[:, None, None] - J
        K, dK = [], []
        F = dF = None
        for s in range(S):
            if s == 0 or tab.new_f[s]:
                Ys = y.copy()
                for j in range(s):
                    if tab.A(s, j) != 0.0:
                        Ys[:, var] += tab.A(s, j) * K[j]
                F = mech.rates(Ys, k)[:, var]
                if dy is not None:
                    dYs = dy.copy()
                    for j in range(s):
                        if tab.A(s, j) != 0.0:
                            dYs[:, var] += tab.A(s, j) * dK[j]
                    dF = mech.drates(Ys, k, dYs, dk)[:, var]
            rhs = F.copy()
            for j in range(s):
                if tab.C(s, j) != 0.0:
                    rhs += (tab.C(s, j) / h)[:, None] * K[j]
            Ks = np.linalg.solve(W, rhs[..., None])[..., 0]
            K.append(Ks)
            if dy is not None:
                drhs = dF.copy()
                for j in range(s):
                    if tab.C(s, j) != 0.0:
                        drhs += (tab.C(s, j) / h)[:, None] * dK[j]
                Kfull = np.zeros_like(y)
                Kfull[:, var] = Ks
                # K = W^-1 rhs, W = I/(g h) - J(y):  dK = W^-1 (drhs + dJ K)
                dJK = mech.djac_dir(y, k, Kfull, dy, dk)[:, var]
                dK.append(np.linalg.solve(W, (drhs + dJK)[..., None])[..., 0])
        ynew = y.copy()
        for j in range(S):
            if tab.m[j] != 0.0:
                ynew[:, var] += tab.m[j] * K[j]
        y = np.where(act[:, None], ynew, y)
        if dy is not None:
            dnew = dy.copy()
            for j in range(S):
                if tab.m[j] != 0.0:
                    dnew[:, var] += tab.m[j] * dK[j]
            dy = np.where(act[:, None], dnew, dy)
    return y, dy
