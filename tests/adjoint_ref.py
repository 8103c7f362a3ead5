"""FP64 reference for the adjoint (test infrastructure).

A torch restatement of the oracle's step (oracle/sph_oracle.cpp: density, alpha, gravity, the
DF/CD loops, the viscosity/surface-tension solve, collision and advection) for one scene with
either kernel set (cubic spline, or poly6 density with the spiky gradient), per-particle radii
(support h_i = 4 r_i, neighbours within (h_i + h_j)/2, every kernel of particle i evaluated
with h_i) and no Z-sort, with its control flow frozen: the DF/CD iteration
counts are given per step (from the run being checked), the neighbour lists are recomputed from
the positions (the kernels vanish smoothly at the support, so list changes do not matter), the
viscosity solve is solved exactly (dense), which is the system the forward solve converges to:
A_f v* = b + f v_in. Gradients come from torch.autograd in float64; tests compare them with
central finite differences of this map and with the gfx950 hand-written adjoint.
"""
from __future__ import annotations

import numpy as np
import torch

KW6 = 1.5666814710608447114749495456982
KGS = 14.323944878270580219199538703526
KCW = 2.5464790894703253723021402139602
KCG = 15.278874536821952233812841283761
KSTW = 7.2756545413437867780061148970292
KSTG = 21.826963624031360334018344691088
EPS = 1.0e-20
FLT_EPS = 1.19209290e-7
T = torch.float64


def _safe(d):
    return torch.where(d > 0, d, torch.ones_like(d))


def W(d, h, cubic=True):
    q = d / h
    if not cubic:  # poly6
        return torch.where(q >= 1, torch.zeros_like(d), KW6 * ((1 - q * q) / h) ** 3)
    inner = KCW / h ** 3 * ((6 * q - 6) * q * q + 1)
    outer = KCW * 2 * ((1 - q) / h) ** 3
    return torch.where(q >= 1, torch.zeros_like(d), torch.where(q >= 0.5, outer, inner))


def G(d, h, cubic=True):
    """grad W_ij = G x_ij: cubic spline or spiky."""
    q = d / h
    ds = _safe(d)
    if not cubic:
        val = -KGS / ds * ((1 - q) / h ** 2) ** 2
        return torch.where((q < EPS) | (q >= 1), torch.zeros_like(d), val)
    inner = KCG / ds / h ** 4 * (3 * q - 2) * q
    outer = -KCG / ds * ((1 - q) / h ** 2) ** 2
    val = torch.where(q > 0.5, outer, inner)
    return torch.where((q < EPS) | (q >= 1), torch.zeros_like(d), val)


def STW(d, h):
    q = torch.clamp(d / h, min=0.5)
    return torch.where(q >= 1, torch.zeros_like(d), KSTW * ((1 - q) / h) ** 3)


def STG(d, h):
    q = d / h
    val = -KSTG / _safe(d) * ((1 - q) / h ** 2) ** 2
    return torch.where((q <= 0.5) | (q >= 1), torch.zeros_like(d), val)


def _pairs(a, b, cut, self_pairs):
    """Index pairs with |a_i - b_j| < cut (a scalar or an (len(a), len(b)) matrix)."""
    with torch.no_grad():
        D = torch.cdist(a, b)
        m = D < cut
        if self_pairs:
            m.fill_diagonal_(False)
        I, J = torch.nonzero(m, as_tuple=True)
    return I, J


class Replay:
    """One scene: fluid (pos, radius r and mass m: scalars or per particle), walls (bpos,
    bradius), physics, material; `cubic` picks the kernel set; `accumulate_volumes` is the
    reference's wall volume V <- 1/(V + S) carried across steps (faithful_quirks), else
    V = 1/S every step."""

    def __init__(self, pos, r, m, bpos, br, physics, material, cg_factor=0.01, cubic=True, accumulate_volumes=True):
        self.x0 = torch.as_tensor(np.asarray(pos, np.float64), dtype=T)
        self.N = self.x0.shape[0]
        self.cubic = bool(cubic)
        self.accumulate = bool(accumulate_volumes)
        r32 = np.broadcast_to(np.asarray(r, np.float32), (self.N,))
        self.h = torch.as_tensor(4.0 * r32.astype(np.float64), dtype=T)  # h_i = 4 r_i (exact in float)
        self.m = torch.as_tensor(np.broadcast_to(np.asarray(m, np.float64), (self.N,)).copy(), dtype=T)
        self.xb = torch.as_tensor(np.asarray(bpos, np.float64), dtype=T)
        self.hb = 4.0 * float(br)
        self.dt = float(np.float32(physics["dt"]))
        self.g = torch.tensor([float(np.float32(c)) for c in physics["gravity"]], dtype=T)
        self.bmin = torch.tensor([float(np.float32(c)) for c in physics["box_min"]], dtype=T)
        self.bmax = torch.tensor([float(np.float32(c)) for c in physics["box_max"]], dtype=T)
        R = material["radius"]
        self.V = float(np.float32(R * R * R / 48.0 * 3.141592))
        self.rho0 = float(np.float32(material["density"]))
        self.xsph = float(np.float32(material["xsph"]))
        self.f = float(np.float32(cg_factor))
        # walls: S_b = W(0) + sum W_bk (static), volumes V <- 1/(V + S) every step (never reset)
        with torch.no_grad():
            Ib, Jb = _pairs(self.xb, self.xb, self.hb, True)
            db = torch.linalg.norm(self.xb[Ib] - self.xb[Jb], dim=1)
            S = torch.full((self.xb.shape[0],), float(W(torch.zeros(1, dtype=T), self.hb, cubic)[0]), dtype=T)
            S.index_add_(0, Ib, W(db, self.hb, cubic))
        self.Sb = S

    def step(self, x, v, Vb, mu, sigma, ldf, lcd):
        h, dt, V, rho0, cub = self.h, self.dt, self.V, self.rho0, self.cubic
        Vb = 1.0 / (Vb + self.Sb) if self.accumulate else 1.0 / self.Sb
        I, J = _pairs(x, x, 0.5 * (h[:, None] + h[None, :]), True)
        Ib, Jb = _pairs(x, self.xb, 0.5 * (h[:, None] + self.hb), False)
        hI, hB = h[I], h[Ib]  # each pair's kernels use the support of its own particle
        xij = x[I] - x[J]
        d = torch.linalg.norm(xij, dim=1)
        xib = x[Ib] - self.xb[Jb]
        db = torch.linalg.norm(xib, dim=1)
        Gij, Gib = G(d, hI, cub), G(db, hB, cub)
        N = self.N
        zero = torch.zeros(N, dtype=T)
        # density and alpha
        rho = rho0 * (V * W(torch.zeros(N, dtype=T), h, cub) + zero.index_add(0, I, V * W(d, hI, cub))
                      + zero.index_add(0, Ib, Vb[Jb] * W(db, hB, cub)))
        grad = (V * Gij)[:, None] * xij
        s = zero.index_add(0, I, (grad * grad).sum(1))
        gvec = torch.zeros(N, 3, dtype=T).index_add(0, I, grad).index_add(0, Ib, (Vb[Jb] * Gib)[:, None] * xib)
        a = s + (gvec * gvec).sum(1)
        alpha = torch.where(a > EPS, 1.0 / torch.where(a > EPS, a, torch.ones_like(a)), torch.zeros_like(a))
        # gravity
        v = v + dt * self.g[None, :]

        def solve(v, df, L):
            for _ in range(max(L - 1, 0)):
                vij = v[I] - v[J]
                delta = zero.index_add(0, I, V * Gij * (xij * vij).sum(1)) + \
                    zero.index_add(0, Ib, Vb[Jb] * Gib * (xib * v[Ib]).sum(1))
                if df:
                    e = torch.minimum(delta * dt, rho / rho0 + dt * delta - 0.8)
                else:
                    e = rho / rho0 + dt * delta - 1.0
                kap = torch.where(e > 0, e * alpha / (dt * dt), torch.zeros_like(e))
                F = -torch.zeros(N, 3, dtype=T).index_add(0, I, (0.5 * V * (kap[I] + kap[J]) * Gij)[:, None] * xij) \
                    - torch.zeros(N, 3, dtype=T).index_add(0, Ib, (0.5 * Vb[Jb] * kap[Ib] * Gib)[:, None] * xib)
                v = v + dt * F
            return v

        v = solve(v, True, ldf)
        # viscosity + surface tension: A_f v* = b + f v_in
        m = self.m
        vin = v
        uij = vin[I] - vin[J]
        fst = zero.index_add(0, I, m[J] * (uij * (Gij[:, None] * xij)).sum(1))
        invD = 2.0 / (rho[I] + rho[J])
        stw = STW(d, hI)
        stg = STG(d, hI)[:, None] * xij
        g = (stw + dt * ((uij * stg).sum(1) - (fst[I] + fst[J]) * invD * 0.5 * stw)) * invD
        sij = 0.5 * (m[I] + m[J]) * g
        wij = m[J] / rho[J] * W(d, hI, cub)
        cij = 0.5 * (m[I] + m[J]) / rho[J] / (d * d + 0.01 * hI * hI) * Gij / (d + FLT_EPS)
        cs = dt * sigma / m[I]
        aij = dt * (cs * sij + (self.xsph / dt) * wij)
        bij = -dt * (10.0 * rho0 * mu / rho[I]) * cij
        bvec = (dt * sigma / m)[:, None] * (-torch.zeros(N, 3, dtype=T).index_add(0, I, sij[:, None] * xij)) + vin
        Mij = aij[:, None, None] * torch.eye(3, dtype=T)[None] + bij[:, None, None] * xij[:, :, None] * xij[:, None, :]
        A4 = torch.zeros(N, N, 3, 3, dtype=T)
        A4 = A4.index_put((I, I), Mij, accumulate=True).index_put((I, J), -Mij, accumulate=True)
        A = A4.permute(0, 2, 1, 3).reshape(3 * N, 3 * N) + (1.0 + self.f) * torch.eye(3 * N, dtype=T)
        v = torch.linalg.solve(A, (bvec + self.f * vin).reshape(3 * N)).reshape(N, 3)
        v = solve(v, False, lcd)
        # collision + advection (branch frozen from the values)
        pk = x + v * dt
        with torch.no_grad():
            hit = ((pk < self.bmin) & (v < 0)) | ((pk > self.bmax) & (v > 0))
            c = torch.where(hit, torch.where(v.abs() < 1.0e-4, torch.zeros_like(v), torch.full_like(v, -0.4)),
                            torch.ones_like(v))
        v = c * v
        x = x + dt * v
        return x, v, Vb

    def rollout(self, v0, mu, sigma, counts, observe=None):
        """counts: (steps, 3) = (l_df, l_cg, l_cd) per step. Returns final (x, v), or with
        `observe` (steps in 1..len(counts)) the positions after each of those steps (k, N, 3)."""
        x, v = self.x0, v0
        Vb = torch.zeros(self.xb.shape[0], dtype=T)
        seen = {}
        for t, (ldf, _, lcd) in enumerate(counts):
            x, v, Vb = self.step(x, v, Vb, mu, sigma, int(ldf), int(lcd))
            seen[t + 1] = x
        if observe is not None:
            return torch.stack([seen[int(t)] for t in observe])
        return x, v


VARIANTS = ("cubic", "poly6", "radii")


def droplet(variant="cubic", steps=4, n=300):
    """The small droplet of the adjoint tests (radius 0.08 m, h 0.04 m, seed 3) in one of
    VARIANTS: the cubic spline with equal radii, the poly6/spiky kernel set, or two fluid radii
    (half of the particles, seeded, at 0.8 r)."""
    from mps_cuda_amd import _capi, scenes
    sc = scenes.diff_droplet(n=n, radius=0.08, seed=3, steps=steps, h=0.04)
    if variant == "poly6":
        sc.config["kernel"] = _capi.KERNEL_POLY6_SPIKY
    elif variant == "radii":
        small = np.random.default_rng(11).random(sc.n) < 0.5
        sc.radius = np.where(small, np.float32(0.8) * sc.radius, sc.radius).astype(np.float32)
    elif variant != "cubic":
        raise ValueError(variant)
    return sc


def replay_for(sc, mass, bpos, bradius):
    """Replay of a droplet() scene: fluid masses and wall particles as the run produced them."""
    from mps_cuda_amd import _capi
    return Replay(sc.pos, sc.radius, mass, bpos, bradius, sc.physics, sc.material,
                  cubic=sc.config["kernel"] == _capi.KERNEL_CUBIC,
                  accumulate_volumes=bool(sc.config.get("faithful_quirks", 1)))
