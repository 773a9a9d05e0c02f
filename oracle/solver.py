"""Explicit multiblock driver (oracle).  Follows reference `wmles/solver.py:16-232`.

State = tuple of float64 arrays (5, Ni, Nj, Nk) including ghosts.  Nothing is
mutated: every call returns new arrays, as in the reference.
"""
import numpy as np

from .bc import WallModel
from .gas import _sum3
from .mesh import Interface, exchange, sl
from .numerics import green_gauss, viscous_flux, ducros, make_sgs


class ConstantMassFlux:
    """Shift the volume-averaged rho u_dir of all block interiors to `target`,
    keeping the internal energy (solver.py:16-36)."""

    def __init__(self, direction, target):
        self.dir, self.target = direction, target

    def delta(self, Q, geoms, blocks):
        num, den = 0.0, 0.0
        for q, g, b in zip(Q, geoms, blocks):
            vol = b.interior(g.vol)
            num = num + np.sum(b.interior(q[1 + self.dir]) * vol)
            den = den + np.sum(vol)
        return self.target - num / den

    def __call__(self, Q, geoms, blocks):
        delta = self.delta(Q, geoms, blocks)
        out = []
        for q in Q:
            z = np.zeros_like(q[0])
            mom = [z, z, z]
            mom[self.dir] = z + delta
            dE = (2.0 * q[1 + self.dir] * delta + delta ** 2) / (2.0 * q[0])
            out.append(q + np.stack([z] + mom + [dE]))
        return tuple(out)


class Solver:
    """solver.py:39-232 (without the JIT wrappers)."""

    def __init__(self, mesh, gas, scheme, sgs="vreman", cfl=0.8, sgs_kwargs=None,
                 constraints=(), sources=(), bc_passes=2):
        self.mesh, self.blocks, self.gas = mesh, mesh.blocks, gas
        self.geoms = mesh.geoms
        self.scheme, self.cfl = scheme, cfl
        self.sgs = make_sgs(sgs, **(sgs_kwargs or {}))
        self.constraints, self.sources = tuple(constraints), tuple(sources)
        self.bc_passes = bc_passes

    def init_state(self, prim_fn):
        # solver.py:54-56: prim_fn at every cell centre, ghosts included
        return tuple(self.gas.prim_to_cons(np.asarray(prim_fn(g.xc), float)) for g in self.geoms)

    def apply_bcs(self, Q, geoms=None, time=0.0):
        """Sequential sweep: passes x axes x blocks x sides (solver.py:59-73)."""
        geoms = self.geoms if geoms is None else geoms
        Q = list(Q)
        for _ in range(self.bc_passes):
            for a in range(3):
                for ib, blk in enumerate(self.blocks):
                    q = Q[ib]
                    for s in (0, 1):
                        f = blk.faces[(a, s)]
                        if isinstance(f, Interface):
                            src = q if f.nb == ib else Q[f.nb]
                            q = exchange(q, src, blk, (a, s), self.blocks[f.nb], f, "cell")
                        else:
                            q = f.fill_at_time(q, blk, geoms[ib], self.gas, a, s, time)
                    Q[ib] = q
        return tuple(Q)

    def block_rhs(self, ib, q, geom):
        """solver.py:76-95."""
        blk, gas = self.blocks[ib], self.gas
        w, T = gas.cons_to_prim(q)
        mu = gas.mu(T)
        phi = np.concatenate([w[1:4], T[None]], axis=0)
        g = green_gauss(phi, geom)
        mut = self.sgs(g[:3], geom.vol ** (1.0 / 3.0), w[0])
        mu_eff = mu + mut
        k_eff = gas.cp * (mu / gas.Pr + mut / gas.Pr_t)
        theta = ducros(g[:3])
        rhs = 0.0
        for d in range(3):
            F = self.scheme(blk, geom, w, gas, d, theta) + viscous_flux(blk, geom, d, phi, g,
                                                                         mu_eff, k_eff)
            for s in (0, 1):
                bc = blk.faces[(d, s)]
                if getattr(bc, "flux_override", False):
                    Fw, _, _ = bc.wall_flux(w, T, blk, geom, gas, d, s)
                    F = F.copy()
                    F[sl(F.ndim, d, 0 if s == 0 else blk.n[d])] = Fw
            rhs = rhs - (F[sl(F.ndim, d, slice(1, None))] - F[sl(F.ndim, d, slice(None, -1))])
        return rhs / blk.interior(geom.vol)

    def rhs(self, Q, time=0.0):
        """solver.py:97-106: ghosts of the result are 0."""
        Q = self.apply_bcs(Q, self.geoms, time)
        ng = self.mesh.ng
        out = []
        for ib, (q, g) in enumerate(zip(Q, self.geoms)):
            r = self.block_rhs(ib, q, g)
            for src in self.sources:
                r = r + src(self.blocks[ib], q, g, self.gas)
            out.append(np.pad(r, [(0, 0)] + [(ng, ng)] * 3))
        return tuple(out)

    @staticmethod
    def rk1(Q, R, dt):
        return tuple(q + dt * r for q, r in zip(Q, R))                     # solver.py:110-111

    @staticmethod
    def rk2(Q0, Q1, R1, dt):
        return tuple(0.75 * q0 + 0.25 * (q1 + dt * r1) for q0, q1, r1 in zip(Q0, Q1, R1))

    @staticmethod
    def rk3(Q0, Q2, R2, dt):
        return tuple((1.0 / 3.0) * q0 + (2.0 / 3.0) * (q2 + dt * r2) for q0, q2, r2 in zip(Q0, Q2, R2))

    def advance(self, Q, dt, time=0.0):
        """Shu-Osher SSP-RK3, stage times t, t+dt, t+dt/2, then constraints
        (solver.py:123-133)."""
        R0 = self.rhs(Q, time)
        Q1 = self.rk1(Q, R0, dt)
        R1 = self.rhs(Q1, time + dt)
        Q2 = self.rk2(Q, Q1, R1, dt)
        R2 = self.rhs(Q2, time + 0.5 * dt)
        Q3 = self.rk3(Q, Q2, R2, dt)
        for c in self.constraints:
            Q3 = c(Q3, self.geoms, self.blocks)
        return Q3

    def max_dt(self, Q, time=0.0):
        """CFL time step over all interiors (solver.py:135-154).  The reference's
        ghost fill before it does not reach the interior-only result, so it is
        not repeated here."""
        gas = self.gas
        dts = []
        for blk, q, g in zip(self.blocks, Q, self.geoms):
            w, T = gas.cons_to_prim(q)
            c = gas.sound_speed(T)
            nu = max(4.0 / 3.0, gas.gamma / gas.Pr) * gas.mu(T) / w[0]
            lam, lamv = 0.0, 0.0
            for d in range(3):
                Sd = g.S[d]
                Sb = 0.5 * (Sd[sl(4, d, slice(0, -1))] + Sd[sl(4, d, slice(1, None))])
                Sm = np.sqrt(_sum3(Sb * Sb))
                lam = lam + np.abs(_sum3(w[1:4] * Sb)) + c * Sm
                lamv = lamv + nu * Sm ** 2 / g.vol
            with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
                dts.append(np.min(blk.interior(self.cfl / ((lam + 4.0 * lamv) / g.vol))))
        return np.min(np.stack(dts))

    def wall_stats(self, Q):
        """{(block, face): (tau_w, q_w)} for WallModel faces (solver.py:163-179)."""
        out = {}
        for blk, q, g in zip(self.blocks, Q, self.geoms):
            w, T = self.gas.cons_to_prim(q)
            for (a, s), bc in blk.faces.items():
                if isinstance(bc, WallModel):
                    _, tau, qw = bc.wall_flux(w, T, blk, g, self.gas, a, s)
                    out[(blk.name, "ijk"[a] + ("min", "max")[s])] = (tau, qw)
        return out

    def wall_loads(self, Q, moment_center=(0.0, 0.0, 0.0)):
        """Integrated wall force and moment per WallModel face (solver.py:181-232)."""
        center = np.asarray(moment_center, dtype=float)
        if center.shape != (3,) or not np.all(np.isfinite(center)):     # solver.py:187-189
            raise ValueError("moment_center must contain three finite coordinates")
        out = {}
        for blk, q, g in zip(self.blocks, Q, self.geoms):
            w, T = self.gas.cons_to_prim(q)
            for (axis, side), bc in blk.faces.items():
                if not isinstance(bc, WallModel):
                    continue
                ng, n = blk.ng, blk.n[axis]
                fi = ng if side == 0 else ng + n
                mi = ng + bc.n_match - 1 if side == 0 else ng + n - bc.n_match
                ci = ng if side == 0 else ng + n - 1
                Sv = blk.take(g.S[axis], axis, fi)
                area = np.sqrt(_sum3(Sv * Sv))
                nrm = Sv / area
                if side == 0:
                    nrm = -nrm
                vel = blk.take(w[1:4], axis, mi)
                ut = vel - _sum3(vel * nrm)[None] * nrm
                speed = np.sqrt(_sum3(ut ** 2) + 1e-20)
                tang = ut / speed
                _, tau, _ = bc.wall_flux(w, T, blk, g, self.gas, axis, side)
                pw = blk.take(w[4], axis, ci)
                traction = (pw[None] * nrm + tau[None] * tang) * area[None]
                fc = blk.take(g.xf[axis], axis, fi)
                arm = fc - center.reshape((3,) + (1,) * (fc.ndim - 1))
                mom = np.stack([arm[1] * traction[2] - arm[2] * traction[1],
                                arm[2] * traction[0] - arm[0] * traction[2],
                                arm[0] * traction[1] - arm[1] * traction[0]])
                ax = tuple(range(1, traction.ndim))
                out[(blk.name, "ijk"[axis] + ("min", "max")[side])] = (
                    np.sum(traction, axis=ax), np.sum(mom, axis=ax))
        return out
