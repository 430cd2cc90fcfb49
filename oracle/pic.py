"""Oracle: FP64 particle-in-cell coupling for the 1-D Ex/Hy grid.  TEST INFRASTRUCTURE ONLY.

The reference has no particle code: its current sink is V.Jx, consumed by
ADE_ExUpdate (BaseFDTD11.py:663-669) as Ex += (dHy - Jx)*UpExMat, i.e. in
units of J_x*dz (A/m).  This module restates the textbook algorithms the
build adds in front of that sink, in the normalised units of
oracle/physical.py (E [V/m], H' = Z0*Hy_ref, J' = Z0*Jx_ref):

Geometry.  E = (Ex, 0, 0) on integer nodes x = n (cell units), H' on
half-integer nodes x = n + 1/2.  The reference's Hy update
(BaseFDTD11.py:640-656) has the sign of ``dHy/dt = +(1/mu0) dEx/dz``, so its
Hy is minus the physical H_y and the physical magnetic field is
B_y = -H'/c.

Step n (before the field update of step n, which consumes J^{n+1/2}):
  gather   E^n at x^n (linear between nodes), B^n from the average of H'
           before and after the last H update (H^{n-1/2}, H^{n+1/2}), linear
           between half-integer nodes;
  push     relativistic Boris (Boris 1970; Birdsall & Langdon 15-4) on
           u = gamma*v/c with a = q dt/(2 m c):
             u- = u + a E x^,  t_y = -a H'/gamma(u-),  s_y = 2 t_y/(1+t_y^2)
             u' = u- + u- x t, u+ = u- + u' x s,  u^{n+1/2} = u+ + a E x^
           x^{n+1} = x^n + Sc * u_z/gamma   (c dt/dz = Sc)
  deposit  CIC of J' = Z0 q w c * u_x/gamma at x^{n+1/2} = x^n + Sc*u_z/(2 gamma)
           onto the E nodes floor(x) and floor(x)+1 with weights 1-f, f.
Absorption: live particles occupy cells [2, N-3]; a particle whose new cell
leaves that range is parked in cell 1 (left) or N-2 (right), stops moving
and deposits nothing from then on.
"""
import math

import numpy as np

from . import physical as ph

E_CHARGE = 1.602176634e-19
M_E = 9.1093837139e-31          # CODATA 2022 (scipy.constants.m_e in scipy 1.15)
Z0 = 376.730313668
C0 = ph.C0


class Beam:
    """FP64 particle state: position x (cell units), u = gamma v / c, alive flag, id."""

    def __init__(self, x, u, q, m, w, dt, N, Sc, ids=None):
        self.x = np.asarray(x, dtype=np.float64).copy()
        self.u = np.asarray(u, dtype=np.float64).reshape(-1, 3).copy()
        self.alive = np.ones(len(self.x), dtype=bool)
        self.ids = np.arange(len(self.x)) if ids is None else np.asarray(ids).copy()
        self.a = q * dt / (2.0 * m * C0)
        self.jscale = Z0 * q * w * C0
        self.N, self.Sc = N, Sc
        cell = np.floor(self.x)
        if np.any(cell < 2) or np.any(cell > N - 3):
            raise ValueError("particles must start in cells [2, N-3]")


def gather(x, E, Hprev, Hcur):
    """Linear gather of E (integer nodes) and the time-averaged H' (half-integer nodes)."""
    i = np.floor(x).astype(np.int64)
    f = x - i
    Eg = (1.0 - f) * E[i] + f * E[i + 1]
    xh = x - 0.5
    j = np.floor(xh).astype(np.int64)
    g = xh - j
    Hav = 0.5 * (Hprev + Hcur)
    Hg = (1.0 - g) * Hav[j] + g * Hav[j + 1]
    return Eg, Hg


def boris(u, Eg, Hg, a):
    """One relativistic Boris step for E = (Eg,0,0), B = (0,-Hg/c,0); returns u^{n+1/2}."""
    um = u.copy()
    um[:, 0] += a * Eg
    gm = np.sqrt(1.0 + np.sum(um * um, axis=1))
    ty = -a * Hg / gm
    sy = 2.0 * ty / (1.0 + ty * ty)
    upx = um[:, 0] - um[:, 2] * ty
    upz = um[:, 2] + um[:, 0] * ty
    un = um.copy()
    un[:, 0] = um[:, 0] - upz * sy
    un[:, 2] = um[:, 2] + upx * sy
    un[:, 0] += a * Eg
    return un


def push_deposit(b, E, Hprev, Hcur):
    """Advance the live particles one step; return J' (length N+1) for the field step."""
    N, Sc = b.N, b.Sc
    J = np.zeros(N + 1)
    idx = np.nonzero(b.alive)[0]
    if len(idx) == 0:
        return J
    x = b.x[idx]
    Eg, Hg = gather(x, E, Hprev, Hcur)
    un = boris(b.u[idx], Eg, Hg, b.a)
    gam = np.sqrt(1.0 + np.sum(un * un, axis=1))
    bx, bz = un[:, 0] / gam, un[:, 2] / gam
    xh = x + 0.5 * Sc * bz
    xn = x + Sc * bz
    m0 = np.floor(xh).astype(np.int64)
    f = xh - m0
    q = b.jscale * bx
    J += np.bincount(m0, weights=q * (1.0 - f), minlength=N + 1)[:N + 1]
    J += np.bincount(m0 + 1, weights=q * f, minlength=N + 1)[:N + 1]
    b.u[idx] = un
    cell = np.floor(xn)
    left, right = cell < 2, cell > N - 3
    xn = np.where(left, 1.5, np.where(right, N - 2 + 0.5, xn))
    b.x[idx] = xn
    b.alive[idx[left | right]] = False
    return J


def run(m, beam, steps=None, record=True):
    """Physical-model run with the beam coupled: push/deposit, then the field step."""
    steps = m["steps"] if steps is None else steps
    s = ph.State(m["N"], m["R"] - m["F"])
    Hprev = np.zeros(m["N"] + 1)
    probes = np.zeros((len(m["probes"]), steps))
    for t in range(steps):
        J = push_deposit(beam, s.E, Hprev, s.H)
        Hprev = s.H.copy()
        ph.step(m, s, t, J)
        if record:
            probes[:, t] = s.E[m["probes"]]
    return s, probes


def beam_params(p, n, dz):
    """(q, m, w): electrons with the macro-particle weight per unit transverse area
    w = density * extent / n (m^-2)."""
    bm = p["beam"]
    extent = (bm["cell_hi"] - bm["cell_lo"]) * dz
    return -E_CHARGE, M_E, bm["density"] * extent / n
