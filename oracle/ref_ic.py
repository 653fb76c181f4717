"""Initial conditions of the reference -- TEST INFRASTRUCTURE ONLY.

A restatement of set_initial_conditions (src/initial_conditions.f90:46-208)
for the interior levels 1..nz and of its seeded noise (add_noise1,
:265-306, drawing from Numerical Recipes' ran3, src/ran3.f90), used by the
tests as the checker of the Fortran host's initial conditions.  The
generator is checked against the reference's own ran3 compiled into
oracle/_ref/ (tests/test_ref_ic.py).
"""
import math

import numpy as np

MBIG, MSEED = 1000000000, 161803398


class Ran3:
    """Knuth's subtractive generator as ran3 runs it: a negative seed (or the
    first call) restarts the 55-value table; every call returns
    (t[i] - t[i+31 mod 55]) mod MBIG / MBIG and stores it back in t[i]."""

    def __init__(self):
        self.t = [0] * 56
        self.i = self.j = 0
        self.started = False

    def __call__(self, seed):
        """One draw; returns (value, seed after the call) like ran3(idum)."""
        if seed < 0 or not self.started:
            self.started = True
            t = self.t
            mj = (MSEED - abs(seed)) % MBIG
            t[55] = mj
            mk = 1
            for n in range(1, 55):
                ii = (21 * n) % 55
                t[ii] = mk
                mk = mj - mk
                if mk < 0:
                    mk += MBIG
                mj = t[ii]
            for _ in range(4):
                for n in range(1, 56):
                    t[n] -= t[1 + (n + 30) % 55]
                    if t[n] < 0:
                        t[n] += MBIG
            self.i, self.j = 0, 31
            seed = 1
        self.i = 1 if self.i == 55 else self.i + 1
        self.j = 1 if self.j == 55 else self.j + 1
        mj = self.t[self.i] - self.t[self.j]
        if mj < 0:
            mj += MBIG
        self.t[self.i] = mj
        return mj * (1.0 / MBIG), seed


def reference_ic(prm, ic_type="log_law", u_star=1.0, z_turb=0.5, vnf=1.0, u0=0.0, v0=0.0, w0=0.0, coriolis=False,
                 z0=None):
    """u, v, w (nz, ny, nx) on levels 1..nz as set_initial_conditions leaves
    them before the boundary conditions; prm supplies nx, ny, nz, lz, z0,
    pfx, pfy, ug, vg and re (Poiseuille's Re).  As in the reference, the
    Coriolis start replaces every type but 'constant' (also 'files', whose
    fields are then never read), and 'files' gets no noise.  z0: a
    roughness map (ny, nx) (surface_files) instead of prm.z0."""
    nx, ny, nz = prm.nx, prm.ny, prm.nz
    dz = prm.lz / nz
    u = np.zeros((nz, ny, nx))
    v = np.zeros_like(u)
    w = np.zeros_like(u)
    norm = (prm.pfx ** 2 + prm.pfy ** 2) ** 0.5
    dx, dy = (prm.pfx / norm, prm.pfy / norm) if norm > 0.0 else (1.0, 0.0)
    for k in range(1, nz + 1):
        z = (k - 0.5) * dz
        if ic_type == "constant":
            u[k - 1], v[k - 1], w[k - 1] = u0, v0, w0
        elif coriolis:
            u[k - 1] = prm.ug
            if z <= 0.5:
                v[k - 1] = prm.vg
        elif ic_type == "log_law":
            if z0 is None:
                arg = (1.0 / 0.4) * math.log(z / prm.z0)
            else:
                arg = (1.0 / 0.4) * np.array([[math.log(z / q) for q in row] for row in z0])
            u[k - 1] = dx * arg * u_star
            v[k - 1] = dy * arg * u_star
        elif ic_type == "poiseuille":
            u[k - 1] = prm.re * z * (1.0 - 0.5 * z)
        else:
            raise ValueError(ic_type)
    if u_star > 0.0 and ic_type != "files":
        gen = Ran3()
        vf = 3.0 * vnf * (u_star * u_star)
        zturb = z_turb * prm.lz
        for rnd, x in ((4, u), (8, v), (12, w)):
            for k in range(1, nz + 1):
                z = (k - 0.5) * dz
                if z <= zturb:
                    seed = -80 - k - rnd
                    for j in range(ny):
                        for i in range(nx):
                            r, seed = gen(seed)
                            noise = r * 2.0 - 1.0
                            x[k - 1, j, i] = x[k - 1, j, i] + vf * (noise * (prm.lz - z * 0.9) / prm.lz)
    return u, v, w
