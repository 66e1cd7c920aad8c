"""Initial conditions for the hydro-advance workloads.

These mirror CAMR's problem set-up (``CAMR_initdata``) and generate the
synthetic seeded fields used by the benchmarks.  They produce the conserved
state U in Array4 order: a float64 array of shape (NVAR, nz, ny, nx).

Component map (Source/Utils/IndexDefines.H:47-92, 3-D, NUM_SPECIES=3):
URHO 0, UMX 1, UMY 2, UMZ 3, UEDEN 4, UEINT 5, UTEMP 6, UFS 7..9.
"""
from __future__ import annotations

import numpy as np

GAMMA = 1.4  # Source/EOS/GammaLaw/EOS.H:16
NVAR = 10
URHO, UMX, UMY, UMZ, UEDEN, UEINT, UTEMP, UFS = 0, 1, 2, 3, 4, 5, 6, 7

SEED_FIELD = 20261202
SEED_TARGET = 20261203
SEED_DIRECTIONS = 20261204


def _blank(lo, hi):
    n = [hi[d] - lo[d] + 1 for d in range(3)]
    return np.zeros((NVAR, n[2], n[1], n[0]))


def sod(lo, hi, dx, prob_lo=(0.0, 0.0, 0.0), ptype=0, p_l=1.0, p_r=0.1,
        rho_l=1.0, rho_r=0.125, u_l=0.0, u_r=0.0, interface=0.5):
    """Sod shock tube, Exec/Sod/prob.H:15-107 (``prob.type`` 0/1/2 picks the
    axis).  Exec/SodPlusSphere/prob.H:15-49 is ``ptype=0`` with
    ``interface`` moved."""
    U = _blank(lo, hi)
    d = ptype
    idx = np.arange(lo[d], hi[d] + 1)
    x = prob_lo[d] + (idx + 0.5) * dx[d]
    left = x < interface
    Pt = np.where(left, p_l, p_r)
    rhot = np.where(left, rho_l, rho_r)
    ut = np.where(left, u_l, u_r)
    shape = [1, 1, 1]
    shape[2 - d] = len(idx)
    Pt, rhot, ut = (a.reshape(shape) for a in (Pt, rhot, ut))
    full = U.shape[1:]
    rho = np.broadcast_to(rhot, full)
    U[URHO] = rho
    mom = np.broadcast_to(rhot * ut, full)
    # prob.type 2 writes UMY instead of UMZ (Exec/Sod/prob.H:94); harmless
    # since the reference decks use u = 0, and kept for fidelity.
    U[UMX if d == 0 else UMY] = mom
    et = np.broadcast_to(Pt / (GAMMA - 1.0), full)
    U[UEINT] = et
    U[UEDEN] = et + 0.5 * rho * np.broadcast_to(ut, full) * np.broadcast_to(ut, full)
    U[UTEMP] = 0.0
    U[UFS] = 1.0 * U[URHO]
    U[UFS + 1] = 0.0 * U[URHO]
    U[UFS + 2] = 0.0 * U[URHO]
    return U


class SmoothField:
    """Seeded smooth periodic field on [0,1]^3.

    rho = 1 + 0.1 * sum_{m=1..8} sin(2 pi k_m . x + phi_m), k_m components
    uniform in {-4..4} (k_m != 0), phi_m ~ U[0, 2 pi).
    u   = Gaussian random field, |u_hat(k)| ~ |k|^-2 for 1 <= |k| <= kmax,
          scaled so that sqrt(<|u|^2 / c^2>) = mach with c^2 = gamma p / rho,
          the scale fixed on a 64^3 reference grid so every resolution sees
          the same continuous field.
    p   = 1.
    """

    def __init__(self, seed=SEED_FIELD, kmax=16, mach=0.5, p0=1.0):
        rng = np.random.Generator(np.random.PCG64(seed))
        ks = []
        while len(ks) < 8:
            k = rng.integers(-4, 5, size=3)
            if np.any(k != 0):
                ks.append(k)
        self.k_rho = np.array(ks, dtype=np.int64)
        self.phi_rho = rng.uniform(0.0, 2.0 * np.pi, size=8)
        K = kmax
        r = np.arange(-K, K + 1)
        kx, ky, kz = np.meshgrid(r, r, r, indexing="ij")
        kk = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
        sel = (kk >= 1.0) & (kk <= K)
        self.kvec = np.stack([kx[sel], ky[sel], kz[sel]], axis=1)
        amp = kk[sel] ** -2.0
        self.coef = []
        for _ in range(3):
            c = (rng.standard_normal(len(amp)) + 1j * rng.standard_normal(len(amp))) * amp
            self.coef.append(c)
        self.p0 = p0
        self.mach = mach
        self.kmax = K
        self.scale = 1.0
        rho, u = self._eval(64, (0, 0, 0), (63, 63, 63))
        c2 = GAMMA * p0 / rho
        m2 = np.mean((u[0] ** 2 + u[1] ** 2 + u[2] ** 2) / c2)
        self.scale = mach / np.sqrt(m2)

    def _eval(self, n, lo, hi):
        """Evaluate rho and velocity at the cell centres (i+0.5)/n of the
        index box [lo, hi] of an n^3 periodic grid."""
        idx = [np.arange(lo[d], hi[d] + 1) for d in range(3)]
        xs = [(i + 0.5) / n for i in idx]
        X = xs[0][None, None, :]
        Y = xs[1][None, :, None]
        Z = xs[2][:, None, None]
        rho = np.ones((len(idx[2]), len(idx[1]), len(idx[0])))
        for (a, b, c), ph in zip(self.k_rho, self.phi_rho):
            rho = rho + 0.1 * np.sin(2.0 * np.pi * (a * X + b * Y + c * Z) + ph)
        # band-limited GRF evaluated with an m^3 inverse FFT, m = r*n the
        # smallest multiple of n with kmax < m/2 (exact), phase-shifted so
        # that grid point r*i sits at the cell centre (i+0.5)/n, then
        # sampled on the box
        r = 1
        while self.kmax >= (r * n) // 2:
            r += 1
        m = r * n
        u = []
        ph = np.exp(1j * np.pi * (self.kvec.sum(axis=1)) / n)
        for c in self.coef:
            F = np.zeros((m, m, m), dtype=np.complex128)
            F[self.kvec[:, 2] % m, self.kvec[:, 1] % m, self.kvec[:, 0] % m] = c * ph
            f = np.fft.ifftn(F).real * (m ** 3)
            f = f[np.ix_(r * (idx[2] % n), r * (idx[1] % n), r * (idx[0] % n))]
            u.append(self.scale * f)
        return rho, u

    def state(self, n, lo=None, hi=None):
        """Conserved state on the index box [lo, hi] of an n^3 grid."""
        lo = (0, 0, 0) if lo is None else lo
        hi = (n - 1, n - 1, n - 1) if hi is None else hi
        rho, u = self._eval(n, lo, hi)
        U = _blank(lo, hi)
        U[URHO] = rho
        U[UMX] = rho * u[0]
        U[UMY] = rho * u[1]
        U[UMZ] = rho * u[2]
        et = self.p0 / (GAMMA - 1.0)
        U[UEINT] = et
        U[UEDEN] = et + 0.5 * rho * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2])
        U[UTEMP] = 0.0
        U[UFS] = 1.0 * rho
        U[UFS + 1] = 0.0 * rho
        U[UFS + 2] = 0.0 * rho
        return U
