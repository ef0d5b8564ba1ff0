"""Grid: sizes, cell-centred coordinates and the geometry mask.

Restates core/grid.py:12-140 for one subdomain (npx = npy = 1): the mask is
int8, 0 in solid cells (walls are the halo rows/columns of the closed
directions), and ``domain_integration`` sums the interior in NumPy's order.
"""
import numpy as np


def _celltocorner(xr, xp):
    """core/fortran_operators.f90:44-64 on host arrays (set-up only): the
    mean of the four cells around each upper-right corner; the last row and
    column are left as they are."""
    xp[:-1, :-1] = 0.25 * (((xr[:-1, :-1] + xr[:-1, 1:]) + xr[1:, :-1]) + xr[1:, 1:])


class Island:
    """core/island.py: islands are solid cell sets (``idx``, a NumPy index
    as from ``np.where``) whose walls carry a fixed streamfunction ``psi0``.
    ``finalize`` builds the inversion's rhs correction ``rhsp`` and the
    boundary values ``psi`` (core/island.py:21-43)."""

    def __init__(self, grid):
        self.nxl, self.nyl, self.dx, self.dy = grid.nxl, grid.nyl, grid.dx, grid.dy
        self.rhsp = np.zeros((self.nyl, self.nxl))
        self.psi = np.zeros((self.nyl, self.nxl))
        self.nbisland = 0
        self.data = []

    def add(self, idx, psi0):
        self.data.append({"idx": idx, "psi0": psi0})
        self.nbisland += 1

    def finalize(self, mskp_model=None):
        mskp = np.zeros((self.nyl, self.nxl), dtype=np.int8)
        work = np.zeros((self.nyl, self.nxl))
        mskr = np.zeros((self.nyl, self.nxl))
        for d in self.data:
            idx, psi0 = d["idx"], d["psi0"]
            mskr[:, :] = 1.0
            mskp[:, :] = 0
            mskr[idx] = 0.0
            _celltocorner(mskr, work)
            mskp[work == 1] = 1
            mskp = 1 - mskp  # the island's corners
            vort = (np.roll(mskp, -1, axis=1) + np.roll(mskp, -1, axis=0)
                    + np.roll(mskp, +1, axis=1) + np.roll(mskp, +1, axis=0))
            z = vort * psi0 / (self.dx * self.dy)
            self.rhsp[vort > 0] = z[vort > 0]
            self.psi[mskp == 1] = psi0


class Grid:
    def __init__(self, param):
        param.checkall()
        self.nx, self.ny = int(param.nx), int(param.ny)
        self.Lx, self.Ly = float(param.Lx), float(param.Ly)
        self.nh = int(param.nh)
        self.geometry = param.geometry
        self.hydroepsilon = param.hydroepsilon
        self.nxl = self.nx + 2 * self.nh
        self.nyl = self.ny + 2 * self.nh
        self.dx = self.Lx / self.nx
        self.dy = self.Ly / self.ny
        nh = self.nh
        self.x1d = (np.arange(self.nxl) + 0.5 - nh) * self.dx
        self.y1d = (np.arange(self.nyl) + 0.5 - nh) * self.dy
        xr, yr = np.meshgrid(self.x1d, self.y1d)
        self.xr, self.yr = np.ascontiguousarray(xr), np.ascontiguousarray(yr)
        self.set_msk()
        if getattr(param, "isisland", False):
            self.island = Island(self)

    def set_msk(self):
        """core/grid.py:67-102"""
        nh = self.nh
        self.msk = np.ones((self.nyl, self.nxl), dtype=np.int8)
        if self.geometry in ("xperio", "xchannel", "closed", "disc"):
            self.msk[-nh:, :] = 0
            self.msk[:nh, :] = 0
        if self.geometry in ("yperio", "ychannel", "closed", "disc"):
            self.msk[:, -nh:] = 0
            self.msk[:, :nh] = 0
        if self.geometry == "disc":
            r = np.sqrt((self.xr / self.Lx - 0.5) ** 2 + (self.yr / self.Ly - 0.5) ** 2)
            self.msk[r >= 0.5] = 0
        self.msknoslip = self.msk.copy()
        self.finalize_msk()

    def finalize_msk(self):
        """core/grid.py:104-130 (area, centred coordinates, second moments)."""
        msk = self.msk
        self.area = self.domain_integration(msk)
        x0 = self.domain_integration(self.xr * msk) / self.area
        y0 = self.domain_integration(self.yr * msk) / self.area
        self.xr0 = (self.xr - x0) * msk
        self.yr0 = (self.yr - y0) * msk
        self.x0, self.y0 = x0, y0
        self.x2 = self.domain_integration(self.xr0 ** 2 * msk) / self.area
        self.y2 = self.domain_integration(self.yr0 ** 2 * msk) / self.area
        self.r2 = self.xr0 ** 2 + self.yr0 ** 2

    def boundary_mask(self):
        """Operators.set_boundary_msk (core/operators.py:157-185): mskbc = 1
        on the no-slip fluid cells next to a wall (periodic halo), its area
        and the second moments x2bc, y2bc."""
        msk = self.msknoslip
        z = (np.roll(msk, -1, axis=1) + np.roll(msk, -1, axis=0)
             + np.roll(msk, +1, axis=1) + np.roll(msk, +1, axis=0) - 4 * msk)
        z = z * msk
        mskbc = self.msk * 0
        mskbc[z < 0] = 1
        mskbc *= self.msknoslip
        self.fill_halo(mskbc)
        bcarea = self.domain_integration(mskbc)
        x2bc = self.domain_integration(self.xr0 ** 2 * mskbc * self.msknoslip)
        y2bc = self.domain_integration(self.yr0 ** 2 * mskbc * self.msknoslip)
        return mskbc, bcarea, x2bc, y2bc

    def domain_integration(self, z2d):
        nh = self.nh
        return float(np.sum(z2d[nh:-nh, nh:-nh]) * 1.0)

    def fill_halo(self, x):
        """Periodic halo (core/gmg/fortran_multigrid.f90:365-412) on a host array."""
        nh = self.nh
        M, N = x.shape[-2] - 2 * nh, x.shape[-1] - 2 * nh
        x[..., :nh, :] = x[..., M:M + nh, :]
        x[..., -nh:, :] = x[..., nh:2 * nh, :]
        x[..., :, :nh] = x[..., :, N:N + nh]
        x[..., :, -nh:] = x[..., :, nh:2 * nh]
