"""Writes the golden vectors in tests/golden/*.npz.

Runs in the build container only, where /root/reference exists: it calls the
reference's own Fortran kernels, compiled from their sources by
oracle/build_ref.sh into oracle/_ref/ (LLVM flang, -O2 -ffp-contract=off),
through ctypes, on small seeded inputs, and stores inputs and outputs.

Two kinds of fixtures:

* kernel fixtures (``kernels_*.npz``): one call of each Fortran routine on the
  path with random fields and masks (periodic, closed box and an irregular
  mask with islands);
* composed fixtures (``composed_*.npz``): the reference's Python driver logic
  (core/gmg/hierarchy.py Vcycle/Fcycle/solve/twoVcycle, core/gmg/level.py
  set-up, core/operators.py invert_vorticity, core/timescheme.py RK3_SSP,
  core/euler.py / core/boussinesq.py) restated here in NumPy around the
  *reference's* Fortran kernels, run for a few steps on 16x16 / 32x16 grids.
  This second restatement of the orchestration is independent of the C oracle
  (oracle/f2d_oracle.c), which the tests compare against these fixtures.

Usage:  python tests/golden/make_golden.py   (after ``make -C oracle ref``)
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REFDIR = os.path.join(REPO, "oracle", "_ref")

_libs = {}


def L(name):
    if name not in _libs:
        _libs[name] = C.CDLL(os.path.join(REFDIR, "lib%s.so" % name))
    return _libs[name]


def F(a, dt=np.float64):
    """Fortran-ordered copy: x(j,i) of the Fortran routine == a[j-1, i-1]."""
    return np.asfortranarray(np.asarray(a, dtype=dt))


def ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def fcall(lib, name, *args):
    """Call a Fortran routine: arrays by address, scalars by reference."""
    cargs, keep = [], []
    for a in args:
        if isinstance(a, np.ndarray):
            cargs.append(ptr(a))
        elif isinstance(a, (int, np.integer)):
            c = C.c_int(int(a))
            keep.append(c)
            cargs.append(C.byref(c))
        elif isinstance(a, (float, np.floating)):
            c = C.c_double(float(a))
            keep.append(c)
            cargs.append(C.byref(c))
        elif isinstance(a, C._SimpleCData):
            keep.append(a)
            cargs.append(C.byref(a))
        else:
            raise TypeError(type(a))
    f = getattr(L(lib), name)
    f.restype = None
    f(*cargs)


# ---------------------------------------------------------------------------
# thin wrappers over the reference Fortran (same argument meaning as the
# f2py signatures the reference Python uses)
# ---------------------------------------------------------------------------
def r_adv_upwind(msk, x, y, u, v, cst, nh, method, order, flux=False):
    m, n = x.shape
    Fm, Fx, Fy, Fu, Fv = F(msk, np.int8), F(x), F(y), F(u), F(v)
    cst = np.asarray(cst, dtype=np.float64).copy()
    if flux:
        xf, yf = F(np.zeros_like(x)), F(np.zeros_like(x))
        fcall("fortran_fluxes", "adv_upwind_", Fm, Fx, Fy, Fu, Fv, xf, yf, cst, nh, method, order, m, n)
        return np.ascontiguousarray(Fy), np.ascontiguousarray(xf), np.ascontiguousarray(yf)
    fcall("fortran_advection", "adv_upwind_", Fm, Fx, Fy, Fu, Fv, cst, nh, method, order, m, n)
    return np.ascontiguousarray(Fy)


def r_adv_centered(msk, x, y, u, v, cst, nh, order, flux=False):
    m, n = x.shape
    Fm, Fx, Fy, Fu, Fv = F(msk, np.int8), F(x), F(y), F(u), F(v)
    cst = np.asarray(cst, dtype=np.float64).copy()
    if flux:
        xf, yf = F(np.zeros_like(x)), F(np.zeros_like(x))
        fcall("fortran_fluxes", "adv_centered_", Fm, Fx, Fy, Fu, Fv, xf, yf, cst, nh, 0, order, m, n)
        return np.ascontiguousarray(Fy), np.ascontiguousarray(xf), np.ascontiguousarray(yf)
    fcall("fortran_advection", "adv_centered_", Fm, Fx, Fy, Fu, Fv, cst, nh, 0, order, m, n)
    return np.ascontiguousarray(Fy)


def r_celltocorner(xr, xp):
    m, n = xr.shape
    Fp = F(xp)
    fcall("fortran_operators", "celltocorner_", F(xr), Fp, n, m)
    xp[...] = Fp


def r_orthogradient(msk, psi, dx, dy, nh, u, v):
    m, n = psi.shape
    Fu, Fv = F(u), F(v)
    fcall("fortran_operators", "computeorthogradient_", F(msk, np.int8), F(psi), float(dx), float(dy),
          nh, n, m, Fu, Fv)
    u[...] = Fu
    v[...] = Fv


def r_add_diffusion(msk, trac, dx, nh, K, dtrac):
    m, n = trac.shape
    Fd = F(dtrac)
    fcall("fortran_operators", "add_diffusion_", F(msk, np.int8), F(trac), float(dx), nh, float(K), Fd, n, m)
    dtrac[...] = Fd


def r_add_torque(msk, buoy, dx, nh, g, domega):
    m, n = buoy.shape
    Fd = F(domega)
    fcall("fortran_operators", "add_torque_", F(msk, np.int8), F(buoy), float(dx), nh, float(g), Fd, n, m)
    domega[...] = Fd


def r_noslip(msk, x, y, dx, dy, nh):
    """computenoslipsourceterm(msk, x, y, total, dx, dy, nh, m, n)."""
    m, n = x.shape
    Fy = F(y)
    tot = C.c_double(0.0)
    fcall("fortran_operators", "computenoslipsourceterm_", F(msk, np.int8), F(x), Fy, tot,
          float(dx), float(dy), nh, m, n)
    y[...] = Fy
    return tot.value


def r_fillhalo(x, nh):
    m, n = x.shape
    Fx = F(x)
    fcall("fortran_multigrid", "fillhalo_", Fx, nh, m, n)
    x[...] = Fx


def r_smoothtwice(msk, A5, x, b, omega):
    m, n = x.shape
    Fx = F(x)
    FA = F(np.transpose(A5, (1, 2, 0)))
    yo = F(np.zeros((3, n)))
    fcall("fortran_multigrid", "smoothtwicewitha_", F(msk, np.int8), FA, Fx, F(b), float(omega), m, n, yo)
    x[...] = Fx


def r_residual(msk, A5, x, b, r):
    m, n = x.shape
    Fr = F(r)
    FA = F(np.transpose(A5, (1, 2, 0)))
    fcall("fortran_multigrid", "computeresidualwitha_", F(msk, np.int8), FA, F(x), F(b), Fr, m, n)
    r[...] = Fr


def r_restrict(msk2, x1, nh, x2):
    m1, n1 = x1.shape
    m2, n2 = x2.shape
    F2 = F(x2)
    fcall("fortran_multigrid", "restrict_", F(msk2, np.int8), F(x1), nh, n2, m2, n1, m1, F2)
    x2[...] = F2


def r_interpolate(msk1, msk2, x2, nh, x1):
    m1, n1 = x1.shape
    m2, n2 = x2.shape
    F1 = F(x1)
    fcall("fortran_multigrid", "interpolate_", F(msk1, np.int8), F(msk2, np.int8), F(x2), nh,
          n2, m2, n1, m1, F1)
    x1[...] = F1


def r_norm(msk, x, nh):
    m, n = x.shape
    y = C.c_double(0.0)
    fcall("fortran_multigrid", "computenorm_", F(msk, np.int8), F(x), nh, n, m, y)
    return y.value


def r_coarsenmatrix(Af9, msk1, msk2, nh):
    m1, n1 = msk1.shape
    m2, n2 = msk2.shape
    FAf = F(np.transpose(Af9, (1, 2, 0)))
    FAc = F(np.zeros((m2, n2, 9)))
    fcall("fortran_multigrid", "coarsenmatrix_", FAf, FAc, F(msk1, np.int8), F(msk2, np.int8), nh,
          n1, m1, n2, m2)
    return np.ascontiguousarray(np.transpose(np.asarray(FAc), (2, 0, 1)))


def r_diag(msk, u, v, x, y, nh):
    m, n = u.shape
    Fm = F(msk, np.int8)
    ke, mx = C.c_double(), C.c_double()
    fcall("fortran_diag", "computekemaxu_", Fm, F(u), F(v), nh, n, m, ke, mx)
    s, s2 = C.c_double(), C.c_double()
    fcall("fortran_diag", "computesumandnorm_", Fm, F(x), nh, n, m, s, s2)
    z = C.c_double()
    fcall("fortran_diag", "computedotprod_", Fm, F(x), F(y), nh, n, m, z)
    sm = C.c_double()
    fcall("fortran_diag", "computesum_", Fm, F(x), nh, n, m, sm)
    return np.array([ke.value, mx.value, s.value, s2.value, z.value, sm.value])


# ---------------------------------------------------------------------------
# masks and fields
# ---------------------------------------------------------------------------
def masks(nx, ny, nh, rng):
    m, n = ny + 2 * nh, nx + 2 * nh
    perio = np.ones((m, n), np.int8)
    closed = perio.copy()
    closed[:nh, :] = 0
    closed[-nh:, :] = 0
    closed[:, :nh] = 0
    closed[:, -nh:] = 0
    islands = closed.copy()
    # a few solid blocks and isolated solid cells inside the box
    islands[nh + 4:nh + 7, nh + 3:nh + 5] = 0
    islands[nh + 9, nh + 9] = 0
    islands[nh + 2:nh + 4, n - nh - 5:n - nh - 2] = 0
    pick = rng.random((m, n)) < 0.04
    islands[pick] = 0
    return {"perio": perio, "closed": closed, "islands": islands}


def kernel_fixtures(nx, ny, seed):
    nh = 3
    rng = np.random.default_rng(seed)
    m, n = ny + 2 * nh, nx + 2 * nh
    out = {}
    for mname, msk in masks(nx, ny, nh, rng).items():
        x = rng.standard_normal((m, n))
        u = rng.standard_normal((m, n))
        v = rng.standard_normal((m, n))
        u[rng.random((m, n)) < 0.2] *= 0.01  # exercise |u| < u1 branch
        v[rng.random((m, n)) < 0.2] *= 0.01
        y0 = rng.standard_normal((m, n))
        dx, dy = 1.0 / nx, 1.3 / ny
        umax = float(np.abs(u).max())
        cst = np.array([dx, dy, 0.05, umax, 0.05])
        p = "%s_" % mname
        out[p + "msk"] = msk
        out[p + "x"], out[p + "u"], out[p + "v"], out[p + "y0"] = x, u, v, y0
        out[p + "cst"] = cst
        for order in (1, 3, 5):
            for method in (0, 1):
                out[p + "up_o%d_m%d" % (order, method)] = r_adv_upwind(msk, x, y0, u, v, cst, nh, method, order)
            yy, xf, yf = r_adv_upwind(msk, x, y0, u, v, cst, nh, 1, order, flux=True)
            out[p + "upflx_o%d" % order] = yy
            out[p + "upflx_o%d_xf" % order] = xf
            out[p + "upflx_o%d_yf" % order] = yf
        for order in (2, 4, 6):
            out[p + "ce_o%d" % order] = r_adv_centered(msk, x, y0, u, v, cst, nh, order)
            yy, xf, yf = r_adv_centered(msk, x, y0, u, v, cst, nh, order, flux=True)
            out[p + "ceflx_o%d" % order] = yy
            out[p + "ceflx_o%d_xf" % order] = xf
            out[p + "ceflx_o%d_yf" % order] = yf
        xp = y0.copy()
        r_celltocorner(x, xp)
        out[p + "corner"] = xp
        uu, vv = y0.copy(), (y0 * 0.5).copy()
        r_orthogradient(msk, x, dx, dy, nh, uu, vv)
        out[p + "ortho_u"], out[p + "ortho_v"] = uu, vv
        dd = y0.copy()
        r_add_diffusion(msk, x, dx, nh, 1.7e-3, dd)
        out[p + "diffusion"] = dd
        dw = y0.copy()
        r_add_torque(msk, x, dx, nh, 9.81, dw)
        out[p + "torque"] = dw
        hh = x.copy()
        r_fillhalo(hh, nh)
        out[p + "fillhalo"] = hh
        out[p + "diag"] = r_diag(msk, u, v, x, y0, nh)
        # multigrid kernels with a random symmetric-looking operator
        A5 = rng.standard_normal((5, m, n)) * 0.1
        A5[4] = -(3.0 + rng.random((m, n)))
        b = rng.standard_normal((m, n))
        out[p + "A5"], out[p + "b"] = A5, b
        xs = x.copy()
        r_smoothtwice(msk, A5, xs, b, 8.0 / 9.0)
        out[p + "smooth"] = xs
        rr = y0.copy()
        r_residual(msk, A5, x, b, rr)
        out[p + "residual"] = rr
        out[p + "norm"] = np.array(r_norm(msk, x, nh))
        # transfers between this level and one twice coarser
        m2, n2 = ny // 2 + 2 * nh, nx // 2 + 2 * nh
        msk2 = (rng.random((m2, n2)) < 0.85).astype(np.int8)
        xc = rng.standard_normal((m2, n2))
        out[p + "msk2"], out[p + "xc"] = msk2, xc
        x2 = rng.standard_normal((m2, n2))
        r_restrict(msk2, x, nh, x2)
        out[p + "restrict"] = x2
        x1 = y0.copy()
        r_interpolate(msk, msk2, xc, nh, x1)
        out[p + "interpolate"] = x1
        A9 = rng.standard_normal((9, m, n))
        out[p + "A9"] = A9
        mskc = (rng.random((m2, n2)) < 0.9).astype(np.int8)
        out[p + "mskc"] = mskc
        out[p + "coarsen"] = r_coarsenmatrix(A9, msk, mskc, nh)
    return out


def noslip_fixtures(nx, ny, seed):
    """computenoslipsourceterm on each mask (y pre-filled: the routine leaves
    the last nh-1 rows and columns alone)."""
    nh = 3
    rng = np.random.default_rng(seed)
    m, n = ny + 2 * nh, nx + 2 * nh
    out = {}
    for mname, msk in masks(nx, ny, nh, rng).items():
        x = rng.standard_normal((m, n))
        y0 = rng.standard_normal((m, n))
        dx, dy = 1.0 / nx, 1.3 / ny
        y = y0.copy()
        tot = r_noslip(msk, x, y, dx, dy, nh)
        p = "%s_" % mname
        out[p + "msk"], out[p + "x"], out[p + "y0"] = msk, x, y0
        out[p + "dxdy"] = np.array([dx, dy])
        out[p + "noslip"] = y
        out[p + "total"] = np.array(tot)
    return out


# ---------------------------------------------------------------------------
# composed reference: NumPy restatement of the Python driver around the
# reference Fortran kernels
# ---------------------------------------------------------------------------
class RefGmg:
    """core/gmg/hierarchy.py + core/gmg/level.py with one subdomain."""

    def __init__(self, nx, ny, dx, dy, mskp, nh=3, omega=8.0 / 9.0):
        self.nh, self.omega = nh, omega
        self.levels = []
        n, m, lev = nx, ny, 0
        while True:  # level.py:62-115 with np=mp=1
            if lev > 0:
                n, m = n // 2, m // 2
            self.levels.append((m + 2 * nh, n + 2 * nh))
            lev += 1
            if n <= 4 or m <= 4:
                break
        self.nlevs = len(self.levels)
        self.msk, self.A = [], []
        for lev, (mv, nv) in enumerate(self.levels):
            if lev == 0:
                msk = mskp.astype(np.int8)
                A = self._afinest(msk, dx, dy)
            else:
                mskf = self.msk[lev - 1] * 1.0
                coef = np.ones((mv, nv))
                r_restrict(np.ones((mv, nv), np.int8), mskf, nh, coef)
                r_fillhalo(coef, nh)
                msk = np.ones((mv, nv), np.int8)
                msk[coef <= 0.5] = 0
                A = r_coarsenmatrix(self.A[lev - 1], self.msk[lev - 1], msk, nh)
                for k in range(9):
                    r_fillhalo(A[k], nh)
            self.msk.append(msk)
            self.A.append(A)
        self.A = [np.ascontiguousarray(a[:5]) for a in self.A]
        self.x = [np.zeros(s) for s in self.levels]
        self.b = [np.zeros(s) for s in self.levels]
        self.r = [np.zeros(s) for s in self.levels]

    def _afinest(self, msk, dx, dy):  # level.py:261-302
        mv, nv = msk.shape
        if dx == dy:
            a, b, c = -6.0 / 2, 1.0 / 2, 0.5 / 2
            st = np.array([[c, b, c], [b, a, b], [c, b, c]])
        else:
            bx, by = dy / dx, dx / dy
            st = np.array([[0, by, 0], [bx, -2 * (bx + by), bx], [0, by, 0]], dtype=float)
        A = np.zeros((9, mv, nv))
        I, J = np.meshgrid(np.arange(nv - 2) + 1, np.arange(mv - 2) + 1)
        coef = 1.0 / (dx * dy)
        for k in range(9):
            Ak = np.zeros((mv, nv))
            i, j = (k % 3) - 1, (k // 3) - 1
            Ak[J, I] = st[j + 1, i + 1] * coef * msk[J + j, I + i] * msk[J, I]
            r_fillhalo(Ak, self.nh)
            A[k] = Ak
        return A

    def smooth(self, lev, x, b, nite):
        for _ in range(nite):
            r_smoothtwice(self.msk[lev], self.A[lev], x, b, self.omega)
            r_fillhalo(x, self.nh)

    def residual(self, lev, x, b, r):
        r_residual(self.msk[lev], self.A[lev], x, b, r)
        r_fillhalo(r, self.nh)

    def f2c(self, lev, xf, xc):
        r_restrict(self.msk[lev + 1], xf, self.nh, xc)
        r_fillhalo(xc, self.nh)

    def c2f(self, lev, xc, xf):
        r_interpolate(self.msk[lev], self.msk[lev + 1], xc, self.nh, xf)

    def norm(self, lev, x):
        return np.sqrt(r_norm(self.msk[lev], x, self.nh) / 1.0)

    def Vcycle(self, lev1):
        deep = self.nlevs - 1
        for lev in range(lev1, deep):
            if lev > lev1:
                self.x[lev][:, :] = 0.0
            self.smooth(lev, self.x[lev], self.b[lev], 1)
            self.residual(lev, self.x[lev], self.b[lev], self.r[lev])
            self.f2c(lev, self.r[lev], self.b[lev + 1])
        self.x[deep][:, :] = 0.0
        self.smooth(deep, self.x[deep], self.b[deep], 16)
        for lev in range(deep - 1, lev1 - 1, -1):
            self.c2f(lev, self.x[lev + 1], self.r[lev])
            self.x[lev] += self.r[lev]
            self.smooth(lev, self.x[lev], self.b[lev], 1)

    def Fcycle(self, lev1):
        deep = self.nlevs - 1
        for lev in range(lev1, deep):
            self.f2c(lev, self.b[lev], self.b[lev + 1])
        self.x[deep][:, :] = 0.0
        self.smooth(deep, self.x[deep], self.b[deep], 16)
        for lev in range(deep - 1, lev1 - 1, -1):
            self.c2f(lev, self.x[lev + 1], self.x[lev])
            self.Vcycle(lev)

    def solve(self, x, b, maxite=4, tol=1e-11):
        self.residual(0, x, b, self.b[0])
        normeb = self.norm(0, b)
        if not normeb > 0:
            return 0, 0.0
        res0 = self.norm(0, self.b[0]) / normeb
        res, nite = res0, 0
        while nite < maxite and res0 > tol:
            self.Fcycle(0)
            x += self.x[0]
            self.residual(0, x, b, self.b[0])
            res = self.norm(0, self.b[0]) / normeb
            res0 = res
            nite += 1
        return nite, res

    def twoVcycle(self, x, b):
        self.x[0][:, :] = x
        self.b[0][:, :] = b
        for _ in range(2):
            self.residual(0, self.x[0], self.b[0], self.r[0])
            self.Vcycle(0)
        x[:, :] = self.x[0]


class RefModel:
    """Euler / Boussinesq RK3_SSP step: core/euler.py, core/boussinesq.py,
    core/operators.py, core/timescheme.py:167-180."""

    def __init__(self, kind, nx, ny, Lx, Ly, msk, Kdiff=0.0, diffusion=False, gravity=1.0,
                 scheme="RK3_SSP"):
        self.kind, self.nh = kind, 3
        # core/timescheme.py:30-33, 61-75: the stage of the single-stage
        # terms and their coefficient
        self.scheme = scheme
        self.kforcing = 2 if scheme == "RK3_SSP" else 0
        self.dtcoef = 1.5 if scheme == "RK3_SSP" else 1.0
        self.dx, self.dy = Lx / nx, Ly / ny
        self.msk = msk
        self.nvar = 5 if kind == "euler" else 6
        m, n = msk.shape
        self.work = np.zeros((m, n))
        r_celltocorner(msk * 1.0, self.work)
        self.work[self.work < 1.0] = 0.0
        self.mskp = msk * 0
        self.mskp[self.work == 1.0] = 1
        self.gmg = RefGmg(nx, ny, self.dx, self.dy, self.work)
        self.cst = np.array([self.dx, self.dy, 0.05, 0.0, 0.05])
        self.state = np.zeros((self.nvar, m, n))
        self.dx0, self.dx1, self.dx2 = (np.zeros_like(self.state) for _ in range(3))
        self.Kdiff, self.diffusion, self.gravity = Kdiff, diffusion, gravity
        self.area = np.sum(msk[3:-3, 3:-3]) * 1.0
        self.perio = bool(np.all(msk == 1))
        self.nite = []

    def rhs_adv(self, x, dxdt):
        tr = [0] if self.kind == "euler" else [0, 4]
        for k in tr:
            y = r_adv_upwind(self.msk, x[k], dxdt[k], x[2], x[3], self.cst, 3, 1, 5)
            r_fillhalo(y, 3)
            dxdt[k][:, :] = y

    def invert(self, x, full):
        psi = x[1].copy()
        r_celltocorner(x[0], self.work)
        if full:
            nite, res = self.gmg.solve(psi, self.work)
            self.nite.append(nite)
            if self.perio:
                psi -= np.sum(psi[3:-3, 3:-3]) * 1.0 / self.area
        else:
            self.gmg.twoVcycle(psi, self.work)
        psi = psi * self.mskp
        u, v = x[2].copy(), x[3].copy()
        r_orthogradient(self.msk, psi, self.dx, self.dy, 3, u, v)
        x[2], x[3], x[1] = u, v, psi

    def dynamics(self, x, dxdt, kstage):
        self.rhs_adv(x, dxdt)
        if self.kind == "euler":
            if kstage != self.kforcing:
                self.invert(dxdt, False)
            return
        y = dxdt[0].copy()
        y *= self.msk
        r_add_torque(self.msk, x[4], self.dx, 3, self.gravity, y)
        r_fillhalo(y, 3)
        dxdt[0][:, :] = y
        if kstage == self.kforcing and self.diffusion:
            for k in (0, 4):
                y = dxdt[k].copy()
                r_add_diffusion(self.msk, x[k], self.dx, 3, self.dtcoef * self.Kdiff, y)
                r_fillhalo(y, 3)
                dxdt[k] = y
        self.invert(dxdt, False)

    def step(self, dt, maxspeed):
        self.cst[3] = maxspeed
        x = self.state
        if self.scheme == "RK3_SSP":  # core/timescheme.py:167-180
            self.dynamics(x, self.dx0, 0)
            xs = x + dt * self.dx0
            self.dynamics(xs, self.dx1, 1)
            xs = x + (0.25 * dt) * (self.dx0 + self.dx1)
            self.dynamics(xs, self.dx2, 2)
            x += (dt / 6.0) * (self.dx0 + self.dx1 + 4 * self.dx2)
        else:  # core/timescheme.py:152-164
            self.dynamics(x, self.dx0, 0)
            xs = x + (dt / 3.0) * self.dx0
            self.dynamics(xs, self.dx1, 1)
            xs = x + (0.5 * dt) * self.dx1
            self.dynamics(xs, self.dx2, 2)
            x += dt * self.dx2
        if self.kind == "euler":
            self.invert(x, True)
        else:
            x[5][:, :] = x[4] - 0.0

    def maxspeed(self):
        d = r_diag(self.msk, self.state[2], self.state[3], self.state[0], self.state[0], 3)
        return d[1]


def composed_fixture(kind, nx, ny, Lx, Ly, geometry, seed, nsteps, scheme="RK3_SSP"):
    rng = np.random.default_rng(seed)
    nh = 3
    m, n = ny + 2 * nh, nx + 2 * nh
    msk = np.ones((m, n), np.int8)
    if geometry in ("closed", "xchannel"):
        msk[:nh, :] = 0
        msk[-nh:, :] = 0
    if geometry == "closed":
        msk[:, :nh] = 0
        msk[:, -nh:] = 0
    M = RefModel(kind, nx, ny, Lx, Ly, msk, Kdiff=1e-3, diffusion=(kind == "boussinesq"),
                 scheme=scheme)
    w = rng.standard_normal((m, n)) * msk
    if geometry == "perio":
        w -= np.mean(w[nh:-nh, nh:-nh])
    r_fillhalo(w, nh)
    M.state[0] = w
    if kind == "boussinesq":
        b = (rng.random((m, n)) - 0.5) * msk
        r_fillhalo(b, nh)
        M.state[4] = b
    out = {"msk": msk, "w0": w.copy(), "state0": M.state.copy()}
    # initial full inversion (set_psi_from_vorticity)
    M.invert(M.state, True)
    out["psi_init"] = M.state[1].copy()
    out["u_init"], out["v_init"] = M.state[2].copy(), M.state[3].copy()
    ms = M.maxspeed()
    dt = 0.8 * min(M.dx, M.dy) / ms
    out["dt"] = np.array(dt)
    # fast inversion of a tendency (twoVcycle from a non-zero guess)
    t = np.zeros_like(M.state)
    t[0] = rng.standard_normal((m, n)) * msk
    r_fillhalo(t[0], nh)
    t[1] = rng.standard_normal((m, n)) * 1e-3
    out["tend_in"] = t.copy()
    M.invert(t, False)
    out["tend_out"] = t.copy()
    for k in range(nsteps):
        M.step(dt, ms)
        ms = M.maxspeed()
        out["state_%d" % (k + 1)] = M.state.copy()
    out["nite"] = np.array(M.nite)
    out["levels"] = np.array(M.gmg.levels)
    for lev in range(M.gmg.nlevs):
        out["gmg_msk_%d" % lev] = M.gmg.msk[lev]
        out["gmg_A_%d" % lev] = M.gmg.A[lev]
    return out


def main():
    if not os.path.isdir(REFDIR):
        sys.exit("oracle/_ref missing: run `make -C oracle ref` first")
    if sys.argv[1:] == ["rk3"]:
        np.savez_compressed(os.path.join(HERE, "composed_euler_closed_rk3_32x16.npz"),
                            **composed_fixture("euler", 32, 16, 2.0, 1.0, "closed", 24, 3, "RK3"))
        np.savez_compressed(os.path.join(HERE, "composed_bous_xchannel_rk3_32x16.npz"),
                            **composed_fixture("boussinesq", 32, 16, 2.0, 1.0, "xchannel", 25, 3,
                                               "RK3"))
        return
    np.savez_compressed(os.path.join(HERE, "noslip_24x16.npz"), **noslip_fixtures(24, 16, 13))
    if sys.argv[1:] == ["noslip"]:
        return
    np.savez_compressed(os.path.join(HERE, "kernels_16x16.npz"), **kernel_fixtures(16, 16, 11))
    np.savez_compressed(os.path.join(HERE, "kernels_24x16.npz"), **kernel_fixtures(24, 16, 12))
    np.savez_compressed(os.path.join(HERE, "composed_euler_perio_16.npz"),
                        **composed_fixture("euler", 16, 16, 1.0, 1.0, "perio", 21, 3))
    np.savez_compressed(os.path.join(HERE, "composed_euler_closed_32x16.npz"),
                        **composed_fixture("euler", 32, 16, 2.0, 1.0, "closed", 22, 3))
    np.savez_compressed(os.path.join(HERE, "composed_bous_xchannel_32x16.npz"),
                        **composed_fixture("boussinesq", 32, 16, 2.0, 1.0, "xchannel", 23, 3))
    np.savez_compressed(os.path.join(HERE, "composed_euler_closed_rk3_32x16.npz"),
                        **composed_fixture("euler", 32, 16, 2.0, 1.0, "closed", 24, 3, "RK3"))
    np.savez_compressed(os.path.join(HERE, "composed_bous_xchannel_rk3_32x16.npz"),
                        **composed_fixture("boussinesq", 32, 16, 2.0, 1.0, "xchannel", 25, 3, "RK3"))
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
