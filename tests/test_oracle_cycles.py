"""The oracle's multigrid cycles (core/gmg/hierarchy.py:98-218: Vcycle,
Fcycle, solve, twoVcycle) against the textbook cycle algebra on dense
matrices, independently of the oracle's cycle code.

Every map on one level is affine and is assembled column by column from the
oracle's kernels, which tests/test_oracle_golden.py pins bit for bit to the
reference's Fortran (core/gmg/fortran_multigrid.f90):

  smoothing    x <- Sx x + Sb b      smoothtwicewithA + fill (level.py:340-368)
  residual     r  = Rx x + Rb b      computeresidualwithA + fill (level.py:370-387)
  restriction  b' = R r              restrict + fill (level.py:473-496)
  prolongation r  = P x'             interpolate (level.py:447-470)

The V(1,1)-cycle with one pre- and one post-smoothing and 16 smoothings from
zero on the coarsest level (hierarchy.py:29-32) is then, as matrices,
    x1 = Sx x + Sb b,  x2 = x1 + P B' R (Rx x1 + Rb b),  x3 = Sx x2 + Sb b,
with B' the coarse V-cycle from zero (B_deep = sum_{k<16} Sx^k Sb); the
F-cycle restricts b down, smooths the coarsest from zero and on the way up
takes x_l = V_l(P x_{l+1}, b_l); solve iterates x <- x + F (Rx x + Rb b);
twoVcycle is two V-cycles from the given guess.  A slip in the cycle order
(a missing correction, a smoothing count, residual and restriction
swapped, interpolation added instead of assigned in the F-cycle) changes
these matrices.  The same maps also satisfy what the mathematics fixes: a
V-cycle iteration converges (spectral radius below 1) on a closed box, and
its fixed point solves the masked system Rx x + Rb b = 0.
"""
import numpy as np
import pytest

from oracle import oracle as orc

NH = 3
OMEGA = 8.0 / 9.0


def corner_mask(msk):
    m, n = msk.shape
    w = np.zeros((m, n))
    orc.celltocorner(msk * 1.0, w)
    return (w == 1.0).astype(np.int8)


def make(nx, ny, geometry, islands=False):
    g = orc.make_grid(nx, ny, 1.0 * nx / ny, 1.0, geometry)
    msk = g["msk"].copy()
    if islands:
        msk[NH + 5:NH + 8, NH + 4:NH + 6] = 0
        msk[NH + 11, NH + 12] = 0
    mskp = corner_mask(msk)
    gm = orc.OracleGmg(nx, ny, g["dx"], g["dy"], mskp)
    return gm


class Level:
    """Dense affine maps of one level, assembled from the kernels."""

    def __init__(self, gm, lev):
        self.gm, self.lev = gm, lev
        self.shape = gm.shape(lev)
        self.N = self.shape[0] * self.shape[1]
        self.msk = gm.msk(lev).copy()
        self.A = gm.A(lev).copy()

    def smooth(self, x, b):
        x = x.reshape(self.shape).copy()
        orc.smoothtwice(self.msk, self.A, x, b.reshape(self.shape), OMEGA)
        orc.fillhalo(x, NH)
        return x.ravel()

    def residual(self, x, b):
        r = np.zeros(self.shape)
        orc.residual(self.msk, self.A, x.reshape(self.shape), b.reshape(self.shape), r)
        orc.fillhalo(r, NH)
        return r.ravel()


def columns(f, nin):
    e = np.zeros(nin)
    cols = []
    for k in range(nin):
        e[k] = 1.0
        cols.append(f(e))
        e[k] = 0.0
    return np.array(cols).T


@pytest.fixture(scope="module", params=[(16, 16, "closed", False), (16, 16, "perio", False),
                                        (32, 16, "closed", True), (16, 8, "xchannel", False)],
                ids=["closed16", "perio16", "islands32x16", "channel16x8"])
def mats(request):
    nx, ny, geom, isl = request.param
    gm = make(nx, ny, geom, isl)
    L = [Level(gm, l) for l in range(gm.nlevs)]
    M = []
    for l, lv in enumerate(L):
        z = np.zeros(lv.N)
        d = {"Sx": columns(lambda e: lv.smooth(e, z), lv.N),
             "Sb": columns(lambda e: lv.smooth(z, e), lv.N),
             "Rx": columns(lambda e: lv.residual(e, z), lv.N),
             "Rb": columns(lambda e: lv.residual(z, e), lv.N)}
        if l + 1 < len(L):
            c = L[l + 1]

            def restrict(e, lv=lv, c=c):
                out = np.zeros(c.shape)
                orc.restrict(c.msk, e.reshape(lv.shape), NH, out)
                orc.fillhalo(out, NH)
                return out.ravel()

            def prolong(e, lv=lv, c=c):
                out = np.zeros(lv.shape)
                orc.interpolate(lv.msk, c.msk, e.reshape(c.shape), NH, out)
                return out.ravel()

            d["R"] = columns(restrict, lv.N)
            d["P"] = columns(prolong, c.N)
        M.append(d)
    return gm, L, M


def vcycle_maps(M, l):
    """(Vx, Vb) of a V-cycle entered at level l (core/gmg/hierarchy.py:98-126)."""
    deep = len(M) - 1
    d = M[l]
    if l + 1 == deep:
        Bc = coarsest(M[deep])
    else:
        Bc = vcycle_maps(M, l + 1)[1]  # entered from zero: V(0, b) = Vb b
    I = np.eye(d["Sx"].shape[0])
    G = I + d["P"] @ Bc @ d["R"] @ d["Rx"]
    H = d["P"] @ Bc @ d["R"] @ d["Rb"]
    Vx = d["Sx"] @ G @ d["Sx"]
    Vb = d["Sx"] @ G @ d["Sb"] + d["Sx"] @ H + d["Sb"]
    return Vx, Vb


def coarsest(d, n=16):
    B = np.zeros_like(d["Sb"])
    for _ in range(n):
        B = d["Sx"] @ B + d["Sb"]
    return B


def fcycle_map(M):
    """b0 -> x0 of an F-cycle from level 0 (core/gmg/hierarchy.py:131-151)."""
    deep = len(M) - 1
    down = [np.eye(M[0]["Sx"].shape[0])]
    for l in range(deep):
        down.append(M[l]["R"] @ down[-1])
    X = coarsest(M[deep]) @ down[deep]
    for l in range(deep - 1, -1, -1):
        Vx, Vb = vcycle_maps(M, l)
        X = Vx @ M[l]["P"] @ X + Vb @ down[l]
    return X


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_maps_are_affine(mats):
    gm, L, M = mats
    rng = np.random.default_rng(1)
    for l, lv in enumerate(L):
        x, b = rng.standard_normal(lv.N), rng.standard_normal(lv.N)
        assert rel(lv.smooth(x, b), M[l]["Sx"] @ x + M[l]["Sb"] @ b) < 1e-13
        assert rel(lv.residual(x, b), M[l]["Rx"] @ x + M[l]["Rb"] @ b) < 1e-13


def test_vcycle_matches_textbook_algebra(mats):
    gm, L, M = mats
    rng = np.random.default_rng(2)
    for l in range(len(L) - 1):
        Vx, Vb = vcycle_maps(M, l)
        for trial in range(2):
            x = rng.standard_normal(L[l].shape)
            b = rng.standard_normal(L[l].shape)
            orc.fillhalo(x, NH)
            orc.fillhalo(b, NH)
            gm.x(l)[...] = x
            gm.b(l)[...] = b
            gm.Vcycle(l)
            got = gm.x(l).ravel().copy()
            want = Vx @ x.ravel() + Vb @ b.ravel()
            assert rel(got, want) < 1e-12, (l, trial, rel(got, want))


def test_fcycle_matches_textbook_algebra(mats):
    gm, L, M = mats
    rng = np.random.default_rng(3)
    X = fcycle_map(M)
    b = rng.standard_normal(L[0].shape)
    orc.fillhalo(b, NH)
    gm.b(0)[...] = b
    gm.x(0)[...] = rng.standard_normal(L[0].shape)  # overwritten by the cycle
    gm.Fcycle(0)
    got = gm.x(0).ravel().copy()
    assert rel(got, X @ b.ravel()) < 1e-12


@pytest.mark.parametrize("maxite", [1, 2, 3])
def test_solve_is_defect_correction_with_fcycles(mats, maxite):
    gm, L, M = mats
    rng = np.random.default_rng(4 + maxite)
    X = fcycle_map(M)
    x0 = rng.standard_normal(L[0].shape)
    b = rng.standard_normal(L[0].shape)
    orc.fillhalo(x0, NH)
    orc.fillhalo(b, NH)
    x = x0.copy()
    nite, _ = gm.solve(x, b, maxite=maxite, tol=0.0)
    assert nite == maxite
    want = x0.ravel()
    for _ in range(maxite):
        want = want + X @ (M[0]["Rx"] @ want + M[0]["Rb"] @ b.ravel())
    assert rel(x.ravel(), want) < 1e-11


def test_twovcycle_is_two_vcycles_from_the_guess(mats):
    gm, L, M = mats
    rng = np.random.default_rng(9)
    Vx, Vb = vcycle_maps(M, 0)
    x0 = rng.standard_normal(L[0].shape)
    b = rng.standard_normal(L[0].shape)
    orc.fillhalo(x0, NH)
    orc.fillhalo(b, NH)
    x = x0.copy()
    gm.twoVcycle(x, b)
    want = Vx @ (Vx @ x0.ravel() + Vb @ b.ravel()) + Vb @ b.ravel()
    assert rel(x.ravel(), want) < 1e-12


def test_vcycle_converges_to_the_masked_solution(mats):
    """What the mathematics fixes, with no reference to the cycle order: on
    a domain with walls the V-cycle iteration contracts (spectral radius of
    Vx below 1) and its fixed point x* = (I - Vx)^-1 Vb b has zero
    residual on the fluid points."""
    gm, L, M = mats
    if np.all(L[0].msk[NH:-NH, NH:-NH] == 1):
        pytest.skip("periodic: the constant mode is a null space of A")
    Vx, Vb = vcycle_maps(M, 0)
    rho = np.abs(np.linalg.eigvals(Vx)).max()
    assert rho < 0.5, rho
    rng = np.random.default_rng(5)
    b = rng.standard_normal(L[0].shape)
    orc.fillhalo(b, NH)
    xs = np.linalg.solve(np.eye(Vx.shape[0]) - Vx, Vb @ b.ravel())
    r = M[0]["Rx"] @ xs + M[0]["Rb"] @ b.ravel()
    assert np.abs(r).max() < 1e-9 * np.abs(b).max() * np.abs(M[0]["Rx"]).max()
