"""Initial conditions of the benchmark configurations (seeded, synthetic).

* ``vortex_pair``: two Gaussian vortices (radius 0.05, amplitudes +1/-1,
  centres 0.2 apart) plus seeded uniform noise 1e-3 with its domain mean
  removed (periodic solvability, as experiments/TaylorInstability does for
  its noise), after the vortex profile of core/ana_profiles.py.
* ``turbulence``: random-phase vorticity whose shell energy spectrum is
  k^-3 above k0 = 8 (|w_k| ~ 1/k), normalised to rms 1, after
  experiments/Twodim_turbulence/freedecay/freedecay.py.
* ``rayleigh_taylor``: buoyancy +0.5 below / -0.5 above an interface at
  mid-height perturbed by 1e-3 (sin(2 pi x / Lx) + seeded noise), after
  experiments/TaylorInstability/taylor_instability.py.

Each returns host arrays with the periodic halo filled and the mask applied.
"""
import numpy as np

from .grid import Grid
from .param import Param

NH = 3


def _fill(grid, a):
    grid.fill_halo(a)
    return a


def vortex_pair(grid, seed=1):
    x, y = grid.xr, grid.yr
    r0 = 0.05
    rp = (x - 0.4 * grid.Lx) ** 2 + (y - 0.5 * grid.Ly) ** 2
    rm = (x - 0.6 * grid.Lx) ** 2 + (y - 0.5 * grid.Ly) ** 2
    w = np.exp(-rp / r0 ** 2) - np.exp(-rm / r0 ** 2)
    rng = np.random.default_rng(seed)
    noise = rng.uniform(-1.0, 1.0, w.shape) * grid.msk
    noise -= grid.domain_integration(noise) * grid.msk / grid.area
    w = (w + 1e-3 * noise) * grid.msk
    return _fill(grid, w)


def turbulence(grid, seed=2, k0=8.0):
    nx, ny = grid.nx, grid.ny
    rng = np.random.default_rng(seed)
    kx = np.fft.fftfreq(nx, 1.0 / nx)
    ky = np.fft.fftfreq(ny, 1.0 / ny)
    KX, KY = np.meshgrid(kx, ky)
    k = np.sqrt(KX ** 2 + KY ** 2)
    amp = np.zeros_like(k)
    sel = k > k0
    amp[sel] = 1.0 / k[sel]
    phase = np.exp(2j * np.pi * rng.random((ny, nx)))
    wi = np.real(np.fft.ifft2(amp * phase))
    wi /= np.sqrt(np.mean(wi ** 2))
    w = np.zeros((ny + 2 * NH, nx + 2 * NH))
    w[NH:-NH, NH:-NH] = wi
    w *= grid.msk
    return _fill(grid, w)


def rayleigh_taylor(grid, seed=3):
    rng = np.random.default_rng(seed)
    x, y = grid.xr, grid.yr
    eta = 0.5 * grid.Ly + 1e-3 * (np.sin(2 * np.pi * x / grid.Lx) + rng.uniform(-1, 1, x.shape))
    b = np.where(y < eta, 0.5, -0.5) * grid.msk
    return _fill(grid, b)


CONFIGS = {
    # name: (param kwargs, init, steps)
    "euler-perio-128": (dict(modelname="euler", nx=128, ny=128, Lx=1.0, Ly=1.0, geometry="perio"),
                        "vortex_pair", 200),
    "euler-closed-1024": (dict(modelname="euler", nx=1024, ny=1024, Lx=1.0, Ly=1.0,
                               geometry="closed"), "turbulence", 1000),
    "boussinesq-rt-2048x1024": (dict(modelname="boussinesq", nx=2048, ny=1024, Lx=2.0, Ly=1.0,
                                     geometry="xchannel", diffusion=True, Kdiff=1e-5,
                                     gravity=1.0), "rayleigh_taylor", 1000),
}


def ensemble_specs(n=256, seed=4, sizes=(64, 128, 256, 512), geometries=("perio", "closed")):
    """Batched ensemble: n independent Euler runs, grid size and geometry
    drawn (seeded) from ``sizes`` and ``geometries``; initial field k^-3
    turbulence with seed 1000 + i."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        s = int(rng.choice(sizes))
        geo = str(rng.choice(geometries))
        kw = dict(modelname="euler", nx=s, ny=s, Lx=1.0, Ly=1.0, geometry=geo,
                  adaptable_dt=False, order=5, flux_splitting_method="parabolic",
                  timestepping="RK3_SSP")
        out.append((kw, 1000 + i))
    return out


def make_param(name, **over):
    kw, _, _ = CONFIGS[name]
    kw = dict(kw)
    kw.update(adaptable_dt=False, order=5, flux_splitting_method="parabolic",
              timestepping="RK3_SSP")
    kw.update(over)
    return Param(**kw)


def initial_fields(name, grid, seed=None):
    """{'vorticity': ..., ['buoyancy': ...]} for configuration ``name``."""
    _, init, _ = CONFIGS[name]
    if init == "vortex_pair":
        return {"vorticity": vortex_pair(grid, 1 if seed is None else seed)}
    if init == "turbulence":
        return {"vorticity": turbulence(grid, 2 if seed is None else seed)}
    return {"vorticity": np.zeros((grid.nyl, grid.nxl)),
            "buoyancy": rayleigh_taylor(grid, 3 if seed is None else seed)}


def time_step(name, grid, maxspeed0):
    """Fixed dt at CFL 0.8: of the initial maxspeed, or for the
    Rayleigh-Taylor start (at rest) of the buoyancy velocity scale
    sqrt(g * db * Ly) = 1."""
    kw, _, _ = CONFIGS[name]
    u = maxspeed0 if kw["modelname"] == "euler" else 1.0
    return 0.8 * min(grid.dx, grid.dy) / u


def build_grid(name, **over):
    return Grid(make_param(name, **over))
