"""Problem descriptions and the synthetic state points the benchmarks and tests use.

A :class:`ProblemSpec` carries exactly the inputs the reference fill reads from
its globals (grid, boundary conditions, functional switches, fluid parameters,
``Vext`` and ``Zero_density_TF``) plus a seeded density field.  The builders
below construct the five state points of the hot-path workload:

* ``hard_wall_1d``     -- 1D Rosenfeld hard spheres against a hard planar wall
* ``periodic_3d``      -- 3D periodic Rosenfeld + WCA mean-field fluid
* ``slit_pores_1d``    -- binary mixture in LJ 9-3 slit pores (one pore each)
* ``rough_wall_2d``    -- 2D fluid at a rough LJ 9-3 wall
* ``solute_3d``        -- 3D White Bear mixture around a spherical solute

Random inputs come from SplitMix64 so that every harness (C, HIP, numpy)
reproduces them bit for bit.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Dict, List, Optional

import numpy as np

IN_WALL, IN_BULK, PERIODIC, REFLECT = -1, 0, 1, 2
FMT1, FMT3 = 0, 2
NCOMP_MAX = 4


# --------------------------------------------------------------------------- RNG
class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014): 64-bit state, golden-gamma increment."""

    GAMMA = np.uint64(0x9E3779B97F4A7C15)
    M1 = np.uint64(0xBF58476D1CE4E5B9)
    M2 = np.uint64(0x94D049BB133111EB)

    def __init__(self, seed: int):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)

    def next_u64(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            steps = np.arange(1, n + 1, dtype=np.uint64)
            z = self.state + steps * self.GAMMA
            self.state = self.state + np.uint64(n) * self.GAMMA
            z = (z ^ (z >> np.uint64(30))) * self.M1
            z = (z ^ (z >> np.uint64(27))) * self.M2
            z = z ^ (z >> np.uint64(31))
        return z

    def uniform(self, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
        u = (self.next_u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u


# --------------------------------------------------------------------------- spec
@dataclasses.dataclass
class ProblemSpec:
    name: str
    ndim: int
    ncomp: int
    nx: List[int]
    dx: List[float]
    bc: List[List[int]]
    type_func: int = FMT1
    type_attr: int = -1
    nwall: int = 0
    vext_max: float = 10.0
    hs_diam: List[float] = dataclasses.field(default_factory=lambda: [1.0])
    sigma_ff: Optional[np.ndarray] = None
    eps_ff: Optional[np.ndarray] = None     # epsilon / kT
    cut_ff: Optional[np.ndarray] = None
    rho_b: List[float] = dataclasses.field(default_factory=lambda: [0.5])
    vext: Optional[np.ndarray] = None       # [N, C]
    zero_tf: Optional[np.ndarray] = None    # [N, C] uint8
    rho0: Optional[np.ndarray] = None       # [C, N] initial densities
    att_in_a22: int = 0                     # 1: no mean-field unknowns (the reference's ATTInA22Block)
    lhard_surf: int = 0                     # 1: hard surfaces (Lhard_surf)
    hw_mask: Optional[np.ndarray] = None    # [ncomp][N] uint8 fluid-element bits of boundary nodes

    @property
    def nnodes(self) -> int:
        return int(np.prod(self.nx[: self.ndim]))

    @property
    def nrb(self) -> int:
        return 4 + 2 * self.ndim

    @property
    def nmf(self) -> int:
        """Mean-field unknowns of the layout (none with the attraction in the A22 block)."""
        return self.ncomp if (self.type_attr >= 0 and not self.att_in_a22) else 0

    @property
    def nunk(self) -> int:
        return self.ncomp + self.nrb + self.nmf

    def as_params(self) -> Dict:
        C = self.ncomp
        pad = lambda v: list(v) + [0.0] * (NCOMP_MAX - len(v))
        mat = lambda m: [[float(m[i][j]) if (i < C and j < C) else 0.0 for j in range(NCOMP_MAX)] for i in range(NCOMP_MAX)]
        sig = self.sigma_ff if self.sigma_ff is not None else np.ones((C, C))
        eps = self.eps_ff if self.eps_ff is not None else np.zeros((C, C))
        cut = self.cut_ff if self.cut_ff is not None else np.zeros((C, C))
        return dict(
            ndim=self.ndim, ncomp=C,
            nx=list(self.nx) + [1] * (3 - len(self.nx)),
            dx=list(self.dx) + [1.0] * (3 - len(self.dx)),
            bc=[list(b) for b in self.bc] + [[IN_BULK, IN_BULK]] * (3 - len(self.bc)),
            type_func=self.type_func, type_attr=self.type_attr, nwall=self.nwall, lhard_surf=int(self.lhard_surf),
            vext_max=self.vext_max, hs_diam=pad(self.hs_diam), sigma_ff=mat(sig), eps_ff=mat(eps),
            cut_ff=mat(cut), rho_b=pad(self.rho_b), att_in_a22=int(self.att_in_a22),
        )

    def initial_x(self) -> np.ndarray:
        """x[nunk, N] with the seeded densities; rho-bar / mean-field rows zero
        (fill them with the library's init_fields)."""
        x = np.zeros((self.nunk, self.nnodes))
        x[: self.ncomp] = self.rho0
        return x

    def positions(self) -> List[np.ndarray]:
        return [np.arange(self.nx[d]) * self.dx[d] for d in range(self.ndim)]


def lorentz_berthelot(sig, eps, cut):
    """Mixing rules of pot_parameters (src/dft_potential_params.c:61-63)."""
    C = len(sig)
    S = np.zeros((C, C)); E = np.zeros((C, C)); R = np.zeros((C, C))
    for i in range(C):
        for j in range(C):
            S[i, j] = 0.5 * (sig[i] + sig[j])
            E[i, j] = math.sqrt(eps[i] * eps[j])
            R[i, j] = 0.5 * (cut[i] + cut[j])
    return S, E, R


def nodes_along(size: float, dx: float, periodic: bool) -> int:
    """Nodes_x = round(Size_x/Esize_x) (+1 unless periodic)."""
    n = int(math.floor(size / dx + 0.5))
    return n if periodic else n + 1


def lj93_wall(z: np.ndarray, eps_w: float, sigma_w: float, cut: float) -> np.ndarray:
    """Cut-and-shifted LJ 9-3 wall potential (src/dft_vext_1D.c:42-56 form:
    eps_w [ (2/15)(s/z)^9 - (s/z)^3 ] shifted to zero at the cut-off)."""
    def u(zz):
        r = sigma_w / zz
        return eps_w * ((2.0 / 15.0) * r ** 9 - r ** 3)
    with np.errstate(divide="ignore", over="ignore", invalid="ignore"):
        out = np.where(z <= cut, u(np.maximum(z, 1e-300)) - u(cut), 0.0)
    return out


# --------------------------------------------------------------------------- builders
def hard_wall_1d(seed: int = 1, size: float = 20.0, dx: float = 0.05, rho_b: float = 0.7) -> ProblemSpec:
    """1D Rosenfeld hard spheres (sigma = 1) against a hard planar wall at x = 0,
    bulk on the right.  Centres are excluded within sigma/2 of the wall:
    V_ext = VEXT_MAX and Zero_density_TF there; seeded densities
    rho_i = rho_b (1 + 0.1 U(-1,1)) beyond 0.5 sigma."""
    nx = nodes_along(size, dx, False)
    x = np.arange(nx) * dx
    excl = x < 0.5 - 1e-12
    vext = np.where(excl, 10.0, 0.0)[:, None]
    zero = excl.astype(np.uint8)[:, None]
    rng = SplitMix64(seed)
    rho = rho_b * (1.0 + 0.1 * rng.uniform(nx, -1.0, 1.0))
    rho[excl] = 0.0
    return ProblemSpec(name="hard_wall_1d", ndim=1, ncomp=1, nx=[nx], dx=[dx], bc=[[IN_WALL, IN_BULK]],
                       type_func=FMT1, type_attr=-1, nwall=1, vext_max=10.0, hs_diam=[1.0],
                       sigma_ff=np.ones((1, 1)), rho_b=[rho_b], vext=vext, zero_tf=zero, rho0=rho[None, :])


def periodic_3d(seed: int = 2, n=64, dx: float = 0.1, rho_b: float = 0.6, cut: float = 2.5,
                eps: float = 1.0, kT: float = 1.0, func: int = FMT1, modes: int = 8) -> ProblemSpec:
    """3D periodic Rosenfeld + WCA mean-field fluid (LJ 12-6, eps, cut-off `cut`) on n^3 nodes, or
    on n = (nx, ny, nz) nodes.  Seeded field rho = clip(rho_b + 0.2 f, 0.05, 0.9) with the plain sum
    f = sum_m cos(k_m . r + phi_m) of `modes` random-phase modes; each k_m = 2 pi (n_m / L) is a
    lattice vector of the periodic box with wavelength 2 pi / |k_m| in [2, 6] sigma.  With 8 modes
    f spans [-8, 8], so the clip bounds are reached over much of the box."""
    ns = [n, n, n] if np.isscalar(n) else [int(v) for v in n]
    L = np.asarray(ns, dtype=float) * dx
    if L.max() < 2.0:
        raise ValueError("periodic_3d: the box (n dx) must be at least 2 sigma for modes of wavelength 2-6 sigma")
    rng = SplitMix64(seed)
    ks = []
    tries = 0
    while len(ks) < modes:
        tries += 1
        if tries > 100000:
            raise ValueError("periodic_3d: no lattice modes of wavelength 2-6 sigma in this box")
        v = np.round(rng.uniform(3, -4.0, 4.0)).astype(int)
        nrm = float(np.linalg.norm(v / L))
        if nrm == 0:
            continue
        lam = 1.0 / nrm
        if 2.0 <= lam <= 6.0:
            ks.append(v)
    phis = rng.uniform(modes, 0.0, 2 * math.pi)
    gx, gy, gz = (np.arange(m) * dx for m in ns)
    X, Y, Z = np.meshgrid(gx, gy, gz, indexing="ij")
    f = np.zeros_like(X)
    for v, ph in zip(ks, phis):
        k = 2 * math.pi * np.asarray(v, dtype=float) / L
        f += np.cos(k[0] * X + k[1] * Y + k[2] * Z + ph)
    rho = np.clip(rho_b + 0.2 * f, 0.05, 0.9)
    # node index = i + j nx + k nx ny: x fastest
    rho = np.transpose(rho, (2, 1, 0)).reshape(-1)
    attr = 0 if cut > 0 else -1
    return ProblemSpec(name="periodic_3d", ndim=3, ncomp=1, nx=ns, dx=[dx, dx, dx],
                       bc=[[PERIODIC, PERIODIC]] * 3, type_func=func, type_attr=attr, nwall=0,
                       hs_diam=[1.0], sigma_ff=np.ones((1, 1)), eps_ff=np.full((1, 1), eps / kT),
                       cut_ff=np.full((1, 1), cut), rho_b=[rho_b], vext=None, zero_tf=None,
                       rho0=rho[None, :])


def periodic_mix3(seed: int = 9, n=(128, 12, 12), dx: float = 0.25, sig=(1.0, 0.9, 0.7), eta_b: float = 0.35,
                  cut_fac: float = 1.2) -> ProblemSpec:
    """Three-component White Bear + WCA fluid (sigma_i = sig, eps = 1, cut-off cut_fac sigma_i,
    Lorentz-Berthelot cross terms) on a periodic box: bulk packing fraction eta_b split over the
    species by seeded Dirichlet weights (as solute_3d), and each species' density the bulk value
    times (1 + 0.5 f), f the seeded mode sum of periodic_3d scaled into [-1, 1]."""
    C = len(sig)
    eta = dirichlet_split(eta_b, C, seed)
    rho_b = [float(eta[c] / (math.pi * sig[c] ** 3 / 6.0)) for c in range(C)]
    base = periodic_3d(seed=seed, n=n, dx=dx, rho_b=0.0, cut=1.0)
    f = np.clip(base.rho0[0] / 0.2, -8.0, 8.0) / 8.0
    S, E, R = lorentz_berthelot(list(sig), [1.0] * C, [cut_fac * s for s in sig])
    rho0 = np.stack([rho_b[c] * (1.0 + 0.5 * f) for c in range(C)])
    return ProblemSpec(name="periodic_mix3", ndim=3, ncomp=C, nx=list(base.nx), dx=[dx] * 3,
                       bc=[[PERIODIC, PERIODIC]] * 3, type_func=FMT3, type_attr=0, nwall=0, hs_diam=list(sig),
                       sigma_ff=S, eps_ff=E, cut_ff=R, rho_b=rho_b, rho0=rho0)


def slit_pore_1d(H: float, dx: float, sig=(1.0, 0.8), rho_b=(0.2, 0.2), eps_w: float = 1.0,
                 cut_fac: float = 2.5, seed: int = 3, vext_max: float = 10.0, eps_ff: float = 1.0) -> ProblemSpec:
    """Binary Rosenfeld + WCA fluid (fluid-fluid eps_ff/kT, cut-off cut_fac
    sigma_i) in a slit of width H between two LJ 9-3 walls (eps_w,
    sigma_wf = sigma_i), IN_WALL outside.  Nodes with V_ext >= VEXT_MAX are
    zero-density nodes.  Seeded initial densities rho_b exp(-max(V_ext, 0))
    (1 + 0.05 U(-1, 1)): the Boltzmann factor only where the wall repels, since
    amplifying dense states in the attractive well can start Newton past
    packing (n3 > 1)."""
    C = len(sig)
    nx = nodes_along(H, dx, False)
    z = np.arange(nx) * dx
    vext = np.zeros((nx, C))
    for c in range(C):
        v = lj93_wall(z, eps_w, sig[c], 2.5 * sig[c] + 100.0) + lj93_wall(H - z, eps_w, sig[c], 2.5 * sig[c] + 100.0)
        v = np.where(np.isfinite(v), v, vext_max)
        v = np.where(v >= vext_max, vext_max, v)
        vext[:, c] = v
    zero = (vext >= vext_max).astype(np.uint8)
    S, E, R = lorentz_berthelot(list(sig), [eps_ff] * C, [cut_fac * s for s in sig])
    rng = SplitMix64(seed)
    rho0 = np.zeros((C, nx))
    for c in range(C):
        rho0[c] = rho_b[c] * np.exp(-np.clip(vext[:, c], 0.0, 50.0)) * (1.0 + 0.05 * rng.uniform(nx, -1, 1))
        rho0[c][zero[:, c] == 1] = 0.0
    return ProblemSpec(name="slit_pore_1d", ndim=1, ncomp=C, nx=[nx], dx=[dx], bc=[[IN_WALL, IN_WALL]],
                       type_func=FMT1, type_attr=0, nwall=2, vext_max=vext_max, hs_diam=list(sig),
                       sigma_ff=S, eps_ff=E, cut_ff=R, rho_b=list(rho_b), vext=vext, zero_tf=zero, rho0=rho0)


def gaussian_rough_profile(n: int, dx: float, amplitude: float, corr_len: float, seed: int) -> np.ndarray:
    """Periodic random height profile h(y) on n nodes: SplitMix64 white noise
    filtered by a Gaussian of correlation length corr_len (FFT, periodic), zero
    mean, scaled so that max |h| = amplitude."""
    rng = SplitMix64(seed)
    white = rng.uniform(n, -1.0, 1.0)
    k = 2 * math.pi * np.fft.fftfreq(n, d=dx)
    filt = np.exp(-0.25 * (k * corr_len) ** 2)
    h = np.real(np.fft.ifft(np.fft.fft(white) * filt))
    h -= h.mean()
    return amplitude * h / np.abs(h).max()


def rough_wall_2d(seed: int = 4, nx: int = 256, ny: int = 256, dx: float = 0.05, eps_w: float = 1.5,
                  amplitude: float = 0.5, corr_len: float = 2.0, rho_b: float = 0.5, cut: float = 2.5,
                  wall_pos: float = 1.0, wall_cut: float = 3.0, z_excl: float = 0.62,
                  vext_max: float = 10.0) -> ProblemSpec:
    """2D Rosenfeld + WCA fluid (sigma = 1, eps = 1, cut-off `cut`, kT = 1) at a
    rough LJ 9-3 wall on the x = 0 face.  The wall surface sits at
    x_s(y) = wall_pos + h(y), h a Gaussian-correlated periodic profile
    (gaussian_rough_profile); V_ext(x, y) = lj93_wall(x - x_s(y); eps_w,
    sigma_w = 1, cut wall_cut).  Nodes with x - x_s(y) < z_excl are excluded
    (Zero_density_TF, V_ext = VEXT_MAX); elsewhere V_ext < VEXT_MAX, so the
    excluded set does not depend on eps_w.  Boundaries: IN_WALL at x = 0,
    IN_BULK at x = L_x, periodic in y."""
    yv = np.arange(ny) * dx
    xv = np.arange(nx) * dx
    h = gaussian_rough_profile(ny, dx, amplitude, corr_len, seed)
    zs = xv[None, :] - (wall_pos + h[:, None])          # [ny, nx]
    excl = zs < z_excl
    v = lj93_wall(np.where(excl, 1.0, zs), eps_w, 1.0, wall_cut)
    if np.any(v[~excl] >= vext_max):
        raise ValueError("V_ext reaches VEXT_MAX outside the excluded region; raise z_excl")
    v = np.where(excl, vext_max, v)
    vext = v.reshape(-1)[:, None]                        # node = i + j*nx: x fastest
    zero = excl.reshape(-1).astype(np.uint8)[:, None]
    rng = SplitMix64(seed + 1)
    rho = rho_b * np.exp(-np.minimum(vext[:, 0], 30.0)) * (1.0 + 0.02 * rng.uniform(nx * ny, -1.0, 1.0))
    rho[zero[:, 0] == 1] = 0.0
    return ProblemSpec(name="rough_wall_2d", ndim=2, ncomp=1, nx=[nx, ny], dx=[dx, dx],
                       bc=[[IN_WALL, IN_BULK], [PERIODIC, PERIODIC]], type_func=FMT1, type_attr=0, nwall=1,
                       vext_max=vext_max, hs_diam=[1.0], sigma_ff=np.ones((1, 1)), eps_ff=np.ones((1, 1)),
                       cut_ff=np.full((1, 1), cut), rho_b=[rho_b], vext=vext, zero_tf=zero, rho0=rho[None, :])


def rough_slit_pore_2d(seed: int = 3, nx: int = 64, ny: int = 61, dx: float = 0.1, eps_w: float = 1.0,
                       amplitude: float = 0.3, corr_len: float = 1.5, rho_b: float = 0.5, cut: float = 2.5,
                       wall_pos: float = 0.3, wall_cut: float = 3.0, z_excl: float = 0.62,
                       vext_max: float = 10.0, eps_ff: float = 1.0) -> ProblemSpec:
    """2D Rosenfeld + WCA fluid (sigma = 1, eps = eps_ff, cut-off `cut`, kT = 1) in a slit between two
    rough LJ 9-3 walls, periodic along x.  The lower surface sits at y_l(x) = wall_pos + h_l(x),
    the upper at y_u(x) = (ny - 1) dx - wall_pos - h_u(x), with h_l, h_u seeded Gaussian-correlated
    periodic profiles (gaussian_rough_profile); V_ext = lj93_wall(y - y_l) + lj93_wall(y_u - y)
    (eps_w, sigma_w = 1, cut wall_cut).  Nodes closer than z_excl to either surface are excluded
    (Zero_density_TF, V_ext = VEXT_MAX).  Boundaries: periodic in x, IN_WALL at both y ends."""
    xv = np.arange(nx) * dx
    yv = np.arange(ny) * dx
    hl = gaussian_rough_profile(nx, dx, amplitude, corr_len, seed)
    hu = gaussian_rough_profile(nx, dx, amplitude, corr_len, seed + 7)
    zl = yv[:, None] - (wall_pos + hl[None, :])                      # [ny, nx]
    zu = ((ny - 1) * dx - wall_pos - hu[None, :]) - yv[:, None]
    excl = (zl < z_excl) | (zu < z_excl)
    v = lj93_wall(np.where(excl, 1.0, zl), eps_w, 1.0, wall_cut) + lj93_wall(np.where(excl, 1.0, zu), eps_w, 1.0,
                                                                             wall_cut)
    if np.any(v[~excl] >= vext_max):
        raise ValueError("V_ext reaches VEXT_MAX outside the excluded region; raise z_excl")
    v = np.where(excl, vext_max, v)
    vext = v.reshape(-1)[:, None]                                   # node = i + j*nx: x fastest
    zero = excl.reshape(-1).astype(np.uint8)[:, None]
    rng = SplitMix64(seed + 1)
    # exp(-max(V, 0)): the attractive wells are not amplified into an over-packed start
    rho = rho_b * np.exp(-np.clip(vext[:, 0], 0.0, 30.0)) * (1.0 + 0.02 * rng.uniform(nx * ny, -1.0, 1.0))
    rho[zero[:, 0] == 1] = 0.0
    return ProblemSpec(name="rough_slit_pore_2d", ndim=2, ncomp=1, nx=[nx, ny], dx=[dx, dx],
                       bc=[[PERIODIC, PERIODIC], [IN_WALL, IN_WALL]], type_func=FMT1, type_attr=0, nwall=1,
                       vext_max=vext_max, hs_diam=[1.0], sigma_ff=np.ones((1, 1)), eps_ff=np.full((1, 1), eps_ff),
                       cut_ff=np.full((1, 1), cut), rho_b=[rho_b], vext=vext, zero_tf=zero, rho0=rho[None, :])


def dirichlet_split(total: float, k: int, seed: int) -> np.ndarray:
    """total * (w_1..w_k), w ~ Dirichlet(1,...,1) from SplitMix64 (normalised exponentials)."""
    e = -np.log(SplitMix64(seed).uniform(k, 1e-300, 1.0))
    return total * e / e.sum()


def solute_3d(seed: int = 5, n: int = 128, dx: float = 0.1, sig=(1.0, 0.9, 0.7), eta_b: float = 0.35,
              radius: float = 4.0, eps_w: float = 1.0, cut_fac: float = 3.0, wall_cut: float = 3.0,
              z_excl: float = 0.62, vext_max: float = 10.0) -> ProblemSpec:
    """3D White Bear + WCA mixture (sigma_i = sig, eps = 1, cut-off 3 sigma_i,
    Lorentz-Berthelot cross terms) around a spherical solute of `radius` at the
    box centre, periodic box of n^3 nodes.  Solute-fluid potential: LJ 9-3 in
    the surface distance s = r - radius (eps_w, sigma_w = sigma_i, cut wall_cut
    sigma_i); nodes with s < z_excl sigma_i are excluded for that species.
    Bulk packing fraction eta_b split over the species by seeded Dirichlet
    weights, rho_i = eta_i / (pi sigma_i^3 / 6)."""
    C = len(sig)
    eta = dirichlet_split(eta_b, C, seed)
    rho_b = [float(eta[c] / (math.pi * sig[c] ** 3 / 6.0)) for c in range(C)]
    g = (np.arange(n) - 0.5 * (n - 1)) * dx
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")        # node = i + j n + k n^2: x fastest
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2).reshape(-1)
    vext = np.zeros((r.size, C))
    zero = np.zeros((r.size, C), dtype=np.uint8)
    for c in range(C):
        s = r - radius
        excl = s < z_excl * sig[c]
        v = lj93_wall(np.where(excl, sig[c], s), eps_w, sig[c], wall_cut * sig[c])
        if np.any(v[~excl] >= vext_max):
            raise ValueError("V_ext reaches VEXT_MAX outside the excluded region; raise z_excl")
        vext[:, c] = np.where(excl, vext_max, v)
        zero[:, c] = excl
    S, E, R = lorentz_berthelot(list(sig), [1.0] * C, [cut_fac * s for s in sig])
    rho0 = np.zeros((C, r.size))
    for c in range(C):
        rho0[c] = rho_b[c] * np.exp(-np.minimum(vext[:, c], 30.0))
        rho0[c][zero[:, c] == 1] = 0.0
    return ProblemSpec(name="solute_3d", ndim=3, ncomp=C, nx=[n, n, n], dx=[dx] * 3,
                       bc=[[PERIODIC, PERIODIC]] * 3, type_func=FMT3, type_attr=0, nwall=1,
                       vext_max=vext_max, hs_diam=list(sig), sigma_ff=S, eps_ff=E, cut_ff=R, rho_b=rho_b,
                       vext=vext, zero_tf=zero, rho0=rho0)


def reflect_slab_1d(half: bool, seed: int = 7, M: int = 60, dx: float = 0.05, a: int = 2,
                    rho_b: float = 0.6, attraction: bool = True) -> ProblemSpec:
    """A hard slab of 2a elements centred on node M of a 1D line of 2M + 1 nodes (IN_BULK at both
    ends), with hard surfaces (Lhard_surf) and a symmetric seeded density; `half` keeps nodes
    M..2M with a REFLECT boundary at the slab's centre instead.  The two are the same system:
    every stencil point that crosses the REFLECT plane in the half line lands on the mirror node of
    the full line, and at the slab's surface node a the fluid element lies on the other side of the
    mirror image (mask bit 0 in the half line, bit 1 at node M - a of the full line)."""
    n_full = 2 * M + 1
    own = np.zeros(n_full - 1, dtype=bool)            # wall elements
    centre = (np.arange(n_full - 1) + 0.5) - M
    own[np.abs(centre) < a] = True
    rng = SplitMix64(seed)
    r = rho_b * (1.0 + 0.1 * rng.uniform(M + 1, -1.0, 1.0))
    rho = np.concatenate([r[:0:-1], r])               # symmetric about node M
    zero = np.zeros(n_full, dtype=np.uint8)
    mask = np.zeros(n_full, dtype=np.uint8)
    for n in range(n_full):
        left = own[n - 1] if n >= 1 else False
        right = own[n] if n < n_full - 1 else False
        if left and right:
            zero[n] = 1
        elif left != right:
            mask[n] = 1 if not right else 2           # bit 0: element to the right is fluid
    rho[zero == 1] = 0.0
    vext = np.where(zero == 1, 10.0, 0.0)
    sl = slice(M, None) if half else slice(None)
    nx = M + 1 if half else n_full
    bc = [[REFLECT, IN_BULK]] if half else [[IN_BULK, IN_BULK]]
    return ProblemSpec(name="reflect_slab_1d" + ("_half" if half else ""), ndim=1, ncomp=1, nx=[nx], dx=[dx],
                       bc=bc, type_func=FMT1, type_attr=0 if attraction else -1, nwall=1, vext_max=10.0,
                       hs_diam=[1.0], sigma_ff=np.ones((1, 1)), eps_ff=np.full((1, 1), 0.8),
                       cut_ff=np.full((1, 1), 1.5), rho_b=[rho_b], vext=vext[sl, None].copy(),
                       zero_tf=zero[sl, None].copy(), rho0=rho[None, sl].copy(), lhard_surf=1,
                       hw_mask=mask[None, sl].copy())


def small_case(kind: str, seed: int = 7) -> ProblemSpec:
    """Small state points the oracle evaluates in seconds (parity tests)."""
    if kind == "1d_reflect_slab":
        return reflect_slab_1d(True, seed=seed)
    if kind == "1d_hw":
        return hard_wall_1d(seed=seed, size=6.0, dx=0.05)
    if kind == "1d_hw_full":
        return hard_wall_1d(seed=seed)
    if kind == "1d_reflect_wca":
        sp = hard_wall_1d(seed=seed, size=5.0, dx=0.05, rho_b=0.6)
        sp.bc = [[REFLECT, IN_BULK]]
        sp.nwall = 0
        sp.vext = None
        sp.zero_tf = None
        rng = SplitMix64(seed + 11)
        sp.rho0 = (0.6 * (1 + 0.1 * rng.uniform(sp.nnodes, -1, 1)))[None, :]
        sp.type_attr = 0
        sp.eps_ff = np.full((1, 1), 1.0)
        sp.cut_ff = np.full((1, 1), 2.5)
        return sp
    if kind == "1d_pore_mix":
        return slit_pore_1d(H=4.0, dx=0.02, seed=seed)
    if kind == "2d_periodic":
        n = 32
        sp = periodic_3d(seed=seed, n=n, dx=0.1, cut=1.5)
        # reduce to 2D: take the z = 0 plane
        rho = sp.rho0.reshape(n, n, n)[0].reshape(-1)
        return ProblemSpec(name="2d_periodic", ndim=2, ncomp=1, nx=[n, n], dx=[0.1, 0.1],
                           bc=[[PERIODIC, PERIODIC]] * 2, type_func=FMT1, type_attr=0,
                           hs_diam=[1.0], sigma_ff=np.ones((1, 1)), eps_ff=np.ones((1, 1)),
                           cut_ff=np.full((1, 1), 1.5), rho_b=[0.6], rho0=rho[None, :])
    if kind == "2d_wall_reflect":
        nxx, nyy = 20, 16
        dx = 0.1
        rng = SplitMix64(seed)
        x = np.arange(nxx) * dx
        excl = np.repeat((x < 0.5 - 1e-12)[None, :], nyy, axis=0).reshape(-1)
        rho = 0.55 * (1 + 0.1 * rng.uniform(nxx * nyy, -1, 1))
        rho[excl] = 0.0
        vext = np.where(excl, 10.0, 0.0)[:, None]
        return ProblemSpec(name="2d_wall_reflect", ndim=2, ncomp=1, nx=[nxx, nyy], dx=[dx, dx],
                           bc=[[IN_WALL, IN_BULK], [REFLECT, REFLECT]], type_func=FMT1, type_attr=0,
                           nwall=1, hs_diam=[1.0], sigma_ff=np.ones((1, 1)), eps_ff=np.full((1, 1), 0.8),
                           cut_ff=np.full((1, 1), 1.5), rho_b=[0.55], vext=vext,
                           zero_tf=excl.astype(np.uint8)[:, None], rho0=rho[None, :])
    if kind == "3d_periodic_small":
        return periodic_3d(seed=seed, n=12, dx=0.25, cut=1.2)
    if kind == "3d_periodic_wb":
        return periodic_3d(seed=seed, n=12, dx=0.25, cut=1.2, func=FMT3)
    if kind == "3d_mix_bulk":
        n = 14
        sp = periodic_3d(seed=seed, n=n, dx=0.2, cut=1.2, func=FMT3)
        sig = [1.0, 0.8]
        S, E, R = lorentz_berthelot(sig, [1.0, 1.0], [1.2, 1.0])
        rng = SplitMix64(seed + 5)
        rho0 = np.stack([0.3 * (1 + 0.2 * rng.uniform(sp.nnodes, -1, 1)), 0.25 * (1 + 0.2 * rng.uniform(sp.nnodes, -1, 1))])
        return ProblemSpec(name="3d_mix_bulk", ndim=3, ncomp=2, nx=[n, n, n], dx=[0.2] * 3,
                           bc=[[IN_BULK, IN_BULK], [PERIODIC, PERIODIC], [REFLECT, IN_WALL]],
                           type_func=FMT3, type_attr=0, nwall=0, hs_diam=sig, sigma_ff=S, eps_ff=E,
                           cut_ff=R, rho_b=[0.3, 0.25], rho0=rho0)
    if kind == "3d_mix_periodic":
        n = 12
        sp = periodic_3d(seed=seed, n=n, dx=0.25, cut=1.2, func=FMT3)
        sig = [1.0, 0.8]
        S, E, R = lorentz_berthelot(sig, [1.0, 1.0], [1.2, 1.0])
        rho0 = np.stack([0.5 * sp.rho0[0], 0.4 * sp.rho0[0]])
        return ProblemSpec(name="3d_mix_periodic", ndim=3, ncomp=2, nx=[n, n, n], dx=[0.25] * 3,
                           bc=[[PERIODIC, PERIODIC]] * 3, type_func=FMT3, type_attr=0, nwall=0, hs_diam=sig,
                           sigma_ff=S, eps_ff=E, cut_ff=R, rho_b=[0.3, 0.24], rho0=rho0)
    if kind == "2d_rough_wall":
        return rough_wall_2d(seed=seed, nx=40, ny=32, dx=0.1, cut=1.5, wall_cut=2.0)
    if kind == "3d_solute_small":
        return solute_3d(seed=seed, n=20, dx=0.25, sig=(1.0, 0.8), radius=1.2, cut_fac=1.2, wall_cut=2.0)
    if kind == "3d_periodic_fmt2":
        return periodic_3d(seed=seed, n=12, dx=0.25, cut=1.2, func=1)
    if kind == "2d_wall_fmt4":
        sp = small_case("2d_wall_reflect", seed)
        sp.type_func = 3
        return sp
    if kind == "3d_mix_fmt4":   # White Bear mixture case with FMT4 (reflect, wall and periodic edges)
        sp = small_case("3d_mix_bulk", seed)
        sp.type_func = 3
        return sp
    if kind == "3d_solute3_small":   # the 128^3 solute's three species, small box
        return solute_3d(seed=seed, n=20, dx=0.25, sig=(1.0, 0.9, 0.7), radius=1.2, cut_fac=1.2, wall_cut=2.0)
    # periodic boxes whose axis lengths select the register FFT kernels (x 64 / 128, strided y, z)
    if kind == "3d_x64":
        return periodic_3d(seed=seed, n=(64, 12, 12), dx=0.25, cut=1.2)
    if kind == "3d_x128":
        return periodic_3d(seed=seed, n=(128, 12, 12), dx=0.25, cut=1.2)
    if kind == "3d_yz64":
        return periodic_3d(seed=seed, n=(12, 64, 64), dx=0.25, cut=1.2)
    if kind == "3d_y128_z32":
        return periodic_3d(seed=seed, n=(12, 128, 32), dx=0.25, cut=1.2, func=FMT3)
    if kind == "3d_mix3_x128":
        return periodic_mix3(seed=seed, n=(128, 12, 12), dx=0.25)
    raise KeyError(kind)
