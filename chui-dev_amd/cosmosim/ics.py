"""Synthetic seeded workloads for the TreePM force evaluation.

Units: length Mpc/h, mass 1e10 Msun/h, velocity km/s, so G = 43.0091 and
H0 = 100 (time unit (Mpc/h)/(km/s)). Particles are a regular lattice plus a
Zel'dovich displacement of a Gaussian random field with P(k) ~ k^n below the
particle Nyquist frequency, scaled to a stated RMS displacement in units of
the mean spacing. A fraction of the particles is moved into NFW halos
(concentration 8, masses log-uniform over two decades, uniform centres) with
the radial law of tools/generate_clustered_dmo_box.py:31-50 (inverse CDF of
ln(1+x) - x/(1+x)); the virial radius follows that tool's law
0.035 * box * (n/2000)^(1/3) for its 64^3 box, i.e. 2.24 mean spacings *
(n/2000)^(1/3), which keeps the tool's halo overdensity at any particle count.
Velocities are u = a H(a) f Psi with f = Omega_m(a)^0.55
(tests/validation/test_dmo_zeldovich_workflow.cpp:451-472).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

G_CODE = 43.0091          # (Mpc/h) (km/s)^2 / (1e10 Msun/h)
H0_CODE = 100.0           # km/s / (Mpc/h)
OMEGA_M = 0.3
BASE_SEED = 20261202


@dataclass
class Workload:
    name: str
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    vx: np.ndarray
    vy: np.ndarray
    vz: np.ndarray
    mass: np.ndarray
    box: float
    mesh: int
    softening: float
    scale_factor: float
    n_side: int
    halo_fraction: float
    halo_count: int


def hubble(a: float, omega_m: float = OMEGA_M) -> float:
    return H0_CODE * np.sqrt(omega_m / a**3 + (1.0 - omega_m))


def growth_rate(a: float, omega_m: float = OMEGA_M) -> float:
    om_a = omega_m / a**3 / (omega_m / a**3 + 1.0 - omega_m)
    return om_a**0.55


def zeldovich_displacement(n_side: int, box: float, spectral_index: float, rms_spacing: float,
                           rng: np.random.Generator) -> np.ndarray:
    """(3, n^3) displacement field, RMS (over components) = rms_spacing * box/n."""
    noise = rng.standard_normal((n_side, n_side, n_side))
    dk = np.fft.rfftn(noise)
    k1 = np.fft.fftfreq(n_side, d=1.0 / n_side) * (2.0 * np.pi / box)
    kz = np.fft.rfftfreq(n_side, d=1.0 / n_side) * (2.0 * np.pi / box)
    kx, ky, kzz = np.meshgrid(k1, k1, kz, indexing="ij")
    k2 = kx * kx + ky * ky + kzz * kzz
    k = np.sqrt(k2)
    k_ny = np.pi * n_side / box
    amp = np.zeros_like(k)
    sel = (k > 0) & (k <= k_ny)
    amp[sel] = k[sel] ** (0.5 * spectral_index)
    dk *= amp
    disp = np.empty((3, n_side**3))
    for axis, ka in enumerate((kx, ky, kzz)):
        psi_k = np.zeros_like(dk)
        psi_k[sel] = 1j * ka[sel] / k2[sel] * dk[sel]
        disp[axis] = np.fft.irfftn(psi_k, s=(n_side, n_side, n_side), axes=(0, 1, 2)).reshape(-1)
    rms = np.sqrt(np.mean(disp**2))
    spacing = box / n_side
    if rms > 0:
        disp *= rms_spacing * spacing / rms
    return disp


def nfw_radii(count: int, r_virial: float, concentration: float, rng: np.random.Generator) -> np.ndarray:
    grid = np.linspace(1.0e-4, concentration, 4000)
    cdf = np.log1p(grid) - grid / (1.0 + grid)
    cdf /= cdf[-1]
    return (r_virial / concentration) * np.interp(rng.uniform(0.0, 1.0, count), cdf, grid)


def make_box(name: str, n_side: int, box: float, mesh: int, spectral_index: float, rms_spacing: float,
             halo_fraction: float, halo_count: int, scale_factor: float, seed: int,
             mass_decades: float = 2.0) -> Workload:
    rng = np.random.default_rng(seed)
    n = n_side**3
    spacing = box / n_side
    disp = zeldovich_displacement(n_side, box, spectral_index, rms_spacing, rng)
    idx = np.arange(n_side, dtype=np.float64)
    gx, gy, gz = np.meshgrid(idx, idx, idx, indexing="ij")
    pos = np.stack([(gx.reshape(-1) + 0.5) * spacing, (gy.reshape(-1) + 0.5) * spacing,
                    (gz.reshape(-1) + 0.5) * spacing]) + disp
    vel_factor = scale_factor * hubble(scale_factor) * growth_rate(scale_factor)
    vel = vel_factor * disp
    n_halo = int(round(halo_fraction * n))
    if n_halo > 0 and halo_count > 0:
        members = rng.choice(n, size=n_halo, replace=False)
        weights = 10.0 ** (mass_decades * rng.uniform(0.0, 1.0, halo_count))
        counts = rng.multinomial(n_halo, weights / weights.sum())
        start = 0
        for c in counts:
            if c == 0:
                continue
            sel = members[start:start + c]
            start += c
            centre = rng.uniform(0.0, box, 3)
            r_vir = 2.24 * spacing * (c / 2000.0) ** (1.0 / 3.0)
            r = nfw_radii(c, r_vir, 8.0, rng)
            d = rng.standard_normal((c, 3))
            d /= np.linalg.norm(d, axis=1)[:, None]
            pos[:, sel] = (centre[None, :] + r[:, None] * d).T
            vel[:, sel] = 0.0
    pos = np.mod(pos, box)
    pos[pos >= box] = 0.0
    rho_crit = 3.0 * H0_CODE**2 / (8.0 * np.pi * G_CODE)
    m = np.full(n, rho_crit * OMEGA_M * box**3 / n)
    return Workload(name, pos[0].copy(), pos[1].copy(), pos[2].copy(), vel[0].copy(), vel[1].copy(),
                    vel[2].copy(), m, box, mesh, spacing / 30.0, scale_factor, n_side, halo_fraction, halo_count)


def config(k: int, n_side: int | None = None) -> Workload:
    """The benchmark's five workloads (seed BASE_SEED + k). n_side overrides the
    particle count for reduced test sizes, keeping box/mesh ratios."""
    if k == 0:
        ns = n_side or 32
        return make_box("plain_fp64_32^3_box10", ns, 10.0 * ns / 32, 2 * ns, -2.5, 0.2, 0.0, 0, 0.04,
                        BASE_SEED + 0)
    if k == 1:
        ns = n_side or 128
        return make_box("treepm_128^3_box100_mesh256", ns, 100.0 * ns / 128, 2 * ns, -2.5, 0.5, 0.10,
                        max(1, 200 * ns**3 // 128**3), 1.0, BASE_SEED + 1)
    if k == 2:
        ns = n_side or 256
        return make_box("treepm_256^3_box200_mesh512", ns, 200.0 * ns / 256, 2 * ns, -2.5, 0.5, 0.20,
                        max(1, 1000 * ns**3 // 256**3), 1.0, BASE_SEED + 2)
    if k == 3:
        ns = n_side or 64
        return make_box("adjoint_64^3_box50_mesh128", ns, 50.0 * ns / 64, 2 * ns, -2.5, 1.0, 0.0, 0, 0.05,
                        BASE_SEED + 3)
    raise ValueError(f"unknown workload config {k}")
