"""FP64 direct-sum reference of the differentiable forward, in PyTorch on the CPU.

Test infrastructure only. It restates the forward of cosmosim/adjoint.py with
no tree: the short-range force is the residual pair sum over every pair within
r_cut (minimum image), which is what the strict walk with monopole moments
computes (tree_pm_acceptance.hpp:62-66); the PM part is the CIC deposit,
Green's function -4 pi G exp(-k^2 r_s^2) / k^2, spectral gradient -i k_a with
the even-n Nyquist derivative zeroed, and CIC interpolation
(pm_solver.cpp:925-951, 2557-2705, 5900-5957), written with torch.fft. The
gradient then comes from torch.autograd, independently of the hand-written
adjoint kernels it checks.
"""
from __future__ import annotations

import math

import numpy as np
import torch

INV_SQRT_PI = 1.0 / math.sqrt(math.pi)


def _long_inv_r3(r, a):
    """L(r)/r^3, L(r) = erf(q) - 2 q exp(-q^2)/sqrt(pi), q = r/(2a); series below q = 1/8."""
    q = 0.5 * r / a
    u = q * q
    c = 4.0 * INV_SQRT_PI
    series = 0.125 / a**3 * c * (1 / 3 + u * (-1 / 5 + u * (1 / 14 + u * (-1 / 54 + u * (1 / 264 + u * (
        -1 / 1560 + u * (1 / 10800 - u / 85680)))))))
    rs = torch.where(q < 0.125, torch.ones_like(r), r)
    qs = 0.5 * rs / a
    closed = (torch.special.erf(qs) - 2.0 * INV_SQRT_PI * qs * torch.exp(-qs * qs)) / rs**3
    return torch.where(q < 0.125, series, closed)


def residual_inv_r3(r2, eps, a):
    soft = (r2 + eps * eps) ** -1.5
    return soft - _long_inv_r3(torch.sqrt(r2), a)


def _cic(x, L, n):
    g = torch.remainder(x, L) * (n / L)
    base = torch.floor(g)
    t = g - base
    i0 = base.long() % n
    return i0, (i0 + 1) % n, 1.0 - t, t


def cic_weights(pos, L, n):
    """(8, N) flat cell indices and weights of the CIC stencil."""
    ax = [_cic(pos[a], L, n) for a in range(3)]
    idx, w = [], []
    for i in range(2):
        for j in range(2):
            for k in range(2):
                idx.append((ax[0][i] * n + ax[1][j]) * n + ax[2][k])
                w.append(ax[0][2 + i] * ax[1][2 + j] * ax[2][2 + k])
    return torch.stack(idx), torch.stack(w)


def pm_accel(pos, mass, L, n, G, rs):
    idx, w = cic_weights(pos, L, n)
    rho = torch.zeros(n**3, dtype=pos.dtype).index_add(0, idx.reshape(-1), (w * mass[None, :]).reshape(-1))
    rho_k = torch.fft.rfftn(rho.reshape(n, n, n))
    m = torch.fft.fftfreq(n, d=1.0 / n).to(pos.dtype)
    m[n // 2] = n // 2  # Nyquist mode as +n/2
    mz = torch.arange(n // 2 + 1, dtype=pos.dtype)
    k1 = 2.0 * math.pi / L * m
    kz = 2.0 * math.pi / L * mz
    dk1 = k1.clone()
    dk1[n // 2] = 0.0
    dkz = kz.clone()
    dkz[n // 2] = 0.0
    KX, KY, KZ = torch.meshgrid(k1, k1, kz, indexing="ij")
    k2 = KX**2 + KY**2 + KZ**2
    vcell = (L / n) ** 3
    green = torch.where(k2 > 0, -4.0 * math.pi * G * torch.exp(-k2 * rs * rs) / torch.where(k2 > 0, k2, 1.0) /
                        vcell, torch.zeros_like(k2))
    phi = rho_k * green
    DX, DY, DZ = torch.meshgrid(dk1, dk1, dkz, indexing="ij")
    acc = []
    for d in (DX, DY, DZ):
        field = torch.fft.irfftn(-1j * d * phi, s=(n, n, n)).reshape(-1)
        acc.append((field[idx] * w).sum(0))
    return torch.stack(acc)


def sr_accel(pos, mass, L, G, rs, rcut, eps):
    d = pos[:, None, :] - pos[:, :, None]  # d[a, i, j] = x_j - x_i
    d = d - L * torch.round(d / L)
    r2 = (d * d).sum(0)
    inside = (r2 <= rcut * rcut) & (r2 > 0)
    r2s = torch.where(inside, r2, torch.ones_like(r2))
    f = torch.where(inside, residual_inv_r3(r2s, eps, rs), torch.zeros_like(r2))
    return G * (f[None, :, :] * mass[None, None, :] * d).sum(2)


def forward_loss(psi, lattice, mass, target, box, mesh, G, rs, rcut, eps, factors, vel_factor, loss_mesh):
    """Differentiable FP64 forward: Zel'dovich start, KDK steps, CIC density loss."""
    def accel(x):
        return pm_accel(x, mass, box, mesh, G, rs) + sr_accel(x, mass, box, G, rs, rcut, eps)

    x = torch.remainder(lattice + psi, box)
    v = vel_factor * psi
    g = accel(x)
    for f in factors:
        vh = f["first_hubble_drag"] * v + f["first_kick_factor"] * g
        x = torch.remainder(x + f["drift_factor"] * vh, box)
        g = accel(x)
        v = f["second_hubble_drag"] * vh + f["second_kick_factor"] * g
    idx, w = cic_weights(x, box, loss_mesh)
    n = x.shape[1]
    count = torch.zeros(loss_mesh**3, dtype=x.dtype).index_add(0, idx.reshape(-1), w.reshape(-1))
    delta = count * (loss_mesh**3 / n) - 1.0
    return ((delta - target.reshape(-1)) ** 2).sum()


def reference_gradient(prob, psi64_cpu, target_cpu):
    """d loss / d psi of the direct-sum model for a cosmosim.adjoint.KdkProblem."""
    psi = psi64_cpu.clone().requires_grad_(True)
    o = prob.options
    mesh_spacing = prob.box / prob.mesh
    rs = o.asmth_cells * mesh_spacing
    rcut = o.rcut_cells * mesh_spacing
    loss = forward_loss(psi, prob.lattice.cpu(), prob.mass.cpu(), target_cpu, prob.box, prob.mesh,
                        o.gravitational_constant, rs, rcut, o.softening_epsilon, prob.factors, prob.vel_factor,
                        prob.loss_mesh)
    (grad,) = torch.autograd.grad(loss, psi)
    return float(loss.detach()), grad


def reference_accel(prob, x_cpu):
    o = prob.options
    mesh_spacing = prob.box / prob.mesh
    rs = o.asmth_cells * mesh_spacing
    rcut = o.rcut_cells * mesh_spacing
    m = prob.mass.cpu()
    return (pm_accel(x_cpu, m, prob.box, prob.mesh, o.gravitational_constant, rs) +
            sr_accel(x_cpu, m, prob.box, o.gravitational_constant, rs, rcut, o.softening_epsilon))


__all__ = ["forward_loss", "reference_gradient", "reference_accel", "residual_inv_r3", "np"]
