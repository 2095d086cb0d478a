"""Synthetic inputs of the solver workloads (device tensors, seeded).

* poisson_bumps: Dirichlet Poisson problem on a 256² Cartesian grid with
  spacing 1/256, right-hand side a sum of three Gaussian bumps (σ = 0.05,
  centres uniform in [0.2, 0.8]², amplitudes uniform in [0.5, 2]).
* magnetostatics_inductor: curl-curl problem on 192³ uniform cells (193³
  nodes) of the unit cube: relative permeability 1000 in a centred box core
  of 48 × 48 × 96 cells, 1 elsewhere; a 1 A/m² current circulating around z
  in the square coil annulus between half-widths 32 and 40 cells, 64 cells
  high, centred.
* uneven_poisson_batch: 512 problems with sides log-uniform in [32, 512],
  κ = exp(g), g a Gaussian random field (σ = 1, correlation 8 cells; disjoint
  windows of large periodic fields), N(0, 1) right-hand sides.
The learned-preconditioner problems are in preconditioner_training.
"""
import math

import numpy as np
import torch

MU0 = 4e-7 * math.pi


def poisson_bumps(n=256, seed=2024):
    rng = np.random.default_rng(seed)
    x = np.arange(n) / 256.0
    X, Y = np.meshgrid(x, x, indexing="ij")
    b = np.zeros((n, n))
    for _ in range(3):
        cx, cy = rng.uniform(0.2, 0.8, 2)
        amp = rng.uniform(0.5, 2.0)
        b += amp * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2 * 0.05 ** 2))
    b[0, :] = b[-1, :] = 0.0
    b[:, 0] = b[:, -1] = 0.0
    return x, x, b


def magnetostatics_inductor(cells=192, core=(48, 48, 96), coil=(32, 40, 64)):
    """(coordinates, cell μ [cells³] in H/m, cell current [cells³][3] in A/m²);
    cells are numbered x fastest."""
    n = cells + 1
    coords = np.arange(n) / cells
    c = cells / 2.0
    mid = np.arange(cells) + 0.5 - c  # cell centres relative to the centre, in cells
    dx = mid[None, None, :]  # x fastest: index [k][j][i]
    dy = mid[None, :, None]
    dz = mid[:, None, None]
    in_core = (np.abs(dx) < core[0] / 2) & (np.abs(dy) < core[1] / 2) & (np.abs(dz) < core[2] / 2)
    mu = np.where(in_core, 1000.0, 1.0) * MU0
    cheb = np.maximum(np.abs(dx), np.abs(dy))
    in_coil = (cheb >= coil[0]) & (cheb < coil[1]) & (np.abs(dz) < coil[2] / 2)
    jx = np.zeros((cells, cells, cells))
    jy = np.zeros((cells, cells, cells))
    east = in_coil & (np.abs(dx) >= np.abs(dy))
    north = in_coil & (np.abs(dy) > np.abs(dx))
    jy = np.where(east, np.sign(dx), 0.0)      # +y on the +x side, −y on the −x side
    jx = np.where(north, -np.sign(dy), 0.0)    # −x on the +y side, +x on the −y side
    current = np.stack([jx, jy, np.zeros_like(jx)], axis=-1).reshape(-1, 3)
    return [coords] * 3, mu.reshape(-1), current


def magnetostatics_nonlinear_inductor(cells=192, current=2.0e3):
    """The inductor of magnetostatics_inductor on unit-spaced nodes, its core
    a nonlinear material: (coordinates, cell μ, cell nonlinear-material flag,
    cell current scaled by `current`).  The B-H curve acts on the discrete
    induction (face fluxes), so unit spacing makes them Tesla-sized; with
    current 2000 the core reaches |B| ≈ 0.3, several segments into a
    typical curve."""
    coords, mu, j = magnetostatics_inductor(cells)
    flags = (mu > 1.5 * MU0).astype(np.float64)
    return [np.arange(cells + 1, dtype=np.float64)] * 3, mu, flags, j * current


def magnetostatics_2d_inductor(cells=2048, current=187.5):
    """A 2D cross-section of an inductor on unit-spaced nodes, for the A_z
    operator: a core of cells/4 × cells/2 in the middle (μ_r 1000, or the
    nonlinear material), two coil slabs cells/8 wide beside it carrying ±J_z
    = ±current.  (coordinates, cell μ, cell nonlinear flag, cell J_z), cells
    x fastest.  The induction scales with current × cells.  At 2048 cells,
    187.5 takes the nonlinear core to |B| ≈ 0.76 and Newton needs five
    steps; from 250 up, Newton with its halving line search stagnates near
    a relative residual of 1e-3."""
    mid = np.arange(cells) + 0.5 - cells / 2.0
    dx = mid[None, :]
    dy = mid[:, None]
    core = (np.abs(dx) < cells / 8) & (np.abs(dy) < cells / 4)
    mu = np.where(core, 1000.0, 1.0) * MU0
    slab = (np.abs(dx) >= cells / 6) & (np.abs(dx) < cells / 6 + cells / 8) & (np.abs(dy) < cells / 4)
    jz = np.where(slab, np.sign(dx) * current, 0.0)
    coords = [np.arange(cells + 1, dtype=np.float64)] * 2
    return coords, mu.reshape(-1), core.astype(np.float64).reshape(-1), jz.reshape(-1)


def gaussian_random_field(shape, sigma, corr, generator, device="cuda"):
    """One field of the given 2D shape with covariance σ² exp(−r²/(2ℓ²)),
    spectral synthesis on a periodic grid twice as large, cropped and
    rescaled to mean 0 / standard deviation σ."""
    n0, n1 = shape
    m0, m1 = 2 * n0, 2 * n1
    white = torch.randn((m0, m1), generator=generator, dtype=torch.float64, device=device)
    k0 = torch.fft.fftfreq(m0, device=device, dtype=torch.float64) * 2 * math.pi
    k1 = torch.fft.fftfreq(m1, device=device, dtype=torch.float64) * 2 * math.pi
    amp = torch.exp(-0.25 * (k0[:, None] ** 2 + k1[None, :] ** 2) * corr * corr)
    g = torch.fft.ifft2(torch.fft.fft2(white) * amp).real[:n0, :n1]
    g = (g - g.mean()) / g.std()
    return sigma * g


def periodic_random_fields(count, size, sigma, corr, generator, device="cuda"):
    """`count` periodic fields of size × size (one FFT size for all: the FFT
    library compiles a kernel per transform size), covariance
    σ² exp(−r²/(2ℓ²)), each rescaled to mean 0 / standard deviation σ."""
    white = torch.randn((count, size, size), generator=generator, dtype=torch.float64, device=device)
    k = torch.fft.fftfreq(size, device=device, dtype=torch.float64) * 2 * math.pi
    amp = torch.exp(-0.25 * (k[:, None] ** 2 + k[None, :] ** 2) * corr * corr)
    g = torch.fft.ifft2(torch.fft.fft2(white) * amp).real
    g = (g - g.mean(dim=(1, 2), keepdim=True)) / g.std(dim=(1, 2), keepdim=True)
    return sigma * g


def uneven_poisson_batch(count=512, lo=32, hi=512, sigma=1.0, corr=8.0, seed=4242, device="cuda", field=2048):
    """(sides, κ concatenated, rhs concatenated) in the layout of
    sil_batched_poisson2d (each problem n_p × n_p, x slowest).

    The coefficient fields are disjoint windows of a few large periodic
    Gaussian random fields (shelf-packed, largest first): windows far apart
    compared with the correlation length are independent samples of the same
    field, and one transform size serves every problem."""
    rng = np.random.default_rng(seed)
    sides = np.exp(rng.uniform(math.log(lo), math.log(hi), count)).round().astype(np.int32)
    sides = np.clip(sides, lo, hi)
    # shelf packing of the windows into field × field sheets
    place = [None] * count
    sheet, x, y, shelf = 0, 0, 0, 0
    for p in sorted(range(count), key=lambda q: -int(sides[q])):
        n = int(sides[p])
        if x + n > field:
            x, y, shelf = 0, y + shelf, 0
        if y + n > field:
            sheet, x, y, shelf = sheet + 1, 0, 0, 0
        place[p] = (sheet, x, y)
        x += n
        shelf = max(shelf, n)
    gen = torch.Generator(device=device).manual_seed(seed)
    fields = periodic_random_fields(sheet + 1, field, sigma, corr, gen, device)
    kappas, rhss = [], []
    for p in range(count):
        n = int(sides[p])
        s_, x0, y0 = place[p]
        kappas.append(torch.exp(fields[s_, y0:y0 + n, x0:x0 + n]).reshape(-1))
        b = torch.randn((n, n), generator=gen, dtype=torch.float64, device=device)
        b[0, :] = b[-1, :] = 0
        b[:, 0] = b[:, -1] = 0
        rhss.append(b.reshape(-1))
    return sides, torch.cat(kappas), torch.cat(rhss)
