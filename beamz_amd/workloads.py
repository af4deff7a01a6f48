"""Builders for the benchmark and parity workloads, expressed with the public API.

* ``waveguide_2d``: 2D TM straight Si waveguide in SiO2 (dx = lambda/20 at 1.55 um),
  20-cell cubic-graded PML (R = 1e-6) on all sides, Courant 0.99/sqrt(2),
  Gaussian-pulse point source, DFT line monitor. Default size 800 x 400 cells.
* ``soi_mmi_3d``: 3D six-component SOI 1x2 MMI (Si 220 nm on SiO2, air above),
  25 nm cells, seeded +-20% jitter of guide widths and body size, 12-cell PML
  (x/y from the reference profile, z as the documented extension), dipole
  Gaussian-pulse source, two x-normal DFT flux ports. Default 480 x 240 x 160.

Gaussian pulse convention: carrier f0 = c/lambda0; spectral sigma
df = c*dlambda/lambda0^2; time sigma 1/(2 pi df); centre 4 time sigmas.
"""
from __future__ import annotations

import numpy as np

from .const import LIGHT_SPEED, µm
from .design import Design, Material, PML, Rectangle
from .devices import GaussianSource, Monitor
from .devices.signals import gaussian_pulse
from .simulation.layered import layered_grid_3d, pml_sigma_factor

EPS_SI = 3.476 ** 2
EPS_OX = 1.444 ** 2


def pulse_signal(t, lam0=1.55 * µm, dlam=0.2 * µm):
    f0 = LIGHT_SPEED / lam0
    df = LIGHT_SPEED * dlam / lam0 ** 2
    st = 1.0 / (2 * np.pi * df)
    return gaussian_pulse(t, 1.0, 4 * st, st, f0, 0.0)


def edge_pmls(W, H, cells, dx, order=3.0, reflection=1e-6):
    s = pml_sigma_factor(cells, reflection, order)
    L = cells * dx
    return [PML("rect", (0, 0), (L, H), "left", polynomial_order=order, sigma_factor=s),
            PML("rect", (W - L, 0), (L, H), "right", polynomial_order=order, sigma_factor=s),
            PML("rect", (0, H - L), (W, L), "top", polynomial_order=order, sigma_factor=s),
            PML("rect", (0, 0), (W, L), "bottom", polynomial_order=order, sigma_factor=s)]


def waveguide_2d(nx=800, ny=400, steps=4000, npml=20, lam0=1.55 * µm, src_x=50, mon_x=700, n_freq=16):
    """Returns (design, dx, time array, monitor). Sizes in cells."""
    dx = lam0 / 20
    W, H = nx * dx, ny * dx
    d = Design(W, H, material=Material(EPS_OX), auto_pml=False)
    d.boundaries += edge_pmls(W, H, npml, dx)
    d += Rectangle(position=(0, H / 2 - 0.25 * µm), width=W, height=0.5 * µm, material=Material(EPS_SI))
    dt = 0.99 * dx / (LIGHT_SPEED * np.sqrt(2))
    t = np.arange(steps) * dt
    d += GaussianSource(position=(src_x * dx, H / 2), width=dx / 4, signal=pulse_signal(t, lam0))
    freqs = LIGHT_SPEED / (np.linspace(1.45, 1.65, n_freq) * µm)
    mon = Monitor(design=d, start=(min(mon_x, nx - 1) * dx, npml * dx), end=(min(mon_x, nx - 1) * dx, H - npml * dx),
                  dft_frequencies=freqs, name="line")
    d += mon
    return d, dx, t, mon


def binary_density(n, rng, sigma_px=3.0, fill=0.5):
    """Seeded Bernoulli(fill) pixel map, Gaussian-blurred (sigma in pixels), thresholded at 0.5."""
    from scipy.ndimage import gaussian_filter
    r = (rng.random((n, n)) < fill).astype(np.float64)
    return gaussian_filter(r, sigma_px, mode="nearest") > 0.5


_BLUR_POOL = None


def binary_densities(m, n, rng, sigma_px=3.0, fill=0.5):
    """m maps of binary_density at once: the same maps, and the same state of rng afterwards, as m calls in a
    row. The m x n x n uniforms are drawn in one call (the generator fills them in the same order), and the
    blur runs on slices of the stack in a thread pool (scipy.ndimage releases the GIL; 256 maps of 128 x 128:
    150 ms on one thread, 23 ms on eight)."""
    import os
    from concurrent.futures import ThreadPoolExecutor
    from scipy.ndimage import gaussian_filter
    global _BLUR_POOL
    r = (rng.random((m, n, n)) < fill).astype(np.float64)
    out = np.empty((m, n, n), dtype=bool)
    if m == 0:
        return out
    if _BLUR_POOL is None:
        _BLUR_POOL = ThreadPoolExecutor(max(1, min(8, os.cpu_count() or 1)))

    def blur(idx):
        out[idx] = gaussian_filter(r[idx], (0, sigma_px, sigma_px), mode="nearest") > 0.5

    list(_BLUR_POOL.map(blur, np.array_split(np.arange(m), min(m, 16))))
    return out


def batch_member_2d(n, rng, steps=2000, npml=10, dx=25e-9, n_freq=8, design_um=2.0):
    """One member of the batched 2D workload: an n x n cell domain (SiO2), a design_um x design_um
    binary Si/SiO2 design region of one cell per pixel at the centre between 0.5 um input and output
    Si guides, 10-cell PML, Gaussian-pulse point source in the input guide, input and output DFT line
    monitors across the guide. Returns (FDTD-ready mesh, time array, (monitor_in, monitor_out))."""
    from .simulation.meshing import RegularGrid
    # the grid has int(W / dx) cells (meshing.py:74-75 truncates); n*dx / dx rounds below n for some n
    W = H = (n + 1e-6) * dx
    npx = int(round(design_um * µm / dx))
    x0 = y0 = (n - npx) // 2
    d = Design(W, H, material=Material(EPS_OX), auto_pml=False)
    d.boundaries += edge_pmls(W, H, npml, dx)
    d += Rectangle(position=(0, H / 2 - 0.25 * µm), width=x0 * dx, height=0.5 * µm, material=Material(EPS_SI))
    d += Rectangle(position=((x0 + npx) * dx, H / 2 - 0.25 * µm), width=W - (x0 + npx) * dx, height=0.5 * µm,
                   material=Material(EPS_SI))
    dt = 0.99 * dx / (LIGHT_SPEED * np.sqrt(2))
    t = np.arange(steps) * dt
    d += GaussianSource(position=((npml + 10) * dx, H / 2), width=dx / 4, signal=pulse_signal(t))
    freqs = LIGHT_SPEED / (np.linspace(1.50, 1.60, n_freq) * µm)
    span = min(1.0 * µm, H / 2 - (npml + 1) * dx)
    m_in = Monitor(design=d, start=((npml + 20) * dx, H / 2 - span), end=((npml + 20) * dx, H / 2 + span),
                   dft_frequencies=freqs, record_fields=False, accumulate_power=False, name="in")
    xo = (n - npml - 10) * dx
    m_out = Monitor(design=d, start=(xo, H / 2 - span), end=(xo, H / 2 + span), dft_frequencies=freqs,
                    record_fields=False, accumulate_power=False, name="out")
    d += m_in
    d += m_out
    mesh = RegularGrid(d, dx)
    if mesh.permittivity.shape != (n, n):
        raise AssertionError(f"grid {mesh.permittivity.shape} != ({n}, {n})")
    eps = np.array(mesh.permittivity)
    dens = binary_density(npx, rng)
    eps[y0:y0 + npx, x0:x0 + npx] = np.where(dens, EPS_SI, EPS_OX)
    mesh = RegularGrid.from_arrays(d, eps, mesh.conductivity, dx)
    return mesh, t, (m_in, m_out)


def batch_2d(nsims=64, seed=0, nmin=256, nmax=1024, steps=2000, **kw):
    """The batched 2D workload: nsims members with sides drawn uniformly from {nmin..nmax} cells
    (side-uniform), each built by batch_member_2d with its own seeded density. Returns a list of
    (mesh, time, monitors)."""
    rng = np.random.default_rng(seed)
    sides = rng.integers(nmin, nmax + 1, size=nsims)
    return [batch_member_2d(int(n), rng, steps=steps, **kw) for n in sides]


def topo_bend_2d(n=600, design_px=120, steps=3000, npml=20, dx=20e-9, lam0=1.55 * µm):
    """The adjoint topology-optimisation workload: an n x n cell domain (SiO2), a 0.5 um Si input guide
    along +x at mid-height ending at the design box, a 0.5 um Si output guide along +y from the top of the
    design box to the top edge (a 90-degree bend, as in the reference's topology example), a
    design_px x design_px cell design box at the centre, 20-cell PML, Gaussian-pulse point source in the
    input guide and a horizontal DFT line monitor across the output guide at lam0.
    Returns (mesh, time, monitor, box, mode_weights)."""
    from .optimization.adjoint import mode_weights
    from .simulation.meshing import RegularGrid
    W = H = (n + 1e-6) * dx
    x0 = y0 = (n - design_px) // 2
    d = Design(W, H, material=Material(EPS_OX), auto_pml=False)
    d.boundaries += edge_pmls(W, H, npml, dx)
    d += Rectangle(position=(0, H / 2 - 0.25 * µm), width=x0 * dx, height=0.5 * µm, material=Material(EPS_SI))
    d += Rectangle(position=(W / 2 - 0.25 * µm, (y0 + design_px) * dx), width=0.5 * µm,
                   height=H - (y0 + design_px) * dx, material=Material(EPS_SI))
    dt = 0.99 * dx / (LIGHT_SPEED * np.sqrt(2))
    t = np.arange(steps) * dt
    d += GaussianSource(position=((npml + 20) * dx, H / 2), width=dx / 4, signal=pulse_signal(t, lam0))
    ym = (n - npml - 20) * dx
    mon = Monitor(design=d, start=(W / 2 - 1.0 * µm, ym), end=(W / 2 + 1.0 * µm, ym),
                  dft_frequencies=[LIGHT_SPEED / lam0], record_fields=False, accumulate_power=False, name="out")
    d += mon
    mesh = RegularGrid(d, dx)
    if mesh.permittivity.shape != (n, n):
        raise AssertionError(f"grid {mesh.permittivity.shape} != ({n}, {n})")
    c = mode_weights(mon, dx, dx, W / 2, 0.5 * µm, 3.476, 1.444, lam0)
    return mesh, t, mon, (x0, y0, design_px, design_px), c


def mmi_rectangles(W, H, rng, jitter=0.2, scale=1.0):
    """Input guide, 2.5 x 6 um body and two output guides (in-plane sizes times ``scale``), each size scaled by
    1 + U(-jitter, jitter)."""
    j = lambda: 1.0 + rng.uniform(-jitter, jitter)
    gw_in, gw_o1, gw_o2 = 0.5 * µm * scale * j(), 0.5 * µm * scale * j(), 0.5 * µm * scale * j()
    body_w, body_l = 2.5 * µm * scale * j(), 6.0 * µm * scale * j()
    x0 = (W - body_l) / 2
    yc = H / 2
    sep = body_w / 4 * j()
    rects = [(0.0, yc - gw_in / 2, x0, gw_in),
             (x0, yc - body_w / 2, body_l, body_w),
             (x0 + body_l, yc + sep - gw_o1 / 2, W - (x0 + body_l), gw_o1),
             (x0 + body_l, yc - sep - gw_o2 / 2, W - (x0 + body_l), gw_o2)]
    return rects, (yc + sep, yc - sep), x0 + body_l


def soi_mmi_3d(nx=480, ny=240, nz=160, steps=3000, npml=12, dx=25e-9, seed=0, port_freqs=8, z_pml=True, scale=1.0,
               record_ports=False):
    """Returns (mesh, design3d, time array, monitors). ``scale`` shrinks the in-plane MMI sizes and the port
    planes (not the 220 nm Si layer), for parity runs on grids small enough for the FP64 oracle;
    record_ports=True keeps the port planes' time series too (the DFT sums are always kept)."""
    rng = np.random.default_rng(seed)
    W, H, D = nx * dx, ny * dx, nz * dx
    z_si0 = D / 2 - 0.11 * µm
    z_si1 = z_si0 + 0.22 * µm
    rects, port_y, body_end = mmi_rectangles(W, H, rng, scale=scale)

    def layer(bg, with_mmi):
        d2 = Design(W, H, material=Material(bg), auto_pml=False)
        d2.boundaries += edge_pmls(W, H, npml, dx)
        if with_mmi:
            for (x, y, w, h) in rects:
                d2 += Rectangle(position=(x, y), width=w, height=h, material=Material(EPS_SI))
        return d2

    d3 = Design(W, H, depth=D, material=Material(1.0), auto_pml=False)
    mesh = layered_grid_3d(d3, [(0.0, z_si0, layer(EPS_OX, False)), (z_si0, z_si1, layer(1.0, True)),
                                (z_si1, D, layer(1.0, False))], dx,
                           z_pml_cells=npml if z_pml else 0, z_pml_sigma_factor=pml_sigma_factor(npml))
    dt = 0.99 * dx / (LIGHT_SPEED * np.sqrt(3))
    t = np.arange(steps) * dt
    zc = (z_si0 + z_si1) / 2
    d3 += GaussianSource(position=((npml + 20) * dx, H / 2, zc), width=dx, signal=pulse_signal(t))
    freqs = LIGHT_SPEED / (np.linspace(1.5, 1.6, port_freqs) * µm)
    mons = []
    xp = W - (npml + 16) * dx
    for q, yp in enumerate(port_y):
        hy, hz = 0.6 * µm * scale, 0.4 * µm * min(1.0, 2 * scale)
        m = Monitor(design=d3, plane_normal="x", plane_position=xp, start=(xp, yp - hy, zc - hz),
                    size=(2 * hy, 2 * hz), dft_frequencies=freqs, record_fields=record_ports,
                    accumulate_power=False, name=f"port{q + 1}")
        d3 += m
        mons.append(m)
    return mesh, d3, t, mons
