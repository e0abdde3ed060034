"""Synthetic initial states of the benchmark configurations.

Laser-foil (256 x 128 x 128 cells at dx = lambda/20, Mur walls, a0 = 3
Gaussian beam of waist 3 lambda and 10-period intensity FWHM injected on the
x = 20 plane, fully ionised 20 n_c foil 0.5 lambda thick at x = 128 cells,
27 markers per cell per species, electrons Maxwellian vth = 0.01 c, cold
protons) and the 32^3 periodic two-stream box.  Code units as the
reference: lengths in lambda, time in lambda/c, momenta gamma*v/c, critical
density rho = 4 pi^2 (source/misc_units.c:96-115).

Every random number of a marker state comes from RANMAR (the reference's
generator, source/misc_ranmar.c:78-146) seeded by one (ij, kl) pair per run
(mandor3_amd/rng.py): thermal momenta are Box-Muller deviates of its uniform
stream, random in-cell positions are its uniforms.  The parameters of scan
requests are drawn with numpy from the scan's own seed.
"""
import math

import numpy as np

from . import BC_OPEN, BC_PERIODIC, MARKER_DTYPE
from .rng import Ranmar

RHO_CRIT = 4 * math.pi ** 2
QM_PROTON = 1.0 / 1836.0


def courant_tau(h, frac=0.45):
    """tau = frac * tau_max, the negative-tau convention of
    source/setup/tag_mesh.c:59-60."""
    return frac / math.sqrt(sum(1.0 / (v * v) for v in h))


def gaussian_map(n, h, plane, a0, waist, focus):
    """Incident-field map of a y-polarised paraxial Gaussian beam for the
    one-plane soft source (E map: Ey^ at x = (I-1)h driving Hz on plane I;
    H map: -Hz^ at x = (I-1/2)h driving Ey on plane I-1), j, k in [5, n-5]."""
    I = plane
    k0 = 2 * math.pi
    w0 = waist
    xf, yf, zf = focus
    zr = math.pi * w0 * w0
    js = np.arange(5, n[1] - 5 + 1)
    ks = np.arange(5, n[2] - 5 + 1)
    jj, kk = np.meshgrid(js, ks, indexing="ij")

    def beam(x, y, z):
        xi = x - xf
        w = w0 * math.sqrt(1 + (xi / zr) ** 2)
        invR = xi / (xi * xi + zr * zr)
        gouy = math.atan2(xi, zr)
        r2 = (y - yf) ** 2 + (z - zf) ** 2
        amp = (w0 / w) * np.exp(-r2 / (w * w))
        th = -k0 * r2 * invR / 2 + gouy - k0 * (x - I * h[0])
        return amp * np.cos(th), -amp * np.sin(th)

    mapE = np.zeros(jj.shape + (4,))
    mapH = np.zeros(jj.shape + (4,))
    c, s = beam((I - 1) * h[0], (jj - 0.5) * h[1], kk * h[2])
    mapE[..., 0], mapE[..., 1] = c, s
    c, s = beam((I - 0.5) * h[0], (jj - 0.5) * h[1], kk * h[2])
    mapH[..., 2], mapH[..., 3] = -c, -s
    return np.ascontiguousarray(np.stack([mapE.reshape(-1, 4), mapH.reshape(-1, 4)])).reshape(-1)


def laser_params(a0, plane, h, focus_x, fwhm_periods, peak_time):
    """(A, T0, delay, dt, phi0, forward, plane) of the soft source; dt from
    the intensity FWHM as source/scpic/em_mirror.c:305-308 does."""
    dt = fwhm_periods * math.sqrt(2.0 / math.log(2))
    return (2 * math.pi * a0, peak_time, focus_x - plane * h[0], dt, 0.0, 1, plane)


def _quiet_cells(cells, sub, h, rng):
    """Markers at the sub-cell centres of every listed cell (the placement
    of source/setup/tag_plasma.c:92-110)."""
    sx, sy, sz = sub
    off = np.stack(np.meshgrid((np.arange(sx) + 0.5) / sx, (np.arange(sy) + 0.5) / sy,
                               (np.arange(sz) + 0.5) / sz, indexing="ij"), -1).reshape(-1, 3)
    ppc = off.shape[0]
    pos = (cells[:, None, :] + off[None, :, :]).reshape(-1, 3) * np.asarray(h)
    return pos, ppc


def laser_foil(scale=1.0, seed=12345, ppc_sub=(3, 3, 3)):
    """The 256x128x128 laser-foil state.  scale < 1 shrinks the transverse
    box (tests); markers are generated for the whole foil slab."""
    n = (256, max(16, int(round(128 * scale))), max(16, int(round(128 * scale))))
    h = (0.05, 0.05, 0.05)
    L = tuple(n[a] * h[a] for a in range(3))
    tau = courant_tau(h)
    rng = Ranmar(seed)
    i0, i1 = 123, 133                      # 10 cells = 0.5 lambda centred at 128
    ii, jj, kk = np.meshgrid(np.arange(i0, i1), np.arange(n[1]), np.arange(n[2]), indexing="ij")
    cells = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], -1).astype(np.float64)
    pos, ppc = _quiet_cells(cells, ppc_sub, h, rng)
    N = pos.shape[0]
    M = np.zeros(2 * N, dtype=MARKER_DTYPE)
    w = 20.0 * RHO_CRIT / ppc
    for s, (q, qm, vth) in enumerate(((-w, -1.0, 0.01), (w, QM_PROTON, 0.0))):
        sl = slice(s * N, (s + 1) * N)
        M["x"][sl], M["y"][sl], M["z"][sl] = pos[:, 0], pos[:, 1], pos[:, 2]
        if vth > 0:
            u = rng.normal(3 * N, vth).reshape(N, 3)
            M["vx"][sl], M["vy"][sl], M["vz"][sl] = u[:, 0], u[:, 1], u[:, 2]
        M["rho"][sl] = q
        M["qDivM"][sl] = qm
    M["id"] = np.arange(1, 2 * N + 1)
    plane = 20
    focus_x = 128 * h[0]
    lp = laser_params(3.0, plane, h, focus_x, 10.0, focus_x - plane * h[0] + 2.0)
    lmap = gaussian_map(n, h, plane, 3.0, 3.0, (focus_x, L[1] / 2, L[2] / 2))
    return dict(name="laser-foil 256x128x128 h=lambda/20, 20 n_c foil 0.5 lambda, "
                     "27 ppc x 2 species, a0=3 w0=3 lambda, Mur",
                n=n, h=h, L=L, tau=tau, bc=(BC_OPEN,) * 6, markers=M,
                laser=(lp, lmap), ppc=2 * ppc, plasma_cells=cells.shape[0])


def large_foil(seed=4000, planes_per_chunk=8):
    """The 512x256x256 single-GPU state, h = lambda/24, Mur walls: 16
    markers per cell per species (4x2x2 sub-lattice) in every cell one cell
    or more from the walls (510 x 254 x 254 cells, 1 052 949 120 markers:
    the full-box fill used for throughput), weighted by the physical
    profile along x: a 40 n_c foil 1 lambda thick from x = 256 cells, a
    2 lambda exponential preplasma in front of it (scale length 0.5 lambda)
    and a 0.1 n_c floor elsewhere; electrons Maxwellian vth = 0.01 c, cold
    protons; a0 = 10, waist 4 lambda, 15-period beam injected on x = 20.
    The markers are not built here: 'chunks' yields them a few i planes at
    a time (about 1 GB of records each), in id order."""
    n = (512, 256, 256)
    h = (1.0 / 24,) * 3
    L = tuple(n[a] * h[a] for a in range(3))
    sub = _SUB[16]
    ppc = sub[0] * sub[1] * sub[2]
    i_lo, i_hi = 1, n[0] - 2                   # the open walls delete markers within a cell
    jk = (n[1] - 2) * (n[2] - 2)
    foil0, foil1 = 256, 256 + 24
    pre0 = foil0 - 48

    def density(i):
        """n / n_c of cell plane i (profile at the cell centre)."""
        x = (i + 0.5) * h[0]
        if foil0 <= i < foil1:
            return 40.0
        if pre0 <= i < foil0:
            return max(40.0 * np.exp(-(foil0 * h[0] - x) / 0.5), 0.1)
        return 0.1

    def chunks():
        """Records of planes [i0, i0 + planes_per_chunk), alternately in two
        reused buffers (a chunk stays valid while the next one is built):
        positions on the sub-lattice by broadcasting per axis (the order of
        _quiet_cells: cells i, j, k, then sub-points, last axis fastest),
        electrons then protons."""
        nj, nk = n[1] - 2, n[2] - 2
        sx, sy, sz = sub
        bufs = [np.zeros(2 * planes_per_chunk * jk * ppc, dtype=MARKER_DTYPE) for _ in range(2)]
        oj = (np.arange(1, n[1] - 1)[:, None] + (np.arange(sy) + 0.5) / sy) * h[1]      # (nj, sy)
        ok = (np.arange(1, n[2] - 1)[:, None] + (np.arange(sz) + 0.5) / sz) * h[2]      # (nk, sz)
        nid = 1
        rng = Ranmar(seed)                        # one stream, continued chunk by chunk
        for c, i0 in enumerate(range(i_lo, i_hi + 1, planes_per_chunk)):
            i1 = min(i0 + planes_per_chunk, i_hi + 1)
            P = i1 - i0
            N = P * jk * ppc
            M = bufs[c % 2][:2 * N]
            oi = (np.arange(i0, i1)[:, None] + (np.arange(sx) + 0.5) / sx) * h[0]      # (P, sx)
            w = (np.array([density(i) for i in range(i0, i1)]) * (RHO_CRIT / ppc))
            for s, (sign, qm, vth) in enumerate(((-1.0, -1.0, 0.01), (1.0, QM_PROTON, 0.0))):
                V = M[s * N:(s + 1) * N].reshape(P, nj, nk, sx, sy, sz)
                V["x"] = oi[:, None, None, :, None, None]
                V["y"] = oj[None, :, None, None, :, None]
                V["z"] = ok[None, None, :, None, None, :]
                V["rho"] = (sign * w)[:, None, None, None, None, None]
                V["qDivM"] = qm
                Vf = M[s * N:(s + 1) * N]
                if vth > 0:
                    for f in ("vx", "vy", "vz"):
                        Vf[f] = rng.normal(N, vth)
                else:
                    for f in ("vx", "vy", "vz"):
                        Vf[f] = 0.0
            M["id"] = np.arange(nid, nid + 2 * N, dtype=np.int32)
            nid += 2 * N
            yield M

    cells_filled = (i_hi - i_lo + 1) * jk
    plane = 20
    focus_x = foil0 * h[0]
    lp = laser_params(10.0, plane, h, focus_x, 15.0, focus_x - plane * h[0] + 2.0)
    lmap = gaussian_map(n, h, plane, 10.0, 4.0, (focus_x, L[1] / 2, L[2] / 2))
    return dict(name="large 512x256x256 h=lambda/24, full-box fill 16 ppc x 2 species, 40 n_c foil "
                     "1 lambda + 2 lambda preplasma, a0=10 w0=4 lambda, Mur",
                n=n, h=h, L=L, tau=courant_tau(h), bc=(BC_OPEN,) * 6, markers=None, chunks=chunks,
                density=density,
                n_markers=2 * ppc * cells_filled, laser=(lp, lmap), ppc=2 * ppc,
                plasma_cells=cells_filled)


def two_stream_box(n=32, seed=7):
    """32^3 periodic box, h = 0.05, electrons (vth 0.05 + beams +-0.2 c
    along x) and cold ions, 8 markers per cell each at uniform random
    in-cell positions, n_e = n_i = 1 n_c.  Each ion sits where an electron
    sits, so the box starts neutral (rho = 0) with E = 0 and div E = rho
    holds from the first step: the Gauss residual measures only what the
    scheme adds."""
    h = (0.05, 0.05, 0.05)
    nn = (n, n, n)
    L = tuple(nn[a] * h[a] for a in range(3))
    rng = Ranmar(seed)
    ii, jj, kk = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    cells = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], -1).astype(np.float64)
    N = cells.shape[0] * 8
    M = np.zeros(2 * N, dtype=MARKER_DTYPE)
    w = RHO_CRIT / 8
    pos = (np.repeat(cells, 8, axis=0) + rng.uniform(3 * N).reshape(N, 3)) * 0.05
    u = rng.normal(3 * N, 0.05).reshape(N, 3)
    u[:, 0] += np.where(np.arange(N) % 2 == 0, 0.2, -0.2)
    for s, (q, qm) in enumerate(((-w, -1.0), (w, QM_PROTON))):
        sl = slice(s * N, (s + 1) * N)
        M["x"][sl], M["y"][sl], M["z"][sl] = pos[:, 0], pos[:, 1], pos[:, 2]
        if s == 0:
            M["vx"][sl], M["vy"][sl], M["vz"][sl] = u[:, 0], u[:, 1], u[:, 2]
        M["rho"][sl] = q
        M["qDivM"][sl] = qm
    M["id"] = np.arange(1, 2 * N + 1)
    return dict(name="two-stream 32^3 periodic h=0.05, 8 ppc x 2 species",
                n=nn, h=h, L=L, tau=courant_tau(h), bc=(BC_PERIODIC,) * 6, markers=M,
                laser=None, ppc=16, plasma_cells=cells.shape[0])


# markers per cell per species -> in-cell sub-lattice
_SUB = {4: (2, 2, 1), 8: (2, 2, 2), 16: (4, 2, 2), 27: (3, 3, 3), 32: (4, 4, 2)}


def foil_request(n, ppc, a0, density, thickness, seed=0, steps=100):
    """One laser-foil scan point: grid n (h = lambda/20, Mur walls), a fully
    ionised foil of `density` n_c and `thickness` lambda centred at 5/8 of
    the box along x, electrons (vth 0.01 c) and cold protons at `ppc`
    markers per cell each, a y-polarised Gaussian beam of amplitude a0
    focused on the foil's front face and injected near the low-x wall."""
    n = tuple(int(v) for v in n)
    h = (0.05, 0.05, 0.05)
    L = tuple(n[a] * h[a] for a in range(3))
    rng = Ranmar(seed)
    xc = int(round(0.625 * n[0]))
    half = max(1, int(round(thickness / h[0] / 2)))
    i0, i1 = max(xc - half, 8), min(xc + half, n[0] - 3)
    ii, jj, kk = np.meshgrid(np.arange(i0, i1), np.arange(2, n[1] - 2), np.arange(2, n[2] - 2),
                             indexing="ij")
    cells = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], -1).astype(np.float64)
    pos, p = _quiet_cells(cells, _SUB[ppc], h, rng)
    N = pos.shape[0]
    M = np.zeros(2 * N, dtype=MARKER_DTYPE)
    w = density * RHO_CRIT / p
    for s, (q, qm, vth) in enumerate(((-w, -1.0, 0.01), (w, QM_PROTON, 0.0))):
        sl = slice(s * N, (s + 1) * N)
        M["x"][sl], M["y"][sl], M["z"][sl] = pos[:, 0], pos[:, 1], pos[:, 2]
        if vth > 0:
            u = rng.normal(3 * N, vth).reshape(N, 3)
            M["vx"][sl], M["vy"][sl], M["vz"][sl] = u[:, 0], u[:, 1], u[:, 2]
        M["rho"][sl] = q
        M["qDivM"][sl] = qm
    M["id"] = np.arange(1, 2 * N + 1)
    plane = max(6, n[0] // 8)
    focus_x = i0 * h[0]
    waist = 0.25 * min(L[1], L[2])
    lp = laser_params(a0, plane, h, focus_x, 3.0, focus_x - plane * h[0] + 1.5)
    lmap = gaussian_map(n, h, plane, a0, waist, (focus_x, L[1] / 2, L[2] / 2))
    return dict(name="foil %dx%dx%d ppc %d a0 %.2f n %.1f d %.2f" % (n + (ppc, a0, density, thickness)),
                n=n, h=h, L=L, tau=courant_tau(h), bc=(BC_OPEN,) * 6, markers=M,
                laser=(lp, lmap), ppc=2 * p, plasma_cells=cells.shape[0], steps=int(steps))


def scan_request(rng, large_fraction=0.0, steps=(500, 3000)):
    """Parameters of one ensemble / serving request: grid uniformly from
    {32^3, 64x32x32, 64^3, 128x64x64}, markers per cell from {4, 8, 16, 32},
    a0 log-uniform in [0.5, 5], density uniform in [5, 50] n_c, thickness
    uniform in [0.2, 1.0] lambda, steps uniform in `steps`; with probability
    large_fraction a 256x128x128 run at 16 markers per cell."""
    grids = [(32, 32, 32), (64, 32, 32), (64, 64, 64), (128, 64, 64)]
    if rng.random() < large_fraction:
        n, ppc = (256, 128, 128), 16
    else:
        n, ppc = grids[rng.integers(len(grids))], int(rng.choice([4, 8, 16, 32]))
    return dict(n=n, ppc=ppc, a0=float(np.exp(rng.uniform(np.log(0.5), np.log(5.0)))),
                density=float(rng.uniform(5, 50)), thickness=float(rng.uniform(0.2, 1.0)),
                seed=int(rng.integers(1 << 30)), steps=int(rng.integers(steps[0], steps[1] + 1)))
