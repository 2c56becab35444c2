"""Synthetic grating workloads (structures, sources, media) for the five
benchmark shapes of the RCWA solve path.

Every structure is rasterised on the host at cell centres,
x = (ix + 0.5) * Lx / Nx, the convention of rcwa3d/src/main.cpp:51-56, and
stored as er/ur grids [layer][iy][ix] (rcwa3d/include/rcwa.hpp:56).
Lengths are in nm; angles in radians at the API boundary.  All randomness
comes from numpy default_rng(seed).

The reference uses the e^{+j w t} convention (rcwa.cpp:129-133), so a loss
is Im(eps) < 0; the +Im values below are fed verbatim to every solver.
"""
import math
import numpy as np


class Problem:
    """One solve: a layered device, one or more (pte, ptm) polarisations at
    one wavelength/angle, and the two half-space media."""

    def __init__(self, er, ur, thick, Lx, Ly, lam0, theta, phi, pols, hx, hy,
                 er_ref=1.0, ur_ref=1.0, er_trn=2.25, ur_trn=1.0, name=""):
        self.er = np.ascontiguousarray(er, dtype=np.complex128)
        self.ur = np.ascontiguousarray(ur, dtype=np.complex128)
        self.thick = np.ascontiguousarray(thick, dtype=np.float64)
        self.Lx, self.Ly = float(Lx), float(Ly)
        self.lam0, self.theta, self.phi = float(lam0), float(theta), float(phi)
        self.pols = [(float(a), float(b)) for a, b in pols]
        self.hx, self.hy = int(hx), int(hy)
        self.er_ref, self.ur_ref = complex(er_ref), complex(ur_ref)
        self.er_trn, self.ur_trn = complex(er_trn), complex(ur_trn)
        self.name = name

    @property
    def L(self):
        return self.er.shape[0]

    @property
    def N(self):
        return (2 * self.hx + 1) * (2 * self.hy + 1)

    def np_args(self, pol=0):
        pte, ptm = self.pols[pol]
        return (self.er, self.ur, self.thick, self.Lx, self.Ly, self.lam0, self.theta,
                self.phi, pte, ptm, self.hx, self.hy, self.er_ref, self.ur_ref,
                self.er_trn, self.ur_trn)


# shape kinds of rcwa_rasterize (include/rcwa_amd.h)
SHAPE_ELLIPSE, SHAPE_RECT, SHAPE_CIRCLE = 0, 1, 2


def _centres(n, L):
    return (np.arange(n) + 0.5) * L / n


def circle_hole_grid(nx, ny, Lx, Ly, r, eps_in, eps_out):
    x = _centres(nx, Lx)[None, :]
    y = _centres(ny, Ly)[:, None]
    inside = (x - Lx / 2) ** 2 + (y - Ly / 2) ** 2 <= r * r
    return np.where(inside, eps_in, eps_out).astype(np.complex128)


def c1_single(grid=256, h=5):
    """Square lattice 500x500 nm, one 200 nm layer of eps=15+0.1i with a
    centred r=150 nm air hole, air above, eps=2.25 below, lambda=600 nm,
    normal incidence, TE."""
    er = circle_hole_grid(grid, grid, 500.0, 500.0, 150.0, 1.0, 15.0 + 0.1j)[None]
    ur = np.ones_like(er)
    return Problem(er, ur, [200.0], 500.0, 500.0, 600.0, 0.0, 0.0, [(1.0, 0.0)],
                   h, h, 1.0, 1.0, 2.25, 1.0, name="c1")


def c2_materials(lams, nmat=5, seed=2):
    """Seeded two-term Lorentz dispersion per material, mapped into
    Re eps in [1.5, 20] and Im eps in [0, 5] over 250-800 nm."""
    rng = np.random.default_rng(seed)
    lams = np.asarray(lams, dtype=np.float64)
    out = np.zeros((nmat, lams.size), np.complex128)
    for k in range(nmat):
        lc = rng.uniform(200.0, 900.0, 2)       # resonance wavelengths
        g = rng.uniform(40.0, 200.0, 2)         # widths
        f = rng.uniform(0.2, 1.0, 2)            # strengths
        lo = rng.uniform(1.5, 8.0)              # background Re eps
        span = rng.uniform(2.0, 20.0 - lo)      # Re excursion
        imax = rng.uniform(0.5, 5.0)            # peak Im eps
        w = 2 * np.pi / lams
        z = np.zeros_like(lams, dtype=np.complex128)
        for j in range(2):
            wj = 2 * np.pi / lc[j]
            gj = 2 * np.pi * g[j] / lc[j] ** 2
            z += f[j] * wj * wj / (wj * wj - w * w - 1j * gj * w)
        re = z.real - z.real.min()
        re = re / max(re.max(), 1e-30)
        im = np.abs(z.imag) / max(np.abs(z.imag).max(), 1e-30)
        out[k] = (lo + span * re) + 1j * (imax * im)
    return out


def c2_halfwidths(period=120.0, cd=40.0, height=100.0, nslice=20, sidewall_deg=85.0):
    """Half-widths of the trapezoid line's slices (layer 0 on top)."""
    dz = height / nslice
    cot = 1.0 / math.tan(math.radians(sidewall_deg))
    return [(cd + 2.0 * ((k + 0.5) * dz - height / 2) * cot) / 2 for k in range(nslice)]


def c2_profile(grid=512, period=120.0, cd=40.0, height=100.0, nslice=20, sidewall_deg=85.0):
    """Trapezoid line, y-invariant, cut into nslice slabs (layer 0 on top).
    Returns a boolean mask [nslice][grid] along x."""
    x = _centres(grid, period)
    hw = c2_halfwidths(period, cd, height, nslice, sidewall_deg)
    masks = np.zeros((nslice, grid), bool)
    for k in range(nslice):
        masks[k] = np.abs(x - period / 2) <= hw[k]
    return masks, np.full(nslice, height / nslice)


def c2_sweep(nlam=200, grid=512, h=10, nslice=20, lam_lo=250.0, lam_hi=800.0, seed=2,
             theta_deg=65.0, indices=None):
    """Problems of the 200-wavelength sweep (or the given wavelength
    indices), each carrying TE and TM.  All problems share one all-ones
    permeability array."""
    lams = np.linspace(lam_lo, lam_hi, nlam)
    mats = c2_materials(lams, 5, seed)
    masks, thick = c2_profile(grid, nslice=nslice)
    ur = np.ones((nslice, grid, grid), np.complex128)
    probs = []
    for i in (range(nlam) if indices is None else indices):
        er = np.ones((nslice, grid, grid), np.complex128)
        for k in range(nslice):
            er[k][:, masks[k]] = mats[(k * 5) // nslice, i]
        probs.append(Problem(er, ur, thick, 120.0, 120.0, lams[i], math.radians(theta_deg), 0.0,
                             [(1.0, 0.0), (0.0, 1.0)], h, h, 1.0, 1.0, 2.25, 1.0,
                             name="c2"))
    return probs


def c2_structure(nlam=200, grid=512, nslice=20, lam_lo=250.0, lam_hi=800.0, seed=2):
    """The line grating of c2_sweep as material indices for a sweep plan:
    material 0 is air, 1..5 the five dispersive materials.  Returns
    (material uint8 [nslice][grid][grid], thick, lams, eps [nlam][6])."""
    lams = np.linspace(lam_lo, lam_hi, nlam)
    mats = c2_materials(lams, 5, seed)
    masks, thick = c2_profile(grid, nslice=nslice)
    material = np.zeros((nslice, grid, grid), np.uint8)
    for k in range(nslice):
        material[k][:, masks[k]] = 1 + (k * 5) // nslice
    eps = np.concatenate([np.ones((nlam, 1), np.complex128), mats.T], axis=1)
    return material, thick, lams, eps


def random_shapes_layer(rng, grid, Lx, Ly, shapes=None, layer=0):
    """1-3 ellipses/rectangles with eps Re in [1,16], Im in [0,2].  If shapes
    is a list, the layer's (layer, kind, cx, cy, ax, ay, eps) entries are
    appended to it and its background is returned with the grid."""
    x = _centres(grid, Lx)[None, :]
    y = _centres(grid, Ly)[:, None]
    eb = rng.uniform(1, 16) + 1j * rng.uniform(0, 2)
    g = np.full((grid, grid), eb, np.complex128)
    for _ in range(int(rng.integers(1, 4))):
        cx, cy = rng.uniform(0, Lx), rng.uniform(0, Ly)
        ax, ay = rng.uniform(0.05, 0.35) * Lx, rng.uniform(0.05, 0.35) * Ly
        es = rng.uniform(1, 16) + 1j * rng.uniform(0, 2)
        if rng.integers(0, 2) == 0:
            inside = ((x - cx) / ax) ** 2 + ((y - cy) / ay) ** 2 <= 1.0
            kind = SHAPE_ELLIPSE
        else:
            inside = (np.abs(x - cx) <= ax) & (np.abs(y - cy) <= ay)
            kind = SHAPE_RECT
        g = np.where(inside, es, g)
        if shapes is not None:
            shapes.append((layer, kind, cx, cy, ax, ay, es))
    return g if shapes is None else (g, eb)


def min_abs_kz(hx, hy, Lx, Ly, lam0, theta, phi, media):
    """Smallest |k_z| over harmonics in the given media (Wood-anomaly guard)."""
    kxi = math.sin(theta) * math.cos(phi)
    kyi = math.sin(theta) * math.sin(phi)
    kx = kxi - np.arange(-hx, hx + 1) * lam0 / Lx
    ky = kyi - np.arange(-hy, hy + 1) * lam0 / Ly
    k2 = kx[None, :] ** 2 + ky[:, None] ** 2
    return min(np.abs(np.sqrt(np.complex128(e) - k2)).min() for e in media)


def c4_stream(count=2000, seed=4, grid=256, rate_hz=2000.0, hs=(3, 5, 7, 10, 12),
              max_layers=30, period=500.0):
    """Seeded Poisson-arrival request stream.  Returns (arrival_s, Problem)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    while len(out) < count:
        t += rng.exponential(1.0 / rate_hz)
        h = int(hs[rng.integers(0, len(hs))])
        L = int(rng.integers(1, max_layers + 1))
        lam = rng.uniform(250.0, 800.0)
        th = math.radians(rng.uniform(0.0, 70.0))
        if min_abs_kz(h, h, period, period, lam, th, 0.0, (1.0, 2.25)) < 1e-6:
            continue
        er = np.stack([random_shapes_layer(rng, grid, period, period) for _ in range(L)])
        thick = rng.uniform(20.0, 200.0, L)
        out.append((t, Problem(er, np.ones_like(er), thick, period, period, lam, th, 0.0,
                               [(1.0, 0.0)], h, h, 1.0, 1.0, 2.25, 1.0, name="c4")))
    return out


def c1_shapes():
    """The C1 hole as a shape list: (Lx, Ly, background[1], shapes)."""
    return 500.0, 500.0, [15.0 + 0.1j], [(0, SHAPE_CIRCLE, 500.0 / 2, 500.0 / 2, 150.0, 0.0, 1.0)]


def c2_shapes(lam_index, nlam=200, nslice=20, lam_lo=250.0, lam_hi=800.0, seed=2):
    """The C2 line at one wavelength of the sweep as a shape list: each slice
    a line (rectangle of infinite height) of its material over air."""
    mats = c2_materials(np.linspace(lam_lo, lam_hi, nlam), 5, seed)
    hw = c2_halfwidths(nslice=nslice)
    shapes = [(k, SHAPE_RECT, 120.0 / 2, 0.0, hw[k], math.inf, mats[(k * 5) // nslice, lam_index])
              for k in range(nslice)]
    return 120.0, 120.0, [1.0] * nslice, shapes


def c5_shapes(nlayer=50):
    """The C5 holes as a shape list: per layer the centre hole and the four
    corner images, radius 60 -> 45 nm, over the layer's eps."""
    Lx, Ly = 150.0, 260.0
    bg, shapes = [], []
    for k in range(nlayer):
        r = 60.0 - 15.0 * k / (nlayer - 1)
        bg.append(2.1 if k % 2 == 0 else 4.0 + 0.05j)
        for cx, cy in [(Lx / 2, Ly / 2), (0.0, 0.0), (0.0, Ly), (Lx, 0.0), (Lx, Ly)]:
            shapes.append((k, SHAPE_CIRCLE, cx, cy, r, 0.0, 1.0))
    return Lx, Ly, bg, shapes


def c5_large(grid=1024, h=20, nlayer=50):
    """Hexagonal-equivalent rectangular cell 150x260 nm (two holes per cell),
    alternating eps 2.1 / 4.0+0.05i, 40 nm layers, hole radius 60 -> 45 nm."""
    Lx, Ly = 150.0, 260.0
    x = _centres(grid, Lx)[None, :]
    y = _centres(grid, Ly)[:, None]
    er = np.empty((nlayer, grid, grid), np.complex128)
    for k in range(nlayer):
        r = 60.0 - 15.0 * k / (nlayer - 1)
        eps = 2.1 if k % 2 == 0 else 4.0 + 0.05j
        d1 = (x - Lx / 2) ** 2 + (y - Ly / 2) ** 2
        d2 = np.minimum.reduce([(x - cx) ** 2 + (y - cy) ** 2
                                for cx in (0.0, Lx) for cy in (0.0, Ly)])
        inside = (d1 <= r * r) | (d2 <= r * r)
        er[k] = np.where(inside, 1.0, eps)
    return Problem(er, np.ones_like(er), np.full(nlayer, 40.0), Lx, Ly, 400.0,
                   math.radians(30.0), 0.0, [(1.0, 0.0)], h, h, 1.0, 1.0, 2.25, 1.0,
                   name="c5")


def random_small(rng, hx=1, hy=1, L=2, grid=8, lossless=False, uniform_mu=True):
    """Tiny random stack for parity tests."""
    Lx, Ly = rng.uniform(300, 700), rng.uniform(300, 700)
    er = np.empty((L, grid, grid), np.complex128)
    ur = np.ones((L, grid, grid), np.complex128)
    for l in range(L):
        im = 0.0 if lossless else rng.uniform(-1, 1)
        base = rng.uniform(1, 12) + 1j * im
        er[l] = base
        m = rng.random((grid, grid)) < 0.4
        er[l][m] = rng.uniform(1, 12) + 1j * (0.0 if lossless else rng.uniform(-1, 1))
        if not uniform_mu:
            ur[l] = 1.0 + rng.uniform(0, 0.5) * (rng.random((grid, grid)) < 0.3)
    thick = rng.uniform(20, 300, L)
    lam = rng.uniform(400, 900)
    th = rng.uniform(0, 1.0)
    ph = rng.uniform(0, 1.0)
    erf = rng.uniform(1, 3) if lossless else rng.uniform(1, 3) - 1j * rng.uniform(0, 0.2)
    return Problem(er, ur, thick, Lx, Ly, lam, th, ph, [(1.0, 0.0), (0.0, 1.0), (0.6, 0.8)],
                   hx, hy, 1.0, 1.0, erf, 1.0, name="small")
