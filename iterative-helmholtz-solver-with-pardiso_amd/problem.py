"""Problem descriptors, grid geometry, media and sources (host-side input synthesis).

The descriptor mirrors ``hz_desc`` of include/hz.h.  Two operator modes:

* ``MODE_REF`` -- the reference's own path: x/y stretched-coordinate PML, a
  periodic z direction handled spectrally, a Gaussian-tapered z sponge, and the
  implicit true operator ``A = L0 - diag(deltaL)`` (FGMRES.cpp:10,
  functions.cpp:5554-5575).
* ``MODE_S7`` -- the 7-point PML stencil on every face (beta3D,
  functions.cpp:2242-2303) with the FFT-in-z shifted-Laplacian preconditioner
  on the z-averaged background medium.

Layout: ``l = i + nx*j + nx*ny*k`` (x fastest), node ``i`` at ``(i+1)*h``
(functions.cpp:225-230, 4655-4661).
"""
from dataclasses import dataclass, field, replace
import math

import numpy as np

MODE_REF = 0
MODE_S7 = 1

KZ_REFERENCE = 0   # FGMRES.cpp:182-190 map (k < nz/2-1 -> k, else nz-k)
KZ_SPECTRAL = 1    # corrected map min(k, nz-k)
KZ_FD = 2          # symbol of the periodic 3-point second difference

C0_REFERENCE = 0   # SetSoundSpeed2D running sum (functions.cpp:1716-1728)
C0_HARMONIC = 1    # per-column harmonic mean over the physical z range


@dataclass
class HzDesc:
    n: tuple
    h: tuple
    l: tuple
    pml: tuple = (0, 0, 0)
    spg: tuple = (0, 0, 0)
    omega: float = 1.0
    beta: float = 0.5
    pml_c: float = 1000.0
    mode: int = MODE_S7
    kz_symbol: int = KZ_FD
    c0_rule: int = C0_HARMONIC
    c_bg: float = 0.0
    skip_ratio: int = 0      # SAVE_MEM plane skipping (FGMRES.cpp:196-206): 0 = off
    skip_k0: float = 0.0     # wavenumber of the skip test (0: omega / c_bg or over the slowest c0)
    # extra host-side facts (not part of the ABI struct)
    meta: dict = field(default_factory=dict)

    @property
    def N(self):
        return int(self.n[0]) * int(self.n[1]) * int(self.n[2])

    @property
    def shape_zyx(self):
        return (int(self.n[2]), int(self.n[1]), int(self.n[0]))

    def replace(self, **kw):
        return replace(self, **kw)

    def phys_bounds(self):
        """[lo, hi) per axis of the physical box (layers removed)."""
        return [(self.pml[a] + self.spg[a], self.n[a] - self.pml[a] - self.spg[a]) for a in range(3)]

    def phys_mask(self):
        (x0, x1), (y0, y1), (z0, z1) = self.phys_bounds()
        m = np.zeros(self.shape_zyx, dtype=bool)
        m[z0:z1, y0:y1, x0:x1] = True
        return m.reshape(-1)


# ---------------------------------------------------------------------------
# reference geometry (main.cpp:120-211, default build without SOURCE_SHIFT)
def ref_geometry(n_nopml=24, pml=8, spg=6, nz_nopml=None, length=(3200.0, 3200.0, 3200.0),
                 nu=4.0, beta=0.5, c_z=1280.0, homo=True, kz_symbol=KZ_REFERENCE,
                 c0_rule=C0_REFERENCE):
    """Descriptor of the reference's grid: x/y PML of ``pml`` points, z sponge of
    ``spg`` points, domain lengths enlarged by the layers, h = l/n, w = 2 pi nu."""
    nz_nopml = n_nopml if nz_nopml is None else nz_nopml
    nx = n_nopml + 2 * pml
    nz = nz_nopml + 2 * spg
    lx = length[0] + 2.0 * pml * length[0] / n_nopml
    ly = length[1] + 2.0 * pml * length[1] / n_nopml
    lz = length[2] + 2.0 * spg * length[2] / nz_nopml
    d = HzDesc(n=(nx, nx, nz), h=(lx / nx, ly / nx, lz / nz), l=(lx, ly, lz),
               pml=(pml, pml, 0), spg=(0, 0, spg), omega=2.0 * math.pi * nu, beta=beta,
               pml_c=1000.0, mode=MODE_REF, kz_symbol=kz_symbol, c0_rule=c0_rule,
               c_bg=(c_z if homo else 0.0))
    d.meta.update(dict(c_z=c_z, homo=homo, nu=nu))
    return d


def ref_sound_speed(desc):
    """SetSoundSpeed3D + HeteroSoundSpeed3DExtensionToPML (functions.cpp:1362-1430,
    2675-2716): c_z everywhere (HOMO) or c = c1 x + c2 y + c3 with c(0,0) = 600,
    c(0, Ly) = 1500, c(Lx, Ly) = 1800 on the physical columns, copied outward."""
    nx, ny, nz = desc.n
    px, py = desc.pml[0], desc.pml[1]
    if desc.meta.get("homo", True):
        return np.full(desc.N, float(desc.meta.get("c_z", 1280.0)))
    lx, ly = desc.l[0], desc.l[1]
    hx, hy = desc.h[0], desc.h[1]
    c3 = 600.0
    c2 = (1500.0 - c3) / ly
    c1 = (1800.0 - c3 - c2 * ly) / lx
    c2d = np.zeros((ny, nx))
    i = np.arange(nx)
    j = np.arange(ny)
    X = (i[None, :] + 1.0 - px) * hx
    Y = (j[:, None] + 1.0 - py) * hy
    c2d[:, :] = c1 * X + c2 * Y + c3
    c2d = _extend_xy(c2d, px, py)
    return np.broadcast_to(c2d[None], (nz, ny, nx)).reshape(-1).copy()


def _extend_xy(a, px, py):
    ny, nx = a.shape
    a = a.copy()
    if px:
        a[:, :px] = a[:, px:px + 1]
        a[:, nx - px:] = a[:, nx - px - 1:nx - px]
    if py:
        a[:py, :] = a[py:py + 1, :]
        a[ny - py:, :] = a[ny - py - 1:ny - py, :]
    return a


def source_index(desc):
    """Centre node of the reference (main.cpp:204-211): (n+1)//2 - 1 per axis."""
    return tuple((int(desc.n[a]) + 1) // 2 - 1 for a in range(3))


def flat_index(desc, i, j, k):
    return int(i) + int(desc.n[0]) * (int(j) + int(desc.n[1]) * int(k))


def point_source(desc, idx=None, amplitude=None):
    """Point source; the reference sets f = -1/(hx hy hz) at the source node
    (F3D_ex_complex, functions.cpp:4564-4578)."""
    if idx is None:
        idx = source_index(desc)
    if amplitude is None:
        amplitude = -1.0 / (desc.h[0] * desc.h[1] * desc.h[2])
    f = np.zeros(desc.N, np.complex128)
    f[flat_index(desc, *idx)] = amplitude
    return f


def green_function(desc, idx=None, c=None):
    """e^{ikr}/(4 pi r) at every node with r -> r + hx at the source
    (u_ex_complex_sound3D, functions.cpp:4530-4562)."""
    if idx is None:
        idx = source_index(desc)
    if c is None:
        c = desc.meta.get("c_z", 1.0)
    nx, ny, nz = desc.n
    hx, hy, hz = desc.h
    k = np.arange(nz)[:, None, None]
    j = np.arange(ny)[None, :, None]
    i = np.arange(nx)[None, None, :]
    X = (i - idx[0]) * hx
    Y = (j - idx[1]) * hy
    Z = (k - idx[2]) * hz
    r = np.sqrt(X * X + Y * Y + Z * Z)
    r = np.where(r == 0.0, hx, r)
    kw = desc.omega / c
    return (np.exp(1j * kw * r) / (4.0 * math.pi * r)).reshape(-1)


# ---------------------------------------------------------------------------
# unit-cube 7-point configurations
def cube_geometry(n, npml, omega, beta=0.5, pml_c=1000.0, kz_symbol=KZ_FD, c0_rule=C0_HARMONIC):
    """Unit cube, ``n`` nodes per axis including ``npml`` PML nodes per face,
    h = 1/n, node i at (i+1) h."""
    n3 = (n, n, n) if np.isscalar(n) else tuple(n)
    p3 = (npml, npml, npml) if np.isscalar(npml) else tuple(npml)
    h = tuple(1.0 / v for v in n3)
    l = tuple(v * hh for v, hh in zip(n3, h))
    return HzDesc(n=n3, h=h, l=l, pml=p3, spg=(0, 0, 0), omega=float(omega), beta=beta,
                  pml_c=pml_c, mode=MODE_S7, kz_symbol=kz_symbol, c0_rule=c0_rule, c_bg=0.0)


def omega_for_ppw(c_min, h, ppw=10.0):
    """Angular frequency giving ``ppw`` points per minimum wavelength."""
    return 2.0 * math.pi * c_min / (ppw * h)


# ---------------------------------------------------------------------------
# seeded Gaussian random field (SplitMix64 -> Box-Muller -> Gaussian filter)
_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, count):
    """``count`` SplitMix64 outputs of stream ``seed`` (vectorised)."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (np.arange(1, count + 1, dtype=np.uint64) * _GOLD)
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed, count):
    z = splitmix64(seed, count)
    return ((z >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def white_noise(seed, count):
    m = (count + 1) // 2
    u = uniform01(seed, 2 * m)
    u1, u2 = u[0::2], u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.empty(2 * m)
    z[0::2] = r * np.cos(2.0 * math.pi * u2)
    z[1::2] = r * np.sin(2.0 * math.pi * u2)
    return z[:count]


def gaussian_random_field(shape_zyx, seed, corr_len, h):
    """Zero-mean unit-variance field with covariance exp(-r^2 / (2 l^2)):
    white noise filtered by exp(-|k|^2 l^2 / 4) with a periodic FFT, then
    normalised empirically."""
    nz, ny, nx = shape_zyx
    w = white_noise(seed, nx * ny * nz).reshape(nz, ny, nx)
    W = np.fft.rfftn(w)
    kz = 2.0 * math.pi * np.fft.fftfreq(nz, d=h[2])[:, None, None]
    ky = 2.0 * math.pi * np.fft.fftfreq(ny, d=h[1])[None, :, None]
    kx = 2.0 * math.pi * np.fft.rfftfreq(nx, d=h[0])[None, None, :]
    W *= np.exp(-(kx * kx + ky * ky + kz * kz) * corr_len * corr_len / 4.0)
    g = np.fft.irfftn(W, s=(nz, ny, nx), axes=(0, 1, 2))
    g -= g.mean()
    g /= g.std()
    return g.reshape(-1)


def random_velocity(desc, seed, corr_len=0.1, amp=0.5, cmin=0.7, cmax=1.5):
    """c = clip(1 + amp*g, cmin, cmax) over the whole grid."""
    g = gaussian_random_field(desc.shape_zyx, seed, corr_len, desc.h)
    return np.clip(1.0 + amp * g, cmin, cmax)


def gaussian_blobs(desc, centres, amps, width):
    """c = 1 + sum_b a_b exp(-|x - x_b|^2 / (2 w^2)) at node coordinates."""
    nx, ny, nz = desc.n
    hx, hy, hz = desc.h
    Z = ((np.arange(nz) + 1) * hz)[:, None, None]
    Y = ((np.arange(ny) + 1) * hy)[None, :, None]
    X = ((np.arange(nx) + 1) * hx)[None, None, :]
    c = np.ones((nz, ny, nx))
    for (cx, cy, cz), a in zip(centres, amps):
        c += a * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2) / (2.0 * width * width))
    return c.reshape(-1)


def gaussian_smooth(field, desc, sigma):
    """Periodic Gaussian smoothing (std ``sigma`` in length units)."""
    nz, ny, nx = desc.shape_zyx
    F = np.fft.rfftn(field.reshape(nz, ny, nx))
    kz = 2.0 * math.pi * np.fft.fftfreq(nz, d=desc.h[2])[:, None, None]
    ky = 2.0 * math.pi * np.fft.fftfreq(ny, d=desc.h[1])[None, :, None]
    kx = 2.0 * math.pi * np.fft.rfftfreq(nx, d=desc.h[0])[None, None, :]
    F *= np.exp(-(kx * kx + ky * ky + kz * kz) * sigma * sigma / 2.0)
    return np.fft.irfftn(F, s=(nz, ny, nx), axes=(0, 1, 2)).reshape(-1)
