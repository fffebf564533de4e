"""CPU restatement of the reference's velocity-model ingestion and text output
formats, in plain numpy / Python loops.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the library's
file entry points; the product package never imports it.

  sound_from_box      SetSoundSpeed3DFromFile placement (functions.cpp:1496-1501)
                      + HeteroSoundSpeed3DExtensionToPML (functions.cpp:2675-2716),
                      z layers filled from the nearest physical z node (the
                      reference leaves them at zero)
  reduce_pml          reducePML3D order (functions.cpp:238-263)
  solution_text       "%18.16lf %18.16lf\\n" lines (main.cpp:830-831)
  convergence_text    "%d %e %e %lf\\n" lines (FGMRES.cpp:646)
  convergence_name    FGMRES.cpp:50
"""
import numpy as np


def layers(desc, a):
    return int(desc.pml[a]) + int(desc.spg[a])


def sound_from_box(desc, box_yxz):
    """box_yxz: float32 array (n2, n1, n3) in file order [y][x][z]."""
    nx, ny, nz = (int(v) for v in desc.n)
    lx, ly, lz = layers(desc, 0), layers(desc, 1), layers(desc, 2)
    n1, n2, n3 = nx - 2 * lx, ny - 2 * ly, nz - 2 * lz
    flat = np.asarray(box_yxz, dtype=np.float32).reshape(-1)
    s = np.zeros(nx * ny * nz)
    # functions.cpp:1496-1501: sound3D[k Nx Ny + j Nx + i] = data[(j-py) n3 n1 + (i-px) n3 + (k-sz)]
    for j in range(ly, ny - ly):
        for i in range(lx, nx - lx):
            for k in range(lz, nz - lz):
                s[k * nx * ny + j * nx + i] = flat[(j - ly) * n3 * n1 + (i - lx) * n3 + (k - lz)]
    for k in list(range(0, lz)) + list(range(nz - lz, nz)):
        ks = lz if k < lz else nz - lz - 1
        for j in range(ly, ny - ly):
            for i in range(lx, nx - lx):
                s[k * nx * ny + j * nx + i] = s[ks * nx * ny + j * nx + i]
    size2d = nx * ny
    # functions.cpp:2682-2698: x layers, every k and every j
    for k in range(nz):
        for j in range(ny):
            for i in range(lx):
                s[i + nx * j + k * size2d] = s[lx + nx * j + k * size2d]
            for i in range(nx - lx, nx):
                s[i + nx * j + k * size2d] = s[nx - lx - 1 + nx * j + k * size2d]
    # functions.cpp:2700-2715: y layers, every k and every i
    for k in range(nz):
        for i in range(nx):
            for j in range(ly):
                s[i + nx * j + k * size2d] = s[i + nx * ly + k * size2d]
            for j in range(ny - ly, ny):
                s[i + nx * j + k * size2d] = s[i + nx * (ny - ly - 1) + k * size2d]
    return s


def reduce_pml(desc, u):
    nx, ny, nz = (int(v) for v in desc.n)
    lx, ly, lz = layers(desc, 0), layers(desc, 1), layers(desc, 2)
    out = []
    for l in range(nx * ny * nz):
        i, j, k = l % nx, (l // nx) % ny, l // (nx * ny)
        if lx <= i < nx - lx and ly <= j < ny - ly and lz <= k < nz - lz:
            out.append(u[l])
    return np.array(out)


def solution_text(desc, u):
    return "".join("%18.16f %18.16f\n" % (v.real, v.imag) for v in reduce_pml(desc, u))


def convergence_text(iters, res, relres, diff):
    return "".join("%d %e %e %f\n" % (a, b, c, d) for a, b, c, d in zip(iters, res, relres, diff))


def convergence_name(nx_phys, pml_x, length_x, nu, sponge_width, beta):
    return "convergence_N%d_PML%d_Lx%d_FREQ%d_SPG%6.f_BETA%5.3f_FGMRES.dat" % (
        nx_phys, pml_x, int(length_x), int(nu), sponge_width, beta)
