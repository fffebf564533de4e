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
  runge_ratio         Runge, 3D branch (functions.cpp:1061-1122) with IntegralNorm3D
                      (functions.cpp:1348-1360)
  beta3d_norms        Beta3D (functions.cpp:1203-1217) with
                      compute_and_print_circle_norm_nopml (:1290-1302) and
                      check_norm_circle_3D_nopml (:629-674), one shell at a time
  beta_norm_text      "%lf %lf\n" lines (functions.cpp:1301)
  beta_norm_name      functions.cpp:1295
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


def restrict_box(uf, nf, nc):
    """Fine box (nf per side, x fastest) at the coarse nodes: fine node
    s (i + 1) - 1 per axis, s = (nf + 1) / (nc + 1) (functions.cpp:1102-1112)."""
    s = (nf + 1) // (nc + 1)
    idx = s * (np.arange(nc) + 1) - 1
    u3 = np.asarray(uf).reshape(nf, nf, nf)  # [k][j][i]
    return u3[np.ix_(idx, idx, idx)].reshape(-1)


def runge_ratio(u1, u2, u3, n1, n2, n3, h):
    f1 = np.sum(np.abs(restrict_box(u3, n3, n1) - np.asarray(u1))) * h
    f2 = np.sum(np.abs(restrict_box(u3, n3, n2) - np.asarray(u2))) * h
    return f1, f2, f1 / f2


def beta3d_norms(uc, uf, grid, h, source):
    n1, n2 = grid - 1, 2 * grid - 1
    orig = restrict_box(uf, n2, n1)  # x_orig: the fine solution at the coarse nodes
    sol = np.asarray(uc)             # x_sol: the coarse solution
    k, j, i = np.meshgrid(np.arange(n1), np.arange(n1), np.arange(n1), indexing="ij")
    x = (i.reshape(-1) + 1) * h - source[0]
    y = (j.reshape(-1) + 1) * h - source[1]
    z = (k.reshape(-1) + 1) * h - source[2]
    r = np.sqrt(x * x + y * y + z * z)
    r0 = 2 * h
    radius, norm = [], []
    for e in range(2, n1):
        m = (r >= r0) & (r <= e * h)
        with np.errstate(invalid="ignore", divide="ignore"):
            norm.append(np.linalg.norm(sol[m] - orig[m]) / np.linalg.norm(orig[m]))
        radius.append(e * h)
    return np.array(radius), np.array(norm)


def beta_norm_text(radius, norm):
    return "".join("%f %f\n" % (a, b) for a, b in zip(radius, norm))


def beta_norm_name(n, spg):
    return "beta_3D_norm2_Nx%dNy%d_SPG%d.dat" % (n, n, spg)
