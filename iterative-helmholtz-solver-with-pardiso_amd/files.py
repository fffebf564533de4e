"""Velocity-model files and the reference's solution / convergence text files
through the C ABI (host memory).

  velocity_from_file     SetSoundSpeed3DFromFile + HeteroSoundSpeed3DExtensionToPML
                         (functions.cpp:1458-1531, 2675-2716)
  write_solution         sol3D_N%d.dat (main.cpp:827-834), physical box in reducePML3D order
  read_solution          the same format read back
  write_convergence      convergence_*.dat lines "%d %e %e %lf" (FGMRES.cpp:646)
  convergence_file_name  the reference's log file name (FGMRES.cpp:50)
  runge_ratio            Runge's ratio of three nested solutions (functions.cpp:932-1134)
  beta3d_norms           Beta3D's shell norms of two nested solutions
                         (functions.cpp:1136-1229, 629-674)
  write_beta_norms       beta_3D_norm2_*.dat lines "%lf %lf" (functions.cpp:1295-1302)
  beta_norm_file_name    that file's name (functions.cpp:1295)
"""
import ctypes as C

import numpy as np

from .abi import check, lib, make_desc


def velocity_file_name(h):
    buf = C.create_string_buffer(256)
    check(lib().hz_velocity_file_name(float(h), buf, 256), "hz_velocity_file_name")
    return buf.value.decode()


def velocity_from_file(path, desc):
    """Speeds on the whole grid (N doubles, x fastest) from a float32 [y][x][z]
    physical-box file, extended into the layers."""
    c = np.empty(desc.N, dtype=np.float64)
    dd = make_desc(desc)
    check(lib().hz_velocity_from_file(str(path).encode(), C.byref(dd), c.ctypes.data_as(C.POINTER(C.c_double))),
          "hz_velocity_from_file")
    return c


def write_solution(path, desc, u):
    u = np.ascontiguousarray(np.asarray(u, dtype=np.complex128).reshape(-1))
    if u.size != desc.N:
        raise ValueError("solution has %d values, grid %d" % (u.size, desc.N))
    dd = make_desc(desc)
    check(lib().hz_write_solution(str(path).encode(), C.byref(dd), u.ctypes.data), "hz_write_solution")


def read_solution(path, desc):
    u = np.empty(desc.N, dtype=np.complex128)
    dd = make_desc(desc)
    check(lib().hz_read_solution(str(path).encode(), C.byref(dd), u.ctypes.data), "hz_read_solution")
    return u


def write_convergence(path, iters, res, relres, diff):
    it = np.ascontiguousarray(iters, dtype=np.intc)
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (res, relres, diff)]
    n = len(it)
    if any(len(a) != n for a in arrs):
        raise ValueError("convergence columns differ in length")
    P = C.POINTER(C.c_double)
    check(lib().hz_write_convergence(str(path).encode(), n, it.ctypes.data_as(C.POINTER(C.c_int)),
                                     *[a.ctypes.data_as(P) for a in arrs]), "hz_write_convergence")


def convergence_file_name(desc, length_x, nu, beta):
    buf = C.create_string_buffer(512)
    dd = make_desc(desc)
    check(lib().hz_convergence_file_name(C.byref(dd), float(length_x), float(nu), float(beta), buf, 512),
          "hz_convergence_file_name")
    return buf.value.decode()


def runge_ratio(coarse, mid, fine, n_coarse, n_mid, n_fine, h):
    """(f1, f2, f1 / f2): h * sum |u_fine - u| over the coarse and the middle box
    of three nested sol3D files; the order is log2(f1 / f2)."""
    out = (C.c_double * 3)()
    check(lib().hz_runge_ratio(str(coarse).encode(), str(mid).encode(), str(fine).encode(), int(n_coarse),
                               int(n_mid), int(n_fine), float(h), out), "hz_runge_ratio")
    return tuple(out)


def beta3d_norms(coarse, fine, grid, h, source):
    """(radius, norm) arrays of the grid - 3 shells 2h <= r <= e h, e = 2 .. grid - 2."""
    count = int(grid) - 3
    src = (C.c_double * 3)(*[float(v) for v in source])
    radius = np.empty(max(count, 0))
    norm = np.empty(max(count, 0))
    P = C.POINTER(C.c_double)
    check(lib().hz_beta3d_norms(str(coarse).encode(), str(fine).encode(), int(grid), float(h), src, count,
                                radius.ctypes.data_as(P), norm.ctypes.data_as(P)), "hz_beta3d_norms")
    return radius, norm


def write_beta_norms(path, radius, norm):
    r = np.ascontiguousarray(radius, dtype=np.float64)
    v = np.ascontiguousarray(norm, dtype=np.float64)
    if len(r) != len(v):
        raise ValueError("radius and norm differ in length")
    P = C.POINTER(C.c_double)
    check(lib().hz_write_beta_norms(str(path).encode(), len(r), r.ctypes.data_as(P), v.ctypes.data_as(P)),
          "hz_write_beta_norms")


def beta_norm_file_name(n, spg):
    buf = C.create_string_buffer(256)
    check(lib().hz_beta_norm_file_name(int(n), int(spg), buf, 256), "hz_beta_norm_file_name")
    return buf.value.decode()
