"""Velocity-model files and the reference's solution / convergence text files
through the C ABI (host memory).

  velocity_from_file     SetSoundSpeed3DFromFile + HeteroSoundSpeed3DExtensionToPML
                         (functions.cpp:1458-1531, 2675-2716)
  write_solution         sol3D_N%d.dat (main.cpp:827-834), physical box in reducePML3D order
  read_solution          the same format read back
  write_convergence      convergence_*.dat lines "%d %e %e %lf" (FGMRES.cpp:646)
  convergence_file_name  the reference's log file name (FGMRES.cpp:50)
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
