"""MI355X-native 3D Helmholtz solver hot path: FFT-in-z shifted-Laplacian
preconditioner with block-tridiagonal plane factors, FGMRES/BiCGStab drivers and
the receiver-misfit adjoint, behind the C ABI of include/hz.h (libhz.so)."""
from . import problem  # noqa: F401
from . import abi  # noqa: F401
from . import files  # noqa: F401


def Solver(*args, **kw):
    from .solver import Solver as _S
    return _S(*args, **kw)


def factor_many(*args, **kw):
    from .solver import factor_many as _f
    return _f(*args, **kw)


def fgmres_many(*args, **kw):
    from .solver import fgmres_many as _f
    return _f(*args, **kw)
