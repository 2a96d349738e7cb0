"""Direct C-ABI calls with explicit leading dimensions, for tests of the
seam's padding and argument contracts."""
import ctypes as C

import numpy as np


def padded(rows, cols, ld, dtype=np.float64, fill=np.nan):
    """A column-major rows x cols view with leading dimension ld inside a
    buffer whose padding rows hold `fill`."""
    buf = np.full((ld, max(cols, 1)), fill, dtype=dtype, order="F")
    return buf, buf[:rows, :cols]


def addr(a):
    return a.ctypes.data if a.size else None


def cplx(z):
    from csharp_linearalgebra_amd._lib import Complex

    z = complex(z)
    return Complex(z.real, z.imag)
