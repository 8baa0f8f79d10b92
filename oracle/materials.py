"""ORACLE -- test infrastructure, not product code.

CPU restatement of the material preparation in ``Assembler.assemble_freq_matrix``
(src/emerge/_emerge/physics/microwave/assembly/assembler.py:439-469; the same lines in
``assemble_scattering_matrix``, :819-868, and ``assemble_eig_matrix``, :1119-1142), which
runs before the volume assembly:

* ``er' = er * (1 - 1j*tand) - 1j*cond / (W0*EPS0)`` on the (3,3,T) complex128 tensors
  (:453), with ``W0 = 2*pi*f`` (:423) and ``EPS0`` of const.py:23, evaluated as NumPy
  evaluates that expression;
* ``is_frequency_dependent``: some component with ``0 < cond < mw_3d_peclim`` (:455-457);
* the conductor tets, loop for loop (:461-469): ``cond[0,0,t] > mw_3d_peclim`` or
  ``> mw_3d_surfimplim`` (settings.py:35, :38: 1e8 and 1e4). NumPy compares complex
  numbers lexicographically (real part, then imaginary part), and so does the loop.

Pins (tests/test_materials_oracle.py): the textbook complex permittivity
``eps'(1 - j tan d) - j sigma/(w eps0)`` for real isotropic data; lossless data unchanged
bit for bit; a complex conductivity alone gives ``(Im s - j Re s)/(w eps0)``; anisotropic
complex loss tangents against Python's own complex arithmetic; the conductor and
frequency-dependence rules on a table of boundary cases written out by hand.
"""
from __future__ import annotations

import numpy as np

EPS0 = 8.854187818814e-12  # const.py:23
PECLIM = 1e8               # settings.py:35 (mw_3d_peclim)
SURFIMPLIM = 1e4           # settings.py:38 (mw_3d_surfimplim)


def complex_gt(a: complex, b: float) -> bool:
    """``a > b`` as NumPy orders complex numbers: real part first, then imaginary part."""
    a = complex(a)
    return a.real > b or (a.real == b and a.imag > 0.0)


def complex_lt(a: complex, b: float) -> bool:
    a = complex(a)
    return a.real < b or (a.real == b and a.imag < 0.0)


def prepare_materials(er, tand, cond, frequency, peclim=PECLIM, surfimplim=SURFIMPLIM):
    """(er', conductor_tets int64, is_frequency_dependent) of (3,3,T) tensors."""
    er = np.asarray(er, dtype=np.complex128)
    tand = np.asarray(tand, dtype=np.complex128)
    cond = np.asarray(cond, dtype=np.complex128)
    W0 = 2 * np.pi * frequency
    out = er * (1 - 1j * tand) - 1j * cond / (W0 * EPS0)
    freq_dep = bool(np.any((cond > 0) & (cond < peclim)))
    conductor = [t for t in range(er.shape[2])
                 if complex_gt(cond[0, 0, t], peclim) or complex_gt(cond[0, 0, t], surfimplim)]
    return out, np.array(conductor, dtype=np.int64), freq_dep
