"""Device-resident volume part of ``Assembler.assemble_freq_matrix``.

Follows src/emerge/_emerge/physics/microwave/assembly/assembler.py:423-494 (the same
steps in ``assemble_scattering_matrix``, :808-868; ``assemble_eig_matrix``, :1115-1153,
converts K and M separately instead of combining them):

1. ``W0 = 2*pi*f``, ``K0 = W0 / C0`` (:423-424, const.py:20);
2. the material preparation ``er(1 - j tand) - j cond/(W0 EPS0)``, the frequency
   dependence and the conductor tets (:439-469) -- ``em_material_prep``;
3. the element matrices without the conductor tets (:479-485), cached as the reference's
   ``cached_matrices`` and reused while nothing is frequency dependent (:471-478);
4. ``cached_cscmap.to_csc(Evec - Bvec * K0**2)`` (:494) -- while the element matrices
   are reused, ``K_csc - K0**2 M_csc`` from CSC sums of K and M kept with them
   (``FreqCsc``: the same matrix up to the order of each entry's sum); with
   frequency-dependent materials the element kernel forms ``E - K0**2 B`` itself and one
   gather converts it (the same matrix up to rounding: ``-K0**2`` is folded into the mass
   metric tensor); with caching off, the fused CSC combine of K and M. When a
   frequency-dependent call brings the same er, ur, tand, cond arrays as the previous one
   (a lossy medium swept in frequency), ``er' = er (1 - j tand) - j cond/(W0 EPS0)`` is
   affine in 1/W0: ``A = K - K0**2 M(er (1 - j tand)) + j K0**2/(W0 EPS0) M(cond)``, three
   CSC arrays assembled once and one streaming pass per frequency (the same matrix up to
   rounding; the reference reassembles every frequency).

Everything stays on the GPU; the result is the CSC value array of ``A`` with the map's
``indptr``/``indices``. The cached matrices and CSC map belong to the conductor set they
were built for (the reference reuses its cached map whatever the conductor set; here a
new set gets its own plan and map). The boundary-condition steps that follow in the
reference (PEC, Robin, ports) are in ``emerge_amd.pec``, ``emerge_amd.robin`` and
``emerge_amd.ports``.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .assembly import CscMap, FreqCsc, Material, csc_freq_combine, ref_structure
from .curlcurl import _device, plan_for
from .materials import PECLIM, SURFIMPLIM, _dev_tensor, prepare_materials

C0 = 299792458  # const.py:20
EPS0 = 8.854187818814e-12  # const.py:23


class _LossyCache:
    """The inputs of the last frequency-dependent call (device copies of er, ur, tand,
    cond) and, once a later call brings the same arrays, the CSC values of K, of
    M(er (1 - j tand)) and of M(cond) in its plan's pattern."""

    def __init__(self, inputs, plan, cscmap, conductor_tets):
        self.inputs = inputs
        self.plan = plan
        self.cscmap = cscmap
        self.conductor_tets = conductor_tets
        self.K = self.Ma = self.Mc = None

    def matches(self, er, ur, tand, cond) -> bool:
        """Equal arrays (device tensors; ur may still be on the host, it is compared last
        and uploaded only then)."""
        def same(a, b):
            return a.shape == b.shape and bool(torch.equal(a, b))
        e, u, t, c = self.inputs
        return same(er, e) and same(tand, t) and same(cond, c) and \
            same(_dev_tensor(ur, e.device), u)

def _prepared_kind(er, tand, cond) -> int:
    """Structure of er(1 - j tand) - j cond/(w eps0): diagonal when er and cond are (an
    off-diagonal er of 0 stays 0 whatever tand is), scalar when every input's three
    diagonal entries are equal. Checked where each input lives (host or device)."""
    s = [ref_structure(x if isinstance(x, torch.Tensor) else np.asarray(x))
         for x in (er, tand, cond)]
    if not (s[0][0] and s[2][0]):
        return _lib.EM_MAT_FULL
    if all(d for _, d in s):
        return _lib.EM_MAT_SCALAR
    return _lib.EM_MAT_DIAG


# Every 20x20 block the element kernels write equals its transpose bit for bit, whatever
# the tensors: the reference computes each element matrix for i <= j only and mirrors it
# (curlcurl.py:357-371), and the kernels do the same (both entries of a mirrored pair come
# from one canonical pair, with the same coefficients and the same arithmetic). So every
# CSC conversion here takes the mirrored-read column gather.
_BLOCKS_SYMMETRIC = True


@dataclass
class FreqMatrix:
    """``A = K - K0^2 M`` in CSC form on the device."""
    data: torch.Tensor          # CSC values, complex128 (fp64) or complex64 (mixed)
    cscmap: CscMap              # indptr / indices via cscmap.export
    conductor_tets: np.ndarray  # int64, ascending
    is_frequency_dependent: bool
    k0: float

    def to_scipy(self):
        from .pec import to_scipy
        ex = self.cscmap.export("indptr", "indices")
        return to_scipy(ex["indptr"], ex["indices"], self.data, self.cscmap.N)


class VolumeAssembler:
    """Holds the reference assembler's caches for one field: the element matrices of the
    last conductor set (``cached_matrices``) and each plan's CSC map (``cached_cscmap``).
    ``assemblies`` counts the element-matrix launches (reuse leaves it unchanged)."""

    def __init__(self, field, precision: str = "fp64", peclim: float = PECLIM,
                 surfimplim: float = SURFIMPLIM):
        self.field = field
        self.precision = precision
        self.peclim = peclim
        self.surfimplim = surfimplim
        self.cached_matrices = None  # (plan, K, M)
        self._map = None  # (plan, CscMap)
        self._freq = None  # (cached K, FreqCsc of the cached K and M)
        self._lossy = None  # _LossyCache of the last frequency-dependent call
        self.assemblies = 0

    def assemble_freq_matrix(self, er, ur, tand, cond, frequency: float,
                             cache_matrices: bool = True,
                             frequency_dependent: bool = False) -> FreqMatrix:
        """``er``, ``ur``, ``tand``, ``cond``: complex (3, 3, T) tensors as the reference
        fills them from its materials (:440-450); ``frequency_dependent`` stands for the
        materials' own ``frequency_dependent`` flags (:433-436)."""
        W0 = 2 * np.pi * frequency
        K0 = W0 / C0
        dev = False
        if cache_matrices:
            # er, tand, cond on the device (uploaded once, reused below)
            er, tand, cond = (_dev_tensor(x, _device()) for x in (er, tand, cond))
            dev = True
            lc = self._lossy
            if lc is not None and lc.matches(er, ur, tand, cond):
                # the same arrays as the last frequency-dependent call: er' is affine in
                # 1/W0, A = K - K0^2 M(er (1 - j tand)) + j K0^2/(W0 EPS0) M(cond)
                if lc.K is None:
                    self._build_lossy(lc, frequency)
                data = csc_freq_combine(lc.K, lc.Ma, K0, C=lc.Mc, g=K0 ** 2 / (W0 * EPS0))
                return FreqMatrix(data, lc.cscmap, lc.conductor_tets, True, K0)
        plan, er_d, conductor_tets, fd = self._prepare_er(er, tand, cond, frequency)
        fd = fd or frequency_dependent
        c = self.cached_matrices
        if fd:
            er_m, ur_m = self._descriptors(plan, er_d, er, ur, tand, cond)
            # frequency-dependent materials: K and M cannot be reused, so A = K - K0^2 M is
            # formed inside the element kernel (one COO array written and gathered instead
            # of two) and converted once. K and M are not cached: the reference keeps this
            # call's E and B, which only a later call with frequency-independent materials
            # (built from other materials) would reuse; such a call assembles afresh here.
            # The inputs are kept instead: a later call with the same arrays (a lossy
            # medium at another frequency) takes the affine path above.
            A = plan.assemble_a(er_m, ur_m, K0, self.precision)
            self.assemblies += 1
            self.cached_matrices = None
            self._freq = None
            cm = self._cscmap(plan)
            if dev:
                # a later call with the same arrays takes the affine path
                inputs = (er, _dev_tensor(ur, er.device), tand, cond)
                self._lossy = _LossyCache(tuple(x.clone() for x in inputs), plan, cm,
                                          conductor_tets)
            return FreqMatrix(cm.to_csc(A, block_symmetric=_BLOCKS_SYMMETRIC), cm,
                              conductor_tets, fd, K0)
        self._lossy = None
        if cache_matrices and c is not None and c[0] is plan:
            K, M = c[1], c[2]
        else:
            er_m, ur_m = self._descriptors(plan, er_d, er, ur, tand, cond)
            K, M = plan.assemble_km(er_m, ur_m, self.precision)
            self.assemblies += 1
            self.cached_matrices = (plan, K, M)
        cm = self._cscmap(plan)
        if cache_matrices:
            # while K and M are reused, so are their CSC sums: each frequency is then one
            # streaming pass A = K_csc - K0^2 M_csc instead of a gather over the COO values
            if self._freq is None or self._freq[0] is not K:
                self._freq = (K, FreqCsc(cm, K, M, block_symmetric=_BLOCKS_SYMMETRIC))
            data = self._freq[1].A(K0)
        else:
            data = cm.to_csc_combine(K, M, K0, block_symmetric=_BLOCKS_SYMMETRIC)
        return FreqMatrix(data, cm, conductor_tets, fd, K0)

    def assemble_scattering_matrix(self, er, ur, tand, cond, frequency: float,
                                   cache_matrices: bool = True,
                                   frequency_dependent: bool = False) -> FreqMatrix:
        """The volume step of ``assemble_scattering_matrix`` (assembler.py:776-868). Its
        lines 808-868 repeat :423-494 step for step (material preparation, conductor tets,
        the ``cached_matrices`` test, ``to_csc(E - B K0**2)``) on the same caches, so the
        two entry points reuse each other's element matrices as the reference's do. The
        scattered-field excitation that follows it (:878 on) is the scattered form of
        ``emerge_amd.robin``'s port force."""
        return self.assemble_freq_matrix(er, ur, tand, cond, frequency, cache_matrices,
                                         frequency_dependent)

    def assemble_eig_matrices(self, er, ur, tand, cond, frequency: float):
        """The volume step of ``assemble_eig_matrix`` (assembler.py:1115-1153): the same
        material preparation at the compilation frequency, the element matrices without the
        conductor tets, then ``to_csc`` of the stiffness and of the mass separately
        (:1151-1152). Returns ``(K, M, cscmap, conductor_tets)`` with ``K`` and ``M`` CSC
        value arrays on the device; no caching (the reference rebuilds them each call)."""
        plan, er_m, ur_m, conductor_tets, _ = self._prepare(er, ur, tand, cond, frequency)
        K, M = plan.assemble_km(er_m, ur_m, self.precision)
        self.assemblies += 1
        cm = self._cscmap(plan)
        sym = _BLOCKS_SYMMETRIC
        return (cm.to_csc(K, block_symmetric=sym), cm.to_csc(M, block_symmetric=sym), cm,
                conductor_tets)

    def _build_lossy(self, lc: _LossyCache, frequency: float) -> None:
        """CSC values of K, M(er (1 - j tand)) and M(cond) of the cached inputs, assembled
        straight into the pattern (one element-matrix pass each, no COO arrays)."""
        er, ur, tand, cond = lc.inputs
        plan, cm = lc.plan, lc.cscmap
        zero = torch.zeros_like(cond)
        er_a, _, _ = prepare_materials(er, tand, zero, frequency, self.peclim, self.surfimplim)
        er_am, ur_m = self._descriptors(plan, er_a, er, ur, tand, zero)
        cond_m = plan.material(cond, layout="ref")
        p = self.precision
        lc.K = plan.assemble_csc(cm, er_am, ur_m, 0.0, "K", p)
        lc.Ma = plan.assemble_csc(cm, er_am, ur_m, 0.0, "M", p)
        lc.Mc = plan.assemble_csc(cm, cond_m, ur_m, 0.0, "M", p)
        self.assemblies += 1

    def _prepare(self, er, ur, tand, cond, frequency):
        """Prepared materials on the device, the plan of their conductor set and the
        kernels' descriptors of er' and ur."""
        plan, er_d, conductor_tets, fd = self._prepare_er(er, tand, cond, frequency)
        er_m, ur_m = self._descriptors(plan, er_d, er, ur, tand, cond)
        return plan, er_m, ur_m, conductor_tets, fd

    def _prepare_er(self, er, tand, cond, frequency):
        """er' = er(1 - j tand) - j cond/(w eps0) on the device, the plan of the conductor
        set and the frequency dependence."""
        er_d, conductor_tets, fd = prepare_materials(er, tand, cond, frequency, self.peclim,
                                                     self.surfimplim)
        return plan_for(self.field, conductor_tets), er_d, conductor_tets, fd

    def _descriptors(self, plan, er_d, er, ur, tand, cond):
        """The kernels' descriptors of er' and ur, with their structure (scalar, diagonal
        or full) checked where the inputs live; only needed when element matrices are
        assembled, not when cached ones are reused."""
        T = plan.n_tets
        kind = _prepared_kind(er, tand, cond)
        comp = {_lib.EM_MAT_FULL: T, _lib.EM_MAT_DIAG: 4 * T, _lib.EM_MAT_SCALAR: 0}[kind]
        # ur is a (3, 3, T) tensor whether it lives on the host or on the device
        ur_m = plan.material(ur, layout="ref")
        return Material(er_d, kind, 1, comp), ur_m

    def _cscmap(self, plan) -> CscMap:
        """The CSC map of ``plan`` (the reference's ``cached_cscmap``)."""
        if self._map is None or self._map[0] is not plan:
            self._map = (plan, CscMap(plan))
        return self._map[1]
