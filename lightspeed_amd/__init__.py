"""lightspeed_amd: the reference's direct Coulomb/exchange build (IntBox::coulomb /
IntBox::exchange over a Schwarz-screened PairList), MI355X-native.

Usage mirrors the reference's Python API (reference src/python/__init__.py, core.py,
export_intbox.cpp):

    import lightspeed_amd as ls
    mol = ls.Molecule.from_xyz_file("h2o.xyz")
    bas = ls.Basis.from_gbs_file(mol, "6-31gs")
    res = ls.ResourceList.build()
    pairs = ls.PairList.build_schwarz(bas, bas, True, 1.0e-14)
    J = ls.IntBox.coulomb(res, ls.Ewald([1.0], [-1.0]), pairs, pairs, D, 1.0e-14, 1.0e-14)
    K = ls.IntBox.exchange(res, ls.Ewald([1.0], [-1.0]), pairs, pairs, D, True, 1.0e-14, 1.0e-14)

The J/K builds run in the HIP library liblsjk.so (gfx950); there is no CPU fallback.  Importing
fails loudly if the native extension has not been built (`python __graft_entry__.py build` or
`make -C lightspeed_amd/csrc`).
"""
import os as _os

from . import loaders as _loaders
from .loaders import Molecule, Atom, BOHR_PER_ANG, ANG_PER_BOHR

try:
    from . import _lightspeed as _native
except ImportError as _e:  # pragma: no cover - exercised only on an unbuilt tree
    raise ImportError(
        "lightspeed_amd: native extension not built (%s); run `make -C %s`"
        % (_e, _os.path.join(_os.path.dirname(__file__), "csrc"))) from _e

Tensor = _native.Tensor
Primitive = _native.Primitive
Basis = _native.Basis
Ewald = _native.Ewald
PairList = _native.PairList
PairListL = _native.PairListL
ResourceList = _native.ResourceList
PureTransform = _native.PureTransform
IntBox = _native.IntBox
DeviceArray = _native.DeviceArray
JKBatch = _native.JKBatch
Boys = _native.Boys

hip_init = _native.hip_init
hip_finalize = _native.hip_finalize
device_count = _native.device_count
set_counting = _native.set_counting
last_counts = _native.last_counts
last_counts_fp32 = _native.last_counts_fp32
last_wave_steps = _native.last_wave_steps
set_streams = _native.set_streams
last_timing = _native.last_timing
synchronize = _native.synchronize
set_profiling = _native.set_profiling
last_launches = _native.last_launches
count_quartets = _native.count_quartets
autotune = _native.autotune
set_kgroup_rows = _native.set_kgroup_rows

NATIVE_LIBRARY = _os.path.join(_os.path.dirname(__file__), "liblsjk.so")


def _basis_from_records(name, recs):
    return Basis(name, [Primitive(*r) for r in recs])


def _basis_from_gbs_file(molecule, filename, normalize=True, spherical=None):
    """Basis for `molecule` from a GBS file (reference src/python/core.py:505-548)."""
    name, recs = _loaders.basis_records_from_gbs_file(molecule, filename, normalize=normalize,
                                                      spherical=spherical)
    return _basis_from_records(name, recs)


def _basis_from_gbs_str(molecule, strval, name="bas", normalize=True, spherical=None):
    recs = _loaders.basis_records_from_gbs_lines(molecule, strval.split("\n"), name=name,
                                                 normalize=normalize, spherical=spherical)
    return _basis_from_records(name, recs)


def _basis_concatenate(bases):
    """Union of sub-bases with offset indices (reference Basis::concatenate, basis.hpp:528-542)."""
    prims = []
    ao = cart = prim = shell = atom = 0
    for b in bases:
        for p in b.primitives:
            prims.append(Primitive(p.c, p.e, p.x, p.y, p.z, p.c0, p.L, p.is_pure, p.aoIdx + ao,
                                   p.cartIdx + cart, p.primIdx + prim, p.shellIdx + shell, p.atomIdx + atom))
        ao += b.nao
        cart += b.ncart
        prim += b.nprim
        shell += b.nshell
        atom += b.natom
    return Basis("-".join(b.name for b in bases), prims)


Basis.from_gbs_file = staticmethod(_basis_from_gbs_file)
Basis.from_gbs_str = staticmethod(_basis_from_gbs_str)
Basis.concatenate = staticmethod(_basis_concatenate)
Basis.basisdir = _loaders.BASISDIR

__all__ = [
    "Molecule", "Atom", "Tensor", "Primitive", "Basis", "Ewald", "PairList", "PairListL",
    "ResourceList", "PureTransform", "IntBox", "DeviceArray", "JKBatch", "Boys", "hip_init", "hip_finalize",
    "device_count",
]
