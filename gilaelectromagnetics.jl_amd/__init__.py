"""MI355X-native Green-function operator with GilaElectromagnetics.jl's surface.

The discrete electric dyadic Green function matvec (the reference's
egoOpr!, src/glaAct.jl:33) runs in hand-written gfx950 HIP kernels behind
the C ABI of include/gila_amd.h (libgila_amd.so); this package is the
host-side mirror of the reference's Julia interface over that ABI.

The directory name is not a Python identifier; load it under the name
``gila_amd``:

    import importlib.util, sys
    spec = importlib.util.spec_from_file_location(
        "gila_amd", "<repo>/gilaelectromagnetics.jl_amd/__init__.py",
        submodule_search_locations=["<repo>/gilaelectromagnetics.jl_amd"])
    gila = importlib.util.module_from_spec(spec); sys.modules["gila_amd"] = gila
    spec.loader.exec_module(gila)
"""
from ._capi import GlaError, lib  # noqa: F401
from .gla import (  # noqa: F401
    ArgumentError, DimensionMismatch, GlaExtInf, GlaKerOpt, GlaOpr, GlaOprMem, GlaVol, adjoint, egoOpr_, eltype,
    glaSze, glaVolEveGen, isadjoint, isdiag, isexternaloperator, ishermitian, isposdef, isselfoperator,
    issymmetric, mul_, show, size,
)
from . import serve, solve  # noqa: F401,E402
