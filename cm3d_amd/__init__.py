"""cm3d_amd -- MI355X (gfx950) engine for cm3d's nonbonded force path.

The compute lives in libcm3d_nb.so (hand-written HIP kernels behind the C ABI
of include/cm3d_nb.h); this package is its Python host side:

  nb        ctypes binding of the C ABI (NBContext: build_lists, need_rebuild,
            force -- cm3d's get_ngbr / nrefsh / force(...,iflg))
  systems   seeded synthetic cm3d systems (argon FCC, flexible water, the
            LJ + charge binary mixture)
  build     in-tree hipcc build of libcm3d_nb.so for gfx950
"""
from . import nb  # noqa: F401

__all__ = ["nb"]
