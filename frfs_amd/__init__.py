"""MI355X-native fast spectral Boltzmann collision operator Q(f,f).

Drop-in for the jaisw7/frfs scattering-model plugins ``vhs-gll``,
``vhs-gll-nodal`` and ``vss-gll``; the work runs in ``libfrfs_fq.so``
(HIP kernels for gfx950) behind the C ABI of ``include/frfs_fq.h``.
"""
__version__ = '0.1.0'
