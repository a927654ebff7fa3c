"""MI355X-native direct spherical harmonic transform of weighted point sets.

Drop-in for martinjameswhite/directsht's ``from sht.sht import DirectSHT``:

    from directsht_amd.sht import DirectSHT
    alm = DirectSHT(Nl, Nx)(thetas, phis, weights)

The compute lives in libdsht.so (HIP kernels for gfx950, C ABI in
include/dsht.h); build it with ``python -m directsht_amd.build``.
"""
__all__ = ["DirectSHT"]


def __getattr__(name):
    if name == "DirectSHT":
        from .sht import DirectSHT
        return DirectSHT
    raise AttributeError(name)
