"""MI355X-native implementation of CCMpred's pseudo-likelihood path.

Native code lives in ``ccmpred_amd/csrc`` (HIP kernels for gfx950 and a C
host layer) and is built into ``ccmpred_amd/lib/libccmpred_amd.so``; this
package is its Python front end.
"""
from . import msa  # noqa: F401

__all__ = ["msa", "PLL", "init_bias", "contact_map", "top_pairs", "nvar_padded", "nsingle_padded"]


def __getattr__(name):
    if name in ("PLL", "init_bias", "contact_map", "top_pairs", "nvar_padded", "nsingle_padded"):
        from . import pll
        return getattr(pll, name)
    raise AttributeError(name)
