"""COMMON-block mirror of the hot path: the subset of the reference's module-level state
(src/epsr26/common/) that atogen, gensq, stog, the energy routines, molfit, fmolec1,
generate_rp/ep, fit_expansion and ptgnsq read and write, with the reference's names,
1-based indices and column-major shapes.  The drop-in routines in
epsr_python_translation_amd.library and .programs read and write it; the GPU mirror of
it lives in epsr_python_translation_amd.common.gpu."""

from . import clustercom, epsr16, molecom, potential_par, ran1com, weights  # noqa: F401


def reset():
    for m in (molecom, clustercom, epsr16, potential_par, ran1com, weights):
        m.reset()
    from . import gpu
    gpu.reset()
