"""µSpectre's Newton-CG path on MI355X (gfx950 HIP).

Same Python surface as the reference package muSpectre
(language_bindings/python/muSpectre/__init__.py:43-103): ``Cell``,
``Formulation``, ``material``, ``solvers``, ``fft``, and the grid helpers.
The compiled module ``_muSpectre`` drives the HIP library libmugpu.so
through its C-ABI (include/mugpu.h); there is no CPU fallback.
"""
from . import _muSpectre
from ._muSpectre import (Formulation, FFT_PlanFlags, Precision, StrainMeasure,
                         StressMeasure, OptimizeResult, get_domain_ccoord,
                         get_domain_index, get_hermitian_sizes, material,
                         solvers, ConvergenceError, SolverError,
                         ProjectionError)
from . import fft

# the reference's engine table (muSpectre/__init__.py:48-51); 'hip' is this
# build's engine and 'fftw' names it too, so existing scripts run unchanged
_factories = {'fftw': ('CellFactory', False),
              'hip': ('CellFactory', False)}


def Cell(resolutions, lengths, formulation=Formulation.finite_strain,
         fft='fftw', communicator=None, precision=None, allow_even=False):
    """Instantiate a cell (muSpectre/__init__.py:54-103).

    resolutions: grid resolutions; lengths: physical cell size;
    formulation: Formulation.finite_strain (default) or small_strain;
    fft: 'fftw' or 'hip' (both select the HIP engine); communicator: must be
    None (one GPU).  Extensions: precision 'fp64' (default) or 'fp32'
    (FP32 operator inside an FP64-refined CG), allow_even to accept even
    grids with the Nyquist components of the projection set to zero.
    """
    try:
        factory_name, is_parallel = _factories[fft]
    except KeyError:
        raise KeyError("Unknown FFT engine '{}'.".format(fft))
    if communicator is not None:
        raise ValueError("FFT engine '{}' does not support parallel "
                         "execution.".format(fft))
    factory = getattr(_muSpectre, factory_name)
    return factory(resolutions, lengths, formulation, precision=precision,
                   allow_even=allow_even)
