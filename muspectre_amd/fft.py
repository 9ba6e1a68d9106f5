"""FFT engines and projections (muSpectre/fft.py:63-151)."""
from . import _muSpectre

_factories = {'fftw': ('FFTW_2d', 'FFTW_3d', False),
              'hip': ('HIP_2d', 'HIP_3d', False)}

_projections = {_muSpectre.Formulation.finite_strain: 'FiniteStrainFast',
                _muSpectre.Formulation.small_strain: 'SmallStrain'}

fft_engines = [(name, par) for name, (_, _, par) in _factories.items()]


def FFT(resolutions, nb_components, fft='fftw', communicator=None):
    """Instantiate an FFT engine (muSpectre/fft.py:63-113)."""
    try:
        name_2d, name_3d, is_parallel = _factories[fft]
    except KeyError:
        raise KeyError("Unknown FFT engine '{}'.".format(fft))
    if len(resolutions) == 2:
        name = name_2d
    elif len(resolutions) == 3:
        name = name_3d
    else:
        raise ValueError('{}-d transforms are not supported'
                         .format(len(resolutions)))
    if communicator is not None:
        raise ValueError("FFT engine '{}' does not support parallel "
                         "execution.".format(fft))
    return getattr(_muSpectre.fft, name)(resolutions, nb_components)


def Projection(resolutions, lengths,
               formulation=_muSpectre.Formulation.finite_strain,
               fft='fftw', communicator=None):
    """Instantiate a projection (muSpectre/fft.py:116-151)."""
    name = 'Projection{}_{}d'.format(_projections[formulation],
                                      len(resolutions))
    try:
        factory = getattr(_muSpectre.fft, name)
    except AttributeError:
        raise KeyError("Projection engine '{}' has not been compiled into "
                       "the muSpectre library.".format(name))
    if communicator is not None:
        raise ValueError("one GPU: no communicator")
    return factory(resolutions, lengths, fft)
