"""stim_amd: MI355X-native detection-event sampler with the reference's sampler API.

    import stim_amd
    c = stim_amd.Circuit(open('surface.stim').read())
    dets, obs = c.compile_detector_sampler(seed=5).sample(2**20, bit_packed=True, separate_observables=True)
"""

from .sampler import Circuit, CompiledDetectorSampler, CompiledMeasurementSampler  # noqa: F401
from .dem import DetectorErrorModel, CompiledDemSampler  # noqa: F401
from .m2d import CompiledMeasurementsToDetectionEventsConverter  # noqa: F401

__all__ = ['Circuit', 'CompiledDetectorSampler', 'CompiledMeasurementSampler', 'DetectorErrorModel',
           'CompiledDemSampler', 'CompiledMeasurementsToDetectionEventsConverter']
