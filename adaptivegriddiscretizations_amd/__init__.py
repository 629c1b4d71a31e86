"""MI355X-native GPU eikonal solver with the interface of AGD's
`Eikonal.dictIn(...).Run()` (mode='gpu').

    from adaptivegriddiscretizations_amd import Eikonal, Metrics
    hfmIn = Eikonal.dictIn({'model': 'Riemann2', 'metric': Metrics.Riemann(m), 'seed': [0., 0.]})
    hfmIn.SetRect([[-1, 1], [-1, 1]], dims=[512, 512])
    hfmOut = hfmIn.Run()

The compute path is libagd_hip.so (hand-written HIP kernels for gfx950),
reached through the C ABI of include/agd_hip.h.
"""
from . import AutomaticDifferentiation, Eikonal, Metrics, FiniteDifferences  # noqa: F401

__version__ = "0.1.0"
