"""pypomp_amd: an MI355X-native engine for pypomp's particle-filter hot path.

The bootstrap particle filter, the MOP(alpha) filter and the IF2 perturbed
filter run as hand-written HIP kernels for gfx950 in ``_lib/libmipomp.so``
(C ABI: include/mipomp.h), behind the reference's ``Pomp`` / functional API.
"""

from . import autotune, functional, models, proposals, random
from .maths import logmeanexp, logmeanexp_se
from .models import ParTrans
from .optim import BFGS, SGD, Adam, FullMatrixAdam, LearningRate, Newton, Optimizer, WeightedNewton
from .proposals import MVNDiagRW, MVNRWAdaptive, MVNRWFull
from .panel import PanelParameters, PanelPomp
from .pomp import Pomp
from .rw_sigma import RWSigma

functional_api = functional
F = functional

__all__ = ["Pomp", "PanelPomp", "PanelParameters", "ParTrans", "RWSigma", "Adam", "SGD", "FullMatrixAdam", "BFGS", "Newton", "WeightedNewton", "Optimizer", "LearningRate", "MVNDiagRW", "MVNRWFull", "MVNRWAdaptive", "proposals", "autotune", "functional", "F", "models", "random", "logmeanexp", "logmeanexp_se"]
