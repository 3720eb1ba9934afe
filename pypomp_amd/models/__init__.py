"""Native models: each is a device functor in the library plus this host spec."""

from .base import Component, NativeModel, ParTrans
from .lg import LG, LGModel, lg_theta
from .seir import SEIR_DEFAULTS, SEIRModel
from .spx import SPX_THETA, SPXModel, spx, spx_covars
from .sv import SV_DEFAULTS, SVModel

__all__ = ["Component", "NativeModel", "ParTrans", "LG", "LGModel", "lg_theta", "SEIRModel", "SEIR_DEFAULTS",
           "SVModel", "SV_DEFAULTS", "SPXModel", "SPX_THETA", "spx", "spx_covars"]
