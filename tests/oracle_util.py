"""Helpers for the oracle tests: a small domain with a controllable state."""
import numpy as np

from oracle.pyorc import Oracle
from xbeach_gpu_amd import default_params

RHO, G = 1025.0, 9.81


def params(nx=24, ny=12, ndeg=10.0, span=60.0, **kw):
    base = dict(nx=nx, ny=ny, dx=2.0, thetamin=-span, thetamax=span, dtheta=ndeg, sedtrans=1, morphology=1,
                sedstart=0.0)
    base.update(kw)
    return default_params(**base)


def domain(prec=8, depth=5.0, **kw):
    """Flat-bed domain of uniform depth, still water, zero waves."""
    p = params(**kw)
    o = Oracle(p, prec)
    o.set("zb", np.full((p.ny, p.nx), depth))
    o.set("zs", np.zeros((p.ny, p.nx)))
    o.set("hh", np.full((p.ny, p.nx), depth))
    o.set("stdep", np.full((p.ny, p.nx), 5.0))
    o.set_forcing(Trep=10.0)
    o.init_state(10.0)
    return p, o


def theta(o):
    return o.get("theta").astype(np.float64)
