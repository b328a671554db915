"""Coordinate-type tags for coefficient and source callables
(fealpy/decorator/coordinates.py:10-22): a ``cartesian`` function receives
physical points (NC, NQ, GD); otherwise it is called as ``f(bcs, index=...)``
with barycentric points (NQ, TD+1)."""
from functools import wraps


def cartesian(func):
    @wraps(func)
    def wrapper(*args, **kwargs):
        return func(*args, **kwargs)
    wrapper.coordtype = 'cartesian'
    return wrapper


def barycentric(func):
    @wraps(func)
    def wrapper(*args, **kwargs):
        return func(*args, **kwargs)
    wrapper.coordtype = 'barycentric'
    return wrapper
