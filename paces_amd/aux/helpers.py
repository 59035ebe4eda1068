"""Helpers shared by the framework (paces/aux/helpers.py)."""

import numpy

from ..config import SEARCHSORTED_TIMING_LEVEL


def cartesian_product(*arrays):
    """Cartesian product of 1-D arrays as an (n, len(arrays)) host array (helpers.py:10-17)."""
    la = len(arrays)
    dtype = numpy.result_type(*arrays)
    arr = numpy.empty([len(a) for a in arrays] + [la], dtype=dtype)
    for i, a in enumerate(numpy.ix_(*arrays)):
        arr[..., i] = a
    return arr.reshape(-1, la)


def print_searchsorted_timing(verb, delta_t, size1, size2):
    if verb > SEARCHSORTED_TIMING_LEVEL:
        print(f"This application of searchsorted took {delta_t * 1000} ms"
              f" (arg sizes {size1}, {size2}).")


def obs_attrs(**kwargs):
    """Attach ``fname`` and ``header`` to an observable method (helpers.py:27-42)."""
    def wrapper(f):
        for attr, val in kwargs.items():
            setattr(f, attr, val)
        return f
    return wrapper


def debug_lister(dbg_list):
    """Declare the debug columns a generate_mel_<term> returns (helpers.py:45-55)."""
    def wrapper(f):
        f.dbg_list = dbg_list
        return f
    return wrapper


def flatten_dbg_dict(d):
    """Flatten a dict of debug lists into one list (helpers.py:58-60)."""
    return [x for k in d for x in d[k]]


def unique_rows(rows, ws=None):
    """Sorted unique rows of an (n, W) device tensor (cupy_unique, paces/aux/helpers.py:63-75),
    by the library's LSD radix sort."""
    from .. import _hip
    from ..core import engine
    return engine.sort_rows(rows.contiguous(), ws if ws is not None else _hip.Workspace(rows.device),
                            unique=True)
