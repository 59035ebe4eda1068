"""Autotuner for launch shapes (no counterpart in the reference).

It picks, by timing the candidates on the caller's own data, three things:
* the workgroup size and rows per thread of the Taylor-term kernels (plain FP64, plain mixed
  precision and batched): paces_*taylor_term*_tuned with paces_launch_cfg_t;
* the rows per workgroup of the scorer's fused inference: paces_mlp_infer_tiled.

Every candidate gives bit-identical results, so tuning changes only the speed.

Timing uses HIP events on the launch stream, with one untimed warm-up call and ``reps`` timed
calls per candidate. Choices are cached per (kernel, size bucket); a bucket is half an octave of
the row count. With ``cache_file`` set, the choices persist as JSON across processes.
"""

from __future__ import annotations

import json
import math
import os

import torch

from .. import _hip

SHAPES = [(b, r) for b in (128, 256, 512, 1024) for r in (1, 2, 4)]
DEFAULT_SHAPE = (256, 1)
MLP_TILES = (64, 128, 256)


def bucket(n):
    """Half-octave bucket of a row count."""
    return int(round(2.0 * math.log2(max(int(n), 1))))


def launch_cfg(shape):
    return None if shape is None else _hip.LaunchCfg(int(shape[0]), int(shape[1]))


class Autotuner:
    def __init__(self, reps=3, cache_file=None):
        self.reps = int(reps)
        self.cache_file = cache_file
        self.choices = {}
        self.timings = {}
        if cache_file and os.path.exists(cache_file):
            with open(cache_file, encoding="utf-8") as f:
                for k, v in json.load(f).items():
                    self.choices[k] = tuple(v) if isinstance(v, list) else v

    @staticmethod
    def key(kind, n, *extra):
        return "|".join([kind, str(bucket(n))] + [str(e) for e in extra])

    def _time(self, launch):
        st = torch.cuda.current_stream()
        launch()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(self.reps):
            launch()
        e1.record(st)
        e1.synchronize()
        return e0.elapsed_time(e1) * 1e-3 / self.reps

    def pick(self, key, candidates, make_launch):
        """The cached choice for key, or the fastest of candidates; make_launch(c) returns a
        callable that launches the kernel once with candidate c."""
        if key in self.choices:
            return self.choices[key]
        times = {}
        for c in candidates:
            times[c] = self._time(make_launch(c))
        best = min(times, key=times.get)
        self.choices[key] = best
        self.timings[key] = {str(c): t for c, t in times.items()}
        if self.cache_file:
            with open(self.cache_file, "w", encoding="utf-8") as f:
                json.dump({k: list(v) if isinstance(v, tuple) else v
                           for k, v in self.choices.items()}, f, indent=1)
        return best


_default = None


def default_tuner():
    """One tuner per process, shared by every engine that enables autotuning."""
    global _default
    if _default is None:
        _default = Autotuner(cache_file=os.environ.get("PACES_AMD_TUNE_CACHE"))
    return _default
