"""The oracle's transforms and stencils on several host threads.

TEST INFRASTRUCTURE ONLY, like the rest of ``oracle/``: used by ``tests/``
to run the oracle at the workload sizes in seconds and by the
``cpu_baseline`` leg of ``bench.py`` to time it on every host core.  Inside
the ``with`` block the module-level ``r2c``, ``c2r`` and ``deriv`` of
``hetddft_oracle`` are replaced by threaded forms; every value is formed as
in the single-threaded oracle (the FFTs are pocketfft through scipy.fft, the
stencils the same sums over slabs), and the originals come back on exit.
"""
from __future__ import annotations

import numpy as np


class ThreadedOracle:
    """The oracle's transforms and stencils spread over `workers` threads:
    scipy.fft (pocketfft, the same algorithm as numpy.fft) with `workers`
    threads, and each 13-point stencil pass cut into slabs along an axis it
    does not differentiate (numpy releases the GIL), so every value is formed
    as in the oracle.  The remaining whole-array expressions stay numpy's
    single-threaded ones."""

    def __init__(self, orc, workers):
        import scipy.fft
        from concurrent.futures import ThreadPoolExecutor
        self.orc, self.workers = orc, workers
        self.pool = ThreadPoolExecutor(workers)
        self.saved = (orc.r2c, orc.c2r, orc.deriv)
        plain = orc.deriv

        def r2c(f):
            return scipy.fft.rfftn(f, workers=workers)

        def c2r(F, shape, renor):
            return scipy.fft.irfftn(F, s=shape, axes=(0, 1, 2), norm="forward", workers=workers) * renor

        def deriv(f, axis, order, h):
            split = 0 if orc._AX[axis] != 0 else 1
            n = f.shape[split]
            out = np.empty_like(f)

            def run(lo, hi):
                sl = [slice(None)] * 3
                sl[split] = slice(lo, hi)
                out[tuple(sl)] = plain(f[tuple(sl)], axis, order, h)
            cuts = [(i * n // workers, (i + 1) * n // workers) for i in range(workers)]
            list(self.pool.map(lambda c: run(*c), [c for c in cuts if c[1] > c[0]]))
            return out
        self.patched = (r2c, c2r, deriv)

    def __enter__(self):
        self.orc.r2c, self.orc.c2r, self.orc.deriv = self.patched
        return self

    def __exit__(self, *exc):
        self.orc.r2c, self.orc.c2r, self.orc.deriv = self.saved
        self.pool.shutdown()
