"""Batches of independent sampling problems of uneven size on one GPU.

Each instance (its own lattice, measurements and chains) is a handle of the
library with its own HIP stream and its own captured cycle graph; run()
enqueues every instance's cycles before waiting, so the instances' kernels
overlap on the device (up to the hardware queues the runtime gives the
process). Observables are written straight into device buffers.
"""
import numpy as np

from . import _lib
from ._lib import check
from .sampler import MultigridMCSampler


class SamplerBatch:
    def __init__(self, instances):
        """instances: list of dicts with keys op, params, nchains and optional
        precision (default "fp32"), seed, f (right-hand side, default zero) and
        observables (list of (vertex indices, weights))."""
        import torch  # device buffers for the observables

        self._torch = torch
        self.samplers = []
        self.nobs = []
        for k, inst in enumerate(instances):
            op = inst["op"]
            s = MultigridMCSampler(op, inst["params"], nchains=inst["nchains"],
                                   precision=inst.get("precision", "fp32"), seed=inst.get("seed", 1000 + k))
            N = op.lattice.Nvertex
            s.fix_rhs(inst.get("f", np.zeros(N)))
            s.set_state(np.zeros((inst["nchains"], N)))
            obs = inst.get("observables", [])
            if obs:
                s.set_observables(obs)
            self.samplers.append(s)
            self.nobs.append(len(obs))

    def run(self, nsamples, record=True):
        """nsamples cycles of every instance, enqueued on all streams before one
        device-wide wait; returns a list of obs[nsamples, nchains, nobs]."""
        torch = self._torch
        bufs = []
        for s, no in zip(self.samplers, self.nobs):
            buf = None
            if record and no:
                buf = torch.zeros((nsamples, s.nchains, no), dtype=torch.float64, device="cuda")
            bufs.append(buf)
        torch.cuda.synchronize()
        for s, buf in zip(self.samplers, bufs):
            check(_lib.lib().mgmc_run_dev(s.h, nsamples, buf.data_ptr() if buf is not None else None, 1))
        torch.cuda.synchronize()
        return [b.cpu().numpy() if b is not None else None for b in bufs]
