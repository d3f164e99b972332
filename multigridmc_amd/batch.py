"""Batches of independent sampling problems of uneven size on one GPU.

mode="merged" (default): the instances' handles form one library batch
(mgmc_batch_*): at every point of the V-cycle the launches all instances
make of one kernel run as a single launch whose blocks are the instances'
blocks side by side (the lattices are columns of one launch per level
phase), and one cycle of the whole batch is one captured hipGraph; with
persistent=w the merged launches hold at most w workgroups per CU, which loop
over the blocks.

mode="streams": each instance keeps its own HIP stream and captured cycle
graph; run() enqueues every instance's cycles before waiting, so the
instances' kernels overlap on the device as far as the hardware queues the
runtime gives the process allow.

Either way every instance's samples equal those of its own sampler run, and
observables are written straight into device buffers.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check
from .sampler import MultigridMCSampler


class SamplerBatch:
    def __init__(self, instances, mode="merged", persistent=0):
        """instances: list of dicts with keys op, params, nchains and optional
        precision (default "fp32"), seed, f (right-hand side, default zero) and
        observables (list of (vertex indices, weights))."""
        import torch  # device buffers for the observables

        if mode not in ("merged", "streams"):
            raise ValueError("mode is 'merged' or 'streams'")
        self._torch = torch
        self.mode = mode
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
        self.b = None
        if mode == "merged":
            # long z-segments for the fused level sweeps: a merged launch
            # already has the other instances' workgroups to fill the GPU
            # (16-instance mixed batch: 1.97 ms per cycle against 2.04 with the
            # handles' short-segment default)
            for s in self.samplers:
                s.set_tuning("rb_min_planes", 8)
                s.set_tuning("rr_min_planes", 4)
            hs = (C.c_void_p * len(self.samplers))(*[s.h.value for s in self.samplers])
            b = C.c_void_p()
            check(_lib.lib().mgmc_batch_create(len(self.samplers), hs, C.byref(b)))
            self.b = b
            if persistent:
                check(_lib.lib().mgmc_batch_set_persistent(self.b, int(persistent)))

    def close(self):
        if self.b is not None and _lib._LIB is not None:
            _lib._LIB.mgmc_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stats(self):
        """(launches per batch cycle, launches the instances issue on their own)."""
        if self.b is None:
            return None
        a, s = C.c_int64(), C.c_int64()
        check(_lib.lib().mgmc_batch_stats(self.b, C.byref(a), C.byref(s)))
        return a.value, s.value

    def run_dev(self, nsamples, bufs, use_graph=True):
        """nsamples cycles of every instance into the given device buffers
        (torch tensors [nsamples, nchains, nobs] or None), without waiting."""
        if self.mode == "merged":
            ptrs = (C.c_void_p * len(self.samplers))(*[b.data_ptr() if b is not None else None for b in bufs])
            check(_lib.lib().mgmc_batch_run_dev(self.b, int(nsamples), ptrs, int(use_graph)))
        else:
            for s, buf in zip(self.samplers, bufs):
                check(_lib.lib().mgmc_run_dev(s.h, int(nsamples), buf.data_ptr() if buf is not None else None,
                                              int(use_graph)))

    def run(self, nsamples, record=True, use_graph=True):
        """nsamples cycles of every instance, then one device-wide wait;
        returns a list of obs[nsamples, nchains, nobs] (None without
        observables or with record=False)."""
        torch = self._torch
        bufs = []
        for s, no in zip(self.samplers, self.nobs):
            buf = None
            if record and no:
                buf = torch.zeros((nsamples, s.nchains, no), dtype=torch.float64, device="cuda")
            bufs.append(buf)
        torch.cuda.synchronize()
        self.run_dev(nsamples, bufs, use_graph)
        torch.cuda.synchronize()
        return [b.cpu().numpy() if b is not None else None for b in bufs]
