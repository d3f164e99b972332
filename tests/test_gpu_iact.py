"""IACT of the MGMC sampler against the reference algorithm: the batched
multicolour sampler's integrated autocorrelation times at seeded probe
vertices are within 10 % of those of the oracle's lexicographic MGMC chains
(the reference's own visiting order and CLCG + polar noise,
src/sampler/multigridmc_sampler.cc, src/sampler/sor_sampler.cc) at the same
cycle and smoother settings; the oracle's multicolour chains match too."""
import numpy as np
import pytest

import multigridmc_amd as mg
from oracle import pyoracle as po

from gpu_common import make_problem

pytestmark = pytest.mark.gpu


def _taus(obs):
    t = np.array([mg.tau_int_chains(obs[:, :, p])[0] for p in range(obs.shape[2])])
    return t.max(), np.median(t)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_iact_matches_lexicographic_reference(precision):
    n, nlevel = (16, 16, 16), 3
    op, ora = make_problem(3, n, 4)
    N = op.lattice.Nvertex
    rng = np.random.default_rng(7)
    probes = rng.choice(N, size=8, replace=False)
    f = np.zeros(N)
    K, ns = 128, 2000
    s = mg.MultigridMCSampler(op, mg.MultigridParameters(nlevel=nlevel), nchains=K, precision=precision, seed=99)
    s.fix_rhs(f)
    s.set_state(np.zeros((K, N)))
    s.set_probes(probes)
    s.run(20, record=False)
    gpu = _taus(s.run(ns))
    ref = {}
    for name, ordering in (("lex", po.LEX), ("mc", po.MULTICOLOUR)):
        omg = po.Multigrid(ora, nlevel, ordering=ordering)
        xw, _, _ = omg.chains(f, np.zeros((48, N)), 20, nthreads=8, probes=probes)
        _, obs, _ = omg.chains(f, np.asarray(xw).reshape(48, N), 1500, nthreads=8, probes=probes)
        ref[name] = _taus(obs)
    print(f"tau max/median: gpu {gpu}, oracle lexicographic {ref['lex']}, oracle multicolour {ref['mc']}")
    for k in range(2):
        assert abs(gpu[k] / ref["lex"][k] - 1.0) < 0.10
        assert abs(gpu[k] / ref["mc"][k] - 1.0) < 0.10
