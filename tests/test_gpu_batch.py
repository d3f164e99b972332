"""A mixed batch of 16 posterior sampling problems on lattices drawn from
{32^3, 48^3, 64^3, 96^3, 128^3} with 8-32 ball measurements each, run as
one batch whose launches are merged across instances (one launch per kernel
and point of the V-cycle, the lattices side by side as columns, one captured
graph per cycle): every instance equals its own sequential run, and its
observables have the exact posterior means and variances from FP64
multigrid solves. The persistent form of the merged launches and the
one-stream-per-instance form give the same samples."""
import numpy as np
import pytest

import multigridmc_amd as mg
from multigridmc_amd import measurements as ms
from multigridmc_amd.batch import SamplerBatch

pytestmark = pytest.mark.gpu


def _instances(seed=2024, count=16, nchains=4, sizes=(32, 48, 64, 96, 128)):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(rng.choice(sizes))
        lat = mg.Lattice(n, n, n)
        m = int(rng.integers(8, 33))
        centres, y, var = mg.seeded_ball_measurements(lat, m, 0.05, 1e-4, seed + k)
        op = mg.MeasuredOperator(mg.ShiftedLaplaceFDOperator(lat, 10.0), centres, 0.05, var)
        f = op.get_B() @ (y / var)
        obs = [mg.measurement_vector(lat, rng.uniform(0.2, 0.8, 3), 0.1), mg.measurement_vector(lat, [0.5] * 3, 0.0)]
        nlevel = ms._nlevel(op)
        out.append(dict(op=op, params=mg.MultigridParameters(nlevel=nlevel), nchains=nchains, f=f,
                        observables=obs, seed=77 + k, y=y))
    return out


def _sequential(i, nwarm, nsamples):
    s = mg.MultigridMCSampler(i["op"], i["params"], nchains=i["nchains"], precision=i.get("precision", "fp32"),
                              seed=i["seed"])
    s.fix_rhs(i["f"])
    s.set_state(np.zeros((i["nchains"], i["op"].lattice.Nvertex)))
    s.set_observables(i["observables"])
    s.run(nwarm, record=False)
    return s.run(nsamples)


def test_mixed_batch_equals_sequential_and_matches_posterior():
    inst = _instances()
    sizes = sorted({i["op"].lattice.n[0] for i in inst})
    assert len(sizes) >= 3
    batch = SamplerBatch(inst)
    merged, separate = batch.stats()
    print(f"mixed batch of 16: {merged} launches per cycle merged, {separate} separate", flush=True)
    assert merged < separate / 8
    batch.run(30, record=False)
    z = batch.run(300)
    # the same instances one after the other (same seeds): identical samples
    for k in (0, 5, 15):
        assert np.array_equal(_sequential(inst[k], 30, 300), z[k]), k
    # posterior moments of the observables, within Monte Carlo errors
    for k, i in enumerate(inst):
        means, cov = ms.observed_moments(i["op"], np.zeros(i["op"].lattice.Nvertex), i["y"], i["observables"],
                                         rtol=1e-11)
        for p in range(len(i["observables"])):
            series = z[k][:, :, p]
            tau = mg.tau_int_chains(series)[0]
            se = np.sqrt(cov[p, p] * tau / series.size)
            assert abs(series.mean() - means[p]) < 5 * se, (k, p)
            assert abs(series.var() / cov[p, p] - 1.0) < 6 * np.sqrt(2 * tau / series.size) + 0.02, (k, p)


def test_batch_forms_agree():
    """Merged launches replayed as a graph, issued eagerly, in the persistent
    form (one and two workgroups per CU), and one stream per instance: the
    same samples for FP32 and FP64 instances with different chain counts."""
    inst = _instances(seed=99, count=6, nchains=3, sizes=(32, 48, 64))
    inst[1]["precision"] = "fp64"
    inst[2]["nchains"] = 1
    inst[4]["precision"] = "fp64"
    inst[4]["nchains"] = 5
    inst[5]["nchains"] = 12  # separate low-rank solve launches (more than 8 chains)
    ref = [_sequential(i, 3, 20) for i in inst]
    for mode, persistent, graph in (("merged", 0, True), ("merged", 0, False), ("merged", 1, True),
                                    ("merged", 2, True), ("streams", 0, True)):
        b = SamplerBatch(inst, mode=mode, persistent=persistent)
        b.run(3, record=False, use_graph=graph)
        z = b.run(20, use_graph=graph)
        for k in range(len(inst)):
            assert np.array_equal(ref[k], z[k]), (mode, persistent, graph, k)
        b.close()


def test_batch_follows_tuning_and_rejects_unbatchable_handles():
    """A member's launch shapes changed after the batch was made (a tuning
    key) are picked up by the next run, with unchanged samples; a handle in
    the lexicographic order (wavefront kernels that cannot be merged) is
    refused with MGMC_ERR_UNSUPPORTED."""
    import ctypes as C

    from multigridmc_amd import _lib

    inst = _instances(seed=5, count=3, nchains=2, sizes=(32, 48))
    ref = [_sequential(i, 2, 10) for i in inst]
    b = SamplerBatch(inst)
    b.samplers[0].set_tuning("rb_threads", 256)
    b.samplers[2].set_tuning("rb_blocks", 64)
    b.run(2, record=False)
    z = b.run(10)
    for k in range(len(inst)):
        assert np.array_equal(ref[k], z[k]), k
    b.close()
    i = inst[0]
    s = mg.MultigridMCSampler(i["op"], i["params"], nchains=1, precision="fp32", seed=1, ordering="lexicographic")
    hs = (C.c_void_p * 1)(s.h.value)
    out = C.c_void_p()
    assert _lib.lib().mgmc_batch_create(1, hs, C.byref(out)) == 2
    # refused after level 0 was recorded (the single-pass box march on the
    # coarser levels): the handle's own cycles stay those of a fresh handle
    runs = []
    for offer in (True, False):
        t = mg.MultigridMCSampler(i["op"], i["params"], nchains=2, precision="fp32", seed=3)
        t.set_tuning("box_march", 1)
        t.fix_rhs(i["f"])
        t.set_state(np.zeros((2, i["op"].lattice.Nvertex)))
        t.set_observables(i["observables"])
        if offer:
            hs = (C.c_void_p * 1)(t.h.value)
            assert _lib.lib().mgmc_batch_create(1, hs, C.byref(out)) == 2
        runs.append(t.run(4))
    assert np.array_equal(runs[0], runs[1])


def test_batch_mixes_2d_and_3d_instances():
    """2D instances (colour-phase sweeps, 2D transfers, a 49-unknown coarsest
    level) merged with 3D ones: every instance equals its own run."""
    inst = _instances(seed=41, count=2, nchains=2, sizes=(32,))
    for k, n in enumerate((64, 48)):
        lat = mg.Lattice(n, n)
        centres, y, var = mg.seeded_ball_measurements(lat, 5, 0.1, 1e-3, 90 + k)
        op = mg.MeasuredOperator(mg.ShiftedLaplaceFDOperator(lat, 10.0), centres, 0.1, var)
        obs = [mg.measurement_vector(lat, [0.4, 0.6], 0.1), mg.measurement_vector(lat, [0.5, 0.5], 0.0)]
        inst.append(dict(op=op, params=mg.MultigridParameters(nlevel=ms._nlevel(op)), nchains=3 - k,
                         f=op.get_B() @ (y / var), observables=obs, seed=500 + k,
                         precision="fp64" if k else "fp32"))
    ref = [_sequential(i, 2, 12) for i in inst]
    b = SamplerBatch(inst)
    merged, separate = b.stats()
    assert merged < separate
    b.run(2, record=False)
    z = b.run(12)
    for k in range(len(inst)):
        assert np.array_equal(ref[k], z[k]), k
    b.close()
