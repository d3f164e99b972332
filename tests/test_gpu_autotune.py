"""The autotuner (mgmc_autotune) searches launch shapes that change no
computed value: after tuning, the chains' state and the sample path equal
those of an untuned handle with the same seed bit for bit, in both
precisions."""
import numpy as np
import pytest

import multigridmc_amd as mg

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_autotune_keeps_state_and_sample_path(precision):
    lat = mg.Lattice(32, 32, 32)
    centres, y, var = mg.seeded_ball_measurements(lat, 4, 0.1, 1e-3, 11)
    op = mg.MeasuredOperator(mg.ShiftedLaplaceFDOperator(lat, 10.0), centres, 0.1, var)
    f = op.get_B() @ (y / var)
    p = mg.MultigridParameters(nlevel=3)
    K = 8
    x0 = np.random.default_rng(3).standard_normal((K, lat.Nvertex))
    plain = mg.MultigridMCSampler(op, p, nchains=K, precision=precision, seed=5)
    tuned = mg.MultigridMCSampler(op, p, nchains=K, precision=precision, seed=5)
    for s in (plain, tuned):
        s.fix_rhs(f)
        s.set_state(x0)
    # start the search away from the defaults
    tuned.set_tuning("rb_threads", 256)
    tuned.set_tuning("rb_blocks", 2048)
    tuned.set_tuning("rb_balanced", 0)
    tuned.set_tuning("plane_sweep_stage3", 0)
    before = tuned.get_state()
    ms = tuned.autotune(ncycles=2)
    assert ms > 0.0
    assert np.array_equal(tuned.get_state(), before)
    plain.run(3, record=False)
    tuned.run(3, record=False)
    assert np.array_equal(tuned.get_state(), plain.get_state())


def test_tuning_keys_read_back_and_reject_bad_values():
    """The launch-shape keys read back what was set (defaults: balanced split
    on, carried residual term on, 512 entries of B per dot workgroup at 8
    chains and more, 128 below); out-of-range values and unknown keys raise
    MGMC_ERR_ARG and leave the setting as it was."""
    lat = mg.Lattice(16, 16, 16)
    centres, y, var = mg.seeded_ball_measurements(lat, 3, 0.1, 1e-3, 5)
    op = mg.MeasuredOperator(mg.ShiftedLaplaceFDOperator(lat, 10.0), centres, 0.1, var)
    p = mg.MultigridParameters(nlevel=3)
    for K, split in ((8, 512), (2, 128)):
        s = mg.MultigridMCSampler(op, p, nchains=K, precision="fp32", seed=1)
        assert s.get_tuning("rb_balanced") == 1
        assert s.get_tuning("lr_carry") == 1
        assert s.get_tuning("lr_split") == split
        s.set_tuning("lr_split", 64)
        assert s.get_tuning("lr_split") == 64
        for key, bad in (("lr_split", 63), ("rb_min_planes", 0), ("rb_threads", 2048), ("lr_dot_chains", 3)):
            before = s.get_tuning(key)
            with pytest.raises(mg.MGMCError) as e:
                s.set_tuning(key, bad)
            assert "ERR_ARG" in str(e.value)
            assert s.get_tuning(key) == before
        with pytest.raises(mg.MGMCError):
            s.set_tuning("no_such_key", 1)
