"""Parity of every kernel family of the HIP path with the CPU oracle (FP64).

The oracle runs in its multicolour ordering, the order the kernels use, so
single sweeps, transfers and coarse solves agree to rounding; noise is
injected through the library's component entry points.
"""
import numpy as np
import pytest

import multigridmc_amd as mg
from multigridmc_amd import measurements as ms
from gpu_common import make_problem, normals_reference, oracle_mg, params
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

CASES = [
    # dim, n, nmeas, nlevel
    (2, (16, 16), 0, 3),
    (2, (16, 32), 3, 3),
    (3, (8, 8, 8), 0, 2),
    (3, (16, 16, 16), 3, 3),
    # x extents that are not multiples of 8: the last 8-vertex group of a row
    # holds the boundary vertex and padding
    (2, (24, 20), 3, 2),
    (3, (12, 16, 8), 3, 2),
]


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("case", CASES)
def test_level_stencils_match_galerkin(case):
    dim, n, nmeas, nlevel = case
    op, ora = make_problem(dim, n, nmeas)
    p = params(nlevel)
    s = mg.MultigridMCSampler(op, p, nchains=1)
    omg = oracle_mg(ora, p)
    for level in range(nlevel):
        st, ncol = s.level_stencil(level)
        A = omg.level_csr(level, 0).toarray()
        lat = mg.Lattice(*[v >> level for v in n])
        N = lat.Nvertex
        idx = lat.vertex_euclidean(np.arange(N))
        deep = np.where(np.all((idx > 1) & (idx < np.array(lat.n) - 1), axis=1))[0]
        r = deep[len(deep) // 2]
        for o in np.ndindex(*([3] * dim)):
            off = np.array(o) - 1
            j = lat.vertex_index(idx[r] + off)
            flat = sum((off[d] + 1) * 3 ** d for d in range(dim))
            assert abs(st[flat] - A[r, j]) <= 1e-14 * np.abs(A[r]).max()
        assert ncol == (2 if level == 0 else 2 ** dim)


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("omega", [1.0, 0.8])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 2e-6)])
def test_sweep_parity(case, omega, precision, tol):
    """One SOR sweep (+SMW) per level and direction, deterministic and with
    injected Gibbs noise, against the oracle's multicolour SORSmoother /
    SORSampler (src/smoother/sor_smoother.cc:41-77, src/sampler/sor_sampler.cc:37-59):
    FP64 to 1e-12, FP32 cycles (the fields in single precision, the low-rank
    algebra in FP64) to 2e-6 of the largest value."""
    dim, n, nmeas, nlevel = case
    op, ora = make_problem(dim, n, nmeas)
    p = params(nlevel, omega=omega)
    K = 2
    s = mg.MultigridMCSampler(op, p, nchains=K, precision=precision)
    omg = oracle_mg(ora, p)
    rng = np.random.default_rng(1)
    for level in range(nlevel):
        N = s.ndof(level)
        b, x = rng.standard_normal((K, N)), rng.standard_normal((K, N))
        xi, xl = rng.standard_normal((K, N)), rng.standard_normal((K, max(nmeas, 1)))
        for fw in (True, False):
            got = s.sweep(level, fw, False, b, x)
            for c in range(K):
                ref = omg.sor_smoother(level, fw, b[c], x[c])
                assert _rel(got[c], ref) < tol, (level, fw, c)
            got = s.sweep(level, fw, True, b, x, xi=xi, xi_lr=xl[:, :nmeas] if nmeas else None)
            for c in range(K):
                ref = omg.sor_sampler_noise(level, fw, b[c], x[c], xi[c], xl[c, :nmeas] if nmeas else None)
                assert _rel(got[c], ref) < tol, (level, fw, c)


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("precision,tol,tolp", [("fp64", 1e-12, 1e-13), ("fp32", 2e-6, 1e-6)])
def test_transfer_parity(case, precision, tol, tolp):
    """f_c = R(f - Q x) and x += alpha P x_c (multigridmc_sampler.cc:125-133),
    FP64 to 1e-12 / 1e-13, FP32 to 2e-6 / 1e-6 of the largest value."""
    dim, n, nmeas, nlevel = case
    op, ora = make_problem(dim, n, nmeas)
    p = params(nlevel)
    K = 2
    s = mg.MultigridMCSampler(op, p, nchains=K, precision=precision)
    omg = oracle_mg(ora, p)
    rng = np.random.default_rng(2)
    for level in range(nlevel - 1):
        N, Nc = s.ndof(level), s.ndof(level + 1)
        f, x, xc = rng.standard_normal((K, N)), rng.standard_normal((K, N)), rng.standard_normal((K, Nc))
        got = s.residual_restrict(level, f, x)
        got_p = s.prolongate_add(level, 0.7, xc, x)
        for c in range(K):
            ref = omg.restrict(level, f[c] - omg.level_apply(level, x[c]))
            assert _rel(got[c], ref) < tol
            ref_p = omg.prolongate_add(level, 0.7, xc[c], x[c])
            assert _rel(got_p[c], ref_p) < tolp


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-11), ("fp32", 1e-6)])
def test_coarse_parity(case, precision, tol):
    """Coarsest level: Cholesky sample with injected xi and the deterministic
    Q_c^{-1} b (cholesky_sampler.hh:52-66, cholesky_solver.cc:30-40); the
    factors are FP64 in both precisions, FP32 rounds the level's fields."""
    dim, n, nmeas, nlevel = case
    op, ora = make_problem(dim, n, nmeas)
    p = params(nlevel)
    K = 3
    s = mg.MultigridMCSampler(op, p, nchains=K, precision=precision)
    omg = oracle_mg(ora, p)
    Nc = s.ndof(nlevel - 1)
    rng = np.random.default_rng(3)
    f, xi = rng.standard_normal((K, Nc)), rng.standard_normal((K, Nc))
    got = s.coarse(f, xi_mode=2, xi=xi)
    got_d = s.coarse(f, xi_mode=0)
    for c in range(K):
        assert _rel(got[c], omg.coarse_sample_noise(f[c], xi[c])) < tol
        assert _rel(got_d[c], omg.coarse_solve(f[c])) < tol


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-13), ("fp32", 1e-6)])
@pytest.mark.parametrize("n,nlevel,K", [((16, 16, 24), 2, 40), ((64, 64, 64), 4, 17), ((24, 20), 2, 1)])
def test_coarse_mfma_matches_vector_kernel(precision, tol, n, nlevel, K):
    """The coarsest-level products on the FP64 matrix cores equal the vector
    kernel's for sample (injected xi and Philox normals) and solve, with
    partial 16-row and 16-chain tiles (539, 343 and 99 coarse unknowns; K not
    a multiple of 16)."""
    op, ora = make_problem(len(n), n, 3)
    p = params(nlevel)
    a = mg.MultigridMCSampler(op, p, nchains=K, precision=precision, seed=9)
    b = mg.MultigridMCSampler(op, p, nchains=K, precision=precision, seed=9)
    b.set_tuning("coarse_mfma", 0)
    Nc = a.ndof(nlevel - 1)
    rng = np.random.default_rng(13)
    f, xi = rng.standard_normal((K, Nc)), rng.standard_normal((K, Nc))
    for mode, x in ((2, xi), (0, None), (1, None)):
        ga, gb = a.coarse(f, xi_mode=mode, xi=x), b.coarse(f, xi_mode=mode, xi=x)
        assert np.abs(ga - gb).max() <= tol * np.abs(gb).max(), mode


@pytest.mark.parametrize("n,radius", [((16, 16, 16), 0.15), ((32, 32, 32), 0.3)])
def test_apply_operator(n, radius):
    """Q x with the low-rank term; radius 0.3 at 32^3 gives B columns of
    several thousand entries, which k_lr_dot splits over several blocks."""
    op, ora = make_problem(3, n, 3, radius=radius)
    s = mg.MultigridMCSampler(op, params(3), nchains=2)
    x = np.random.default_rng(4).standard_normal((2, op.lattice.Nvertex))
    y = s.apply_operator(x)
    for c in range(2):
        assert _rel(y[c], ora.apply(x[c])) < 1e-13


def test_normals_match_philox():
    """The Gaussians the sweeps draw equal Philox4x32-10 + Box-Muller computed
    independently in NumPy (tolerance covers the fast log/sin/cos), and pass the
    reference's normal-CDF test (test_random.hh:82-109)."""
    lat = mg.Lattice(64, 64, 64)
    op = mg.ShiftedLaplaceFDOperator(lat, 10.0)
    seed = 0x1234567890ABCDEF
    s = mg.MultigridMCSampler(op, params(4), nchains=4, seed=seed)
    s.set_counter(77)
    z = s.normals(5)
    ref = normals_reference(lat, 4, 5, 77, seed)
    # the integer stream is exact; the FP32 v_log/v_sin/v_cos Box-Muller differs
    # from the FP64 NumPy evaluation by a few ulp of the result
    err = np.abs(z - ref)
    assert err.max() < 1e-4 and err.mean() < 1e-6
    flat = z.reshape(-1)
    import math
    h, edges = np.histogram(flat, bins=64, range=(-10, 10))
    cdf = np.cumsum(h) / flat.size
    true = np.array([0.5 * (1 + math.erf(v / math.sqrt(2))) for v in edges[1:]])
    assert np.abs(cdf - true).max() < 2e-3
    assert abs(flat.mean()) < 5e-3 and abs(flat.var() - 1) < 5e-3


def test_csr_operator_input_matches_fd():
    """A supplied as the reference's CSR matrix is recognised as the same stencil."""
    import ctypes as C
    from multigridmc_amd import _lib
    op, ora = make_problem(3, (16, 16, 16), 0)
    s_fd = mg.MultigridMCSampler(op, params(3))
    A = op.get_sparse().tocsr()
    A.sort_indices()
    prob, keep = mg.sampler._problem(op)
    rp, ci, v = _lib.i64(A.indptr), _lib.i64(A.indices), _lib.f64(A.data)
    prob.A_rowptr, prob.A_col, prob.A_val = _lib.ptr(rp), _lib.ptr(ci), _lib.ptr(v)
    par = _lib.Params(nlevel=3, smoother=1, npresmooth=1, npostsmooth=1, ncoarsesmooth=1, omega=1.0, cycle=1,
                      coarse_scaling=1.0, coarse_solver=0, precision=0, nchains=1, seed=1, stream=None)
    h = C.c_void_p()
    _lib.check(_lib.lib().mgmc_create(C.byref(prob), C.byref(par), C.byref(h)))
    try:
        for level in range(3):
            st = np.zeros(27)
            nc = C.c_int()
            _lib.check(_lib.lib().mgmc_level_stencil(h, level, _lib.ptr(st), C.byref(nc)))
            assert np.allclose(st, s_fd.level_stencil(level)[0], rtol=1e-14, atol=0)
    finally:
        _lib.lib().mgmc_destroy(h)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-13), ("fp32", 2e-6)])
@pytest.mark.parametrize("threads,blocks,balanced", [(512, 1024, 0), (128, 1, 0), (64, 100000, 0), (512, 1, 1),
                                                     (64, 1, 1)])
def test_fused_sweep_matches_colour_phases(precision, tol, threads, blocks, balanced):
    """The single-pass red-black kernel (tiles of full x-rows, z-segments with
    redundant halo updates) equals two colour-phase launches, for Philox and
    injected noise, in both directions, over several tile / segment shapes,
    also with the (chain, tile, plane) space split evenly over the resident
    workgroups (runs that cross from one column into the next)."""
    op, _ = make_problem(3, (32, 16, 24), 3)
    p = params(3)
    K = 3
    ref = mg.MultigridMCSampler(op, p, nchains=K, precision=precision, seed=5)
    ref.set_tuning("fused_sweep", 0)
    fus = mg.MultigridMCSampler(op, p, nchains=K, precision=precision, seed=5)
    fus.set_tuning("rb_threads", threads)
    fus.set_tuning("rb_blocks", blocks)
    fus.set_tuning("rb_balanced", balanced)
    N = op.lattice.Nvertex
    rng = np.random.default_rng(6)
    b, x, xi = rng.standard_normal((3, K, N))
    xl = rng.standard_normal((K, 3))
    for fw in (True, False):
        for kw in ({}, {"xi": xi, "xi_lr": xl}):
            for st in (False, True):
                if not st and kw:
                    continue
                a = ref.sweep(0, fw, st, b, x, **kw)
                c = fus.sweep(0, fw, st, b, x, **kw)
                assert np.abs(a - c).max() <= tol * np.abs(a).max(), (fw, st, bool(kw))


@pytest.mark.parametrize("precision,n,K,threads", [("fp32", (128, 128, 128), 64, 512), ("fp64", (128, 128, 128), 1, 512),
                                                   ("fp32", (48, 40, 36), 5, 64), ("fp64", (48, 40, 36), 3, 128),
                                                   ("fp32", (32, 24, 32), 300, 1024)])
def test_balanced_sweep_split_gives_the_same_samples(precision, n, K, threads):
    """Whole sampling cycles (folded coarse correction, low-rank terms) with
    the level-0 sweep's (chain, tile, plane) space split evenly over one round
    of resident workgroups equal those of its z-segment split bit for bit,
    among them the 128^3 posterior with 32 measurements and 64 FP32 chains,
    and 300 chains on a small lattice (more columns than resident workgroups:
    a run crosses several columns)."""
    lat = mg.Lattice(*n)
    centres, y, var = mg.seeded_ball_measurements(lat, 32, 0.05, 1e-4, 3)
    op = mg.MeasuredOperator(mg.ShiftedLaplaceFDOperator(lat, 10.0), centres, 0.05, var)
    p = mg.MultigridParameters(nlevel=ms._nlevel(op))
    f = op.get_B() @ (y / var)
    states = []
    for balanced in (0, 1):
        s = mg.MultigridMCSampler(op, p, nchains=K, precision=precision, seed=44)
        s.set_tuning("rb_threads", threads)
        s.set_tuning("rb_balanced", balanced)
        assert s.get_tuning("rb_balanced") == balanced
        s.fix_rhs(f)
        s.set_state(np.zeros((K, lat.Nvertex)))
        s.run(3, record=False)
        states.append(s.get_state())
    assert np.array_equal(states[0], states[1])


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 2e-5)])
@pytest.mark.parametrize("K", [3, 12])
def test_carried_residual_term_and_dot_split_match_the_gather(precision, tol, K):
    """The residual's low-rank term taken from the SMW step before it (B^T x =
    z_p + B^T W u_p, lr_carry) and other splits of B's columns over the dot's
    workgroups (lr_split) give the samples of the direct gather over x, to
    rounding; with 3 chains the solves run inside the update launches, with
    12 in launches of their own. Whole V- and W-cycles of a 3D posterior."""
    op, _ = make_problem(3, (32, 24, 32), 4)
    N = op.lattice.Nvertex
    f = np.random.default_rng(14).standard_normal(N)
    x0 = np.random.default_rng(15).standard_normal((K, N))
    for cycle in (1, 2):
        p = params(4, cycle=cycle)
        states = []
        for carry, split in ((0, 512), (1, 512), (1, 64), (1, 4096)):
            s = mg.MultigridMCSampler(op, p, nchains=K, precision=precision, seed=31)
            s.set_tuning("lr_carry", carry)
            s.set_tuning("lr_split", split)
            assert s.get_tuning("lr_carry") == carry and s.get_tuning("lr_split") == split
            s.fix_rhs(f)
            s.set_state(x0)
            s.run(3, record=False)
            states.append(s.get_state())
        for k in range(1, len(states)):
            assert np.abs(states[k] - states[0]).max() <= tol * np.abs(states[0]).max(), (cycle, k)


@pytest.mark.parametrize("n,K", [((32, 16, 24), 4), ((24, 20, 16), 6)])
@pytest.mark.parametrize("threads,blocks", [(512, 1024), (64, 1), (64, 100000)])
def test_chain_blocked_sweep_matches_one_chain_per_workgroup(n, K, threads, blocks):
    """Two chains per workgroup of the level-0 red-black sweep (f shared by
    the chains, FP32) give the sample path of one chain per workgroup: whole
    sampling cycles (the folded coarse correction and the low-rank terms
    included) over several tile and z-segment shapes. The arithmetic per
    chain is the same (on gfx950 the states come out bit-equal); the bound
    allows for the compiler contracting the two kernels' products apart."""
    op, _ = make_problem(3, n, 3)
    p = params(3)
    N = op.lattice.Nvertex
    rng = np.random.default_rng(8)
    f = rng.standard_normal(N)
    x0 = rng.standard_normal((K, N))
    states = []
    for cb in (1, 2):
        s = mg.MultigridMCSampler(op, p, nchains=K, precision="fp32", seed=21)
        s.set_tuning("rb_chains", cb)
        s.set_tuning("rb_threads", threads)
        s.set_tuning("rb_blocks", blocks)
        assert s.get_tuning("rb_chains") == cb
        s.fix_rhs(f)
        s.set_state(x0)
        s.run(3, record=False)
        states.append(s.get_state())
    a, c = states
    print(f"chain-blocked sweep, {n} x {K} chains, rb_threads {threads}: max |diff| {np.abs(a - c).max():.3e} "
          f"(bit-equal: {np.array_equal(a, c)})")
    assert np.abs(a - c).max() <= 2e-6 * np.abs(a).max()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-13), ("fp32", 2e-6)])
@pytest.mark.parametrize("stage3", [0, 1])
def test_plane_sweep_matches_colour_phases(precision, tol, stage3):
    """Box-stencil levels swept in two plane-parity launches equal the eight
    colour-phase launches (Philox and injected noise, both directions)."""
    op, _ = make_problem(3, (32, 32, 16), 3)
    p = params(3)
    K = 2
    ref = mg.MultigridMCSampler(op, p, nchains=K, precision=precision, seed=9)
    ref.set_tuning("box_march", 0)
    ref.set_tuning("plane_sweep", 0)
    new = mg.MultigridMCSampler(op, p, nchains=K, precision=precision, seed=9)
    new.set_tuning("box_march", 0)
    new.set_tuning("plane_sweep_min", 1)
    new.set_tuning("plane_sweep_stage3", stage3)
    rng = np.random.default_rng(10)
    for level in (1, 2):
        N = ref.ndof(level)
        b, x, xi = rng.standard_normal((3, K, N))
        xl = rng.standard_normal((K, 3))
        for fw in (True, False):
            for st, kw in ((False, {}), (True, {}), (True, {"xi": xi, "xi_lr": xl})):
                a = ref.sweep(level, fw, st, b, x, **kw)
                c = new.sweep(level, fw, st, b, x, **kw)
                assert np.abs(a - c).max() <= tol * np.abs(a).max(), (level, fw, st, bool(kw))


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
@pytest.mark.parametrize("smoother", ["SSOR", "SOR"])
def test_folded_prolongation_matches_separate(precision, tol, smoother):
    """The coarse-grid correction added while the first post-sweep reads its
    planes equals a separate prolongate-add launch (whole MGMC cycles)."""
    op, _ = make_problem(3, (32, 32, 32), 4)
    p = params(4, smoother=smoother)
    N = op.lattice.Nvertex
    f = np.random.default_rng(12).standard_normal(N)
    out = []
    for fold in (0, 1):
        s = mg.MultigridMCSampler(op, p, nchains=2, precision=precision, seed=21)
        s.set_tuning("fold_prolongation", fold)
        x = np.zeros((2, N))
        for _ in range(3):
            s.apply(f, x)
        out.append(x)
    assert np.abs(out[0] - out[1]).max() <= tol * np.abs(out[0]).max()


@pytest.mark.gpu
@pytest.mark.parametrize("n,threads", [((16, 24, 32), 64), ((16, 24, 32), 512), ((40, 24, 16), 128)])
def test_fused_residual_restriction_matches_oracle(n, threads):
    """The single-pass residual + restriction kernel of 3D star levels equals
    R(f - Q x) of the oracle for tiles of a few coarse rows up to whole planes."""
    op, ora = make_problem(3, n, 3)
    p = params(3)
    K = 3
    omg = oracle_mg(ora, p)
    rng = np.random.default_rng(31)
    s = mg.MultigridMCSampler(op, p, nchains=K)
    s.set_tuning("rr_threads", threads)
    ref = mg.MultigridMCSampler(op, p, nchains=K)
    ref.set_tuning("fused_residual", 0)
    N = s.ndof(0)
    f, x = rng.standard_normal((K, N)), rng.standard_normal((K, N))
    got = s.residual_restrict(0, f, x)
    sep = ref.residual_restrict(0, f, x)
    for c in range(K):
        want = omg.restrict(0, f[c] - omg.level_apply(0, x[c]))
        assert _rel(got[c], want) < 1e-12
        assert _rel(got[c], sep[c]) < 1e-13


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-13), ("fp32", 2e-6)])
@pytest.mark.parametrize("n,blocks,min_planes", [((32, 32, 16), 512, 1), ((32, 32, 16), 512, 2),
                                                  ((16, 48, 32), 6, 3), ((32, 32, 16), 1, 1)])
def test_box_march_matches_colour_phases(precision, tol, n, blocks, min_planes):
    """The single-pass 8-colour sweep of box levels (z-segments with
    recomputed boundary planes, LDS ring of four planes) equals the eight
    colour-phase launches: Philox and injected noise, both directions,
    segments of one plane up to whole chains."""
    op, _ = make_problem(3, n, 3)
    p = params(3)
    K = 3
    ref = mg.MultigridMCSampler(op, p, nchains=K, precision=precision, seed=19)
    ref.set_tuning("box_march", 0)
    ref.set_tuning("plane_sweep", 0)
    new = mg.MultigridMCSampler(op, p, nchains=K, precision=precision, seed=19)
    new.set_tuning("box_march", 1)
    new.set_tuning("box_march_blocks", blocks)
    new.set_tuning("box_march_min_planes", min_planes)
    rng = np.random.default_rng(21)
    for level in (1, 2):
        N = ref.ndof(level)
        b, x, xi = rng.standard_normal((3, K, N))
        xl = rng.standard_normal((K, 3))
        for fw in (True, False):
            for st, kw in ((False, {}), (True, {}), (True, {"xi": xi, "xi_lr": xl})):
                a = ref.sweep(level, fw, st, b, x, **kw)
                c = new.sweep(level, fw, st, b, x, **kw)
                assert np.abs(a - c).max() <= tol * np.abs(a).max(), (level, fw, st, bool(kw))
