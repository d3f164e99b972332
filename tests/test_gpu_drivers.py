"""driver_mgmc and driver_mg end to end on the GPU: the reference's output
files (src/driver_mgmc.cc, src/driver_mg.cc, src/measurements/*.cc,
src/auxilliary/vtk_writer2d.cc) with statistics checked against the exact
posterior of the configured problem."""
import numpy as np
import pytest

import multigridmc_amd as mg
from multigridmc_amd import measurements as ms
from driver_common import fixture_text, run_driver

pytestmark = pytest.mark.gpu


def _read_vtk(path):
    lines = path.read_text().splitlines()
    assert lines[0] == "# vtk DataFile Version 2.0" and lines[3] == "DATASET STRUCTURED_POINTS"
    npts = int([ln for ln in lines if ln.startswith("POINT_DATA")][0].split()[1])
    fields = {}
    for i, ln in enumerate(lines):
        if ln.startswith("SCALARS"):
            name = ln.split()[1]
            fields[name] = np.array([float(v) for v in lines[i + 2: i + 2 + npts]])
    return fields


def _problem():
    """The posterior of tests/golden/drivers/posterior_2d.cfg, built in Python."""
    lat = mg.Lattice(32, 32)
    prior = mg.ShiftedLaplaceFDOperator(lat, 1.0 / 0.3 ** 2)
    centres = np.array([[0.25, 0.25], [0.25, 0.75], [0.75, 0.25], [0.75, 0.75]])
    op = mg.MeasuredOperator(prior, centres, 0.05, np.array([1.0e-3, 2.0e-3, 1.5e-3, 1.0e-3]))
    y = np.array([1.5, -0.5, 0.25, 2.0])
    return op, y


def test_driver_mgmc_outputs_and_statistics(tmp_path):
    rc, out = run_driver("driver_mgmc", tmp_path)
    assert rc == 0, out[-3000:]
    op, y = _problem()
    b = mg.measurement_vector(op.lattice, [0.45, 0.55], 0.05)
    xbar = np.zeros(op.lattice.Nvertex)
    mean_exact, var_exact = ms.observed_mean_and_variance(op, xbar, y, b)
    x_post = ms.posterior_mean(op, xbar, y)
    # exact values printed by the harness (sampling_time.cc:62-66)
    exact_lines = [ln for ln in out.splitlines() if ln.strip().startswith("exact")]
    assert float(exact_lines[0].split("=")[1]) == pytest.approx(mean_exact, rel=1e-3)
    assert float(exact_lines[1].split("=")[1]) == pytest.approx(var_exact, rel=1e-3)
    for sfx in ("cholesky", "ssor", "multigridmc"):
        series = np.loadtxt(tmp_path / f"timeseries_{sfx}.txt")
        assert series.shape == (2000,)
        tau = mg.tau_int_sokal(series)[0]
        assert abs(series.mean() - mean_exact) < 5 * np.sqrt(var_exact * tau / len(series)), sfx
        assert series.var() / var_exact == pytest.approx(1.0, abs=6 * np.sqrt(2 * tau / len(series))), sfx
        conv = (tmp_path / f"convergence_{sfx}.txt").read_text().splitlines()
        assert conv[0].startswith("**** q_k = |E[z^k] - E[z]|")
        assert len([ln for ln in conv if ln.strip().startswith("mean")]) == 7
        mse = (tmp_path / f"meansquarederror_{sfx}.txt").read_text().splitlines()
        assert len(mse) == 6 and mse[0].split()[0] == "0"
    # the exact Cholesky sampler has no autocorrelation
    chol = np.loadtxt(tmp_path / "timeseries_cholesky.txt")
    assert mg.tau_int_sokal(chol)[0] < 1.3
    # posterior.vtk: sampled mean / variance fields and the exact mean
    f = _read_vtk(tmp_path / "posterior.vtk")
    assert set(f) == {"mean", "variance", "mean_exact"}
    grid = f["mean_exact"].reshape(33, 33)
    assert np.all(grid[0] == 0) and np.all(grid[:, -1] == 0)
    inner = grid[1:-1, 1:-1].reshape(-1)
    assert np.abs(inner - x_post).max() < 1e-5 * np.abs(x_post).max() + 1e-12
    assert (tmp_path / "sample_location.vtk").exists()


def _residual_table(out):
    lines = out.splitlines()
    i = lines.index("%5s   %8s   %12s   %6s" % ("iter", "||r||", "||r||/||r_0||", "rho"))
    rows = []
    for ln in lines[i + 1:]:
        parts = ln.split()
        if len(parts) != 4 or not parts[0].isdigit():
            break
        rows.append(parts)
    return rows


def _oracle_solve(ordering):
    """The oracle's loop solver on driver_mg's problem and right-hand side
    (the same libstdc++ std::mt19937_64(1482817) normal draws)."""
    from oracle import pyoracle as po
    op, _ = _problem()
    ora = po.Operator.fd(2, (32, 32), 1.0 / 0.3 ** 2)
    ora.add_measurements(np.array([[0.25, 0.25], [0.25, 0.75], [0.75, 0.25], [0.75, 0.75]]), 0.05,
                         np.array([1.0e-3, 2.0e-3, 1.5e-3, 1.0e-3]))
    omg = po.Multigrid(ora, 3, ordering=ordering)
    b = po.driver_mg_rhs(op.lattice.Nvertex)
    return omg.loop_solve(b, rtol=1e-10, atol=1e-12, maxiter=50)


@pytest.mark.parametrize("ordering", ["multicolour", "lexicographic"])
def test_driver_mg_matches_oracle_solve(tmp_path, ordering):
    """driver_mg (through HipMultigridPreconditioner and HipLoopSolver) prints
    the residual table of the oracle's loop solver on the same right-hand
    side, digit for digit down to residuals of 1e-9 relative (multicolour: the oracle's multicolour twin;
    lexicographic, gpu.ordering: the reference's own order, so the table is
    the reference's), and writes its solution to solution.vtk."""
    from oracle import pyoracle as po
    cfg = fixture_text("posterior_2d.cfg").replace("gpu = {", f'gpu = {{\n    ordering = "{ordering}";', 1)
    rc, out = run_driver("driver_mg", tmp_path, cfg_text=cfg)
    assert rc == 0, out[-3000:]
    rows = _residual_table(out)
    x_ref, hist, it = _oracle_solve(po.LEX if ordering == "lexicographic" else po.MULTICOLOUR)
    assert it >= 0 and len(rows) == it + 1
    # digit for digit while the residual is well above FP64 rounding of its
    # evaluation (relative 1e-9); below that within rounding of the residual
    big = hist / hist[0] > 1e-9
    assert [r[1] for r, g in zip(rows, big) if g] == ["%8.3e" % h for h, g in zip(hist, big) if g]
    assert np.allclose([float(r[1]) for r in rows], hist, rtol=2e-3, atol=1e-12 * hist[0])
    assert any(ln.startswith("Solver converged after %5d iterations" % it) for ln in out.splitlines())
    x = _read_vtk(tmp_path / "solution.vtk")["numerical"].reshape(33, 33)[1:-1, 1:-1].reshape(-1)
    assert np.abs(x - x_ref).max() <= 1e-5 * np.abs(x_ref).max()
    if ordering == "lexicographic":
        rho = np.array([float(r[3]) for r in rows[1:]])
        print("driver_mg lexicographic rho:", rho)


def test_driver_mg_solves_and_reports(tmp_path):
    rc, out = run_driver("driver_mg", tmp_path)
    assert rc == 0, out[-3000:]
    rows = np.array([[float(v) for v in r] for r in _residual_table(out)])
    assert rows[-1, 2] < 1e-10 and np.all(rows[2:, 3] < 0.3)
    assert any(ln.startswith("Solver converged after") for ln in out.splitlines())
    x = _read_vtk(tmp_path / "solution.vtk")["numerical"].reshape(33, 33)[1:-1, 1:-1].reshape(-1)
    # the solution solves Q x = b for the driver's standard-normal right-hand side
    op, _ = _problem()
    s = mg.MultigridMCSampler(op, mg.MultigridParameters(nlevel=3))
    r = s.apply_operator(x)
    assert 0.8 < r.std() < 1.2 and abs(r.mean()) < 0.15
