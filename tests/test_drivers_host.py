"""Configuration handling of the command-line drivers (the reference's
libconfig parameter and measurement files, src/auxilliary/parameters.cc):
everything here happens before any device work."""
import pytest

from driver_common import fixture_text, run_driver


def test_parameters_are_read_and_echoed(tmp_path):
    rc, out = run_driver("driver_mgmc", tmp_path)
    for line in ["  dimension = 2", "  operator = posterior", "  lattice size = 32 x 32 x 32", "    levels = 3",
                 "    cycle = 1 ( V-cycle )", "  timeseries: number of samples         = 2000",
                 "  convergence test: number of samples   = 512", "    Lambda = 0.3",
                 "    number of measurement points = 4", "    radius of individual measurements = 0.05",
                 "  gpu: chains = 1, convergence chains per batch = 256, precision = fp64, ordering = multicolour, seed = 7"]:
        assert line in out.splitlines(), line


def test_libconfig_syntax_variants(tmp_path):
    cfg = fixture_text("posterior_2d.cfg").replace(
        'gpu = {\n    nchains = 1;\n    convergence_chains = 256;\n    precision = "fp64";\n    seed = 7;\n}',
        'gpu : { nchains = 2, convergence_chains = 0x20; precision = "fp" "32"; seed = 12L; };')
    assert "0x20" in cfg
    rc, out = run_driver("driver_mgmc", tmp_path, cfg_text=cfg)
    assert ("  gpu: chains = 2, convergence chains per batch = 32, precision = fp32, ordering = multicolour, "
            "seed = 12") in out


@pytest.mark.parametrize("old,new,message", [
    ('smoother = "SSOR";\n    coarse_solver', 'smoother = "Jacobi";\n    coarse_solver', "invalid multigrid smoother"),
    ('pdemodel = "shiftedlaplace_fd";\n    correlationlengthmodel = "constant"',
     'pdemodel = "shiftedlaplace_fem";\n    correlationlengthmodel = "periodic"', "finite-difference prior only"),
    ("measure_global = false", "measure_global = true", "measure_global"),
    ('pdemodel = "shiftedlaplace_fd"', 'pdemodel = "squared_shiftedlaplace_fd"', "squared shifted-Laplace"),
    ("    nlevel = 3;\n", "", "cannot read group 'multigrid'"),
    ('operator = "posterior"', 'operator = "likelihood"', "operator has to be"),
    ("dim = 2;", "dim = ;", "value expected"),
    ('sample_location = [0.45, 0.55];', 'sample_location = [0.45, {}];', "arrays hold scalars only"),
])
def test_configuration_errors(tmp_path, old, new, message):
    cfg = fixture_text("posterior_2d.cfg")
    assert old in cfg
    rc, out = run_driver("driver_mgmc", tmp_path, cfg_text=cfg.replace(old, new, 1))
    assert rc != 0
    assert "ERROR" in out and message in out


def test_measurement_file_errors(tmp_path):
    meas = fixture_text("measurements_2d.cfg").replace("dim =  2 ;", "dim =  3 ;")
    rc, out = run_driver("driver_mgmc", tmp_path, meas_text=meas)
    assert rc != 0 and ("fewer entries" in out or "dimension of measurement locations" in out)
    rc, out = run_driver("driver_mg", tmp_path, meas_text="dim = 2;\nn = 4;\n")
    assert rc != 0 and "measurements" in out
