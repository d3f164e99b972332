"""Kernel parity on the CPU: the library built with -DMGMC_EMULATION (its
kernels run on the host through hip_emu.h) against the oracle, FP64, for
the operator, one sweep per level and direction (deterministic and with
injected noise), residual + restriction, prolongation and the coarsest-level
solve and sample. Run by tests/test_host_sanitizers.py in a subprocess with
the emulated library's path as argument; prints one line per case."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

import multigridmc_amd._lib as L  # noqa: E402

L.LIB_PATH = sys.argv[1]  # the emulated build

import multigridmc_amd as mg  # noqa: E402
from gpu_common import make_problem, oracle_mg, params  # noqa: E402


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def main():
    worst = 0.0
    for dim, n, nmeas, nlevel in [(3, (16, 16, 16), 3, 3), (3, (12, 16, 8), 2, 2), (2, (24, 20), 3, 2)]:
        op, ora = make_problem(dim, n, nmeas)
        p = params(nlevel)
        K = 2
        s = mg.MultigridMCSampler(op, p, nchains=K, precision="fp64")
        omg = oracle_mg(ora, p)
        rng = np.random.default_rng(1)
        x = rng.standard_normal((K, op.lattice.Nvertex))
        errs = {"apply": rel(s.apply_operator(x)[1], ora.apply(x[1]))}
        for level in range(nlevel):
            N = s.ndof(level)
            b, xx, xi = rng.standard_normal((3, K, N))
            xl = rng.standard_normal((K, max(nmeas, 1)))
            for fw in (True, False):
                got = s.sweep(level, fw, False, b, xx)
                errs[f"sweep{level}{'f' if fw else 'b'}"] = max(
                    rel(got[c], omg.sor_smoother(level, fw, b[c], xx[c])) for c in range(K))
                got = s.sweep(level, fw, True, b, xx, xi=xi, xi_lr=xl[:, :nmeas] if nmeas else None)
                errs[f"gibbs{level}{'f' if fw else 'b'}"] = max(
                    rel(got[c], omg.sor_sampler_noise(level, fw, b[c], xx[c], xi[c], xl[c, :nmeas] if nmeas else None))
                    for c in range(K))
        for level in range(nlevel - 1):
            N, Nc = s.ndof(level), s.ndof(level + 1)
            f, xx, xc = rng.standard_normal((K, N)), rng.standard_normal((K, N)), rng.standard_normal((K, Nc))
            got = s.residual_restrict(level, f, xx)
            errs[f"restrict{level}"] = rel(got[0], omg.restrict(level, f[0] - omg.level_apply(level, xx[0])))
            got = s.prolongate_add(level, 0.7, xc, xx)
            errs[f"prolong{level}"] = rel(got[0], omg.prolongate_add(level, 0.7, xc[0], xx[0]))
        Nc = s.ndof(nlevel - 1)
        f, xi = rng.standard_normal((K, Nc)), rng.standard_normal((K, Nc))
        errs["coarse_solve"] = rel(s.coarse(f, xi_mode=0)[0], omg.coarse_solve(f[0]))
        errs["coarse_sample"] = rel(s.coarse(f, xi_mode=2, xi=xi)[1], omg.coarse_sample_noise(f[1], xi[1]))
        case_worst = max(errs.values())
        worst = max(worst, case_worst)
        print(f"case {dim}d {n} m={nmeas} levels={nlevel}: worst relative error {case_worst:.2e} "
              f"({max(errs, key=errs.get)})", flush=True)
    print(f"worst {worst:.3e}")


if __name__ == "__main__":
    main()
