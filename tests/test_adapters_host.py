"""The adapter classes on the CPU: the header binds its vector views to any
type with data() and size() (as Eigen::VectorXd has), so a reference-side
subclass can pass its vectors unchanged; the built library and example
program are in place (no device work)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(os.path.dirname(HERE), "multigridmc_amd")

SNIPPET = r"""
#include "mgmc_adapters.hh"
// a vector type shaped like Eigen::VectorXd (Index size(), data())
struct VectorXdLike {
  double* d; long n;
  double* data() { return d; }
  const double* data() const { return d; }
  long size() const { return n; }
};
struct ReferenceStyleSampler {
  std::shared_ptr<mgmc_amd::Sampler> impl;
  void apply(const VectorXdLike& f, VectorXdLike& x) const { impl->apply(f, x); }
  void fix_rhs(const VectorXdLike& f) { impl->fix_rhs(f); }
};
struct ReferenceStylePreconditioner {
  std::shared_ptr<mgmc_amd::Preconditioner> impl;
  void apply(const VectorXdLike& b, VectorXdLike& x) { impl->apply(b, x); }
};
int main() { return 0; }
"""


def test_adapter_header_binds_eigen_shaped_vectors(tmp_path):
    src = tmp_path / "use.cc"
    src.write_text(SNIPPET)
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I",
                        os.path.join(PKG, "csrc", "adapters"), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_adapter_library_and_example_built():
    assert os.path.exists(os.path.join(PKG, "libmgmc_adapters.so"))
    assert os.access(os.path.join(PKG, "bin", "adapter_demo"), os.X_OK)
