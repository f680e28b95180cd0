"""Compile checks of the C++ host mirror against the reference's own interface headers.

* ``odehip::HipState`` satisfies the reference's ``ODEState`` concept
  (src/state/ODEState.hpp:25-43) and ``ResizedODEState`` maps it (:45-56);
* the reference's tableau structs (butcher_tableaus/BTRKF45.hpp, BTDOPRI5.hpp,
  BTRKF78.hpp) are accepted as they are by ``odehip::RKEmbeddedParallel`` -- host-driven
  and device-resident ``integrate`` instantiate with them -- and by ``solve_ensemble``.

Compile only (g++ -fsyntax-only): nothing runs and no GPU is needed.  Skipped where the
reference tree is absent (it is read in place, never copied).
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/src"

PROGRAM = r"""
#include "ODEState.hpp"
#include "BTDOPRI5.hpp"
#include "BTRKF45.hpp"
#include "BTRKF78.hpp"
#include "odehip.hpp"

static_assert(ODEState<odehip::HipState<double, 12>>);  // FP64 states only, as the library
static_assert(std::same_as<ResizedODEState<odehip::HipState<double, 12>, 84>,
                           odehip::HipState<double, 84>>);
static_assert(odehip::ReferenceTableau<BTRKF45> && odehip::ReferenceTableau<BTDOPRI5> &&
              odehip::ReferenceTableau<BTRKF78>);
static_assert(!odehip::LibraryTableau<BTRKF45> && odehip::LibraryTableau<odehip::BTRKF45>);

using VdP = odehip::VanDerPolODE<double>;
[[maybe_unused]] auto host_driven =
    &odehip::RKEmbeddedParallel<BTRKF45>::integrate<VdP, odehip::RawOutput<2>, 2>;
[[maybe_unused]] auto on_device =
    static_cast<odehip::DeviceSolveStats (*)(odehip::HipState<double, 2>&, double, double,
                                             const odehip::HipState<double, 2>&,
                                             const odehip::HipState<double, 2>&, const VdP&,
                                             odehip::HipExecutor&, std::vector<double>*, long,
                                             bool)>(
        &odehip::RKEmbeddedParallel<BTRKF78>::integrate_on_device<VdP, 2>);
[[maybe_unused]] auto ensemble = &odehip::solve_ensemble<BTDOPRI5>;
"""


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree absent")
def test_mirror_satisfies_reference_concepts(tmp_path):
    src = tmp_path / "concepts.cpp"
    src.write_text(PROGRAM)
    inc = [os.path.join(REF, d) for d in
           ("state", "integrator/RKEmbeddedParallel/butcher_tableaus")]
    cmd = (["g++", "-std=c++20", "-fsyntax-only", "-Wall", "-Wextra"] +
           [f"-I{d}" for d in inc] +
           [f"-I{os.path.join(ROOT, 'include')}",
            f"-I{os.path.join(ROOT, 'ode_prototyping_amd', 'cpp')}", str(src)])
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr[-4000:]
