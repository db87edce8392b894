"""Python side of the MI355X Yee FDTD backend: problem files, the C-ABI binding and a
driver that runs a problem on the GPU.  The product host is the Node.js backend in
../js; this package is what the test-suite and the benchmark use to reach the same
native library."""

from .problem import Problem, read_problem, build_problem, load_scenario_problem  # noqa: F401
from . import native  # noqa: F401
