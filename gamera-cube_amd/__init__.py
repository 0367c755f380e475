"""MI355X-native six-face GAMERA-Cube step: host-side package.

`config` describes a problem, `native.GpuProblem` runs it on the GPU through
libgcs.so, and `problem.Problem` wraps any library exporting the same
create/advance/field API.
"""
from . import config  # noqa: F401
from . import problem  # noqa: F401
