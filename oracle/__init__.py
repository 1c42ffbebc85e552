"""CPU FP64 oracle for the forced-DNS + droplet RK4 step.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package may import, call or
execute anything under ``oracle/``.  It is used by ``tests/`` (as the checker),
by ``__graft_entry__.smoke()`` (as the checker) and by ``bench.py``'s
``cpu_baseline`` leg (as the timed CPU baseline).  It is never the thing that
is measured on the GPU and never a fallback for the GPU path.

The oracle is a single-rank restatement (numpy + scipy.fft, float64) of the
reference's algorithm in ``forced-dns-sm.py`` and ``particles.py``.  Every
function cites the reference lines it follows.  With one MPI rank the
reference's ``Alltoall`` transposes are identities, the particle migration is a
self-send and the halo requests resolve to periodic index wrapping; those are
restated directly.

Pinning (what checks each oracle function against something other than its own
formula) is listed in ``oracle/PINNING.md``.
"""
from .flow import Flow, Grid  # noqa: F401
from .droplets import Droplets  # noqa: F401
