"""peakpatch_hip -- MI355X (gfx950) implementation of PeakPatch.jl's
halo-finding hot path behind a C ABI (include/peakpatch_hip.h).

Modules:
  api        Context: the C-ABI front end (filter bank, stencil, shell analysis)
  reference  functions named and shaped like the Julia reference's hot-path API
  collapse   collapse-table file I/O (read_homeltab / write_homeltab)
  cosmology  host-side Float64 growth factor / chi(z) used for per-peak fcrit
  catalog    .pksc record packing and I/O
  synth      seeded synthetic inputs (Threefry noise, EH98 P(k), LPT fields)
  pipeline   tile drivers (run_tile / run_multitile_split shaped)
  serve      continuous-batching tile server
"""
from ._lib import PPError, SHELL_PERSISTENT, SHELL_PLAIN, LIB_PATH  # noqa: F401
from .api import Context, PeakList  # noqa: F401
