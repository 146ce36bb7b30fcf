"""MI355X (gfx950) triplet-seeding library: the grid build, pair generation,
triplet prediction and filtering, connectivity and cellular automaton of the
trax track finder as hand-written HIP kernels behind a C ABI
(include/trax_hip.h), with C++ host classes named after the reference's and a
batched, graph-captured pipeline plus a serving loop."""
from .config import Workload, workload
from .generator import Event, make_event, make_events
