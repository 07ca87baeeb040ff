"""RunConfig (src/alb_sim/config/run.py:4-36): photon counts per pass are
photons_per_batch_* x batches_*.  Not frozen, as in the reference."""
from dataclasses import dataclass

from ._meta import param


@dataclass
class RunConfig:
    photons_per_batch_forward: int = param(10_000, description="How many photons to process in each forward pass")
    photons_per_batch_backward: int = param(5_000, description="How many photons to process in each backward pass")
    batches_forward: int = param(10, description="Number of forward passes")
    batches_backward: int = param(5, description="Number of backward passes")
    processes: int = param(8, description="Number of processes to use in multiprocessing, ignored in linear execution")
