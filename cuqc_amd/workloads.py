"""The five benchmark workloads: generator config, seed, mining parameters and
the search window each one is measured on (a negative root_begin counts
branches back from the last one).

Exact maximal-quasi-clique enumeration with the reference's algorithm is
exponential in the size of dense regions; on configs 1-4 the full search
runs for hours on one CPU core (and far beyond the time a benchmark may hold
the GPU).  Each workload is therefore pinned to a deterministic window of the
real search: the root branches [root_begin, root_end) in the reference's
branch order (main.cu:1094) and, where needed, a depth limit.  A window is a
pure function of the graph, so the oracle and the GPU expand exactly the same
task records inside it and their outputs are compared bit for bit.
"""
from __future__ import annotations

from dataclasses import dataclass


@dataclass(frozen=True)
class Workload:
    name: str
    config: int
    seed: int
    root_begin: int = 0
    root_end: int = -1      # -1: all root branches
    max_depth: int = -1     # -1: no depth limit
    note: str = ""
    instances: int = 1      # > 1: a batch of generated instances mined together

    @property
    def window(self) -> tuple[int, int, int]:
        return (self.root_begin, self.root_end, self.max_depth)

    @property
    def full(self) -> bool:
        return self.root_begin == 0 and self.root_end < 0 and self.max_depth < 0


WORKLOADS: dict[int, Workload] = {
    0: Workload("er200-planted", 0, 1, note="full search"),
    1: Workload("coexpr8k", 1, 2, 0, -1, 1, note="all 4163 root branches, depth <= 1"),
    2: Workload("chunglu100k", 2, 3, -32, -1, 2, note="last 32 of 13087 root branches, depth <= 2"),
    3: Workload("batch256", 3, 4, -32, -1, 2, instances=256,
                note="256 instances, each: its last 32 root branches, depth <= 2"),
    4: Workload("chunglu1m", 4, 5, -8, -1, 2, note="last 8 of 66981 root branches, depth <= 2"),
}


def graph(w: Workload, instance: int = 0):
    from . import Graph
    return Graph.generate(w.config, w.seed, instance)


def graphs(w: Workload) -> list:
    """Every instance of the workload (one graph unless it is a batch)."""
    return [graph(w, i) for i in range(w.instances)]
