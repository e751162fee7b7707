"""The benchmark workloads: generator config, seed, mining parameters and
the search window each one is measured on (a negative root_begin counts
branches back from the last one).

Exact maximal-quasi-clique enumeration with the reference's algorithm is
exponential in the size of dense regions.  On config 1 a single root branch
below 3612 runs for more than 40 minutes on one CPU core (its first branch
at every level keeps nearly all candidates), and the full searches of
configs 1, 2 and 4 would run for days.  Each workload is therefore pinned to
a deterministic window of the real search: the root branches
[root_begin, root_end) in the reference's branch order (main.cu:1094) and,
where needed, a depth limit; children past the depth limit are built and
pruned but neither written nor expanded (counted as `beyond`).  A window is
a pure function of the graph, so the oracle and the GPU expand exactly the
same task records inside it and their outputs are compared bit for bit.
Windows of configs 2 and 4 hold root branches whose subtrees contain planted
quasi-cliques (Graph.planted()).
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
    first: int = 0          # generator instance of the first graph (config 3)

    @property
    def window(self) -> tuple[int, int, int]:
        return (self.root_begin, self.root_end, self.max_depth)

    @property
    def full(self) -> bool:
        return self.root_begin == 0 and self.root_end < 0 and self.max_depth < 0

    @property
    def label(self) -> str:
        """The workload with its window, e.g. coexpr8k[3600:3612,depth<=3]."""
        if self.full and self.instances == 1:
            return f"{self.name}[full]"
        end = "" if self.root_end < 0 else str(self.root_end)
        depth = "full depth" if self.max_depth < 0 else f"depth<={self.max_depth}"
        return f"{self.name}[{self.root_begin}:{end},{depth}]"


WORKLOADS: dict[int, Workload] = {
    0: Workload("er200-planted", 0, 1, note="full search"),
    1: Workload("coexpr8k", 1, 2, 3600, 3612, 3,
                note="root branches 3600-3611 of 4163, depth <= 3 (234,654 tasks; the branches below hold "
                     "subtrees beyond any search)"),
    2: Workload("chunglu100k", 2, 3, 9263, 9273, -1,
                note="root branches 9263-9272 of 13087 at full depth; two planted quasi-cliques (53 and 51 "
                     "vertices) are found in branches 9263 and 9264"),
    3: Workload("batch256", 3, 4, -32, -1, 2, instances=256,
                note="256 instances, each: its last 32 root branches, depth <= 2"),
    4: Workload("chunglu1m", 4, 5, 41828, 41846, 3,
                note="root branches 41828-41845 of 66981, depth <= 3; planted quasi-cliques of 80 vertices "
                     "in branches 41828 and 41845"),
    5: Workload("batch256-i15", 3, 4, 0, -1, -1, first=15,
                note="instance 15 of the config-3 batch (700 vertices), full search: 1.46M tasks"),
}


def graph(w: Workload, instance: int = 0):
    from . import Graph
    return Graph.generate(w.config, w.seed, w.first + instance)


def graphs(w: Workload) -> list:
    """Every instance of the workload (one graph unless it is a batch)."""
    return [graph(w, i) for i in range(w.instances)]
