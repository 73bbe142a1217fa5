"""The `pgd` driver cases shared by the driver test and its fixture writer."""
from __future__ import annotations

# the reference prints its wall-clock time on this line; it is the only line
# that differs between runs
TIMING_PREFIX = "graphlet decomposition time:"

CASES = [
    ("4-clique-and-edge.txt", []), ("4-chordal-cycle.txt", []), ("DD21.mtx", []), ("DD21.mtx", ["-o", "kcore"]),
    ("DD21.mtx", ["--s2l", "-n", "kcore"]), ("sample_graph.csv", []), ("rd_csv_header.csv", []),
    ("rd_weighted.txt", []), ("rd_triple_space.txt", []), ("rd_gap_ids.mtx", []), ("rd_tab.edges", []),
    ("C250-9.mtx", []), ("gnp_2000_0.01", ["--s2l_neigh"]), ("ba_3000_4", ["-o", "natural"]),
    # the representation: csc -> the degree-oriented engine, adj and hybrid ->
    # the dense tier, both with the reference's counts and files
    ("DD21.mtx", ["-r", "csc"]), ("C250-9.mtx", ["-r", "adj"]), ("gnp_2000_0.01", ["-r", "hybrid"]),
    ("ba_3000_4", ["-r", "csc"]),
]


def case_id(name: str, flags: list[str]) -> str:
    return name + "".join(flags)
