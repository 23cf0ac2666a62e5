"""Shared fixtures: built libraries, seeded synthetic inputs and oracle outputs."""
from __future__ import annotations

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

import oracle_lib  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def built():
    """The product library and the oracle; builds them when missing."""
    import gsnp_amd

    if not gsnp_amd.LIB_PATH.exists():
        import __graft_entry__

        __graft_entry__.build()
    oracle_lib.ensure_built()
    return gsnp_amd


# (config, seed, scale, extra calling flags) of the shared data sets
DATASETS = {
    "c1": (1, 11, 1.0, {}),
    "c2w3": (2, 12, 0.06, {"refine_mode": 1}),          # 600 kb: 2 full windows + a partial one
    "stress": (6, 13, 1.0, {"refine_mode": 1, "rank_sum_mode": 1, "transition_dominant": 1}),
    "c3s": (3, 14, 0.004, {"rank_sum_mode": 1}),        # 24 small contigs
    "c5s": (5, 15, 0.0012, {"rank_sum_mode": 1}),       # 150 bp reads
    "c2x2": (2, 16, 0.0524288001, {"refine_mode": 1}),  # exactly 2 windows: an empty third window
}


class DataSet:
    def __init__(self, root: Path, name: str):
        cfg, seed, scale, flags = DATASETS[name]
        import gsnp_amd

        self.name = name
        self.prefix = root / name
        n, rl = gsnp_amd.synth(cfg, seed, scale, self.prefix)
        self.n_reads = n
        self.read_len = rl
        self.fasta = Path(str(self.prefix) + ".fa")
        self.soap = Path(str(self.prefix) + ".soap")
        snp = Path(str(self.prefix) + ".snp")
        self.dbsnp = snp if snp.exists() else None
        self.flags = dict(flags, read_length=rl)
        self._oracle = {}

    def oracle_text(self, **extra) -> Path:
        """The oracle's .cns text for this data set (cached per flag set)."""
        flags = dict(self.flags, **extra)
        key = tuple(sorted(flags.items()))
        if key not in self._oracle:
            out = Path(f"{self.prefix}.oracle{len(self._oracle)}.cns")
            oracle_lib.run(self.fasta, self.soap, out, dbsnp=self.dbsnp, **flags)
            self._oracle[key] = out
        return self._oracle[key]


@pytest.fixture(scope="session")
def datasets(built, tmp_path_factory):
    root = tmp_path_factory.mktemp("gsnp_data")
    cache = {}

    def get(name: str) -> DataSet:
        if name not in cache:
            cache[name] = DataSet(root, name)
        return cache[name]

    return get
