"""The launch-shape autotuner (mgbrain_amd/tune.py) on a stand-in simulator (no GPU):
the best shape names the tile size that was asked for, which rebuilds the same
tiling, not the simulator's widest tile."""
from mgbrain_amd import tune as T


class FakeSim:
    """Widest tile 1024 whatever is asked; time per step set by (tile, block, chunk)."""

    def __init__(self, tile):
        self.tile, self.block, self.chunk = tile, 0, 0
        self.steps = 0

    def run(self, n):
        self.steps = n

    def set_launch(self, block, chunk):
        self.block, self.chunk = block, chunk

    def info(self):
        cost = abs(self.tile - 430) + abs(self.block - 1024) / 64 + abs(self.chunk - 8)
        return {"tile_size": 1024, "last_run_ms": self.steps * (1.0 + cost) * 1e-3}

    def close(self):
        pass


def test_autotune_reports_the_requested_tile_size():
    best, res = T.autotune(FakeSim, [200, 430, 900], blocks=(512, 1024), chunks=(4, 8, 16), steps=40, warmup=4)
    assert len(res) == 3 * 2 * 3
    assert (best["tile_size"], best["block"], best["chunk"]) == (430, 1024, 8)
    assert best["widest"] == 1024
    assert sorted({r["tile_size"] for r in res}) == [200, 430, 900]


def test_tile_candidates_respect_the_ring_and_the_word_range():
    # 100k targets, 16-slot ring, one ring word per slot: capped by the 16-bit word range
    c = T.tile_candidates(100_000, 16)
    assert max(c) == 782 and min(c) == 49 and c == sorted(c, reverse=True)
    assert all(b <= 1024 for b in T.tile_candidates(10_000_000, 4))
    assert max(T.tile_candidates(10_000_000, 4, max_width=2048)) == 2048
    # a long ring with two words per slot (conductance rings) leaves room for fewer targets
    room = (160 * 1024 - 1024) // 4 - (3 * 1024 + 4096)
    assert max(T.tile_candidates(10_000_000, 40, rw=2)) == room // 80
    assert max(T.tile_candidates(10_000_000, 40, rw=1)) == min(1024, room // 40)
    assert T.tile_candidates(10, 16) == [32]
