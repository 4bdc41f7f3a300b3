"""The launch-shape autotuner: its cache and lookup rules on the CPU, and one
small tuning run on the GPU that must leave every residue unchanged."""
from __future__ import annotations

import json

import pytest


def test_cache_round_trip(tmp_path):
    from ofp import autotune
    path = tmp_path / "tune.json"
    assert autotune.load_cache(path) == {}
    autotune.save_cache({"n=12,k=3": {"threads": 128, "blocks_per_cu": 8}}, path)
    assert json.loads(path.read_text()) == {"n=12,k=3": {"threads": 128, "blocks_per_cu": 8}}
    assert autotune.load_cache(path)["n=12,k=3"]["threads"] == 128
    path.write_text("{not json")
    assert autotune.load_cache(path) == {}  # a damaged cache reads as empty


def test_lookup_prefers_user_cache_then_packaged(tmp_path):
    from ofp import autotune
    path = tmp_path / "tune.json"
    packaged = autotune.load_cache(autotune.PACKAGED)
    assert "n=22,k=14" in packaged and "n=22,k=1" in packaged
    # exact packaged key
    assert autotune.lookup(22, 14, path) == packaged["n=22,k=14"]
    # the user cache wins over the packaged table
    autotune.save_cache({"n=22,k=14": {"threads": 64, "blocks_per_cu": 4}}, path)
    assert autotune.lookup(22, 14, path)["threads"] == 64


def test_lookup_falls_back_to_nearest_n_same_kind(tmp_path):
    """With no exact entry: the packaged entry whose moduli count is of the same
    kind (one modulus, or several), nearest in n, then in k."""
    from ofp import autotune
    path = tmp_path / "none.json"
    packaged = autotune.load_cache(autotune.PACKAGED)
    keys = [(int(key.split(",")[0][2:]), int(key.split(",")[1][2:])) for key in packaged]

    def nearest(n, k):
        c = min((abs(kn - n), abs(kk - k), kn, kk) for kn, kk in keys if (kk > 1) == (k > 1))
        return packaged[f"n={c[2]},k={c[3]}"]
    for n, k in ((3, 1), (16, 1), (16, 7), (27, 4), (21, 2), (13, 3), (23, 12)):
        want = packaged.get(f"n={n},k={k}") or nearest(n, k)
        assert autotune.lookup(n, k, path) == want, (n, k)


def test_lookup_ties_in_n_go_to_nearest_k(tmp_path):
    from ofp import autotune
    path = tmp_path / "none.json"
    pk = tmp_path / "packaged.json"
    autotune.save_cache({"n=20,k=2": {"threads": 1}, "n=20,k=8": {"threads": 2}, "n=20,k=1": {"threads": 3}}, pk)
    old = autotune.PACKAGED
    try:
        autotune.PACKAGED = pk
        assert autotune.lookup(20, 7, path)["threads"] == 2
        assert autotune.lookup(21, 3, path)["threads"] == 1
        assert autotune.lookup(19, 1, path)["threads"] == 3
    finally:
        autotune.PACKAGED = old


def test_pair_mask_only_from_an_entry_of_the_same_n(tmp_path):
    """A tuned pair mask names cell positions: it is used for its own (n, k) and not
    carried to another n by the nearest-entry fallback (0 = the library default)."""
    from ofp import autotune
    path = tmp_path / "user.json"
    autotune.save_cache({"n=19,k=3": {"threads": 256, "pair_mask": 0x1234}}, path)
    assert autotune.pair_mask_for(19, 3, path) == 0x1234
    packaged = autotune.load_cache(autotune.PACKAGED)
    assert autotune.pair_mask_for(22, 14, path) == int(packaged["n=22,k=14"].get("pair_mask", 0))
    assert autotune.lookup(23, 8, path) is not None  # launch settings fall back to n=22 ...
    assert autotune.pair_mask_for(23, 8, path) == 0  # ... the pair mask does not
    assert autotune.pair_mask_for(27, 4, path) == 0
    assert autotune.pair_mask_for(19, 5, path) == 0x1234  # same n, another moduli count


@pytest.mark.gpu
def test_tiny_tune_keeps_residues(gpu_ctx, tmp_path):
    """A reduced tuning run at n=12 (every stage, a few candidates each) writes its
    cache entry, and solves with the chosen settings give the same residues."""
    from ofp import autotune
    mods = [4294967291, 4294966997, 2147483647]
    want, _ = gpu_ctx.solve(12, mods)
    path = tmp_path / "tune.json"
    print("case autotune n=12 k=3", flush=True)
    res = autotune.tune(gpu_ctx, 12, mods, rows=1, threads=(256,), blocks_per_cu=(8, 32), layouts=(1,),
                        path=path)
    ent = autotune.load_cache(path)["n=12,k=3"]
    assert ent["threads"] == 256 and ent["blocks_per_cu"] in (8, 32)
    assert res["best"]["cap_percent"] in autotune.CAP_PERCENTS
    assert res["best"]["pair_mode"] in (0, 1) and "pair_mask" in ent
    got, _ = gpu_ctx.solve(12, mods)
    assert got == want
    # restore the library defaults for later tests
    gpu_ctx.set_launch(0, 0)
    gpu_ctx.set_variant(2)
    gpu_ctx.set_low_cells(0, 16)
    gpu_ctx.set_xcd_pieces(5)
    gpu_ctx.set_sorted_cells(-1)
    gpu_ctx.set_height_caps(True)
    gpu_ctx.set_pair_passes(1)
