"""Occupancy groups of the high cells (CPU only): group counts through the C
ABI against enumeration, the partition of the state, argument checks, and the
per-group cell body run on the host against the oracle's full-state cells."""
from __future__ import annotations

import shutil
import subprocess

import numpy as np
import pytest

import ofp
from conftest import ROOT
from group_masks import group_masks

HIPCC = "/opt/rocm/bin/hipcc"
SRC = ROOT / "tests" / "native" / "group_host_check.cpp"


@pytest.mark.parametrize("n,low,high", [(5, 2, False), (9, 4, False), (13, 6, False), (13, 1, False),
                                         (20, 10, False), (27, 14, False), (27, 20, False), (5, 2, True),
                                         (9, 4, True), (13, 11, True), (20, 10, True), (27, 14, True),
                                         (27, 6, True)])
def test_groups_partition_the_state(n, low, high):
    nm, nb = ofp.state_counts(n)
    fixed = n - low if high else low
    info, mc, bc = ofp.group_plan(n, 4, 0, 1 << fixed, low, high)
    assert int(mc.sum()) == nm and int(bc.sum()) == nb
    assert info["main_count"] == nm and info["blocked_count"] == nb
    assert info["largest_group"] == int((mc + bc).max())
    assert info["cells"] == (low if high else n - low) and info["low"] == low
    # windows of consecutive masks add up
    half = 1 << (fixed - 1)
    a, _, _ = ofp.group_plan(n, 4, 0, half, low, high)
    b, _, _ = ofp.group_plan(n, 4, half, half, low, high)
    assert a["main_count"] + b["main_count"] == nm and a["blocked_count"] + b["blocked_count"] == nb


def test_n27_groups_fit_32bit_ranks():
    """The W=28 LOW14 / HIGH13 split (16,384 and 8,192 groups, none empty):
    every group ranks in 32 bits."""
    info, mc, bc = ofp.group_plan(27, 4, 0, 1 << 14, 14)
    assert int(mc.max()) == 961466716 and int(mc.max()) < 2**32 and int(bc.max()) < 2**32
    assert int((mc > 0).sum()) == 1 << 14
    assert info["state_bytes"] == (info["main_count"] + info["blocked_count"]) * 16 + (1 << 14) * 15360
    info, mc, bc = ofp.group_plan(27, 4, 0, 1 << 13, 14, high=True)
    assert int(mc.max()) == 1471935235 and int(bc.max()) == 510468519
    assert int((mc > 0).sum()) == 1 << 13


@pytest.mark.parametrize("n,low,high", [(9, 4, False), (11, 3, False), (12, 7, False), (9, 4, True),
                                         (11, 3, True), (12, 7, True)])
def test_group_counts_match_enumeration(oracle, n, low, high):
    nm, nb = ofp.state_counts(n)
    fixed = n - low if high else low
    _, mc, bc = ofp.group_plan(n, 1, 0, 1 << fixed, low, high)
    lib = oracle.load()
    mm, mb = group_masks(n, low, high)
    assert np.array_equal(np.bincount(mm, minlength=1 << fixed), mc.astype(np.int64))
    assert np.array_equal(np.bincount(mb, minlength=1 << fixed), bc.astype(np.int64))
    # the numpy unrank agrees with the oracle's on a sample
    lo = low + 1 if high else 0
    for i in range(0, nm, max(1, nm // 97)):
        w = lib.or_unrank(i, n + 1)
        assert mm[i] == sum(1 << (p - lo) for p in range(lo, lo + fixed) if (w >> (2 * p)) & 3)


@pytest.mark.parametrize("kw,msg", [({"low": -1}, "low"), ({"low": 9}, "low"), ({"low": 26, "n": 27}, "low"),
                                    ({"first": 16, "groups": 1}, "groups"), ({"first": 8, "groups": 9}, "groups")])
def test_group_plan_refuses(kw, msg):
    n, low = kw.get("n", 9), kw.get("low", 4)
    with pytest.raises(ofp.OfpError) as e:
        ofp.group_plan(n, 1, kw.get("first", 0), kw.get("groups", 1), low)
    assert e.value.status == -1


@pytest.fixture(scope="module")
def checker(tmp_path_factory, oracle):
    if not shutil.which(HIPCC) and not (ROOT / HIPCC).exists():
        pytest.skip("hipcc not available")
    out = tmp_path_factory.mktemp("group") / "group_host_check"
    lib = ROOT / "oracle"
    cmd = [HIPCC, "-O1", "-std=c++17", "-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           str(SRC), "-o", str(out), f"-L{lib}", "-loracle", f"-Wl,-rpath,{lib}"]
    subprocess.run(cmd, check=True, capture_output=True, timeout=600)
    return out


def test_group_body_equals_oracle_cells(checker):
    """Every group of every split at n = 2..8, each swept alone on group-local
    ranks, reproduces the oracle's full-state cells: the high cells over LOW
    groups, the low cells (with blocked input) over HIGH groups -- one cell per
    launch, and as pair passes with each group's own class tables (classes in
    rank order and reversed, every launch's out-of-group guard intact)."""
    res = subprocess.run([str(checker), "8"], capture_output=True, text=True, timeout=600)
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-2000:]
    assert res.stdout.strip().endswith("ALL OK")
    assert res.stdout.count(" low V=") >= 40 and res.stdout.count(" high V=") >= 40
    assert res.stdout.count(" pairs groups") >= 50 and res.stdout.count(" pairs reversed") >= 50


def test_group_body_at_full_size(checker):
    """Single groups of the n = 27 LOW14 / HIGH13 split at their real size (and
    of n = 13): the depth-first enumeration of the group comes out in global
    rank order, and the body on group-local ranks equals the oracle's
    transitions applied to the group's words alone, cell by cell (27..15 over
    LOW groups, 14..1 with blocked input over HIGH groups), one cell per launch
    and as pair passes with the group's class tables."""
    for mode, args in (("big", ["13", "6", "5", "17", "63"]), ("big", ["27", "14", "3", "100"]),
                       ("bighigh", ["13", "6", "0", "5", "9"]), ("bighigh", ["27", "14", "0", "4"]),
                       ("bigpair", ["13", "6", "5", "63"]), ("bigpair", ["27", "14", "3", "100"]),
                       ("bighighpair", ["13", "6", "0", "9"]), ("bighighpair", ["27", "14", "0", "4"])):
        res = subprocess.run([str(checker), mode, *args], capture_output=True, text=True, timeout=600)
        assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-2000:]
        assert res.stdout.strip().endswith("ALL OK")
        assert res.stdout.count("ok big") == len(args) - 2
