"""The C ABI library: it loads, exports every entry point include/ofp.h declares,
and answers the device-free queries.  Runs without a GPU (no compute calls)."""
from __future__ import annotations

import re
import subprocess

import pytest

from conftest import PKG, ROOT

HEADER = ROOT / "include" / "ofp.h"
CLI = PKG / "lib" / "ofp_solver"


def declared_functions() -> list[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.findall(r"\b(ofp_[a-z_]+)\s*\(", text)


def test_header_declarations_are_exported():
    import ofp
    lib = ofp.load()
    names = declared_functions()
    assert len(names) >= 12
    for name in names:
        assert hasattr(lib, name), f"libofp.so does not export {name}"
    assert sorted(set(names)) == sorted(ofp.EXPORTS)


def test_state_counts_match_oracle(oracle):
    import ofp
    lib = oracle.load()
    for n in range(1, 28):
        assert ofp.state_counts(n) == (lib.or_main_count(n), lib.or_blocked_count(n))
    for bad in (0, -1, 28):
        with pytest.raises(ofp.OfpError):
            ofp.state_counts(bad)


def test_version_string():
    import ofp
    assert "gfx950" in ofp.version()


def test_context_without_gpu_fails_cleanly():
    import ofp
    try:
        ctx = ofp.Context()
    except ofp.OfpError as e:
        assert e.status == -2  # OFP_E_DEVICE: no usable GPU here
    else:  # on a GPU host the context simply works
        ctx.close()


@pytest.mark.parametrize("argv,msg", [
    (["abc"], "n must be"),
    (["0"], "n must be"),
    (["28"], "n must be"),
    (["12", "x"], "target_mib"),
    (["12", "16384", "0"], "max_window"),
    (["12", "16384", "14", "-3"], "ngpu"),
    (["12", "16384", "14", "1", "1"], "modulus must be in [2, 4294967295]"),
    (["12", "16384", "14", "1", "4294967296"], "modulus must be in [2, 4294967295]"),
    (["12", "16384", "14", "1", "12abc"], "modulus must be in [2, 4294967295]"),
])
def test_cli_rejects_bad_arguments(argv, msg):
    """Argument errors exit 1 before any device work (row6crt20.cu:1697-1704)."""
    res = subprocess.run([str(CLI), *argv], capture_output=True, text=True, timeout=60)
    assert res.returncode == 1
    assert msg in res.stderr
    assert res.stdout == ""


def test_read_main_takes_a_capacity():
    """ofp_read_main carries the caller's buffer size, so a buffer shorter than the
    prepared moduli count is refused instead of overrun (the GPU side of this is
    tests/test_gpu_parity.py::test_read_main_refuses_short_buffer)."""
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    decl = re.search(r"ofp_read_main\s*\(([^)]*)\)", text).group(1)
    assert [a.split()[-1].lstrip("*") for a in decl.split(",")] == ["ctx", "index", "out", "capacity"]


@pytest.mark.parametrize("k,g,want", [
    (14, 14, [14]),
    (14, 16, [14]),
    (15, 8, [8, 7]),
    (15, 9, [8, 7]),
    (14, 8, [8, 6]),
    (14, 7, [7, 7]),
    (14, 6, [6, 6, 2]),
    (20, 6, [6, 6, 6, 2]),
    (13, 4, [4, 4, 4, 1]),
    (5, 2, [2, 2, 1]),
    (1, 1, [1]),
    (7, 1, [1] * 7),
    (9, 5, [5, 4]),
])
def test_moduli_groups(k, g, want):
    """Moduli that do not fit one sweep: fewest sweeps of at most g, near-equal
    sizes, even sizes where that adds no sweep."""
    import ofp
    got = ofp.moduli_groups(k, g)
    assert got == want
    assert sum(got) == k and max(got) <= g
    assert len(got) == -(-k // g)


def test_moduli_groups_rejects_bad_arguments():
    import ofp
    for k, g in ((0, 4), (4, 0), (-1, 2)):
        with pytest.raises(ofp.OfpError):
            ofp.moduli_groups(k, g)
