"""The C-ABI library loads and exports every entry point declared in
include/tsx.h (no device work)."""

import ctypes
import os
import re

import torch_sim_amd as tsa

HEADER = os.path.join(os.path.dirname(os.path.dirname(__file__)), "include", "tsx.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|int64_t|const char\*)\s+(tsx_\w+)\(", text, re.M)))


def test_header_declares_entry_points():
    names = declared()
    assert "tsx_nl_build" in names and "tsx_pair_forces" in names and len(names) >= 15


def test_library_exports_every_declared_symbol():
    lib = tsa._lib.lib()
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_table_covers_header():
    assert set(declared()) <= set(tsa._lib._SIGS), set(declared()) - set(tsa._lib._SIGS)


def test_host_side_queries():
    lib = tsa._lib.lib()
    assert lib.tsx_version() >= 100
    assert lib.tsx_nl_ws_bytes(1000, 4, 100) > 1000 * 4 * 8
    assert lib.tsx_scan_ws_bytes(10**6) >= 8 * 2 * (10**6 // 2048)
    assert lib.tsx_last_error() is not None


def _param_counts():
    text = open(HEADER).read()
    out = {}
    for m in re.finditer(r"^(?:int|int64_t|const char\*)\s+(tsx_\w+)\(([^;]*)\);", text, re.M | re.S):
        args = m.group(2).strip()
        out[m.group(1)] = 0 if args in ("", "void") else args.count(",") + 1
    return out


def test_binding_arity_matches_header():
    counts = _param_counts()
    bad = {n: (c, len(tsa._lib._SIGS[n][1])) for n, c in counts.items()
           if len(tsa._lib._SIGS[n][1]) != c}
    assert not bad, bad


def _param_types():
    """ctypes code of every parameter of every header declaration."""
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for m in re.finditer(r"^(?:int|int64_t|const char\*)\s+(tsx_\w+)\(([^;]*)\);", text, re.M | re.S):
        args = m.group(2).strip()
        codes = []
        if args not in ("", "void"):
            for a in args.split(","):
                a = " ".join(a.split())
                if "*" in a:
                    codes.append("P")
                elif a.startswith("int32_t") or a.startswith("int "):
                    codes.append("I")
                elif a.startswith("int64_t"):
                    codes.append("L")
                elif a.startswith("uint64_t"):
                    codes.append("U")
                elif a.startswith("double"):
                    codes.append("D")
                else:
                    codes.append("?" + a)
        out[m.group(1)] = codes
    return out


def test_binding_types_match_header():
    import ctypes

    code = {ctypes.c_void_p: "P", ctypes.c_int32: "I", ctypes.c_int64: "L", ctypes.c_uint64: "U",
            ctypes.c_double: "D"}
    bad = {}
    for name, want in _param_types().items():
        got = [code.get(t, "?") for t in tsa._lib._SIGS[name][1]]
        if got != want:
            bad[name] = (want, got)
    assert not bad, bad


def test_fused_step_refuses_lists_past_its_32bit_offsets():
    """The fused step checks its arguments before any device work: a list
    whose word offsets (8 row strides of ld * tpa words) or record offsets
    (16 bytes per atom) pass 2^32 bytes, or an ld below n_atoms, returns
    TSX_ERR_ARG with a message instead of launching (tsx.h, fused step)."""
    lib = tsa._lib.lib()
    buf = [ctypes.create_string_buffer(64) for _ in range(8)]
    p = [ctypes.addressof(b) for b in buf]

    def call(n_atoms, ld, tpa):
        # kind Morse, one species, NVE; pointers are never dereferenced
        return lib.tsx_md_fused_step(
            1, p[0], 1, p[1], p[2], None, None, p[3], p[3], p[4], p[3], p[3], n_atoms, 11, tpa,
            ld, 6.0, p[5], p[6], p[7], p[7], None, 0.002, 0, None, None, None, 0, None, 0, 1,
            0.3, None, p[1], p[2], p[4], p[5], 1, 0, None)

    assert call(10, 2**27, 2) == -1
    assert "too large" in lib.tsx_last_error().decode()
    assert call(10, 5, 2) == -1
    assert "ld must be at least n_atoms" in lib.tsx_last_error().decode()
    assert call(2**28, 2**28, 1) == -1
