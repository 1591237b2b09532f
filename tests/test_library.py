"""The C-ABI library loads and exports every entry point include/sidon.h
declares; argument validation runs before any device call.  No GPU."""
from __future__ import annotations

import ctypes
import importlib
import re

import pytest

from conftest import REPO


def declared_entry_points():
    text = (REPO / "include" / "sidon.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sidon_[a-z0-9_]+)\s*\(", text)))


def test_header_and_module_agree():
    S = importlib.import_module("sidon-autocorrelation-bound_amd")
    assert declared_entry_points() == sorted(S.EXPORTS)


def test_library_exports_every_declared_symbol():
    S = importlib.import_module("sidon-autocorrelation-bound_amd")
    L = ctypes.CDLL(str(S.LIB_PATH))
    for name in declared_entry_points():
        assert hasattr(L, name), name


def test_argument_validation_without_device():
    S = importlib.import_module("sidon-autocorrelation-bound_amd")
    L = S.lib()
    fake = ctypes.c_void_p(0x1000)
    # n < 1, null population, bad schedule -> SIDON_ERR_ARG before any HIP call
    assert L.sidon_local_search(fake, None, None, None, None, 0, 4, 10, 4.0, 8.0, 0.1, 0.01,
                                0.9, 0.999, 0, None) == S.SIDON_ERR_ARG
    assert L.sidon_local_search(None, None, None, None, None, 64, 4, 10, 4.0, 8.0, 0.1, 0.01,
                                0.9, 0.999, 0, None) == S.SIDON_ERR_ARG
    assert L.sidon_local_search(fake, None, None, None, None, 64, 4, 10, -4.0, 8.0, 0.1, 0.01,
                                0.9, 0.999, 0, None) == S.SIDON_ERR_ARG
    # the moments' buffers may not overlap the population (they can be the launch's scratch)
    fake2 = ctypes.c_void_p(0x1000 + 4 * 64 * 4 - 4)
    assert L.sidon_local_search(fake, fake, None, None, None, 64, 4, 10, 4.0, 8.0, 0.1, 0.01,
                                0.9, 0.999, 0, None) == S.SIDON_ERR_ARG
    assert L.sidon_local_search(fake, None, fake2, None, None, 64, 4, 10, 4.0, 8.0, 0.1, 0.01,
                                0.9, 0.999, 0, None) == S.SIDON_ERR_ARG
    assert L.sidon_local_search_f64(fake, fake, None, None, None, 64, 4, 10, 4.0, 8.0, 0.1, 0.01,
                                    0.9, 0.999, 0, None) == S.SIDON_ERR_ARG
    # P == 0 / iters == 0 are no-ops
    assert L.sidon_local_search(fake, None, None, None, None, 64, 0, 10, 4.0, 8.0, 0.1, 0.01,
                                0.9, 0.999, 0, None) == S.SIDON_OK
    assert L.sidon_eval_f64(fake, 0, 1, fake, None, 1e-4, None) == S.SIDON_ERR_ARG
    assert L.sidon_frank_wolfe(fake, None, 64, 1, 1, 10, 0, 0.3, None) == S.SIDON_ERR_ARG
    assert L.sidon_select(fake, None, None, 0, 10, 3, 1, 0, 0.05, 1, fake, fake, fake, fake,
                          fake, None) == S.SIDON_ERR_ARG  # P not divisible by NI
    assert L.sidon_set_ls_block(4096, 300) == S.SIDON_ERR_ARG
    assert L.sidon_set_ls_tile(4096, 512, 3, 1) == S.SIDON_ERR_ARG    # depth 2 or 4
    assert L.sidon_set_ls_tile(4096, 1024, 4, 2) == S.SIDON_ERR_ARG   # two a CU: 512 only
    assert L.sidon_set_ls_tile(600, 256, 4, 2) == S.SIDON_ERR_ARG     # four a CU: 256 only
    assert L.sidon_set_ls_tile(600, 512, 2, 4) == S.SIDON_ERR_ARG
    assert L.sidon_ls_tile_config(S.LS_TILES, None, None, None) == S.SIDON_ERR_ARG
    # the persistent launch needs a caller-owned workspace of the stated size
    assert L.sidon_local_search_ragged(fake, 64, fake, 4, 64, None, None, None, 10, 4.0, 8.0,
                                       0.1, 0.01, 0.9, 0.999, 0, 1, None, 0,
                                       None) == S.SIDON_ERR_ARG
    need = L.sidon_local_search_ragged_workspace(4, 64)
    assert need > 0
    assert L.sidon_local_search_ragged(fake, 64, fake, 4, 64, None, None, None, 10, 4.0, 8.0,
                                       0.1, 0.01, 0.9, 0.999, 0, 1, fake, need - 1,
                                       None) == S.SIDON_ERR_ARG
    assert L.sidon_local_search_ragged(fake, 64, fake, 4, 64, None, None, None, 10, 4.0, 8.0,
                                       0.1, 0.01, 0.9, 0.999, 0, 2, fake, need,
                                       None) == S.SIDON_ERR_ARG
    assert S.lib().sidon_status_string(S.SIDON_ERR_LDS).startswith(b"length needs more LDS")


def test_lds_footprints():
    S = importlib.import_module("sidon-autocorrelation-bound_amd")
    ls = S.lds_bytes(0, 8400, 1024)
    assert 130_000 < ls <= 160 * 1024          # fits one MI355X workgroup (160 KiB)
    assert S.lds_bytes(3, 8400, 1024) <= 160 * 1024
    assert S.lds_bytes(2, 8400, 256) <= 160 * 1024
    assert S.lds_bytes(1, 8400, 512) > 160 * 1024   # fp64 local search does not fit at 8400
    # above 9296 the correlation output moves to global memory: the local
    # search with the plain 16-float layout to 13216, the gradient with the
    # widest that fits (24-float segments to 11280, 20-float to 12176, 16)
    for n in (9297, 11280, 11281, 12176, 12177, 13216):
        assert 0 < S.lds_bytes(0, n, 1024) <= 160 * 1024, n
        assert 0 < S.lds_bytes(4, n, 1024) <= 160 * 1024, n
    assert S.lds_bytes(0, 13217, 1024) > 160 * 1024
    assert S.lds_bytes(4, 13217, 1024) > 160 * 1024
    assert S.lds_bytes(0, 11281, 1024) > S.lds_bytes(0, 11280, 1024)   # one layout throughout
    assert S.lds_bytes(4, 11281, 1024) < S.lds_bytes(4, 11280, 1024)   # narrower segments
    assert S.lds_bytes(4, 12176, 1024) == S.lds_bytes(0, 12176, 1024) + 4 * (
        (12176 + 15) // 16 + 32) * 4 - 16   # 20- against 16-float segments
    assert S.lds_bytes(9, 100, 256) == -1


def test_generation_desc_matches_header():
    """The ctypes mirror of sidon_generation_desc lists the header's fields in
    the header's order with the same C types."""
    S = importlib.import_module("sidon-autocorrelation-bound_amd")
    text = (REPO / "include" / "sidon.h").read_text()
    body = re.search(r"typedef struct \{(.*?)\} sidon_generation_desc;", text, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    kinds = {"float*": ctypes.c_void_p, "int*": ctypes.c_void_p, "double*": ctypes.c_void_p,
             "void*": ctypes.c_void_p,
             "unsigned long long*": ctypes.c_void_p, "int": ctypes.c_int,
             "float": ctypes.c_float, "double": ctypes.c_double,
             "unsigned long long": ctypes.c_uint64}
    want = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        m = re.match(r"((?:unsigned long long|int|float|double|void))\s*(\*?)\s*(.*)", decl)
        base, star, names = m.group(1), m.group(2), m.group(3)
        for part in names.split(","):
            part = part.strip()
            ptr = star or part.startswith("*")
            name = part.lstrip("*").strip()
            if part.startswith("*") is False and "*" in part:
                ptr = True
            want.append((name, kinds[base + ("*" if ptr else "")]))
    got = [(f[0], f[1]) for f in S.GenerationDesc._fields_]
    assert got == want


def test_tile_rules_are_host_state():
    """The launch-configuration rules live on the host: the built-in rule, an
    override per length, and its removal (no GPU call)."""
    S = importlib.import_module("sidon-autocorrelation-bound_amd")
    assert S.get_ls_tile(300) == (256, 2, 4)
    assert S.get_ls_tile(1023) == (256, 2, 4)
    assert S.get_ls_tile(1024) == (512, 4, 2)
    assert S.get_ls_tile(4096) == (512, 4, 2)
    assert S.get_ls_tile(4097) == (1024, 4, 1)
    assert S.get_ls_tile(2100) == (512, 4, 2)
    assert S.get_ls_tile(8400) == (1024, 4, 1)
    S.set_ls_tile(2100, 1024, 2, 1)
    try:
        assert S.get_ls_tile(2100) == (1024, 2, 1)
        assert S.get_ls_block(2100) == 1024
        assert S.get_ls_tile(2101) == (512, 4, 2)
    finally:
        S.set_ls_tile(2100, 0)
    assert S.get_ls_tile(2100) == (512, 4, 2)
    assert S.get_ls_tile(600) == (256, 2, 4)
    S.set_ls_block(600, 256)                  # the built-in configuration of that size
    try:
        assert S.get_ls_tile(600) == (256, 2, 4)
    finally:
        S.set_ls_block(600, 0)
    S.set_ls_block(600, 512)                  # another size: depth 4, CU-sharing form
    try:
        assert S.get_ls_tile(600) == (512, 4, 2)
    finally:
        S.set_ls_block(600, 0)
    S.set_ls_block(2100, 256)                 # past the four-a-CU lengths: one a CU
    try:
        assert S.get_ls_tile(2100) == (256, 4, 1)
    finally:
        S.set_ls_block(2100, 0)
    S.set_ls_block(600, 1024)
    try:
        assert S.get_ls_tile(600) == (1024, 4, 1)
    finally:
        S.set_ls_block(600, 0)
    tiles = [S.ls_tile_config(i) for i in range(S.LS_TILES)]
    assert len(set(tiles)) == S.LS_TILES
    for n in (300, 600, 1200, 2100, 4096, 8400):      # the built-in rule is a trial
        assert S.get_ls_tile(n) in tiles
