"""The C-ABI library loads on a CPU-only host and exports every function
declared in include/gtps.h; the ctypes binding covers exactly that set; the
error path works without a GPU."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gtps.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gtps_[a-z0-9_]+)\s*\(", text)))


def test_header_parses():
    names = declared_functions()
    assert "gtps_mg_pcg_solve" in names and "gtps_laplace_vmult" in names
    assert len(names) >= 35


def test_library_exports_every_symbol():
    from gputensorproductsmoother_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_matches_header():
    from gputensorproductsmoother_amd import _lib
    assert sorted(_lib.SIGNATURES) == declared_functions()


def test_version_and_error_path_without_gpu():
    from gputensorproductsmoother_amd import _lib
    lib = _lib.load()
    assert lib.gtps_version().startswith(b"gtps")
    # NULL output pointer is rejected before any device call
    rc = lib.gtps_ctx_create(0, None, None)
    assert rc == _lib.GTPS_EINVAL
    assert b"NULL" in lib.gtps_last_error()
    # destroying NULL handles is a no-op
    for fn in ("gtps_level_destroy", "gtps_smoother_destroy", "gtps_transfer_destroy",
               "gtps_mg_destroy", "gtps_ctx_destroy"):
        assert getattr(lib, fn)(None) == _lib.GTPS_OK
    assert lib.gtps_level_n_dofs(None) == -1


def test_missing_library_fails_loudly(tmp_path):
    from gputensorproductsmoother_amd import _lib
    with pytest.raises(RuntimeError, match="not found"):
        _lib.load(str(tmp_path / "libgtps.so"))


def _fnv1a(data: bytes) -> int:
    h = 0xcbf29ce484222325
    for byte in data:
        h = ((h ^ byte) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


def test_solution_file_header_without_gpu(tmp_path):
    """gtps_vector_info parses the documented 64-byte header (written here
    independently with struct) and rejects a wrong magic or size."""
    import struct
    from gputensorproductsmoother_amd import core, _lib as L
    dim, k, nc = 2, 3, 4
    n = (k * nc + 1) ** dim
    data = struct.pack(f"<{n}d", *[0.5 * i for i in range(n)])
    hdr = b"GTPSVEC1" + struct.pack("<6iqQ", L.GTPS_FP64, dim, k, nc, L.GTPS_LAYOUT_CG, 0, n,
                                    _fnv1a(data)) + bytes(16)
    assert len(hdr) == 64
    good = tmp_path / "good.gtps"
    good.write_bytes(hdr + data)
    info = core.vector_info(good)
    assert (info["dim"], info["fe_degree"], info["n_cells"], info["n"], info["layout"]) == \
        (dim, k, nc, n, L.GTPS_LAYOUT_CG)
    bad = tmp_path / "bad.gtps"
    bad.write_bytes(b"GTPSVEC0" + hdr[8:] + data)
    with pytest.raises(RuntimeError, match="not a valid solution file"):
        core.vector_info(bad)
    wrong_n = tmp_path / "wrong.gtps"
    wrong_n.write_bytes(hdr[:32] + struct.pack("<q", n + 1) + hdr[40:] + data)
    with pytest.raises(RuntimeError):
        core.vector_info(wrong_n)


def documented_tuning_keys():
    text = open(HEADER).read()
    block = text[text.index("Process-wide tuning knobs"):]
    block = block[:block.index("*/")]
    return re.findall(r'^\s*\*\s+"([a-z0-9_]+)"', block, flags=re.M)


def test_tuning_knobs_round_trip_without_gpu():
    """Every knob documented in gtps.h reads and writes back on a CPU-only
    host (host state only); unknown keys and out-of-range values are
    rejected with GTPS_EINVAL and leave the knob unchanged."""
    from gputensorproductsmoother_amd import _lib
    lib = _lib.load()
    keys = documented_tuning_keys()
    assert {"mfma_smoother", "mfma_waves", "vmult_tile", "patches_per_wave"} <= set(keys)
    for k in keys:
        v = lib.gtps_get_tuning(k.encode())
        assert v >= 0, (k, lib.gtps_last_error())
        assert lib.gtps_set_tuning(k.encode(), v) == v
        assert lib.gtps_get_tuning(k.encode()) == v
    assert lib.gtps_get_tuning(b"no_such_knob") == -1
    assert lib.gtps_set_tuning(b"no_such_knob", 1) == -1
    for k, bad in (("mfma_waves", 5), ("patches_per_wave", 0), ("mfma_smoother", 2),
                   ("group_waves", 3)):
        before = lib.gtps_get_tuning(k.encode())
        assert lib.gtps_set_tuning(k.encode(), bad) == -1
        assert lib.gtps_get_tuning(k.encode()) == before
