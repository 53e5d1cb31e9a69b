"""CPU checks of the boundary: libpst_hip.so loads and exports every symbol
include/pst/pst_capi.h declares; host-side size helpers agree with the
reference's struct layouts. No device calls."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pst", "pst_capi.h")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pst_[a-z0-9_]+)\s*\(", txt)))


@pytest.fixture(scope="module")
def L():
    import pst_amd
    return pst_amd.lib()


def test_header_declares_the_mirror_list():
    import pst_amd
    assert sorted(pst_amd.EXPORTED) == declared()


def test_library_exports_every_declared_symbol(L):
    missing = [s for s in declared() if not hasattr(L, s)]
    assert not missing, missing


def test_point_size_matches_reference_layout(L):
    import oracle as O
    import pst_amd
    for t in ("f32", "f64", "i32", "i64"):
        for i in (None, "f32", "f64", "i32", "i64", "u32", "u64"):
            assert L.pst_point_size(pst_amd.DT[t], pst_amd.DT[i]) == O.point_dtype(t, i).itemsize


def test_query_record_layout():
    import pst_amd
    assert pst_amd.query_dtype("f32").itemsize == 16
    assert pst_amd.query_dtype("f64").itemsize == 32
    assert pst_amd.query_dtype("i32").itemsize == 16
    assert pst_amd.query_dtype("i64").itemsize == 32


def test_calc_global_mem_needed_covers_tree(L):
    import pst_amd
    n = 1 << 20
    need = pst_amd.StaticPSTGPU.calcGlobalMemNeeded(n, "f32", "u32")
    S = (1 << 21) - 1
    assert need >= S * 16 + n * 12


def test_invalid_type_codes_rejected_without_device(L):
    h = ctypes.c_void_p()
    rc = L.pst_gpu_create(None, 0, 9, 0, 1, 0, 1, None, ctypes.byref(h))
    assert rc == 1
    assert b"unsupported" in L.pst_last_error()


def test_calc_num_metacells(L):
    """calcNumMetacells: ceil((p - 1) / m) metacells per axis
    (helpers/isosurface-data-processing.h:21-36)."""
    import pst_amd as P
    n, g = P.calcNumMetacells((1024, 1024, 1024), (4, 4, 4))
    assert g == (256, 256, 256) and n == 256 ** 3
    n, g = P.calcNumMetacells((10, 7, 2), (3, 3, 1))
    assert g == (3, 2, 1) and n == 6
    n, g = P.calcNumMetacells((5, 5, 5), (4, 4, 4))
    assert g == (1, 1, 1) and n == 1
    assert P.calcNumMetacells((5, 5, 5), (0, 4, 4))[0] == 0


def test_raw_vertex_reader_host(L, tmp_path):
    """readInVertices (helpers/isosurface-data-processing.h:205-229): the
    values of the file; the reference's two messages for a missing and a
    short file. Host I/O only."""
    import pst_amd as P
    vol = np.arange(7 * 5 * 3, dtype=np.float32) * 0.5 - 3
    path = tmp_path / "v.raw"
    vol.tofile(path)
    got = P.readInVertices(str(path), vol.size, "f32")
    assert np.array_equal(got, vol)
    assert np.array_equal(P.readInVertices(str(path), 4, "f32"), vol[:4])  # a prefix is fine, as in the reference
    with pytest.raises(P.PSTError) as e:
        P.readInVertices(str(path), vol.size + 1, "f32")
    assert e.value.code == P.ERR_IO
    assert f"Failure: read {vol.nbytes} bytes, expected {vol.nbytes + 4} bytes" in str(e.value)
    with pytest.raises(P.PSTError) as e:
        P.readInVertices(str(tmp_path / "missing.raw"), 3, "f32")
    assert e.value.code == P.ERR_IO and "Failed to open" in str(e.value) and "for reading" in str(e.value)


def _header(n=5, S=7, kind=0, dtype=0, id_dtype=4, spt=0, magic=b"PSTAMD\x00\x01", version=1):
    import struct
    h = magic + struct.pack("<IIiiQQQQ", version, kind, dtype, id_dtype, n, S, spt, 0)
    return h + b"\x00" * (64 - len(h))


@pytest.mark.parametrize("blob,what", [
    (b"not a tree file at all", "not a tree file"),
    (_header(magic=b"XSTAMD\x00\x01") + b"\x00" * 7 * 16, "not a tree file"),
    (_header(version=2) + b"\x00" * 7 * 16, "version 2"),
    (_header(S=15) + b"\x00" * 15 * 16, "corrupt header"),
    (_header(kind=1, spt=0) + b"\x00" * 7 * 16, "corrupt header"),
    (_header(kind=0, spt=63) + b"\x00" * 7 * 16, "corrupt header"),
    (_header(dtype=9) + b"\x00" * 7 * 16, "corrupt header"),
    (_header() + b"\x00" * (7 * 16 - 1), "expected 176"),
])
def test_tree_file_header_rejected_without_device(L, tmp_path, blob, what):
    """pst_gpu_load checks the header (magic, version, kind, types, slot
    count from n, file size) before any device work."""
    import pst_amd as P
    path = tmp_path / "bad.pst"
    path.write_bytes(blob)
    h = ctypes.c_void_p()
    rc = L.pst_gpu_load(os.fsencode(str(path)), 0, 1, None, ctypes.byref(h))
    assert rc == P.ERR_IO and not h.value
    assert what in L.pst_last_error().decode()
