"""C ABI checks that need no GPU: the library loads, exports every entry
point include/bsar.h declares, host-only entry points (chirp, file formats)
behave like the reference, and device entry points fail cleanly without a
device."""
import re
import struct

import numpy as np
import pytest

from conftest import GOLDEN, ROOT


def declared_functions():
    text = (ROOT / "include" / "bsar.h").read_text()
    return sorted(set(re.findall(r"BSAR_API\s+[\w\s\*]+?\b(bsar_\w+)\s*\(", text)))


def test_every_declared_symbol_is_exported_and_bound(bsar):
    names = declared_functions()
    assert len(names) >= 20
    L = bsar.lib()
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(bsar.SIGNATURES), set(names) ^ set(bsar.SIGNATURES)


def test_internal_symbols_are_hidden(bsar):
    L = bsar.lib()
    with pytest.raises(AttributeError):
        getattr(L, "bsar_r1tab_free")


def test_version_status_flops(bsar):
    L = bsar.lib()
    assert L.bsar_abi_version() == 1
    assert L.bsar_status_string(0) == b"ok"
    assert L.bsar_status_string(-6) == b"capacity exceeded"
    assert bsar.flops_per_voxel_pulse(2048) == 106366.0


def test_context_without_device_fails_cleanly(bsar):
    import ctypes as C
    L = bsar.lib()
    h = C.c_void_p()
    n = C.c_int()
    st = L.bsar_ctx_create(0, C.byref(h))
    if st == 0:
        L.bsar_ctx_destroy(h)
        pytest.skip("a GPU is present")
    assert st == -2 and not h.value
    assert L.bsar_ctx_create(-1, C.byref(h)) in (-1, -2)


def test_server_without_device_fails_cleanly(bsar):
    import ctypes as C
    L = bsar.lib()
    h = C.c_void_p()
    st = L.bsar_server_create(0, None, C.byref(h))
    if st == 0:
        L.bsar_server_destroy(h)
        pytest.skip("a GPU is present")
    assert st == -2 and not h.value
    L.bsar_server_destroy(None)                      # null is a no-op
    assert L.bsar_server_submit(None, None, None) == -1
    assert L.bsar_server_wait(None, 1, 0.0, None) == -1
    assert L.bsar_server_wait_any(None, 0.0, None) == -1
    assert L.bsar_server_drain(None) == -1
    assert L.bsar_server_last_error(None) == b"null server"


def test_server_structs_match_header(bsar):
    """ctypes mirrors of bsar_server_config / bsar_job_desc / bsar_job_info /
    bsar_server_stats have the C layouts (x86-64 SysV)."""
    import ctypes as C
    assert C.sizeof(bsar.ServerConfigC) == 4 + 4 + 8 + 8 + 4 * 8
    assert C.sizeof(bsar.JobDescC) == 8 + 4 + 4 + 3 * 16 + 8
    assert C.sizeof(bsar.JobInfoC) == 8 + 4 + 4 + 5 * 8 + 8
    assert C.sizeof(bsar.ServerStatsC) == 6 * 8 + 2 * 8 + 2 * 8 + 2 * 4


def test_chirp_matches_reference(bsar, golden):
    for tag in ("refc", "cfg"):
        f0, f1, T, fs = golden[f"chirp_{tag}_args"]
        c = bsar.chirp(f0, f1, T, fs)
        assert np.array_equal(c.view(np.float64), golden[f"chirp_{tag}_out"])
    assert len(bsar.chirp(-150e6, 150e6, 1e-6, 800e6)) == 801


@pytest.mark.parametrize("tag", ["random", "target"])
def test_image_writer_is_byte_identical_to_reference(bsar, tmp_path, tag):
    g = np.load(GOLDEN / f"ref_bp_{tag}.npz")
    blob = g["out_bin"].tobytes()
    nx, ny, nz = struct.unpack_from("<QQQ", blob, 28)
    img = np.frombuffer(blob[52:], np.complex128).reshape(nz, ny, nx)
    path = tmp_path / "output_3D_matrix.bin"
    bsar.write_image(path, img, g["gx"], g["gy"], g["gz"])
    assert path.read_bytes() == blob
    for a, text in zip("xyz", g["grid_csv"]):
        assert (tmp_path / f"output_3D_matrix_grid_{a}.csv").read_text() == str(text)


def test_matrix_round_trip_and_errors(bsar, tmp_path):
    rng = np.random.default_rng(0)
    m = rng.normal(size=(5, 7)) + 1j * rng.normal(size=(5, 7))
    p = tmp_path / "m.bin"
    bsar.write_matrix2d(p, m)
    flat, dims = bsar.read_matrix(p)
    assert dims == [5, 7] and np.array_equal(flat, m.ravel())
    raw = p.read_bytes()
    assert len(raw) == 44 + 16 * 35
    # truncated payload / header -> format error; missing file -> i/o error
    (tmp_path / "t.bin").write_bytes(raw[:-8])
    with pytest.raises(bsar.BsarError) as e:
        bsar.read_matrix(tmp_path / "t.bin")
    assert e.value.status == -5
    (tmp_path / "h.bin").write_bytes(raw[:20])
    with pytest.raises(bsar.BsarError) as e:
        bsar.read_matrix(tmp_path / "h.bin")
    assert e.value.status == -5
    with pytest.raises(bsar.BsarError) as e:
        bsar.read_matrix(tmp_path / "missing.bin")
    assert e.value.status == -4
    # real payload reads as (v, 0)
    hdr = struct.pack("<III", 2, 0, 6) + b"double".ljust(16, b"\0") + struct.pack("<QQ", 2, 3)
    (tmp_path / "r.bin").write_bytes(hdr + np.arange(6.0).tobytes())
    flat, dims = bsar.read_matrix(tmp_path / "r.bin")
    assert dims == [2, 3] and np.array_equal(flat, np.arange(6.0) + 0j)


def test_reference_input_file_reads_range_fastest(bsar, tmp_path):
    """rootData_complex.bin dims are [S, N]; the payload flat index r + p*S
    is a pulse-major [N][S] array (proc_test_script.cpp:330)."""
    root = np.load(GOLDEN / "ref_bp_random.npz")["root"]
    N, S = root.shape
    p = tmp_path / "rootData_complex.bin"
    p.write_bytes(struct.pack("<III", 2, 1, 6) + b"double".ljust(16, b"\0") + struct.pack("<QQ", S, N)
                  + np.ascontiguousarray(root).tobytes())
    flat, dims = bsar.read_matrix(p)
    assert dims == [S, N]
    assert np.array_equal(flat.reshape(N, S), root)
