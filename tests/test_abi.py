"""CPU checks of the C ABI boundary (no compute call needs a GPU here).

Mirrors the reference's contract: names and validation order of
tokura_zgeev_batched_library/src/tokura_zgeev_batched.cu:192-264 and the
prototypes of tokura_zgeev_batched_library/include/tokura_blas.h:129-151.
"""
import ctypes
import os
import re

import pytest

import tokura_zgeev_batched_amd as tk

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "tokura_blas.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tokura\w+)\s*\(", src)))


def test_library_exports_every_header_symbol():
    L = tk.lib()
    names = header_functions()
    assert "tokura_zgeev_batched_gpu" in names and "tokuraCreate" in names
    for name in names:
        assert hasattr(L, name), name


def test_reference_entry_points_present():
    for name in ("tokuraCreate", "tokuraDestroy", "tokura_get_zgeeveigenvaluesgetworspacesize",
                 "tokura_zgeev_batched_gpu"):
        assert name in header_functions()


def test_workspace_size_formula_and_errors():
    f = tk.lib().tokura_get_zgeeveigenvaluesgetworspacesize
    assert f(16, 10000) == 16 * (16 * 16 * 10000 + 16 * 10000) == 43_520_000
    assert f(64, 100000) == 6_656_000_000
    M = (1 << 64) - 1
    assert f(0, 5) == M and f(65, 5) == M            # (size_t)-1
    assert f(4, 0) == M - 1 and f(4, -3) == M - 1     # (size_t)-2


def make_handle():
    h = ctypes.c_void_p()
    rc = tk.lib().tokuraCreate(ctypes.byref(h))
    assert rc in (0, 1)        # 1: no HIP device in this container
    assert h.value
    return h


def test_validation_codes_in_reference_order():
    L = tk.lib()
    h = make_handle()
    P = ctypes.c_void_p
    a, e, w, f = P(0x1000), P(0x2000), P(0x3000), P(0x4000)
    s = P(0)
    assert L.tokura_zgeev_batched_gpu(None, 4, 1, a, e, w, f, s) == -1
    assert L.tokura_zgeev_batched_gpu(h, 0, 1, a, e, w, f, s) == -2
    assert L.tokura_zgeev_batched_gpu(h, 65, 1, a, e, w, f, s) == -2
    assert L.tokura_zgeev_batched_gpu(h, 4, 1, None, e, w, f, s) == -3
    assert L.tokura_zgeev_batched_gpu(h, 4, 1, a, None, w, f, s) == -4
    assert L.tokura_zgeev_batched_gpu(h, 4, 1, a, e, None, f, s) == -5
    assert L.tokura_zgeev_batched_gpu(h, 4, 0, a, e, w, f, s) == -6
    assert L.tokura_zgeev_batched_gpu(h, 4, 1, a, e, w, None, s) == -7
    # order: n is checked before A, A before batch
    assert L.tokura_zgeev_batched_gpu(h, 0, 0, None, None, None, None, s) == -2
    assert L.tokura_zgeev_batched_gpu(h, 4, 0, None, e, w, f, s) == -3
    assert L.tokuraDestroy(h) == 0
    assert L.tokuraDestroy(None) == 0


def test_handle_controls():
    L = tk.lib()
    h = make_handle()
    v, lpm, mpb = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    for n in (1, 4, 8, 16, 32, 48, 64):
        assert L.tokuraGetLaunchConfig(h, n, ctypes.byref(v), ctypes.byref(lpm), ctypes.byref(mpb)) == 0
        assert lpm.value * mpb.value == 64
    assert L.tokuraGetLaunchConfig(h, 65, None, None, None) == -2
    assert L.tokuraSetMaxSweeps(h, 3) == 0
    assert L.tokuraSetVariant(h, 64, 0) == -2          # variant 0 holds n <= 4 only
    assert L.tokuraSetVariant(h, 4, 0) == 0
    assert L.tokuraDestroy(h) == 0


def test_tuning_cache_roundtrip(tmp_path):
    L = tk.lib()
    h = make_handle()
    # the CPU handle has an empty arch key; entries for other devices are ignored
    p = tmp_path / "tune.txt"
    p.write_text("# arch cus n variant\ngfx950 256 16 5\n 0 16 4\n")
    assert L.tokuraLoadTuning(h, str(p).encode()) == 0
    assert L.tokuraLoadTuning(h, b"/nonexistent/file") < 0
    L.tokuraDestroy(h)


def test_tuning_cache_save_and_ragged_lines(tmp_path):
    """tokuraSaveTuning writes the 64 same-size lines under the handle's
    device key; ragged lines ("ragged arch cus signature config multiplier")
    of another device, malformed ones and out-of-range configurations are
    ignored by tokuraLoadTuning."""
    L = tk.lib()
    h = make_handle()
    p = tmp_path / "save.txt"
    assert L.tokuraSaveTuning(h, str(p).encode()) == 0
    body = [ln for ln in p.read_text().splitlines() if not ln.startswith("#")]
    assert len(body) == 64 and not any(ln.startswith("ragged ") for ln in body)
    assert L.tokuraSaveTuning(None, str(p).encode()) == -1
    assert L.tokuraSaveTuning(h, None) == -3
    # other devices' lines survive a save; this device's lines are replaced,
    # not duplicated; the file is replaced whole (no temporary left behind)
    foreign = ["gfx942 304 64 20", "ragged gfx942 304 abc 3 2"]
    p.write_text(p.read_text() + "\n".join(foreign) + "\n")
    assert L.tokuraSaveTuning(h, str(p).encode()) == 0
    assert L.tokuraSaveTuning(h, str(p).encode()) == 0
    body2 = [ln for ln in p.read_text().splitlines() if not ln.startswith("#")]
    assert all(ln in body2 for ln in foreign) and len(body2) == 64 + len(foreign)
    assert sorted(ln for ln in body2 if ln not in foreign) == sorted(body)
    assert [x.name for x in tmp_path.iterdir()] == ["save.txt"]
    assert L.tokuraLoadTuning(h, str(p).encode()) == 64
    q = tmp_path / "ragged.txt"
    q.write_text("ragged gfx950 256 10842108421 5 4\nragged gfx950 256 zz 5 4\nragged gfx950 256 1 99 4\n")
    assert L.tokuraLoadTuning(h, str(q).encode()) == 0
    m = ctypes.c_int()
    assert L.tokuraGetRaggedHessGrid(h, ctypes.byref(m)) == 0 and m.value == 4
    L.tokuraDestroy(h)


def test_python_layer_refuses_host_tensors():
    torch = pytest.importorskip("torch")
    with pytest.raises(tk.TokuraError):
        tk._dev_check(torch.zeros(2, 2, dtype=torch.complex128), "M", torch.complex128)
    # the ragged wrapper's caller-supplied buffers (eig, flag, gbar, schur)
    with pytest.raises(tk.TokuraError):
        tk._buf(torch.zeros(8, dtype=torch.complex128), "gbar", torch.complex128, 8, torch.device("cuda", 0))
    with pytest.raises(tk.TokuraError):   # raw wrappers check before any launch
        tk.zgeev_batched(torch.zeros(16, dtype=torch.complex128), 4, 1, torch.zeros(4, dtype=torch.complex128),
                         torch.zeros(512, dtype=torch.uint8), torch.zeros(1, dtype=torch.int8))
    assert not tk._shape_ok(0, 1) and not tk._shape_ok(65, 1) and not tk._shape_ok(4, 0) and tk._shape_ok(64, 1)


def test_vectors_entry_validation():
    L = tk.lib()
    h = make_handle()
    P = ctypes.c_void_p
    a, e, vl, vr, f, w = P(0x1000), P(0x2000), P(0x3000), P(0x4000), P(0x5000), P(0x6000)
    s = P(0)
    g = L.tokura_zgeev_batched_vectors_gpu
    assert g(None, 4, 1, a, e, vl, vr, w, f, s) == -1
    assert g(h, 0, 1, a, e, vl, vr, w, f, s) == -2
    assert g(h, 65, 1, a, e, vl, vr, None, f, s) == -2
    assert g(h, 4, 1, None, e, vl, vr, w, f, s) == -3
    assert g(h, 4, 1, a, None, vl, vr, w, f, s) == -4
    assert g(h, 4, 1, a, e, None, None, w, f, s) == -5     # neither left nor right vectors asked for
    assert g(h, 4, 0, a, e, vl, vr, w, f, s) == -6
    assert g(h, 4, 1, a, e, vl, vr, w, None, s) == -7
    assert L.tokuraDestroy(h) == 0
    ws = L.tokura_get_zgeev_vectors_worksize
    assert ws(64, 1000) == 16 * 64 * 64 * 1000            # T of every matrix, n*n complex
    assert ws(0, 10) == ctypes.c_size_t(-1).value
    assert ws(65, 10) == ctypes.c_size_t(-1).value
    assert ws(8, 0) == ctypes.c_size_t(-2).value


def test_balancing_switch():
    L = tk.lib()
    h = make_handle()
    for job in (b"N", b"P", b"S", b"B", b"b"):
        assert L.tokuraSetBalancing(h, job) == 0
    assert L.tokuraSetBalancing(h, b"X") == -2
    assert L.tokuraSetBalancing(None, b"B") == -1
    assert L.tokuraDestroy(h) == 0


def test_real_entry_validation_and_workspace():
    L = tk.lib()
    h = make_handle()
    P = ctypes.c_void_p
    a, wr, wi, w, f = P(0x1000), P(0x2000), P(0x2800), P(0x3000), P(0x4000)
    s = P(0)
    g = L.tokura_dgeev_batched_gpu
    assert g(None, 4, 1, a, wr, wi, w, f, s) == -1
    assert g(h, 0, 1, a, wr, wi, w, f, s) == -2
    assert g(h, 65, 1, a, wr, wi, w, f, s) == -2
    assert g(h, 4, 1, None, wr, wi, w, f, s) == -3
    assert g(h, 4, 1, a, None, wi, w, f, s) == -4
    assert g(h, 4, 1, a, wr, None, w, f, s) == -4
    assert g(h, 4, 1, a, wr, wi, None, f, s) == -5
    assert g(h, 4, 0, a, wr, wi, w, f, s) == -6
    assert g(h, 4, 1, a, wr, wi, w, None, s) == -7
    assert L.tokuraDestroy(h) == 0
    ws = L.tokura_get_dgeev_worksize
    # packed upper Hessenberg, (n-1)(n+2)/2 + n doubles per matrix
    assert ws(64, 1000) == 8 * (63 * 66 // 2 + 64) * 1000
    assert ws(1, 3) == 8 * 3
    M = (1 << 64) - 1
    assert ws(0, 5) == M and ws(65, 5) == M and ws(4, 0) == M - 1
