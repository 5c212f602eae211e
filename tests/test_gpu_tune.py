"""GPU checks of the runtime autotuner (tokuraAutotune / tokuraLoadTuning),
which replaces the reference's build-time tuner
(TOKURA_ZHSEQR_BATCHED_GEHRD_TUNING_main.cu:89-242).

* The cache it writes is keyed by the device: every line carries the gfx arch
  name and the CU count of the GPU it was timed on.
* A fresh handle that loads the cache gets the same variant for every n, and
  its results are bit for bit those of a handle forced to that variant.
* Lines for another device are ignored.
"""
import numpy as np
import pytest
import torch

import tokura_zgeev_batched_amd as tk

pytestmark = pytest.mark.gpu


def _eigs(M, h):
    e, f = tk.eigvals(M, h)
    torch.cuda.synchronize()
    assert int(f.abs().sum()) == 0
    return e


def test_autotune_cache_roundtrip(tmp_path):
    lo, hi = 30, 34
    path = str(tmp_path / "tune.txt")
    h = tk.Handle()
    print("[case] autotune n 30..34", flush=True)
    h.autotune(lo, hi, 2048, path)
    props = torch.cuda.get_device_properties(0)
    arch = props.gcnArchName.split(":")[0]
    lines = [ln.split() for ln in open(path) if not ln.startswith("#")]
    assert len(lines) == 64
    assert all(ln[0] == arch and int(ln[1]) == props.multi_processor_count for ln in lines)
    table = {int(ln[2]): int(ln[3]) for ln in lines}
    h2 = tk.Handle()
    assert h2.load_tuning(path) == 64
    g = torch.Generator(device="cuda").manual_seed(77)
    for n in range(lo, hi + 1):
        v = h.launch_config(n)[0]
        assert table[n] == v and h2.launch_config(n)[0] == v
        h3 = tk.Handle()
        h3.set_variant(n, v)
        M = torch.complex(torch.randn(300, n, n, generator=g, device="cuda", dtype=torch.float64),
                          torch.randn(300, n, n, generator=g, device="cuda", dtype=torch.float64))
        e2, e3 = _eigs(M, h2), _eigs(M, h3)
        assert torch.equal(e2, e3), n
        h3.close()
    h.close()
    h2.close()


def test_tuning_lines_of_other_devices_are_ignored(tmp_path):
    props = torch.cuda.get_device_properties(0)
    arch = props.gcnArchName.split(":")[0]
    path = tmp_path / "other.txt"
    path.write_text(f"gfx000 {props.multi_processor_count} 33 5\n{arch} {props.multi_processor_count + 1} 33 5\n"
                    f"{arch} {props.multi_processor_count} 33 19\n")
    h = tk.Handle()
    before = h.launch_config(40)[0]
    assert h.load_tuning(str(path)) == 1
    assert h.launch_config(33)[0] == 19 and h.launch_config(40)[0] == before
    h.close()
