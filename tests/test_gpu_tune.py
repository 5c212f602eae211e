"""GPU checks of the runtime autotuner (tokuraAutotune / tokuraLoadTuning),
which replaces the reference's build-time tuner
(TOKURA_ZHSEQR_BATCHED_GEHRD_TUNING_main.cu:89-242).

* The cache it writes is keyed by the device: every line carries the gfx arch
  name and the CU count of the GPU it was timed on.
* A fresh handle that loads the cache gets the same variant for every n, and
  its results are bit for bit those of a handle forced to that variant.
* Lines for another device are ignored.
* The ragged tuning table (tokuraAutotuneRagged, tokuraSaveTuning): the
  chosen configuration and small-class Hessenberg grid are saved under the
  batch's size signature, and a fresh handle that loads the cache prepares
  that batch with them; a batch of another signature keeps the handle's
  default, and a configuration set by hand wins over the table.
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


def test_ragged_autotune_cache_roundtrip(tmp_path):
    from tokura_zgeev_batched_amd.ragged import RaggedBatch
    rng = np.random.default_rng(5)
    sizes = rng.integers(4, 65, 6000)
    offs = np.zeros(len(sizes) + 1, np.int64)
    offs[1:] = np.cumsum(sizes.astype(np.int64) ** 2)
    g = torch.Generator(device="cuda").manual_seed(21)
    flat = torch.complex(torch.randn(int(offs[-1]), generator=g, device="cuda", dtype=torch.float64),
                         torch.randn(int(offs[-1]), generator=g, device="cuda", dtype=torch.float64))
    h = tk.Handle()
    rb = RaggedBatch(offs, handle=h)
    print("[case] ragged autotune 6000 matrices", flush=True)
    cfg, _ = rb.autotune(flat)
    hm = rb.hess_grid()
    assert hm in (1, 2, 4, 8)
    assert rb.prepared_config() == (cfg, hm)
    e1, f1 = rb.solve(flat)
    e1 = e1.clone()
    path = str(tmp_path / "tune.txt")
    h.save_tuning(path)
    props = torch.cuda.get_device_properties(0)
    arch = props.gcnArchName.split(":")[0]
    rl = [ln.split() for ln in open(path) if ln.startswith("ragged ")]
    assert len(rl) == 1
    assert rl[0][1] == arch and int(rl[0][2]) == props.multi_processor_count
    assert int(rl[0][4]) == cfg and int(rl[0][5]) == hm
    h2 = tk.Handle()
    assert h2.load_tuning(path) == 65
    rb2 = RaggedBatch(offs, handle=h2)          # prepare takes the table's entry
    assert rb2.prepared_config() == (cfg, hm)
    e2, f2 = rb2.solve(flat, graph=True)
    torch.cuda.synchronize()
    assert int(f1.abs().sum()) == 0 and int(f2.abs().sum()) == 0
    assert torch.equal(e1, e2)
    # another size signature: the handle's default configuration
    offs8 = np.arange(2001, dtype=np.int64) * 64
    rb3 = RaggedBatch(offs8, handle=h2)
    assert rb3.prepared_config() == (0, 4)
    # a configuration set by hand wins over the table
    alt = 9 if cfg != 9 else 5
    assert tk.lib().tokuraSetRaggedConfig(h2.ptr, alt) == 0
    rb2.prepare()
    assert rb2.prepared_config()[0] == alt
    h.close()
    h2.close()
