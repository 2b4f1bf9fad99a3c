"""bench.py's roofline bookkeeping (host logic, no GPU): the pair kernel's
name for a row set and the per-launch HBM traffic looked up in the
committed counter summary."""
import importlib.util
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def _rows(norb, kmax=2):
    # (k, a, b, c, d) rows holding every pair of norb orbitals on both sides
    r = [(k, a, a, c, c) for k in range(kmax + 1) for a in range(norb) for c in range(a, norb)]
    return np.array(r)


def test_pair_kernel_names():
    b = _bench()
    # 30 orbitals: 465 densities fit one 512-thread workgroup at one per thread
    assert b.pair_kernel(_rows(30), [0, 256, 16, 4096, 256, 1]) == ("k_pair_rot", "false", ", 1")
    # 40 orbitals: 820 densities, two per thread at W = 8, the general kernel at W = 16
    assert b.pair_kernel(_rows(40), [0, 256, 8, 4096, 256, 1]) == ("k_pair_rot", "false", ", 2")
    assert b.pair_kernel(_rows(40), [0, 256, 16, 4096, 256, 1])[0] == "k_pair_panel"


def test_traffic_lookup(tmp_path):
    b = _bench()
    f = tmp_path / "pmc.md"
    f.write_text("| kernel | tuning | dispatches | fetch KiB | write KiB | bytes |\n|---|---|---:|---:|---:|---:|\n"
                 "| k_x<16, false, true, 1> | 0,128,16,4096,256,1 | 8 | 1 | 2 | 1000 |\n"
                 "| k_x<16, false, true, 1> | 0,256,16,4096,256,1 | 8 | 1 | 2 | 2000 |\n")
    assert b.measured_traffic(str(f), "k_x<16, false, true, 1>", [0, 256, 16, 4096, 256, 1]) == (2000.0, "0,256,16,4096,256,1")
    assert b.measured_traffic(str(f), "k_x<16, false, true, 1>", [0, 64, 16, 4096, 256, 1]) == (1000.0, "0,128,16,4096,256,1")
    assert b.measured_traffic(str(f), "k_y", [0, 64, 16, 4096, 256, 1]) == (None, None)
    assert b.measured_traffic(str(tmp_path / "absent.md"), "k_x", [0]) == (None, None)
