"""The on-disk cache of compiled model kernels survives damaged files: a
truncated file, and a file whose header and checksum are intact but whose
code object the runtime will not load, are both compiled again and
replaced, and the model builds and steps."""
import struct

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAGIC = b"DCGCOv2\x00"


def _fnv1a(data: bytes) -> int:
    h = 1469598103934665603
    for c in data:
        h = ((h ^ c) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def _build_and_step(dca, w):
    # one point-apply tile shape for every build: shapes picked by timing
    # may differ between builds, and with them the order of the dot sums
    m = dca.Model(w.ini, "model.time_step_operator.linear_solver.point_tile = 0\n", 0)
    m.set_state(w.initial_state(m), 0.0)
    assert m.step(0.0, w.dt).status == 0
    u, _ = m.get_state()
    m.close()  # the process-wide module is released with its last model
    return u


def test_damaged_cache_files_are_recompiled(dca, wl, tmp_path, monkeypatch):
    monkeypatch.setenv("DCOPASI_RTC_CACHE", str(tmp_path))
    # a model no other test builds (its own source, so no module is shared)
    w = wl.gray_scott(n=7, F=0.0371, k=0.0613, precond="Jacobi", nl_rtol=1e-14)
    ref = _build_and_step(dca, w)
    files = sorted(p for p in tmp_path.iterdir() if p.suffix == ".co")
    assert len(files) == 1
    good = files[0].read_bytes()
    assert good[:8] == MAGIC
    size, summ = struct.unpack("<QQ", good[8:24])
    assert size == len(good) - 24 and summ == _fnv1a(good[24:])
    # truncated
    files[0].write_bytes(good[: len(good) // 2])
    np.testing.assert_array_equal(_build_and_step(dca, w), ref)
    assert files[0].read_bytes() == good
    # intact header and checksum around a code object the runtime refuses
    junk = bytes(range(256)) * 64
    files[0].write_bytes(MAGIC + struct.pack("<QQ", len(junk), _fnv1a(junk)) + junk)
    np.testing.assert_array_equal(_build_and_step(dca, w), ref)
    assert files[0].read_bytes() == good
    # a file of the old layout (bare code object) is not trusted either
    files[0].write_bytes(good[24:])
    np.testing.assert_array_equal(_build_and_step(dca, w), ref)
    assert files[0].read_bytes() == good
    assert not [p for p in tmp_path.iterdir() if ".tmp" in p.name]
