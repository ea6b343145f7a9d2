"""The C-ABI library loads without a GPU and exports every symbol of include/sph.h."""
import ctypes
import os
import re

import numpy as np
import pytest

import sph_system
from conftest import ROOT


def header_functions(name="sph.h"):
    text = open(os.path.join(ROOT, "include", name)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sph_[A-Za-z0-9_]+)\s*\(", text)))


def test_header_matches_entry_point_list():
    assert header_functions() == sorted(sph_system.ENTRY_POINTS)
    assert header_functions("sph_surrogate.h") == sorted(sph_system.SURROGATE_ENTRY_POINTS)


def test_library_exports_every_header_symbol():
    lib = ctypes.CDLL(sph_system.LIB_PATH)
    for name in header_functions() + header_functions("sph_surrogate.h"):
        assert hasattr(lib, name), name


def test_create_rejects_bad_precision():
    lib = sph_system.library()
    assert not lib.sph_create(7)


def test_params_struct_size_is_multiple_of_8():
    assert ctypes.sizeof(sph_system.SphParams) % 8 == 0
    names = [f[0] for f in sph_system.SphParams._fields_]
    assert len(names) == len(set(names))


def test_device_calls_before_initialize_fail_cleanly():
    s = sph_system.System("fp64")
    for call in (s.neighbours_search, s.calculate_forces, s.prepare, lambda: s.step(1)):
        with pytest.raises(sph_system.SphError) as e:
            call()
        assert e.value.code == 8  # SPH_ERR_STATE


def test_missing_input_is_an_error(tmp_path):
    s = sph_system.System("fp32")
    with pytest.raises(sph_system.SphError) as e:
        s.read_input(str(tmp_path / "nope"))
    assert e.value.code == 2  # SPH_ERR_IO


def test_unknown_field_and_wrong_count(tmp_path):
    import helpers

    s = helpers.make_system(helpers.cfg_couette(n=16), str(tmp_path), noise=0)
    lib = s.lib
    buf = (ctypes.c_double * 4)()
    assert lib.sph_get_field(s.h, b"nope", buf, 4) == 1
    assert lib.sph_get_field(s.h, b"x", buf, 4) == 1
    assert "entries" in lib.sph_last_error(s.h).decode()


def test_batch_host_side_checks(tmp_path):
    """Batch calls that need no GPU: membership rules and calls before prepare."""
    import helpers
    lib = sph_system.library()
    assert not lib.sph_batch_create(5)
    b = sph_system.Batch("fp32")
    with pytest.raises(sph_system.SphError, match="SPH_ERR_STATE"):
        b.step(1)
    with pytest.raises(sph_system.SphError, match="empty"):
        b.prepare()
    fresh = sph_system.System("fp32")
    with pytest.raises(sph_system.SphError, match="not initialized"):
        b.add(fresh)
    s2 = helpers.make_system(helpers.cfg_couette(n=16), str(tmp_path), noise=0, name="a")
    s3 = helpers.make_system(helpers.cfg_sine(n=10), str(tmp_path), noise=0, name="b")
    assert b.add(s2) == 0
    with pytest.raises(sph_system.SphError, match="share dim and wall"):
        b.add(s3)
    assert len(b) == 1
    with pytest.raises(sph_system.SphError, match="64, 128 or 256"):
        b.set_launch(100)
    b.set_launch(128, False)
    assert b.get_launch() == (128, False)
    # before prepare the batch serves its host copy
    np.testing.assert_array_equal(b.get(0, "x"), s2.get("x"))
