"""Checkpoint / restart files (CPU): the reference's RestartObject unit tests restated, the
Python and C++ writers producing identical files, and the error paths.

The reference layout (src/gpu/core/Restart/RestartObject.cpp:123-201): binary = int32 level
count, then per level int32 count and the float values; ASCII = one line per level, each value
printed with the C++ stream defaults (%g) and followed by a space.
"""
import os
import struct
import subprocess
import warnings

import numpy as np
import pytest

from virtualfluids_amd.restart import (ASCIIRestartObject, BinaryRestartObject, RestartObject,
                                       restart_file_name)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST_TEST = os.path.join(ROOT, "tests", "cpp", "test_host")


@pytest.mark.parametrize("cls", [ASCIIRestartObject, BinaryRestartObject])
def test_save_and_load(tmp_path, cls):
    """RestartObjectTests.saveAndLoad_ascii / _binary."""
    w = cls()
    w.fs = [np.array([1, 2, 3], np.float32), np.array([4, 5, 6], np.float32)]
    name = str(tmp_path / "test_do_check_point")
    w.serialize_internal(name)
    r = cls()
    r.deserialize_internal(name)
    assert [a.tolist() for a in r.fs] == [[1, 2, 3], [4, 5, 6]]
    w.delete_restart_file(name)
    assert not os.path.exists(name + cls.extension)


@pytest.mark.parametrize("cls", [ASCIIRestartObject, BinaryRestartObject])
def test_deleting_a_missing_file_warns(tmp_path, cls):
    with warnings.catch_warnings(record=True) as rec:
        warnings.simplefilter("always")
        cls().delete_restart_file(str(tmp_path / "does_not_exist"))
    assert any("Could not delete" in str(w.message) for w in rec)


def test_binary_layout(tmp_path):
    w = BinaryRestartObject()
    w.fs = [np.array([1.5, -2.0], np.float32), np.array([], np.float32), np.array([3.25], np.float32)]
    w.serialize_internal(str(tmp_path / "x"))
    raw = (tmp_path / "x.bin").read_bytes()
    assert raw == struct.pack("<i", 3) + struct.pack("<iff", 2, 1.5, -2.0) + struct.pack("<i", 0) + \
        struct.pack("<if", 1, 3.25)


def test_ascii_values_have_six_significant_digits(tmp_path):
    v = np.array([0.0123456789, -1e-7 / 3, 123456.789, 0.0, -0.0, 2.5e-5], np.float32)
    w = ASCIIRestartObject()
    w.fs = [v]
    w.serialize_internal(str(tmp_path / "a"))
    text = (tmp_path / "a.txt").read_text()
    assert text == "0.0123457 -3.33333e-08 123457 0 -0 2.5e-05 \n"
    r = ASCIIRestartObject()
    r.deserialize_internal(str(tmp_path / "a"))
    np.testing.assert_array_equal(r.fs[0], np.array([float("%g" % x) for x in v.astype(np.float64)], np.float32))


def test_python_and_cpp_writers_give_identical_files(tmp_path):
    d = str(tmp_path)
    r = subprocess.run([HOST_TEST, "restart", d], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Could not delete" in r.stdout  # the reference prints a warning for a missing file
    src = BinaryRestartObject()
    src.deserialize_internal(os.path.join(d, "cpp"))
    assert [a.size for a in src.fs] == [27 * 7, 27 * 3]
    for cls in (BinaryRestartObject, ASCIIRestartObject):
        w = cls()
        w.fs = src.fs
        w.serialize_internal(os.path.join(d, "py"))
        assert open(os.path.join(d, "py" + cls.extension), "rb").read() == \
            open(os.path.join(d, "cpp" + cls.extension), "rb").read(), cls.__name__
    # and the C++ text parses to the same values the Python text does
    a = ASCIIRestartObject()
    a.deserialize_internal(os.path.join(d, "cpp"))
    for x, y in zip(a.fs, src.fs):
        np.testing.assert_array_equal(x, np.array([float("%g" % v) for v in y.astype(np.float64)], np.float32))


def test_file_name():
    assert restart_file_name("out/cavity", 500, 0) == "out/cavity_Restart_0_500"


class _Level:
    def __init__(self, n):
        self.N = n
        self.got = None

    def set_populations(self, f):
        self.got = f


def test_deserialize_errors(tmp_path):
    with pytest.raises(FileNotFoundError):
        BinaryRestartObject().deserialize(str(tmp_path / "missing"), [_Level(1)])
    w = BinaryRestartObject()
    w.fs = [np.zeros(27 * 2, np.float32)]
    w.serialize_internal(str(tmp_path / "one"))
    with pytest.raises(ValueError, match="holds 1 level"):
        BinaryRestartObject().deserialize(str(tmp_path / "one"), [_Level(2), _Level(1)])
    with pytest.raises(ValueError, match="expected 81"):
        BinaryRestartObject().deserialize(str(tmp_path / "one"), [_Level(3)])
    ok = _Level(2)
    BinaryRestartObject().deserialize(str(tmp_path / "one"), [ok])
    assert ok.got.size == 54
    raw = (tmp_path / "one.bin").read_bytes()
    (tmp_path / "cut.bin").write_bytes(raw[:-4])
    with pytest.raises(ValueError, match="truncated"):
        BinaryRestartObject().deserialize_internal(str(tmp_path / "cut"))


def test_base_class_is_abstract(tmp_path):
    with pytest.raises(NotImplementedError):
        RestartObject().serialize_internal(str(tmp_path / "x"))
