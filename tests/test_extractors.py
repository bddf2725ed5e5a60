"""Extractor CSV output (MassExtractor.cpp, MinMaxExtractor.cpp, Extractor.cpp) on the CPU:
the write trigger, the number format and the argument checks, with a stand-in case object
(the device values themselves are checked against the oracle in the GPU tests)."""
import ctypes
import ctypes.util

import pytest

from pfem_amd import PfemError
from pfem_amd.extractors import MassExtractor, MinMaxExtractor


class StubCase:
    """Duck-typed case: returns a value that changes with each call."""

    def __init__(self, dim=2):
        self.dim = dim
        self.calls = 0

    def mass(self):
        self.calls += 1
        return 1.5 * self.calls

    def minmax(self, coordinate, kind):
        self.calls += 1
        return (-1.0 if kind == "min" else 1.0) * (coordinate + 0.25 * self.calls)


def c_to_string(x: float) -> str:
    """std::to_string(double) is specified as sprintf("%f"): ask the C library."""
    libc = ctypes.CDLL(ctypes.util.find_library("c"))
    buf = ctypes.create_string_buffer(512)
    libc.snprintf(buf, 512, b"%f", ctypes.c_double(x))
    return buf.value.decode()


def test_number_format_matches_c_printf(tmp_path):
    vals = [0.0, 1.0, -2.5, 1e-7, 4.9999995e-7, 123456.7890125, 1e300, -0.0, 0.1 + 0.2]

    class Seq(StubCase):
        def mass(self):
            self.calls += 1
            return vals[self.calls - 1]

    path = tmp_path / "m.csv"
    with MassExtractor(Seq(), str(path), 1.0) as ex:
        for v in vals:
            assert ex.update(0.0, force=True)
    lines = path.read_text().splitlines()
    assert lines == [c_to_string(0.0) + "," + c_to_string(v) for v in vals]


def test_trigger_advances_one_interval_per_write(tmp_path):
    path = tmp_path / "mass.csv"
    case = StubCase()
    ex = MassExtractor(case, str(path), 0.1)
    times = [0.0, 0.05, 0.1, 0.12, 0.35, 0.36, 0.38, 0.5]
    written = [ex.update(t) for t in times]
    # trigger 0 -> write at 0.0 (next 0.1); 0.05 skip; 0.1 write (next 0.2); 0.12 skip;
    # 0.35 write (next 0.3: one interval, no catch-up); 0.36 write (next 0.4); 0.38 skip;
    # 0.5 write (next 0.5)
    assert written == [True, False, True, False, True, True, False, True]
    # a forced write does not move the trigger
    assert ex.update(0.45, force=True)
    assert not ex.update(0.49)
    ex.close()
    lines = path.read_text().splitlines()
    assert [line.split(",")[0] for line in lines] == ["%f" % t for t in (0.0, 0.1, 0.35, 0.36, 0.5, 0.45)]
    assert [float(line.split(",")[1]) for line in lines] == [1.5 * k for k in range(1, 7)]
    assert case.calls == 6  # the value is read only when a line is written


def test_minmax_extractor_writes_requested_coordinate(tmp_path):
    path = tmp_path / "mm.csv"
    case = StubCase(dim=3)
    with MinMaxExtractor(case, str(path), 1.0, 2, "min") as ex:
        ex.update(0.0)
        ex.update(2.0)
    assert path.read_text().splitlines() == ["0.000000,-2.250000", "2.000000,-2.500000"]


def test_file_is_truncated_on_creation(tmp_path):
    path = tmp_path / "m.csv"
    path.write_text("old\nold\n")
    MassExtractor(StubCase(), str(path), 1.0).close()
    assert path.read_text() == ""


@pytest.mark.parametrize(
    "make",
    [
        lambda p: MassExtractor(StubCase(), p, 0.0),
        lambda p: MassExtractor(StubCase(), p, -1.0),
        lambda p: MinMaxExtractor(StubCase(dim=2), p, 1.0, 2, "max"),
        lambda p: MinMaxExtractor(StubCase(dim=3), p, 1.0, -1, "max"),
        lambda p: MinMaxExtractor(StubCase(dim=3), p, 1.0, 0, "mean"),
        lambda p: MassExtractor(StubCase(), p + "/no/such/dir/m.csv", 1.0),
    ],
)
def test_invalid_arguments_raise(tmp_path, make):
    with pytest.raises(PfemError):
        make(str(tmp_path / "x.csv"))
