"""Time-series extractors over a device-resident case: one CSV line ``time,value`` per write.

Mirrors the reference's extractors (``srcs/simulation/extractors/Extractor.cpp:5-13``,
``MassExtractor.cpp:8-38``, ``MinMaxExtractor.cpp:6-61``): the file is created (truncated) when
the extractor is built, a line is written when the simulation time has reached the next
trigger (or when ``force`` is given), and the trigger then advances by one interval (never
on a forced write). Both numbers are printed as C++'s ``std::to_string(double)`` prints them,
which is ``printf("%f")``: six digits after the point.

The values come from the case on the GPU: ``Case.mass()`` (``pfem_case_mass``) and
``Case.minmax()`` (``pfem_case_minmax``). The reference reads the simulation time from its
``Problem``; the caller owns the time loop here, so ``update`` takes the time.
"""

from __future__ import annotations

from ._lib import PfemError, PFEM_E_INVALID


def _fmt(x: float) -> str:
    """``std::to_string(double)``: ``%f`` formatting (nan / inf spelled as C prints them)."""
    return "%f" % x


class Extractor:
    """Common trigger logic and the output file (``Extractor.cpp:5-13``)."""

    _file = None  # set once the file is open; a constructor that raises leaves it unset

    def __init__(self, case, out_file_name: str, time_between_writing: float):
        if not time_between_writing > 0:
            raise PfemError(PFEM_E_INVALID, "the interval between write should be strictly greater than 0!")
        self.case = case
        self.out_file_name = out_file_name
        self.time_between_writing = float(time_between_writing)
        self.next_write_trigger = 0.0

    def _open(self) -> None:
        try:
            self._file = open(self.out_file_name, "w")
        except OSError as e:
            raise PfemError(PFEM_E_INVALID, "cannot open file to write point extractor: " + self.out_file_name) from e

    def value(self) -> float:
        raise NotImplementedError

    def update(self, sim_time: float, force: bool = False) -> bool:
        """Write ``time,value`` if ``sim_time`` reached the trigger or ``force``; True if written."""
        if sim_time < self.next_write_trigger and not force:
            return False
        self._file.write(_fmt(sim_time) + "," + _fmt(self.value()) + "\n")
        self._file.flush()  # std::endl flushes every line
        if not force:
            self.next_write_trigger += self.time_between_writing
        return True

    def close(self) -> None:
        if self._file is not None:
            self._file.close()
            self._file = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()


class MassExtractor(Extractor):
    """The problem's global mass (``MassExtractor.cpp:8-38``); both PSPG and WC cases have one
    (``IncompNewton/Problem.cpp:168-180``, ``WCompNewton/Problem.cpp:211-224``)."""

    def __init__(self, case, out_file_name: str, time_between_writing: float):
        super().__init__(case, out_file_name, time_between_writing)
        self._open()

    def value(self) -> float:
        return self.case.mass()


class MinMaxExtractor(Extractor):
    """Smallest or largest coordinate ``coordinate`` over the nodes neither bound nor free
    (``MinMaxExtractor.cpp:6-61``; the maximum starts from 0, as there)."""

    def __init__(self, case, out_file_name: str, time_between_writing: float, coordinate: int, min_max: str):
        super().__init__(case, out_file_name, time_between_writing)
        if coordinate < 0 or coordinate > case.dim - 1:
            raise PfemError(PFEM_E_INVALID, "bad coordinate to write!")
        if min_max not in ("min", "max"):
            raise PfemError(PFEM_E_INVALID, "unknown data to compute for MinMaxExtractor.")
        self.coordinate = int(coordinate)
        self.min_max = min_max
        self._open()

    def value(self) -> float:
        return self.case.minmax(self.coordinate, self.min_max)
