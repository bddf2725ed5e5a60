"""pfem_case_mass_noalloc: the allocation-free variant of pfem_case_mass (partials in the
library's fixed device scratch, the sum written to a caller's device double) gives the same
bits as pfem_case_mass, on two calls in a row with the second on a smaller mesh, for a PSPG
case and a WC case (density-weighted, FP32 state)."""
import ctypes as C
import sys

import numpy as np
import pytest

import pfem_amd
from pfem_amd import _lib
from pfem_amd import meshgen as mg

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class DeviceDouble:
    """One double of device memory from the HIP runtime the library uses (the caller's output
    buffer; the test allocates it, the entry point does not)."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), 8) == 0

    def read(self):
        v = C.c_double(0.0)
        assert self.hip.hipMemcpy(C.byref(v), self.ptr, 8, 2) == 0  # hipMemcpyDeviceToHost
        return v.value

    def free(self):
        self.hip.hipFree(self.ptr)


def mass_noalloc(case, out):
    _lib.check(_lib.lib().pfem_case_mass_noalloc(case._h, out.ptr))
    _lib.check(_lib.lib().pfem_case_sync(case._h))
    return out.read()


def test_mass_noalloc_equals_mass_two_calls_second_smaller(ctx):
    out = DeviceDouble()
    big = mg.dambreak_3d(h=0.02, col=(0.5, 0.5, 1.0), seed=2)
    small = mg.dambreak_2d(h=0.1)
    for mesh in (big, small):  # the second call has the smaller input
        c = pfem_amd.Case(ctx, mesh.dim, kind="PSPG")
        c.set_mesh(mesh)
        ref = c.mass()
        got = mass_noalloc(c, out)
        say(f"[no-alloc PSPG] E={mesh.n_elems}: {got!r} vs {ref!r}")
        assert got == ref
        c.close()
    K0, K0P, RS = 2.2e6, 7.0, 1000.0
    for mesh in (mg.tank_3d(h=0.05, size=(0.6, 0.4, 0.3), fill=0.7, seed=4),
                 mg.tank_2d(h=0.05, size=(1.0, 0.5), fill=0.6, seed=2)):
        c = pfem_amd.Case(ctx, mesh.dim, kind="WC", K0=K0, K0p=K0P, rho_star=RS, wc_state="fp32")
        c.set_mesh(mesh)
        v, p, rho, a = mg.wc_tait_hydrostatic(mesh, K0=K0, K0p=K0P, rho_star=RS)
        c.set_state(np.concatenate([v.reshape(-1), p, rho, a.reshape(-1)]))
        ref = c.mass()
        got = mass_noalloc(c, out)
        say(f"[no-alloc WC] E={mesh.n_elems}: {got!r} vs {ref!r}")
        assert got == ref
        c.close()
    out.free()
