"""The allocation-free occupied-factor J/K entry point
(vibeqc_fock_plan_evaluate_occupied_noalloc) against the allocating one
(vibeqc_fock_plan_evaluate_occupied) on the same plan: identical kernels and
split-K schedule, so J and K agree bit for bit.  Two calls in a row on one
plan (the second with fewer occupied columns), then a smaller plan in BF16X3
that reuses the same per-device scratch array, then a workspace that does not
fit (INVALID_ARGUMENT before any device work)."""

from __future__ import annotations

import ctypes as ct

import numpy as np
import pytest

from _cases import molecule_case

pytestmark = pytest.mark.gpu


class _DeviceBuffer:
    """Caller-side device memory from the HIP runtime the engine links (the
    test allocates the inputs and outputs; the entry point allocates nothing)."""

    hip = None

    def __init__(self, host: np.ndarray):
        if _DeviceBuffer.hip is None:
            lib = ct.CDLL("/opt/rocm/lib/libamdhip64.so")
            lib.hipMalloc.argtypes = [ct.POINTER(ct.c_void_p), ct.c_size_t]
            lib.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
            lib.hipFree.argtypes = [ct.c_void_p]
            _DeviceBuffer.hip = lib
        self.host = np.ascontiguousarray(host, dtype=np.float64)
        self.ptr = ct.c_void_p()
        assert self.hip.hipMalloc(ct.byref(self.ptr), self.host.nbytes) == 0
        assert self.hip.hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, 1) == 0  # host to device

    def read(self) -> np.ndarray:
        out = np.empty_like(self.host)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def _noalloc(plan, C):
    n, nocc = C.shape
    dc, J, K = _DeviceBuffer(C), _DeviceBuffer(np.full((n, n), np.nan)), _DeviceBuffer(np.full((n, n), np.nan))
    try:
        status = plan._lib.vibeqc_fock_plan_evaluate_occupied_noalloc(plan._handle, dc.ptr, n, nocc, J.ptr, K.ptr)
        return status, J.read(), K.read()
    finally:
        for b in (dc, J, K):
            b.free()


def _plan(case, precision="fp64"):
    from vibeqc_amd import FockPlan

    return FockPlan(case.orbital(), auxiliary=case.auxiliary(), precision=precision)


def test_noalloc_matches_the_allocating_entry_point():
    from vibeqc_amd.geometry import lowdin_occupied, water_cluster

    case = molecule_case("water4", water_cluster(4, 6.0, seed=7))
    with _plan(case) as plan:
        S, _ = plan.one_electron()
        for nocc in (20, 12):  # the second call has the smaller input
            C = lowdin_occupied(S, nocc, seed=nocc)
            status, J, K = _noalloc(plan, C)
            assert status == 0, status
            j, k = plan.evaluate_occupied(C)
            print(f"(H2O)4 nocc={nocc}: max |dJ| {np.abs(J - j).max():.1e} max |dK| {np.abs(K - k).max():.1e}")
            assert np.array_equal(J, j) and np.array_equal(K, k)
    small = molecule_case("water", water_cluster(1, 3.0, seed=3))
    with _plan(small, "bf16x3") as plan:
        S, _ = plan.one_electron()
        C = lowdin_occupied(S, 5, seed=1)
        status, J, K = _noalloc(plan, C)
        assert status == 0, status
        j, k = plan.evaluate_occupied(C)
        assert np.array_equal(J, j) and np.array_equal(K, k)


def test_workspace_beyond_the_scratch_is_refused():
    from vibeqc_amd.geometry import lowdin_occupied, water_cluster

    case = molecule_case("water16", water_cluster(16, 3.0 * 16 ** (1 / 3), seed=4))
    with _plan(case) as plan:
        S, _ = plan.one_electron()
        C = lowdin_occupied(S, 40, seed=2)
        status, J, K = _noalloc(plan, C)
        assert status == 1
        assert np.isnan(J).all() and np.isnan(K).all()
        assert b"scratch" in plan._lib.generativeqc_fock_plan_last_error(plan._handle)
