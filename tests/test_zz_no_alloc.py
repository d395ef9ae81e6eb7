"""The no-allocation LU entry point (hsbem_dense_lu_solve_noalloc) against the
allocating one (hsbem_dense_lu_solve), both through the C ABI.

The no-allocation form takes caller-owned device buffers; the test allocates
them with the HIP runtime the library itself links (libamdhip64.so.7), so the
pointers belong to the same runtime instance. Both forms run the same kernels,
so their solutions agree to the last bits; the bar written here is 1e-13
relative max-norm (FP64 arithmetic), and numpy's solve is checked to 1e-10.

Calls run back to back on one context so the second (smaller) call reuses the
scratch array the first left behind, a singular matrix reports its first
singular pivot, and the call after it starts from a cleared status word.
"""
import ctypes

import numpy as np
import pytest

from hornsculpt_amd import bem

pytestmark = pytest.mark.gpu

_H2D, _D2H = 1, 2


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libhsbem.so is linked against."""

    def __init__(self):
        bem.library()
        self.lib = ctypes.CDLL("libamdhip64.so.7")
        self.lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.lib.hipFree.argtypes = [ctypes.c_void_p]
        self.lib.hipDeviceSynchronize.argtypes = []

    def upload(self, a: np.ndarray) -> ctypes.c_void_p:
        ptr = ctypes.c_void_p()
        assert self.lib.hipMalloc(ctypes.byref(ptr), a.nbytes) == 0
        assert self.lib.hipMemcpy(ptr, a.ctypes.data, a.nbytes, _H2D) == 0
        return ptr

    def download(self, ptr: ctypes.c_void_p, out: np.ndarray) -> np.ndarray:
        assert self.lib.hipMemcpy(out.ctypes.data, ptr, out.nbytes, _D2H) == 0
        return out

    def free(self, ptr: ctypes.c_void_p) -> None:
        assert self.lib.hipFree(ptr) == 0


def _noalloc_solve(hip, ctx, A, B):
    """Column-major interleaved copies on the device, the no-allocation call, solutions back."""
    lib = bem.library()
    n, nrhs = B.shape
    Af = np.ascontiguousarray(A.T).astype(np.complex128)
    Bf = np.ascontiguousarray(B.T).astype(np.complex128)
    dA, dB = hip.upload(Af), hip.upload(Bf)
    try:
        assert hip.lib.hipDeviceSynchronize() == 0
        status = lib.hsbem_dense_lu_solve_noalloc(ctx.handle, n, ctypes.cast(dA, ctypes.POINTER(ctypes.c_double)),
                                                  nrhs, ctypes.cast(dB, ctypes.POINTER(ctypes.c_double)))
        X = hip.download(dB, np.empty_like(Bf)).T.copy()
        message = lib.hsbem_last_error().decode() if status else ""
    finally:
        hip.free(dA)
        hip.free(dB)
    return status, message, X


def _system(n, nrhs, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    B = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))
    return A, B


def test_noalloc_lu_equals_allocating_lu_on_consecutive_calls():
    hip = _Hip()
    ctx = bem.Context.default()
    for n, nrhs, seed in ((300, 3, 11), (130, 2, 12)):  # second call smaller than the first
        A, B = _system(n, nrhs, seed)
        print(f"no-alloc LU n={n} nrhs={nrhs}", flush=True)
        status, message, X = _noalloc_solve(hip, ctx, A, B)
        assert status == bem.HSBEM_OK, message
        Xa = bem.dense_lu_solve(A, B, ctx)
        assert np.max(np.abs(X - Xa)) / np.max(np.abs(Xa)) <= 1e-13
        Xn = np.linalg.solve(A, B)
        assert np.max(np.abs(X - Xn)) / np.max(np.abs(Xn)) <= 1e-10

    # A zero column: the status word reports pivot 41 ...
    A, B = _system(200, 1, 13)
    A[:, 40] = 0
    print("no-alloc LU singular n=200", flush=True)
    status, message, _ = _noalloc_solve(hip, ctx, A, B)
    assert status == bem.HSBEM_RUNTIME_ERROR and "pivot 41" in message
    # ... and the next call starts from a cleared status word.
    A, B = _system(64, 1, 14)
    print("no-alloc LU n=64 after the singular call", flush=True)
    status, message, X = _noalloc_solve(hip, ctx, A, B)
    assert status == bem.HSBEM_OK, message
    Xa = bem.dense_lu_solve(A, B, ctx)
    assert np.max(np.abs(X - Xa)) / np.max(np.abs(Xa)) <= 1e-13


def test_noalloc_lu_refuses_orders_beyond_its_scratch():
    lib = bem.library()
    ctx = bem.Context.default()
    dummy = ctypes.cast(ctypes.c_void_p(256), ctypes.POINTER(ctypes.c_double))  # never dereferenced: refused first
    status = lib.hsbem_dense_lu_solve_noalloc(ctx.handle, 65537, dummy, 1, dummy)
    assert status == bem.HSBEM_INVALID_ARGUMENT
    assert "no-allocation limit" in lib.hsbem_last_error().decode()
