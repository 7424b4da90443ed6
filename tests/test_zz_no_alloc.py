"""aar_hip_pinv_noalloc: the allocation-free form of aar_hip_pinv (the m x m
Golub-Reinsch pseudo-inverse solve, tools.c:52-105 semantics) on device
pointers, with its work area in a static device buffer.

Through the C ABI, on two calls in a row (the second with a smaller
system), its x, singular values and V equal aar_hip_pinv's bit for bit, on
the reference's known-answer systems (SPD, a zero row and column,
near-singular); a call leaves the device's free memory unchanged.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _kats(golden):
    _, g = golden
    ms, A, b = g["pinv_m"], g["pinv_A"], g["pinv_b"]
    out = {}
    ia = ib = 0
    for m in ms:
        m = int(m)
        out.setdefault(m, []).append((A[ia:ia + m * m].reshape(m, m).copy(), b[ib:ib + m].copy()))
        ia += m * m
        ib += m
    return out


def _noalloc(L, A, b):
    import torch
    m = b.size
    Ad = torch.from_numpy(np.ascontiguousarray(A)).cuda()
    bd = torch.from_numpy(np.ascontiguousarray(b)).cuda()
    xd = torch.full((m,), np.nan, dtype=torch.float64, device="cuda")
    wd = torch.full((m,), np.nan, dtype=torch.float64, device="cuda")
    Vd = torch.full((m * m,), np.nan, dtype=torch.float64, device="cuda")
    sd = torch.full((1,), -1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    V = ctypes.c_void_p
    rc = L.aar_hip_pinv_noalloc(V(Ad.data_ptr()), V(bd.data_ptr()), V(xd.data_ptr()), m, V(wd.data_ptr()),
                                V(Vd.data_ptr()), V(sd.data_ptr()))
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    assert rc == 0, L.aar_hip_last_error()
    return xd.cpu().numpy(), wd.cpu().numpy(), Vd.cpu().numpy(), int(sd.cpu()[0]), free1 - free0


def _alloc(L, A, b):
    m = b.size
    x, w, V = np.zeros(m), np.zeros(m), np.zeros(m * m)
    A = np.ascontiguousarray(A)
    rc = L.aar_hip_pinv(A.ctypes.data, b.ctypes.data, x.ctypes.data, m, w.ctypes.data, V.ctypes.data)
    assert rc == 0
    return x, w, V


def test_pinv_noalloc_matches_pinv(gpu, golden):
    from aar_hip import _lib
    L = _lib.load()
    kats = _kats(golden)
    # the larger system first, then smaller ones: a call must not see what
    # the previous one left in the shared work area
    order = [(16, 0), (7, 0), (13, 1), (3, 2), (16, 2), (1, 0)]
    deltas = []
    for m, k in order:
        A, b = kats[m][k % len(kats[m])]
        x, w, V, st, dfree = _noalloc(L, A, b)
        xr, wr, Vr = _alloc(L, A, b)
        print(f"m={m} kat={k}: status {st}, free-memory change {dfree}", flush=True)
        assert st == 0
        assert np.array_equal(x, xr) and np.array_equal(w, wr) and np.array_equal(V, Vr), (m, k)
        deltas.append(dfree)
    # after the first call (which may load the code object) no call changes
    # the device's free memory
    assert all(d == 0 for d in deltas[1:]), deltas
