"""The allocation-free control number (mfsv_control_number_noalloc) against
mfsv_control_number on the same states: bit-identical, FP64 and FP32, on two
calls in a row (the second smaller, with a smaller maximum, so that a block
maximum left in the shared scratch array by the first call would show), and
on a grid that needs more than one scratch block per row of blocks.  The
test allocates the states and the result word itself with the HIP runtime
(ctypes; plumbing only, the entry point allocates nothing), on the null
stream and on a stream of its own, and reads the result back after the
stream has finished."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-noalloc]", *a, flush=True)


def _state(nx, ny, seed, vscale):
    rng = np.random.default_rng(seed)
    h = np.where(rng.uniform(size=(ny, nx)) < 0.7, rng.uniform(0.0, 3.0, (ny, nx)), 0.0)
    u = rng.normal(0, vscale, (ny, nx))
    v = rng.normal(0, vscale, (ny, nx))
    h[0, 0] = -0.25           # a negative depth: c = sqrt(g max(h, 0)) = 0
    return h, u, v


def _ctx_cn(prec, nx, ny, dx, dy, h, u, v, dt):
    from mfcm_amd import mfsv
    g = mfsv.Grid(nx, ny, dx, dy, 0.0, 0.0)
    ctx = C.c_void_p()
    L = mfsv.lib()
    assert L.mfsv_create(C.byref(g), prec, 0, C.byref(ctx)) == 0
    for f, a in ((0, h), (1, u), (2, v)):
        a = np.ascontiguousarray(a, np.float64).ravel()
        assert L.mfsv_upload(ctx, f, a.ctypes.data_as(C.c_void_p), a.size) == 0
    cn = C.c_double()
    assert L.mfsv_control_number(ctx, 0, dt, C.byref(cn)) == 0
    L.mfsv_destroy(ctx)
    return cn.value


_HIP = None


def _hip():
    """The HIP runtime the library links (libamdhip64.so.7), for the test's
    own device buffers."""
    global _HIP
    if _HIP is None:
        H = C.CDLL("libamdhip64.so.7")
        for name in ("hipMalloc", "hipFree", "hipMemcpy", "hipStreamCreate", "hipStreamDestroy",
                     "hipStreamSynchronize"):
            getattr(H, name).restype = C.c_int
        H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        H.hipFree.argtypes = [C.c_void_p]
        H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        H.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        H.hipStreamDestroy.argtypes = [C.c_void_p]
        H.hipStreamSynchronize.argtypes = [C.c_void_p]
        _HIP = H
    return _HIP


def _dev(a):
    H = _hip()
    p = C.c_void_p()
    assert H.hipMalloc(C.byref(p), a.nbytes) == 0
    assert H.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0   # host to device
    return p


def _noalloc_cn(prec, nx, ny, dx, dy, h, u, v, dt, own_stream):
    from mfcm_amd import mfsv
    H = _hip()
    td = np.float64 if prec == mfsv.F64 else np.float32
    bufs = [_dev(np.ascontiguousarray(a, np.float64).astype(td)) for a in (h, u, v)]
    out = _dev(np.array([-1.0]))
    stream = C.c_void_p()
    if own_stream:
        assert H.hipStreamCreate(C.byref(stream)) == 0
    try:
        rc = mfsv.lib().mfsv_control_number_noalloc(prec, 0, nx, ny, dx, dy, bufs[0], bufs[1],
                                                    bufs[2], dt, out, stream)
        assert rc == 0
        assert H.hipStreamSynchronize(stream) == 0
        res = np.zeros(1)
        assert H.hipMemcpy(res.ctypes.data_as(C.c_void_p), out, 8, 2) == 0     # device to host
    finally:
        if own_stream:
            H.hipStreamDestroy(stream)
        for b in bufs + [out]:
            H.hipFree(b)
    return float(res[0])


@pytest.mark.parametrize("prec_name", ["f64", "f32"])
def test_noalloc_control_number_equals_context_call(gpu_lib, prec_name):
    from mfcm_amd import mfsv
    prec = mfsv.F64 if prec_name == "f64" else mfsv.F32
    cases = [
        (600, 500, 5.0, 3.0, 11, 4.0, 0.37),     # 300 000 cells: 19 scratch blocks
        (70, 30, 2.0, 2.5, 12, 0.5, 0.21),       # then a smaller grid with a smaller maximum
        (4100, 33, 1.0, 1.0, 13, 1.0, 0.05),     # rows longer than a block
    ]
    prev = None
    for k, (nx, ny, dx, dy, seed, vs, dt) in enumerate(cases):
        h, u, v = _state(nx, ny, seed, vs)
        want = _ctx_cn(prec, nx, ny, dx, dy, h, u, v, dt)
        got = _noalloc_cn(prec, nx, ny, dx, dy, h, u, v, dt, own_stream=k == 1)
        _say(f"{prec_name} {nx}x{ny}: noalloc {got!r} context {want!r}")
        assert got == want
        if prev is not None and nx * ny < prev[0]:
            assert want < prev[1]  # the case is meaningful: a stale maximum would be larger
        prev = (nx * ny, want)
