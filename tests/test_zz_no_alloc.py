"""drape_vertex_normals_noalloc (the allocation-free variant of drape_vertex_normals: caller device arrays in and
out, working memory in a fixed device array of the library) against drape_vertex_normals, through the C-ABI.
Two calls in a row, the second with a smaller mesh, must each equal the allocating entry point bit for bit."""
import ctypes as C
import os

import numpy as np
import pytest

from drape_amd import scenes

pytestmark = pytest.mark.gpu

_HIP = None


def _hip():
    global _HIP
    if _HIP is None:
        _HIP = C.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
        _HIP.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _HIP.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _HIP.hipFree.argtypes = [C.c_void_p]
    return _HIP


class _Dev:
    """A device copy of a host array (the caller's allocation, not the library's)."""

    def __init__(self, host: np.ndarray):
        self.nbytes = max(host.nbytes, 4)
        self.p = C.c_void_p()
        assert _hip().hipMalloc(C.byref(self.p), self.nbytes) == 0
        if host.nbytes:
            assert _hip().hipMemcpy(self.p, host.ctypes.data, host.nbytes, 1) == 0  # host to device

    def read(self, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        assert _hip().hipMemcpy(out.ctypes.data, self.p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        _hip().hipFree(self.p)


def _noalloc(hip, x, tri, welds):
    from drape_amd import _ffi
    lib = _ffi.load_library()
    dx, dt, dw = _Dev(x), _Dev(tri), _Dev(welds)
    out = np.zeros_like(x)
    do = _Dev(out)
    try:
        code = lib.drape_vertex_normals_noalloc(dx.p, len(x), dt.p, len(tri), dw.p, len(welds), do.p, None)
        if code != 0:
            raise {1: ValueError, 2: IndexError}.get(code, RuntimeError)(lib.drape_last_error().decode())
        return do.read(out)
    finally:
        for d in (dx, dt, dw, do):
            d.free()


def _welded_pair(rng, n=7):
    m = scenes.grid_mesh(n, n, 0.01)
    k = m.n
    tri = np.concatenate([m.triangles, m.triangles + k]).astype(np.uint32)
    x = np.concatenate([np.stack([m.rest[:, 0], np.zeros(k), m.rest[:, 1]], 1),
                        np.stack([m.rest[:, 0] + 0.06, np.zeros(k), m.rest[:, 1]], 1)])
    x = (x + rng.normal(scale=2e-3, size=x.shape)).astype(np.float32)
    welds = np.array([(j * n + n - 1, k + j * n) for j in range(n)], np.uint32)
    return x, tri, welds


def test_noalloc_normals_equal_the_allocating_entry_point_on_two_calls(hip):
    rng = np.random.default_rng(11)
    x, tri, welds = _welded_pair(rng)
    want = hip.vertex_normals(x, tri, welds)
    got = _noalloc(hip, x, tri, welds)
    print(f"call 1: {len(x)} vertices, {len(tri)} triangles, {len(welds)} welds", flush=True)
    np.testing.assert_array_equal(got, want)
    m = scenes.grid_mesh(5, 4, 0.02)  # second call, smaller mesh, no welds
    x2 = (np.stack([m.rest[:, 0], 0.1 * np.sin(8 * m.rest[:, 0]), m.rest[:, 1]], 1)).astype(np.float32)
    tri2 = m.triangles.astype(np.uint32)
    no = np.zeros((0, 2), np.uint32)
    print(f"call 2: {len(x2)} vertices, {len(tri2)} triangles", flush=True)
    np.testing.assert_array_equal(_noalloc(hip, x2, tri2, no), hip.vertex_normals(x2, tri2, no))


def test_noalloc_normals_at_the_sheet_size_and_its_limits(hip):
    sc = scenes.c2_sheet()
    g = sc.garments[0]
    x = g.initial_positions.astype(np.float32)
    x[:, 1] += (0.01 * np.sin(30 * x[:, 0]) * np.cos(20 * x[:, 2])).astype(np.float32)
    tri = g.mesh.triangles.astype(np.uint32)
    no = np.zeros((0, 2), np.uint32)
    print(f"sheet: {len(x)} vertices, {len(tri)} triangles", flush=True)
    np.testing.assert_array_equal(_noalloc(hip, x, tri, no), hip.vertex_normals(x, tri, no))
    big = np.zeros((262_145, 3), np.float32)
    with pytest.raises(ValueError, match="at most 262144 vertices"):
        _noalloc(hip, big, tri, no)
    bad = tri.copy()
    bad[17, 2] = len(x) + 5
    with pytest.raises(IndexError, match="out of range"):
        _noalloc(hip, x, bad, no)
