"""The entry point that allocates nothing (blpg_solve_shared_A_noalloc) against the shared-A
entry point it varies (blpg_solve_shared_A, FP64, no check): two calls in a row, the
second with fewer LPs, each bit-identical; and the scratch limit reported as an error.

The caller's device buffers come from the HIP runtime libblpg.so itself loaded (through
ctypes): a second HIP runtime in the process (torch bundles its own) would not see the
device once the first one holds it."""
import ctypes as C

import numpy as np
import pytest

import simplexprojects_amd as sa

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libblpg.so is linked against."""

    def __init__(self):
        sa.load_library()
        path = None
        with open("/proc/self/maps") as f:
            for line in f:
                if "libamdhip64.so" in line and "torch" not in line:
                    path = line.split()[-1]
                    break
        assert path, "libblpg.so did not load a HIP runtime"
        self.rt = C.CDLL(path)
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.held = []

    def put(self, a: np.ndarray) -> int:
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.held.append(p)
        assert self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p.value

    def get(self, p: int, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        assert self.rt.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def sync(self):
        assert self.rt.hipDeviceSynchronize() == 0

    def free(self):
        for p in self.held:
            self.rt.hipFree(p)
        self.held = []


def _noalloc(hip, lib, A, b, Cm):
    m, n = A.shape
    k = Cm.shape[0]
    dA, db, dC = hip.put(A), hip.put(b), hip.put(Cm)
    like = {"st": np.zeros(k, np.int32), "obj": np.zeros(k), "x": np.zeros((k, n)),
            "bs": np.zeros((k, m), np.int32), "piv": np.zeros(k, np.int32)}
    # outputs start as garbage (not zero): every entry must be written by the call
    d = {key: hip.put(np.full_like(v, -7)) for key, v in like.items()}
    rc = lib.blpg_solve_shared_A_noalloc(m, n, dA, db, None, dC, k, d["st"], d["obj"], d["x"],
                                         d["bs"], d["piv"], None)
    assert rc == 0, lib.blpg_last_error(None)
    hip.sync()
    out = {key: hip.get(d[key], like[key]) for key in like}
    hip.free()
    return out["st"], out["obj"], out["x"], out["bs"], out["piv"]


@pytest.mark.parametrize("shape", [(27, 32, 1), (105, 103, 4)])
def test_noalloc_matches_shared_A(shape):
    m, n, cfg = shape
    lib = sa.load_library()
    hip = _Hip()
    if cfg == 1:
        lp0 = sa.gen_config(1, 0, m, n, 1, batch_count=1024)[0]
    else:
        lp0 = sa.gen_config(4, 4, m, n, 1, shape_idx=3, batch_count=1024)[0]
    A, b = lp0.A, lp0.b
    rng = np.random.default_rng(17)
    S = sa.Solver(precision="fp64", verify=False)
    for k in (700, 131):  # second call smaller: stale scratch from the first must not leak
        Cm = rng.uniform(-0.2, 1, size=(k, n))
        ref = S.solve_shared_A(A, b, Cm)
        st, obj, x, bs, piv = _noalloc(hip, lib, A, b, Cm)
        assert np.array_equal(st, ref.status)
        assert np.array_equal(piv, ref.pivots)
        assert np.array_equal(obj, ref.objective)
        for i in range(k):
            assert np.array_equal(x[i], ref.x[i]), i
            assert np.array_equal(bs[i], ref.basis[i]), i
    S.close()


def test_noalloc_limits():
    lib = sa.load_library()
    cap = lib.blpg_noalloc_max_count(105, 103)
    assert 10000 < cap < 100000
    # both calls are refused before any device access (the pointers are never read)
    z = C.c_void_p(0x1000).value
    rc = lib.blpg_solve_shared_A_noalloc(105, 103, z, z, None, z, cap + 1, z, z, z, z, None, None)
    assert rc == sa.E_BUDGET
    assert b"scratch" in lib.blpg_last_error(None)
    rc = lib.blpg_solve_shared_A_noalloc(300, 300, z, z, None, z, 1, z, z, z, z, None, None)
    assert rc == sa.E_INVALID
