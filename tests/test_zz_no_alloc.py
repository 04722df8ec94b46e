"""dsm_stress_noalloc (no allocation; partial sums in a fixed device array) against
dsm_stress on the same chain state: two calls in a row, the second on fewer chains,
each bit-identical to the engine's result (same kernel, same grid, same summation
order), stalled chains and a pending deformation included."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _DeviceBuffers:
    """Device copies of an ensemble's state, allocated by the test through the HIP runtime
    the library itself runs on (the library was loaded first, so loading the runtime by
    its soname returns that same instance); chain c's strands sit at c * cap."""

    def __init__(self, st):
        import ctypes as C

        import gpu_dsm_amd as g
        g.load_library()
        self.C = C
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        n, cap = st["N"].shape
        host = dict(Z=np.ascontiguousarray(st["Z"], np.int32), stall=np.ascontiguousarray(st["stall"], np.int32),
                    base=np.arange(n, dtype=np.int64) * cap, dt_prev=np.ascontiguousarray(st["dt_prev"], np.float32),
                    Qx=np.ascontiguousarray(st["Q"][:, :, 0], np.float32).ravel(),
                    Qy=np.ascontiguousarray(st["Q"][:, :, 1], np.float32).ravel(),
                    Qz=np.ascontiguousarray(st["Q"][:, :, 2], np.float32).ravel(),
                    N=np.ascontiguousarray(st["N"], np.int32).ravel())
        self.addr = {}
        for k, a in host.items():
            p = C.c_void_p()
            assert self.hip.hipMalloc(C.byref(p), a.nbytes) == 0
            self.addr[k] = p.value
            assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device

    def free(self):
        for p in self.addr.values():
            self.hip.hipFree(p)
        self.addr = {}


def test_stress_noalloc_matches_dsm_stress():
    import gpu_dsm_amd as g
    from gpu_dsm_amd.engine import stress_noalloc
    kappa = np.array([0.01, 0.05, 0, -0.02, 0, 0.03, 0, 0, -0.01], np.float32)
    with g.Engine(0) as e:
        a = e.add_ensemble(5000, 60, 1.0, 1, kappa=kappa, seed=3)
        b = e.add_ensemble(700, 40, 0.5, 1, kappa=kappa, seed=4)
        for ens in (a, b):
            e.init_equilibrium(ens)
            e.advance(ens, 3.0)
        st = e.state(a)
        st["dt_prev"][::3] = 0.0625          # a deformation pending on every third chain
        st["stall"][[5, 17, 400]] = 1        # stalled chains are left out
        e.set_state(a, st["Z"], st["N"], st["Q"], st["w"], t=st["t"], ev=st["ev"], stall=st["stall"],
                    dt_prev=st["dt_prev"], next_sample=st["next_sample"])
        for ens, n in ((a, 5000), (b, 700)):    # the second call has the smaller input
            want, hw = e.stress(ens)
            bufs = _DeviceBuffers(e.state(ens))
            got, hg = stress_noalloc(bufs.addr, n, kappa)
            assert hg == hw == n - (3 if ens == a else 0)
            assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (got, want)
            bufs.free()
