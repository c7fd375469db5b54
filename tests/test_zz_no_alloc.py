"""lulesh_mlp_gradient_noalloc, the gradient entry point that allocates
nothing, called through the C ABI and checked against lulesh_mlp_gradient:
the same gradient and loss bit for bit on two calls in a row (the largest
accepted count, then a smaller one, so the second call runs over what the
first left in the library's buffer), and a count above the maximum refused
with an error.  The test itself allocates its device buffers through the HIP
runtime the library is linked against."""
import ctypes

import numpy as np
import pytest

import mlp_oracle as mo

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def hip_runtime():
    """The HIP runtime already loaded by the library (same soname)."""
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMalloc.restype = ctypes.c_int
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipFree.restype = ctypes.c_int
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                              ctypes.c_int]
    hip.hipMemcpy.restype = ctypes.c_int
    hip.hipDeviceSynchronize.argtypes = []
    hip.hipDeviceSynchronize.restype = ctypes.c_int
    return hip


class DeviceBuffer:
    def __init__(self, hip, nbytes):
        self.hip = hip
        self.ptr = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), nbytes) == 0
        self.nbytes = nbytes

    def upload(self, a: np.ndarray):
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D) == 0

    def download(self, a: np.ndarray):
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(a.ctypes.data, self.ptr, a.nbytes, D2H) == 0

    def free(self):
        if self.ptr:
            assert self.hip.hipFree(self.ptr) == 0
            self.ptr = ctypes.c_void_p()


def test_noalloc_gradient_matches_gradient(lulesh_gpu):
    lib = lulesh_gpu.lib()
    net_params = mo.init_params(3)
    net_params[mo.N_PARAMS - 2:] = [0.1, -0.4]
    rng = np.random.default_rng(11)
    mean = rng.normal(size=12).astype(np.float32)
    std = rng.uniform(0.5, 2.0, 12).astype(np.float32)
    net = lulesh_gpu.Mlp(net_params, mean, std)
    n_params = net.n_params
    cap = 16384  # LULESH_MLP_NOALLOC_MAX
    hip = hip_runtime()
    xd = DeviceBuffer(hip, cap * 12 * 4)
    td = DeviceBuffer(hip, cap * 2 * 4)
    gd = DeviceBuffer(hip, n_params * 4)
    ld = DeviceBuffer(hip, 4)
    weight = (1.0, 0.25)
    try:
        for count in (cap, 777):
            x = (rng.normal(size=(count, 12)) * std + mean).astype(np.float32)
            t = rng.normal(size=(count, 2)).astype(np.float32)
            g_ref, loss_ref = net.gradient(x, t, weight)
            xd.upload(x)
            td.upload(t)
            poison = np.full(n_params, np.nan, np.float32)
            gd.upload(poison)
            ld.upload(np.full(1, np.nan, np.float32))
            print(f"noalloc gradient: {count} samples", flush=True)
            net.gradient_noalloc(xd.ptr.value, td.ptr.value, count, gd.ptr.value,
                                 ld.ptr.value, 0, weight)
            assert hip.hipDeviceSynchronize() == 0
            g = np.empty(n_params, np.float32)
            gd.download(g)
            loss = np.empty(1, np.float32)
            ld.download(loss)
            assert np.array_equal(g.view(np.uint32), g_ref.view(np.uint32)), count
            assert float(loss[0]) == loss_ref, (count, float(loss[0]), loss_ref)
        # above the maximum: refused on the host with a message, no launch
        with pytest.raises(lulesh_gpu.LuleshError) as err:
            net.gradient_noalloc(xd.ptr.value, td.ptr.value, cap + 1, gd.ptr.value,
                                 ld.ptr.value, 0, weight)
        assert err.value.code == lulesh_gpu.BAD_ARG
        assert "maximum" in lib.lulesh_mlp_last_error(net.h).decode()
    finally:
        for b in (xd, td, gd, ld):
            b.free()
        net.close()
