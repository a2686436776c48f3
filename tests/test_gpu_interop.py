"""Render interop step on the GPU: copyTransformsToBVO
(opencl/broadphase_benchmark/findPairsOpenCL.cpp:140-160, kernel
broadphaseKernel.cl:236-247) writes each body's (x, y, z, 1) and quaternion
into a device float4 buffer, which must equal the readback of the same
state."""
import ctypes

import numpy as np
import pytest

from experiments_amd import _lib as L
from experiments_amd import scenes

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def test_copy_transforms_matches_readback():
    torch.zeros(1, device="cuda")  # torch's runtime first: the buffer comes from it
    w = scenes.build_world(scenes.lattice_cubes(4, 3, 4, spacing=1.05, gap=0.5))
    w.step(7)
    n = w.num_bodies
    buf = torch.full((2 + 2 * n + 3, 4), -7.0, dtype=torch.float32, device="cuda")
    w.copy_transforms(buf, start_offset=8)
    torch.cuda.synchronize()
    got = buf.cpu().numpy()
    st = w.get_state()
    assert np.all(got[:2] == -7.0) and np.all(got[2 + 2 * n:] == -7.0)
    assert np.array_equal(got[2:2 + n, :3], st["pos"][:, :3])
    assert np.all(got[2:2 + n, 3] == 1.0)
    assert np.array_equal(got[2 + n:2 + 2 * n], st["quat"])
    # a buffer one float4 short is refused, and a misaligned offset
    small = torch.zeros((2 * n - 1, 4), dtype=torch.float32, device="cuda")
    st_code = L.lib().mi_rb_copy_transforms(w.h, ctypes.c_void_p(small.data_ptr()), 0, 2 * n - 1, None)
    assert st_code == L.CAPACITY
    st_code = L.lib().mi_rb_copy_transforms(w.h, ctypes.c_void_p(buf.data_ptr()), 2, 2 * n + 5, None)
    assert st_code == L.INVALID_ARG
    w.close()
