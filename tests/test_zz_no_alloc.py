"""The allocation-free weight gradient (sfac_conv3x3x3_wgrad_bf16_noalloc)
against the workspace entry point sfac_conv3x3x3_wgrad_bf16, through the C
ABI: bit-identical gradients on two calls in a row (the maximum block count,
then a smaller one, so the second call runs on slabs the first one left), and
a clear error for a block count beyond the static scratch maximum."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_BLOCKS = 16      # SFAC_WGRAD_NOALLOC_MAX_BLOCKS
GSZ = 28 * 16 * 16   # gradient slab: 27 taps + the bias row, 16 x 16 channels
CELLS = 16 ** 3      # cells per block
CH = 16              # channels per cell


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _bind(lib):
    vp, i = C.c_void_p, C.c_int
    lib.sfac_conv3x3x3_wgrad_bf16.argtypes = [vp, vp, vp, vp, i, vp]
    lib.sfac_conv3x3x3_wgrad_partial_size.argtypes = [i]
    lib.sfac_conv3x3x3_wgrad_partial_size.restype = C.c_size_t
    lib.sfac_conv3x3x3_wgrad_bf16_noalloc.argtypes = [vp, vp, vp, i, vp]
    lib.sfac_last_error.restype = C.c_char_p
    return lib


def _bf16_bits(torch, rng, nb):
    a = torch.from_numpy(rng.standard_normal((nb, CELLS, CH)).astype(np.float32))
    return a.to(torch.bfloat16).view(torch.int16).cuda().contiguous()


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def test_noalloc_wgrad_matches_workspace_entry_point(torch_gpu, gpu_lib):
    torch = torch_gpu
    lib = _bind(gpu_lib)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rng = np.random.default_rng(77)
    for call, nb in enumerate((MAX_BLOCKS, 3)):
        _say("no-alloc weight gradient, call %d: %d blocks" % (call, nb))
        x = _bf16_bits(torch, rng, nb)
        gz = _bf16_bits(torch, rng, nb)
        part = torch.empty(lib.sfac_conv3x3x3_wgrad_partial_size(nb), dtype=torch.float32,
                           device="cuda")
        ref = torch.full((GSZ,), float("nan"), dtype=torch.float32, device="cuda")
        got = torch.full((GSZ,), float("nan"), dtype=torch.float32, device="cuda")
        assert lib.sfac_conv3x3x3_wgrad_bf16(_ptr(x), _ptr(gz), _ptr(ref), _ptr(part), nb,
                                             stream) == 0, lib.sfac_last_error()
        assert lib.sfac_conv3x3x3_wgrad_bf16_noalloc(_ptr(x), _ptr(gz), _ptr(got), nb,
                                                     stream) == 0, lib.sfac_last_error()
        torch.cuda.synchronize()
        assert torch.isfinite(ref).all()
        assert torch.equal(got.view(torch.int32), ref.view(torch.int32)), \
            "call %d: gradients differ from the workspace entry point" % call
        assert float(ref.abs().max()) > 0


def test_noalloc_wgrad_rejects_more_blocks_than_its_scratch(torch_gpu, gpu_lib):
    torch = torch_gpu
    lib = _bind(gpu_lib)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    x = torch.zeros(1, dtype=torch.int16, device="cuda")
    grad = torch.zeros(GSZ, dtype=torch.float32, device="cuda")
    # rejected on the host before any launch (the pointers are never read)
    st = lib.sfac_conv3x3x3_wgrad_bf16_noalloc(_ptr(x), _ptr(x), _ptr(grad), MAX_BLOCKS + 1,
                                               stream)
    assert st != 0
    assert b"maximum" in lib.sfac_last_error()
