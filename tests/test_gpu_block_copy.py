"""tbk_block_copy: a device copy of a vector block, bit for bit, at sizes
whose 16-byte element count is not a multiple of the copy kernel's 2048 per
workgroup (and smaller than one workgroup), in both precisions; blocks of
different shapes are refused."""
import numpy as np
import pytest

from tight_binding_amd import _native as Nn
from tight_binding_amd import kpm

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rows,R,prec", [(1, 1, kpm.FP64), (37, 3, kpm.FP64), (1000, 16, kpm.FP64),
                                         (4099, 64, kpm.FP64), (777, 8, kpm.FP32), (5, 1, kpm.FP32)])
def test_block_copy_is_bit_exact(gpu_ctx, rows, R, prec):
    src = gpu_ctx.random_block(rows, R, seed=rows + R, precision=prec)
    dst = gpu_ctx.zeros_block(rows, R, precision=prec)
    ms = src.copy_to(dst)
    a, b = src.download(), dst.download()
    assert a.shape == (rows, R) and np.all(a != 0)
    assert np.array_equal(np.ascontiguousarray(a).view(np.uint64), np.ascontiguousarray(b).view(np.uint64))
    assert ms >= 0.0
    src.close()
    dst.close()


def test_block_copy_refuses_other_shapes(gpu_ctx):
    src = gpu_ctx.random_block(64, 4, seed=1)
    for rows, R, prec in ((65, 4, kpm.FP64), (64, 8, kpm.FP64), (64, 4, kpm.FP32)):
        dst = gpu_ctx.zeros_block(rows, R, precision=prec)
        with pytest.raises(Nn.NativeError) as e:
            src.copy_to(dst)
        assert e.value.code == Nn.TBK_E_INVALID
        dst.close()
    src.close()
