"""The scattered-read calibration entry point (kpt_measure_gather16) returns
src[idx[i]] for every i, record 0 for an index out of range, and refuses bad
arguments (no GPU needed for the refusals)."""
import ctypes as C

import numpy as np
import pytest

from kittlespt_amd import abi
from kittlespt_amd.context import DeviceBuffer


def test_gather_refuses_bad_arguments():
    L = abi.lib()
    one = C.c_void_p(16)
    assert L.kpt_measure_gather16(None, 4, one, 1, one) == abi.KPT_ERR_INVALID_ARG
    assert L.kpt_measure_gather16(one, 0, one, 1, one) == abi.KPT_ERR_INVALID_ARG
    assert L.kpt_measure_gather16(one, 4, one, -1, one) == abi.KPT_ERR_INVALID_ARG
    assert L.kpt_measure_gather16(one, 4, one, 0, one) == abi.KPT_OK


@pytest.mark.gpu
def test_gather_reads_the_indexed_records():
    rng = np.random.default_rng(3)
    src = rng.random((1000, 4), dtype=np.float32)
    idx = rng.integers(0, 1000, 5000).astype(np.int32)
    idx[[7, 100]] = [-3, 1000]          # out of range: record 0
    d_src, d_idx = DeviceBuffer.from_array(src), DeviceBuffer.from_array(idx)
    d_dst = DeviceBuffer(len(idx) * 16)
    print("[gather] 5000 reads of 1000 records", flush=True)
    assert abi.lib().kpt_measure_gather16(d_src.ptr, len(src), d_idx.ptr, len(idx), d_dst.ptr) == abi.KPT_OK
    out = d_dst.download(np.zeros((len(idx), 4), np.float32))
    ref = src[np.where((idx >= 0) & (idx < len(src)), idx, 0)]
    assert np.array_equal(out, ref)
