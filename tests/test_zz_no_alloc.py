"""The no-allocation forward (vd_forward_device_noalloc) against the
allocating one (vd_forward_device), through the C-ABI (needs an MI355X).

With the shared-suffix continuation off, vd_forward_device computes every
pair by full sweeps (prefix sharing is bit-identical to them) or in FP64, as
the no-allocation forward does, so the two agree bit for bit, NaN and -inf
positions included.  Two calls in a row, the second on a smaller batch, in
every mode and precision; then a batch beyond the static scratch is refused
before anything is enqueued."""
import ctypes

import numpy as np
import pytest

from vardepth_amd import _lib, engine, workloads

pytestmark = pytest.mark.gpu


def noalloc(dev, mode, precision, out):
    import torch
    p = _lib.params(dev.batch.gap, dev.batch.ext)
    code = _lib.lib().vd_forward_device_noalloc(ctypes.byref(dev.vb), ctypes.byref(p), engine._mode(mode),
                                                engine._precision(precision), dev.npairs, dev.max_hap,
                                                ctypes.c_void_p(out.data_ptr()),
                                                ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    return code


def test_noalloc_equals_forward_device_twice():
    import torch
    L = _lib.lib()
    prev = L.vd_suffix_sharing(0)
    try:
        big = workloads.uneven_regions(seed=21, regions=64)
        small = workloads.uneven_regions(seed=22, regions=6)
        edge = workloads.from_regions([(["ACGT", "", "A" * 300], [b"\x1e" * 4, b"", b"\x05" * 300],
                                        ["ACGTA", "", "C" * 700])])
        assert small.npairs < big.npairs
        for b in (big, small, edge):
            dev = engine.DeviceBatch(b)
            out = torch.empty(max(dev.npairs, 1), dtype=torch.float64, device="cuda")
            for mode in (0, 1, 2):
                for precision in ("mixed", "float64"):
                    dev.forward(mode=mode, precision=precision)
                    torch.cuda.synchronize()
                    ref = dev.result()
                    out.fill_(float("nan"))
                    code = noalloc(dev, mode, precision, out)
                    assert code == 0, _lib.last_error()
                    torch.cuda.synchronize()
                    got = out[:dev.npairs].cpu().numpy()
                    np.testing.assert_array_equal(got, ref)
                    print("noalloc %d pairs mode %d %s: equal" % (dev.npairs, mode, precision), flush=True)
    finally:
        L.vd_suffix_sharing(prev)


def test_noalloc_refuses_batches_beyond_its_scratch():
    import torch
    # 400,000 one-base regions: 204 bytes of plan scratch per region alone exceed 64 MiB
    n = 400000
    b = workloads.RegionBatch(read_bytes=np.full(n, ord("A"), np.uint8), qual_bytes=np.full(n, 30, np.uint8),
                              read_off=np.arange(n + 1), hap_bytes=np.full(n, ord("A"), np.uint8),
                              hap_off=np.arange(n + 1), region_reads=np.arange(n + 1), region_haps=np.arange(n + 1))
    dev = engine.DeviceBatch(b)
    out = torch.full((dev.npairs,), float("nan"), dtype=torch.float64, device="cuda")
    assert noalloc(dev, 0, "mixed", out) == 2
    assert "bytes of scratch" in _lib.last_error()
    torch.cuda.synchronize()
    assert torch.isnan(out).all()   # nothing was enqueued
