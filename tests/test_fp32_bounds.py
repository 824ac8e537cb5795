"""The FP32 bounds of the batched kernel rest on two measured facts about the device's
FP32 math: log2f's error relative to max(|log2 x|, 1) and exp2f's relative error.  These
scans run every float of the domains the bounds use and check the errors against the
margins the kernel applies (LOG2F_ERR = 2^-20 for the draw's dU bound; ETA_POW = 2^-14
around exp2f(2/p log2f(q)), where an exponent of magnitude <= 128 turns a log2f error of
2^-21 into a relative error below 2^-19)."""
import struct

import pytest


def _scan(device, fn, lo, hi):
    import torch
    out = torch.zeros(1, dtype=torch.int64, device="cuda")
    device.check(device.lib().ecmc_fp32_math_error(fn, lo, hi, device.dptr(out), None))
    torch.cuda.synchronize()
    return struct.unpack("<d", struct.pack("<q", int(out.item())))[0]


@pytest.mark.gpu
def test_log2f_and_exp2f_errors_stay_inside_the_bound_margins(device):
    print("scanning log2f over every positive float, exp2f over [-126, 128)", flush=True)
    e_log = _scan(device, 0, 0x00000001, 0x7F800000)
    e_exp = max(_scan(device, 1, 0x80000000, 0xC2FC0001), _scan(device, 1, 0x00000000, 0x43000000))
    print("log2f max error / max(|log2 x|, 1) 2^%.2f, exp2f max relative error 2^%.2f"
          % (__import__("math").log2(e_log) if e_log else -999, __import__("math").log2(e_exp) if e_exp else -999),
          flush=True)
    assert e_log <= 2.0 ** -21
    assert e_exp <= 2.0 ** -21
