"""Allocation-free dense assembly, tdbem_assemble_step_device_noalloc, called through the C ABI
with device buffers, against the allocating tdbem_assemble_step_device: the same row blocks, bit
for bit, on two calls in a row (the second with a smaller row range and another step), every
output element written, and no device memory taken by the call."""
import ctypes as C

import numpy as np
import pytest

import helpers
from tdbem import api

pytestmark = pytest.mark.gpu


def _entry(name):
    fn = getattr(api.lib(), name)
    fn.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int] + [C.c_void_p] * 4
    return fn


def _blocks(torch, n, fill):
    return [torch.full((n,), fill, dtype=torch.float64, device="cuda:0") for _ in range(4)]


def test_noalloc_matches_the_allocating_entry(gpu):
    import torch
    noalloc = _entry("tdbem_assemble_step_device_noalloc")
    alloc = _entry("tdbem_assemble_step_device")
    p = helpers.small_sphere(numstep=10)
    N = 3 * p.nnd
    with api.Context.from_problem(p) as ctx:
        for step, lo, hi in ((1, 0, p.nnd - 1), (3, 5, 17)):
            print(f"[gpu case] noalloc assembly step {step} rows {lo}..{hi}", flush=True)
            n = 3 * (hi - lo + 1) * N
            got = _blocks(torch, n, float("nan"))
            ref = _blocks(torch, n, 0.0)
            torch.cuda.synchronize()
            free0, _ = api.device_memory()
            rc = noalloc(ctx._h, step, lo, hi, *[C.c_void_p(b.data_ptr()) for b in got])
            assert rc == 0, api.lib().tdbem_last_error(ctx._h)
            ctx.synchronize()
            free1, _ = api.device_memory()
            assert free1 == free0
            rc = alloc(ctx._h, step, lo, hi, *[C.c_void_p(b.data_ptr()) for b in ref])
            assert rc == 0, api.lib().tdbem_last_error(ctx._h)
            ctx.synchronize()
            host = ctx.assemble_step(step, lo, hi)
            for g, r, h in zip(got, ref, host):
                g = g.cpu().numpy()
                assert not np.isnan(g).any()
                assert np.array_equal(g, r.cpu().numpy())
                assert np.array_equal(g.reshape(h.shape), h)
