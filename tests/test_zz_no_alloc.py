"""The allocation-free Gram entry point (sos_anderson_gram_static) against sos_anderson_gram through the C ABI: two
calls in a row on one device, the second smaller (fewer lanes, a shorter residual, another depth and slot), each
bitwise equal to the entry point with a caller work buffer and to an FP64 host product within rounding."""

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _case(n, lanes, depth, slot, filled, seed):
    rng = np.random.default_rng(seed)
    ring = rng.normal(size=(depth, lanes, n)).astype('f4')
    return torch.from_numpy(ring).cuda()


def test_gram_without_allocation_matches_gram_with_work_buffer(gpu):
    from sosfuse2_amd import hip

    stream = torch.cuda.current_stream().cuda_stream
    for n, lanes, depth, slot, filled, seed in ((3_000_000, 32, 5, 3, 0b11111, 1), (70_001, 16, 4, 1, 0b0111, 2)):
        print(f'no-alloc gram: n={n} lanes={lanes} depth={depth} slot={slot}', flush=True)
        ring = _case(n, lanes, depth, slot, filled, seed)
        blocks = min(128, (n + 65535) // 65536)  # the Gram kernel's chunks per lane
        work = torch.empty(blocks * lanes * depth, dtype=torch.float64, device='cuda')
        gram_a = torch.zeros((lanes, depth, depth), dtype=torch.float64, device='cuda')
        gram_b = torch.zeros_like(gram_a)
        hip.call('sos_anderson_gram', n, lanes, depth, slot, filled, hip.ptr(ring), hip.ptr(work), hip.ptr(gram_a),
                 stream)
        hip.call('sos_anderson_gram_static', n, lanes, depth, slot, filled, hip.ptr(ring), hip.ptr(gram_b), stream)
        torch.cuda.synchronize()
        a, b = gram_a.cpu().numpy(), gram_b.cpu().numpy()
        assert np.array_equal(a, b)
        r = ring.cpu().numpy().astype('f8')
        for j in range(depth):
            if (filled >> j) & 1:
                ref = np.einsum('wi,wi->w', r[slot], r[j])
                np.testing.assert_allclose(b[:, slot, j], ref, rtol=1e-10, atol=1e-9)
                np.testing.assert_array_equal(b[:, j, slot], b[:, slot, j])
    with pytest.raises(RuntimeError, match='bad argument'):
        hip.call('sos_anderson_gram_static', 100, 33, 4, 0, 1, hip.ptr(ring), hip.ptr(gram_b), stream)
