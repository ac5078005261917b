"""The shared-divisor division the kernels use in place of a / d (the compiler's
correctly rounded binary32 division) agrees with it bit for bit: every
numerator mantissa in [1, 2) against 1,024 divisors (8.6 G pairs, quotients at
every distance from the rounding midpoints), 2^31 random in-range pairs and
2^31 random bit patterns over the whole range (zeros, subnormals, infinities
and NaN take the compiler's division)."""
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('mode,npairs,fast_min', [(0, 1024 << 23, 1.0), (2, 1 << 31, 0.5), (1, 1 << 31, 0.125)])
def test_shared_divisor_division_is_correctly_rounded(gpu, mode, npairs, fast_min):
    import torch
    from taichi_three_amd import _abi
    print(f'case: division self-test mode {mode}, {npairs} pairs', flush=True)
    counts = torch.zeros(2, dtype=torch.int64, device='cuda')
    _abi.check(_abi.load().tina_selftest_div(mode, npairs, 12345, _abi.ptr(counts), _abi.stream_handle()))
    fast, bad = (int(x) for x in counts.cpu())
    print(f'fast path {fast} of {npairs}, mismatches {bad}', flush=True)
    assert bad == 0
    assert fast >= fast_min * npairs
