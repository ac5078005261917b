"""tina_mlp_train_step_static (the training step that allocates nothing, its
gradient rows in the library's static device array) against
tina_mlp_train_step with a caller workspace: two steps in a row, the second on
a smaller batch, give bit-identical parameters, Adam moments and gradients; a
batch above the static maximum is rejected."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_static_training_step_equals_workspace_step(gpu):
    import torch
    from taichi_three_amd import _abi, neural
    lib = _abi.load()
    print('case: no-alloc training step vs workspace step, batches 16384 then 3001', flush=True)
    dev = torch.device('cuda')
    rng = np.random.default_rng(3)
    n = 40000
    X = torch.from_numpy(rng.uniform(-1, 1, (n, 12)).astype(np.float32)).to(dev)
    Y = torch.from_numpy(rng.uniform(0, 1, (n, 4)).astype(np.float32)).to(dev)
    p0 = torch.from_numpy(neural.init_params(5)).to(dev)
    state = {k: [p0.clone(), torch.zeros_like(p0), torch.zeros_like(p0)] for k in ('ws', 'static')}
    grads = {k: torch.empty_like(p0) for k in state}
    part = torch.empty(256 * neural.NPARAMS, dtype=torch.float32, device=dev)
    for step, (first, batch, lr) in enumerate([(0, 16384, 3e-3), (17000, 3001, 1e-3)], start=1):
        ps, m1, m2 = state['ws']
        _abi.check(lib.tina_mlp_train_step(_abi.ptr(X), _abi.ptr(Y), first, n, batch, _abi.ptr(ps), _abi.ptr(m1),
                                           _abi.ptr(m2), _abi.ptr(part), ctypes.c_float(lr), step,
                                           _abi.ptr(grads['ws']), _abi.stream_handle()))
        ps, m1, m2 = state['static']
        _abi.check(lib.tina_mlp_train_step_static(_abi.ptr(X), _abi.ptr(Y), first, n, batch, _abi.ptr(ps),
                                                  _abi.ptr(m1), _abi.ptr(m2), ctypes.c_float(lr), step,
                                                  _abi.ptr(grads['static']), _abi.stream_handle()))
        torch.cuda.synchronize()
        for a, b in zip(state['ws'], state['static']):
            assert torch.equal(a, b), step
        assert torch.equal(grads['ws'], grads['static']), step
        assert float(grads['ws'].abs().max()) > 0
    with pytest.raises(ValueError):
        ps, m1, m2 = state['static']
        _abi.check(lib.tina_mlp_train_step_static(_abi.ptr(X), _abi.ptr(Y), 0, n, 16385, _abi.ptr(ps),
                                                  _abi.ptr(m1), _abi.ptr(m2), ctypes.c_float(1e-3), 3, None,
                                                  _abi.stream_handle()))
