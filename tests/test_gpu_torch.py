"""torch.autograd.Function over the engine: gradients equal the oracle's
adjoint Jacobian contracted with dy, and central finite differences
(h = 1e-6, FP64)."""
import numpy as np
import pytest

import oracle as O
from helpers import random_hamiltonian, to_oracle

pytestmark = pytest.mark.gpu


def _circuit_ops(n, layers, rng):
    from pennylane_lightning_kokkos_amd.torch_adjoint import Op

    ops = []
    for layer in range(layers):
        for w in range(n):
            ops.append(Op("Rot", [w], list(rng.uniform(0, 2 * np.pi, size=3)), inverse=(w % 3 == 1)))
        for w in range(n):
            ops.append(Op("CNOT", [w, (w + 1 + layer) % n]))
        ops.append(Op("IsingZZ", [0, n - 1], [float(rng.uniform(0, 6))]))
    return ops


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_autograd_matches_oracle_and_fd(dtype):
    torch = pytest.importorskip("torch")
    from pennylane_lightning_kokkos_amd.torch_adjoint import Circuit, expand_ops, make_function

    rng = np.random.default_rng(42)
    n = 6
    ops = _circuit_ops(n, 2, rng)
    obs = [("named", "PauliZ", [0]), random_hamiltonian(n, 5, rng), ("named", "PauliX", [3])]
    circ = Circuit(n, ops, obs, dtype=dtype)
    values = np.concatenate([o.params for o in ops]).astype(np.float64)
    params = torch.tensor(values, dtype=torch.float64, requires_grad=True)
    f = make_function()
    out = f.apply(params, circ)
    dy = torch.tensor([0.3, -1.2, 0.7], dtype=torch.float64)
    (out * dy).sum().backward()
    grad = params.grad.numpy()

    # oracle: forward expvals and the full Jacobian on the serialized circuit
    ser, origin, k = expand_ops(circ._with_params(values))
    ref_exp = [O.expval(O.apply_ops(O.zero_state(n), O.Ops(*ser)), to_oracle(o)) for o in obs]
    jac = O.adjoint_jacobian(O.zero_state(n), [to_oracle(o) for o in obs], O.Ops(*ser),
                             list(range(sum(1 for p in ser[1] if p))), True)
    full = np.zeros((len(obs), k))
    for col, src in enumerate(origin):
        full[:, src] += jac[:, col]
    atol = 1e-10 if dtype == np.complex128 else 2e-5
    np.testing.assert_allclose(out.detach().numpy(), ref_exp, atol=atol)
    np.testing.assert_allclose(grad, dy.numpy() @ full, atol=atol)
    if dtype == np.complex128:
        h = 1e-6
        fd = np.zeros(k)
        for i in range(k):
            vp, vm = values.copy(), values.copy()
            vp[i] += h
            vm[i] -= h
            fd[i] = dy.numpy() @ (circ.expvals(vp) - circ.expvals(vm)) / (2 * h)
        np.testing.assert_allclose(grad, fd, atol=1e-7)
