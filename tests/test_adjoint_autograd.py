"""The autograd wrapper of the rollout (athena_cuda_amd.adjoint.rollout) on
the host: a Rollout handle keeps the checkpoints of its latest forward pass
only, so the backward pass of a graph whose forward pass was followed by
another forward pass on the same handle must raise instead of reversing the
other trajectory. A stand-in handle with the same methods (forward,
generation, backward) on CPU tensors exercises the wrapper itself."""
import importlib

import pytest

from helpers import load_pkg

torch = pytest.importorskip("torch")


class LinearHandle:
    """forward: x -> 2 x; counts its forward passes like the device handle"""

    def __init__(self):
        self.gen = 0

    def forward(self, s):
        self.gen += 1
        return 2.0 * s.detach()

    def generation(self):
        return self.gen

    def backward(self, g):
        return 2.0 * g


def adjoint_module():
    pkg = load_pkg()
    return pkg, importlib.import_module(pkg.__name__ + ".adjoint")


def test_backward_after_the_latest_forward_gives_the_gradient():
    pkg, A = adjoint_module()
    h = LinearHandle()
    x = torch.ones(4, dtype=torch.float64, requires_grad=True)
    (3.0 * A.rollout(x, h)).sum().backward()
    assert torch.equal(x.grad, torch.full((4,), 6.0, dtype=torch.float64))


def test_backward_of_a_superseded_forward_raises():
    pkg, A = adjoint_module()
    h = LinearHandle()
    x1 = torch.ones(4, dtype=torch.float64, requires_grad=True)
    x2 = torch.full((4,), 5.0, dtype=torch.float64, requires_grad=True)
    y1 = A.rollout(x1, h)
    y2 = A.rollout(x2, h)
    with pytest.raises(pkg.AthmiError, match="another forward pass"):
        y1.sum().backward()
    # the latest graph is still valid
    y2.sum().backward()
    assert torch.equal(x2.grad, torch.full((4,), 2.0, dtype=torch.float64))
