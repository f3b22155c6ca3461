"""The FP64 torch restatement of HEOM propagation (oracle/heom_torch.py), which the GPU adjoint's
gradients are checked against: its forward state equals the reference's spin-boson trajectory, and its
autograd gradients equal central differences of itself."""
import numpy as np
import torch

from conftest import bath_of, golden
from oracle import heom_torch as T


def _c1():
    g = golden("traj_c1.npz")
    b = bath_of(g)
    return g, b[:4], b[4]


def test_forward_matches_reference_trajectory():
    g, bath, mode = _c1()
    rT = T.final_state(torch.zeros(0, dtype=torch.float64), torch.zeros(25, dtype=torch.float64), g["H"],
                       np.zeros((0, 2, 2)), np.zeros((2, 2)), g["Q"], bath, mode, 5, g["rho0"], 0.01, 8)
    assert np.abs(rT.numpy() - g["rho"][200]).max() <= 1e-14


def test_autograd_matches_central_differences():
    g, bath, mode = _c1()
    sz = np.diag([1.0, -1.0]).astype(complex)
    sx = np.array([[0, 1], [1, 0]], complex)
    P = np.array([sz / 2, sx / 2])
    args = (np.zeros((2, 2)), P, sx, g["Q"], bath, mode, 3, g["rho0"], 0.02, 5)

    def J(theta, amps):
        rT = T.final_state(theta, amps, *args[:1], *args[1:])
        return rT[0, 0].real

    theta = torch.tensor([1.0, 0.8], dtype=torch.float64, requires_grad=True)
    amps = torch.tensor([0.1, -0.2, 0.05, 0.0], dtype=torch.float64, requires_grad=True)
    J(theta, amps).backward()
    h = 1e-6
    for x, grad in ((theta, theta.grad), (amps, amps.grad)):
        for i in range(x.numel()):
            with torch.no_grad():
                xp, xm = x.detach().clone(), x.detach().clone()
                xp[i] += h
                xm[i] -= h
                if x is theta:
                    fd = (J(xp, amps.detach()) - J(xm, amps.detach())) / (2 * h)
                else:
                    fd = (J(theta.detach(), xp) - J(theta.detach(), xm)) / (2 * h)
            assert abs(float(fd) - float(grad[i])) <= 1e-7 * max(1.0, abs(float(grad[i])))
