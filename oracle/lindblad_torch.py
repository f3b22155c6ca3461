"""FP64 autograd reference for Lindblad gradients -- TEST INFRASTRUCTURE ONLY.

The same propagation as pyqed_amd.adjoint.LindbladAdjoint (generator A_0(theta) + sum_f E_f B_f
per step, RK4 of pyqed/phys.py:1051-1063) written with torch complex128 operations on the CPU, so
torch.autograd supplies the gradients the adjoint kernel is checked against."""
import numpy as np
import torch


def final_state(theta, amps, base, G, B, rho0, dt, steps_per_seg):
    """rho(T) as a torch complex128 tensor ([d, d] when the generator is d^2 x d^2, else the state
    vector); theta [P] and amps [F, nseg] are float64 tensors."""
    base = torch.as_tensor(np.asarray(base, np.complex128))
    G = torch.as_tensor(np.asarray(G, np.complex128).reshape(-1, *base.shape))
    B = torch.as_tensor(np.asarray(B, np.complex128).reshape(-1, *base.shape))
    A0 = base + torch.einsum("j,jab->ab", theta.to(torch.complex128), G) if G.shape[0] else base
    v = torch.as_tensor(np.asarray(rho0, np.complex128).reshape(-1))
    nseg = amps.shape[1]
    for s in range(nseg):
        A = A0 + torch.einsum("f,fab->ab", amps[:, s].to(torch.complex128), B)
        for _ in range(steps_per_seg):
            k1 = A @ v
            k2 = A @ (v + k1 * (dt / 2))
            k3 = A @ (v + k2 * (dt / 2))
            k4 = A @ (v + k3 * dt)
            v = v + (k1 + 2 * k2 + 2 * k3 + k4) / 6.0 * dt
    d = int(round(np.sqrt(v.shape[0])))
    return v.reshape(d, d) if d * d == v.shape[0] else v
