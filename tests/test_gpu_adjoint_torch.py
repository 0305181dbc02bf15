"""torch.autograd.Function harness over the C ABI's objective and adjoint:
backward() returns the discrete-adjoint gradient, checked against central
finite differences of forward() for the scalar parameters and a random
direction of the inlet map.  The tensors stay on the CPU (float64); all GPU
work happens inside the library."""
import sys

import numpy as np
import pytest

import _pkg

pytestmark = pytest.mark.gpu


def make_function():
    L = _pkg.load()
    L.lib()  # the library's HIP runtime is loaded before torch is imported
    import torch

    class SoluteFlux(torch.autograd.Function):
        """J(params, cmap): params = (g_x, g_y, g_z, tau, tau_s) of the flow, cmap = inlet map;
        kind: "flux" (the solute flux through plane LZ - 2) or "permeability"
        """

        @staticmethod
        def forward(ctx, params, cmap, mask, nsteps, window, every, kind="flux"):
            p = params.detach().cpu().numpy()
            nk, nj, ni = mask.shape
            lat = L.Lattice(mask=mask, dims=3, ni=ni, nj=nj, nk=nk, nsubs=2, tau=(p[3], p[4]),
                            gaccel=(p[0], p[1], p[2], 0, 0, 0), zface_pressure=1, rho_bot=1.001, rho_top=1.0,
                            zface_solute=1, c_bot_map=cmap.detach().cpu().numpy(), flux_plane=nk - 2)
            lat.init_rest()
            J = lat.objective(nsteps, window, every, kind=kind)
            ctx.lat = lat
            return torch.tensor(J, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gJ):
            g, gm = ctx.lat.adjoint()
            ctx.lat.close()
            return gJ * torch.from_numpy(g), gJ * torch.from_numpy(gm), None, None, None, None, None

    return torch, SoluteFlux


def test_autograd_gradient_matches_finite_differences():
    print("case torch harness 10x8x16, 60 steps", flush=True)
    torch, SoluteFlux = make_function()
    L = _pkg.load()
    mask, _ = L.spheres3d(10, 8, 16, 11, 2.0, 3.5, 0.6, zbuf=3)
    cmap0 = np.random.default_rng(4).uniform(0.8, 1.2, (8, 10))
    params = torch.tensor([2e-6, -1e-6, 1e-5, 0.8, 0.75], dtype=torch.float64, requires_grad=True)
    cmap = torch.tensor(cmap0, dtype=torch.float64, requires_grad=True)
    J = SoluteFlux.apply(params, cmap, mask, 60, 20, 25)
    J.backward()
    gp, gc = params.grad.numpy().copy(), cmap.grad.numpy().copy()
    with torch.no_grad():
        for q, h in ((2, 1e-8), (3, 1e-6), (4, 1e-6)):
            d = torch.zeros(5, dtype=torch.float64)
            d[q] = h
            fd = (SoluteFlux.apply(params + d, cmap, mask, 60, 20, 25)
                  - SoluteFlux.apply(params - d, cmap, mask, 60, 20, 25)).item() / (2 * h)
            print(f"param {q}: autograd {gp[q]:.10e} fd {fd:.10e}", flush=True)
            assert abs(gp[q] - fd) <= 1e-5 * abs(fd)
        dc = torch.from_numpy(np.random.default_rng(2).uniform(-1, 1, (8, 10)))
        h = 1e-6
        fd = (SoluteFlux.apply(params, cmap + h * dc, mask, 60, 20, 25)
              - SoluteFlux.apply(params, cmap - h * dc, mask, 60, 20, 25)).item() / (2 * h)
        adj = float((torch.from_numpy(gc) * dc).sum())
        print(f"map direction: autograd {adj:.10e} fd {fd:.10e}", flush=True)
        assert abs(adj - fd) <= 1e-5 * abs(fd)
    sys.stdout.flush()


def test_autograd_flux_plus_permeability():
    """a loss combining both objectives of one parameter set: backward()
    sums the two adjoint gradients, checked against central differences"""
    print("case torch harness, flux + 1e-3 permeability", flush=True)
    torch, SoluteFlux = make_function()
    L = _pkg.load()
    mask, _ = L.spheres3d(10, 8, 16, 11, 2.0, 3.5, 0.6, zbuf=3)
    cmap = torch.tensor(np.random.default_rng(4).uniform(0.8, 1.2, (8, 10)), dtype=torch.float64)
    params = torch.tensor([2e-6, -1e-6, 1e-5, 0.8, 0.75], dtype=torch.float64, requires_grad=True)

    def loss(p):
        return SoluteFlux.apply(p, cmap, mask, 60, 20, 25) + 1e-3 * SoluteFlux.apply(p, cmap, mask, 60, 20, 25,
                                                                                     "permeability")
    loss(params).backward()
    gp = params.grad.numpy().copy()
    with torch.no_grad():
        for q, h in ((2, 1e-8), (3, 1e-6), (4, 1e-6)):
            d = torch.zeros(5, dtype=torch.float64)
            d[q] = h
            fd = (loss(params + d) - loss(params - d)).item() / (2 * h)
            print(f"param {q}: autograd {gp[q]:.10e} fd {fd:.10e}", flush=True)
            assert abs(gp[q] - fd) <= 1e-5 * abs(fd)
    sys.stdout.flush()
