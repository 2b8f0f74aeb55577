"""GPU adjoint sweep (complex128) against the FP64 CPU adjoint and finite differences.

The adjoint has no reference counterpart; the oracle adjoint it is compared with is itself pinned
by finite differences (tests/test_oracle.py).  Tolerances: gradients to 1e-9 of max|grad| against
the oracle; against central differences of the GPU forward sweep, 1e-6 relative plus the
differences' own round-off floor (eps * L / h).
"""
import numpy as np
import pytest

from oracle import oracle as O
from recurgreen_amd import binding

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)


ADJ = {
    "dirichlet_absorb": dict(boundtype="dirichlet", nxdiso=30, nxfree=3, wol=3.3, dscat=1.0, ny=17, dos_model=1,
                             dabso=0.2),
    "periodic": dict(boundtype="periodic", nxdiso=25, nxfree=2, wol=3.3, dscat=0.8, ny=15),
    "dirichlet_filtered": dict(boundtype="dirichlet", nxdiso=20, wol=4.3, exfac=1.2, dscat=0.5, napera=0.7,
                               naperb=0.9),
    "N20_ny205": dict(boundtype="dirichlet", nxdiso=57, wol=10.25, aratio=2.85 / 10.25, dscat=0.6, ny=205,
                      dos_model=1),
}


@pytest.mark.parametrize("name", list(ADJ))
def test_adjoint_matches_cpu_adjoint(name):
    p = ADJ[name]
    say(f"adjoint vs oracle {name}")
    ow = O.Waveguide(**p)
    plan = binding.Plan(p, batch_max=3)
    U = np.stack([ow.potential(s).T for s in (4, 5, 6)], axis=2)
    val, grad, st, code = plan.adjoint(U, "g")
    assert code == 0 and not st.any()
    for k in range(3):
        vo, go = ow.adjoint_grad(U[:, :, k].T, "g")
        assert val[k] == pytest.approx(vo, rel=1e-10)
        np.testing.assert_allclose(grad[:, :, k].T, go, rtol=0, atol=1e-9 * np.max(np.abs(go)))
    plan.close()


def test_adjoint_finite_differences_n64():
    """N=64 (W=32.25), L=20 on 400 slices, ny=645, l_s=10: dL/du against finite differences of the
    GPU forward sweep (L = ||t||_F^2 from the t matrices), all perturbed potentials in one batch:
    64 random cells with central differences of step 1e-6 (their round-off, ~1e-14 L / h, is part
    of the bound), and 16 cells with the fourth-order stencil of step 1e-3 to 1e-6 relative."""
    p = dict(boundtype="dirichlet", nxdiso=400, wol=32.25, aratio=20.0 / 32.25, dscat=2.0, ny=645, dos_model=1)
    say("adjoint vs finite differences N64 nx400")
    plan = binding.Plan(p, batch_max=130)
    U = plan.generate_potential(17, 1)[:, :, 0]
    val, grad, st, code = plan.adjoint(U, "g")
    assert code == 0
    t0, _, _ = plan.transmission_potential(U)
    assert val[0] == pytest.approx(np.sum(np.abs(t0) ** 2), rel=1e-12)
    rng = np.random.default_rng(7)

    def losses(perturb):
        batch = np.repeat(U[:, :, None], len(perturb), axis=2)
        for k, (ix, iy, dh) in enumerate(perturb):
            batch[ix, iy, k] += dh
        t, _, code2 = plan.transmission_potential(batch)
        assert code2 == 0
        return np.sum(np.abs(t) ** 2, axis=(1, 2))

    cells = [(int(rng.integers(p["nxdiso"])), int(rng.integers(p["ny"]))) for _ in range(64)]
    h = 1e-6
    L = losses([(ix, iy, s * h) for ix, iy in cells for s in (1, -1)])
    floor = 1e-14 * val[0] / h
    worst = 0.0
    for c, (ix, iy) in enumerate(cells):
        fd = (L[2 * c] - L[2 * c + 1]) / (2 * h)
        err = abs(grad[ix, iy, 0] - fd)
        worst = max(worst, err / (abs(fd) + floor))
        assert err <= 1e-6 * abs(fd) + floor, (ix, iy, grad[ix, iy, 0], fd)
    h = 1e-3
    cells4 = cells[:16]
    L4 = losses([(ix, iy, s * h) for ix, iy in cells4 for s in (2, 1, -1, -2)])
    worst4 = 0.0
    for c, (ix, iy) in enumerate(cells4):
        f2, f1, m1, m2 = L4[4 * c:4 * c + 4]
        fd4 = (-f2 + 8 * f1 - 8 * m1 + m2) / (12 * h)
        rel = abs(grad[ix, iy, 0] - fd4) / abs(fd4)
        worst4 = max(worst4, rel)
        assert rel <= 1e-6, (ix, iy, grad[ix, iy, 0], fd4)
    print(f"[gpu-case] step 1e-6: max err/(|fd|+floor) = {worst:.2e} (floor {floor:.1e}); "
          f"4th-order step 1e-3: max rel err {worst4:.2e}", flush=True)
    plan.close()


def test_autograd_function_backward():
    import torch

    from recurgreen_amd.autograd import transmission_loss

    p = ADJ["N20_ny205"]
    say("autograd TransmissionLoss")
    plan = binding.Plan(p, batch_max=4)
    ow = O.Waveguide(**p)
    U = np.stack([ow.potential(s) for s in (1, 2)], axis=0)  # (nreal, ny, nxdiso)
    u = torch.tensor(U, dtype=torch.float64, device="cuda:0", requires_grad=True)
    L = transmission_loss(u, plan)
    w = torch.tensor([0.3, -1.7], dtype=torch.float64, device="cuda:0")
    (L * w).sum().backward()
    for k in range(2):
        vo, go = ow.adjoint_grad(U[k], "g")
        assert float(L[k]) == pytest.approx(vo, rel=1e-10)
        np.testing.assert_allclose(u.grad[k].cpu().numpy(), float(w[k]) * go, rtol=0,
                                   atol=1e-9 * abs(float(w[k])) * np.max(np.abs(go)))
    plan.close()


@pytest.mark.parametrize("name", ["dirichlet_absorb", "periodic", "dirichlet_filtered", "N20_ny205"])
def test_adjoint_tmax_matches_cpu_adjoint(name):
    """Gradient of the largest transmission eigenvalue (top eigenvector of the Gram matrix by
    inverse iteration on its tridiagonal form, back-transformed) against the CPU adjoint."""
    p = ADJ[name]
    say(f"adjoint T_max vs oracle {name}")
    ow = O.Waveguide(**p)
    plan = binding.Plan(p, batch_max=3)
    U = np.stack([ow.potential(s).T for s in (4, 5, 6)], axis=2)
    val, grad, st, code = plan.adjoint(U, "tmax")
    assert code == 0 and not st.any()
    for k in range(3):
        vo, go = ow.adjoint_grad(U[:, :, k].T, "tmax")
        assert val[k] == pytest.approx(vo, rel=1e-10)
        np.testing.assert_allclose(grad[:, :, k].T, go, rtol=0, atol=1e-8 * np.max(np.abs(go)))
    plan.close()


# ---------------------------------------------------------------------------------------------
# The complex64 adjoint (option adjoint_precision = mixed_fp32): FP32 state and products, FP64
# loss, terminal adjoint and gradient.  Against the FP64 CPU adjoint on identical potentials:
# the loss to 1e-4 relative; the gradient of g to 1e-3 of max|grad| (FP32 rounding over the
# sweep); the gradient of T_max to 1e-3 + 1e-4 T1 / (T1 - T2), because it goes through the top
# eigenvector of t^H t, whose perturbation by the FP32 error of t grows as 1 / (T1 - T2).  At the
# c4 shape (N=64, 400 slices, ny=645), against the complex128 GPU adjoint over the whole grid in
# the relative Frobenius norm, with the same bounds (2e-3 for g).
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("loss", ["g", "tmax"])
@pytest.mark.parametrize("name", ["dirichlet_absorb", "periodic", "N20_ny205"])
def test_fp32_adjoint_against_cpu_adjoint(name, loss):
    p = ADJ[name]
    say(f"fp32 adjoint vs oracle {name} {loss}")
    ow = O.Waveguide(**p)
    plan = binding.Plan(p, batch_max=2)
    plan.set_option("adjoint_precision", binding.PRECISIONS["mixed_fp32"])
    U = np.stack([ow.potential(s).T for s in (7, 8)], axis=2)
    val, grad, st, code = plan.adjoint(U, loss)
    assert code == 0 and not st.any()
    for k in range(2):
        vo, go = ow.adjoint_grad(U[:, :, k].T, loss)
        assert val[k] == pytest.approx(vo, rel=1e-4)
        err = np.max(np.abs(grad[:, :, k].T - go)) / np.max(np.abs(go))
        T = ow.tdistrib_potential(U[:, :, k].T)
        tol = 1e-3 + (1e-4 * T[0] / (T[0] - T[1]) if loss == "tmax" else 0.0)
        print(f"[gpu-case] fp32 adjoint {name} {loss}: loss rel err {abs(val[k] - vo) / abs(vo):.2e}, grad err {err:.2e}, "
              f"T1 {T[0]:.4f} gap {T[0] - T[1]:.2e}, bound {tol:.2e}", flush=True)
        assert err <= tol
    plan.close()


@pytest.mark.parametrize("loss", ["g", "tmax"])
def test_fp32_adjoint_c4_shape_against_fp64(loss):
    import torch

    from recurgreen_amd.autograd import transmission_loss

    p = dict(boundtype="dirichlet", nxdiso=400, wol=32.25, aratio=20.0 / 32.25, dscat=2.0, ny=645, dos_model=1)
    say(f"fp32 vs fp64 adjoint N64 nx400 {loss}")
    plan = binding.Plan(p, batch_max=8)
    plan.adjoint_reserve(8)
    U = plan.generate_potential(31, 8)  # (nxdiso, ny, r)
    u = torch.from_numpy(np.ascontiguousarray(U.transpose(2, 1, 0))).cuda().requires_grad_(True)
    grads, vals = {}, {}
    for prec in ("fp64", "mixed_fp32"):
        if u.grad is not None:
            u.grad = None
        L = transmission_loss(u, plan, loss, prec)
        L.sum().backward()
        grads[prec] = u.grad.detach().cpu().numpy().copy()
        vals[prec] = L.detach().cpu().numpy()
    rel = np.linalg.norm(grads["mixed_fp32"] - grads["fp64"]) / np.linalg.norm(grads["fp64"])
    lrel = np.max(np.abs(vals["mixed_fp32"] - vals["fp64"]) / np.abs(vals["fp64"]))
    T = plan.tdistrib_potential(U, "fp64")[0].reshape(8, -1)
    tol = 2e-3 + (1e-4 * np.max(T[:, 0] / (T[:, 0] - T[:, 1])) if loss == "tmax" else 0.0)
    print(f"[gpu-case] c4 {loss}: gradient rel err {rel:.2e}, loss rel err {lrel:.2e}, bound {tol:.2e}", flush=True)
    assert lrel <= 1e-4 and rel <= tol
    plan.close()
