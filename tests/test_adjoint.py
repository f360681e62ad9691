"""The differentiable run: the hand-written adjoint of push + deposit + FDTD
(vpic_hip_adjoint_bfield_energy, vpic_amd/adjoint.py) against finite
differences.

The observable is W = the magnetic field energy after N steps of one pushed
species on a periodic box.  Checks:
  * the adjoint's forward is the hot path: its FP32 energy equals the main
    path's (vpic_amd.sim.Simulation, the tiled/flat push kernels) to
    rounding, and its FP64 energy the FP64 oracle's (oracle/trace.py, the
    oracle's step source compiled over double, on the CPU);
  * the FP64 adjoint's directional derivatives, for the initial momenta,
    positions and fields, equal FP64 central differences of the FP64
    forward at random directions to 1e-6 relative -- both on the device
    and with the differences taken by the CPU oracle;
  * the Weibel run (bi-Maxwellian electrons, T_perp/T_par = 4 about x,
    uth_par = 0.05c, immobile ions, 200 steps, a checkpoint every 20): the
    FP32 product adjoint, called through a torch.autograd.Function, gives
    d W / d(initial momenta) along 8 random directions and d W / d
    (anisotropy) within 1e-3 relative of FP64 central differences;
  * the full-size Weibel box (64^3 cells, 32 ppc, 8,388,608 electrons) to
    the same bound, with the adjoint / forward time ratio reported.
The finite differences are of the FP64 run from the FP32 initial state
widened exactly, so they measure the gradient of the same model.
"""
import ctypes
import sys
import time

import numpy as np
import pytest

from vpic_amd import types as T

import helpers as H

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def lib():
    import torch
    from vpic_amd import _lib
    assert torch.cuda.is_available()
    _lib.call("vpic_hip_init", 0)
    return _lib


def weibel_box(n, ppc, dx=0.1, seed=3, aniso=4.0, uth_par=0.05):
    """A periodic n^3 box of cells dx (in d_e) with bi-Maxwellian electrons
    (x parallel), weights n0 dV / ppc; returns (grid, FP32 particles, unit
    normals used for the momenta)."""
    L = n * dx
    dt = 0.95 * T.courant_length(L, L, L, n, n, n)
    g = T.grid_periodic(0, 0, 0, L, L, L, n, n, n, dt)
    p = H.thermal_particles(g, ppc, 1.0, seed, w=dx ** 3 / ppc)
    nrm = np.stack([p["ux"], p["uy"], p["uz"]], 1).astype(np.float64)
    s = np.sqrt(aniso)
    p["ux"] = (uth_par * nrm[:, 0]).astype(np.float32)
    p["uy"] = (s * uth_par * nrm[:, 1]).astype(np.float32)
    p["uz"] = (s * uth_par * nrm[:, 2]).astype(np.float32)
    return g, p, nrm


def dev(a, dtype):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda").view(dtype)


def p32_tensor(p):
    import torch
    return dev(p.view(np.float32).reshape(-1, 8), torch.float32)


def p64_tensor(p):
    import torch
    from vpic_amd import adjoint as A
    return dev(A.host_particles64(p), torch.float64)


def f_tensor(f, fp64):
    import torch
    from vpic_amd import adjoint as A
    if fp64:
        return dev(A.fields64(f), torch.float64)
    return dev(f.view(np.float32).reshape(-1, 20), torch.float32)


def energy(part, fields, g, n_step, q=-1.0, m=1.0):
    from vpic_amd import adjoint as A
    return A.run(part, fields, g, q, m, n_step, grads=False)[0]


def test_adjoint_forward_is_the_hot_path(lib):
    """FP32: the adjoint run's energy against the main path's step loop;
    FP64: against the FP64 oracle's."""
    import torch
    from oracle.trace import OracleRun
    from vpic_amd import sim as S
    g, p, _ = weibel_box(8, 16, seed=5)
    f = H.random_fields(g, 6, 0.02, 0.02)
    f["ex"][:] = 0  # E starts at zero; B random (not divergence free: the
    f["ey"][:] = 0  # step does not care)
    f["ez"][:] = 0
    steps = 30
    w32 = energy(p32_tensor(p), f_tensor(f, False), g, steps)
    s = S.Simulation(g)
    s.fields.upload(f)
    S.host_species(s, "electron", -1.0, 1.0, p, sort_interval=10)
    s.load_interpolator_array()
    for _ in range(steps):
        s.advance()
    torch.cuda.synchronize()
    wm = s.energy_f()[3:6].sum()
    say("W fp32 adjoint run %.9e, main path %.9e" % (w32, wm))
    assert abs(w32 - wm) <= 1e-5 * abs(wm)
    w64 = energy(p64_tensor(p), f_tensor(f, True), g, steps)
    r = OracleRun(g, [("electron", -1.0, 1.0, p)], fp64=True, fields=f)
    r.load_interpolator()
    for _ in range(steps):
        r.advance()
    wo = r.energies()[0][3:6].sum()
    r.close()
    say("W fp64 adjoint run %.15e, FP64 oracle %.15e" % (w64, wo))
    assert abs(w64 - wo) <= 1e-12 * abs(wo)


def _fd_dirs(rng, k, shape, scale):
    return [rng.standard_normal(shape) * scale for _ in range(k)]


def test_fp64_adjoint_matches_finite_differences(lib):
    """FP64 adjoint against FP64 central differences (device forward) for
    momenta, positions and fields, 6 random directions each."""
    import torch
    from vpic_amd import adjoint as A
    g, p, _ = weibel_box(6, 8, seed=7)
    f = H.random_fields(g, 8, 0.01, 0.01)
    steps, ckpt = 24, 5
    P = p64_tensor(p)
    F = f_tensor(f, True)
    w, ub, db, fb = A.run(P, F, g, -1.0, 1.0, steps, ckpt,
                          want_positions=True, want_fields=True)
    rng = np.random.default_rng(1)
    n = len(p)
    worst = 0.0
    for kind in ("u", "d", "f"):
        for _ in range(6):
            if kind == "u":
                v = torch.from_numpy(rng.standard_normal((n, 3)) * 0.05).cuda()
                eps = 1e-6
                grad = float((ub * v).sum())

                def at(h):
                    Q = P.clone()
                    Q[:, 4:7] += h * v
                    return energy(Q, F, g, steps)
            elif kind == "d":
                v = torch.from_numpy(rng.standard_normal((n, 3)) * 0.1).cuda()
                eps = 1e-6
                grad = float((db * v).sum())

                def at(h):
                    Q = P.clone()
                    Q[:, 0:3] += h * v
                    return energy(Q, F, g, steps)
            else:
                vv = np.zeros((g.nv, 18))
                vv[:, [0, 1, 2, 4, 5, 6]] = rng.standard_normal((g.nv, 6))
                v = torch.from_numpy(vv * 0.01).cuda()
                eps = 1e-6
                grad = float((fb * v[:, [0, 1, 2, 4, 5, 6]]).sum())

                def at(h):
                    return energy(P, F + h * v, g, steps)
            fd = (at(eps) - at(-eps)) / (2 * eps)
            rel = abs(grad - fd) / abs(fd)
            worst = max(worst, rel)
            assert rel <= 1e-6, (kind, grad, fd, rel)
    say("FP64 adjoint vs FP64 differences: worst rel %.2e" % worst)


def test_fp64_adjoint_matches_oracle_differences(lib):
    """The same derivative against central differences taken by the FP64
    CPU oracle (the oracle's step source compiled over double), 4 random
    momentum directions."""
    from oracle.trace import OracleRun, widen_particles
    from vpic_amd import adjoint as A
    g, p, _ = weibel_box(6, 8, seed=9)
    steps, ckpt = 20, 4
    w, ub, _, _ = A.run(p64_tensor(p), f_tensor(np.zeros(g.nv, T.FIELD),
                                                 True), g, -1.0, 1.0, steps,
                        ckpt)
    ub = ub.cpu().numpy()
    rng = np.random.default_rng(2)

    def oracle_w(dv):
        r = OracleRun(g, [("electron", -1.0, 1.0, p)], fp64=True, threads=4)
        q = r.species[0][3]
        q["ux"] += dv[:, 0]
        q["uy"] += dv[:, 1]
        q["uz"] += dv[:, 2]
        r.load_interpolator()
        for _ in range(steps):
            r.advance()
        out = r.energies()[0][3:6].sum()
        r.close()
        return out

    for _ in range(4):
        v = rng.standard_normal((len(p), 3)) * 0.05
        eps = 1e-6
        fd = (oracle_w(eps * v) - oracle_w(-eps * v)) / (2 * eps)
        grad = float((ub * v).sum())
        rel = abs(grad - fd) / abs(fd)
        say("oracle differences: adjoint %.10e fd %.10e rel %.2e"
            % (grad, fd, rel))
        assert rel <= 1e-5


def _weibel_gradients(n, ppc, steps=200, ckpt=20, ndir=8, seed=3):
    """FP32 adjoint through the autograd Function against FP64 central
    differences; returns the relative errors and the timings."""
    import torch
    from vpic_amd import adjoint as A
    aniso, uth_par = 4.0, 0.05
    g, p, nrm = weibel_box(n, ppc, seed=seed, aniso=aniso, uth_par=uth_par)
    zero_f = np.zeros(g.nv, T.FIELD)
    P32 = p32_tensor(p)
    F32 = f_tensor(zero_f, False)
    normals = torch.from_numpy(nrm.astype(np.float32)).cuda()
    A_t = torch.tensor(aniso, dtype=torch.float32, device="cuda",
                       requires_grad=True)
    u = A.bimaxwellian_momenta(normals, uth_par, A_t)
    u.retain_grad()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    W = A.BFieldEnergy.apply(u, P32, F32, g, -1.0, 1.0, steps, ckpt)
    W.backward()
    torch.cuda.synchronize()
    t_adj = time.perf_counter() - t0
    t0 = time.perf_counter()
    w_fwd = energy(P32, F32, g, steps)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    ub = u.grad.double()
    dA = float(A_t.grad)
    # FP64 central differences of the FP64 run from the widened state
    P64 = p64_tensor(p)
    F64 = f_tensor(zero_f, True)
    base_u = P64[:, 4:7].clone()
    nrm_t = torch.from_numpy(nrm).cuda()

    def w_at_u(du):
        Q = P64.clone()
        Q[:, 4:7] = base_u + du
        return energy(Q, F64, g, steps)

    def w_at_aniso(a):
        Q = P64.clone()
        s = np.sqrt(a / aniso)
        Q[:, 5:7] = base_u[:, 1:3] * s
        return energy(Q, F64, g, steps)

    rng = np.random.default_rng(11)
    errs = []
    for _ in range(ndir):
        v = torch.from_numpy(rng.standard_normal((len(p), 3))).cuda() * \
            (uth_par * 1e-4)
        fd = (w_at_u(v) - w_at_u(-v)) / 2
        grad = float((ub * v).sum())
        errs.append(abs(grad - fd) / abs(fd))
        say("direction: adjoint %.8e fd %.8e rel %.2e" % (grad, fd, errs[-1]))
    h = aniso * 1e-5
    fdA = (w_at_aniso(aniso + h) - w_at_aniso(aniso - h)) / (2 * h)
    errA = abs(dA - fdA) / abs(fdA)
    say("anisotropy: adjoint %.8e fd %.8e rel %.2e" % (dA, fdA, errA))
    say("W %.8e; forward %.3f s, forward+adjoint %.3f s, ratio %.2f"
        % (w_fwd, t_fwd, t_adj, t_adj / t_fwd))
    return errs, errA, t_adj / t_fwd, float(W)


def test_weibel_gradient_reduced(lib):
    """The Weibel run on a 16^3 box, 32 ppc, 200 steps, checkpoints every
    20: FP32 adjoint within 1e-3 of FP64 differences."""
    errs, errA, ratio, _ = _weibel_gradients(16, 32)
    assert max(errs) <= 1e-3 and errA <= 1e-3


@pytest.mark.timeout(900)
def test_weibel_gradient_full_size(lib):
    """The Weibel run at full size: 64^3 cells, 32 ppc (8,388,608
    electrons), 200 steps, a checkpoint every 20; 8 momentum directions and
    the anisotropy within 1e-3 of FP64 differences."""
    errs, errA, ratio, _ = _weibel_gradients(64, 32)
    assert max(errs) <= 1e-3 and errA <= 1e-3
