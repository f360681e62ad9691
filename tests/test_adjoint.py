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
  * d W / d(anisotropy) through the torch.autograd.Function (the
    bi-Maxwellian loader's chain into the momentum gradient) against FP64
    central differences, on a box where they are well posed;
  * the Weibel physics (bi-Maxwellian electrons, T_perp/T_par = 4 about x,
    uth_par = 0.05c, immobile ions), on a 16^3 box and at full size (64^3
    cells, 32 ppc, 8,388,608 electrons): the FP64 adjoint against FP64
    central differences along 8 sparse random directions (64 particles
    each) over 50 and 20 steps within 1e-3, and the FP32 product adjoint
    against the FP64 adjoint over 200 steps (checkpoint every 20) within
    5e-2, with the FP32 adjoint / forward time at full size at most 4.
Why not differences over all 200 steps: the push interpolates E_x constant
in x within a cell and B_x constant in y and z (interpolator_array_pipeline
.cc:67-121), so the force on a particle jumps where it crosses a face.  A
perturbation moves some crossings from one step to another -- the
perturbed particles' and, through the fields, every other particle's -- and
those jumps make central differences differ from the derivative at any step
size; the number of moved crossings grows with the run (the plasma is
unstable).  On top, the deposits' float atomics make the run's summation
order vary, and the instability amplifies that rounding noise.  Measured on
the 16^3 box: dense directions O(1) relative from 50 steps on for h from 1e-3
to 1e-7; sparse directions ~1e-6 over 50 steps, but 2e-4 to O(1) over 200.
So the differences are taken where their quotient has converged (h and 10 h
agree), and the 200-step run compares the FP32 adjoint (the product) with
the FP64 one (the same kernels over double): the two trajectories part by
rounding, 3e-3 in relative norm after 50 steps, 1.6e-2 after 200.
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


def test_fp64_anisotropy_gradient_matches_differences(lib):
    """d W / d(anisotropy) through the autograd Function (the chain from
    the bi-Maxwellian loader into the adjoint's momentum gradient) against
    FP64 central differences of the anisotropy, on a box small and short
    enough that no particle changes cell sequence between the difference
    points (a dense direction: every particle's momentum moves)."""
    import torch
    from vpic_amd import adjoint as A
    aniso, uth_par = 4.0, 0.05
    g, p, nrm = weibel_box(6, 8, seed=7, aniso=aniso, uth_par=uth_par)
    steps, ckpt = 24, 5
    P = p64_tensor(p)
    F = f_tensor(H.random_fields(g, 8, 0.01, 0.01), True)
    normals = torch.from_numpy(nrm).cuda()
    A_t = torch.tensor(aniso, dtype=torch.float64, device="cuda",
                       requires_grad=True)
    u = A.bimaxwellian_momenta(normals, uth_par, A_t)
    W = A.BFieldEnergy.apply(u, P, F, g, -1.0, 1.0, steps, ckpt)
    W.backward()
    grad = float(A_t.grad)

    def w_at(a):
        uu = A.bimaxwellian_momenta(normals, uth_par,
                                    torch.tensor(a, dtype=torch.float64,
                                                 device="cuda"))
        return energy(A.set_momenta(P, uu), F, g, steps)
    h = aniso * 1e-6
    fd = (w_at(aniso + h) - w_at(aniso - h)) / (2 * h)
    rel = abs(grad - fd) / abs(fd)
    say("anisotropy: adjoint %.12e fd %.12e rel %.2e" % (grad, fd, rel))
    assert rel <= 1e-5


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


def _sparse_dirs(rng, n, npart=64, scale=0.05):
    """A random direction over the momenta of `npart` random particles (the
    other particles unperturbed)."""
    import torch
    v = np.zeros((n, 3))
    idx = rng.choice(n, npart, replace=False)
    v[idx] = rng.standard_normal((npart, 3)) * scale
    return torch.from_numpy(v).cuda()


def _gradients(g, p, nrm, steps, ckpt, aniso, uth_par):
    """FP32 and FP64 adjoint gradients of W through the autograd Function
    (d W / d u and d W / d anisotropy) and their timings."""
    import torch
    from vpic_amd import adjoint as A
    zero_f = np.zeros(g.nv, T.FIELD)
    out = {}
    for name, P, F, dt in (("fp32", p32_tensor(p), f_tensor(zero_f, False),
                            torch.float32),
                           ("fp64", p64_tensor(p), f_tensor(zero_f, True),
                            torch.float64)):
        normals = torch.from_numpy(nrm).to("cuda", dt)
        A_t = torch.tensor(aniso, dtype=dt, device="cuda", requires_grad=True)
        u = A.bimaxwellian_momenta(normals, uth_par, A_t)
        u.retain_grad()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        W = A.BFieldEnergy.apply(u, P, F, g, -1.0, 1.0, steps, ckpt)
        W.backward()
        torch.cuda.synchronize()
        t_adj = time.perf_counter() - t0
        t0 = time.perf_counter()
        energy(P, F, g, steps)
        torch.cuda.synchronize()
        t_fwd = time.perf_counter() - t0
        out[name] = {"ub": u.grad.double(), "dA": float(A_t.grad),
                     "W": float(W), "ratio": t_adj / t_fwd}
        say("%s: %d steps, W %.9e dW/dA %.6e; forward %.3f s, "
            "forward+adjoint %.3f s, ratio %.2f"
            % (name, steps, float(W), float(A_t.grad), t_fwd, t_adj,
               t_adj / t_fwd))
    a, b = out["fp32"], out["fp64"]
    out["fp32_vs_fp64_norm"] = float((a["ub"] - b["ub"]).norm() /
                                     b["ub"].norm())
    out["fp32_vs_fp64_aniso"] = abs(a["dA"] - b["dA"]) / abs(b["dA"])
    say("FP32 against FP64 adjoint: momenta %.2e (relative norm), "
        "anisotropy %.2e" % (out["fp32_vs_fp64_norm"],
                            out["fp32_vs_fp64_aniso"]))
    return out


def _fd_directions(g, p, steps, grads, ndir, seed, h=1e-7, tries=24,
                   npart=64):
    """FP64 central differences of the FP64 run along sparse random
    directions, against both adjoints.  A direction counts only where the
    difference quotient has converged (h and 10 h agree to 1e-4): where it
    has not, some particle changed cell sequence between the difference
    points, or the run's summation-order noise exceeds the difference,
    and the quotient measures neither the derivative nor the adjoint.
    Returns the relative errors (FP64, FP32) of the converged directions
    and the number skipped."""
    P64 = p64_tensor(p)
    F64 = f_tensor(np.zeros(g.nv, T.FIELD), True)
    base_u = P64[:, 4:7].clone()

    def w_at(du):
        Q = P64.clone()
        Q[:, 4:7] = base_u + du
        return energy(Q, F64, g, steps)

    rng = np.random.default_rng(seed)
    e64, e32, skipped = [], [], 0
    for _ in range(tries):
        if len(e64) == ndir:
            break
        v = _sparse_dirs(rng, len(p), npart)
        fd = (w_at(h * v) - w_at(-h * v)) / (2 * h)
        fd10 = (w_at(10 * h * v) - w_at(-10 * h * v)) / (20 * h)
        g64 = float((grads["fp64"]["ub"] * v).sum())
        g32 = float((grads["fp32"]["ub"] * v).sum())
        if abs(fd - fd10) > 1e-4 * abs(fd):
            skipped += 1
            say("direction skipped: fd(h) %.9e fd(10h) %.9e" % (fd, fd10))
            continue
        e64.append(abs(g64 - fd) / abs(fd))
        e32.append(abs(g32 - fd) / abs(fd))
        say("direction: fd %.9e adjoint fp64 %.9e (%.2e) fp32 %.9e (%.2e)"
            % (fd, g64, e64[-1], g32, e32[-1]))
    return e64, e32, skipped


def test_weibel_gradient_reduced(lib):
    """The Weibel physics on a 16^3 box, 32 ppc (131,072 electrons).  Over
    50 steps (checkpoint every 10), the FP64 adjoint within 1e-3 of FP64
    central differences along 8 sparse random directions whose difference
    quotient has converged (measured ~1e-6), and the FP32 product adjoint
    within 1e-2 of them (measured <= 3e-4 on three directions).  Over 200
    steps (checkpoint every 20), the FP32 adjoint within 5e-2 of the FP64
    adjoint in relative norm and in d W / d(anisotropy) (measured 1.6e-2 and
    6.9e-3: the two trajectories of this unstable run part by rounding)."""
    aniso, uth_par = 4.0, 0.05
    g, p, nrm = weibel_box(16, 32, seed=3, aniso=aniso, uth_par=uth_par)
    r50 = _gradients(g, p, nrm, 50, 10, aniso, uth_par)
    e64, e32, skipped = _fd_directions(g, p, 50, r50, 8, 11)
    say("50 steps: %d converged directions, %d skipped; worst FP64 %.2e, "
        "FP32 %.2e" % (len(e64), skipped, max(e64), max(e32)))
    assert len(e64) == 8
    assert max(e64) <= 1e-3
    assert max(e32) <= 1e-2
    r200 = _gradients(g, p, nrm, 200, 20, aniso, uth_par)
    assert r200["fp32_vs_fp64_norm"] <= 5e-2
    assert r200["fp32_vs_fp64_aniso"] <= 5e-2


@pytest.mark.timeout(900)
def test_weibel_gradient_full_size(lib):
    """The Weibel run at full size: 64^3 cells, 32 ppc (8,388,608
    electrons), 200 steps, a checkpoint every 20: the FP32 product adjoint
    within 5e-2 of the FP64 adjoint (relative norm and d W /
    d(anisotropy)), at most 4 forwards of cost; over 20 steps the FP64
    adjoint within 1e-3 of FP64 central differences along 8 sparse random
    directions (512 particles each) whose difference quotient has
    converged."""
    aniso, uth_par = 4.0, 0.05
    g, p, nrm = weibel_box(64, 32, seed=3, aniso=aniso, uth_par=uth_par)
    r = _gradients(g, p, nrm, 200, 20, aniso, uth_par)
    assert r["fp32_vs_fp64_norm"] <= 5e-2
    assert r["fp32_vs_fp64_aniso"] <= 5e-2
    assert r["fp32"]["ratio"] <= 4.0
    r20 = _gradients(g, p, nrm, 20, 5, aniso, uth_par)
    e64, e32, skipped = _fd_directions(g, p, 20, r20, 8, 12, npart=512)
    say("20 steps: %d converged directions, %d skipped; worst FP64 %.2e, "
        "FP32 %.2e" % (len(e64), skipped, max(e64), max(e32)))
    assert len(e64) == 8
    assert max(e64) <= 1e-3
