"""The differentiable run as a torch.autograd.Function.

`BFieldEnergy.apply(u, part, fields, grid, q, m, n_step, ckpt)` runs n_step
steps of the hot path (vpic_hip_adjoint_bfield_energy: one pushed species
on a periodic box) from particles whose momenta are the (np, 3) tensor `u`
and whose positions, cells and weights come from `part` (a device particle
array of the run's precision), and returns the magnetic field energy at the
end.  Its backward is the library's hand-written adjoint (computed with the
forward, checkpointed), so d(energy)/du flows into whatever produced u --
e.g. the anisotropy of a bi-Maxwellian loader.

torch holds the device memory and chains the gradients; all compute is the
C-ABI's.
"""
import ctypes

import numpy as np
import torch

from . import _lib


def particle_tensor(np_, precision, device="cuda"):
    """An empty device particle array: (np, 8) float32 rows for precision 4
    (vpic_particle_t), (np, 8) float64 rows for precision 8 (dx, dy, dz,
    (i, pad) as one double's bytes, ux, uy, uz, w)."""
    dt = torch.float32 if precision == 4 else torch.float64
    return torch.empty((np_, 8), dtype=dt, device=device)


def set_momenta(part, u):
    """part with columns 4..6 replaced by u (a new tensor)."""
    out = part.clone()
    out[:, 4:7] = u
    return out


def run(part, fields, grid, q, m, n_step, ckpt=20, grads=True,
        want_positions=False, want_fields=False):
    """The raw call: returns (energy, u_bar, d_bar, f_bar) with None for the
    gradients not asked for."""
    precision = 4 if part.dtype == torch.float32 else 8
    np_ = part.shape[0]
    en = ctypes.c_double()
    ub = db = fb = None
    if grads:
        ub = torch.empty((np_, 3), dtype=part.dtype, device=part.device)
        if want_positions:
            db = torch.empty_like(ub)
        if want_fields:
            fb = torch.empty((grid.nv, 6), dtype=part.dtype,
                             device=part.device)
    _lib.call("vpic_hip_adjoint_bfield_energy", part.data_ptr(), np_,
              float(q), float(m), fields.data_ptr(), ctypes.byref(grid),
              int(n_step), int(ckpt), precision, ctypes.byref(en),
              ub.data_ptr() if ub is not None else None,
              db.data_ptr() if db is not None else None,
              fb.data_ptr() if fb is not None else None, None)
    return en.value, ub, db, fb


class BFieldEnergy(torch.autograd.Function):
    """W = magnetic field energy after n_step steps, differentiable in the
    initial momenta u (np, 3)."""

    @staticmethod
    def forward(ctx, u, part, fields, grid, q, m, n_step, ckpt):
        p = set_momenta(part, u.detach())
        need = u.requires_grad
        w, ub, _, _ = run(p, fields, grid, q, m, n_step, ckpt, grads=need)
        ctx.save_for_backward(ub if need else u.new_zeros(0))
        return torch.tensor(w, dtype=torch.float64, device=u.device)

    @staticmethod
    def backward(ctx, gw):
        (ub,) = ctx.saved_tensors
        gu = ub.to(torch.float64) * gw
        return (gu.to(ub.dtype), None, None, None, None, None, None, None)


def bimaxwellian_momenta(normals, uth_par, anisotropy):
    """u = (uth_par n_x, sqrt(A) uth_par n_y, sqrt(A) uth_par n_z): the
    bi-Maxwellian with T_perp / T_par = A about x, from fixed standard
    normals (np, 3); differentiable in A (a 0-d tensor)."""
    s = torch.sqrt(anisotropy)
    scale = torch.stack([torch.ones_like(s), s, s]) * uth_par
    return normals * scale.to(normals.dtype)


def host_particles64(p32):
    """FP32 particles (numpy, vpic_amd.types.PARTICLE) widened to the FP64
    layout as a (np, 8) float64 array."""
    out = np.zeros((len(p32), 8), dtype=np.float64)
    for k, c in enumerate(("dx", "dy", "dz")):
        out[:, k] = p32[c]
    iv = out.view(np.int32)  # (np, 16): words 6, 7 are column 3
    iv[:, 6] = p32["i"]
    iv[:, 7] = 0
    for k, c in enumerate(("ux", "uy", "uz", "w")):
        out[:, 4 + k] = p32[c]
    return out


def fields64(f32):
    """A field array (vpic_amd.types.FIELD) in the FP64 layout as (nv, 18)
    float64 (16 components, then the material ids' bytes)."""
    out = np.zeros((len(f32), 18), dtype=np.float64)
    names = f32.dtype.names
    for k in range(16):
        out[:, k] = f32[names[k]]
    return out
