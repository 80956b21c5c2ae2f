"""Python side of the pseudo-spectral Navier-Stokes solver (the caller of the
distributed plan in ``examples/navier-stokes-MPI/main.cpp`` and
``examples/navier-stokes-MPI-batched/main.cpp``) and a
``torch.autograd.Function`` over its checkpointed adjoint.

All arithmetic runs in ``libdfft.so`` (HIP kernels); this module only holds
per-rank torch buffers and passes their device pointers through the C ABI.
Field sets are lists of P tensors, one per virtual rank: physical blocks are
float64 ``(3, nx_r, ny_r, n)`` in the (1, px, py, 1) topology, spectral blocks
complex128 ``(3, n, ny_r, nz_r)`` in the (1, 1, px, py) topology with
``nz = n // 2 + 1`` (the batched example's layout, main.cpp:102-107).
"""

from __future__ import annotations

import ctypes

import dfft
from dfft import DFFTError, _check, _ptrs, _stream_handle, load_library

DIVERGENCE = 0  # u.grad(u) as div(u u): 3 backward + 6 forward transforms per stage
ADVECTIVE = 1   # the reference's form: 12 backward + 3 forward transforms per stage


class NavierStokes:
    """``NavierStokes`` of main.cpp:484-986 on a px x py grid of virtual ranks."""

    def __init__(self, n, px=1, py=1, lx=2.0, nu=1.0 / 1600.0, dt=1e-3, form=DIVERGENCE):
        self._lib = load_library()
        self.n, self.px, self.py = int(n), int(px), int(py)
        self.lx, self.nu, self.dt, self.form = float(lx), float(nu), float(dt), int(form)
        self.nranks = self.px * self.py
        d = dfft.NsDesc(self.n, self.px, self.py, self.lx, self.nu, self.dt, self.form)
        self._h = ctypes.c_void_p()
        _check(self._lib.dfft_ns_create(ctypes.byref(d), ctypes.byref(self._h)))

    # -- blocks -----------------------------------------------------------
    def local_extents(self, vrank, which):
        """(extents, starts) of rank ``vrank``'s physical (0) or spectral (1) block, without the field axis."""
        e = (ctypes.c_size_t * 3)()
        s = (ctypes.c_size_t * 3)()
        _check(self._lib.dfft_ns_local_extents(self._h, int(vrank), int(which), e, s))
        return [int(v) for v in e], [int(v) for v in s]

    def physical_fields(self, nf=3, device="cuda"):
        import torch
        return [torch.zeros([nf] + self.local_extents(r, 0)[0], dtype=torch.float64, device=device)
                for r in range(self.nranks)]

    def spectral_fields(self, nf=3, device="cuda"):
        import torch
        return [torch.zeros([nf] + self.local_extents(r, 1)[0], dtype=torch.complex128, device=device)
                for r in range(self.nranks)]

    def _check_set(self, ts, which):
        if len(ts) != self.nranks:
            raise DFFTError(dfft.ERROR_INVALID_ARGUMENT, f"expected {self.nranks} blocks, got {len(ts)}")
        import torch
        dt = torch.float64 if which == 0 else torch.complex128
        for r, t in enumerate(ts):
            want = [3] + self.local_extents(r, which)[0]
            if list(t.shape) != want or t.dtype != dt or not t.is_contiguous() or not t.is_cuda:
                raise DFFTError(dfft.ERROR_INVALID_EXTENTS,
                                f"rank {r}: expected a contiguous {dt} device block of shape {want}, "
                                f"got {t.dtype} {list(t.shape)}")

    # -- state ------------------------------------------------------------
    def init_taylor_green(self, v0=1.0, stream=None):
        _check(self._lib.dfft_ns_init_taylor_green(self._h, float(v0), _stream_handle(stream)))

    def set_physical(self, u, stream=None):
        self._check_set(u, 0)
        _check(self._lib.dfft_ns_set_physical(self._h, _ptrs(u), _stream_handle(stream)))

    def set_state(self, y, stream=None):
        self._check_set(y, 1)
        _check(self._lib.dfft_ns_set_state(self._h, _ptrs(y), _stream_handle(stream)))

    def get_state(self, stream=None):
        y = self.spectral_fields()
        _check(self._lib.dfft_ns_get_state(self._h, _ptrs(y), _stream_handle(stream)))
        return y

    # -- time stepping ----------------------------------------------------
    def step(self, nsteps=1, use_graph=False, stream=None):
        _check(self._lib.dfft_ns_step(self._h, int(nsteps), int(bool(use_graph)), _stream_handle(stream)))

    def autotune(self, stream=None):
        """Tune the launch shapes of every plan of the solver (the state is kept)."""
        _check(self._lib.dfft_ns_autotune(self._h, _stream_handle(stream)))

    def set_option(self, key, value):
        """"fuse_products" (1 default / 0): form the 6 products inside the
        forward transform's first pass instead of storing them."""
        _check(self._lib.dfft_ns_set_option(self._h, key.encode(), int(value)))

    def get_option(self, key):
        v = ctypes.c_long()
        _check(self._lib.dfft_ns_get_option(self._h, key.encode(), ctypes.byref(v)))
        return int(v.value)

    def energy(self, stream=None) -> float:
        e = ctypes.c_double()
        _check(self._lib.dfft_ns_energy(self._h, ctypes.byref(e), _stream_handle(stream)))
        return float(e.value)

    # -- diagnostics (main.cpp:812-935) --------------------------------------
    def velocity(self, stream=None):
        u = self.physical_fields()
        _check(self._lib.dfft_ns_velocity(self._h, _ptrs(u), _stream_handle(stream)))
        return u

    def vorticity(self, stream=None):
        w = self.physical_fields()
        _check(self._lib.dfft_ns_vorticity(self._h, _ptrs(w), _stream_handle(stream)))
        return w

    def energy_spectrum(self, num_bins=50, stream=None):
        """(E_spec, k_vals): radial bins of 0.5 f |u_k|^2 (unscaled, as the reference's E_spec)."""
        e = (ctypes.c_double * num_bins)()
        k = (ctypes.c_double * num_bins)()
        _check(self._lib.dfft_ns_energy_spectrum(self._h, int(num_bins), e, k, _stream_handle(stream)))
        return list(e), list(k)

    def diag(self, out_dir, diag_it, stream=None) -> float:
        """Writes the reference's per-rank .dat files; returns the kinetic energy."""
        e = ctypes.c_double()
        _check(self._lib.dfft_ns_diag(self._h, str(out_dir).encode(), int(diag_it), ctypes.byref(e),
                                      _stream_handle(stream)))
        return float(e.value)

    # -- derivatives (divergence form) --------------------------------------
    def rhs(self, y, stream=None):
        self._check_set(y, 1)
        out = self.spectral_fields()
        _check(self._lib.dfft_ns_rhs(self._h, _ptrs(y), _ptrs(out), _stream_handle(stream)))
        return out

    def rhs_tangent(self, y, d, stream=None):
        self._check_set(y, 1)
        self._check_set(d, 1)
        out = self.spectral_fields()
        _check(self._lib.dfft_ns_rhs_tangent(self._h, _ptrs(y), _ptrs(d), _ptrs(out), _stream_handle(stream)))
        return out

    def rhs_adjoint(self, y, lam, stream=None):
        self._check_set(y, 1)
        self._check_set(lam, 1)
        out = self.spectral_fields()
        _check(self._lib.dfft_ns_rhs_adjoint(self._h, _ptrs(y), _ptrs(lam), _ptrs(out), _stream_handle(stream)))
        return out

    def gradient(self, nsteps, u0, checkpoint_every=5, stream=None):
        """J = kinetic energy after ``nsteps`` RK4 steps from physical velocity ``u0``; returns (J, dJ/du0)."""
        self._check_set(u0, 0)
        grad = self.physical_fields()
        j = ctypes.c_double()
        _check(self._lib.dfft_ns_gradient(self._h, int(nsteps), int(checkpoint_every), _ptrs(u0), _ptrs(grad),
                                          ctypes.byref(j), _stream_handle(stream)))
        return float(j.value), grad

    def final_energy(self, nsteps, u0, stream=None) -> float:
        """Forward only: kinetic energy after ``nsteps`` eager RK4 steps from ``u0``."""
        self.set_physical(u0, stream)
        self.step(nsteps, False, stream)
        return self.energy(stream)

    def destroy(self):
        if getattr(self, "_h", None) and self._h.value:
            self._lib.dfft_ns_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def _energy_function():
    import torch

    class FinalEnergy(torch.autograd.Function):
        """J(u0) = kinetic energy after ``nsteps`` RK4 steps; backward runs the checkpointed adjoint."""

        @staticmethod
        def forward(ctx, solver, nsteps, checkpoint_every, *u0):
            u0 = [t.detach().contiguous() for t in u0]
            j, grad = solver.gradient(nsteps, u0, checkpoint_every)
            ctx.save_for_backward(*grad)
            return torch.tensor(j, dtype=torch.float64, device=u0[0].device)

        @staticmethod
        def backward(ctx, g):
            return (None, None, None) + tuple(gr * g for gr in ctx.saved_tensors)

    return FinalEnergy


_FINAL_ENERGY = None


def final_energy(solver: NavierStokes, nsteps: int, u0, checkpoint_every: int = 5):
    """Differentiable kinetic energy after ``nsteps`` steps from the per-rank physical blocks ``u0``."""
    global _FINAL_ENERGY
    if _FINAL_ENERGY is None:
        _FINAL_ENERGY = _energy_function()
    return _FINAL_ENERGY.apply(solver, int(nsteps), int(checkpoint_every), *u0)
