"""DGSEM operator over the HIP library, mirroring Raven.jl's balance-law API.

``DGSEM(law=..., grid=..., surface_numericalflux=...)`` (dgsem.jl:46-81) and
its call ``dg(dq, q, t, increment=True)`` (dgsem.jl:84-126).  State arrays are
torch tensors on the HIP device shaped ``(K, Ns, Np)`` (C-order), which is the
memory of Raven's GridArray parent ``(Np, Ns, K)``.  torch only provides device
memory and the stream; all arithmetic happens in libravenhip's kernels.
"""
import ctypes

import numpy as np

from . import _lib as L


class AdvectionLaw:
    """AdvectionLaw{FT,D}(u⃗) (balancelaws/advection.jl:8-16); u⃗ defaults to ones."""

    law_id = L.RV_LAW_ADVECTION

    def __init__(self, dim, u=None):
        self.dim = dim
        self.u = tuple(float(v) for v in (u if u is not None else (1.0,) * dim))
        if len(self.u) != dim:
            raise ValueError("velocity needs one component per dimension")
        self.gamma = 1.4
        self.ns = 1


class EulerLaw:
    """EulerLaw{FT,D}(; γ = 7//5) (balancelaws/euler.jl:9-19)."""

    law_id = L.RV_LAW_EULER

    def __init__(self, dim, gamma=7 / 5):
        self.dim = dim
        self.gamma = float(gamma)
        self.u = (0.0,) * dim
        self.ns = dim + 2


class RusanovFlux:
    flux_id = L.RV_FLUX_RUSANOV


class CentralFlux:
    flux_id = L.RV_FLUX_CENTRAL


class EntropyConservativeFlux:
    """EntropyConservativeFlux (numericalfluxes.jl:56; Euler: euler.jl:150-184);
    a volume (two-point) flux of FluxDifferencingForm, and with that form also
    a surface flux, twopointflux(q-, q+)' n (numericalfluxes.jl:27-37)."""
    flux_id = L.RV_FLUX_EC


class WeakForm:
    """WeakForm() (dgsem.jl:7): the flux-divergence volume term."""
    form_id = L.RV_FORM_WEAK
    flux_id = 0


class FluxDifferencingForm:
    """FluxDifferencingForm(volume_numericalflux, kernel_type = :per_dir)
    (dgsem.jl:8-17): the volume term from a two-point flux (CentralFlux, or
    EntropyConservativeFlux for Euler).  Raven's three kernel types (:naive,
    :per_dir, :per_dir_symmetric; dgsem.jl:163-255) compute the same operator
    in different orders; the library evaluates every symmetric pair once, as
    :per_dir_symmetric does, for all of them.  :naive keeps Raven's
    restriction to all directions (dgsem.jl:180-181)."""
    form_id = L.RV_FORM_FLUX_DIFFERENCING
    KERNEL_TYPES = ("naive", "per_dir", "per_dir_symmetric")

    def __init__(self, volume_numericalflux, kernel_type="per_dir"):
        kernel_type = kernel_type.lstrip(":")
        if kernel_type not in self.KERNEL_TYPES:
            raise ValueError(f"unknown kernel_type {kernel_type!r}")
        self.volume_numericalflux = volume_numericalflux
        self.flux_id = volume_numericalflux.flux_id
        self.kernel_type = kernel_type


_DTYPES = {"f64": L.RV_F64, "f32": L.RV_F32}
_GEOMS = {"auto": L.RV_GEOM_AUTO, "affine": L.RV_GEOM_AFFINE, "general": L.RV_GEOM_GENERAL}


def _torch():
    import torch
    return torch


def _current_stream():
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class DGSEM:
    """Fused DGSEM right-hand side on one MI355X.

    grid      -- a Grid from generate() (explicit arrays, the drop-in path)
    law       -- AdvectionLaw / EulerLaw
    dtype     -- "f64" or "f32" (FP32 state/flux arithmetic, FP64 geometry)
    geometry  -- "auto" (affine cells detected), "affine" or "general"
    """

    def __init__(self, *, law, grid, surface_numericalflux=None, dtype="f64", geometry="auto",
                 directions=None, stream=None, volume_form=None, _handle=None):
        self.law = law
        self.grid = grid
        self.dtype = dtype
        self.volume_form = volume_form or WeakForm()
        flux = surface_numericalflux or RusanovFlux()
        if (getattr(self.volume_form, "kernel_type", None) == "naive" and directions is not None
                and sorted(directions) != list(range(1, law.dim + 1))):
            # esvolumeterm! (:naive) has no directions (dgsem.jl:180-181)
            raise ValueError("the :naive flux-differencing kernel needs all directions")
        if _handle is not None:
            self._h = _handle
        else:
            if grid.dim != law.dim:
                raise ValueError("law and grid dimensions differ")
            d = L.RvDesc()
            d.abi_version = L.ABI_VERSION
            d.dim, d.nq = grid.dim, grid.nq
            d.law, d.flux = law.law_id, flux.flux_id
            d.dtype = _DTYPES[dtype]
            d.geometry = _GEOMS[geometry]
            d.directions = 0 if directions is None else sum(1 << (k - 1) for k in directions)
            d.ncells = grid.K
            d.gamma = law.gamma
            for k in range(3):
                d.velocity[k] = law.u[k] if k < len(law.u) else 0.0
            cell = grid.referencecell
            keep = [np.ascontiguousarray(a) for a in (
                cell.D_colmajor, cell.weights_1d, grid.metrics, grid.surface)]
            ikeep = [None if a is None else np.ascontiguousarray(a, dtype=np.int64) for a in (
                grid.vmapM, grid.vmapP, grid.boundarycodes, grid.dtoc)]
            d.D, d.weights, d.metrics, d.surface_metrics = (L.ptr(a, ctypes.c_double) for a in keep)
            d.vmapM, d.vmapP, d.boundary_codes, d.dtoc = (L.ptr(a, ctypes.c_int64) for a in ikeep)
            d.ncontinuous = grid.ncontinuous
            d.stream = stream if stream is not None else _current_stream()
            d.volume_form, d.volume_flux = self.volume_form.form_id, self.volume_form.flux_id
            h = L.H()
            L.check(L.lib().rv_create(ctypes.byref(d), ctypes.byref(h)))
            self._h = h
        info = L.RvInfo()
        L.check(L.lib().rv_get_info(self._h, ctypes.byref(info)), self._h)
        self.info = info
        self.ns = info.ns
        self.K = info.ncells
        self.Np = info.nq ** info.dim

    @classmethod
    def brick(cls, gm, law, surface_numericalflux=None, dtype="f64", stream=None,
              volume_form=None):
        """DGSEM on a library-generated uniform brick (no per-node host arrays)."""
        flux = surface_numericalflux or RusanovFlux()
        vf = volume_form or WeakForm()
        desc = gm.brick_desc(law=law.law_id, dtype=_DTYPES[dtype], flux=flux.flux_id,
                             gamma=law.gamma, velocity=law.u,
                             stream=stream if stream is not None else _current_stream(),
                             volume_form=vf.form_id, volume_flux=vf.flux_id)
        h = L.H()
        L.check(L.lib().rv_create_brick(ctypes.byref(desc), ctypes.byref(h)))
        return cls(law=law, grid=None, surface_numericalflux=flux, dtype=dtype, volume_form=vf,
                   _handle=h)

    @classmethod
    def forest(cls, gm, law, surface_numericalflux=None, dtype="f64", stream=None,
               volume_form=None, points=None):
        """DGSEM on an adapted GridManager's grid generated on the device
        (rv_create_forest): points (a device tensor (K, dim, Np), e.g. warped)
        or the unwarped brick points, curved metrics and geometry computed on
        the GPU, no per-node host arrays."""
        flux = surface_numericalflux or RusanovFlux()
        vf = volume_form or WeakForm()
        if not gm.adapted:
            gm._ensure_forest()
        desc = gm.brick_desc(law=law.law_id, dtype=_DTYPES[dtype], flux=flux.flux_id,
                             gamma=law.gamma, velocity=law.u,
                             stream=stream if stream is not None else _current_stream(),
                             volume_form=vf.form_id, volume_flux=vf.flux_id)
        pts = None
        if points is not None:
            if points.dtype != _torch().float64 or not points.is_contiguous() or \
                    points.device.type != "cuda":
                raise ValueError("points must be a contiguous float64 device tensor")
            pts = ctypes.c_void_p(points.data_ptr())
        h = L.H()
        L.check(L.lib().rv_create_forest(gm._forest, ctypes.byref(desc), pts, ctypes.byref(h)))
        dg = cls(law=law, grid=None, surface_numericalflux=flux, dtype=dtype, volume_form=vf,
                 _handle=h)
        dg._points_keep = points
        return dg

    # -- lifetime ----------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            L.lib().rv_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    @property
    def stream(self):
        return L.lib().rv_stream(self._h)

    def synchronize(self):
        L.check(L.lib().rv_synchronize(self._h), self._h)

    # -- state helpers ------------------------------------------------------
    @property
    def torch_dtype(self):
        torch = _torch()
        return torch.float64 if self.dtype == "f64" else torch.float32

    def GridArray(self, fill=0.0, device="cuda"):
        torch = _torch()
        return torch.full((self.K, self.ns, self.Np), fill, dtype=self.torch_dtype, device=device)

    def _check_state(self, t):
        if t.dtype != self.torch_dtype or not t.is_cuda or not t.is_contiguous():
            raise ValueError("state must be a contiguous device tensor of the operator's dtype")
        if t.numel() != self.K * self.ns * self.Np:
            raise ValueError("state has the wrong number of entries")
        return ctypes.c_void_p(t.data_ptr())

    # -- the hot path ------------------------------------------------------
    def __call__(self, dq, q, time=0.0, increment=True):
        """dg(dq, q, t; increment) (dgsem.jl:84-126)."""
        L.check(L.lib().rv_rhs(self._h, self._check_state(dq), self._check_state(q),
                               float(time), 1 if increment else 0), self._h)

    def lsrk_stage(self, q_out, q_in, dq, a, b_dt, time=0.0):
        L.check(L.lib().rv_lsrk_stage(self._h, self._check_state(q_out), self._check_state(q_in),
                                      self._check_state(dq), float(a), float(b_dt), float(time)),
                self._h)

    def enable_mortars(self, grid=None, from_nctables=False):
        """Couple the hanging faces of an adapted grid through mortars
        (rv_enable_mortars) instead of Raven's self-mapped faces.  The face
        kinds and partners come from the forest, or with from_nctables from
        Raven's own facemaps tables vmapNC / nctoface (rv_mortars_from_nctables)."""
        from .grids import mortars_from_nctables
        g = grid if grid is not None else self.grid
        if from_nctables:
            kind, partner = mortars_from_nctables(g.dim, g.nq, g.K, g.vmapNC, g.nctoface)
        else:
            kind = np.ascontiguousarray(g.face_kind, dtype=np.int32)
            partner = np.ascontiguousarray(g.mortar_partner, dtype=np.int64)
        L.check(L.lib().rv_enable_mortars(self._h, L.ptr(kind, ctypes.c_int32),
                                          L.ptr(partner, ctypes.c_int64)), self._h)

    # -- launch shape ------------------------------------------------------
    @property
    def waves_per_block(self):
        return L.lib().rv_get_waves_per_block(self._h)

    @waves_per_block.setter
    def waves_per_block(self, w):
        L.check(L.lib().rv_set_waves_per_block(self._h, int(w)), self._h)

    def autotune(self, q, choices=(1, 2, 4), repeats=5):
        """Time one fused LSRK stage on state q for each launch shape and keep
        the fastest; returns {waves_per_block: mean ms}.  q is not modified."""
        torch = _torch()
        out, dq = torch.empty_like(q), torch.zeros_like(q)
        stream = torch.cuda.ExternalStream(self.stream) if self.stream else \
            torch.cuda.default_stream()
        times = {}
        for w in choices:
            self.waves_per_block = w
            self.lsrk_stage(out, q, dq, 0.0, 0.0)  # warm-up
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
            ev[0].record(stream)
            for _ in range(repeats):
                self.lsrk_stage(out, q, dq, 0.0, 0.0)
            ev[1].record(stream)
            ev[1].synchronize()
            times[w] = ev[0].elapsed_time(ev[1]) / repeats
        self.waves_per_block = min(times, key=times.get)
        return times

    # -- diagnostics ------------------------------------------------------
    def weightednorm_components(self, q):
        out = np.zeros(self.ns)
        L.check(L.lib().rv_weightednorm(self._h, self._check_state(q),
                                        L.ptr(out, ctypes.c_double)), self._h)
        return out

    def entropy_integral(self, q):
        out = np.zeros(1)
        L.check(L.lib().rv_entropy_integral(self._h, self._check_state(q),
                                            L.ptr(out, ctypes.c_double)), self._h)
        return float(out[0])

    def entropy_product(self, p, q):
        out = np.zeros(1)
        L.check(L.lib().rv_entropy_product(self._h, self._check_state(p), self._check_state(q),
                                           L.ptr(out, ctypes.c_double)), self._h)
        return float(out[0])

    def integrals(self, q):
        out = np.zeros(self.ns)
        L.check(L.lib().rv_integrals(self._h, self._check_state(q), L.ptr(out, ctypes.c_double)),
                self._h)
        return out

    # -- GatherScatter -------------------------------------------------------
    def gather(self, dest, src, nf):
        L.check(L.lib().rv_gather(self._h, ctypes.c_void_p(dest.data_ptr()),
                                  ctypes.c_void_p(src.data_ptr()), int(nf)), self._h)

    def scatter(self, dest, src, nf):
        L.check(L.lib().rv_scatter(self._h, ctypes.c_void_p(dest.data_ptr()),
                                   ctypes.c_void_p(src.data_ptr()), int(nf)), self._h)

    def maps(self):
        n = 2 * self.info.dim * self.info.nq ** (self.info.dim - 1) * self.K
        out = [np.zeros(n, dtype=np.int64) for _ in range(4)]
        L.check(L.lib().rv_get_maps(self._h, *(L.ptr(a, ctypes.c_int64) for a in out)), self._h)
        return out


def entropyintegral(dg, q):
    """entropyintegral(dg, q) = sum(MJ .* entropy.(q)) (measures.jl:23-27)."""
    return dg.entropy_integral(q)


def entropyproduct(dg, p, q):
    """entropyproduct(dg, p, q) = sum(MJ .* v(p) . q) (measures.jl:28-33)."""
    return dg.entropy_product(p, q)


def weightednorm(dg, q, p=2, componentwise=False):
    """weightednorm(dg, q, p = 2; componentwise = false) (measures.jl:6-21):
    per state (sum MJ |q_s|^p)^(1/p) (p = inf: max |q_s|), compensated FP64
    sums, then the p-norm over the states unless componentwise."""
    if p == 2:
        comp = dg.weightednorm_components(q)
    else:
        comp = np.zeros(dg.ns)
        L.check(L.lib().rv_weightednorm_p(dg.handle, dg._check_state(q), float(p),
                                          L.ptr(comp, ctypes.c_double)), dg.handle)
    if componentwise:
        return comp
    if p == 2:
        return float(np.sqrt(np.sum(comp * comp)))
    return float(np.linalg.norm(comp, p))
