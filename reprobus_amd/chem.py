"""Host-side interface of the GPU chemistry step.

Mirrors the model's chemistry sub-step (src/reprobus_l137_20230723.f:398-440):
``chem3d`` over every latitude row (F:398-415), then ``conserv`` (F:427) and the
1e-30 floor (F:431-440).  Arrays are in the model's Fortran layout, held as
C-ordered numpy arrays / torch tensors of shape (nspecies, niv, nlat, nlon).

Two ways in:
  * host arrays (numpy float64): ``chem3d`` / ``conserv`` copy to the device, run
    and copy back, updating the caller's arrays in place like the Fortran routines;
  * device-resident torch tensors: ``step_device`` enqueues the kernels on the
    current torch stream, the path the model's time loop and the benchmark use.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import Config, DayConfig, Grid, RbError, Stats


def _numel(a) -> int:
    return int(a.size) if isinstance(a, np.ndarray) else int(a.numel())


def _ptr(a, n: int | None = None, what: str = "array") -> int:
    """Address of a C-contiguous float64 numpy array or GPU tensor holding at least n
    values (the extent the C ABI reads or writes)."""
    if isinstance(a, np.ndarray):
        if a.dtype != np.float64 or not a.flags.c_contiguous:
            raise TypeError(f"{what}: arrays must be C-contiguous float64")
        p = a.ctypes.data
    else:   # torch tensor
        if a.dtype.__repr__() != "torch.float64" or not a.is_contiguous() or a.device.type != "cuda":
            raise TypeError(f"{what}: tensors must be contiguous float64 on a GPU")
        p = a.data_ptr()
    if n is not None and _numel(a) < n:
        raise ValueError(f"{what}: {_numel(a)} values, the grid needs {n}")
    return p


def _iptr(a, n: int, what: str) -> int:
    """Address of a contiguous int32 GPU tensor of at least n values."""
    if a.dtype.__repr__() != "torch.int32" or not a.is_contiguous() or a.device.type != "cuda":
        raise TypeError(f"{what}: must be a contiguous int32 tensor on a GPU")
    if _numel(a) < n:
        raise ValueError(f"{what}: {_numel(a)} values, the grid needs {n}")
    return a.data_ptr()


def _extents(grid) -> tuple[int, int, int]:
    """(field, nbox, plane): values per (nlon, nlat, niv) field, chemistry boxes, columns."""
    plane = int(grid.nlon) * int(grid.nlat)
    return plane * int(grid.niv), plane * int(grid.nivbas), plane


def _step_ptrs(grid, T, p, sza, alt, qj1, hc, surfarea, vsed3d, parth2osed, parthno3sed):
    """Checked addresses of a chemistry step's ten arrays (inputs read on the chemistry
    levels, the state on every level)."""
    field, nbox, _ = _extents(grid)
    sizes = (nbox, nbox, nbox, nbox, 43 * field, 15 * field, field, nbox, nbox, nbox)
    names = ("T", "p", "sza", "alt", "qj1", "hc", "surfarea", "vsed3d", "parth2osed",
             "parthno3sed")
    return [_ptr(a, n, w) for a, n, w in zip((T, p, sza, alt, qj1, hc, surfarea, vsed3d,
                                              parth2osed, parthno3sed), sizes, names)]


def make_grid(nlon: int, nlat: int, niv: int, nivbas: int, nivhau: int) -> Grid:
    return Grid(nlon, nlat, niv, nivbas, nivhau)


class ChemistryStep:
    """One library context: device, stream, photolysis table, work buffers."""

    def __init__(self, ajl: np.ndarray, o3up: np.ndarray, device: int = 0, niter: int = 5,
                 block: int = 256, check_finite: bool = False, profile: bool = False,
                 conv_refill: int = 16, conv_aa_start: int = 60, conv_aa_budget: int = 240,
                 newton_waves: int = 2, conv_order: int = 0):
        self._L = _lib.load()
        self._ctx = C.c_void_p()
        self.conv_refill, self.conv_aa_start = conv_refill, conv_aa_start
        self.conv_aa_budget = conv_aa_budget
        cfg = Config(niter, 1 if check_finite else 0, block, 1 if profile else 0, conv_refill,
                     conv_aa_start, conv_aa_budget, newton_waves, conv_order)
        rc = self._L.rb_init(C.byref(self._ctx), device, C.byref(cfg))
        if rc != 0:
            raise RbError(rc, f"rb_init(device={device}) failed")
        ajl_f = np.ascontiguousarray(np.asarray(ajl, dtype=np.float64).transpose(3, 2, 1, 0))
        o3 = np.ascontiguousarray(o3up, dtype=np.float64)
        if ajl_f.size != 40 * 101 * 27 * 7 or o3.size != 80:
            raise ValueError("ajl must be (40,101,27,7) and o3up (80,)")
        self._check(self._L.rb_load_jtable(self._ctx, ajl_f.ctypes.data, o3.ctypes.data))

    # -- plumbing -------------------------------------------------------------------
    def _check(self, rc: int) -> None:
        if rc != 0:
            msg = self._L.rb_last_error(self._ctx)
            raise RbError(rc, msg.decode() if msg else "")

    def close(self) -> None:
        if self._ctx:
            self._L.rb_finalize(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def stats(self) -> dict:
        s = Stats()
        self._check(self._L.rb_get_stats(self._ctx, C.byref(s)))
        return {"boxes": s.boxes, "nonfinite": s.nonfinite, "kernel_ms": s.last_kernel_ms,
                "h2d_ms": s.h2d_ms, "d2h_ms": s.d2h_ms, "chem_ms_total": s.chem_ms_total,
                "chem_launches": s.chem_launches}

    def set_stream(self, stream_handle: int | None) -> None:
        self._check(self._L.rb_set_stream(self._ctx, stream_handle))

    def sync(self) -> None:
        self._check(self._L.rb_sync(self._ctx))

    # -- host arrays (drop-in for the Fortran calls) ------------------------------
    def chem3d(self, grid: Grid, dt: float, T, p, sza, alt, qj1, hc, surfarea, vsed3d,
               parth2osed, parthno3sed) -> None:
        """All rows of chem3d (F:398-415) on host arrays, in place."""
        self._check(self._L.rb_chem_step(
            self._ctx, C.byref(grid), float(dt), *_step_ptrs(
                grid, T, p, sza, alt, qj1, hc, surfarea, vsed3d, parth2osed, parthno3sed)))

    def conserv(self, grid: Grid, qj1, hc) -> None:
        """conserv (F:2526-2616) + floor (F:431-440) on host arrays, in place."""
        field = _extents(grid)[0]
        self._check(self._L.rb_conserv_floor(self._ctx, C.byref(grid), _ptr(qj1, 43 * field, "qj1"),
                                             _ptr(hc, 15 * field, "hc")))

    # -- device tensors -----------------------------------------------------------
    def step_device(self, grid: Grid, dt: float, T, p, sza, alt, qj1, hc, surfarea, vsed3d,
                    parth2osed, parthno3sed, fuse_conserv: bool = True) -> None:
        """Enqueue one step on device tensors (conserv + floor fused when asked)."""
        self._check(self._L.rb_chem_step_dev(
            self._ctx, C.byref(grid), float(dt), *_step_ptrs(
                grid, T, p, sza, alt, qj1, hc, surfarea, vsed3d, parth2osed, parthno3sed),
            1 if fuse_conserv else 0))

    def converge_device(self, grid: Grid, dt: float, tol: float, maxiter: int, T, p, sza, alt,
                        qj1, hc, surfarea, vsed3d, parth2osed, parthno3sed, guess_qj1=None,
                        guess_hc=None, iters=None, fuse_conserv: bool = True,
                        accel: int = 0) -> None:
        """Converged mode: sweep every box to the fixed point (optionally from a guess;
        accel = 2: Anderson-accelerated sweeps)."""
        field, nbox, _ = _extents(grid)
        g_q = _ptr(guess_qj1, 42 * field, "guess_qj1") if guess_qj1 is not None else None
        g_h = _ptr(guess_hc, 15 * field, "guess_hc") if guess_hc is not None else None
        it = _iptr(iters, nbox, "iters") if iters is not None else None
        self._check(self._L.rb_chem_converge_dev(
            self._ctx, C.byref(grid), float(dt), float(tol), int(maxiter), int(accel),
            *_step_ptrs(grid, T, p, sza, alt, qj1, hc, surfarea, vsed3d, parth2osed,
                        parthno3sed), g_q, g_h, it, 1 if fuse_conserv else 0))

    def newton_device(self, grid: Grid, dt: float, tol: float, max_newton: int, maxiter: int,
                      T, p, sza, alt, qj1, hc, surfarea, vsed3d, parth2osed, parthno3sed,
                      guess_qj1=None, guess_hc=None, iters=None, fuse_conserv: bool = True,
                      sweeps_first: int = 0) -> None:
        """Converged mode by Newton iterations with the generated sparse LU (optionally from
        a guess; with sweeps_first > 0 after that many plain sweeps, for the boxes still
        unconverged).  iters: k sweeps, 10000 + Newton iterations, 20000 + sweeps (boxes
        Newton handed back to the sweeps), negative: fallback."""
        field, nbox, _ = _extents(grid)
        g_q = _ptr(guess_qj1, 42 * field, "guess_qj1") if guess_qj1 is not None else None
        g_h = _ptr(guess_hc, 15 * field, "guess_hc") if guess_hc is not None else None
        it = _iptr(iters, nbox, "iters") if iters is not None else None
        self._check(self._L.rb_chem_newton_dev(
            self._ctx, C.byref(grid), float(dt), float(tol), int(sweeps_first), int(max_newton),
            int(maxiter),
            *_step_ptrs(grid, T, p, sza, alt, qj1, hc, surfarea, vsed3d, parth2osed,
                        parthno3sed), g_q, g_h, it, 1 if fuse_conserv else 0))

    def surrogate_device(self, grid: Grid, dt: float, accept_tol: float, tol: float,
                         maxiter: int, T, p, sza, alt, qj1, hc, surfarea, vsed3d, parth2osed,
                         parthno3sed, sur_qj1, sur_hc, iters=None, fuse_conserv: bool = True,
                         accel: int = 0) -> None:
        """Surrogate step: keep the surrogate's answer (sur_qj1, sur_hc) where one FP64
        sweep from it changes every species by less than accept_tol (iters 0), sweep the
        other boxes on to tol from it."""
        field, nbox, _ = _extents(grid)
        it = _iptr(iters, nbox, "iters") if iters is not None else None
        self._check(self._L.rb_chem_surrogate_dev(
            self._ctx, C.byref(grid), float(dt), float(accept_tol), float(tol), int(maxiter),
            int(accel), *_step_ptrs(grid, T, p, sza, alt, qj1, hc, surfarea, vsed3d,
                                    parth2osed, parthno3sed),
            _ptr(sur_qj1, 42 * field, "sur_qj1"), _ptr(sur_hc, 15 * field, "sur_hc"), it,
            1 if fuse_conserv else 0))

    def sediment(self, grid: Grid, dt: float, nivh2o: int, T, p, qj1, vsed3d, parth2osed,
                 parthno3sed) -> None:
        """Sedimentation on host arrays (qj1 updated in place), F:2416-2522."""
        for a in (T, p, qj1, vsed3d, parth2osed, parthno3sed):
            assert a.dtype == np.float64 and a.flags.c_contiguous
        self._check(self._L.rb_sediment(self._ctx, C.byref(grid), float(dt), int(nivh2o),
                                        *[a.ctypes.data for a in (T, p, qj1, vsed3d, parth2osed,
                                                                  parthno3sed)]))

    def sediment_device(self, grid: Grid, dt: float, nivh2o: int, T, p, qj1, vsed3d, parth2osed,
                        parthno3sed) -> None:
        self._check(self._L.rb_sediment_dev(self._ctx, C.byref(grid), float(dt), int(nivh2o),
                                            *[_ptr(a) for a in (T, p, qj1, vsed3d, parth2osed,
                                                                parthno3sed)]))

    def zenith(self, grid: Grid, day: float, trajlon, trajlat, sza) -> None:
        """SZA of levels 1..nivbas (host arrays), F:5308-5406."""
        for a in (trajlon, trajlat, sza):
            assert a.dtype == np.float64 and a.flags.c_contiguous
        assert sza.size >= grid.nlon * grid.nlat * grid.nivbas
        self._check(self._L.rb_zenith(self._ctx, C.byref(grid), float(day), trajlon.ctypes.data,
                                      trajlat.ctypes.data, sza.ctypes.data))

    def zenith_device(self, grid: Grid, day: float, trajlon, trajlat, sza) -> None:
        self._check(self._L.rb_zenith_dev(self._ctx, C.byref(grid), float(day), _ptr(trajlon),
                                          _ptr(trajlat), _ptr(sza)))

    def altitude(self, grid: Grid, relief, psol, aaa, bbb, t, h2o, trajlon, trajlat, trajpre, alt,
                 dlon: float = 2.0, dlat: float = 2.0) -> None:
        """Geometric altitude of the trajectory points (host arrays), altitude.f90."""
        arrs = (relief, psol, aaa, bbb, t, h2o, trajlon, trajlat, trajpre, alt)
        for a in arrs:
            assert a.dtype == np.float64 and a.flags.c_contiguous
        self._check(self._L.rb_altitude(self._ctx, C.byref(grid), float(dlon), float(dlat),
                                        *[a.ctypes.data for a in arrs]))

    def altitude_device(self, grid: Grid, relief, psol, aaa, bbb, t, h2o, trajlon, trajlat, trajpre,
                        alt, dlon: float = 2.0, dlat: float = 2.0) -> None:
        arrs = (relief, psol, aaa, bbb, t, h2o, trajlon, trajlat, trajpre, alt)
        self._check(self._L.rb_altitude_dev(self._ctx, C.byref(grid), float(dlon), float(dlat),
                                            *[_ptr(a) for a in arrs]))

    # -- semi-Lagrangian transport interpolation -------------------------------------
    def backtra_device(self, grid: Grid, aaa, bbb, m1: dict, p1: dict, daynum: float,
                       dtdyn: float, nrapp: int, ntime: int, nintecmwf: int, x0, y0, z0, p0,
                       idx, w, traj=None) -> None:
        """Back-trajectories (backtra) of every node of the transport grid: m1, p1 are the
        forcing records {"u", "v", "w", "t": (niv, nlat, nlon), "lnps": (nlat, nlon),
        "day": float} as device tensors; outputs the departure points x0..p0 (n), cells idx
        (3, n) int32 and weights w (12, n), and with traj = (lon, lat, pre, tem) tensors of
        (nrapp, niv, nlat, nlon) the positions and temperatures along the way."""
        n = _extents(grid)[0]
        plane = int(grid.nlon) * int(grid.nlat)
        niv = int(grid.niv)
        recs = []
        for name, r in (("m1", m1), ("p1", p1)):
            recs.append(_lib.Forcing(*[_ptr(r[k], n, f"{name}.{k}") for k in ("u", "v", "w", "t")],
                                     _ptr(r["lnps"], plane, f"{name}.lnps"), float(r["day"])))
        tp = [None] * 4
        if traj is not None:
            tp = [_ptr(a, int(nrapp) * n, f"traj[{j}]") for j, a in enumerate(traj)]
        self._check(self._L.rb_transport_backtra_dev(
            self._ctx, C.byref(grid), _ptr(aaa, niv, "aaa"), _ptr(bbb, niv, "bbb"),
            C.byref(recs[0]), C.byref(recs[1]), float(daynum), float(dtdyn), int(nrapp),
            int(ntime), int(nintecmwf),
            *[_ptr(a, n, w_) for a, w_ in ((x0, "x0"), (y0, "y0"), (z0, "z0"), (p0, "p0"))],
            _iptr(idx, 3 * n, "idx"), _ptr(w, 12 * n, "w"), *tp))

    def transport_weights_device(self, grid: Grid, aaa, bbb, x0, y0, z0, p0, idx, w) -> None:
        """Cell indices idx (3, n) int32 and Lagrange weights w (12, n) of the departure
        points (device tensors; posx..poidz)."""
        n = _extents(grid)[0]
        niv = int(grid.niv)
        self._check(self._L.rb_transport_weights_dev(
            self._ctx, C.byref(grid), _ptr(aaa, niv, "aaa"), _ptr(bbb, niv, "bbb"),
            *[_ptr(a, n, w_) for a, w_ in ((x0, "x0"), (y0, "y0"), (z0, "z0"), (p0, "p0"))],
            _iptr(idx, 3 * n, "idx"), _ptr(w, 12 * n, "w")))

    def transport_interp_device(self, grid: Grid, idx, w, q_in, q_out) -> None:
        """The ntr fields of q_in (ntr, niv, nlat, nlon) interpolated at the departure points
        into q_out (internew for each tracer)."""
        ntr = int(q_in.shape[0])
        n = _extents(grid)[0]
        if tuple(q_out.shape) != tuple(q_in.shape):
            raise ValueError("q_out must have the shape of q_in")
        self._check(self._L.rb_transport_interp_dev(
            self._ctx, C.byref(grid), _iptr(idx, 3 * n, "idx"), _ptr(w, 12 * n, "w"), ntr,
            _ptr(q_in, ntr * n, "q_in"), _ptr(q_out, ntr * n, "q_out")))

    def transport(self, grid: Grid, aaa, bbb, x0, y0, z0, p0, q) -> None:
        """Host arrays: weights and interpolation of the ntr fields of q (ntr, niv, nlat,
        nlon), in place."""
        arrs = (aaa, bbb, x0, y0, z0, p0, q)
        for a in arrs:
            assert a.dtype == np.float64 and a.flags.c_contiguous
        self._check(self._L.rb_transport(self._ctx, C.byref(grid),
                                         *[a.ctypes.data for a in arrs[:6]], int(q.shape[0]),
                                         q.ctypes.data))

    def run_day_device(self, grid: Grid, nsteps: int, day0: float, dt: float, trajlon, trajlat,
                       T, p, alt, sza, qj1, hc, surfarea, vsed3d, parth2osed, parthno3sed,
                       use_graph: bool = True, converged: bool = False, tol: float = 1e-10,
                       maxiter: int = 2000, iters=None, accel: int = 0) -> None:
        """nsteps sub-steps (device zenith + chemistry, conserv fused), optionally as one
        captured HIP graph."""
        field, nbox, _ = _extents(grid)
        dc = DayConfig(int(nsteps), 1 if use_graph else 0, 1 if converged else 0, int(maxiter),
                       float(day0), float(dt), float(tol), int(accel), 0)
        self._check(self._L.rb_run_day_dev(
            self._ctx, C.byref(grid), C.byref(dc),
            *[_ptr(a, n, w) for a, n, w in zip(
                (trajlon, trajlat, T, p, alt, sza, qj1, hc, surfarea, vsed3d, parth2osed,
                 parthno3sed),
                (nbox,) * 6 + (43 * field, 15 * field, field, nbox, nbox, nbox),
                ("trajlon", "trajlat", "T", "p", "alt", "sza", "qj1", "hc", "surfarea",
                 "vsed3d", "parth2osed", "parthno3sed"))],
            _iptr(iters, nbox, "iters") if iters is not None else None))

    def conserv_device(self, grid: Grid, qj1, hc) -> None:
        field = _extents(grid)[0]
        self._check(self._L.rb_conserv_floor_dev(self._ctx, C.byref(grid),
                                                 _ptr(qj1, 43 * field, "qj1"),
                                                 _ptr(hc, 15 * field, "hc")))

    # -- convenience on synth.StepInputs -------------------------------------------
    def run_inputs(self, inp, conserv: bool = True) -> dict:
        """Run one step on a synth.StepInputs via the host-array path; returns outputs."""
        g = inp.grid
        grid = make_grid(g.nlon, g.nlat, g.niv, g.nivbas, g.nivhau)
        c = lambda a: np.ascontiguousarray(a, dtype=np.float64).copy()
        q, h = c(inp.qj1), c(inp.hc)
        surf = np.zeros(g.shape3)
        shb = (g.nivbas, g.nlat, g.nlon)
        vsed, p2o, pno3 = np.zeros(shb), np.zeros(shb), np.zeros(shb)
        self.chem3d(grid, inp.dt, c(inp.T), c(inp.p), c(inp.sza), c(inp.alt), q, h, surf, vsed,
                    p2o, pno3)
        out = {"qj1_chem": q.copy(), "hc_chem": h.copy(), "surfarea": surf, "vsed3d": vsed,
               "parth2osed": p2o, "parthno3sed": pno3}
        if conserv:
            self.conserv(grid, q, h)
            out["qj1"], out["hc"] = q, h
        return out
