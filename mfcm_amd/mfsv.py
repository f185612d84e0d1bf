"""Python front end of libmfsv_hip.so (ctypes over the C ABI in include/mfsv.h).

This is the host-side mirror of the reference's stepping interfaces for the
Saint-Venant hot path:

* ``Catchment`` owns one device context (SaintVenantSolver + its Fields).
* ``Catchment.do_timestep`` / ``accept_timestep`` / ``step`` / ``solve`` follow
  RungeKuttaTemporalScheme::do_timestep (TemporalScheme/RungeKutta.cpp:143-167)
  and TypedTemporalScheme::step/solve (TemporalScheme/TemporalScheme.hpp:112-149,
  244-276) with the controller on the host, one synchronisation per timestep.
* ``Catchment.run`` is the same loop with the controller on the device and
  the timesteps replayed from a hipGraph.

The library must be built (``python -m mfcm_amd.build``); there is no CPU
fallback: every call goes to the HIP library or raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmfsv_hip.so")

F64, F32 = 0, 1
E1_RAIN_GUARD, E2_WEST_OUTFLOW, E4_STORM, E7_Y_ORIENT = 0x01, 0x02, 0x08, 0x40
E8_DRY_VELOCITY, E9_VELOCITY_CAP = 0x80, 0x100
FIELD = {"h": 0, "u": 1, "v": 2, "zb": 3, "manning_n": 4, "qbdy0": 5, "qbdy1": 6,
         "hbdy0": 7, "hbdy1": 8, "dhdt": 9, "dudt": 10, "dvdt": 11,
         "branch_x": 12, "branch_y": 13, "eloss_fx": 14, "eloss_fy": 15,
         "infil_rate": 16, "infil_cap": 17, "sbdy0": 18, "sbdy1": 19, "eddy_mu": 20,
         "n_deep": 21, "d_shallow": 22, "d_deep": 23,
         "flux_h": 24, "flux_u": 25, "flux_v": 26, "flux_z": 27,
         "mannings_n": 28, "friction_slope": 29, "dzdx_bed": 30, "dzdy_bed": 31,
         "stage": 32}
FACE_FIELDS = ("flux_h", "flux_u", "flux_v", "flux_z")
DIAG_FLUXES, DIAG_MANNING, DIAG_BED_SLOPES = 1, 2, 4

ERRORS = {-1: "bad argument", -2: "HIP runtime error", -3: "invalid state",
          -4: "out of device memory", -5: "unsupported"}

# Butcher tables of TemporalScheme/RungeKutta.cpp:54-116 (rows as the a_ vectors).
TABLEAUX = {
    "euler": [[0.0], [0.0, 1.0]],
    "midpoint": [[0.0], [0.5, 0.5], [0.0, 0.0, 1.0]],
    "heun2": [[0.0], [1.0, 1.0], [0.0, 0.5, 0.5]],
    "ralston2": [[0.0], [2.0 / 3.0, 2.0 / 3.0], [0.0, 0.25, 0.75]],
    "kutta3": [[0.0], [0.5, 0.5], [1.0, -1.0, 2.0], [0.0, 1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0]],
    "heun3": [[0.0], [1.0 / 3.0, 1.0 / 3.0], [2.0 / 3.0, 0.0, 2.0 / 3.0], [0.0, 0.25, 0.0, 0.75]],
    "wray3": [[0.0], [8.0 / 15.0, 8.0 / 15.0], [2.0 / 3.0, 0.25, 5.0 / 12.0], [0.0, 0.25, 0.0, 0.75]],
    "ralston3": [[0.0], [0.5, 0.5], [0.75, 0.0, 0.75], [0.0, 2.0 / 9.0, 1.0 / 3.0, 4.0 / 9.0]],
    "ssprk3": [[0.0], [1.0, 1.0], [0.5, 0.25, 0.25], [0.0, 1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0]],
    "classic": [[0.0], [0.5, 0.5], [0.5, 0.0, 0.5], [1.0, 0.0, 0.0, 1.0],
                [0.0, 1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0]],
    "ralston4": [[0.0], [0.4, 0.4], [0.45573725, 0.29697761, 0.15875964],
                 [1.0, 0.21810040, -3.05096516, 3.83286476],
                 [0.0, 0.17476028, -0.55148066, 1.20553560, 0.17118478]],
    "3/8rule": [[0.0], [1.0 / 3.0, 1.0 / 3.0], [2.0 / 3.0, -1.0 / 3.0, 1.0],
                [1.0, 1.0, -1.0, 1.0], [0.0, 1.0 / 8.0, 3.0 / 8.0, 3.0 / 8.0, 1.0 / 8.0]],
}


class MfsvError(RuntimeError):
    pass


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int64), ("ny", C.c_int64), ("dx", C.c_double), ("dy", C.c_double),
                ("origin_x", C.c_double), ("origin_y", C.c_double)]


class Params(C.Structure):
    _fields_ = [("flags", C.c_int32), ("cn_target", C.c_double), ("manning", C.c_int32),
                ("n_shallow", C.c_double), ("n_deep", C.c_double), ("d_shallow", C.c_double),
                ("d_deep", C.c_double), ("discharge", C.c_int32), ("head", C.c_int32),
                ("head_x0", C.c_int64), ("head_x1", C.c_int64),
                ("storm_peak", C.c_double), ("storm_sigma", C.c_double),
                ("storm_cx0", C.c_double), ("storm_cy0", C.c_double),
                ("storm_vx", C.c_double), ("storm_vy", C.c_double),
                ("h_dry", C.c_double), ("u_max", C.c_double),
                ("energy_loss", C.c_int32), ("infiltration", C.c_int32),
                ("stage", C.c_int32), ("eddy", C.c_int32),
                ("mu_const", C.c_double), ("mu_smag", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("time", C.c_double), ("t_local", C.c_double), ("dt", C.c_double),
                ("cn", C.c_double), ("volume", C.c_double), ("source_volume", C.c_double),
                ("volume0", C.c_double), ("step", C.c_int64), ("nsteps", C.c_int64),
                ("accepted", C.c_int64), ("rejected", C.c_int64), ("done", C.c_int32),
                ("current", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# Every entry point of include/mfsv.h with its ctypes signature.
_P, _I, _D, _I64 = C.c_void_p, C.c_int, C.c_double, C.c_int64
SIGNATURES = {
    "mfsv_version": (C.c_char_p, []),
    "mfsv_device_count": (_I, [C.POINTER(C.c_int)]),
    "mfsv_create": (_I, [C.POINTER(Grid), _I, _I, C.POINTER(_P)]),
    "mfsv_destroy": (_I, [_P]),
    "mfsv_last_error": (C.c_char_p, [_P]),
    "mfsv_synchronize": (_I, [_P]),
    "mfsv_device_bytes": (_I, [_P, C.POINTER(C.c_int64)]),
    "mfsv_set_params": (_I, [_P, C.POINTER(Params)]),
    "mfsv_set_scheme": (_I, [_P, _I, _P]),
    "mfsv_set_time": (_I, [_P, _D, _D, _D, _D]),
    "mfsv_set_tile_rows": (_I, [_P, _I]),
    "mfsv_get_tile_rows": (_I, [_P, C.POINTER(C.c_int)]),
    "mfsv_start_new_step": (_I, [_P, _D, _D, _D, _D, _D]),
    "mfsv_set_forcing_table": (_I, [_P, _I64, _P]),
    "mfsv_upload": (_I, [_P, _I, _P, _I64]),
    "mfsv_download": (_I, [_P, _I, _P, _I64]),
    "mfsv_dudt": (_I, [_P, _D, _D, _I]),
    "mfsv_diagnose": (_I, [_P, _I, _D, _D, _I]),
    "mfsv_update_stage_field": (_I, [_P]),
    "mfsv_control_number": (_I, [_P, _I, _D, C.POINTER(C.c_double)]),
    "mfsv_control_number_noalloc": (_I, [_I, _I, _I64, _I64, _D, _D, _P, _P, _P, _D, _P, _P]),
    "mfsv_spatial_derivative": (_I, [_P, _I, _P, _P, _I64]),
    "mfsv_timestep": (_I, [_P, _D, _D, C.POINTER(C.c_double)]),
    "mfsv_accept": (_I, [_P]),
    "mfsv_volume": (_I, [_P, C.POINTER(C.c_double)]),
    "mfsv_run_begin": (_I, [_P]),
    "mfsv_run": (_I, [_P, _I64, _I]),
    "mfsv_get_stats": (_I, [_P, C.POINTER(Stats)]),
    "mfsv_run_step": (_I, [_P, _I]),
    "mfsv_last_run_ms": (_I, [_P, C.POINTER(C.c_double)]),
    "mfsv_last_stage_ms": (_I, [_P, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "mfsv_batch_create": (_I, [_I, _I, C.POINTER(_P)]),
    "mfsv_batch_destroy": (_I, [_P]),
    "mfsv_batch_last_error": (C.c_char_p, [_P]),
    "mfsv_batch_add": (_I, [_P, _P]),
    "mfsv_batch_remove": (_I, [_P, _P]),
    "mfsv_batch_size": (_I, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "mfsv_batch_run": (_I, [_P, _I64, _I]),
    "mfsv_batch_last_run_ms": (_I, [_P, C.POINTER(C.c_double)]),
}

_lib = None


def lib():
    """Load libmfsv_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MfsvError(f"{LIB_PATH} is missing: run `python -m mfcm_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def device_count() -> int:
    n = C.c_int(0)
    rc = lib().mfsv_device_count(C.byref(n))
    return n.value if rc == 0 else 0


def ts_linterp(t, v, time):
    """Input/TimeSeries.cpp:27-46 (FP64), NaN past the last sample."""
    i1 = 1
    while i1 < len(t) and not (t[i1] > time):
        i1 += 1
    if i1 >= len(t):
        return float("nan")
    i0 = i1 - 1
    dvdt = (v[i1] - v[i0]) / (t[i1] - t[i0])
    return v[i0] + dvdt * (time - t[i0])


def num_steps(start, end, step):
    """Config/TimeParameters.cpp:40-58."""
    eps = 1e-6
    dur = end - start
    n = int(dur / step)
    if step * n >= dur - eps and step * n <= dur + eps:
        return n
    return n + 1


def tableau_matrix(name):
    rows = TABLEAUX[name] if isinstance(name, str) else name
    n = len(rows) - 1
    m = np.zeros((n + 1, n + 1))
    for i, r in enumerate(rows):
        m[i, :len(r)] = r
    return n, m


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class Catchment:
    """One catchment on the device: fields, parameters, the stepping loop."""

    def __init__(self, scenario, precision=F64, device=0, rows=None):
        sc = scenario
        self.sc = sc
        self.nx, self.ny = int(sc.nx), int(sc.ny)
        self.nc = self.nx * self.ny
        self.precision = precision
        g = Grid(self.nx, self.ny, float(sc.dx), float(sc.dy), 0.0, 0.0)
        h = C.c_void_p()
        self._check(lib().mfsv_create(C.byref(g), precision, device, C.byref(h)), None)
        self._h = h
        nst, tab = tableau_matrix(sc.scheme)
        self.nstages = nst
        self._tab = np.ascontiguousarray(tab)
        self._check(lib().mfsv_set_scheme(self._h, nst, _ptr(self._tab)))
        p = Params()
        p.flags = int(sc.flags)
        p.cn_target = float(sc.cn_target)
        if getattr(sc, "manning_fields", None) is not None:
            # per-cell n_shallow, n_deep, d_shallow, d_deep
            p.manning = 1
            ns, nd, ds, dd = (np.broadcast_to(np.asarray(a, np.float64), (self.ny, self.nx))
                              for a in sc.manning_fields)
            self.upload("manning_n", ns)
            self.upload("n_deep", nd)
            self.upload("d_shallow", ds)
            self.upload("d_deep", dd)
            p.n_shallow, p.n_deep = float(np.mean(ns)), float(np.mean(nd))
            p.d_shallow, p.d_deep = float(np.mean(ds)), float(np.mean(dd))
        elif sc.n_manning is not None:
            p.manning = 1
            if isinstance(sc.n_manning, np.ndarray):
                self.upload("manning_n", sc.n_manning)
                p.n_shallow = p.n_deep = float(np.mean(sc.n_manning))
            else:
                p.n_shallow = p.n_deep = float(sc.n_manning)
            p.d_shallow, p.d_deep = float(sc.d_shallow), float(sc.d_deep)
        p.discharge = 1 if sc.rain_t is not None else 0
        if sc.head_cols:
            p.head, p.head_x0, p.head_x1 = 1, 0, int(sc.head_cols)
        if sc.storm is not None:
            st = sc.storm
            p.storm_peak, p.storm_sigma = st.peak, st.sigma
            p.storm_cx0, p.storm_cy0, p.storm_vx, p.storm_vy = st.cx0, st.cy0, st.vx, st.vy
        p.h_dry = float(getattr(sc, "h_dry", 1e-3))
        p.u_max = float(getattr(sc, "u_max", 10.0))
        cells = lambda a: np.broadcast_to(np.asarray(a, np.float64), (self.ny, self.nx))
        if getattr(sc, "energy_loss", None) is not None:
            p.energy_loss = 1
            self.upload("eloss_fx", cells(sc.energy_loss[0]))
            self.upload("eloss_fy", cells(sc.energy_loss[1]))
        if getattr(sc, "infiltration", None) is not None:
            p.infiltration = 1
            self.upload("infil_rate", cells(sc.infiltration[0]))
            self.upload("infil_cap", cells(sc.infiltration[1]))
        if getattr(sc, "eddy", None) is not None:
            if "mu" in sc.eddy:
                p.eddy = 1
                self.upload("eddy_mu", cells(sc.eddy["mu"]))
            else:
                p.eddy = 2
                p.mu_const, p.mu_smag = (float(x) for x in sc.eddy["smagorinsky"])
        if getattr(sc, "stage_t", None) is not None:
            p.stage = 1
        self.params = p
        self._check(lib().mfsv_set_params(self._h, C.byref(p)))
        self._check(lib().mfsv_set_time(self._h, 0.0, float(sc.t_end), float(sc.step_duration),
                                         float(sc.initial_dt)))
        self.upload("zb", sc.zb)
        self.set_state(sc.h0, sc.u0, sc.v0)
        if p.stage:
            self.set_stage_levels(0.0)
        self.nsteps = num_steps(0.0, float(sc.t_end), float(sc.step_duration))
        self._table = self.forcing_table()
        self._check(lib().mfsv_set_forcing_table(self._h, self.nsteps, _ptr(self._table)))
        self.dt = float(sc.initial_dt)
        self.n_acc = 0
        self.n_rej = 0

    # -- plumbing ---------------------------------------------------------
    def _check(self, rc, h="self"):
        if rc != 0:
            msg = ""
            if h is not None and getattr(self, "_h", None):
                msg = lib().mfsv_last_error(self._h).decode()
            raise MfsvError(f"mfsv error {rc} ({ERRORS.get(rc, '?')}): {msg}")

    def close(self):
        if getattr(self, "_h", None):
            lib().mfsv_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        self._check(lib().mfsv_upload(self._h, FIELD[name], _ptr(a), a.size))

    def download(self, name):
        if name == "branch_x":
            n, shp = (self.nx + 1) * self.ny, (self.ny, self.nx + 1)
        elif name == "branch_y":
            n, shp = self.nx * (self.ny + 1), (self.ny + 1, self.nx)
        elif name in FACE_FIELDS:  # x-faces then y-faces, flat
            n = (self.nx + 1) * self.ny + self.nx * (self.ny + 1)
            shp = (n,)
        else:
            n, shp = self.nc, (self.ny, self.nx)
        a = np.empty(n, dtype=np.float64)
        self._check(lib().mfsv_download(self._h, FIELD[name], _ptr(a), n))
        return a.reshape(shp)

    def set_state(self, h, u=None, v=None):
        z = np.zeros((self.ny, self.nx))
        self.upload("h", h)
        self.upload("u", z if u is None else u)
        self.upload("v", z if v is None else v)

    def state(self):
        return self.download("h"), self.download("u"), self.download("v")

    def volume(self):
        v = C.c_double()
        self._check(lib().mfsv_volume(self._h, C.byref(v)))
        return v.value

    def set_tile_rows(self, rows):
        self._check(lib().mfsv_set_tile_rows(self._h, int(rows)))

    def tile_rows(self):
        r = C.c_int()
        self._check(lib().mfsv_get_tile_rows(self._h, C.byref(r)))
        return r.value

    def device_bytes(self):
        b = C.c_int64()
        self._check(lib().mfsv_device_bytes(self._h, C.byref(b)))
        return b.value

    def set_stage_levels(self, time_now):
        """StageBoundarySourceTerm.hpp:114-142: water-level endpoints of the
        outer step starting at time_now on the stage cells (NaN elsewhere)."""
        sc = self.sc
        sel = np.asarray(sc.stage_sel, bool) if sc.stage_sel is not None else np.ones(
            (self.ny, self.nx), bool)
        a0 = ts_linterp(sc.stage_t, sc.stage_v, time_now)
        a1 = ts_linterp(sc.stage_t, sc.stage_v, time_now + sc.step_duration)
        self.upload("sbdy0", np.where(sel, a0, np.nan))
        self.upload("sbdy1", np.where(sel, a1, np.nan))

    # -- forcing (Solver.cpp:119-129 start_new_step) ----------------------
    def forcing_at(self, time_now):
        sc = self.sc
        q0 = q1 = 0.0
        hb0 = hb1 = float("nan")
        if sc.rain_t is not None:
            q0 = 0.0 + ts_linterp(sc.rain_t, sc.rain_v, time_now)
            q1 = 0.0 + ts_linterp(sc.rain_t, sc.rain_v, time_now + sc.step_duration)
        if sc.head_cols:
            hb0 = hb1 = 0.0
        return q0, q1, hb0, hb1

    def forcing_table(self):
        t = np.zeros((max(self.nsteps, 1), 4))
        for i in range(self.nsteps):
            t[i] = self.forcing_at(0.0 + i * float(self.sc.step_duration))
        return np.ascontiguousarray(t)

    # -- reference-shaped host-driven loop --------------------------------
    def start_new_step(self, time_now):
        if self.params.stage:
            self.set_stage_levels(time_now)
        q0, q1, hb0, hb1 = self.forcing_at(time_now)
        self._check(lib().mfsv_start_new_step(self._h, time_now, q0, q1, hb0, hb1))

    def dudt(self, t_sub, dt, branches=False):
        self._check(lib().mfsv_dudt(self._h, t_sub, dt, 1 if branches else 0))
        return self.download("dhdt"), self.download("dudt"), self.download("dvdt")

    def diagnose(self, source=0, t_sub=0.0, dt=1.0, what=7):
        """Output-only fields (face fluxes, Manning n and friction slope, bed
        slopes) of U0 (source 0) or of the last RK stage's input (source 1)."""
        self._check(lib().mfsv_diagnose(self._h, source, t_sub, dt, what))
        out = {}
        if what & DIAG_FLUXES:
            for f in FACE_FIELDS:
                out[f] = self.download(f)
        if what & DIAG_MANNING and self.params.manning:
            out["mannings_n"] = self.download("mannings_n")
            out["friction_slope"] = self.download("friction_slope")
        if what & DIAG_BED_SLOPES:
            out["dzdx_bed"] = self.download("dzdx_bed")
            out["dzdy_bed"] = self.download("dzdy_bed")
        return out

    def control_number(self, state, dt):
        """State.cpp:199-229 on state 0 (U0), 1..n-1 (RK rows) or n (result)."""
        cn = C.c_double()
        self._check(lib().mfsv_control_number(self._h, state, dt, C.byref(cn)))
        return cn.value

    def spatial_derivative(self, field, axis):
        """Minmod3 slope of an FP64 cell field (axis 0: x; 1: the reference's y)."""
        a = np.ascontiguousarray(field, dtype=np.float64).ravel()
        out = np.empty_like(a)
        self._check(lib().mfsv_spatial_derivative(self._h, axis, _ptr(a), _ptr(out), a.size))
        return out.reshape(self.ny, self.nx)

    def stage_field(self):
        """z_bed + h of U0 (the solver's stage field)."""
        self._check(lib().mfsv_update_stage_field(self._h))
        return self.download("stage")

    def control_timestep(self, t_local, dt):
        cn = C.c_double()
        self._check(lib().mfsv_timestep(self._h, t_local, dt, C.byref(cn)))
        return cn.value

    def do_timestep(self, t_local, dt):
        """RungeKutta.cpp:143-167 (controller on the host)."""
        cn_raw = self.control_timestep(t_local, dt)
        tgt = float(self.sc.cn_target)
        cn = cn_raw if tgt == 1.0 else cn_raw / tgt
        if cn > 1.0:
            return (dt / 5.0 if cn > 5.0 else dt / (cn * 1.1)), True
        if cn > 0.9:
            return dt, False
        return (dt / cn if cn != 0.0 else float("inf")), False

    def accept_timestep(self):
        self._check(lib().mfsv_accept(self._h))

    def step(self):
        """TemporalScheme.hpp:112-149."""
        t_local = 0.0
        sd = float(self.sc.step_duration)
        while t_local < sd:
            new_dt, repeat = self.do_timestep(t_local, self.dt)
            if repeat:
                self.n_rej += 1
            else:
                self.n_acc += 1
                t_local += self.dt
                self.accept_timestep()
            self.dt = new_dt
            if t_local >= sd:
                return
            elif t_local + self.dt > sd:
                self.dt = sd - t_local
            elif t_local + 2.0 * self.dt > sd:
                self.dt = (sd - t_local) * 0.4

    def run_steps(self, first, count):
        for i in range(first, first + count):
            self.start_new_step(0.0 + i * float(self.sc.step_duration))
            self.step()

    def solve(self):
        """TemporalScheme.hpp:244-276 with the host controller."""
        self.dt = float(self.sc.initial_dt)
        self.run_steps(0, self.nsteps)

    # -- device-controlled run ---------------------------------------------
    def run_begin(self):
        self._check(lib().mfsv_run_begin(self._h))

    def run(self, max_timesteps=1 << 40, graph=True):
        """graph: True/1 hipGraph replay, False/0 eager, 2 eager with
        per-timestep HIP events around the stage kernels."""
        mode = int(graph) if not isinstance(graph, bool) else (1 if graph else 0)
        self._check(lib().mfsv_run(self._h, int(max_timesteps), mode))
        return self.stats()

    def advance(self, max_timesteps, graph=True):
        """run() without the statistics read back afterwards (one host wait
        per call: the end of the work)."""
        mode = int(graph) if not isinstance(graph, bool) else (1 if graph else 0)
        self._check(lib().mfsv_run(self._h, int(max_timesteps), mode))

    def run_step(self, graph=True):
        """The current outer step's inner loop on the device."""
        mode = int(graph) if not isinstance(graph, bool) else (1 if graph else 0)
        self._check(lib().mfsv_run_step(self._h, mode))
        return self.stats()

    def last_stage_ms(self):
        m, n = C.c_double(), C.c_int64()
        self._check(lib().mfsv_last_stage_ms(self._h, C.byref(m), C.byref(n)))
        return m.value, n.value

    def stats(self):
        s = Stats()
        self._check(lib().mfsv_get_stats(self._h, C.byref(s)))
        return s.as_dict()

    def last_run_ms(self):
        m = C.c_double()
        self._check(lib().mfsv_last_run_ms(self._h, C.byref(m)))
        return m.value
