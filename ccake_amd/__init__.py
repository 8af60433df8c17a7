"""MI355X-native per-timestep SPH path of CCAKE.

Python front-end over the C-ABI of `libccake_amd.so` (include/ccake_amd.h).
The HIP library is the only compute path: importing works on any machine,
but every call that computes requires the built library and a HIP device,
and raises otherwise.
"""
from __future__ import annotations

import ctypes as C
import os
import re
import weakref

import numpy as np

from . import config  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(_HERE)
# The gfx950 library.  CCAKE_AMD_LIBRARY names another build of the same
# sources; the tests use it for the host emulation build (tests/hostemu),
# which runs the kernels on the CPU under AddressSanitizer to check memory
# accesses.
LIB_PATH = os.environ.get("CCAKE_AMD_LIBRARY") or os.path.join(_HERE, "libccake_amd.so")
HEADER = os.path.join(ROOT, "include", "ccake_amd.h")


class CcakeError(RuntimeError):
    def __init__(self, status, what):
        self.status = status
        super().__init__(f"{what}: status {status} ({STATUS.get(status, '?')}): {last_error()}")


def _parse_enum(name):
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    body = text[text.index("enum " + name):]
    body = body[body.index("{") + 1: body.index("}")]
    out, cur = {}, -1
    for item in body.split(","):
        item = item.strip()
        if not item:
            continue
        if "=" in item:
            nm, expr = [x.strip() for x in item.split("=", 1)]
            for k, v in sorted(out.items(), key=lambda kv: -len(kv[0])):
                expr = re.sub(r"\b%s\b" % k, str(v), expr)
            cur = int(eval(expr))
        else:
            nm = item
            cur += 1
        out[nm] = cur
    return out


FIELDS = _parse_enum("ccake_field")
IFIELDS = _parse_enum("ccake_ifield")
KERNELS = _parse_enum("ccake_kernel_id")
REDUCTIONS = _parse_enum("ccake_reduction")
STATUS = {v: k for k, v in _parse_enum("ccake_status").items()}
FO_ROW = 42       # CCAKE_FO_ROW: id, step, then the 40 freeze_out.dat columns

PARAM_FIELDS = [
    ("dim", C.c_int), ("rk_order", C.c_int),
    ("hT", C.c_double), ("hEta", C.c_double),
    ("t0", C.c_double), ("dt", C.c_double), ("e_freeze", C.c_double),
    ("particlization", C.c_int), ("using_shear", C.c_int), ("using_vorticity", C.c_int),
    ("input_initial_shear", C.c_int), ("bulk_from_trace", C.c_int),
    ("prohibit_unstable_cs2", C.c_int), ("prohibit_acausal_cs2", C.c_int),
    ("eta_mode", C.c_int), ("constant_eta_over_s", C.c_double),
    ("shear_relax_mode", C.c_int),
    ("zeta_mode", C.c_int), ("constant_zeta_over_s", C.c_double),
    ("cs2_dependent_zeta_A", C.c_double), ("cs2_dependent_zeta_p", C.c_double),
    ("modulate_zeta_with_tanh", C.c_int),
    ("zeta_gauss_A", C.c_double), ("zeta_gauss_T0", C.c_double), ("zeta_gauss_width", C.c_double),
    ("delta_pipi_mode", C.c_int), ("tau_pipi_mode", C.c_int), ("lambda_piPi_mode", C.c_int),
    ("phi6_mode", C.c_int), ("phi7_mode", C.c_int),
    ("delta_PiPi_mode", C.c_int), ("lambda_Pipi_mode", C.c_int), ("phi1_mode", C.c_int),
    ("phi3_mode", C.c_int),
    ("cells_per_rc", C.c_int),
    ("using_diffusion", C.c_int), ("diffusion_mode", C.c_int), ("diffusion_relax_mode", C.c_int),
    ("kappa_over_T2", C.c_double * 9), ("C_B", C.c_double),
    ("diffusion_regulator", C.c_int), ("regulator_chi0", C.c_double), ("regulator_e0", C.c_double),
    ("regulator_xi0", C.c_double), ("regulator_rq_max", C.c_double),
    ("pair_precision", C.c_int), ("factorized_kernel", C.c_int), ("x_subcells", C.c_int),
]


class Params(C.Structure):
    _fields_ = PARAM_FIELDS


class EosDesc(C.Structure):
    _fields_ = [("kind", C.c_int), ("T_freeze_out", C.c_double), ("dims", C.c_int * 4),
                ("mins", C.c_double * 4), ("maxs", C.c_double * 4),
                ("table", C.POINTER(C.c_double))]


# exported C-ABI: name -> (restype, argtypes)
_vp, _dp, _ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32)
ABI = {
    "ccake_dev_create": (C.c_int, [C.c_int, C.c_size_t, C.POINTER(_vp)]),
    "ccake_dev_destroy": (C.c_int, [_vp]),
    "ccake_dev_sync": (C.c_int, [_vp]),
    "ccake_dev_memory": (C.c_int, [_vp, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "ccake_last_error": (C.c_char_p, []),
    "ccake_eos_create": (C.c_int, [_vp, C.POINTER(EosDesc), C.POINTER(_vp)]),
    "ccake_eos_destroy": (C.c_int, [_vp]),
    "ccake_eos_efreeze": (C.c_int, [_vp, C.c_double, _dp]),
    "ccake_eos_full_thermo": (C.c_int, [_vp, C.c_int, _dp, _dp]),
    "ccake_event_create": (C.c_int, [_vp, C.POINTER(Params), _vp, C.c_int, C.POINTER(_vp)]),
    "ccake_event_destroy": (C.c_int, [_vp]),
    "ccake_event_size": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "ccake_event_set_time": (C.c_int, [_vp, C.c_double]),
    "ccake_event_get_time": (C.c_int, [_vp, _dp]),
    "ccake_event_upload": (C.c_int, [_vp, C.c_int, _dp, C.c_size_t]),
    "ccake_event_download": (C.c_int, [_vp, C.c_int, _dp, C.c_size_t]),
    "ccake_event_upload_int": (C.c_int, [_vp, C.c_int, _ip, C.c_size_t]),
    "ccake_event_download_int": (C.c_int, [_vp, C.c_int, _ip, C.c_size_t]),
    "ccake_build_neighbors": (C.c_int, [_vp]),
    "ccake_export_pairs": (C.c_int, [_vp, _ip, _ip, C.c_size_t, C.POINTER(C.c_size_t)]),
    "ccake_export_pairs_noalloc": (C.c_int, [_vp, _ip, _ip, C.c_size_t, C.POINTER(C.c_size_t)]),
    "ccake_gamma_velocity": (C.c_int, [_vp]),
    "ccake_smooth_fields": (C.c_int, [_vp]),
    "ccake_initial_sigma": (C.c_int, [_vp]),
    "ccake_eos_invert": (C.c_int, [_vp, C.c_int]),
    "ccake_reset_pi_tensor": (C.c_int, [_vp]),
    "ccake_viscosities": (C.c_int, [_vp]),
    "ccake_smooth_gradients": (C.c_int, [_vp]),
    "ccake_eom": (C.c_int, [_vp]),
    "ccake_get_time_derivatives": (C.c_int, [_vp]),
    "ccake_cache_state": (C.c_int, [_vp]),
    "ccake_rk_update": (C.c_int, [_vp, C.c_double, C.c_double]),
    "ccake_advance_timestep": (C.c_int, [_vp, C.c_double, C.c_int]),
    "ccake_run_steps": (C.c_int, [_vp, C.c_double, C.c_int, C.c_int, C.POINTER(C.c_int)]),
    "ccake_initialize": (C.c_int, [_vp]),
    "ccake_allocate_k_values": (C.c_int, [_vp]),
    "ccake_update_k": (C.c_int, [_vp, C.c_int]),
    "ccake_update_rk4": (C.c_int, [_vp, C.c_double]),
    "ccake_process_initial_conditions": (C.c_int, [_vp]),
    "ccake_initialize_entropy_and_charge_densities": (C.c_int, [_vp]),
    "ccake_initial_smoothing": (C.c_int, [_vp]),
    "ccake_freeze_out_particles": (C.c_int, [_vp]),
    "ccake_continue_evolution": (C.c_int, [_vp, C.c_double, C.POINTER(C.c_int)]),
    "ccake_batch_create": (C.c_int, [_vp, C.POINTER(Params), _vp, C.c_int, _ip, C.POINTER(_vp)]),
    "ccake_batch_events": (C.c_int, [_vp, C.POINTER(C.c_int), _ip]),
    "ccake_batch_reduce": (C.c_int, [_vp, C.c_int, _dp]),
    "ccake_batch_continue_evolution": (C.c_int, [_vp, C.c_double, _ip]),
    "ccake_freeze_out_surface": (C.c_int, [_vp, _dp, C.c_size_t, C.POINTER(C.c_size_t)]),
    "ccake_event_check": (C.c_int, [_vp]),
    "ccake_reduce": (C.c_int, [_vp, C.c_int, _dp]),
    "ccake_profile_enable": (C.c_int, [_vp, C.c_int]),
    "ccake_profile_read": (C.c_int, [_vp, C.c_int, _dp, C.POINTER(C.c_int64)]),
    "ccake_profile_reset": (C.c_int, [_vp]),
    "ccake_last_pair_count": (C.c_int, [_vp, C.POINTER(C.c_int64)]),
    "ccake_pair_total": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int64)]),
    "ccake_event_pair_slices": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int)]),
}

_lib = None


def lib():
    """Load the HIP library.  There is no fallback: a missing library raises."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in ABI.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error():
    try:
        return lib().ccake_last_error().decode()
    except Exception:  # pragma: no cover - only during failed loads
        return ""


def _ok(st, what):
    if st != 0:
        raise CcakeError(st, what)


OPTIONAL_PARAMS = dict(pair_precision=0, factorized_kernel=0, x_subcells=8)


def make_params(cfg):
    p = Params()
    for name, typ in PARAM_FIELDS:
        v = cfg.get(name, OPTIONAL_PARAMS[name]) if name in OPTIONAL_PARAMS else cfg[name]
        setattr(p, name, typ(*v) if hasattr(typ, "_length_") else v)
    return p


class Device:
    def __init__(self, gpu=0, hbm_budget=0):
        h = C.c_void_p()
        _ok(lib().ccake_dev_create(gpu, hbm_budget, C.byref(h)), "ccake_dev_create")
        self.h = h
        self._children = weakref.WeakSet()   # Eos and Event handles on this device

    def close(self):
        """Destroys the device after any events and EoS objects still open on it."""
        if getattr(self, "h", None):
            kids = list(self._children)
            for c in kids:
                if isinstance(c, Event):
                    c.close()
            for c in kids:
                c.close()
            lib().ccake_dev_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def sync(self):
        _ok(lib().ccake_dev_sync(self.h), "ccake_dev_sync")

    def memory(self):
        u, b = C.c_size_t(), C.c_size_t()
        _ok(lib().ccake_dev_memory(self.h, C.byref(u), C.byref(b)), "ccake_dev_memory")
        return u.value, b.value

    def profile(self, enable=True):
        _ok(lib().ccake_profile_enable(self.h, int(enable)), "ccake_profile_enable")

    def profile_reset(self):
        _ok(lib().ccake_profile_reset(self.h), "ccake_profile_reset")

    def profile_read(self, kernel):
        k = KERNELS[kernel] if isinstance(kernel, str) else int(kernel)
        ms, n = C.c_double(), C.c_int64()
        _ok(lib().ccake_profile_read(self.h, k, C.byref(ms), C.byref(n)), "ccake_profile_read")
        return ms.value, n.value


class Eos:
    def __init__(self, dev, kind=0, T_freeze_out=150.0 / 197.3269804, dims=None, mins=None, maxs=None,
                 table=None):
        d = EosDesc()
        d.kind = kind
        d.T_freeze_out = T_freeze_out
        if table is not None:
            tab = np.ascontiguousarray(table, dtype=np.float64)
            for a in range(4):
                d.dims[a] = int(dims[a]); d.mins[a] = float(mins[a]); d.maxs[a] = float(maxs[a])
            d.table = tab.ctypes.data_as(C.POINTER(C.c_double))
        h = C.c_void_p()
        _ok(lib().ccake_eos_create(dev.h, C.byref(d), C.byref(h)), "ccake_eos_create")
        self.h = h
        self.dev = dev
        self._events = weakref.WeakSet()
        dev._children.add(self)

    def close(self):
        if getattr(self, "h", None):
            for e in list(self._events):
                e.close()
            lib().ccake_eos_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def efreeze(self, T):
        e = C.c_double()
        _ok(lib().ccake_eos_efreeze(self.h, T, C.byref(e)), "ccake_eos_efreeze")
        return e.value

    def full_thermo(self, member, tbqs):
        x = (C.c_double * 4)(*tbqs)
        out = (C.c_double * 17)()
        _ok(lib().ccake_eos_full_thermo(self.h, member, x, out), "ccake_eos_full_thermo")
        return np.array(out[:])


class Event:
    """One particle system (one SystemState) resident on the device."""

    def __init__(self, dev, eos, cfg, n, _counts=None):
        self.dev, self.eos, self.cfg = dev, eos, cfg
        self._p = make_params(cfg)
        h = C.c_void_p()
        if _counts is None:
            self.n = int(n)
            _ok(lib().ccake_event_create(dev.h, C.byref(self._p), eos.h, self.n, C.byref(h)), "ccake_event_create")
        else:
            cnt = np.ascontiguousarray(_counts, dtype=np.int32)
            self.n = int(cnt.sum())
            _ok(lib().ccake_batch_create(dev.h, C.byref(self._p), eos.h, len(cnt),
                                         cnt.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(h)), "ccake_batch_create")
        self.h = h
        dev._children.add(self)
        eos._events.add(self)

    def close(self):
        if getattr(self, "h", None):
            lib().ccake_event_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # ---- copies
    def upload(self, field, values):
        fid = FIELDS[field] if isinstance(field, str) else int(field)
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(values, dtype=np.float64), (self.n,)))
        _ok(lib().ccake_event_upload(self.h, fid, a.ctypes.data_as(_dp), self.n), "upload")

    def download(self, field):
        fid = FIELDS[field] if isinstance(field, str) else int(field)
        a = np.empty(self.n, dtype=np.float64)
        _ok(lib().ccake_event_download(self.h, fid, a.ctypes.data_as(_dp), self.n), "download")
        return a

    def upload_int(self, field, values):
        fid = IFIELDS[field] if isinstance(field, str) else int(field)
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(values, dtype=np.int32), (self.n,)))
        _ok(lib().ccake_event_upload_int(self.h, fid, a.ctypes.data_as(_ip), self.n), "upload_int")

    def download_int(self, field):
        fid = IFIELDS[field] if isinstance(field, str) else int(field)
        a = np.empty(self.n, dtype=np.int32)
        _ok(lib().ccake_event_download_int(self.h, fid, a.ctypes.data_as(_ip), self.n), "download_int")
        return a

    @property
    def t(self):
        v = C.c_double()
        _ok(lib().ccake_event_get_time(self.h, C.byref(v)), "get_time")
        return v.value

    @t.setter
    def t(self, v):
        _ok(lib().ccake_event_set_time(self.h, float(v)), "set_time")

    # ---- phases
    def _call(self, name, *args):
        _ok(getattr(lib(), "ccake_" + name)(self.h, *args), name)

    def build_neighbors(self): self._call("build_neighbors")
    def gamma_velocity(self): self._call("gamma_velocity")
    def smooth_fields(self): self._call("smooth_fields")
    def initial_sigma(self): self._call("initial_sigma")
    def eos_invert(self, mode): self._call("eos_invert", int(mode))
    def reset_pi_tensor(self): self._call("reset_pi_tensor")
    def viscosities(self): self._call("viscosities")
    def smooth_gradients(self): self._call("smooth_gradients")
    def eom(self): self._call("eom")
    def get_time_derivatives(self): self._call("get_time_derivatives")
    def cache_state(self): self._call("cache_state")
    def rk_update(self, dt, t0): self._call("rk_update", float(dt), float(t0))
    def advance_timestep(self, dt, rk_order=2): self._call("advance_timestep", float(dt), int(rk_order))
    def initialize(self): self._call("initialize")
    def process_initial_conditions(self): self._call("process_initial_conditions")
    def initialize_entropy_and_charge_densities(self): self._call("initialize_entropy_and_charge_densities")
    def initial_smoothing(self): self._call("initial_smoothing")
    def freeze_out_particles(self): self._call("freeze_out_particles")
    def check(self): self._call("event_check")

    def run_steps(self, dt, n_steps, rk_order=2):
        done = C.c_int()
        _ok(lib().ccake_run_steps(self.h, float(dt), int(rk_order), int(n_steps), C.byref(done)), "run_steps")
        return done.value

    def continue_evolution(self, max_tau=50.0):
        k = C.c_int()
        _ok(lib().ccake_continue_evolution(self.h, float(max_tau), C.byref(k)), "continue_evolution")
        return bool(k.value)

    def reduce(self, which):
        w = REDUCTIONS[which] if isinstance(which, str) else int(which)
        v = C.c_double()
        _ok(lib().ccake_reduce(self.h, w, C.byref(v)), "reduce")
        return v.value

    def pairs(self):
        n = C.c_size_t()
        st = lib().ccake_export_pairs(self.h, None, None, 0, C.byref(n))
        if st not in (0, STATUS_CAPACITY):
            raise CcakeError(st, "export_pairs")
        pi = np.empty(n.value, dtype=np.int32)
        pj = np.empty(n.value, dtype=np.int32)
        _ok(lib().ccake_export_pairs(self.h, pi.ctypes.data_as(_ip), pj.ctypes.data_as(_ip), n.value,
                                     C.byref(n)), "export_pairs")
        return pi, pj

    def pairs_noalloc(self, cap=None):
        """The pair set through ccake_export_pairs_noalloc (no allocation in the
        library; the caller's arrays hold `cap` pairs, by default the count
        of the last neighbour build)."""
        if cap is None:
            cap = self.last_pair_count()
        pi = np.empty(max(cap, 1), dtype=np.int32)
        pj = np.empty(max(cap, 1), dtype=np.int32)
        n = C.c_size_t()
        _ok(lib().ccake_export_pairs_noalloc(self.h, pi.ctypes.data_as(_ip), pj.ctypes.data_as(_ip), cap,
                                             C.byref(n)), "export_pairs_noalloc")
        return pi[:n.value], pj[:n.value]

    def freeze_out_surface(self):
        """The freeze-out hypersurface so far: one row per frozen-out particle,
        ordered by (step, particle id); columns: id, step, then the 40 columns
        of freeze_out.dat (divEener, divE[2], gsub, u[2], swsub, bulk, shear
        00/11/22/33/12, tau, r[2], s, e, T, muB, muS, muQ, w, cs2, rhoB, rhoS,
        rhoQ, diffusion[12], eos flag)."""
        n = C.c_size_t()
        _ok(lib().ccake_freeze_out_surface(self.h, None, 0, C.byref(n)), "freeze_out_surface")
        rows = np.zeros((n.value, FO_ROW))
        _ok(lib().ccake_freeze_out_surface(self.h, rows.ctypes.data_as(_dp), n.value, C.byref(n)),
            "freeze_out_surface")
        return rows[: n.value]

    def nslices(self, set_to=0):
        """Slices per tile of the pair kernels; set_to > 0 changes it (up to
        the count the event was created with)."""
        v = C.c_int()
        _ok(lib().ccake_event_pair_slices(self.h, int(set_to), C.byref(v)), "pair_slices")
        return v.value

    def pair_total(self, reset=False):
        v = C.c_int64()
        _ok(lib().ccake_pair_total(self.h, int(reset), C.byref(v)), "pair_total")
        return v.value

    def last_pair_count(self):
        v = C.c_int64()
        _ok(lib().ccake_last_pair_count(self.h, C.byref(v)), "last_pair_count")
        return v.value


STATUS_CAPACITY = 7


class Batch(Event):
    """Many independent events stepped together (ccake_batch_create): each
    kernel launches once per stage for all of them.  Event e owns the
    particle ids self.offsets[e] .. self.offsets[e + 1]; columns are the
    concatenation of the events' columns."""

    def __init__(self, dev, eos, cfg, counts):
        super().__init__(dev, eos, cfg, 0, _counts=counts)
        k = C.c_int()
        _ok(lib().ccake_batch_events(self.h, C.byref(k), None), "batch_events")
        off = np.zeros(k.value + 1, dtype=np.int32)
        _ok(lib().ccake_batch_events(self.h, C.byref(k), off.ctypes.data_as(C.POINTER(C.c_int32))), "batch_events")
        self.n_events = k.value
        self.offsets = off

    def event_slice(self, e):
        return slice(int(self.offsets[e]), int(self.offsets[e + 1]))

    def batch_reduce(self, which):
        w = REDUCTIONS[which] if isinstance(which, str) else int(which)
        out = np.zeros(self.n_events)
        _ok(lib().ccake_batch_reduce(self.h, w, out.ctypes.data_as(C.POINTER(C.c_double))), "batch_reduce")
        return out

    def batch_continue_evolution(self, max_tau=50.0):
        out = np.zeros(self.n_events, dtype=np.int32)
        _ok(lib().ccake_batch_continue_evolution(self.h, float(max_tau), out.ctypes.data_as(C.POINTER(C.c_int32))),
            "batch_continue_evolution")
        return out.astype(bool)

    def surface_by_event(self):
        """The freeze-out rows split by event, with each event's own
        particle ids (row[0] - offset)."""
        rows = self.freeze_out_surface()
        ids = rows[:, 0].astype(np.int64)
        ev = np.searchsorted(self.offsets, ids, side="right") - 1
        out = []
        for e in range(self.n_events):
            r = rows[ev == e].copy()
            r[:, 0] -= self.offsets[e]
            out.append(r)
        return out


def concat_ics(ics):
    """One initial-condition dict from several (the batch's columns)."""
    keys = [k for k in ("x", "y", "eta", "e", "ux", "uy", "ueta", "rhoB", "rhoS", "rhoQ", "bulk")
            if all(k in d for d in ics)]
    out = {k: np.concatenate([np.asarray(d[k], dtype=np.float64) for d in ics]) for k in keys}
    if all("shv" in d for d in ics):
        out["shv"] = np.concatenate([np.asarray(d["shv"], dtype=np.float64) for d in ics], axis=1)
    return out


def load_event(ev, ic):
    """Upload an initial condition dict (x, y, e, ux, uy, rhoB, rhoS, rhoQ;
    for a 3+1D event eta and ueta; optionally bulk and shv, the initial Pi
    and pi^{mu nu} that initialize reads with bulk_from_trace /
    input_initial_shear)."""
    ev.upload("CCAKE_F_X", ic["x"])
    ev.upload("CCAKE_F_Y", ic["y"])
    if "eta" in ic:
        ev.upload("CCAKE_F_ETA", ic["eta"])
    if "ueta" in ic:
        ev.upload("CCAKE_F_UETA", ic["ueta"])
    ev.upload("CCAKE_F_IN_E", ic["e"])
    ev.upload("CCAKE_F_UX", ic["ux"])
    ev.upload("CCAKE_F_UY", ic["uy"])
    ev.upload("CCAKE_F_IN_B", ic["rhoB"])
    ev.upload("CCAKE_F_IN_SS", ic["rhoS"])
    ev.upload("CCAKE_F_IN_Q", ic["rhoQ"])
    if "bulk" in ic:
        ev.upload("CCAKE_F_BULK", ic["bulk"])
    if "shv" in ic:
        for c in range(16):
            ev.upload(FIELDS["CCAKE_F_SHV0"] + c, ic["shv"][c])
