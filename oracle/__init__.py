"""ctypes front-end of the CPU oracle (test infrastructure only).

The oracle is a plain-C restatement of CCAKE's per-timestep SPH path
(``oracle/*.c``).  It is the parity checker for the HIP library and the
timed CPU baseline; nothing in the product imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libccake_oracle.so")

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


class OrcParams(C.Structure):
    _fields_ = PARAM_FIELDS


class OrcEosDesc(C.Structure):
    _fields_ = [("kind", C.c_int), ("T_freeze_out", C.c_double), ("dims", C.c_int * 4),
                ("mins", C.c_double * 4), ("maxs", C.c_double * 4),
                ("table", C.POINTER(C.c_double))]


def _parse_enum(header, prefix):
    """Field numbering of the oracle header (a C enum with arithmetic)."""
    import re
    text = re.sub(r"/\*.*?\*/", "", open(header).read(), flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    body = text[text.index("enum " + prefix):]
    body = body[body.index("{") + 1: body.index("}")]
    out, cur = {}, -1
    for item in body.split(","):
        item = item.strip()
        if not item:
            continue
        if "=" in item:
            name, expr = [x.strip() for x in item.split("=", 1)]
            for k, v in sorted(out.items(), key=lambda kv: -len(kv[0])):
                expr = re.sub(r"\b%s\b" % k, str(v), expr)
            cur = int(eval(expr))
        else:
            name = item
            cur += 1
        out[name] = cur
    return out


FIELDS = _parse_enum(os.path.join(_HERE, "ccake_oracle.h"), "orc_field")
IFIELDS = _parse_enum(os.path.join(_HERE, "ccake_oracle.h"), "orc_ifield")
FO_ROW = 42       # hypersurface row: id, step, then the 40 freeze_out.dat columns


def build(force=False):
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-s"], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32)
        sig = {
            "orc_kernel": (C.c_double, [C.c_int, C.c_double, C.c_double]),
            "orc_grad_kernel": (None, [C.c_int, dp, C.c_double, C.c_double, dp]),
            "orc_eos_create": (vp, [C.POINTER(OrcEosDesc)]),
            "orc_eos_destroy": (None, [vp]),
            "orc_eos_members": (C.c_int, [vp]),
            "orc_eos_full_thermo": (None, [vp, C.c_int, dp, dp]),
            "orc_eos_efreeze": (C.c_double, [vp, C.c_double]),
            "orc_preinv_point": (None, [vp, C.c_int, C.c_double, C.c_double, dp]),
            "orc_eos_invert_point": (C.c_int, [vp, C.c_int, dp, dp, C.POINTER(C.c_int), dp]),
            "orc_event_create": (vp, [C.POINTER(OrcParams), vp, C.c_int]),
            "orc_event_destroy": (None, [vp]),
            "orc_field": (dp, [vp, C.c_int]),
            "orc_ifield": (ip, [vp, C.c_int]),
            "orc_set_time": (None, [vp, C.c_double]),
            "orc_get_time": (C.c_double, [vp]),
            "orc_build_neighbors": (C.c_int, [vp, C.c_int]),
            "orc_pair_count": (C.c_int64, [vp]),
            "orc_export_pairs": (C.c_int, [vp, ip, ip, C.c_int64]),
            "orc_gamma_velocity": (None, [vp]),
            "orc_smooth_fields": (None, [vp]),
            "orc_initial_sigma": (None, [vp]),
            "orc_eos_invert": (C.c_int, [vp, C.c_int]),
            "orc_reset_pi_tensor": (None, [vp]),
            "orc_viscosities": (None, [vp]),
            "orc_smooth_gradients": (None, [vp]),
            "orc_eom": (None, [vp]),
            "orc_get_time_derivatives": (C.c_int, [vp]),
            "orc_cache_state": (None, [vp]),
            "orc_rk_update": (C.c_int, [vp, C.c_double, C.c_double]),
            "orc_advance_timestep": (C.c_int, [vp, C.c_double, C.c_int]),
            "orc_initialize": (C.c_int, [vp]),
            "orc_freeze_out_particles": (None, [vp]),
            "orc_freeze_out_surface": (C.c_int, [vp, dp, C.c_int]),
            "orc_continue_evolution": (C.c_int, [vp, C.c_double]),
            "orc_reduce": (C.c_double, [vp, C.c_int]),
            "orc_set_threads": (None, [C.c_int]),
            "orc_get_threads": (C.c_int, []),
            "orc_set_serial_eos": (None, [C.c_int]),
            "orc_phase_reset": (None, []),
            "orc_phase_read": (None, [dp, C.POINTER(C.c_longlong)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def make_params(cfg):
    p = OrcParams()
    for name, typ in PARAM_FIELDS:
        v = cfg.get(name, 0) if name in ("pair_precision", "factorized_kernel", "x_subcells") else cfg[name]
        setattr(p, name, typ(*v) if hasattr(typ, "_length_") else v)
    return p


class Eos:
    def __init__(self, kind=0, T_freeze_out=150.0 / 197.3269804, dims=None, mins=None,
                 maxs=None, table=None):
        d = OrcEosDesc()
        d.kind = kind
        d.T_freeze_out = T_freeze_out
        self._table = None
        if table is not None:
            self._table = np.ascontiguousarray(table, dtype=np.float64)
            for a in range(4):
                d.dims[a] = int(dims[a]); d.mins[a] = float(mins[a]); d.maxs[a] = float(maxs[a])
            d.table = self._table.ctypes.data_as(C.POINTER(C.c_double))
        self.h = lib().orc_eos_create(C.byref(d))
        if not self.h:
            raise ValueError("the oracle rejected this EoS description")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_eos_destroy(self.h)
            self.h = None

    def full_thermo(self, member, tbqs):
        x = (C.c_double * 4)(*tbqs)
        out = (C.c_double * 17)()
        lib().orc_eos_full_thermo(self.h, member, x, out)
        return np.array(out[:])

    def efreeze(self, T):
        return lib().orc_eos_efreeze(self.h, T)

    def preinv_point(self, which, a, b):
        """Pre-inverted table queries: 0 fill_thermodynamics, 1
        query_thermo_host, 2 invert_e_to_s_host (returns s)."""
        out = (C.c_double * 12)()
        lib().orc_preinv_point(self.h, which, a, b, out)
        return out[0] if which == 2 else np.array(out[:12 if which == 0 else 10])

    def invert_point(self, mode, target, seed, member=0):
        tg = (C.c_double * 4)(*target)
        x = (C.c_double * 4)(*seed)
        m = C.c_int(member)
        th = (C.c_double * 31)()
        ok = lib().orc_eos_invert_point(self.h, mode, tg, x, C.byref(m), th)
        return bool(ok), np.array(x[:]), m.value, np.array(th[:])


class Event:
    """One particle system on the CPU oracle."""

    def __init__(self, cfg, eos, n):
        self.cfg = cfg
        self.eos = eos
        self.n = int(n)
        self._p = make_params(cfg)
        self.h = lib().orc_event_create(C.byref(self._p), eos.h, self.n)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_event_destroy(self.h)
            self.h = None

    def field(self, f):
        fid = FIELDS[f] if isinstance(f, str) else int(f)
        ptr = lib().orc_field(self.h, fid)
        return np.ctypeslib.as_array(ptr, shape=(self.n,))

    def ifield(self, f):
        fid = IFIELDS[f] if isinstance(f, str) else int(f)
        ptr = lib().orc_ifield(self.h, fid)
        return np.ctypeslib.as_array(ptr, shape=(self.n,))

    def set(self, f, values):
        self.field(f)[:] = values

    @property
    def t(self):
        return lib().orc_get_time(self.h)

    @t.setter
    def t(self, v):
        lib().orc_set_time(self.h, float(v))

    def build_neighbors(self, brute_force=False):
        return lib().orc_build_neighbors(self.h, int(brute_force))

    def pairs(self):
        n = lib().orc_pair_count(self.h)
        pi = np.empty(n, dtype=np.int32)
        pj = np.empty(n, dtype=np.int32)
        st = lib().orc_export_pairs(self.h, pi.ctypes.data_as(C.POINTER(C.c_int32)),
                                    pj.ctypes.data_as(C.POINTER(C.c_int32)), n)
        assert st == 0
        return pi, pj

    def freeze_out_surface(self):
        """Hypersurface rows (id, step, 40 freeze_out.dat columns)."""
        k = lib().orc_freeze_out_surface(self.h, None, 0)
        rows = np.zeros((k, FO_ROW))
        lib().orc_freeze_out_surface(self.h, rows.ctypes.data_as(C.POINTER(C.c_double)), k)
        return rows

    def __getattr__(self, name):
        fn = getattr(lib(), "orc_" + name)

        def call(*args):
            return fn(self.h, *args)
        return call


def set_threads(n):
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


# The reference's phase names (src/sph_workstation.cpp:1543-1583), indexed
# as orc_phase_read returns them.
PHASE_NAMES = {1: "01_neighbour_list", 2: "02_gamma_velocity", 3: "03_smooth_fields", 4: "04_thermodynamics",
               5: "05_reset_pi_tensor", 6: "06_regulate_diffusion", 8: "08_viscosities",
               9: "09_smooth_gradients", 10: "10_time_derivatives", 11: "rk_update", 12: "freeze_out"}


def set_serial_eos(on):
    """Run the EoS inversion loop on one thread (the reference's serial loop)."""
    lib().orc_set_serial_eos(1 if on else 0)


def phase_reset():
    lib().orc_phase_reset()


def phase_times():
    """{phase name: (seconds, calls)} accumulated since phase_reset()."""
    n = 13
    s = (C.c_double * n)()
    c = (C.c_longlong * n)()
    lib().orc_phase_read(s, c)
    return {PHASE_NAMES[k]: (s[k], c[k]) for k in PHASE_NAMES if c[k]}
