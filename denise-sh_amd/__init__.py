"""Python binding of libdenise_sh.so (ctypes over the C-ABI in include/denise_sh.h).

Plumbing for tests and benchmarks: every computation runs in the HIP library;
this module only marshals numpy arrays.  Loading fails loudly if the library
has not been built (python denise-sh_amd/build.py).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdenise_sh.so")

# Every symbol include/denise_sh.h declares (checked by the CPU test suite).
EXPORTS = [
    "dn_set_globals", "update_v_PML", "update_s_elastic_PML", "update_s_visc_PML", "seismo_ssg", "psource",
    "dn_create", "dn_destroy", "dn_last_error", "dn_set_model", "dn_get_model", "dn_set_receivers",
    "dn_set_shots", "dn_set_observed", "dn_set_batch", "dn_forward", "dn_gradient", "dn_get_gradient",
    "dn_forward_fields", "dn_last_timing", "dn_holbergcoeff", "dn_pml_pro", "dn_relaxation_constants",
    "dn_wavelet", "dn_timedomain_filt", "dn_pass_timing", "dn_set_shots_multi", "dn_set_signals",
    "dn_par_read", "dn_par_free", "dn_par_get_int", "dn_par_get_float", "dn_par_get_floats", "dn_par_get_string",
    "dn_workflow_read", "dn_workflow_free", "dn_workflow_nstages", "dn_workflow_get", "dn_read_sources",
    "dn_read_receivers", "dn_read_model_file", "dn_write_model_file", "dn_read_model", "dn_write_su", "dn_read_su",
    "dn_fwi_start", "dn_fwi_iteration", "dn_fwi_get_model", "dn_fwi_get_direction", "dn_checkfd",
    "dn_set_exec", "dn_autotune", "dn_pseudo_hessian", "dn_get_energy_maps", "dn_hessian_probe",
    "dn_unet_create", "dn_unet_destroy", "dn_unet_num_params", "dn_unet_get_params", "dn_unet_set_params",
    "dn_unet_forward", "dn_unet_loss_grad", "dn_unet_train_step", "dn_fwi_set_learned", "dn_precond_features",
    "dn_smooth_grad", "dn_taper_grad", "dn_gauss_filt_var", "dn_taper_grad_host", "dn_gauss_filt", "dn_get_means", "dn_smooth2",
    "dn_set_datacond", "dn_time_window", "dn_stf_shot", "dn_stf", "dn_fwi_write_files", "dn_set_wave_shots",
    "dn_set_group_layout", "dn_batch_shots", "dn_calc_res",
    "dn_calc_res_noalloc",
    "dn_set_time_filt", "zero_fdveps", "zero_fdveps_visc", "dn_dropin_resident",
    "dn_dropin_sync", "dn_fwi_save_state", "dn_fwi_load_state",
    "dn_fwi_set_learned_fixed", "dn_get_launch_shape", "dn_fwi_set_learned_revert", "dn_fwi_set_learned_power", "dn_set_batch_order",
]


class Globals(C.Structure):
    _fields_ = [("NX", C.c_int), ("NY", C.c_int), ("FDORDER", C.c_int), ("FW", C.c_int), ("FREE_SURF", C.c_int),
                ("BOUNDARY", C.c_int), ("L", C.c_int), ("GRAD_FORM", C.c_int), ("QUELLTYPB", C.c_int),
                ("NT", C.c_int), ("DT", C.c_float), ("DH", C.c_float)]


class Params(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("fdorder", C.c_int), ("maxrelerror", C.c_int), ("fw", C.c_int),
                ("free_surf", C.c_int), ("L", C.c_int), ("fl", C.c_float * 4), ("dt", C.c_float), ("dh", C.c_float),
                ("damping", C.c_float), ("fpml", C.c_float), ("npower", C.c_float), ("k_max_pml", C.c_float),
                ("dtinv", C.c_int), ("snap_precision", C.c_int), ("eprecond", C.c_int), ("boundary", C.c_int)]


class Taper(C.Structure):
    _fields_ = [("sws", C.c_int), ("gradt", C.c_int * 4), ("srtshape", C.c_int), ("srtradius", C.c_float),
                ("filtsize", C.c_int), ("dh", C.c_float), ("map", C.POINTER(C.c_float))]


def _taper(sws=0, gradt=(0, 0, 0, 0), srtshape=1, srtradius=1.0, filtsize=1, dh=0.0, taper_map=None):
    """(Taper, kept array): the map array must outlive the call."""
    t = Taper()
    t.sws = sws
    for k in range(4):
        t.gradt[k] = int(gradt[k])
    t.srtshape, t.srtradius, t.filtsize, t.dh = srtshape, srtradius, filtsize, dh
    keep = None
    if taper_map is not None:
        keep = np.ascontiguousarray(taper_map, np.float32)
        t.map = _f(keep)
    return t, keep


class DataCond(C.Structure):
    _fields_ = [("trkill", C.c_int), ("kill", C.POINTER(C.c_int)), ("kill_nsrc", C.c_int), ("kill_ntr", C.c_int),
                ("offset_mute", C.c_int), ("offsetc", C.c_float), ("src_x", C.POINTER(C.c_float)),
                ("src_y", C.POINTER(C.c_float))]


class StfOpts(C.Structure):
    _fields_ = [("eps_stf", C.c_float), ("offsetc_stf", C.c_float), ("normalize", C.c_int)]


class FwiFiles(C.Structure):
    _fields_ = [("misfit_log", C.c_char_p), ("time_filt", C.c_int), ("fc_start", C.c_float), ("fc", C.c_float),
                ("jacobian", C.c_char_p), ("inv_modelfile", C.c_char_p), ("write_it", C.c_int),
                ("lbfgs_old", C.c_int)]


class FwiOpts(C.Structure):
    _fields_ = [("grad_method", C.c_int), ("nlbfgs", C.c_int), ("eps_scale", C.c_float), ("stepmax", C.c_int),
                ("scalefac", C.c_float), ("testshot_start", C.c_int), ("testshot_end", C.c_int),
                ("testshot_incr", C.c_int), ("vs_lower", C.c_float), ("vs_upper", C.c_float),
                ("rho_lower", C.c_float), ("rho_upper", C.c_float), ("inv_vs_iter", C.c_int),
                ("inv_rho_iter", C.c_int), ("scalerho", C.c_float), ("min_iter", C.c_int), ("pro", C.c_float),
                ("spatfilter", C.c_int), ("wd_damp", C.c_float), ("wd_damp1", C.c_float), ("fc_end", C.c_float),
                ("taper_vert", C.c_int), ("taper_hor", C.c_int), ("taper_sources", C.c_int),
                ("taper_file", C.c_int), ("taper", Taper), ("time_filt", C.c_int), ("fc_start", C.c_float),
                ("order", C.c_int)]


class FwiRecord(C.Structure):
    _fields_ = [("stage", C.c_int), ("iter", C.c_int), ("iter_true", C.c_int), ("opteps_vs", C.c_float),
                ("epst1", C.c_float * 3), ("L2t", C.c_float * 4), ("M2_vs", C.c_float), ("M2_rho", C.c_float),
                ("eps_u", C.c_float), ("eps_rho", C.c_float), ("l2", C.c_double), ("energy", C.c_double),
                ("step1", C.c_int), ("step3", C.c_int), ("test_forwards", C.c_int), ("next_stage", C.c_int),
                ("precond", C.c_int), ("judged", C.c_int), ("drop_learned", C.c_float), ("drop_plain", C.c_float)]


FWI_DEFAULTS = dict(grad_method=2, nlbfgs=20, eps_scale=0.01, stepmax=6, scalefac=2.0, testshot_start=1,
                    testshot_end=1, testshot_incr=1, vs_lower=0.0, vs_upper=1e9, rho_lower=0.0, rho_upper=1e9,
                    inv_vs_iter=0, inv_rho_iter=0, scalerho=0.5, min_iter=0, pro=0.0, spatfilter=0, wd_damp=0.0,
                    wd_damp1=0.0, fc_end=0.0, taper_vert=0, taper_hor=0, taper_sources=0, taper_file=0,
                    time_filt=0, fc_start=0.0, order=0)
# taper parameters of fwi_start (go into FwiOpts.taper)
TAPER_KEYS = ("gradt", "srtshape", "srtradius", "filtsize", "taper_map")

_lib = None


def lib():
    """Load libdenise_sh.so (raises if it is missing)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libdenise_sh.so not built: run python denise-sh_amd/build.py")
        L = C.CDLL(LIB_PATH)
        fp = C.POINTER(C.c_float)
        ip = C.POINTER(C.c_int)
        L.dn_create.restype = C.c_void_p
        L.dn_create.argtypes = [C.POINTER(Params)]
        L.dn_destroy.argtypes = [C.c_void_p]
        L.dn_last_error.restype = C.c_char_p
        L.dn_last_error.argtypes = [C.c_void_p]
        L.dn_set_model.argtypes = [C.c_void_p, fp, fp, fp]
        L.dn_get_model.argtypes = [C.c_void_p, fp, fp]
        L.dn_set_receivers.argtypes = [C.c_void_p, C.c_int, ip, ip]
        L.dn_set_shots.argtypes = [C.c_void_p, C.c_int, C.c_int, ip, ip, ip, fp]
        L.dn_set_observed.argtypes = [C.c_void_p, fp]
        L.dn_set_batch.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.dn_set_batch_order.argtypes = [C.c_void_p, C.c_int]
        L.dn_holbergcoeff.argtypes = [C.c_int, C.c_int, fp]
        L.dn_batch_shots.argtypes = [C.c_void_p, C.c_int, ip]
        L.dn_forward.argtypes = [C.c_void_p, C.c_int, C.c_int, fp]
        L.dn_gradient.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double),
                                  C.POINTER(C.c_double)]
        L.dn_get_gradient.argtypes = [C.c_void_p, fp, fp, fp, fp]
        L.dn_forward_fields.argtypes = [C.c_void_p, C.c_int, C.c_int, fp, fp, fp]
        L.dn_last_timing.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_longlong),
                                     C.POINTER(C.c_longlong)]
        L.dn_set_globals.argtypes = [C.POINTER(Globals)]
        L.dn_set_exec.argtypes = [C.c_void_p, C.c_int]
        L.dn_pseudo_hessian.argtypes = [C.c_void_p, C.c_int, C.c_int, fp, fp]
        L.dn_get_energy_maps.argtypes = [C.c_void_p, fp, fp]
        L.dn_hessian_probe.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_uint, fp]
        L.dn_autotune.argtypes = [C.c_void_p, C.c_int, C.c_int, ip, ip]
        L.dn_set_shots_multi.argtypes = [C.c_void_p, C.c_int, C.c_int, ip, ip, ip, ip, fp]
        L.dn_set_signals.argtypes = [C.c_void_p, fp]
        L.dn_unet_create.restype = C.c_void_p
        L.dn_unet_create.argtypes = [C.c_uint]
        L.dn_unet_destroy.argtypes = [C.c_void_p]
        L.dn_unet_num_params.restype = C.c_longlong
        L.dn_unet_num_params.argtypes = [C.c_void_p]
        L.dn_unet_get_params.argtypes = [C.c_void_p, fp]
        L.dn_unet_set_params.argtypes = [C.c_void_p, fp]
        L.dn_unet_forward.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, fp, fp]
        L.dn_unet_loss_grad.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, fp, fp, C.POINTER(C.c_double), fp]
        L.dn_unet_train_step.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, fp, fp, C.c_float,
                                         C.POINTER(C.c_double)]
        _lib = L
    return _lib


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else None


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class DeniseError(RuntimeError):
    pass


class Context:
    """Handle API wrapper: one dn_ctx on the current HIP device."""

    def __init__(self, nx, ny, fdorder=8, maxrelerror=1, fw=20, free_surf=1, L=0, fl=(), dt=1e-4, dh=0.25,
                 damping=300.0, fpml=40.0, npower=2.0, k_max_pml=1.0, dtinv=1, snap_precision=0, eprecond=1,
                 boundary=0):
        p = Params()
        p.nx, p.ny, p.fdorder, p.maxrelerror, p.fw, p.free_surf, p.L = nx, ny, fdorder, maxrelerror, fw, free_surf, L
        for k, v in enumerate(fl):
            p.fl[k] = v
        p.dt, p.dh, p.damping, p.fpml, p.npower, p.k_max_pml = dt, dh, damping, fpml, npower, k_max_pml
        p.dtinv, p.snap_precision, p.eprecond, p.boundary = dtinv, snap_precision, eprecond, boundary
        self.p = p
        self.nx, self.ny = nx, ny
        self._h = lib().dn_create(C.byref(p))
        if not self._h:
            raise DeniseError(lib().dn_last_error(None).decode())
        self.ntr = 0
        self.nshots = 0
        self.nt_max = 0

    def _chk(self, rc):
        if rc != 0:
            raise DeniseError(lib().dn_last_error(self._h).decode())

    def close(self):
        if self._h:
            lib().dn_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_model(self, vs, rho, qs=None):
        vs = np.ascontiguousarray(vs, np.float32)
        rho = np.ascontiguousarray(rho, np.float32)
        q = None if qs is None else np.ascontiguousarray(qs, np.float32)
        self._chk(lib().dn_set_model(self._h, _f(vs), _f(rho), _f(q)))

    def get_model(self):
        vs = np.zeros((self.ny, self.nx), np.float32)
        rho = np.zeros((self.ny, self.nx), np.float32)
        self._chk(lib().dn_get_model(self._h, _f(vs), _f(rho)))
        return vs, rho

    def set_receivers(self, rec_i, rec_j):
        ri = np.ascontiguousarray(rec_i, np.int32)
        rj = np.ascontiguousarray(rec_j, np.int32)
        self._chk(lib().dn_set_receivers(self._h, len(ri), _i(ri), _i(rj)))
        self.ntr = len(ri)

    def set_shots(self, src_i, src_j, signals, nt=None):
        sig = np.ascontiguousarray(signals, np.float32)
        ns, ntm = sig.shape
        nt = np.full(ns, ntm, np.int32) if nt is None else np.ascontiguousarray(nt, np.int32)
        si = np.ascontiguousarray(src_i, np.int32)
        sj = np.ascontiguousarray(src_j, np.int32)
        self._chk(lib().dn_set_shots(self._h, ns, ntm, _i(nt), _i(si), _i(sj), _f(sig)))
        self.nshots, self.nt_max = ns, ntm
        self.shot_nt = nt.copy()

    def set_shots_multi(self, src_off, src_i, src_j, signals, nt):
        """Shots with several sources: those of shot s are src_off[s]..src_off[s+1]-1."""
        sig = np.ascontiguousarray(signals, np.float32)
        off = np.ascontiguousarray(src_off, np.int32)
        nt = np.ascontiguousarray(nt, np.int32)
        ns, ntm = len(nt), sig.shape[1]
        self._chk(lib().dn_set_shots_multi(self._h, ns, ntm, _i(nt), _i(off), _i(np.ascontiguousarray(src_i, np.int32)),
                                           _i(np.ascontiguousarray(src_j, np.int32)), _f(sig)))
        self.nshots, self.nt_max = ns, ntm
        self.shot_nt = nt.copy()

    def set_signals(self, signals):
        self._chk(lib().dn_set_signals(self._h, _f(np.ascontiguousarray(signals, np.float32))))

    def set_observed(self, obs):
        obs = np.ascontiguousarray(obs, np.float32)
        assert obs.shape == (self.nshots, self.ntr, self.nt_max)
        self._chk(lib().dn_set_observed(self._h, _f(obs)))

    def set_time_filt(self, time_filt, fc_start=0.0, fc=0.0, order=0):
        """Stage frequency filter of the signals and observed data (from
        the copies last set; time_filt 0 restores them)."""
        lib().dn_set_time_filt.argtypes = [C.c_void_p, C.c_int, C.c_float, C.c_float, C.c_int]
        self._chk(lib().dn_set_time_filt(self._h, time_filt, fc_start, fc, order))

    # ---- FWI (dn_fwi_*) ----
    def fwi_start(self, **opts):
        o = FwiOpts()
        tk = {k: opts.pop(k) for k in TAPER_KEYS if k in opts}
        for k, v in dict(FWI_DEFAULTS, **opts).items():
            setattr(o, k, v)
        o.taper, self._taper_keep = _taper(**tk)
        lib().dn_fwi_start.argtypes = [C.c_void_p, C.POINTER(FwiOpts)]
        self._chk(lib().dn_fwi_start(self._h, C.byref(o)))

    def fwi_iteration(self):
        r = FwiRecord()
        lib().dn_fwi_iteration.argtypes = [C.c_void_p, C.POINTER(FwiRecord)]
        self._chk(lib().dn_fwi_iteration(self._h, C.byref(r)))
        out = {k: getattr(r, k) for k, _ in FwiRecord._fields_}
        out["epst1"] = list(r.epst1)
        out["L2t"] = list(r.L2t)
        return out

    def fwi_write_files(self, rec, misfit_log=None, time_filt=0, fc_start=0.0, fc=0.0, jacobian=None,
                        inv_modelfile=None, write_it=0, lbfgs_old=0):
        """Writes the reference's per-iteration files for the record rec of
        fwi_iteration()."""
        r = FwiRecord()
        for k, _ in FwiRecord._fields_:
            if k in ("epst1", "L2t"):
                for i, v in enumerate(rec[k]):
                    getattr(r, k)[i] = v
            else:
                setattr(r, k, rec[k])
        enc = lambda x: None if x is None else x.encode()
        f = FwiFiles(enc(misfit_log), time_filt, fc_start, fc, enc(jacobian), enc(inv_modelfile), write_it,
                     lbfgs_old)
        lib().dn_fwi_write_files.argtypes = [C.c_void_p, C.POINTER(FwiRecord), C.POINTER(FwiFiles)]
        self._chk(lib().dn_fwi_write_files(self._h, C.byref(r), C.byref(f)))

    def fwi_save_state(self, path):
        lib().dn_fwi_save_state.argtypes = [C.c_void_p, C.c_char_p]
        self._chk(lib().dn_fwi_save_state(self._h, path.encode()))

    def fwi_load_state(self, path, jacobian=None, name_fix=1):
        lib().dn_fwi_load_state.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p, C.c_int]
        self._chk(lib().dn_fwi_load_state(self._h, path.encode(), None if jacobian is None else jacobian.encode(),
                                          name_fix))

    def fwi_model(self):
        vs = np.zeros((self.ny, self.nx), np.float32)
        rho = np.zeros((self.ny, self.nx), np.float32)
        lib().dn_fwi_get_model.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        self._chk(lib().dn_fwi_get_model(self._h, _f(vs), _f(rho)))
        return vs, rho

    def set_learned(self, net, judge_every=0, fixed=True, revert=True, power=1.0):
        lib().dn_fwi_set_learned.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
        self._chk(lib().dn_fwi_set_learned(self._h, net._h if net is not None else None, judge_every))
        lib().dn_fwi_set_learned_fixed.argtypes = [C.c_void_p, C.c_int]
        self._chk(lib().dn_fwi_set_learned_fixed(self._h, 1 if fixed else 0))
        lib().dn_fwi_set_learned_revert.argtypes = [C.c_void_p, C.c_int]
        self._chk(lib().dn_fwi_set_learned_revert(self._h, 1 if revert else 0))
        lib().dn_fwi_set_learned_power.argtypes = [C.c_void_p, C.c_float]
        self._chk(lib().dn_fwi_set_learned_power(self._h, power))
        self._net = net      # keep the network alive while the context uses it

    def precond_features(self):
        h8, w8 = (self.ny + 7) // 8 * 8, (self.nx + 7) // 8 * 8
        x = np.zeros((3, h8, w8), np.float32)
        lib().dn_precond_features.argtypes = [C.c_void_p, C.POINTER(C.c_float)]
        self._chk(lib().dn_precond_features(self._h, _f(x)))
        return x

    def set_exec(self, graph_steps=-1):
        self._chk(lib().dn_set_exec(self._h, graph_steps))

    def set_wave_shots(self, n=2):
        lib().dn_set_wave_shots.argtypes = [C.c_void_p, C.c_int]
        self._chk(lib().dn_set_wave_shots(self._h, n))

    def set_group_layout(self, mode):
        """Workgroup layout of the step kernel: 0 eight adjacent strips of one
        shot (default); 2, 4: 8/S adjacent strips of S shots; 1 or 8: one
        strip of eight shots."""
        lib().dn_set_group_layout.argtypes = [C.c_void_p, C.c_int]
        self._chk(lib().dn_set_group_layout(self._h, mode))

    def launch_shape(self):
        """(rows per segment (0: cost model), graph steps, shots per wave, shots per workgroup)."""
        v = [C.c_int() for _ in range(4)]
        lib().dn_get_launch_shape.argtypes = [C.c_void_p] + [C.POINTER(C.c_int)] * 4
        self._chk(lib().dn_get_launch_shape(self._h, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def autotune(self, nshots, nsteps=40):
        rows, graph = C.c_int(), C.c_int()
        self._chk(lib().dn_autotune(self._h, nshots, nsteps, C.byref(rows), C.byref(graph)))
        return rows.value, graph.value

    def set_batch(self, max_shots=0, rows_per_segment=0):
        self._chk(lib().dn_set_batch(self._h, max_shots, rows_per_segment))

    def set_batch_order(self, by_record_length=True):
        self._chk(lib().dn_set_batch_order(self._h, 1 if by_record_length else 0))

    def batch_shots(self, gradient=True):
        """Shots advanced together in one batch (allocates the batch workspace)."""
        n = C.c_int()
        self._chk(lib().dn_batch_shots(self._h, 1 if gradient else 0, C.byref(n)))
        return n.value

    def forward(self, shot0=0, nshots=None, want_seis=True):
        nshots = self.nshots - shot0 if nshots is None else nshots
        seis = np.zeros((nshots, self.ntr, self.nt_max), np.float32) if want_seis else None
        self._chk(lib().dn_forward(self._h, shot0, nshots, _f(seis)))
        return seis

    def gradient(self, shot0=0, nshots=None, accumulate=False):
        nshots = self.nshots - shot0 if nshots is None else nshots
        l2 = C.c_double()
        en = C.c_double()
        self._chk(lib().dn_gradient(self._h, shot0, nshots, 1 if accumulate else 0, C.byref(l2), C.byref(en)))
        return l2.value, en.value

    def energy_maps(self):
        ws = np.zeros((self.ny, self.nx), np.float32)
        wr = np.zeros((self.ny, self.nx), np.float32)
        self._chk(lib().dn_get_energy_maps(self._h, _f(ws), _f(wr)))
        return ws, wr

    def hessian_probe(self, nprobes=8, eps=1e-3, seed=1, shot0=0, nshots=None):
        nshots = self.nshots - shot0 if nshots is None else nshots
        d = np.zeros((self.ny, self.nx), np.float32)
        self._chk(lib().dn_hessian_probe(self._h, shot0, nshots, nprobes, eps, seed, _f(d)))
        return d

    def pseudo_hessian(self, shot0=0, nshots=None):
        nshots = self.nshots - shot0 if nshots is None else nshots
        hv = np.zeros((self.ny, self.nx), np.float32)
        hr = np.zeros((self.ny, self.nx), np.float32)
        self._chk(lib().dn_pseudo_hessian(self._h, shot0, nshots, _f(hv), _f(hr)))
        return hv, hr

    # ---- data conditioning (calc_res.c, calc_energy.c) and stf.c ----
    def set_datacond(self, trkill=0, kill=None, offset_mute=0, offsetc=0.0, src_x=None, src_y=None):
        """kill: the TRKILL_FILE integers as a [nsrc_file][ntr_file] array."""
        d = DataCond()
        keep = []
        d.trkill, d.offset_mute, d.offsetc = trkill, offset_mute, offsetc
        if kill is not None:
            k = np.ascontiguousarray(kill, np.int32)
            keep.append(k)
            d.kill = _i(k)
            d.kill_nsrc, d.kill_ntr = k.shape
        for name, v in (("src_x", src_x), ("src_y", src_y)):
            if v is not None:
                a = np.ascontiguousarray(v, np.float32)
                keep.append(a)
                setattr(d, name, _f(a))
        lib().dn_set_datacond.argtypes = [C.c_void_p, C.POINTER(DataCond)]
        off = trkill == 0 and offset_mute == 0
        self._chk(lib().dn_set_datacond(self._h, None if off else C.byref(d)))

    def stf_shot(self, shot, eps_stf=0.05, offsetc_stf=0.0, normalize=0):
        o = StfOpts(eps_stf, offsetc_stf, normalize)
        out = np.zeros(int(self.shot_nt[shot]), np.float32)
        lib().dn_stf_shot.argtypes = [C.c_void_p, C.c_int, C.POINTER(StfOpts), C.POINTER(C.c_float)]
        self._chk(lib().dn_stf_shot(self._h, shot, C.byref(o), _f(out)))
        return out

    # ---- gradient post-processing (smooth_grad.c, taper_grad.c) ----
    def means(self):
        """(C_vs, C_rho), or None while they are not taken yet."""
        cv, cr = C.c_float(), C.c_float()
        lib().dn_get_means.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        rc = lib().dn_get_means(self._h, C.byref(cv), C.byref(cr))
        return None if rc else (cv.value, cr.value)

    def smooth_grad(self, spatfilter, wd_damp, wd_damp1, fc_end):
        """Smooths the context's gradients in place; returns the device ms."""
        ms = C.c_float()
        lib().dn_smooth_grad.argtypes = [C.c_void_p, C.c_int, C.c_float, C.c_float, C.c_float,
                                         C.POINTER(C.c_float)]
        self._chk(lib().dn_smooth_grad(self._h, spatfilter, wd_damp, wd_damp1, fc_end, C.byref(ms)))
        return ms.value

    def taper_grad(self, which, sws, **kw):
        t, keep = _taper(sws=sws, **kw)
        lib().dn_taper_grad.argtypes = [C.c_void_p, C.c_int, C.POINTER(Taper)]
        self._chk(lib().dn_taper_grad(self._h, which, C.byref(t)))

    def get_gradient(self):
        out = [np.zeros((self.ny, self.nx), np.float32) for _ in range(4)]
        self._chk(lib().dn_get_gradient(self._h, *[_f(a) for a in out]))
        return out  # grad_vs, grad_rho, raw_u, raw_rho

    def forward_fields(self, shot, nsteps):
        out = [np.zeros((self.ny, self.nx), np.float32) for _ in range(3)]
        self._chk(lib().dn_forward_fields(self._h, shot, nsteps, *[_f(a) for a in out]))
        return out  # vy, sxy, syz

    def pass_timing(self, adjoint):
        ms = C.c_double()
        la = C.c_longlong()
        ce = C.c_longlong()
        lib().dn_pass_timing.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_longlong),
                                         C.POINTER(C.c_longlong)]
        self._chk(lib().dn_pass_timing(self._h, 1 if adjoint else 0, C.byref(ms), C.byref(la), C.byref(ce)))
        return ms.value, la.value, ce.value

    def last_timing(self):
        ms = C.c_double()
        la = C.c_longlong()
        ce = C.c_longlong()
        self._chk(lib().dn_last_timing(self._h, C.byref(ms), C.byref(la), C.byref(ce)))
        return ms.value, la.value, ce.value


# ---------------------------------------------------------------------------
# Numerical-Recipes style offset arrays for the reference-named drop-ins.
# ---------------------------------------------------------------------------
class NRMatrix:
    """float matrix m[nrl..nrh][ncl..nch] backed by a contiguous numpy array."""

    def __init__(self, nrl, nrh, ncl, nch, data=None):
        self.nrl, self.nrh, self.ncl, self.nch = nrl, nrh, ncl, nch
        shape = (nrh - nrl + 1, nch - ncl + 1)
        self.a = np.zeros(shape, np.float32) if data is None else np.ascontiguousarray(data, np.float32)
        assert self.a.shape == shape
        nrow = shape[0]
        self._rows = (C.c_void_p * nrow)()
        base = self.a.ctypes.data
        for r in range(nrow):
            self._rows[r] = base + (r * shape[1] - ncl) * 4
        self.ptr = C.c_void_p(C.addressof(self._rows) - nrl * C.sizeof(C.c_void_p))

    def __getitem__(self, ji):
        j, i = ji
        return self.a[j - self.nrl, i - self.ncl]


class NRVector:
    """float vector v[nl..nh]."""

    def __init__(self, nl, nh, data=None):
        self.nl = nl
        self.a = np.zeros(nh - nl + 1, np.float32) if data is None else np.ascontiguousarray(data, np.float32)
        self.ptr = C.c_void_p(self.a.ctypes.data - nl * 4)


class NRTensor3:
    """f3tensor t[nrl..nrh][ncl..nch][ndl..ndh] (last index fastest)."""

    def __init__(self, nrl, nrh, ncl, nch, ndl, ndh, data=None):
        shape = (nrh - nrl + 1, nch - ncl + 1, ndh - ndl + 1)
        self.a = np.zeros(shape, np.float32) if data is None else np.ascontiguousarray(data, np.float32)
        assert self.a.shape == shape
        self._rows = (C.c_void_p * shape[0])()
        self._cols = (C.c_void_p * (shape[0] * shape[1]))()
        base = self.a.ctypes.data
        colbase = C.addressof(self._cols)
        for r in range(shape[0]):
            for cidx in range(shape[1]):
                self._cols[r * shape[1] + cidx] = base + ((r * shape[1] + cidx) * shape[2] - ndl) * 4
            self._rows[r] = colbase + (r * shape[1] - ncl) * C.sizeof(C.c_void_p)
        self.ptr = C.c_void_p(C.addressof(self._rows) - nrl * C.sizeof(C.c_void_p))


class NRIMatrix:
    """int matrix m[nrl..nrh][ncl..nch]."""

    def __init__(self, nrl, nrh, ncl, nch, data):
        shape = (nrh - nrl + 1, nch - ncl + 1)
        self.a = np.ascontiguousarray(data, np.int32)
        assert self.a.shape == shape
        self._rows = (C.c_void_p * shape[0])()
        for r in range(shape[0]):
            self._rows[r] = self.a.ctypes.data + (r * shape[1] - ncl) * 4
        self.ptr = C.c_void_p(C.addressof(self._rows) - nrl * C.sizeof(C.c_void_p))


def set_globals(**kw):
    g = Globals()
    defaults = dict(BOUNDARY=0, GRAD_FORM=2, QUELLTYPB=2, L=0, FREE_SURF=1, NT=1)
    defaults.update(kw)
    for k, v in defaults.items():
        setattr(g, k, v)
    rc = lib().dn_set_globals(C.byref(g))
    if rc != 0:
        raise DeniseError(lib().dn_last_error(None).decode())


# ---------------------------------------------------------------------------
# Host setup routines (no device needed)
# ---------------------------------------------------------------------------
PML_NAMES = ("K_x", "a_x", "b_x", "K_x_half", "a_x_half", "b_x_half", "K_y", "a_y", "b_y", "K_y_half", "a_y_half",
             "b_y_half")


def holbergcoeff(fdorder, maxrelerror):
    hc = np.zeros(7, np.float32)
    if lib().dn_holbergcoeff(fdorder, maxrelerror, _f(hc)) != 0:
        raise DeniseError("wrong FDORDER/MAXRELERROR")
    return hc


def pml_pro(nx, ny, fw, dh, dt, damping, fpml, npower=2.0, k_max_pml=1.0):
    out = np.zeros((12, 2 * fw + 1), np.float32)
    f = C.c_float
    lib().dn_pml_pro.argtypes = [C.c_int, C.c_int, C.c_int, f, f, f, f, f, f, C.POINTER(C.c_float)]
    rc = lib().dn_pml_pro(nx, ny, fw, dh, dt, damping, fpml, npower, k_max_pml, _f(out))
    if rc != 0:
        raise DeniseError("dn_pml_pro failed")
    return dict(zip(PML_NAMES, out))


def relaxation_constants(fl, dt):
    L = len(fl)
    fl = np.ascontiguousarray(fl, np.float32)
    eta, bip, bjm, cip, cjm = [np.zeros(L, np.float32) for _ in range(5)]
    sumu = C.c_float()
    lib().dn_relaxation_constants.argtypes = [C.c_int, C.POINTER(C.c_float), C.c_float] + [C.POINTER(C.c_float)] * 6
    rc = lib().dn_relaxation_constants(L, _f(fl), dt, _f(eta), C.byref(sumu), _f(bip), _f(bjm), _f(cip), _f(cjm))
    if rc != 0:
        raise DeniseError("dn_relaxation_constants failed")
    return dict(eta=eta, sumu=np.float32(sumu.value), bip=bip, bjm=bjm, cip=cip, cjm=cjm)


def wavelet(quellart, nt, dt, tshift, fc, amp):
    out = np.zeros(nt, np.float32)
    f = C.c_float
    lib().dn_wavelet.argtypes = [C.c_int, C.c_int, f, f, f, f, C.POINTER(C.c_float)]
    if lib().dn_wavelet(quellart, nt, dt, tshift, fc, amp, _f(out)) != 0:
        raise DeniseError("unknown source wavelet")
    return out


def timedomain_filt(data, dt, fc, order, method):
    d = np.ascontiguousarray(data, np.float32).copy()
    f = C.c_float
    lib().dn_timedomain_filt.argtypes = [C.POINTER(C.c_float), C.c_int, C.c_int, f, f, C.c_int, C.c_int]
    if lib().dn_timedomain_filt(_f(d), d.shape[0], d.shape[1], dt, fc, order, method) != 0:
        raise DeniseError("dn_timedomain_filt failed")
    return d


def smooth2(grad, wd_damp):
    """smooth2.c (SPATFILTER 2) on a host array [ny][nx] (runs on the GPU)."""
    g = np.ascontiguousarray(grad, np.float32).copy()
    ny, nx = g.shape
    L = lib()
    L.dn_smooth2.argtypes = [C.c_int, C.c_int, C.c_float, C.POINTER(C.c_float)]
    if L.dn_smooth2(nx, ny, wd_damp, _f(g)) != 0:
        raise DeniseError(L.dn_last_error(None).decode())
    return g


def gauss_filt(grad, dh, wd_damp, fc_end, c_vs):
    """gauss_filt.c (SPATFILTER 3) on a host array [ny][nx] (runs on the GPU)."""
    g = np.ascontiguousarray(grad, np.float32).copy()
    ny, nx = g.shape
    L = lib()
    L.dn_gauss_filt.argtypes = [C.c_int, C.c_int] + [C.c_float] * 4 + [C.POINTER(C.c_float)]
    if L.dn_gauss_filt(nx, ny, dh, wd_damp, fc_end, c_vs, _f(g)) != 0:
        raise DeniseError(L.dn_last_error(None).decode())
    return g


def gauss_filt_var(grad, vel, dh, wd_damp, wd_damp1, fc_end):
    """gauss_filt_var.c on host arrays [ny][nx] (runs on the GPU)."""
    g = np.ascontiguousarray(grad, np.float32).copy()
    v = np.ascontiguousarray(vel, np.float32)
    ny, nx = g.shape
    assert v.shape == g.shape
    L = lib()
    L.dn_gauss_filt_var.argtypes = [C.c_int, C.c_int] + [C.c_float] * 4 + [C.POINTER(C.c_float)] * 2
    if L.dn_gauss_filt_var(nx, ny, dh, wd_damp, wd_damp1, fc_end, _f(v), _f(g)) != 0:
        raise DeniseError(L.dn_last_error(None).decode())
    return g


def taper_grad(grad, sws, dh, src_i=(), src_j=(), rec_i=(), rec_j=(), **kw):
    """taper_grad.c on a host array [ny][nx] (runs on the GPU); returns
    (tapered gradient, taper coefficients)."""
    g = np.ascontiguousarray(grad, np.float32).copy()
    ny, nx = g.shape
    tap = np.zeros_like(g)
    t, keep = _taper(sws=sws, dh=dh, **kw)
    si, sj = np.ascontiguousarray(src_i, np.int32), np.ascontiguousarray(src_j, np.int32)
    ri, rj = np.ascontiguousarray(rec_i, np.int32), np.ascontiguousarray(rec_j, np.int32)
    L = lib()
    ip = C.POINTER(C.c_int)
    L.dn_taper_grad_host.argtypes = [C.c_int, C.c_int, C.POINTER(Taper), C.c_int, ip, ip, C.c_int, ip, ip,
                                     C.POINTER(C.c_float), C.POINTER(C.c_float)]
    if L.dn_taper_grad_host(nx, ny, C.byref(t), len(si), _i(si), _i(sj), len(ri), _i(ri), _i(rj), _f(g),
                            _f(tap)) != 0:
        raise DeniseError(L.dn_last_error(None).decode())
    return g, tap


def time_window(data, dt, timewin, gamma, twlength_plus=0.0, picks=None):
    """time_window.c on data [ntr][ns] (host); returns the windowed copy."""
    d = np.ascontiguousarray(data, np.float32).copy()
    ntr, ns = d.shape
    p = None if picks is None else np.ascontiguousarray(picks, np.float32)
    L = lib()
    L.dn_time_window.argtypes = [C.c_int, C.c_int, C.c_float, C.c_int, C.c_float, C.c_float,
                                 C.POINTER(C.c_float), C.POINTER(C.c_float)]
    if L.dn_time_window(ntr, ns, dt, timewin, gamma, twlength_plus, _f(p), _f(d)) != 0:
        raise DeniseError(L.dn_last_error(None).decode())
    return d


def calc_res(syn, obs, dt):
    """calc_res.c + calc_energy.c on one shot's host seismograms [ntr][ns]
    (runs on the GPU): returns (time-reversed residual, l2, energy, obs with
    obs[:, 0] = syn[:, 0] as the reference leaves it)."""
    s = np.ascontiguousarray(syn, np.float32)
    o = np.ascontiguousarray(obs, np.float32).copy()
    ntr, ns = s.shape
    diff = np.zeros_like(s)
    l2, en = C.c_double(), C.c_double()
    L = lib()
    L.dn_calc_res.argtypes = [C.c_int, C.c_int, C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float),
                              C.POINTER(C.c_float), C.POINTER(C.c_double), C.POINTER(C.c_double)]
    if L.dn_calc_res(ntr, ns, dt, _f(s), _f(o), _f(diff), C.byref(l2), C.byref(en)) != 0:
        raise DeniseError(L.dn_last_error(None).decode())
    return diff, l2.value, en.value, o


def stf(obs, syn, signal, mute=None, eps_stf=0.05, normalize=0):
    """stf.c's Wiener deconvolution on host arrays (runs on the GPU)."""
    o = np.ascontiguousarray(obs, np.float32)
    m = np.ascontiguousarray(syn, np.float32)
    s = np.ascontiguousarray(signal, np.float32)
    ntr, ns = o.shape
    mu = None if mute is None else np.ascontiguousarray(mute, np.uint8)
    out = np.zeros(ns, np.float32)
    opts = StfOpts(eps_stf, 0.0, normalize)
    L = lib()
    L.dn_stf.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                         C.POINTER(C.c_ubyte), C.POINTER(StfOpts), C.POINTER(C.c_float)]
    mp = None if mu is None else mu.ctypes.data_as(C.POINTER(C.c_ubyte))
    if L.dn_stf(ntr, ns, _f(o), _f(m), _f(s), mp, C.byref(opts), _f(out)) != 0:
        raise DeniseError(L.dn_last_error(None).decode())
    return out


class CheckFD(C.Structure):
    _fields_ = [("cmin", C.c_float), ("cmax", C.c_float), ("dtstab", C.c_float), ("dhstab", C.c_float),
                ("dt_unstable", C.c_int), ("dh_dispersive", C.c_int), ("fw_too_large", C.c_int),
                ("fw_small", C.c_int)]


def checkfd(vs, rho, qs, fdorder, maxrelerror, fl, dt, dh, ts, fw):
    """checkfd_ssg_visc.c / checkfd_ssg_elastic.c: returns (status, dict)."""
    vs = np.ascontiguousarray(vs, np.float32)
    rho = np.ascontiguousarray(rho, np.float32)
    L = len(fl)
    flv = np.ascontiguousarray(fl if L else [0.0], np.float32)
    qsv = np.ascontiguousarray(qs if qs is not None else np.zeros(1), np.float32)
    ny, nx = vs.shape
    f = C.c_float
    fn = lib().dn_checkfd
    fn.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float), f, f, f, C.c_int,
                   C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(CheckFD)]
    out = CheckFD()
    rc = fn(nx, ny, fdorder, maxrelerror, L, _f(flv), dt, dh, ts, fw, _f(vs), _f(rho), _f(qsv), C.byref(out))
    if rc < 0:
        raise DeniseError("dn_checkfd: bad arguments")
    return rc, {k: getattr(out, k) for k, _ in CheckFD._fields_}


class UNet:
    """The learned diagonal-Hessian preconditioner network (dn_unet_*)."""

    def __init__(self, seed=1):
        self._h = lib().dn_unet_create(seed)
        if not self._h:
            raise DeniseError(lib().dn_last_error(None).decode())
        self.nparams = lib().dn_unet_num_params(self._h)

    def _chk(self, rc):
        if rc != 0:
            raise DeniseError(lib().dn_last_error(None).decode())

    def close(self):
        if self._h:
            lib().dn_unet_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        p = np.zeros(self.nparams, np.float32)
        self._chk(lib().dn_unet_get_params(self._h, _f(p)))
        return p

    def set_params(self, p):
        self._chk(lib().dn_unet_set_params(self._h, _f(np.ascontiguousarray(p, np.float32))))

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        n, c, h, w = x.shape
        y = np.zeros((n, h, w), np.float32)
        self._chk(lib().dn_unet_forward(self._h, n, h, w, _f(x), _f(y)))
        return y

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        n, c, h, w = x.shape
        g = np.zeros(self.nparams, np.float32)
        loss = C.c_double()
        self._chk(lib().dn_unet_loss_grad(self._h, n, h, w, _f(x), _f(t), C.byref(loss), _f(g)))
        return loss.value, g

    def train_step(self, x, t, lr):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        n, c, h, w = x.shape
        loss = C.c_double()
        self._chk(lib().dn_unet_train_step(self._h, n, h, w, _f(x), _f(t), lr, C.byref(loss)))
        return loss.value
