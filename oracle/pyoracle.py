"""ctypes binding of the CPU oracle (oracle/sh_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() to
check its result and by bench.py's cpu_baseline leg.  Never by the product path.

`Oracle("f32")` reproduces the reference's FP32 arithmetic, `Oracle("f64")`
is the FP64 oracle.  Arrays follow sh_oracle.c: padded fields of shape
(NY+2nd, NX+2nd) with nd = FDORDER/2+1 (index [j+nd-1, i+nd-1]), or plain
[ny][nx] row-major model inputs.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


class Cfg(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("fdorder", C.c_int), ("maxrelerror", C.c_int), ("fw", C.c_int),
                ("free_surf", C.c_int), ("L", C.c_int), ("nt", C.c_int), ("dtinv", C.c_int),
                ("dt", C.c_double), ("dh", C.c_double), ("damping", C.c_double), ("fpml", C.c_double),
                ("npower", C.c_double), ("k_max_pml", C.c_double), ("fl", C.c_double * 8), ("eprecond", C.c_int),
                ("c_vs", C.c_double), ("c_rho", C.c_double), ("skip_res", C.c_void_p), ("skip_en", C.c_void_p),
                ("periodic", C.c_int), ("snap_bf16", C.c_int)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def make_cfg(nx, ny, fdorder=8, maxrelerror=1, fw=20, free_surf=1, L=0, nt=100, dtinv=1, dt=1e-4, dh=0.25,
             damping=300.0, fpml=40.0, npower=2.0, k_max_pml=1.0, fl=(), eprecond=1, boundary=0, snap_bf16=0):
    c = Cfg()
    c.eprecond = eprecond
    c.periodic = 1 if boundary else 0
    c.snap_bf16 = 1 if snap_bf16 else 0
    c.nx, c.ny, c.fdorder, c.maxrelerror, c.fw, c.free_surf, c.L, c.nt, c.dtinv = (
        nx, ny, fdorder, maxrelerror, fw, free_surf, L, nt, dtinv)
    # scalar parameters are float32 values in the reference (read as float)
    c.dt, c.dh, c.damping, c.fpml, c.npower, c.k_max_pml = [float(np.float32(v)) for v in
                                                            (dt, dh, damping, fpml, npower, k_max_pml)]
    for k, v in enumerate(fl):
        c.fl[k] = float(np.float32(v))
    return c


class Oracle:
    def __init__(self, kind="f32"):
        assert kind in ("f32", "f64")
        path = os.path.join(HERE, "liboracle_%s.so" % kind)
        if not os.path.exists(path):
            build()
        self.L = C.CDLL(path)
        self.kind = kind
        self.dt = np.float32 if kind == "f32" else np.float64
        self.sfx = "_" + kind
        self.rp = C.POINTER(C.c_float if kind == "f32" else C.c_double)

    def fn(self, name):
        return getattr(self.L, name + self.sfx)

    def _r(self, a):
        return None if a is None else a.ctypes.data_as(self.rp)

    @staticmethod
    def _f32(a):
        return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))

    @staticmethod
    def _i32(a):
        return a.ctypes.data_as(C.POINTER(C.c_int))

    def padded_shape(self, cfg):
        nd = cfg.fdorder // 2 + 1
        return (cfg.ny + 2 * nd, cfg.nx + 2 * nd)

    # ---- setup ----
    def holberg(self, fdorder, maxrelerror):
        hc = np.zeros(7, self.dt)
        rc = self.fn("orc_holberg")(fdorder, maxrelerror, self._r(hc))
        if rc:
            raise ValueError("bad FDORDER/MAXRELERROR")
        return hc

    def pml(self, cfg):
        n = 2 * cfg.fw + 2
        arrs = [np.zeros(n, self.dt) for _ in range(20)]
        self.fn("orc_pml_pro")(C.byref(cfg), *[self._r(a) for a in arrs])
        names = []
        for ax in ("x", "x_half", "y", "y_half"):
            for q in ("d", "K", "alpha", "a", "b"):
                names.append("%s_%s" % (q, ax))
        return dict(zip(names, arrs))

    def wavelet(self, quellart, nt, dt, tshift, fc, amp):
        out = np.zeros(nt, self.dt)
        rc = self.fn("orc_wavelet")(quellart, nt, C.c_double(dt), C.c_double(tshift), C.c_double(fc),
                                    C.c_double(amp), self._r(out))
        assert rc == 0
        return out

    def butterworth(self, x, dt, t0, order, method):
        x = np.ascontiguousarray(x, np.float64).copy()
        self.fn("orc_seife_butterworth")(x.ctypes.data_as(C.POINTER(C.c_double)), len(x), C.c_double(dt),
                                         C.c_double(t0), order, method)
        return x

    def timedomain_filt(self, data, dt, fc, order, method):
        d = np.ascontiguousarray(data, np.float32).copy()
        ntr, ns = d.shape
        self.fn("orc_timedomain_filt")(self._f32(d), ntr, ns, C.c_double(dt), C.c_double(fc), order, method)
        return d

    def dtstab(self, cfg, vs, rho, qs, ts):
        f = self.fn("orc_dtstab")
        f.restype = C.c_double
        return f(C.byref(cfg), self._f32(np.ascontiguousarray(vs, np.float32)),
                 self._f32(np.ascontiguousarray(rho, np.float32)),
                 self._f32(None if qs is None else np.ascontiguousarray(qs, np.float32)), C.c_double(ts))

    # ---- composed simulation ----
    def forward(self, cfg, vs, rho, qs, src_i, src_j, signals, rec_i, rec_j, want_snaps=False):
        vs = np.ascontiguousarray(vs, np.float32)
        rho = np.ascontiguousarray(rho, np.float32)
        qs = None if qs is None else np.ascontiguousarray(qs, np.float32)
        src_i = np.ascontiguousarray(src_i, np.int32)
        src_j = np.ascontiguousarray(src_j, np.int32)
        rec_i = np.ascontiguousarray(rec_i, np.int32)
        rec_j = np.ascontiguousarray(rec_j, np.int32)
        sig = np.ascontiguousarray(signals, np.float32).reshape(len(src_i), cfg.nt)
        seis = np.zeros((len(rec_i), cfg.nt), self.dt)
        snaps = Ws = None
        if want_snaps:
            ntd = -(-cfg.nt // cfg.dtinv)
            snaps = np.zeros((ntd, 3, cfg.ny, cfg.nx), self.dt)
            Ws = np.zeros(self.padded_shape(cfg), self.dt)
        rc = self.fn("orc_forward")(C.byref(cfg), self._f32(vs), self._f32(rho), self._f32(qs), len(src_i),
                                    self._i32(src_i), self._i32(src_j), self._f32(sig), len(rec_i),
                                    self._i32(rec_i), self._i32(rec_j), self._r(seis), self._r(snaps), self._r(Ws))
        if rc:
            raise RuntimeError("oracle forward failed (%d)" % rc)
        return seis, snaps, Ws

    def gradient(self, cfg, vs, rho, qs, src_i, src_j, signals, rec_i, rec_j, obs):
        vs = np.ascontiguousarray(vs, np.float32)
        rho = np.ascontiguousarray(rho, np.float32)
        qs = None if qs is None else np.ascontiguousarray(qs, np.float32)
        src_i = np.ascontiguousarray(src_i, np.int32)
        src_j = np.ascontiguousarray(src_j, np.int32)
        rec_i = np.ascontiguousarray(rec_i, np.int32)
        rec_j = np.ascontiguousarray(rec_j, np.int32)
        sig = np.ascontiguousarray(signals, np.float32)
        # observed data in this oracle's precision (FP64 oracle: float64 data
        # keep the whole chain in double), else as float32
        native = self.kind == "f64" and np.asarray(obs).dtype == np.float64
        obs = np.ascontiguousarray(obs, np.float64 if native else np.float32)
        gvs = np.zeros((cfg.ny, cfg.nx), self.dt)
        grho = np.zeros((cfg.ny, cfg.nx), self.dt)
        wu = np.zeros((cfg.ny, cfg.nx), self.dt)
        wr = np.zeros((cfg.ny, cfg.nx), self.dt)
        l2 = C.c_double()
        en = C.c_double()
        rc = self.fn("orc_gradient_obsreal" if native else "orc_gradient")(
                                     C.byref(cfg), self._f32(vs), self._f32(rho), self._f32(qs), len(src_i),
                                     self._i32(src_i), self._i32(src_j), self._f32(sig), len(rec_i),
                                     self._i32(rec_i), self._i32(rec_j), self._r(obs) if native else self._f32(obs),
                                     self._r(gvs),
                                     self._r(grho), C.byref(l2), C.byref(en), self._r(wu), self._r(wr))
        if rc:
            raise RuntimeError("oracle gradient failed (%d)" % rc)
        return dict(grad_vs=gvs, grad_rho=grho, raw_u=wu, raw_rho=wr, l2=l2.value, energy=en.value)

    def pseudo_hessian(self, cfg, vs, rho, qs, src_i, src_j, signals):
        vs = np.ascontiguousarray(vs, np.float32)
        rho = np.ascontiguousarray(rho, np.float32)
        qs = None if qs is None else np.ascontiguousarray(qs, np.float32)
        src_i = np.ascontiguousarray(src_i, np.int32)
        src_j = np.ascontiguousarray(src_j, np.int32)
        sig = np.ascontiguousarray(signals, np.float32)
        hv = np.zeros((cfg.ny, cfg.nx), self.dt)
        hr = np.zeros((cfg.ny, cfg.nx), self.dt)
        rc = self.fn("orc_pseudo_hessian")(C.byref(cfg), self._f32(vs), self._f32(rho), self._f32(qs), len(src_i),
                                           self._i32(src_i), self._i32(src_j), self._f32(sig), self._r(hv),
                                           self._r(hr))
        if rc:
            raise RuntimeError("oracle pseudo-Hessian failed (%d)" % rc)
        return hv, hr

    # ---- single sweeps (padded arrays, shapes from padded_shape) ----
    def snap_bf16(self, snaps):
        """BF16 snapshot storage of a [nsnap][3][ny][nx] series, in place (orc_snap_bf16)."""
        nsnap = snaps.shape[0]
        nxny = int(np.prod(snaps.shape[2:]))
        assert snaps.dtype == self.dt and snaps.flags.c_contiguous
        self.fn("orc_snap_bf16")(nsnap, C.c_size_t(nxny), self._r(snaps))

    def wrap_x(self, cfg, a):
        """BOUNDARY 1 halo exchange of one padded field (exchange_v / exchange_s on one domain)."""
        self.fn("orc_wrap_x")(C.byref(cfg), self._r(a))

    def update_v(self, cfg, sw, vy, vyp1, syz, sxy, rho, hc, pml, psi_syy_y, psi_sxy_x):
        self.fn("orc_update_v")(C.byref(cfg), sw, self._r(vy), self._r(vyp1), self._r(syz), self._r(sxy),
                                self._r(rho), self._r(hc), self._r(pml["K_x"]), self._r(pml["a_x"]),
                                self._r(pml["b_x"]), self._r(pml["K_y"]), self._r(pml["a_y"]), self._r(pml["b_y"]),
                                self._r(psi_syy_y), self._r(psi_sxy_x))

    def update_s_elastic(self, cfg, vy, uy, uyx, syz, sxy, uip, ujp, hc, pml, psi_vyy, psi_vyx):
        self.fn("orc_update_s_elastic")(C.byref(cfg), self._r(vy), self._r(uy), self._r(uyx), self._r(syz),
                                        self._r(sxy), self._r(uip), self._r(ujp), self._r(hc),
                                        self._r(pml["K_x_half"]), self._r(pml["a_x_half"]), self._r(pml["b_x_half"]),
                                        self._r(pml["K_y_half"]), self._r(pml["a_y_half"]), self._r(pml["b_y_half"]),
                                        self._r(psi_vyy), self._r(psi_vyx))

    def update_s_visc(self, cfg, vy, uy, uyx, syz, sxy, r, q, fipjp, f, dip, d, bip, bjm, cip, cjm, hc, pml,
                      psi_vyy, psi_vyx):
        self.fn("orc_update_s_visc")(C.byref(cfg), self._r(vy), self._r(uy), self._r(uyx), self._r(syz),
                                     self._r(sxy), self._r(r), self._r(q), self._r(fipjp), self._r(f), self._r(dip),
                                     self._r(d), self._r(bip), self._r(bjm), self._r(cip), self._r(cjm),
                                     self._r(hc), self._r(pml["K_x_half"]), self._r(pml["a_x_half"]),
                                     self._r(pml["b_x_half"]), self._r(pml["K_y"]), self._r(pml["a_y"]),
                                     self._r(pml["b_y"]), self._r(pml["K_y_half"]), self._r(pml["a_y_half"]),
                                     self._r(pml["b_y_half"]), self._r(psi_vyy), self._r(psi_vyx))

    def model_setup(self, cfg, vs, rho, qs=None):
        """matcopy + av_mue (+ av_tau + prepare_update_s): padded fields."""
        sh = self.padded_shape(cfg)
        nd = cfg.fdorder // 2 + 1
        P = {k: np.zeros(sh, self.dt) for k in ("vs", "rho", "taus", "uip", "ujp", "tausipjp", "fipjp", "f")}
        P["vs"][nd:nd + cfg.ny, nd:nd + cfg.nx] = np.asarray(vs, np.float32)
        P["rho"][nd:nd + cfg.ny, nd:nd + cfg.nx] = np.asarray(rho, np.float32)
        if cfg.L > 0:
            P["taus"][nd:nd + cfg.ny, nd:nd + cfg.nx] = (2.0 / np.asarray(qs, np.float64)).astype(np.float32)
        for k in ("rho", "vs", "taus"):
            self.fn("orc_matcopy_wrap")(C.byref(cfg), self._r(P[k]))
        self.fn("orc_av_mue")(C.byref(cfg), self._r(P["vs"]), self._r(P["rho"]), self._r(P["uip"]),
                              self._r(P["ujp"]))
        Lm = max(cfg.L, 1)
        P["dip"] = np.zeros((Lm,) + sh, self.dt)
        P["d"] = np.zeros((Lm,) + sh, self.dt)
        for k in ("bip", "bjm", "cip", "cjm"):
            P[k] = np.zeros(8, self.dt)
        if cfg.L > 0:
            self.fn("orc_av_tau")(C.byref(cfg), self._r(P["taus"]), self._r(P["tausipjp"]))
            self.fn("orc_prepare_update_s")(C.byref(cfg), self._r(P["ujp"]), self._r(P["uip"]), self._r(P["taus"]),
                                            self._r(P["tausipjp"]), self._r(P["fipjp"]), self._r(P["f"]),
                                            self._r(P["dip"]), self._r(P["d"]), self._r(P["bip"]),
                                            self._r(P["bjm"]), self._r(P["cip"]), self._r(P["cjm"]))
        return P

    # ---- optimiser pieces ----
    def calc_opt_step(self, L2t, epst):
        f = self.fn("orc_calc_opt_step")
        f.restype = C.c_double
        a = np.ascontiguousarray(L2t, self.dt)
        b = np.ascontiguousarray(epst, self.dt)
        return f(self._r(a), self._r(b))

    def lbfgs(self, y, s, q, cur):
        y = np.ascontiguousarray(y, self.dt)
        s = np.ascontiguousarray(s, self.dt)
        q = np.ascontiguousarray(q, self.dt).copy()
        r = np.zeros_like(q)
        nslots, n = y.shape
        self.fn("orc_lbfgs")(nslots, n, self._r(y), self._r(s), self._r(q), self._r(r), cur)
        return r

    def pcg(self, gnew, gold, d):
        gnew = np.ascontiguousarray(gnew, self.dt)
        gold = np.ascontiguousarray(gold, self.dt)
        d = np.ascontiguousarray(d, self.dt).copy()
        ncl, n = gnew.shape
        self.fn("orc_pcg")(ncl, n, self._r(gnew), self._r(gold), self._r(d))
        return d

    # ---- gradient post-processing ([ny][nx] arrays) ----
    def gauss_filt_var(self, grad, vel, dh, wd_damp, wd_damp1, fc_end):
        g = np.ascontiguousarray(grad, self.dt).copy()
        v = np.ascontiguousarray(vel, np.float32)
        ny, nx = g.shape
        f32 = lambda x: float(np.float32(x))
        self.fn("orc_gauss_filt_var")(nx, ny, C.c_double(f32(dh)), C.c_double(f32(wd_damp)),
                                      C.c_double(f32(wd_damp1)), C.c_double(f32(fc_end)), self._f32(v), self._r(g))
        return g

    def smooth2(self, grad, wd_damp):
        """smooth2.c (SPATFILTER 2), float as the reference; returns the filtered [ny][nx] gradient."""
        g = np.ascontiguousarray(grad, np.float32).copy()
        ny, nx = g.shape
        rc = self.fn("orc_smooth2")(nx, ny, C.c_double(float(np.float32(wd_damp))), self._f32(g))
        if rc:
            raise ValueError("orc_smooth2: grid too small")
        return g

    def gauss_filt(self, grad, dh, wd_damp, fc_end, c_vs):
        """gauss_filt.c (SPATFILTER 3); returns (filtered gradient, status 0 / 1 none / -1 rejected)."""
        g = np.ascontiguousarray(grad, self.dt).copy()
        ny, nx = g.shape
        f32 = lambda x: float(np.float32(x))
        rc = self.fn("orc_gauss_filt")(nx, ny, C.c_double(f32(dh)), C.c_double(f32(wd_damp)),
                                       C.c_double(f32(fc_end)), C.c_double(f32(c_vs)), self._r(g))
        return g, rc

    def taper_grad(self, grad, sws, dh, gradt=(0, 0, 0, 0), srtshape=1, srtradius=1.0, filtsize=1,
                   src_i=(), src_j=(), rec_i=(), rec_j=()):
        """Returns (tapered gradient, taper coefficients)."""
        g = np.ascontiguousarray(grad, self.dt).copy()
        ny, nx = g.shape
        tap = np.zeros((ny, nx), np.float32)
        gt = np.ascontiguousarray(gradt, np.int32)
        si, sj = np.ascontiguousarray(src_i, np.int32), np.ascontiguousarray(src_j, np.int32)
        ri, rj = np.ascontiguousarray(rec_i, np.int32), np.ascontiguousarray(rec_j, np.int32)
        f32 = lambda x: float(np.float32(x))
        rc = self.fn("orc_taper_grad")(nx, ny, C.c_double(f32(dh)), sws, self._i32(gt), srtshape,
                                       C.c_double(f32(srtradius)), filtsize, len(si), self._i32(si), self._i32(sj),
                                       len(ri), self._i32(ri), self._i32(rj), self._r(g), self._f32(tap))
        if rc:
            raise RuntimeError("oracle taper_grad: unknown sws %d" % sws)
        return g, tap

    def time_window(self, data, dt, timewin, gamma, twlength_plus=0.0, picks=None):
        d = np.ascontiguousarray(data, self.dt).copy()
        ntr, ns = d.shape
        p = None if picks is None else np.ascontiguousarray(picks, np.float32)
        f32 = lambda x: float(np.float32(x))
        self.fn("orc_time_window")(ntr, ns, C.c_double(f32(dt)), timewin, C.c_double(f32(gamma)),
                                   C.c_double(f32(twlength_plus)), self._f32(p), self._r(d))
        return d
