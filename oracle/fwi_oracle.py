"""CPU restatement of DENISE-SH's FWI iteration with L-BFGS (GRAD_METHOD 2)
on top of the oracle library (oracle/sh_oracle.c).

TEST INFRASTRUCTURE ONLY: used by tests/ to check the device FWI
loop (dn_fwi_iteration).  Never by the product path.

One iteration follows denise.c:695-2018:
  * gradient of all shots at the current model with the normalisation means
    C_vs, C_rho of the starting model (denise.c:745-816), after eprecond1
    and the chain rule (orc_gradient);
  * L2t[1] = L2t[4] = sum of the test shots' misfits / their energies
    (calc_res.c:255-257, calc_energy.c:57-61, denise.c:1715-1723);
  * smooth_grad / taper_grad when switched on (denise.c:1818-1847;
    orc_gauss_filt_var, orc_taper_grad);
  * GRAD_METHOD 2, store_LBFGS_SH.c:39-120: waveconv *= C, y/s pairs of
    [rho; u] in the slot of LBFGS_pointer; LBFGS.c:26-95 two-loop in slot
    order (orc_lbfgs); extract_LBFGS_SH.c:31-64: direction = C * r for
    iter > 1, else C * g;
  * GRAD_METHOD 1, denise.c:1850-1870: descent.c negates the gradient,
    store_PCG_SH.c stacks [u; rho], PCG.c forms the Polak-Ribiere direction
    per class (orc_pcg, beta reset at < 0), extract_PCG_SH.c and descent.c
    give it back as a gradient, the negated gradient is kept as PCG_old;
  * line_search.c:60-380: trial models by calc_mat_change_test.c (itest 1)
    at eps_scale, forward of the test shots, the step1/step2/step3 state
    machine, calc_opt_step.c:13-109 for the final step;
  * (iter > min_iter_help) and step1 == 0: no update (denise.c:1916-1921);
    calc_mat_change_test with itest 0 commits the model;
  * the stage abort rule of denise.c:1982-2018.
The product sums misfits and L-BFGS dot products in FP64 where the reference
sums in float; the oracle library's misfit sums are FP64 as well, its
L-BFGS dots float (as the reference).  Vectors are ordered row by row here
and x-major in the reference; the two-loop result does not depend on the
order beyond rounding.

Parity: pinned through its parts (each oracle function is pinned in
tests/test_oracle_pins.py) and, as a composition, by
tests/test_fwi_oracle_pins.py: iteration 1 steps along minus the normalised
gradient, iteration 2 along the one-pair BFGS (or Polak-Ribiere) direction
built from independently computed gradients, and the misfit falls.  It is
not run against the reference itself, whose FWI loop needs MPI and cannot
be built here.
"""
import ctypes as C

import numpy as np


def float_mean(a):
    """denise.c:778-816: float sum over i (outer) and j (inner), / (NX*NY)."""
    s = np.cumsum(np.ascontiguousarray(a, np.float32).T.ravel(), dtype=np.float32)[-1]
    return np.float32(s / np.float32(a.size))


class FwiOracle:
    def __init__(self, O, cfg, vs, rho, qs, src_i, src_j, signals, rec_i, rec_j, obs, **opts):
        assert O.kind == "f32"
        self.O, self.cfg = O, cfg
        self.vs = np.ascontiguousarray(vs, np.float32).copy()
        self.rho = np.ascontiguousarray(rho, np.float32).copy()
        self.qs = None if qs is None else np.ascontiguousarray(qs, np.float32)
        self.src_i, self.src_j = np.asarray(src_i, np.int32), np.asarray(src_j, np.int32)
        self.sig = np.ascontiguousarray(signals, np.float32)
        self.rec_i, self.rec_j = np.asarray(rec_i, np.int32), np.asarray(rec_j, np.int32)
        self.obs = np.ascontiguousarray(obs, np.float32)
        self.sig0, self.obs0 = self.sig.copy(), self.obs.copy()
        o = dict(nlbfgs=20, eps_scale=0.01, stepmax=6, scalefac=2.0, testshot_start=1, testshot_end=1,
                 testshot_incr=1, vs_lower=0.0, vs_upper=1e9, rho_lower=0.0, rho_upper=1e9, inv_vs_iter=0,
                 inv_rho_iter=0, scalerho=0.5, min_iter=0, pro=0.0, spatfilter=0, wd_damp=0.0, wd_damp1=0.0,
                 fc_end=0.0, taper_vert=0, gradt=(0, 0, 0, 0), grad_method=2)
        o.update(opts)
        self.o = o
        cfg.c_vs = float(float_mean(self.vs))
        cfg.c_rho = float(float_mean(self.rho))
        self.C_vs, self.C_rho = np.float32(cfg.c_vs), np.float32(cfg.c_rho)
        n = self.vs.size
        self.n = n
        NL = o["nlbfgs"]
        self.y = np.zeros((NL, 2 * n), np.float32)
        self.s = np.zeros((NL, 2 * n), np.float32)
        self.ptr = 1
        self.iter = 1
        self.pcg_old = np.zeros((2, n), np.float32)      # [u; rho], zero_PCG
        self.pcg_dir = np.zeros((2, n), np.float32)
        self.min_iter_help = 0           # denise.c:138
        self.L2_hist = {}
        self.gold = self.mold = None
        inc = max(1, o["testshot_incr"])
        self.tshots = [k - 1 for k in range(o["testshot_start"], min(o["testshot_end"], len(self.src_i)) + 1, inc)
                       if k >= 1]

    def start_stage(self, time_filt=0, fc_start=0.0, order=0, **opts):
        """A new workflow stage (denise.c:682-689 and the optimiser reset of
        denise.c:1998-2005): the stage's options replace the old ones, and
        the wavelets and observed data are the originals low-passed at the
        stage's FC = FC_END with ORDER (TIME_FILT 1), then high-passed at
        FC_START (TIME_FILT 2), by timedomain_filt (denise.c:1099-1111,
        1381-1394)."""
        self.o.update(opts)
        self.sig, self.obs = self.sig0.copy(), self.obs0.copy()
        if time_filt:
            fc = self.o["fc_end"]
            self.sig = self.O.timedomain_filt(self.sig, self.cfg.dt, fc, order, 1)
            self.obs = np.stack([self.O.timedomain_filt(o, self.cfg.dt, fc, order, 1) for o in self.obs])
            if time_filt == 2:
                self.sig = self.O.timedomain_filt(self.sig, self.cfg.dt, fc_start, order, 2)
                self.obs = np.stack([self.O.timedomain_filt(o, self.cfg.dt, fc_start, order, 2) for o in self.obs])
        self.y[:] = 0.0
        self.s[:] = 0.0
        self.ptr = 1
        self.iter = 1
        self.pcg_old[:] = 0.0
        self.pcg_dir[:] = 0.0

    # test-shot misfit of a model (forward + residual, FP64 sums)
    def _misfit(self, vs, rho):
        l2 = C.c_double(0.0)
        en = C.c_double(0.0)
        f = self.O.fn("orc_residual")
        rp = self.O.rp
        for k in self.tshots:
            seis, _, _ = self.O.forward(self.cfg, vs, rho, self.qs, self.src_i[k:k + 1], self.src_j[k:k + 1],
                                        self.sig[k:k + 1], self.rec_i, self.rec_j)
            seis = np.ascontiguousarray(seis, np.float32)
            ob = self.obs[k].copy()
            diff = np.zeros_like(ob)
            f(len(self.rec_i), self.cfg.nt, C.c_double(self.cfg.dt), seis.ctypes.data_as(rp), ob.ctypes.data_as(rp),
              diff.ctypes.data_as(rp), C.byref(l2), C.byref(en))
        return l2.value, en.value

    def _mat_change(self, gu, grho, eps_scale):
        o = self.o
        f = self.O.fn("orc_calc_mat_change")
        f.restype = C.c_double
        rp = self.O.rp
        un = np.zeros_like(self.vs)
        rn = np.zeros_like(self.rho)
        er = C.c_float()
        ny, nx = self.vs.shape
        gu = np.ascontiguousarray(gu, np.float32)
        grho = np.ascontiguousarray(grho, np.float32)
        eu = f(nx, ny, gu.ctypes.data_as(rp), grho.ctypes.data_as(rp), self.vs.ctypes.data_as(rp),
               self.rho.ctypes.data_as(rp), un.ctypes.data_as(rp), rn.ctypes.data_as(rp),
               C.c_double(float(np.float32(eps_scale))), self.iter, o["inv_vs_iter"], o["inv_rho_iter"],
               C.c_double(float(np.float32(o["scalerho"]))), C.c_double(float(np.float32(o["vs_lower"]))),
               C.c_double(float(np.float32(o["vs_upper"]))), C.c_double(float(np.float32(o["rho_lower"]))),
               C.c_double(float(np.float32(o["rho_upper"]))), C.byref(er))
        return un, rn, eu, er.value

    def _line_search(self, gu, grho, L2t, epst1, energy):
        o = self.o
        f32 = np.float32
        eps_scale, scalefac = f32(o["eps_scale"]), f32(o["scalefac"])
        step1 = step2 = step3 = 0
        countstep, itests, iteste = 0, 2, 2
        while step2 != 1 or step1 != 1:
            for itest in range(itests, iteste + 1):
                un, rn, _, _ = self._mat_change(gu, grho, eps_scale)
                l2, _ = self._misfit(un, rn)
                epst1[itest] = eps_scale
                epst1[1] = f32(0.0)
                L2t[itest] = f32(l2 / energy)
            if step1 == 0 and L2t[1] <= L2t[2]:
                eps_scale = f32(eps_scale / scalefac)
                countstep += 1
            if step1 == 1 and L2t[2] < L2t[3]:
                epst1[3] = eps_scale
                step2 = 1
            if step1 == 1 and L2t[2] >= L2t[3]:
                epst1[3] = eps_scale
                eps_scale = f32(eps_scale + f32(eps_scale / scalefac))
                countstep += 1
            if step1 == 0 and L2t[1] > L2t[2]:
                epst1[2] = eps_scale
                step1 = 1
                iteste = itests = 3
                countstep = 0
                eps_scale = f32(eps_scale + f32(eps_scale / scalefac))
            step3 = 0
            if step1 == 0 and countstep > o["stepmax"]:
                step3 = 1
                break
            if step1 == 1 and countstep > o["stepmax"]:
                step1 = 1
                step2 = 1
        if step1 == 1:
            eps_scale = f32(self.O.calc_opt_step(L2t[1:4], epst1[1:4]))
        return eps_scale, step1, step3

    def iteration(self):
        """One iteration; returns a record like dn_fwi_iteration's."""
        o, O, n = self.o, self.O, self.n
        if self.iter > 2:
            self.ptr += 1
        if self.ptr > o["nlbfgs"]:
            self.ptr = 1
        g = O.gradient(self.cfg, self.vs, self.rho, self.qs, self.src_i, self.src_j, self.sig, self.rec_i,
                       self.rec_j, self.obs)
        gvs = np.asarray(g["grad_vs"], np.float32)
        grho = np.asarray(g["grad_rho"], np.float32)
        l2, energy = self._misfit(self.vs, self.rho)
        L2t = np.zeros(5, np.float32)
        epst1 = np.zeros(4, np.float32)
        L2t[1] = L2t[4] = np.float32(l2 / energy)
        if o["spatfilter"] == 4:
            gvs = O.gauss_filt_var(gvs, self.vs, self.cfg.dh, o["wd_damp"], o["wd_damp1"], o["fc_end"])
            grho = O.gauss_filt_var(grho, self.vs, self.cfg.dh, o["wd_damp"], o["wd_damp1"], o["fc_end"])
        if o["taper_vert"]:
            kw = dict(gradt=o["gradt"], src_i=self.src_i, src_j=self.src_j, rec_i=self.rec_i, rec_j=self.rec_j)
            gvs, _ = O.taper_grad(gvs, 1, self.cfg.dh, **kw)
            grho, _ = O.taper_grad(grho, 1, self.cfg.dh, **kw)
        if o["grad_method"] == 1:
            new = np.stack([-gvs.ravel(), -grho.ravel()]).astype(np.float32)
            self.pcg_dir = np.asarray(O.pcg(new, self.pcg_old, self.pcg_dir), np.float32)
            self.pcg_old = new
            du = (-self.pcg_dir[0]).reshape(self.vs.shape)
            dr = (-self.pcg_dir[1]).reshape(self.vs.shape)
            return self._finish(du, dr, L2t, epst1, energy)
        # store_LBFGS_SH.c: normalisation
        gu = (self.C_vs * gvs).astype(np.float32).ravel()
        gr = (self.C_rho * grho).astype(np.float32).ravel()
        if self.iter > 1:
            k = self.ptr - 1
            self.y[k, :n] = gr - self.gold[:n]
            self.y[k, n:] = gu - self.gold[n:]
            self.s[k, :n] = self.rho.ravel() - self.mold[:n]
            self.s[k, n:] = self.vs.ravel() - self.mold[n:]
            r = O.lbfgs(self.y, self.s, np.concatenate([gr, gu]), self.ptr)
        self.gold = np.concatenate([gr, gu])
        self.mold = np.concatenate([self.rho.ravel(), self.vs.ravel()])
        if self.iter > 1:
            gr = (np.asarray(r[:n], np.float32) * self.C_rho).astype(np.float32)
            gu = (np.asarray(r[n:], np.float32) * self.C_vs).astype(np.float32)
        du, dr = gu.reshape(self.vs.shape), gr.reshape(self.vs.shape)
        return self._finish(du, dr, L2t, epst1, energy)

    def _finish(self, du, dr, L2t, epst1, energy):
        """Line search, model update and the stage rule of one iteration."""
        o = self.o
        eps_scale, step1, step3 = self._line_search(du, dr, L2t, epst1, energy)
        if self.iter > self.min_iter_help and step1 == 0:
            eps_scale = np.float32(0.0)
        un, rn, eps_u, eps_rho = self._mat_change(du, dr, eps_scale)
        self.vs, self.rho = un, rn
        rec = dict(iter=self.iter, L2t=[float(v) for v in L2t[1:5]], epst1=[float(v) for v in epst1[1:4]],
                   step1=step1, step3=step3, opteps_vs=float(eps_scale), eps_u=eps_u, eps_rho=eps_rho,
                   next_stage=0)
        self.L2_hist[self.iter] = L2t[4]
        if self.iter > self.min_iter_help:
            k = self.iter - 2
            if k >= 1 or step3 == 1:
                prev = self.L2_hist.get(k, np.float32(0.0)) if k >= 1 else np.float32(0.0)
                diff = abs((prev - L2t[4]) / prev) if k >= 1 else 0.0
                if (k >= 1 and diff <= o["pro"]) or step3 == 1:
                    self.min_iter_help = self.iter + o["min_iter"]
                    self.iter = 0
                    rec["next_stage"] = 1
                    # denise.c:1998-2005: zero_PCG, or zero_LBFGS and LBFGS_pointer = 1
                    self.pcg_old[:] = 0.0
                    self.pcg_dir[:] = 0.0
                    self.y[:] = 0.0
                    self.s[:] = 0.0
                    self.ptr = 1
        self.iter += 1
        return rec
