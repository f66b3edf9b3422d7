"""ctypes view of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, never by the product package.  The C sources
restate rust_qmc's hot path literally; each routine cites the reference
file:line it follows.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

RNG_INIT, RNG_SINGLE, RNG_ALL, RNG_SYS, RNG_DRIFT = 1, 2, 3, 4, 5


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, a few seconds)."""
    if force or not os.path.exists(_LIB_PATH):
        args = ["make", "-s", "-C", _HERE] + (["-B"] if force else [])
        subprocess.run(args, check=True)
    return _LIB_PATH


_lib = None

_d = C.c_double
_i = C.c_int
_i64 = C.c_int64
_u32 = C.c_uint32
_u64 = C.c_uint64
_p = C.c_void_p


class McmcParams(C.Structure):
    _fields_ = [("n_walkers", _i64), ("n_steps", _i64), ("initial_step_size", _d),
                ("max_step_size", _d), ("min_step_size", _d), ("target_acceptance", _d),
                ("adaptation_interval", _i64)]


class McmcResults(C.Structure):
    _fields_ = [("energy", _d), ("error", _d), ("autocorrelation_time", _d),
                ("acceptance", _d), ("final_step", _d)]


class IsdmcParams(C.Structure):
    _fields_ = [("n_walkers", _i64), ("n_steps", _i64), ("n_burnin", _i64), ("time_step", _d),
                ("max_drift", _d), ("e_ref_damping", _d), ("n_vmc_equilibrate", _i64),
                ("n_vmc_force_samples", _i64), ("capacity", _i64), ("matched_branching", _i), ("log_domain", _i)]


class IsdmcResults(C.Structure):
    _fields_ = [("energy", _d), ("error", _d), ("acceptance_rate", _d), ("avg_population", _d),
                ("collapses", _i64), ("dropped", _i64)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        sig = {
            "orc_philox4x32_10": (None, [_p, _p, _p]),
            "orc_u01": (_d, [_u32]),
            "orc_normals3": (None, [_u64, _u32, _u32, _u32, _u32, _p]),
            "orc_lu_det_inv": (_i, [_i, _p, _p, _p, _p, _p]),
            "orc_lu_solve": (_i, [_i, _p, _p, _p]),
            "orc_h2mo_new": (_p, [_d, _d, _d, _d]),
            "orc_h2mo_new_general": (_p, [_p, _p, _d, _d, _d]),
            "orc_lithium_ref_new": (_p, [_i]),
            "orc_lithium_config_new": (_p, [_u64, _i]),
            "orc_methane_gto_new": (_p, [_d, _d, _d]),
            "orc_methane_sto3g_new": (_p, [_d, _d, _d, _u64]),
            "orc_benzene_new": (_p, [_d, _d, _d, _d]),
            "orc_system_free": (None, [_p]),
            "orc_system_n_elec": (_i, [_p]),
            "orc_system_n_nuc": (_i, [_p]),
            "orc_system_n_ao": (_i, [_p]),
            "orc_system_n_mo": (_i, [_p]),
            "orc_system_get_mo": (None, [_p, _i, _p]),
            "orc_system_get_nuclei": (None, [_p, _p, _p]),
            "orc_system_set_nuclei": (None, [_p, _p]),
            "orc_num_params": (_i, [_p]),
            "orc_get_params": (None, [_p, _p]),
            "orc_set_params": (None, [_p, _p]),
            "orc_set_jastrow_series": (None, [_p, _i, _i]),
            "orc_nuclear_repulsion": (_d, [_p]),
            "orc_initialize": (None, [_p, _u64, _u32, _p]),
            "orc_evaluate": (_d, [_p, _p]),
            "orc_derivative": (None, [_p, _p, _p]),
            "orc_laplacian": (None, [_p, _p, _p]),
            "orc_local_energy": (_d, [_p, _p]),
            "orc_log_psi": (None, [_p, _p, _p, _p]),
            "orc_log_derivatives": (None, [_p, _p, _p]),
            "orc_hf_force": (None, [_p, _p, _p]),
            "orc_wfn_nuclear_gradient": (None, [_p, _p, _p]),
            "orc_batch_eval": (None, [_p, _i64, _p, _p, _p, _p, _p, _p]),
            "orc_autocorrelation_time": (_d, [_p, _i64]),
            "orc_blocking_error": (_d, [_p, _i64, _d]),
            "orc_mcmc_run": (None, [_p, _p, _u64, _p, _p]),
            "orc_mcmc_time_steps": (None, [_p, _i64, _i64, _i64, _d, _u64, _p, _p]),
            "orc_mcmc_time_steps_mt": (None, [_p, _i64, _i64, _i64, _d, _u64, _i, _p, _p]),
            "orc_single_electron_steps": (None, [_p, _i64, _p, _p, _u64, _u32, _i, _d, _p, _p]),
            "orc_dd_steps": (None, [_p, _i64, _p, _u64, _u32, _i, _d, _d, _i, _p, _p]),
            "orc_all_electron_steps": (None, [_p, _i64, _p, _p, _p, _u64, _u32, _i, _d, _p]),
            "orc_sr_update": (_i, [_i64, _i, _p, _p, _d, _p, _p, _p]),
            "orc_zv_forces": (None, [_i64, _i, _p, _p, _p, _i, _p, _p, _p, _p]),
            "orc_random_mo": (None, [_u64, _i, _i, _p]),
            "orc_slater_matrix": (_i, [_p, _p, _i, _p]),
            "orc_jastrow_u": (_d, [_p, _p]),
            "orc_sto_system_new": (_p, [_i, _p, _p, _p, _p, _p, _d, _i, _i]),
            "orc_isdmc_branch": (None, [_i64, _p, _p, _d, _d, _u64, _u32, _p]),
            "orc_isdmc_block_error": (_d, [_p, _i64]),
            "orc_isdmc_run": (None, [_p, _p, _u64, _p, _p, _p, _p, _p, _p]),
            "orc_j1_eval": (_d, [_p, _p]),
            "orc_j1_deriv": (None, [_p, _p, _p]),
            "orc_j1_lap": (None, [_p, _p, _p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def philox(ctr, key) -> np.ndarray:
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_ptr(c), _ptr(k), _ptr(out))
    return out


def lu_det_inv(a):
    a = _f64(a)
    n = a.shape[0]
    det = C.c_double()
    la = C.c_double()
    sg = C.c_double()
    inv = np.zeros_like(a)
    sing = lib().orc_lu_det_inv(n, _ptr(a), C.byref(det), _ptr(inv), C.byref(la), C.byref(sg))
    return det.value, inv, la.value, sg.value, sing


class System:
    """An oracle system; mirrors the reference's MultiWfn + EnergyCalculator
    (+ OptimizableWfn, ForceCalculator) trait methods one configuration at a
    time, plus batched helpers."""

    def __init__(self, handle):
        if not handle:
            raise RuntimeError("oracle constructor failed")
        self._h = C.c_void_p(handle)
        L = lib()
        self.n_elec = L.orc_system_n_elec(self._h)
        self.n_nuc = L.orc_system_n_nuc(self._h)

    def __del__(self):
        if getattr(self, "_h", None) is not None and _lib is not None:
            _lib.orc_system_free(self._h)
            self._h = None

    # constructors -----------------------------------------------------------
    @classmethod
    def h2mo(cls, alpha=1.3, half_bond=0.7, F=3.0, init_sigma=1.0):
        return cls(lib().orc_h2mo_new(alpha, half_bond, F, init_sigma))

    @classmethod
    def h2mo_general(cls, c1, c2, alpha, F, init_sigma=1.0):
        a = _f64(c1)
        b = _f64(c2)
        return cls(lib().orc_h2mo_new_general(_ptr(a), _ptr(b), alpha, F, init_sigma))

    @classmethod
    def lithium_ref(cls, exact=False):
        return cls(lib().orc_lithium_ref_new(int(exact)))

    @classmethod
    def lithium_config(cls, seed=0, exact=False):
        return cls(lib().orc_lithium_config_new(seed, int(exact)))

    @classmethod
    def sto_system(cls, orbitals, spins, F=1.0, Z=3, exact=False):
        """STOSlaterDet x Jastrow2 around one nucleus of charge Z at the
        origin; orbitals is a list of (coef, powers, zetas) triples."""
        nprim = np.array([len(o[0]) for o in orbitals], dtype=np.int32)
        coef = _f64(np.concatenate([o[0] for o in orbitals]))
        pw = np.ascontiguousarray(np.concatenate([o[1] for o in orbitals]), dtype=np.int32)
        zeta = _f64(np.concatenate([o[2] for o in orbitals]))
        sp = np.ascontiguousarray(spins, dtype=np.int32)
        return cls(lib().orc_sto_system_new(len(orbitals), _ptr(nprim), _ptr(coef), _ptr(pw),
                                            _ptr(zeta), _ptr(sp), F, Z, int(exact)))

    @classmethod
    def methane_gto(cls, b_ee=1.5, b_en=2.0, ch_bond=2.05):
        return cls(lib().orc_methane_gto_new(b_ee, b_en, ch_bond))

    @classmethod
    def methane_sto3g(cls, b_ee=1.5, b_en=2.0, ch_bond=2.06, mo_seed=0):
        return cls(lib().orc_methane_sto3g_new(b_ee, b_en, ch_bond, mo_seed))

    @classmethod
    def benzene(cls, b_ee=1.5, b_en=2.0, cc_bond=2.64, ch_bond=2.05):
        return cls(lib().orc_benzene_new(b_ee, b_en, cc_bond, ch_bond))

    # properties ---------------------------------------------------------------
    def nuclei(self):
        R = np.zeros((self.n_nuc, 3))
        Z = np.zeros(self.n_nuc)
        lib().orc_system_get_nuclei(self._h, _ptr(R), _ptr(Z))
        return R, Z

    def set_nuclei(self, R):
        R = _f64(R).reshape(self.n_nuc, 3)
        lib().orc_system_set_nuclei(self._h, _ptr(R))

    def mo(self, spin=0):
        L = lib()
        m = np.zeros((L.orc_system_n_mo(self._h), L.orc_system_n_ao(self._h)))
        L.orc_system_get_mo(self._h, spin, _ptr(m))
        return m

    def num_params(self):
        return lib().orc_num_params(self._h)

    def get_params(self):
        p = np.zeros(self.num_params())
        lib().orc_get_params(self._h, _ptr(p))
        return p

    def set_params(self, p):
        p = _f64(p)
        assert p.size == self.num_params()
        lib().orc_set_params(self._h, _ptr(p))

    def set_jastrow_series(self, k_ee, k_en):
        lib().orc_set_jastrow_series(self._h, k_ee, k_en)

    def nuclear_repulsion(self):
        return lib().orc_nuclear_repulsion(self._h)

    # trait methods ------------------------------------------------------------
    def initialize(self, seed, walker):
        r = np.zeros((self.n_elec, 3))
        lib().orc_initialize(self._h, seed, walker, _ptr(r))
        return r

    def initialize_batch(self, seed, W):
        return np.stack([self.initialize(seed, w) for w in range(W)])

    def evaluate(self, r):
        r = _f64(r)
        return lib().orc_evaluate(self._h, _ptr(r))

    def derivative(self, r):
        r = _f64(r)
        g = np.zeros((self.n_elec, 3))
        lib().orc_derivative(self._h, _ptr(r), _ptr(g))
        return g

    def laplacian(self, r):
        r = _f64(r)
        l_ = np.zeros(self.n_elec)
        lib().orc_laplacian(self._h, _ptr(r), _ptr(l_))
        return l_

    def local_energy(self, r):
        r = _f64(r)
        return lib().orc_local_energy(self._h, _ptr(r))

    def log_psi(self, r):
        r = _f64(r)
        la = C.c_double()
        sg = C.c_double()
        lib().orc_log_psi(self._h, _ptr(r), C.byref(la), C.byref(sg))
        return la.value, sg.value

    def slater_matrix(self, r, spin):
        r = _f64(r)
        out = np.zeros(64 * 64)
        n = lib().orc_slater_matrix(self._h, _ptr(r), spin, _ptr(out))
        return out[: n * n].reshape(n, n)

    def jastrow_u(self, r):
        r = _f64(r)
        return lib().orc_jastrow_u(self._h, _ptr(r))

    def log_derivatives(self, r):
        r = _f64(r)
        O = np.zeros(self.num_params())
        lib().orc_log_derivatives(self._h, _ptr(r), _ptr(O))
        return O

    def hf_force(self, r):
        r = _f64(r)
        F = np.zeros((self.n_nuc, 3))
        lib().orc_hf_force(self._h, _ptr(r), _ptr(F))
        return F

    def wfn_nuclear_gradient(self, r):
        r = _f64(r)
        Om = np.zeros((self.n_nuc, 3))
        lib().orc_wfn_nuclear_gradient(self._h, _ptr(r), _ptr(Om))
        return Om

    def batch_eval(self, pos):
        pos = _f64(pos)
        W = pos.shape[0]
        n = self.n_elec
        la = np.zeros(W)
        sg = np.zeros(W)
        g = np.zeros((W, n, 3))
        l_ = np.zeros((W, n))
        el = np.zeros(W)
        lib().orc_batch_eval(self._h, W, _ptr(pos), _ptr(la), _ptr(sg), _ptr(g), _ptr(l_), _ptr(el))
        return dict(logabs=la, sign=sg, grad=g, lap=l_, eloc=el)

    # samplers -----------------------------------------------------------------
    def mcmc_run(self, n_walkers, n_steps, step, seed=0, max_step=None, min_step=None,
                 target=0.5, interval=100):
        p = McmcParams(n_walkers, n_steps, step, step if max_step is None else max_step,
                       step if min_step is None else min_step, target, interval)
        e = np.zeros(n_steps)
        res = McmcResults()
        lib().orc_mcmc_run(self._h, C.byref(p), seed, _ptr(e), C.byref(res))
        return res, e

    def mcmc_time_steps(self, W, warm, steps, step, seed=0):
        sec = C.c_double()
        me = C.c_double()
        lib().orc_mcmc_time_steps(self._h, W, warm, steps, step, seed, C.byref(sec), C.byref(me))
        return sec.value, me.value

    def mcmc_time_steps_mt(self, W, warm, steps, step, seed=0, threads=1):
        sec = C.c_double()
        me = C.c_double()
        lib().orc_mcmc_time_steps_mt(self._h, W, warm, steps, step, seed, threads, C.byref(sec), C.byref(me))
        return sec.value, me.value

    def single_electron_steps(self, pos, seed, step0, nsteps, step):
        pos = _f64(pos).copy()
        W = pos.shape[0]
        lp = np.zeros(W)
        el = np.zeros((nsteps, W))
        acc = C.c_int64()
        lib().orc_single_electron_steps(self._h, W, _ptr(pos), _ptr(lp), seed, step0, nsteps, step,
                                        _ptr(el), C.byref(acc))
        return pos, lp, el, acc.value

    def dd_steps(self, pos, seed, step0, nsteps, tau, max_drift, no_green=False, log_domain=False):
        """DriftDiffusionVMC sweeps (vmc.rs:382-438) with the GPU's counters;
        log_domain: acceptance from ln|psi| without the 1e-30 guard."""
        pos = _f64(pos).copy()
        W = pos.shape[0]
        el = np.zeros((nsteps, W))
        acc = C.c_int64()
        flags = (1 if no_green else 0) | (2 if log_domain else 0)
        lib().orc_dd_steps(self._h, W, _ptr(pos), seed, step0, nsteps, tau, max_drift, flags, _ptr(el),
                           C.byref(acc))
        return pos, el, acc.value

    def isdmc_run(self, seed, n_walkers, n_steps, n_burnin, time_step, max_drift=1.0, e_ref_damping=0.5,
                  n_vmc_equilibrate=500, n_vmc_force_samples=2000, capacity=0, matched_branching=False,
                  log_domain=False):
        """ImportanceSampledDMC::run (is_dmc.rs:376-546) with the GPU's
        counters; capacity 0 means 4 x n_walkers rounded up to 256;
        log_domain: drift-diffusion acceptance from ln|psi|."""
        if capacity <= 0:
            capacity = (4 * n_walkers + 255) // 256 * 256
        p = IsdmcParams(n_walkers, n_steps, n_burnin, time_step, max_drift, e_ref_damping, n_vmc_equilibrate,
                        n_vmc_force_samples, capacity, int(matched_branching), int(log_domain))
        r = IsdmcResults()
        M = self.n_nuc
        fm, fv, fx = np.zeros((M, 3)), np.zeros((M, 3)), np.zeros((M, 3))
        et, pt = np.zeros(n_steps), np.zeros(n_steps)
        lib().orc_isdmc_run(self._h, C.byref(p), seed, C.byref(r), _ptr(fm), _ptr(fv), _ptr(fx), _ptr(et),
                            _ptr(pt))
        return dict(energy=r.energy, error=r.error, acceptance_rate=r.acceptance_rate,
                    avg_population=r.avg_population, collapses=r.collapses, dropped=r.dropped,
                    forces_mixed=fm, forces_vmc=fv, forces_extrapolated=fx, energies=et, populations=pt)

    def all_electron_steps(self, pos, logpsi, eloc, seed, step0, nsteps, step):
        pos = _f64(pos).copy()
        lp = _f64(logpsi).copy()
        el = _f64(eloc).copy()
        acc = C.c_int64()
        lib().orc_all_electron_steps(self._h, pos.shape[0], _ptr(pos), _ptr(lp), _ptr(el), seed,
                                     step0, nsteps, step, C.byref(acc))
        return pos, lp, el, acc.value


def isdmc_branch(el, el_prev, e_ref, tau, seed, step):
    """Copies per walker of the IS-DMC branching (is_dmc.rs:284-289)."""
    el, el_prev = _f64(el), _f64(el_prev)
    out = np.zeros(el.size, dtype=np.int32)
    lib().orc_isdmc_branch(el.size, _ptr(el), _ptr(el_prev), e_ref, tau, seed, step, _ptr(out))
    return out


def isdmc_block_error(e):
    e = _f64(e)
    return lib().orc_isdmc_block_error(_ptr(e), e.size)


def autocorrelation_time(e):
    e = _f64(e)
    return lib().orc_autocorrelation_time(_ptr(e), e.size)


def blocking_error(e, tau):
    e = _f64(e)
    return lib().orc_blocking_error(_ptr(e), e.size, tau)


def sr_update(E, O, eps):
    E = _f64(E)
    O = _f64(O)
    ns, P = O.shape
    dp = np.zeros(P)
    em = C.c_double()
    var = C.c_double()
    sing = lib().orc_sr_update(ns, P, _ptr(E), _ptr(O), eps, _ptr(dp), C.byref(em), C.byref(var))
    return dp, em.value, var.value, sing


def zv_forces(E, F, Om, use_zv=True):
    E = _f64(E)
    F = _f64(F)
    Om = _f64(Om)
    ns, M = F.shape[0], F.shape[1]
    mf = np.zeros((M, 3))
    fv = np.zeros(M)
    me = C.c_double()
    err = C.c_double()
    lib().orc_zv_forces(ns, M, _ptr(E), _ptr(F), _ptr(Om), int(use_zv), _ptr(mf), C.byref(me),
                        C.byref(err), _ptr(fv))
    return mf, me.value, err.value, fv


def jastrow1(F, r):
    """Jastrow1 (jastrow.rs:30-53) value, gradient and Laplacian of J."""
    f = C.c_double(F)
    r = _f64(r)
    g = np.zeros((2, 3))
    l_ = np.zeros(2)
    v = lib().orc_j1_eval(C.byref(f), _ptr(r))
    lib().orc_j1_deriv(C.byref(f), _ptr(r), _ptr(g))
    lib().orc_j1_lap(C.byref(f), _ptr(r), _ptr(l_))
    return v, g, l_


def random_mo(seed, n_mo, n_ao):
    m = np.zeros((n_mo, n_ao))
    lib().orc_random_mo(seed, n_mo, n_ao, _ptr(m))
    return m
