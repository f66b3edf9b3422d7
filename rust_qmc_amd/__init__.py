"""rust_qmc_amd: MI355X-native VMC walker-ensemble engine.

Python mirror of the rust_qmc trait surface over the C ABI of
include/qmc.h.  Class and method names follow the reference crate (paths
relative to its root):

  H2MoleculeMO      src/systems/hydrogen.rs:28-35
  Lithium           src/systems/lithium.rs:9-19 (STOSlaterDet x Jastrow2)
  MethaneGTO        src/systems/methane.rs:601-685
  BenzeneGTO        src/systems/benzene.rs:49-112
  MCMCParams / MCMCSimulation / MCMCResults   src/sampling/vmc.rs:12-189

The trait methods evaluate / derivative / laplacian return the reference's
linear quantities (psi, grad psi, lap psi per electron) computed as
sign*exp(log psi) times the log-domain results of the GPU; log_psi,
grad_log_psi and lap_over_psi return the log-domain values themselves.
Every method runs on the GPU through libqmc_hip.so; nothing here computes
physics on the CPU.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import (QMC_FP64, QMC_MIXED, QMC_FLAGS_REFERENCE, QMC_MOVE_ALL, QMC_MOVE_DRIFT, QMC_MOVE_SINGLE,
                   QmcError, check, lib)

__all__ = [
    "QmcError", "System", "H2MoleculeMO", "Lithium", "STOSlaterDet", "MethaneGTO", "BenzeneGTO", "Ensemble",
    "DeviceBuffer", "check", "lib",
    "MCMCParams", "MCMCResults", "MCMCSimulation", "autocorrelation_time", "blocking_error", "QMC_FP64",
    "QMC_MIXED", "QMC_MOVE_SINGLE", "QMC_MOVE_ALL", "QMC_MOVE_DRIFT", "QMC_FLAGS_REFERENCE", "Batch",
    "DDVMCParams", "DDVMCResults", "DriftDiffusionVMC",
    "ISDMCParams", "ISDMCResults", "ImportanceSampledDMC", "isdmc_block_error", "GeometryOptResult",
]

_PREC = {"fp64": QMC_FP64, "mixed": QMC_MIXED, QMC_FP64: QMC_FP64, QMC_MIXED: QMC_MIXED}


def _ptr(a):
    """ctypes pointer for a numpy array or a torch tensor (host or device)."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return C.c_void_p(a.data_ptr())
    return a.ctypes.data_as(C.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class System:
    """A wavefunction + Hamiltonian handle (MultiWfn + EnergyCalculator +
    OptimizableWfn + ForceCalculator)."""

    def __init__(self, handle: C.c_void_p, precision):
        self._h = handle
        self.precision = _PREC[precision]
        ne, nu, nn, na = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        check(lib().qmc_system_info(self._h, C.byref(ne), C.byref(nu), C.byref(nn), C.byref(na)))
        self.n_elec, self.n_up, self.n_nuc, self.n_ao = ne.value, nu.value, nn.value, na.value
        self.n_dn = self.n_elec - self.n_up

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib._lib is not None:
            _lib._lib.qmc_system_destroy(h)
            self._h = None

    @staticmethod
    def _new(fn, *args):
        h = C.c_void_p()
        check(fn(*args, C.byref(h)))
        return h

    @classmethod
    def from_desc(cls, desc: dict, precision="fp64"):
        """Generic description (see qmc_system_desc in include/qmc.h)."""
        keep = []

        def arr(x, dt):
            if x is None:
                return None
            a = np.ascontiguousarray(x, dtype=dt)
            keep.append(a)
            return a.ctypes.data_as(C.c_void_p)

        aos = (_lib.QmcAO * len(desc["aos"]))(*[_lib.QmcAO(*a) for a in desc["aos"]])
        keep.append(aos)
        mo_up = _f64(desc["mo_up"])
        d = _lib.QmcSystemDesc(
            n_elec=len(desc["spins"]), spins=arr(desc["spins"], np.int32), n_nuc=len(desc["charges"]),
            nuc_pos=arr(desc["nuclei"], np.float64), nuc_charge=arr(desc["charges"], np.float64),
            n_ao=len(desc["aos"]), aos=C.cast(aos, C.c_void_p), prim_exp=arr(desc["prim_exp"], np.float64),
            prim_coef=arr(desc["prim_coef"], np.float64), prim_pow=arr(desc.get("prim_pow"), np.int32),
            n_mo=mo_up.shape[0], mo_up=arr(mo_up, np.float64), mo_dn=arr(desc.get("mo_dn", mo_up), np.float64),
            jastrow_kind=desc.get("jastrow_kind", 0), jas_F=desc.get("F", 1.0), b_ee=desc.get("b_ee", 0.0),
            b_en=desc.get("b_en", 0.0), a_anti=desc.get("a_anti", 0.5), a_para=desc.get("a_para", 0.25),
            k_ee=0, k_en=0, c_anti=None, c_para=None, n_species=0, species=None, d_en=None,
            init_nuc=arr(desc.get("init_nuc", [0] * len(desc["spins"])), np.int32),
            init_sigma=desc.get("init_sigma", 1.0), flags=desc.get("flags", 0), mo_rule=desc.get("mo_rule", 0))
        h = cls._new(lib().qmc_system_create, C.byref(d), _PREC[precision])
        return cls(h, precision)

    # ---- descriptions as JSON / YAML (io/config.rs:20-35) ------------------------
    @classmethod
    def from_json(cls, text, precision="fp64"):
        """A system from a JSON description (the forms qmc_system_from_json
        reads: the H2MoleculeMO serde layout, a named molecule, or the
        generic description)."""
        raw = text.encode() if isinstance(text, str) else bytes(text)
        h = cls._new(lib().qmc_system_from_json, raw, len(raw), _PREC[precision])
        return cls(h, precision)

    @classmethod
    def from_file(cls, path, precision="fp64"):
        """A system from a .json file, or a .yaml/.yml file with the same
        structure (read with yaml.safe_load)."""
        import json
        with open(path, "r", encoding="utf-8") as f:
            text = f.read()
        if str(path).lower().endswith((".yaml", ".yml")):
            import yaml
            text = json.dumps(yaml.safe_load(text))
        return cls.from_json(text, precision)

    def to_json(self) -> str:
        """The generic description of this system (reloads exactly)."""
        n = C.c_int64()
        check(lib().qmc_system_to_json(self._h, None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        check(lib().qmc_system_to_json(self._h, buf, n.value, C.byref(n)))
        return buf.value.decode()

    def save(self, path):
        with open(path, "w", encoding="utf-8") as f:
            f.write(self.to_json())

    # ---- configuration -------------------------------------------------------
    def set_precision(self, precision):
        check(lib().qmc_set_precision(self._h, _PREC[precision]))
        self.precision = _PREC[precision]

    def set_flags(self, flags):
        check(lib().qmc_set_flags(self._h, flags))

    def set_engine(self, engine):
        """'auto', 'register' (one walker per lane), 'wave' (one walker per
        wavefront), 'generic' (register engine with run-time shell loops) or
        'tile' (16 walkers per workgroup, MFMA MO contraction)."""
        check(lib().qmc_set_engine(self._h, {"auto": 0, "register": 1, "wave": 2, "generic": 3, "tile": 4}[engine]))

    def engine_name(self):
        """The kernels this system runs on, e.g. 'register/ch4-sto3g'."""
        buf = C.create_string_buffer(64)
        check(lib().qmc_engine_name(self._h, buf, 64))
        return buf.value.decode()

    def nuclear_repulsion(self):
        v = C.c_double()
        check(lib().qmc_nuclear_repulsion(self._h, C.byref(v)))
        return v.value

    # OptimizableWfn (traits.rs:98-115)
    def num_params(self):
        n = C.c_int32()
        check(lib().qmc_num_params(self._h, C.byref(n)))
        return n.value

    def get_params(self):
        p = np.zeros(max(1, self.num_params()))
        check(lib().qmc_get_params(self._h, _ptr(p)))
        return p[: self.num_params()]

    def set_params(self, p):
        p = _f64(p)
        check(lib().qmc_set_params(self._h, _ptr(p), p.size))

    def set_jastrow_series(self, k_ee, k_en):
        check(lib().qmc_set_jastrow_series(self._h, k_ee, k_en))

    # ForceCalculator geometry (sampling/traits.rs:20-35)
    def num_nuclei(self):
        return self.n_nuc

    def get_nuclei(self):
        R = np.zeros((self.n_nuc, 3))
        check(lib().qmc_get_nuclei(self._h, _ptr(R), None))
        return R

    def get_charges(self):
        Z = np.zeros(self.n_nuc)
        check(lib().qmc_get_nuclei(self._h, None, _ptr(Z)))
        return Z

    def set_nuclei(self, R):
        R = _f64(R).reshape(self.n_nuc, 3)
        check(lib().qmc_set_nuclei(self._h, _ptr(R), self.n_nuc))

    # ---- batched trait methods (positions [W][N][3]) ---------------------------
    def evaluate_batch(self, pos, want=("logabs", "sign", "grad", "lap", "eloc")):
        """Run qmc_evaluate on host positions [W][N][3]."""
        pos = _f64(pos)
        W = pos.shape[0]
        mk = lambda *s: np.zeros(s)
        n = self.n_elec
        out = {}
        out["logabs"] = mk(W) if "logabs" in want else None
        out["sign"] = mk(W) if "sign" in want else None
        out["grad"] = mk(W, n, 3) if "grad" in want else None
        out["lap"] = mk(W, n) if "lap" in want else None
        out["eloc"] = mk(W) if "eloc" in want else None
        check(lib().qmc_evaluate(self._h, W, _ptr(pos), _ptr(out["logabs"]), _ptr(out["sign"]), _ptr(out["grad"]),
                                 _ptr(out["lap"]), _ptr(out["eloc"]), None))
        return {k: v for k, v in out.items() if v is not None}

    def log_psi(self, r):
        o = self.evaluate_batch(_f64(r)[None], ("logabs", "sign"))
        return o["logabs"][0], o["sign"][0]

    def grad_log_psi(self, r):
        return self.evaluate_batch(_f64(r)[None], ("grad",))["grad"][0]

    def lap_over_psi(self, r):
        return self.evaluate_batch(_f64(r)[None], ("lap",))["lap"][0]

    # MultiWfn (traits.rs:48-59), linear convention
    def evaluate(self, r):
        la, sg = self.log_psi(r)
        return sg * np.exp(la)

    def derivative(self, r):
        o = self.evaluate_batch(_f64(r)[None], ("logabs", "sign", "grad"))
        return o["sign"][0] * np.exp(o["logabs"][0]) * o["grad"][0]

    def laplacian(self, r):
        o = self.evaluate_batch(_f64(r)[None], ("logabs", "sign", "lap"))
        return o["sign"][0] * np.exp(o["logabs"][0]) * o["lap"][0]

    # EnergyCalculator (sampling/traits.rs:6-8)
    def local_energy(self, r):
        return self.evaluate_batch(_f64(r)[None], ("eloc",))["eloc"][0]

    # derivatives (OptimizableWfn::log_derivatives, ForceCalculator) -----------
    def derivatives_batch(self, pos):
        """Per configuration: eloc [W], O [W][P], dE/dp [W][P], HF force
        [W][M][3] and d ln psi / dR [W][M][3]."""
        pos = _f64(pos)
        W = pos.shape[0]
        P, M = self.num_params(), self.n_nuc
        el, O, dE = np.zeros(W), np.zeros((W, P)), np.zeros((W, P))
        F, Om = np.zeros((W, M, 3)), np.zeros((W, M, 3))
        check(lib().qmc_derivatives(self._h, W, _ptr(pos), _ptr(el), _ptr(O), _ptr(dE), _ptr(F), _ptr(Om), None))
        return dict(eloc=el, O=O, dE=dE, F=F, Om=Om)

    def log_derivatives(self, r):
        return self.derivatives_batch(_f64(r)[None])["O"][0]

    def hellmann_feynman_force(self, r):
        return self.derivatives_batch(_f64(r)[None])["F"][0]

    def wfn_nuclear_gradient(self, r):
        return self.derivatives_batch(_f64(r)[None])["Om"][0]

    def initialize(self, seed=0, walker=0):
        """MultiWfn::initialize with a Philox stream (walker `walker`)."""
        ens = Ensemble(self, walker + 1, seed)
        return ens.positions()[walker]


class H2MoleculeMO(System):
    """hydrogen.rs:28-35: (phi1+phi2)(r1) (phi1+phi2)(r2) J1(r12)."""

    def __init__(self, alpha=1.3, center1=(0.0, 0.0, 0.7), center2=(0.0, 0.0, -0.7), cusp_param=3.0,
                 init_sigma=1.0, precision="fp64"):
        c1 = _f64(center1)
        c2 = _f64(center2)
        h = self._new(lib().qmc_h2_molecule_mo_create, alpha, _ptr(c1), _ptr(c2), cusp_param, init_sigma,
                      _PREC[precision])
        super().__init__(h, precision)


class STOSlaterDet(System):
    """STOSlaterDet x Jastrow2 (slater.rs:149-271, lithium.rs:9-82).
    orbitals: list of (coefficients, powers, exponents); spins per column."""

    def __init__(self, orbitals, spins, cusp_param=1.0, Z=3.0, flags=QMC_FLAGS_REFERENCE, precision="fp64"):
        nprim = np.array([len(o[0]) for o in orbitals], dtype=np.int32)
        coef = _f64(np.concatenate([o[0] for o in orbitals]))
        pw = np.ascontiguousarray(np.concatenate([o[1] for o in orbitals]), dtype=np.int32)
        zeta = _f64(np.concatenate([o[2] for o in orbitals]))
        sp = np.ascontiguousarray(spins, dtype=np.int32)
        h = self._new(lib().qmc_sto_slater_det_create, len(orbitals), _ptr(nprim), _ptr(coef), _ptr(pw), _ptr(zeta),
                      _ptr(sp), cusp_param, Z, flags, _PREC[precision])
        super().__init__(h, precision)


# slater.rs:68-89 (init_li_sto): 7 primitives, powers [1,1,2,2,2,2,3]
_LI_STO_EXP = [0.72089388, 2.61691643, 0.69257443, 1.37137558, 3.97864549, 13.52900016, 19.30801440]
_LI_STO_POW = [1, 1, 2, 2, 2, 2, 3]
_LI_STO_C = {
    1: [-0.12220686, 1.11273225, 0.04125378, 0.09306499, -0.10260021, -0.00034191, 0.00021963],
    2: [0.47750469, 0.11140449, -1.25954273, -0.18475003, -0.02736293, -0.00025064, 0.00057962],
}


def init_li_sto(n):
    """slater.rs:68-89"""
    return (_LI_STO_C.get(n, [0.0] * 7), _LI_STO_POW, _LI_STO_EXP)


class Lithium(STOSlaterDet):
    """src/main.rs:36-49: orbitals [1s, 1s, 2s], spins [+1, -1, +1], Jastrow2 F = 1.
    seed=None builds the reference's init_li_sto atom; an integer seed builds
    the seeded 3-term STO atom of the Li benchmark workload."""

    def __init__(self, seed=None, flags=QMC_FLAGS_REFERENCE, precision="fp64"):
        if seed is None:
            super().__init__([init_li_sto(1), init_li_sto(1), init_li_sto(2)], [1, -1, 1], 1.0, 3.0, flags, precision)
        else:
            h = self._new(lib().qmc_lithium_seeded_create, seed, flags, _PREC[precision])
            System.__init__(self, h, precision)


class MethaneGTO(System):
    """methane.rs:601-685.  basis='6-31G' is the reference's 17-AO system;
    basis='STO-3G' is the 9-AO system with seeded orthonormal MOs."""

    def __init__(self, b_ee=1.5, b_en=2.0, ch_bond=2.05, basis="6-31G", mo_seed=0, precision="fp64"):
        bk = {"6-31G": _lib.QMC_BASIS_631G, "STO-3G": _lib.QMC_BASIS_STO3G}[basis]
        h = self._new(lib().qmc_methane_gto_create, b_ee, b_en, ch_bond, bk, mo_seed, _PREC[precision])
        super().__init__(h, precision)


class BenzeneGTO(System):
    """benzene.rs:49-112"""

    def __init__(self, b_ee=1.5, b_en=2.0, cc_bond=2.64, ch_bond=2.042, precision="fp64"):
        h = self._new(lib().qmc_benzene_gto_create, b_ee, b_en, cc_bond, ch_bond, _PREC[precision])
        super().__init__(h, precision)


_hip_rt = None


def _hip():
    """The HIP runtime libqmc_hip.so itself links (libamdhip64.so.7)."""
    global _hip_rt
    if _hip_rt is None:
        lib()
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipDeviceSynchronize.argtypes = []
        _hip_rt = h
    return _hip_rt


class DeviceBuffer:
    """Device memory from the library's own HIP runtime, for callers that
    hand device pointers to the ABI.  (PyTorch bundles a second HIP runtime;
    the two are not mixed in one process.)"""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        if _hip().hipMalloc(C.byref(self.ptr), max(self.nbytes, 8)) != 0:
            raise QmcError(_lib.QMC_ERR_HIP, "hipMalloc failed")

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        if _hip().hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) != 0:
            raise QmcError(_lib.QMC_ERR_HIP, "hipMemcpy H2D failed")

    def download(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        assert out.nbytes <= self.nbytes
        if _hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes, 2) != 0:
            raise QmcError(_lib.QMC_ERR_HIP, "hipMemcpy D2H failed")
        return out

    def copy_from(self, other: "DeviceBuffer", nbytes: int):
        """Device-to-device copy (runs as the HIP runtime's copy kernel)."""
        assert nbytes <= self.nbytes and nbytes <= other.nbytes
        if _hip().hipMemcpy(self.ptr, other.ptr, int(nbytes), 3) != 0:
            raise QmcError(_lib.QMC_ERR_HIP, "hipMemcpy D2D failed")

    def free(self):
        if self.ptr:
            _hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()


class Ensemble:
    """W walkers resident in HBM (qmc_ensemble)."""

    def __init__(self, system: System, n_walkers: int, seed: int = 0):
        self.system = system
        self.W = n_walkers
        self._h = C.c_void_p()
        check(lib().qmc_ensemble_create(system._h, n_walkers, seed, C.byref(self._h)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib._lib is not None:
            _lib._lib.qmc_ensemble_destroy(h)
            self._h = None

    def positions(self):
        p = np.zeros((self.W, self.system.n_elec, 3))
        check(lib().qmc_ensemble_get_positions(self._h, _ptr(p)))
        return p

    def set_positions(self, pos):
        pos = _f64(pos)
        check(lib().qmc_ensemble_set_positions(self._h, _ptr(pos)))

    def state(self):
        la, sg, el = np.zeros(self.W), np.zeros(self.W), np.zeros(self.W)
        check(lib().qmc_ensemble_get_state(self._h, _ptr(la), _ptr(sg), _ptr(el)))
        return la, sg, el

    def refresh(self):
        check(lib().qmc_ensemble_refresh(self._h))

    def measure(self, reset=False):
        """Accumulate one SR-grad sample of every walker."""
        check(lib().qmc_ensemble_measure(self._h, int(reset)))

    def moments(self):
        P, M = self.system.num_params(), self.system.n_nuc
        Q = P + 3 * M
        out = dict(o_mean=np.zeros(Q), S=np.zeros((Q, Q)), f=np.zeros(Q), de_mean=np.zeros(max(P, 1)),
                   f_hf=np.zeros((M, 3)), f_zv=np.zeros((M, 3)), f_var=np.zeros(M))
        m = _lib.QmcMoments(0, 0, 0, 0.0, 0.0, *[_ptr(out[k]) for k in
                                                 ("o_mean", "S", "f", "de_mean", "f_hf", "f_zv", "f_var")])
        check(lib().qmc_ensemble_moments(self._h, C.byref(m)))
        out["de_mean"] = out["de_mean"][:P]
        out.update(n_samples=m.n_samples, energy=m.energy, variance=m.variance, P=P, M=M)
        return out

    def set_timing(self, on=True):
        check(lib().qmc_ensemble_set_timing(self._h, int(on)))

    def timing(self):
        """(summed sweep-kernel ms, launches) of the last sweeps() call."""
        ms, n = C.c_double(), C.c_int64()
        check(lib().qmc_ensemble_timing(self._h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def sweeps(self, n_steps, step_size, move=QMC_MOVE_SINGLE, refresh_interval=64, steps_per_launch=0,
               trace=False, max_drift=1.0):
        """Returns (e_sum, e2_sum, accepted[, eloc_trace]) per sweep.  For
        move=QMC_MOVE_DRIFT step_size is the time step tau."""
        p = _lib.QmcSweepParams(move, n_steps, step_size, refresh_interval, steps_per_launch, max_drift)
        es, e2, acc = np.zeros(n_steps), np.zeros(n_steps), np.zeros(n_steps)
        tr = np.zeros((n_steps, self.W)) if trace else None
        check(lib().qmc_ensemble_sweeps(self._h, C.byref(p), _ptr(es), _ptr(e2), _ptr(acc), _ptr(tr)))
        return (es, e2, acc, tr) if trace else (es, e2, acc)


@dataclass
class MCMCParams:
    """vmc.rs:12-21 plus n_burnin / move / refresh_interval."""
    n_walkers: int
    n_steps: int
    initial_step_size: float
    max_step_size: float
    min_step_size: float
    target_acceptance: float = 0.5
    adaptation_interval: int = 100
    n_burnin: int = 0
    move: int = QMC_MOVE_ALL
    refresh_interval: int = 64


@dataclass
class MCMCResults:
    """vmc.rs:30-35 plus acceptance and final step size."""
    energy: float
    error: float
    autocorrelation_time: float
    acceptance: float
    final_step_size: float


class MCMCSimulation:
    """vmc.rs:38-189 on the GPU."""

    def __init__(self, wavefunction: System, params: MCMCParams, seed: int = 0):
        self.wavefunction = wavefunction
        self.params = params
        self.seed = seed
        self.energies = None

    def run(self) -> MCMCResults:
        p = self.params
        cp = _lib.QmcMcmcParams(p.n_walkers, p.n_steps, p.initial_step_size, p.max_step_size, p.min_step_size,
                                p.target_acceptance, p.adaptation_interval, p.n_burnin, p.move, p.refresh_interval)
        res = _lib.QmcVmcResults()
        tr = np.zeros(p.n_steps)
        check(lib().qmc_vmc_run(self.wavefunction._h, C.byref(cp), self.seed, C.byref(res), _ptr(tr)))
        self.energies = tr
        return MCMCResults(res.energy, res.error, res.autocorrelation_time, res.acceptance, res.final_step_size)


def sr_solve(S, f, eps):
    """(S + eps I) dp = -f (sr_optimize.rs:219-242); returns (dp, singular)."""
    S = _f64(S)
    f = _f64(f)
    dp = np.zeros(f.size)
    rc = lib().qmc_sr_solve(f.size, _ptr(S), _ptr(f), eps, _ptr(dp))
    return dp, rc == _lib.QMC_ERR_SINGULAR


@dataclass
class SRResult:
    """sr_optimize.rs:351-366"""
    final_params: np.ndarray
    final_energy: float
    final_variance: float
    energy_history: np.ndarray
    variance_history: np.ndarray
    param_history: np.ndarray


@dataclass
class SROptimizer:
    """sr_optimize.rs:21-56 (defaults as SROptimizer::default)"""
    n_samples: int = 5000
    n_walkers: int = 20
    n_equilibrate: int = 500
    max_iterations: int = 50
    learning_rate: float = 0.05
    sr_epsilon: float = 0.001
    tolerance: float = 1e-5
    step_size: float = 0.5
    decorrelation: int = 5
    move: int = QMC_MOVE_ALL
    include_nuclear: bool = False
    refresh_interval: int = 64
    seed: int = 0

    def optimize(self, wfn: System) -> SRResult:
        """sr_optimize.rs:252-348"""
        P = wfn.num_params()
        p = _lib.QmcSrParams(self.n_samples, self.n_walkers, self.n_equilibrate, self.max_iterations,
                             self.learning_rate, self.sr_epsilon, self.tolerance, self.step_size, self.decorrelation,
                             self.move, int(self.include_nuclear), self.refresh_interval)
        res = _lib.QmcSrResults()
        eh = np.zeros(self.max_iterations)
        vh = np.zeros(self.max_iterations)
        ph = np.zeros((self.max_iterations + 1, P))
        check(lib().qmc_sr_optimize(wfn._h, C.byref(p), self.seed, C.byref(res), _ptr(eh), _ptr(vh), _ptr(ph)))
        n = res.iterations
        return SRResult(wfn.get_params(), res.final_energy, res.final_variance, eh[:n], vh[:n], ph[: n + 1])


@dataclass
class GeometryOptResult:
    """geometry_opt.rs:351-368; histories cover the force samplings made."""
    final_nuclei: np.ndarray
    final_energy: float
    final_max_force: float
    energy_history: np.ndarray
    force_history: np.ndarray
    geometry_history: np.ndarray
    zero_variance: bool
    converged: bool


@dataclass
class GeometryOptimizer:
    """geometry_opt.rs:16-53 (defaults as Default for GeometryOptimizer; the
    verbose printing has no counterpart) plus the sampler's move kind,
    decorrelation sweeps, refresh interval, the displacement cap the
    reference hard-codes (geometry_opt.rs:298) and the Philox seed."""
    n_samples: int = 5000
    n_walkers: int = 20
    n_equilibrate: int = 500
    step_size: float = 0.005
    max_iterations: int = 50
    force_threshold: float = 0.05
    mc_step_size: float = 0.5
    zero_variance: bool = False
    decorrelation: int = 5
    move: int = QMC_MOVE_ALL
    refresh_interval: int = 64
    max_displacement: float = 0.02
    seed: int = 0

    def _force_params(self):
        return _lib.QmcForceParams(self.n_samples, self.n_walkers, self.n_equilibrate, self.mc_step_size,
                                   int(self.zero_variance), self.decorrelation, self.move, self.refresh_interval)

    def sample_forces(self, wfn: System):
        """geometry_opt.rs:105-227: (mean_forces [M][3], mean_energy, error, force_variance [M])."""
        p = self._force_params()
        M = wfn.n_nuc
        F = np.zeros((M, 3))
        fv = np.zeros(M)
        me, err = C.c_double(), C.c_double()
        check(lib().qmc_sample_forces(wfn._h, C.byref(p), self.seed, _ptr(F), C.byref(me), C.byref(err), _ptr(fv)))
        return F, me.value, err.value, fv

    def optimize(self, wfn: System) -> GeometryOptResult:
        """geometry_opt.rs:233-347 on the GPU (qmc_geometry_optimize); moves wfn's nuclei."""
        fp = self._force_params()
        gp = _lib.QmcGeomoptParams(self.step_size, self.max_iterations, self.force_threshold, self.max_displacement)
        res = _lib.QmcGeomoptResults()
        M, K = wfn.n_nuc, self.max_iterations
        eh, fh, gh = np.zeros(max(K, 1)), np.zeros(max(K, 1)), np.zeros((K + 1, M, 3))
        check(lib().qmc_geometry_optimize(wfn._h, C.byref(fp), C.byref(gp), self.seed, C.byref(res), _ptr(eh),
                                          _ptr(fh), _ptr(gh)))
        n = res.iterations
        steps = n - 1 if res.converged else n  # geometries after each displacement
        return GeometryOptResult(wfn.get_nuclei(), res.final_energy, res.final_max_force, eh[:n], fh[:n],
                                 gh[: steps + 1], self.zero_variance, bool(res.converged))


@dataclass
class DDVMCParams:
    """vmc.rs:196-226 (defaults as Default for DDVMCParams) plus refresh_interval."""
    n_walkers: int = 10
    n_steps: int = 100_000
    n_burnin: int = 1000
    time_step: float = 0.01
    max_drift: float = 1.0
    target_acceptance: float = 0.5
    adaptation_interval: int = 100
    refresh_interval: int = 64


@dataclass
class DDVMCResults:
    """vmc.rs:228-239"""
    energy: float
    error: float
    autocorrelation_time: float
    acceptance_rate: float
    final_time_step: float


class DriftDiffusionVMC:
    """vmc.rs:279-535 on the GPU (qmc_ddvmc_run); register-engine systems."""

    def __init__(self, wavefunction, params: DDVMCParams, seed: int = 0):
        self.wavefunction = wavefunction
        self.params = params
        self.seed = seed
        self.energies = None

    def run(self) -> DDVMCResults:
        p = self.params
        c = _lib.QmcDdvmcParams(p.n_walkers, p.n_steps, p.n_burnin, p.time_step, p.max_drift, p.target_acceptance,
                                p.adaptation_interval, p.refresh_interval)
        r = _lib.QmcDdvmcResults()
        self.energies = np.zeros(p.n_steps)
        check(lib().qmc_ddvmc_run(self.wavefunction._h, C.byref(c), self.seed, C.byref(r), _ptr(self.energies)))
        return DDVMCResults(r.energy, r.error, r.autocorrelation_time, r.acceptance_rate, r.final_time_step)


@dataclass
class ISDMCParams:
    """is_dmc.rs:49-84 (defaults as Default for ISDMCParams; print_interval has
    no counterpart) plus the population capacity (0: 4 x n_walkers), the
    FP64 inverse refresh interval of a mixed-precision population, and
    matched_branching: weights and E_ref feedback over the 2 tau the moves
    diffuse for, instead of the reference's tau (is_dmc.rs:215-222, 286)."""
    n_walkers: int = 200
    n_steps: int = 5000
    n_burnin: int = 1000
    time_step: float = 0.005
    max_drift: float = 1.0
    e_ref_damping: float = 0.5
    n_vmc_equilibrate: int = 500
    n_vmc_force_samples: int = 2000
    capacity: int = 0
    refresh_interval: int = 64
    matched_branching: bool = False


@dataclass
class ISDMCResults:
    """is_dmc.rs:90-108; forces are [M][3] arrays.  collapses, dropped,
    capacity and step_ms describe the population run."""
    energy: float
    error: float
    forces_mixed: np.ndarray
    forces_vmc: np.ndarray
    forces_extrapolated: np.ndarray
    acceptance_rate: float
    avg_population: float
    collapses: int = 0
    dropped: int = 0
    capacity: int = 0
    step_ms: float = 0.0


class ImportanceSampledDMC:
    """is_dmc.rs:176-607 on the GPU (qmc_isdmc_run); register-engine systems.
    After run(), energies and populations hold the production series."""

    def __init__(self, wfn, params: ISDMCParams, seed: int = 0):
        self.wfn = wfn
        self.params = params
        self.seed = seed
        self.energies = None
        self.populations = None

    def run(self) -> ISDMCResults:
        p = self.params
        c = _lib.QmcIsdmcParams(p.n_walkers, p.n_steps, p.n_burnin, p.time_step, p.max_drift, p.e_ref_damping,
                                p.n_vmc_equilibrate, p.n_vmc_force_samples, p.capacity, p.refresh_interval,
                                int(p.matched_branching))
        r = _lib.QmcIsdmcResults()
        M = self.wfn.num_nuclei()
        fm, fv, fx = np.zeros((M, 3)), np.zeros((M, 3)), np.zeros((M, 3))
        self.energies = np.zeros(p.n_steps)
        self.populations = np.zeros(p.n_steps)
        check(lib().qmc_isdmc_run(self.wfn._h, C.byref(c), self.seed, C.byref(r), _ptr(fm), _ptr(fv), _ptr(fx),
                                  _ptr(self.energies), _ptr(self.populations)))
        return ISDMCResults(r.energy, r.error, fm, fv, fx, r.acceptance_rate, r.avg_population, r.collapses,
                            r.dropped, r.capacity, r.step_ms)


def isdmc_block_error(e) -> float:
    """ImportanceSampledDMC::block_error (is_dmc.rs:555-607)."""
    e = np.ascontiguousarray(e, dtype=np.float64)
    return lib().qmc_isdmc_block_error(_ptr(e), e.size)


_BATCH_MODES = {"fused": 0, "streams": 1, "sequential": 2}
_BATCH_ORDERS = {"cost": 0, "given": 1}


class Batch:
    """Several ensembles sampled together (qmc_batch): in 'fused' mode one
    persistent launch runs them all from a queue of walker tiles."""

    def __init__(self, ensembles):
        self.ensembles = list(ensembles)
        arr = (C.c_void_p * len(self.ensembles))(*[e._h for e in self.ensembles])
        self._h = C.c_void_p()
        check(lib().qmc_batch_create(arr, len(self.ensembles), C.byref(self._h)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib._lib is not None:
            _lib._lib.qmc_batch_destroy(h)
            self._h = None

    @property
    def fusable(self):
        f = C.c_int32()
        check(lib().qmc_batch_fusable(self._h, C.byref(f)))
        return bool(f.value)

    def config(self):
        c = _lib.QmcBatchConfig()
        check(lib().qmc_batch_get_config(self._h, C.byref(c)))
        inv_m = {v: k for k, v in _BATCH_MODES.items()}
        inv_o = {v: k for k, v in _BATCH_ORDERS.items()}
        return {"mode": inv_m[c.mode], "blocks_per_cu": c.blocks_per_cu, "steps_per_launch": c.steps_per_launch,
                "order": inv_o[c.order], "refresh_interval": c.refresh_interval}

    def set_config(self, mode="fused", blocks_per_cu=0, steps_per_launch=8, order="cost", refresh_interval=0):
        c = _lib.QmcBatchConfig(_BATCH_MODES[mode], blocks_per_cu, steps_per_launch, _BATCH_ORDERS[order],
                                refresh_interval)
        check(lib().qmc_batch_set_config(self._h, C.byref(c)))

    def sweeps(self, n_steps, step_sizes, refresh_interval=64):
        """Per ensemble and sweep: sums of E_L, E_L^2 and accepted moves, each
        [n][n_steps].  refresh_interval < 0 takes the tuned interval."""
        n = len(self.ensembles)
        st = _f64(step_sizes).reshape(n)
        es, e2, acc = (np.zeros((n, n_steps)) for _ in range(3))
        check(lib().qmc_batch_sweeps(self._h, n_steps, _ptr(st), refresh_interval, _ptr(es), _ptr(e2), _ptr(acc)))
        return es, e2, acc

    def timing(self):
        ms, nl = C.c_double(), C.c_int64()
        check(lib().qmc_batch_timing(self._h, C.byref(ms), C.byref(nl)))
        return ms.value, nl.value

    def tune(self, trial_steps, step_sizes, refresh_interval=64):
        """Autotuner: keeps and returns the fastest configuration and its ms
        per sweep; refresh_interval < 0 also tunes the FP64 inverse refresh
        interval (up to -refresh_interval) under the drift guard."""
        st = _f64(step_sizes).reshape(len(self.ensembles))
        c = _lib.QmcBatchConfig()
        ms = C.c_double()
        check(lib().qmc_batch_tune(self._h, trial_steps, _ptr(st), refresh_interval, C.byref(c), C.byref(ms)))
        return self.config(), ms.value


def autocorrelation_time(e):
    e = _f64(e)
    return lib().qmc_autocorrelation_time(_ptr(e), e.size)


def blocking_error(e, tau):
    e = _f64(e)
    return lib().qmc_blocking_error(_ptr(e), e.size, tau)
