"""ctypes binding of libptx.so (the C ABI declared in include/ptx.h).

This is the Python-side host of the engine used by the tests and bench.py; it
mirrors the reference's objects (pticm::Problem, the solve() loop) on top of
the C ABI.  Loading fails loudly when the shared library has not been built:
there is no CPU fallback behind it.
"""
import ctypes as C
import os
import sys

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "build", "libptx.so")

FLAG_ICM = 1 << 0
FLAG_PILOT = 1 << 2
FLAG_NO_EXCHANGE = 1 << 3
FLAG_HISTOGRAM = 1 << 5
FLAG_OBSERVE = 1 << 6
MODE_AUTO, MODE_SPLIT, MODE_FUSED, MODE_COOP = 0, 1, 2, 3
LADDER_EXACT, LADDER_FAST, LADDER_NEWTON = 0, 1, 2
CSR_AUTO, CSR_LDS, CSR_GLOBAL = 0, 1, 2


class PtxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ptx error {code}: {msg}")
        self.code = code


class ProblemInfo(C.Structure):
    _fields_ = [("n", C.c_int32), ("m", C.c_int64), ("nnz", C.c_int64), ("sparse", C.c_int32),
                ("char_scale", C.c_double), ("n_colours", C.c_int32), ("fmax", C.c_int32),
                ("e_min_bound", C.c_int64), ("e_max_bound", C.c_int64), ("has_field", C.c_int32),
                ("lattice_dim", C.c_int32), ("lattice_L", C.c_int32)]


class EngineConfig(C.Structure):
    _fields_ = [("chains", C.c_int32), ("sweeps", C.c_int32), ("flags", C.c_uint32), ("nc", C.c_int32),
                ("hist_start", C.c_int64), ("mode", C.c_int32), ("waves_per_group", C.c_int32),
                ("steps_per_launch", C.c_int32), ("use_graph", C.c_int32), ("spins_per_lane", C.c_int32),
                ("csr_placement", C.c_int32), ("icm_waves", C.c_int32), ("small_batch_coop", C.c_int32),
                ("coop_wait_us", C.c_int32), ("sliced_sweep", C.c_int32)]


class EngineInfo(C.Structure):
    _fields_ = [("n_instances", C.c_int32), ("steps_done", C.c_int64), ("mode", C.c_int32),
                ("waves_per_group", C.c_int32), ("steps_per_launch", C.c_int32), ("total_replicas", C.c_int64),
                ("spin_updates_per_step", C.c_int64), ("lds_bytes_per_step", C.c_int64), ("csr_in_lds", C.c_int32),
                ("icm_waves", C.c_int32), ("coop_capacity", C.c_int64)]


class InstanceInfo(C.Structure):
    _fields_ = [("n", C.c_int32), ("nt", C.c_int32), ("nrep", C.c_int32), ("nc", C.c_int32),
                ("n_colours", C.c_int32), ("hist_bins", C.c_int32), ("e_base", C.c_int64),
                ("lattice_sweep", C.c_int32), ("sliced_sweep", C.c_int32)]


class SlotObservables(C.Structure):
    _fields_ = [("samples", C.c_int64), ("sum_e", C.c_int64), ("sum_e2", C.c_double), ("sum_up", C.c_int64),
                ("sum_up2", C.c_double), ("sum_abs_m", C.c_int64)]


class TuneTrial(C.Structure):
    _fields_ = [("mode", C.c_int32), ("waves_per_group", C.c_int32), ("spins_per_lane", C.c_int32),
                ("csr_placement", C.c_int32), ("icm_waves", C.c_int32), ("us_per_step", C.c_double)]


class Timing(C.Structure):
    _fields_ = [("sweep_ms", C.c_double), ("exchange_ms", C.c_double), ("icm_ms", C.c_double),
                ("fused_ms", C.c_double), ("sweep_launches", C.c_int64), ("exchange_launches", C.c_int64),
                ("icm_launches", C.c_int64), ("fused_launches", C.c_int64)]


_lib = None

# entry points declared in include/ptx.h, with (restype, argtypes)
_vp = C.c_void_p
_P = C.POINTER
_SIGS = {
    "ptx_last_error": (C.c_int32, [C.c_char_p, C.c_size_t]),
    "ptx_version": (C.c_char_p, []),
    "ptx_device_count": (C.c_int32, []),
    "ptx_problem_read_gset": (C.c_int32, [C.c_char_p, _P(_vp)]),
    "ptx_problem_create_coo": (C.c_int32, [C.c_int32, C.c_int64, _P(C.c_int32), _P(C.c_int32), _P(C.c_double),
                                           _P(C.c_double), C.c_int32, _P(_vp)]),
    "ptx_problem_destroy": (None, [_vp]),
    "ptx_problem_info_get": (C.c_int32, [_vp, _P(ProblemInfo)]),
    "ptx_problem_energy": (C.c_int32, [_vp, _P(C.c_uint8), _P(C.c_int64)]),
    "ptx_problem_maxcut": (C.c_int32, [_vp, _P(C.c_uint8), _P(C.c_int64)]),
    "ptx_problem_delta_e": (C.c_int32, [_vp, _P(C.c_uint8), _P(C.c_int64), _P(C.c_int64)]),
    "ptx_problem_colouring": (C.c_int32, [_vp, _P(C.c_int32), _P(C.c_int32)]),
    "ptx_problem_csr": (C.c_int32, [_vp, _P(C.c_int32), _P(C.c_int32), _P(C.c_int32), _P(C.c_int32)]),
    "ptx_beta_linspace": (C.c_int32, [_P(C.c_double), C.c_double, C.c_double, C.c_int64, C.c_int32]),
    "ptx_icm_count": (C.c_int32, [_P(C.c_double), C.c_int32, C.c_double]),
    "ptx_pilot_betas": (C.c_int32, [C.c_double, C.c_double, C.c_int32, C.c_int32, _P(C.c_double),
                                    _P(C.c_double), _P(C.c_double)]),
    "ptx_ladder_from_histograms": (C.c_int32, [_P(C.c_uint32), C.c_int32, C.c_int64, C.c_int64, _P(C.c_double),
                                               C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                               C.c_int32, _P(C.c_double), _P(C.c_int32)]),
    "ptx_ladder_predict_rates": (C.c_int32, [_P(C.c_uint32), C.c_int32, C.c_int64, C.c_int64, _P(C.c_double),
                                             C.c_double, _P(C.c_double), C.c_int32, C.c_int32, _P(C.c_double)]),
    "ptx_engine_config_default": (None, [_P(EngineConfig)]),
    "ptx_engine_create": (C.c_int32, [_P(_vp), C.c_int32, _P(C.c_double), _P(C.c_int32), C.c_uint64,
                                      _P(EngineConfig), _P(_vp)]),
    "ptx_engine_destroy": (None, [_vp]),
    "ptx_engine_autotune": (C.c_int32, [_P(_vp), C.c_int32, _P(C.c_double), _P(C.c_int32), C.c_uint64,
                                        _P(EngineConfig), C.c_int32, _P(EngineConfig), _P(TuneTrial), C.c_int32,
                                        _P(C.c_int32)]),
    "ptx_engine_step": (C.c_int32, [_vp, C.c_int64]),
    "ptx_engine_sync": (C.c_int32, [_vp]),
    "ptx_engine_set_active": (C.c_int32, [_vp, _P(C.c_uint8)]),
    "ptx_engine_info_get": (C.c_int32, [_vp, _P(EngineInfo)]),
    "ptx_engine_instance_info": (C.c_int32, [_vp, C.c_int32, _P(InstanceInfo)]),
    "ptx_engine_read_spins": (C.c_int32, [_vp, C.c_int32, _P(C.c_uint8)]),
    "ptx_engine_write_spins": (C.c_int32, [_vp, C.c_int32, _P(C.c_uint8)]),
    "ptx_engine_read_energies": (C.c_int32, [_vp, C.c_int32, _P(C.c_int64)]),
    "ptx_engine_read_perm": (C.c_int32, [_vp, C.c_int32, _P(C.c_int32)]),
    "ptx_engine_read_swap_stats": (C.c_int32, [_vp, C.c_int32, _P(C.c_uint32), _P(C.c_uint32), _P(C.c_uint64)]),
    "ptx_engine_read_flips": (C.c_int32, [_vp, C.c_int32, _P(C.c_uint64)]),
    "ptx_engine_read_best": (C.c_int32, [_vp, C.c_int32, _P(C.c_int64), _P(C.c_int64), _P(C.c_uint8)]),
    "ptx_engine_read_histograms": (C.c_int32, [_vp, C.c_int32, _P(C.c_uint32)]),
    "ptx_engine_read_min_energies": (C.c_int32, [_vp, _P(C.c_int64)]),
    "ptx_engine_read_observables": (C.c_int32, [_vp, C.c_int32, _P(SlotObservables)]),
    "ptx_engine_state_size": (C.c_int32, [_vp, _P(C.c_int64)]),
    "ptx_engine_save_state": (C.c_int32, [_vp, _vp, C.c_int64]),
    "ptx_engine_load_state": (C.c_int32, [_vp, _vp, C.c_int64]),
    "ptx_engine_timing_enable": (C.c_int32, [_vp, C.c_int32]),
    "ptx_engine_timing_get": (C.c_int32, [_vp, _P(Timing)]),
    "ptx_engine_phase_profile": (C.c_int32, [_vp, C.c_int32, _P(C.c_double)]),
    "ptx_selftest_fp64": (C.c_int32, [_P(C.c_double), C.c_int64, _P(C.c_double), _P(C.c_double)]),
    "ptx_host_fp64": (C.c_int32, [_P(C.c_double), C.c_int64, _P(C.c_double), _P(C.c_double)]),
    "ptx_selftest_philox": (C.c_int32, [_P(C.c_uint32), C.c_int64, _P(C.c_uint32), _P(C.c_uint32)]),
    "ptx_selftest_copy": (C.c_int32, [C.c_int64, C.c_int32, _P(C.c_double)]),
    "ptx_selftest_lds_bandwidth": (C.c_int32, [C.c_int32, C.c_int32, _P(C.c_double)]),
    "ptx_mlp_create": (C.c_int32, [_P(C.c_int32), C.c_int32, C.c_int32, C.c_uint64, _P(_vp)]),
    "ptx_mlp_destroy": (None, [_vp]),
    "ptx_mlp_param_count": (C.c_int32, [_vp, _P(C.c_int64)]),
    "ptx_mlp_set_params": (C.c_int32, [_vp, _P(C.c_float), C.c_int64]),
    "ptx_mlp_get_params": (C.c_int32, [_vp, _P(C.c_float), C.c_int64]),
    "ptx_mlp_get_grads": (C.c_int32, [_vp, _P(C.c_float), C.c_int64]),
    "ptx_mlp_forward": (C.c_int32, [_vp, _P(C.c_float), C.c_int32, _P(C.c_float)]),
    "ptx_mlp_forward_noalloc": (C.c_int32, [_P(C.c_int32), C.c_int32, _P(C.c_float), C.c_int64, _P(C.c_float),
                                            C.c_int32, _P(C.c_float)]),
    "ptx_mlp_loss_grad": (C.c_int32, [_vp, _P(C.c_float), _P(C.c_float), _P(C.c_uint8), C.c_int32,
                                      _P(C.c_double)]),
    "ptx_mlp_adam": (C.c_int32, [_vp, C.c_float, C.c_float, C.c_float, C.c_float]),
    "ptx_mlp_train": (C.c_int32, [_vp, _P(C.c_float), _P(C.c_float), _P(C.c_uint8), C.c_int64, C.c_int32,
                                  C.c_int32, C.c_float, C.c_uint64, _P(C.c_double)]),
    "ptx_selftest_gemm_bf16": (C.c_int32, [C.c_int32, C.c_int32, C.c_int32, _P(C.c_uint16), _P(C.c_uint16),
                                           _P(C.c_float)]),
}


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libptx.so not built at {LIB_PATH}: run `make -C pt-icm_amd` "
                           "(or __graft_entry__.build()); the engine has no CPU fallback")
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def exported_symbols():
    return list(_SIGS.keys())


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def check(code):
    if code != 0:
        buf = C.create_string_buffer(512)
        lib().ptx_last_error(buf, 512)
        raise PtxError(code, buf.value.decode(errors="replace"))


class Problem:
    """pticm::Problem (include/problem.hh:7-98) over the C ABI."""

    def __init__(self, handle):
        self.h = handle
        info = ProblemInfo()
        check(lib().ptx_problem_info_get(self.h, C.byref(info)))
        self.info = info
        self.n = info.n

    @classmethod
    def read(cls, path):
        h = C.c_void_p()
        check(lib().ptx_problem_read_gset(str(path).encode(), C.byref(h)))
        return cls(h)

    @classmethod
    def from_coo(cls, n, src, dst, w, h=None, dense=False):
        src = np.ascontiguousarray(src, np.int32)
        dst = np.ascontiguousarray(dst, np.int32)
        w = np.ascontiguousarray(w, np.float64)
        hp = None if h is None else _ptr(np.ascontiguousarray(h, np.float64), C.c_double)
        hh = C.c_void_p()
        check(lib().ptx_problem_create_coo(n, len(src), _ptr(src, C.c_int32), _ptr(dst, C.c_int32),
                                           _ptr(w, C.c_double), hp, int(dense), C.byref(hh)))
        return cls(hh)

    def close(self):
        """Releases the problem now (its engines keep their own copy)."""
        h, self.h = getattr(self, "h", None), None
        if h:
            lib().ptx_problem_destroy(h)

    def __del__(self):
        # no library calls while the interpreter shuts down: the HIP runtime
        # (or a profiler attached to it) may already be gone
        if not sys.is_finalizing():
            try:
                self.close()
            except Exception:
                pass

    def energy(self, s):
        s = np.ascontiguousarray(s, np.uint8)
        out = C.c_int64()
        check(lib().ptx_problem_energy(self.h, _ptr(s, C.c_uint8), C.byref(out)))
        return out.value

    def maxcut(self, s):
        s = np.ascontiguousarray(s, np.uint8)
        out = C.c_int64()
        check(lib().ptx_problem_maxcut(self.h, _ptr(s, C.c_uint8), C.byref(out)))
        return out.value

    def delta_e(self, s):
        s = np.ascontiguousarray(s, np.uint8)
        d = np.zeros(self.n, np.int64)
        e = C.c_int64()
        check(lib().ptx_problem_delta_e(self.h, _ptr(s, C.c_uint8), _ptr(d, C.c_int64), C.byref(e)))
        return d, e.value

    def colouring(self):
        p2o = np.zeros(self.n, np.int32)
        cs = np.zeros(self.info.n_colours + 1, np.int32)
        check(lib().ptx_problem_colouring(self.h, _ptr(p2o, C.c_int32), _ptr(cs, C.c_int32)))
        return p2o, cs

    def csr(self):
        rp = np.zeros(self.n + 1, np.int32)
        nnz = max(self.info.nnz, 1)
        col = np.zeros(nnz, np.int32)
        w = np.zeros(nnz, np.int32)
        h = np.zeros(self.n, np.int32)
        check(lib().ptx_problem_csr(self.h, _ptr(rp, C.c_int32), _ptr(col, C.c_int32), _ptr(w, C.c_int32),
                                    _ptr(h, C.c_int32)))
        return rp, col[: self.info.nnz], w[: self.info.nnz], h


def beta_linspace(T0, T1, n, inverse=False):
    out = np.zeros(n, np.float64)
    check(lib().ptx_beta_linspace(_ptr(out, C.c_double), T0, T1, n, int(inverse)))
    return out


def icm_count(beta, char_scale):
    b = np.ascontiguousarray(beta, np.float64)
    return lib().ptx_icm_count(_ptr(b, C.c_double), len(b), char_scale)


def pilot_betas(T0, T1, nt, inverse=False):
    out = np.zeros(nt, np.float64)
    bs, be = C.c_double(), C.c_double()
    check(lib().ptx_pilot_betas(T0, T1, nt, int(inverse), _ptr(out, C.c_double), C.byref(bs), C.byref(be)))
    return out, bs.value, be.value


def ladder_from_histograms(counts, e_base, hist_betas, beta_start, beta_end, estep=2.0, step_size=0.1, r=0.4,
                           mode=LADDER_EXACT):
    counts = np.ascontiguousarray(counts, np.uint32)
    nt, nb = counts.shape
    hb = np.ascontiguousarray(hist_betas, np.float64)
    out = np.zeros(nt, np.float64)
    k = C.c_int32()
    check(lib().ptx_ladder_from_histograms(_ptr(counts, C.c_uint32), nt, nb, e_base, _ptr(hb, C.c_double), estep,
                                           step_size, r, beta_start, beta_end, mode, _ptr(out, C.c_double),
                                           C.byref(k)))
    return out[: k.value].copy()


def ladder_predict_rates(counts, e_base, hist_betas, ladder, estep=2.0, mode=LADDER_FAST):
    counts = np.ascontiguousarray(counts, np.uint32)
    nt, nb = counts.shape
    hb = np.ascontiguousarray(hist_betas, np.float64)
    lad = np.ascontiguousarray(ladder, np.float64)
    out = np.zeros(max(len(lad) - 1, 1), np.float64)
    check(lib().ptx_ladder_predict_rates(_ptr(counts, C.c_uint32), nt, nb, e_base, _ptr(hb, C.c_double), estep,
                                         _ptr(lad, C.c_double), len(lad), mode, _ptr(out, C.c_double)))
    return out[: len(lad) - 1]


def default_config():
    c = EngineConfig()
    lib().ptx_engine_config_default(C.byref(c))
    return c


class Engine:
    """A batch of independent PT-ICM instances on the GPU (one is the plain path)."""

    def __init__(self, problems, ladders, seed=0, chains=2, sweeps=1, flags=0, nc=-1, hist_start=0, mode=MODE_AUTO,
                 waves_per_group=0, steps_per_launch=0, use_graph=0, spins_per_lane=0, csr_placement=CSR_AUTO,
                 icm_waves=0, small_batch_coop=0, coop_wait_us=0, sliced_sweep=0):
        if isinstance(problems, Problem):
            problems = [problems]
            ladders = [ladders]
        self.problems = list(problems)
        self.ladders = [np.ascontiguousarray(b, np.float64) for b in ladders]
        cfg = default_config()
        cfg.chains, cfg.sweeps, cfg.flags, cfg.nc = chains, sweeps, flags, nc
        cfg.hist_start, cfg.mode, cfg.waves_per_group = hist_start, mode, waves_per_group
        cfg.steps_per_launch, cfg.use_graph, cfg.spins_per_lane = steps_per_launch, use_graph, spins_per_lane
        cfg.csr_placement, cfg.icm_waves, cfg.small_batch_coop = csr_placement, icm_waves, small_batch_coop
        cfg.coop_wait_us, cfg.sliced_sweep = coop_wait_us, sliced_sweep
        self.cfg = cfg
        arr = (C.c_void_p * len(self.problems))(*[p.h.value if isinstance(p.h, C.c_void_p) else p.h
                                                  for p in self.problems])
        betas = np.ascontiguousarray(np.concatenate(self.ladders), np.float64)
        nts = np.array([len(b) for b in self.ladders], np.int32)
        h = C.c_void_p()
        check(lib().ptx_engine_create(C.cast(arr, C.POINTER(C.c_void_p)), len(self.problems), _ptr(betas, C.c_double),
                                      _ptr(nts, C.c_int32), seed, C.byref(cfg), C.byref(h)))
        self.h = h
        self.ninst = len(self.problems)

    def close(self):
        """Releases the engine's device memory, stream and graphs now."""
        h, self.h = getattr(self, "h", None), None
        if h:
            lib().ptx_engine_destroy(h)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        if not sys.is_finalizing():
            try:
                self.close()
            except Exception:
                pass

    def step(self, n=1):
        check(lib().ptx_engine_step(self.h, n))

    def sync(self):
        check(lib().ptx_engine_sync(self.h))

    def set_active(self, active):
        """Later steps run only the instances with a nonzero flag."""
        a = np.ascontiguousarray(np.asarray(active, bool).astype(np.uint8))
        assert len(a) == self.ninst
        check(lib().ptx_engine_set_active(self.h, _ptr(a, C.c_uint8)))

    def info(self):
        i = EngineInfo()
        check(lib().ptx_engine_info_get(self.h, C.byref(i)))
        return i

    def instance_info(self, inst=0):
        i = InstanceInfo()
        check(lib().ptx_engine_instance_info(self.h, inst, C.byref(i)))
        return i

    def spins(self, inst=0):
        ii = self.instance_info(inst)
        out = np.zeros((ii.nrep, ii.n), np.uint8)
        check(lib().ptx_engine_read_spins(self.h, inst, _ptr(out, C.c_uint8)))
        return out

    def set_spins(self, s, inst=0):
        s = np.ascontiguousarray(s, np.uint8)
        check(lib().ptx_engine_write_spins(self.h, inst, _ptr(s, C.c_uint8)))

    def energies(self, inst=0):
        ii = self.instance_info(inst)
        out = np.zeros(ii.nrep, np.int64)
        check(lib().ptx_engine_read_energies(self.h, inst, _ptr(out, C.c_int64)))
        return out

    def slot2rep(self, inst=0):
        ii = self.instance_info(inst)
        out = np.zeros(ii.nrep, np.int32)
        check(lib().ptx_engine_read_perm(self.h, inst, _ptr(out, C.c_int32)))
        return out

    def swap_stats(self, inst=0):
        ii = self.instance_info(inst)
        a = np.zeros(ii.nrep, np.uint32)
        s = np.zeros(ii.nrep, np.uint32)
        n = C.c_uint64()
        check(lib().ptx_engine_read_swap_stats(self.h, inst, _ptr(a, C.c_uint32), _ptr(s, C.c_uint32), C.byref(n)))
        return a, s, n.value

    def flips(self, inst=0):
        ii = self.instance_info(inst)
        out = np.zeros(ii.nrep, np.uint64)
        check(lib().ptx_engine_read_flips(self.h, inst, _ptr(out, C.c_uint64)))
        return out

    def best(self, inst=0, with_state=True):
        ii = self.instance_info(inst)
        be = np.zeros(ii.nrep, np.int64)
        bs = np.zeros(ii.nrep, np.int64)
        st = np.zeros((ii.nrep, ii.n), np.uint8)
        check(lib().ptx_engine_read_best(self.h, inst, _ptr(be, C.c_int64), _ptr(bs, C.c_int64),
                                         _ptr(st, C.c_uint8) if with_state else None))
        return be, bs, st

    def histograms(self, inst=0):
        ii = self.instance_info(inst)
        out = np.zeros((ii.nrep, ii.hist_bins), np.uint32)
        check(lib().ptx_engine_read_histograms(self.h, inst, _ptr(out, C.c_uint32)))
        return out, ii.e_base

    def observables(self, inst=0):
        """Per-slot sums of --log: dict of arrays samples, sum_e, sum_e2, sum_up, sum_up2, sum_abs_m."""
        ii = self.instance_info(inst)
        arr = (SlotObservables * ii.nrep)()
        check(lib().ptx_engine_read_observables(self.h, inst, arr))
        return {f: np.array([getattr(x, f) for x in arr]) for f, _ in SlotObservables._fields_}

    def save_state(self):
        """The engine's whole mutable state as bytes (checkpoint)."""
        n = C.c_int64()
        check(lib().ptx_engine_state_size(self.h, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        check(lib().ptx_engine_save_state(self.h, buf, n.value))
        return buf.raw

    def load_state(self, blob):
        """Restores a checkpoint taken by an engine of the same shape."""
        buf = C.create_string_buffer(bytes(blob), len(blob))
        check(lib().ptx_engine_load_state(self.h, buf, len(blob)))

    def min_energies(self):
        out = np.zeros(self.ninst, np.int64)
        check(lib().ptx_engine_read_min_energies(self.h, _ptr(out, C.c_int64)))
        return out

    def phase_profile(self, on=None):
        """on=True/False resets and switches; None reads [sweep, exchange, icm, barrier] microseconds
        (workgroup 0)."""
        us = np.zeros(4, np.float64)
        check(lib().ptx_engine_phase_profile(self.h, -1 if on is None else int(on), _ptr(us, C.c_double)))
        return us

    def timing(self, on=None):
        if on is not None:
            check(lib().ptx_engine_timing_enable(self.h, int(on)))
            return None
        t = Timing()
        check(lib().ptx_engine_timing_get(self.h, C.byref(t)))
        return t


def autotune(problems, ladders, seed=0, steps=32, **cfg_kw):
    """Times every launch shape that fits the batch; returns (Engine keyword
    arguments of the fastest shape, list of trial dicts)."""
    if isinstance(problems, Problem):
        problems = [problems]
        ladders = [ladders]
    base = default_config()
    for k, v in cfg_kw.items():
        setattr(base, k, v)
    arr = (C.c_void_p * len(problems))(*[p.h.value if isinstance(p.h, C.c_void_p) else p.h for p in problems])
    betas = np.ascontiguousarray(np.concatenate([np.asarray(b, np.float64) for b in ladders]), np.float64)
    nts = np.array([len(b) for b in ladders], np.int32)
    best = EngineConfig()
    tr = (TuneTrial * 64)()
    n = C.c_int32()
    check(lib().ptx_engine_autotune(C.cast(arr, C.POINTER(C.c_void_p)), len(problems), _ptr(betas, C.c_double),
                                    _ptr(nts, C.c_int32), seed, C.byref(base), int(steps), C.byref(best), tr, 64,
                                    C.byref(n)))
    trials = [{"mode": t.mode, "waves_per_group": t.waves_per_group, "spins_per_lane": t.spins_per_lane,
               "csr_placement": t.csr_placement, "icm_waves": t.icm_waves, "us_per_step": t.us_per_step}
              for t in tr[: n.value]]
    kw = {"mode": best.mode, "waves_per_group": best.waves_per_group, "spins_per_lane": best.spins_per_lane,
          "csr_placement": best.csr_placement, "icm_waves": best.icm_waves}
    return kw, trials


def device_fp64(x):
    x = np.ascontiguousarray(x, np.float64)
    e = np.zeros_like(x)
    lg = np.zeros_like(x)
    check(lib().ptx_selftest_fp64(_ptr(x, C.c_double), len(x), _ptr(e, C.c_double), _ptr(lg, C.c_double)))
    return e, lg


def host_fp64(x):
    x = np.ascontiguousarray(x, np.float64)
    e = np.zeros_like(x)
    lg = np.zeros_like(x)
    check(lib().ptx_host_fp64(_ptr(x, C.c_double), len(x), _ptr(e, C.c_double), _ptr(lg, C.c_double)))
    return e, lg


def device_philox(ctr, key):
    ctr = np.ascontiguousarray(ctr, np.uint32).reshape(-1, 4)
    key = np.ascontiguousarray(key, np.uint32)
    out = np.zeros_like(ctr)
    check(lib().ptx_selftest_philox(_ptr(ctr, C.c_uint32), len(ctr), _ptr(key, C.c_uint32), _ptr(out, C.c_uint32)))
    return out


def copy_ms(nbytes, reps=20):
    """Average milliseconds of the k_copy16 kernel copying nbytes."""
    out = C.c_double()
    check(lib().ptx_selftest_copy(int(nbytes), int(reps), C.byref(out)))
    return out.value


def lds_bandwidth_gbs(blocks=4096, iters=4096):
    """Aggregate LDS read bandwidth (GB/s) of the k_lds_read microbenchmark."""
    out = C.c_double()
    check(lib().ptx_selftest_lds_bandwidth(int(blocks), int(iters), C.byref(out)))
    return out.value


class Mlp:
    """The ladder surrogate network on the GPU (BF16 MFMA kernels, FP32
    accumulate).  Parameters are a flat float32 vector: per layer W[in][out]
    then b[out]."""

    def __init__(self, dims, max_batch=256, seed=0):
        self.dims = [int(d) for d in dims]
        d = np.ascontiguousarray(self.dims, np.int32)
        h = C.c_void_p()
        check(lib().ptx_mlp_create(_ptr(d, C.c_int32), len(d), int(max_batch), int(seed), C.byref(h)))
        self.h = h
        n = C.c_int64()
        check(lib().ptx_mlp_param_count(self.h, C.byref(n)))
        self.nparams = n.value

    def close(self):
        """Releases the network's device memory and stream now."""
        h, self.h = getattr(self, "h", None), None
        if h:
            lib().ptx_mlp_destroy(h)

    def __del__(self):
        if not sys.is_finalizing():
            try:
                self.close()
            except Exception:
                pass

    def params(self):
        out = np.zeros(self.nparams, np.float32)
        check(lib().ptx_mlp_get_params(self.h, _ptr(out, C.c_float), self.nparams))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        check(lib().ptx_mlp_set_params(self.h, _ptr(p, C.c_float), len(p)))

    def grads(self):
        out = np.zeros(self.nparams, np.float32)
        check(lib().ptx_mlp_get_grads(self.h, _ptr(out, C.c_float), self.nparams))
        return out

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((x.shape[0], self.dims[-1]), np.float32)
        check(lib().ptx_mlp_forward(self.h, _ptr(x, C.c_float), x.shape[0], _ptr(y, C.c_float)))
        return y

    def loss_grad(self, x, t, mode):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        m = np.ascontiguousarray(mode, np.uint8)
        loss = C.c_double()
        check(lib().ptx_mlp_loss_grad(self.h, _ptr(x, C.c_float), _ptr(t, C.c_float), _ptr(m, C.c_uint8),
                                      x.shape[0], C.byref(loss)))
        return loss.value

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        check(lib().ptx_mlp_adam(self.h, lr, beta1, beta2, eps))

    def train(self, x, t, mode, batch=128, epochs=10, lr=1e-3, seed=0):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        m = np.ascontiguousarray(mode, np.uint8)
        losses = np.zeros(epochs, np.float64)
        check(lib().ptx_mlp_train(self.h, _ptr(x, C.c_float), _ptr(t, C.c_float), _ptr(m, C.c_uint8), x.shape[0],
                                  int(batch), int(epochs), float(lr), int(seed), _ptr(losses, C.c_double)))
        return losses


def mlp_forward_noalloc(dims, params, x):
    """The network's forward pass on the library's fixed device array (no
    allocation); same result as Mlp.forward with the same parameters."""
    d = np.ascontiguousarray(dims, np.int32)
    p = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    y = np.zeros((x.shape[0], int(d[-1])), np.float32)
    check(lib().ptx_mlp_forward_noalloc(_ptr(d, C.c_int32), len(d), _ptr(p, C.c_float), len(p), _ptr(x, C.c_float),
                                        x.shape[0], _ptr(y, C.c_float)))
    return y


def gemm_bf16(a_bits, b_bits):
    """C = A . B on the MFMA kernel; operands as raw BF16 bit patterns (uint16)."""
    a = np.ascontiguousarray(a_bits, np.uint16)
    b = np.ascontiguousarray(b_bits, np.uint16)
    M, K = a.shape
    K2, N = b.shape
    assert K == K2
    c = np.zeros((M, N), np.float32)
    check(lib().ptx_selftest_gemm_bf16(M, N, K, _ptr(a, C.c_uint16), _ptr(b, C.c_uint16), _ptr(c, C.c_float)))
    return c
