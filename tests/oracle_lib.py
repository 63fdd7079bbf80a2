"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: tests, the smoke check and bench.py's CPU-baseline
leg use it as the checker.  The product never imports this module.
"""
import ctypes as C
import os

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(REPO, "oracle", "liboracle.so")

FLAG_ICM = 1
FLAG_REFERENCE_ORDER = 2
FLAG_PILOT = 4
FLAG_NO_EXCHANGE = 8
FLAG_LOG_ENERGY = 16

_lib = None


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("oracle library missing: run `make -C oracle` (or __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.or_philox4x32_10.argtypes = [C.POINTER(C.c_uint32)] * 3
    L.or_exp.restype = C.c_double
    L.or_exp.argtypes = [C.c_double]
    L.or_log.restype = C.c_double
    L.or_log.argtypes = [C.c_double]
    L.or_uniform_log.restype = C.c_double
    L.or_uniform_log.argtypes = [C.c_uint32]
    L.or_read_gset.restype = vp
    L.or_read_gset.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
    L.or_graph_from_coo.restype = vp
    L.or_graph_from_coo.argtypes = [C.c_int32, C.c_int64, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                    C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int, C.c_char_p, C.c_int]
    L.or_graph_free.argtypes = [vp]
    L.or_graph_n.restype = C.c_int32
    L.or_graph_n.argtypes = [vp]
    L.or_graph_nnz.restype = C.c_int64
    L.or_graph_nnz.argtypes = [vp]
    L.or_graph_m_header.restype = C.c_int64
    L.or_graph_m_header.argtypes = [vp]
    L.or_graph_sparse.restype = C.c_int
    L.or_graph_sparse.argtypes = [vp]
    L.or_graph_char_scale.restype = C.c_double
    L.or_graph_char_scale.argtypes = [vp]
    L.or_graph_csr.argtypes = [vp] + [C.POINTER(C.c_int32)] * 4
    L.or_energy.restype = C.c_int64
    L.or_energy.argtypes = [vp, C.POINTER(C.c_uint8)]
    L.or_maxcut.restype = C.c_int64
    L.or_maxcut.argtypes = [vp, C.POINTER(C.c_uint8)]
    L.or_delta_e.argtypes = [vp, C.POINTER(C.c_uint8), C.POINTER(C.c_int64)]
    L.or_incremental_de.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_uint8), C.c_int32]
    L.or_colour.restype = C.c_int32
    L.or_colour.argtypes = [vp, C.POINTER(C.c_int32)]
    L.or_positions.argtypes = [vp, C.POINTER(C.c_int32), C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    L.or_beta_linspace.argtypes = [C.POINTER(C.c_double), C.c_double, C.c_double, C.c_int64, C.c_int]
    L.or_icm_count.restype = C.c_int32
    L.or_icm_count.argtypes = [C.POINTER(C.c_double), C.c_int32, C.c_double]
    L.or_threshold_table.argtypes = [C.POINTER(C.c_double), C.c_int32, C.c_int32, C.POINTER(C.c_uint32)]
    L.or_fmax.restype = C.c_int32
    L.or_fmax.argtypes = [vp]
    L.or_engine_create.restype = vp
    L.or_engine_create.argtypes = [vp, C.POINTER(C.c_double), C.c_int32, C.c_int32, C.c_uint64, C.c_uint32,
                                   C.c_uint32, C.c_int32, C.c_int32]
    L.or_engine_free.argtypes = [vp]
    L.or_engine_step.argtypes = [vp, C.c_int64]
    L.or_engine_nrep.restype = C.c_int32
    L.or_engine_nrep.argtypes = [vp]
    L.or_engine_nc.restype = C.c_int32
    L.or_engine_nc.argtypes = [vp]
    L.or_engine_spins.argtypes = [vp, C.POINTER(C.c_uint8)]
    L.or_engine_set_spins.argtypes = [vp, C.POINTER(C.c_uint8)]
    L.or_engine_energies.argtypes = [vp, C.POINTER(C.c_int64)]
    L.or_engine_slot2rep.argtypes = [vp, C.POINTER(C.c_int32)]
    L.or_engine_swap_stats.argtypes = [vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)]
    L.or_engine_flips.argtypes = [vp, C.POINTER(C.c_uint64)]
    L.or_engine_best.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_uint8)]
    L.or_engine_log_len.restype = C.c_int64
    L.or_engine_log_len.argtypes = [vp]
    L.or_engine_energy_log.argtypes = [vp, C.POINTER(C.c_int32)]
    L.or_engine_icm_trace.argtypes = [vp] + [C.POINTER(C.c_int32)] * 3
    L.or_make_histogram.restype = C.c_int64
    L.or_make_histogram.argtypes = [C.POINTER(C.c_int32), C.c_int64, C.c_int64, C.c_double,
                                    C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_uint32), C.c_int64]
    L.or_get_beta_from_logs.restype = C.c_int32
    L.or_get_beta_from_logs.argtypes = [C.POINTER(C.c_int32), C.c_int64, C.c_int32, C.POINTER(C.c_double),
                                        C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                        C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int64,
                                        C.POINTER(C.c_int64), C.POINTER(C.c_double)]
    L.or_acceptance_rate.restype = C.c_double
    L.or_acceptance_rate.argtypes = [C.POINTER(C.c_double)] * 3 + [C.c_int64, C.c_double, C.c_double,
                                                                   C.c_double, C.c_double]
    L.or_partition_func.restype = C.c_double
    L.or_partition_func.argtypes = [C.POINTER(C.c_double), C.c_int64, C.c_double, C.c_double, C.c_double]
    L.or_gather_samples.argtypes = [vp, C.POINTER(C.c_double), C.c_int32, C.c_uint64, C.c_uint32, C.c_int64,
                                    C.c_int64, C.POINTER(C.c_int32), C.c_uint32]
    _lib = L
    return L


def philox(ctr, key):
    c = np.array(ctr, dtype=np.uint32)
    k = np.array(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().or_philox4x32_10(_p(c, C.c_uint32), _p(k, C.c_uint32), _p(o, C.c_uint32))
    return o


class Graph:
    def __init__(self, handle):
        if not handle:
            raise ValueError("oracle graph creation failed")
        self.h = C.c_void_p(handle)
        L = lib()
        self.n = L.or_graph_n(self.h)
        self.nnz = L.or_graph_nnz(self.h)

    @classmethod
    def read(cls, path):
        err = C.create_string_buffer(256)
        h = lib().or_read_gset(path.encode(), err, 256)
        if not h:
            raise ValueError(err.value.decode())
        return cls(h)

    @classmethod
    def from_coo(cls, n, u, v, w, h=None, dense=False):
        u = np.ascontiguousarray(u, dtype=np.int32)
        v = np.ascontiguousarray(v, dtype=np.int32)
        w = np.ascontiguousarray(w, dtype=np.float64)
        hh = np.zeros(n, dtype=np.float64) if h is None else np.ascontiguousarray(h, dtype=np.float64)
        err = C.create_string_buffer(256)
        handle = lib().or_graph_from_coo(n, len(u), _p(u, C.c_int32), _p(v, C.c_int32), _p(w, C.c_double),
                                         _p(hh, C.c_double), int(dense), err, 256)
        if not handle:
            raise ValueError(err.value.decode())
        return cls(handle)

    def __del__(self):
        try:
            lib().or_graph_free(self.h)
        except Exception:
            pass

    @property
    def sparse(self):
        return bool(lib().or_graph_sparse(self.h))

    @property
    def char_scale(self):
        return lib().or_graph_char_scale(self.h)

    @property
    def m_header(self):
        return lib().or_graph_m_header(self.h)

    def csr(self):
        rp = np.zeros(self.n + 1, np.int32)
        col = np.zeros(max(self.nnz, 1), np.int32)
        w = np.zeros(max(self.nnz, 1), np.int32)
        h = np.zeros(self.n, np.int32)
        lib().or_graph_csr(self.h, _p(rp, C.c_int32), _p(col, C.c_int32), _p(w, C.c_int32), _p(h, C.c_int32))
        return rp, col[: self.nnz], w[: self.nnz], h

    def energy(self, s):
        s = np.ascontiguousarray(s, dtype=np.uint8)
        return lib().or_energy(self.h, _p(s, C.c_uint8))

    def maxcut(self, s):
        s = np.ascontiguousarray(s, dtype=np.uint8)
        return lib().or_maxcut(self.h, _p(s, C.c_uint8))

    def delta_e(self, s):
        s = np.ascontiguousarray(s, dtype=np.uint8)
        d = np.zeros(self.n, np.int64)
        lib().or_delta_e(self.h, _p(s, C.c_uint8), _p(d, C.c_int64))
        return d

    def incremental_de(self, dE, s, i):
        s = np.ascontiguousarray(s, dtype=np.uint8)
        lib().or_incremental_de(self.h, _p(dE, C.c_int64), _p(s, C.c_uint8), int(i))

    def colour(self):
        c = np.zeros(self.n, np.int32)
        k = lib().or_colour(self.h, _p(c, C.c_int32))
        return k, c

    def positions(self):
        k, c = self.colour()
        p2o = np.zeros(self.n, np.int32)
        cs = np.zeros(k + 1, np.int32)
        lib().or_positions(self.h, _p(c, C.c_int32), k, _p(p2o, C.c_int32), _p(cs, C.c_int32))
        return p2o, cs

    def fmax(self):
        return lib().or_fmax(self.h)


def beta_linspace(T0, T1, n, inverse):
    out = np.zeros(n, np.float64)
    lib().or_beta_linspace(_p(out, C.c_double), T0, T1, n, int(inverse))
    return out


def icm_count(beta, char_scale):
    b = np.ascontiguousarray(beta, dtype=np.float64)
    return lib().or_icm_count(_p(b, C.c_double), len(b), char_scale)


def threshold_table(beta, fmax):
    b = np.ascontiguousarray(beta, dtype=np.float64)
    t = np.zeros((len(b), max(fmax, 1)), np.uint32)
    lib().or_threshold_table(_p(b, C.c_double), len(b), fmax, _p(t, C.c_uint32))
    return t[:, :fmax]


class Engine:
    def __init__(self, graph, betas, chains=2, seed=0, instance=0, flags=0, sweeps=1, nc=-1):
        self.g = graph
        self.betas = np.ascontiguousarray(betas, dtype=np.float64)
        self.nt = len(self.betas)
        self.chains = chains
        h = lib().or_engine_create(graph.h, _p(self.betas, C.c_double), self.nt, chains, seed, instance, flags,
                                   sweeps, nc)
        if not h:
            raise ValueError("oracle engine creation failed")
        self.h = C.c_void_p(h)
        self.nrep = lib().or_engine_nrep(self.h)
        self.nc = lib().or_engine_nc(self.h)

    def __del__(self):
        try:
            lib().or_engine_free(self.h)
        except Exception:
            pass

    def step(self, n=1):
        lib().or_engine_step(self.h, n)

    def spins(self):
        s = np.zeros((self.nrep, self.g.n), np.uint8)
        lib().or_engine_spins(self.h, _p(s, C.c_uint8))
        return s

    def set_spins(self, s):
        s = np.ascontiguousarray(s, dtype=np.uint8)
        lib().or_engine_set_spins(self.h, _p(s, C.c_uint8))

    def energies(self):
        e = np.zeros(self.nrep, np.int64)
        lib().or_engine_energies(self.h, _p(e, C.c_int64))
        return e

    def slot2rep(self):
        s = np.zeros(self.nrep, np.int32)
        lib().or_engine_slot2rep(self.h, _p(s, C.c_int32))
        return s

    def swap_stats(self):
        a = np.zeros(self.nrep, np.uint32)
        s = np.zeros(self.nrep, np.uint32)
        n = C.c_uint64(0)
        lib().or_engine_swap_stats(self.h, _p(a, C.c_uint32), _p(s, C.c_uint32), C.byref(n))
        return a, s, n.value

    def flips(self):
        f = np.zeros(self.nrep, np.uint64)
        lib().or_engine_flips(self.h, _p(f, C.c_uint64))
        return f

    def best(self):
        be = np.zeros(self.nrep, np.int64)
        bs = np.zeros(self.nrep, np.int64)
        st = np.zeros((self.nrep, self.g.n), np.uint8)
        lib().or_engine_best(self.h, _p(be, C.c_int64), _p(bs, C.c_int64), _p(st, C.c_uint8))
        return be, bs, st

    def energy_log(self):
        n = lib().or_engine_log_len(self.h)
        out = np.zeros(max(n, 1), np.int32)
        lib().or_engine_energy_log(self.h, _p(out, C.c_int32))
        return out[:n].reshape(-1, self.nrep)

    def icm_trace(self):
        a = np.zeros(self.nt, np.int32)
        b = np.zeros(self.nt, np.int32)
        c = np.zeros(self.nt, np.int32)
        lib().or_engine_icm_trace(self.h, _p(a, C.c_int32), _p(b, C.c_int32), _p(c, C.c_int32))
        return a, b, c


def gather_samples(graph, hist_betas, seed, burn_in, nsamples, instance=0, extra_flags=0):
    hb = np.ascontiguousarray(hist_betas, dtype=np.float64)
    out = np.zeros((nsamples, len(hb)), np.int32)
    lib().or_gather_samples(graph.h, _p(hb, C.c_double), len(hb), seed, instance, burn_in, nsamples,
                            _p(out, C.c_int32), extra_flags)
    return out


def hist_betas(T0, T1, nt, inverse):
    """hist_betas of get_beta (cae_estimator.cc:309-328): accumulated steps
    from beta_start = max(b0, b1) down to beta_end."""
    if inverse:
        bs, be = T0, T1
    else:
        bs, be = 1.0 / T0, 1.0 / T1
    if bs < be:
        bs, be = be, bs
    step = (be - bs) / (nt - 1)
    out = np.zeros(nt)
    b = bs
    for i in range(nt):
        out[i] = b
        b += step
    return out, bs, be


def get_beta_from_logs(elog, hist_betas_, estep, step_size, r, beta_start, beta_end, logdos_cap=1 << 20):
    elog = np.ascontiguousarray(elog, dtype=np.int32)
    ns, nt = elog.shape
    hb = np.ascontiguousarray(hist_betas_, dtype=np.float64)
    out = np.zeros(nt, np.float64)
    logdos = np.zeros(logdos_cap, np.float64)
    nb = C.c_int64(0)
    logz = np.zeros(nt, np.float64)
    k = lib().or_get_beta_from_logs(_p(elog, C.c_int32), ns, nt, _p(hb, C.c_double), estep, step_size, r,
                                    beta_start, beta_end, _p(out, C.c_double), _p(logdos, C.c_double), logdos_cap,
                                    C.byref(nb), _p(logz, C.c_double))
    return out[:k], logdos[: min(nb.value, logdos_cap)], logz


def acceptance_rate(logdos, p1, p2, b1, b2, estep, emin):
    logdos = np.ascontiguousarray(logdos, np.float64)
    p1 = np.ascontiguousarray(p1, np.float64)
    p2 = np.ascontiguousarray(p2, np.float64)
    return lib().or_acceptance_rate(_p(logdos, C.c_double), _p(p1, C.c_double), _p(p2, C.c_double), len(logdos),
                                    b1, b2, estep, emin)


def partition_func(logdos, beta, emin, estep):
    logdos = np.ascontiguousarray(logdos, np.float64)
    return lib().or_partition_func(_p(logdos, C.c_double), len(logdos), beta, emin, estep)
