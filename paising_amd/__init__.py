"""MI355X population-annealing engine: Python binding of libpaising.so (C ABI, paising.h).

The engine is native (gfx950 HIP kernels + C++ host engine); this module only loads the
in-tree shared library with ctypes and converts arguments.  There is no Python or CPU
fallback: if the library is missing or no GPU is present, calls raise.
"""
import ctypes
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpaising.so")
BIN_DIR = os.path.join(HERE, "bin")

SSC, MSC = 0, 1
SERVE_SHARE, SERVE_SHORTEST = 0, 1  # server tick policies (pa_server_set_policy)
UNIFORM_CUDA101, UNIFORM_CUDA100 = 0, 1
OK, ERR_INVALID, ERR_HIP, ERR_NBMAX, ERR_CAPACITY, ERR_MEMORY, ERR_STATE = range(7)


class PAError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"paising error {code}: {msg}")
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [
        ("L", ctypes.c_int32), ("coding", ctypes.c_int32), ("msbits", ctypes.c_int32),
        ("Rinit", ctypes.c_int32), ("eqsweeps", ctypes.c_uint32), ("dB", ctypes.c_double),
        ("Binit", ctypes.c_double), ("Bfin", ctypes.c_double), ("seed", ctypes.c_uint64),
        ("initial_sequence", ctypes.c_uint64), ("uniform", ctypes.c_int32),
        ("max_steps", ctypes.c_int32), ("trace", ctypes.c_int32), ("slice_bits", ctypes.c_int32),
        ("threads", ctypes.c_int32), ("balance", ctypes.c_int32), ("adaptive", ctypes.c_int32),
        ("schedule_len", ctypes.c_int32), ("min_overlap", ctypes.c_double),
        ("max_overlap", ctypes.c_double), ("schedule", ctypes.POINTER(ctypes.c_double)),
        ("mhr", ctypes.c_int32), ("pad1", ctypes.c_int32),
    ]


class Row(ctypes.Structure):
    _fields_ = [
        ("B", ctypes.c_double), ("E", ctypes.c_double), ("C", ctypes.c_double),
        ("M", ctypes.c_double), ("M2", ctypes.c_double), ("M4", ctypes.c_double),
        ("BF", ctypes.c_double), ("S", ctypes.c_double), ("R", ctypes.c_int32),
        ("Rprev", ctypes.c_int32), ("lnQ", ctypes.c_double), ("q", ctypes.c_double),
        ("sums", ctypes.c_double * 5), ("sumE", ctypes.c_int64), ("sumM", ctypes.c_int64),
        ("roff_hash", ctypes.c_uint64), ("spin_hash", ctypes.c_uint64),
        ("IS_after", ctypes.c_uint64),
    ]


# exported entry points and their ctypes signatures (paising.h)
_P, _I, _D, _U64, _SZ = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_uint64, ctypes.c_size_t
_PP = ctypes.POINTER(ctypes.c_void_p)
SIGNATURES = {
    "pa_last_error": (ctypes.c_char_p, []),
    "pa_version": (_I, []),
    "pa_device_info": (_I, [ctypes.c_char_p, _SZ, _P, _P]),
    "pa_schedule_steps": (_I, [_D, _D, _D]),
    "pa_bandwidth_probe": (_I, [_SZ, _I, _P]),
    "pa_run_create": (_I, [_P, _PP]),
    "pa_run_step": (_I, [_P, _I]),
    "pa_run_sync": (_I, [_P]),
    "pa_run_steps_done": (_I, [_P, _P, _P]),
    "pa_run_results": (_I, [_P, _P, _I, _P]),
    "pa_run_final_sequence": (_I, [_P, _P]),
    "pa_run_population": (_I, [_P, _P, _P]),
    "pa_run_array": (_I, [_P, _I, _P, _P]),
    "pa_run_timing": (_I, [_P, _P, _P]),
    "pa_run_kernel_times": (_I, [_P, _P, _P, _P]),
    "pa_run_set_timing": (_I, [_P, _I]),
    "pa_run_set_graph": (_I, [_P, _I]),
    "pa_run_destroy": (_I, [_P]),
    "pa_run_all": (_I, [_P, _P, _I, _P, _P]),
    "pa_run_all_hist": (_I, [_P, _P, _I, _P, _P, _P, _I]),
    "pa_run_histogram": (_I, [_P, _P, _I]),
    "pa_mhr_analysis": (_I, [_I, _I, _P, _P, _P, _P, _I, _P]),
    "pa_batch_create": (_I, [_P, _I, _I, _PP]),
    "pa_batch_run": (_I, [_P]),
    "pa_batch_results": (_I, [_P, _I, _P, _I, _P]),
    "pa_batch_time": (_I, [_P, _P]),
    "pa_batch_destroy": (_I, [_P]),
    "pa_server_create": (_I, [_SZ, _I, _PP]),
    "pa_server_submit": (_I, [_P, _P, _P]),
    "pa_server_tick": (_I, [_P, _P]),
    "pa_server_poll": (_I, [_P, ctypes.c_int64, _P]),
    "pa_server_results": (_I, [_P, ctypes.c_int64, _P, _I, _P]),
    "pa_server_stats": (_I, [_P, _P, _P, _P]),
    "pa_server_set_chunk": (_I, [_P, _I, _I]),
    "pa_server_set_policy": (_I, [_P, _I]),
    "pa_stage_sweep": (_I, [_P, _I, ctypes.c_uint64, ctypes.c_double, _P, _P, _P]),
    "pa_stage_energy": (_I, [_P, _I, _P, _P, _P]),
    "pa_stage_copy_counts": (_I, [_P, _I, _P, _P, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_uint64, _P, _P, _P, _P, _P]),
    "pa_stage_resample": (_I, [_P, _I, _P, _P, _P, _P, _P, _P]),
    "pa_stage_averages": (_I, [_P, _I, _P, _P, _P, _P, _P, _P]),
    "pa_server_latency": (_I, [_P, ctypes.c_int64, _P, _P]),
    "pa_server_reserve": (_I, [_P, _SZ]),
    "pa_stage_sweep_noalloc": (_I, [_P, _I, ctypes.c_uint64, ctypes.c_double, _P, _P, _P, _P]),
    "pa_server_memory": (_I, [_P, _P, _P]),
    "pa_server_destroy": (_I, [_P]),
    "pa_run_footprint": (_SZ, [_P]),
    "pa_autotune": (_I, [_P, _P, _P]),
    "pa_autotune_clear": (_I, []),
}

_lib = None


def lib():
    """Load libpaising.so from this package directory (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with paising_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != OK:
        raise PAError(rc, lib().pa_last_error().decode(errors="replace"))


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def make_params(L=64, coding=MSC, msbits=32, Rinit=20000, eqsweeps=100, dB=0.005, Binit=0.0,
                Bfin=1.0, seed=100, initial_sequence=0, uniform=UNIFORM_CUDA101, max_steps=0,
                trace=0, slice_bits=0, threads=0, balance=0, adaptive=0, min_overlap=0.0,
                max_overlap=0.0, schedule=None, mhr=0):
    """pa_params.  adaptive=1: the reference's adaptive beta step; schedule: an explicit
    list of betas (Binit first) to replay, e.g. run 0's adaptive steps for later runs."""
    p = Params(L, coding, msbits, Rinit, eqsweeps, dB, Binit, Bfin, seed & (2**64 - 1),
               initial_sequence, uniform, max_steps, trace, slice_bits, threads, balance,
               adaptive, 0, min_overlap, max_overlap, None, mhr, 0)
    if schedule is not None:
        arr = (ctypes.c_double * len(schedule))(*schedule)
        p._schedule_keep = arr  # the library copies it at creation; keep it alive until then
        p.schedule = ctypes.cast(arr, ctypes.POINTER(ctypes.c_double))
        p.schedule_len = len(schedule)
    return p


def autotune(params):
    """Time the sweep variants for this run shape; returns (tuned params, best sweep ms).
    The choice is cached in the library for later runs of the same shape."""
    out = Params()
    ms = ctypes.c_double(0)
    _check(lib().pa_autotune(ctypes.byref(params), ctypes.byref(out), ctypes.byref(ms)))
    return out, ms.value


def schedule_steps(Binit, dB, Bfin):
    return lib().pa_schedule_steps(Binit, dB, Bfin)


def footprint(params):
    return lib().pa_run_footprint(ctypes.byref(params))


def _members(params, R):
    return R if params.coding == SSC else -(-R // params.msbits) * params.msbits


def _pop_shape(params, R):
    N = params.L * params.L
    if params.coding == SSC:
        return (R, N), np.int8
    return (-(-R // params.msbits), N), {8: np.uint8, 16: np.uint16, 32: np.uint32,
                                         64: np.uint64}[params.msbits]


def stage_sweep(params, pop, R, IS, beta):
    """checkKerALL + energyKer on a copy of `pop`: returns (new pop, IE, M)."""
    shape, dt = _pop_shape(params, R)
    out = np.ascontiguousarray(np.array(pop, dtype=dt).reshape(shape))
    IE = np.zeros(_members(params, R), np.int32)
    M = np.zeros_like(IE)
    _check(lib().pa_stage_sweep(ctypes.byref(params), R, IS, beta, _ptr(out), _ptr(IE), _ptr(M)))
    return out, IE, M


def stage_sweep_noalloc(params, R, IS, beta, pop_ptr, ie_ptr, m_ptr, stream=0):
    """pa_stage_sweep_noalloc on caller-owned device buffers (addresses as integers, e.g.
    torch data_ptr()); enqueued on `stream` (hipStream_t as an integer), not waited for."""
    _check(lib().pa_stage_sweep_noalloc(ctypes.byref(params), R, IS, beta, pop_ptr, ie_ptr,
                                        m_ptr, stream))


def stage_energy(params, pop, R):
    shape, dt = _pop_shape(params, R)
    src = np.ascontiguousarray(np.array(pop, dtype=dt).reshape(shape))
    IE = np.zeros(_members(params, R), np.int32)
    M = np.zeros_like(IE)
    _check(lib().pa_stage_energy(ctypes.byref(params), R, _ptr(src), _ptr(IE), _ptr(M)))
    return IE, M


def stage_copy_counts(params, IE, R, dB, Emean, IS, active=None):
    """QKer, lnQ, CalcTauKer, CalcParSum: returns (q, lnQ, Roff, prefix, Rnew)."""
    n = _members(params, R)
    ie = np.ascontiguousarray(np.array(IE, dtype=np.int32)[:n])
    act = None if active is None else np.ascontiguousarray(np.array(active, dtype=np.uint8)[:n])
    q, lnq = ctypes.c_double(0), ctypes.c_double(0)
    Roff = np.zeros(n, np.uint32)
    pre = np.zeros(n, np.uint32)
    rn = ctypes.c_int(0)
    _check(lib().pa_stage_copy_counts(ctypes.byref(params), R, _ptr(ie),
                                      None if act is None else _ptr(act), dB, Emean, IS,
                                      ctypes.byref(q), ctypes.byref(lnq), _ptr(Roff), _ptr(pre),
                                      ctypes.byref(rn)))
    return q.value, lnq.value, Roff, pre, rn.value


def stage_resample(params, pop, R, Roff, active=None):
    """resampleKer: returns (children population, MSC active bytes or None, Rnew)."""
    shape, dt = _pop_shape(params, R)
    src = np.ascontiguousarray(np.array(pop, dtype=dt).reshape(shape))
    n = _members(params, R)
    ro = np.ascontiguousarray(np.array(Roff, dtype=np.uint32)[:n])
    act = None if active is None else np.ascontiguousarray(np.array(active, dtype=np.uint8)[:n])
    Rn_max = int(ro.astype(np.int64).sum())
    shape_n, _ = _pop_shape(params, max(Rn_max, 1))
    out = np.zeros(shape_n, dt)
    an = np.zeros(_members(params, max(Rn_max, 1)), np.uint8) if params.coding == MSC else None
    rn = ctypes.c_int(0)
    _check(lib().pa_stage_resample(ctypes.byref(params), R, _ptr(src),
                                   None if act is None else _ptr(act), _ptr(ro), _ptr(out),
                                   None if an is None else _ptr(an), ctypes.byref(rn)))
    shape_r, _ = _pop_shape(params, rn.value)
    out = out[:shape_r[0]]
    if an is not None:
        an = an[:_members(params, rn.value)]
    return out, an, rn.value


def stage_averages(params, IE, M, R, active=None):
    """CalcAverages: returns (5 FP64 sums, sum E, sum M)."""
    n = _members(params, R)
    ie = np.ascontiguousarray(np.array(IE, dtype=np.int32)[:n])
    m = np.ascontiguousarray(np.array(M, dtype=np.int32)[:n])
    act = None if active is None else np.ascontiguousarray(np.array(active, dtype=np.uint8)[:n])
    sums = (ctypes.c_double * 5)()
    se, sm = ctypes.c_int64(0), ctypes.c_int64(0)
    _check(lib().pa_stage_averages(ctypes.byref(params), R, _ptr(ie), _ptr(m),
                                   None if act is None else _ptr(act), sums, ctypes.byref(se),
                                   ctypes.byref(sm)))
    return list(sums), se.value, sm.value


def device_info():
    name = ctypes.create_string_buffer(256)
    cu = ctypes.c_int(0)
    hbm = ctypes.c_size_t(0)
    _check(lib().pa_device_info(name, 256, ctypes.byref(cu), ctypes.byref(hbm)))
    return name.value.decode(), cu.value, hbm.value


def bandwidth_probe(nbytes=1 << 30, iters=10):
    """Achieved GB/s (read + write) of a streaming device copy."""
    g = ctypes.c_double(0)
    _check(lib().pa_bandwidth_probe(nbytes, iters, ctypes.byref(g)))
    return g.value


def _rows(buf, n):
    return [buf[i] for i in range(n)]


def run_all(params, max_rows=10000):
    """One complete run; returns (rows, final Philox subsequence counter)."""
    buf = (Row * max_rows)()
    n = ctypes.c_int(0)
    fin = ctypes.c_uint64(0)
    _check(lib().pa_run_all(ctypes.byref(params), buf, max_rows, ctypes.byref(n), ctypes.byref(fin)))
    return _rows(buf, n.value), fin.value


def run_all_hist(params, max_rows=10000):
    """run_all of a params.mhr = 1 run; returns (rows, final counter, energy histogram)."""
    buf = (Row * max_rows)()
    n = ctypes.c_int(0)
    fin = ctypes.c_uint64(0)
    nb = params.L * params.L + 1
    hist = (ctypes.c_int32 * nb)()
    _check(lib().pa_run_all_hist(ctypes.byref(params), buf, max_rows, ctypes.byref(n),
                                 ctypes.byref(fin), hist, nb))
    return _rows(buf, n.value), fin.value, list(hist)


def mhr_analysis(N, B, BF, R, hist, niter=1):
    """Multi-histogram reweighting (PAisingSSC.cu:695-764).  B, BF (unnormalised), R per row;
    returns (rows of [B, (E/N, C/N, BF/N) per iteration], BF after the last iteration)."""
    nB = len(B)
    Ba = (ctypes.c_double * nB)(*B)
    BFa = (ctypes.c_double * nB)(*BF)
    Ra = (ctypes.c_int32 * nB)(*R)
    Ha = (ctypes.c_int32 * (N + 1))(*hist)
    cols = 1 + 3 * niter
    out = (ctypes.c_double * (nB * cols))()
    _check(lib().pa_mhr_analysis(N, nB, Ba, BFa, Ra, Ha, niter, out))
    return [list(out[i * cols:(i + 1) * cols]) for i in range(nB)], list(BFa)


def unnormalised_bf(rows, N, bf0=None):
    """BF_i = -N ln 2 - sum_{j<=i} lnQ_j (PAisingSSC.cu:435,627-628) in the reference's order;
    bf0 replaces BF_0 (the reference keeps BF[0] across -M runs, where MHR overwrites it)."""
    out = [-N * math.log(2.0) if bf0 is None else bf0]
    s = 0.0
    for r in rows[1:]:
        s -= r.lnQ
        out.append(-N * math.log(2.0) + s)
    return out


class Run:
    """A run kept on the device between calls (stepwise access for tests and timing)."""

    def __init__(self, params):
        self.params = params
        h = ctypes.c_void_p()
        _check(lib().pa_run_create(ctypes.byref(params), ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().pa_run_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n=-1):
        _check(lib().pa_run_step(self.h, n))

    def sync(self):
        _check(lib().pa_run_sync(self.h))

    def steps_done(self):
        """(completed steps, steps of the schedule); syncs.  For an adaptive run the total is
        the table capacity until the schedule has passed Bfin."""
        d = ctypes.c_int(0)
        t = ctypes.c_int(0)
        _check(lib().pa_run_steps_done(self.h, ctypes.byref(d), ctypes.byref(t)))
        return d.value, t.value

    def set_timing(self, on=True):
        _check(lib().pa_run_set_timing(self.h, int(on)))

    def set_graph(self, on=True):
        _check(lib().pa_run_set_graph(self.h, int(on)))

    def timing(self):
        s = ctypes.c_double(0)
        t = ctypes.c_double(0)
        _check(lib().pa_run_timing(self.h, ctypes.byref(s), ctypes.byref(t)))
        return s.value, t.value

    def kernel_times(self):
        """(resample_ms, sweep_ms, steps) summed over the steps timed with set_timing(True)."""
        a = ctypes.c_double(0)
        b = ctypes.c_double(0)
        n = ctypes.c_int(0)
        _check(lib().pa_run_kernel_times(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(n)))
        return a.value, b.value, n.value

    def results(self, max_rows=10000):
        buf = (Row * max_rows)()
        n = ctypes.c_int(0)
        _check(lib().pa_run_results(self.h, buf, max_rows, ctypes.byref(n)))
        return _rows(buf, n.value)

    def final_sequence(self):
        v = ctypes.c_uint64(0)
        _check(lib().pa_run_final_sequence(self.h, ctypes.byref(v)))
        return v.value

    def population(self):
        """SSC: int8 array (R, N); MSC: uint8/16/32/64 array (groups, N)."""
        nb = ctypes.c_size_t(0)
        _check(lib().pa_run_population(self.h, None, ctypes.byref(nb)))
        N = self.params.L * self.params.L
        if self.params.coding == SSC:
            a = np.empty(nb.value, dtype=np.int8)
        else:
            a = np.empty(nb.value // (self.params.msbits // 8),
                         dtype=_pop_shape(self.params, 1)[1])
        _check(lib().pa_run_population(self.h, _ptr(a), ctypes.byref(nb)))
        return a.reshape(-1, N)

    def array(self, which):
        """0 IE, 1 M (int32); 2 Roff, 3 prefix (uint32); 4 MSC active masks (uint64)."""
        nb = ctypes.c_size_t(0)
        _check(lib().pa_run_array(self.h, which, None, ctypes.byref(nb)))
        dt = {0: np.int32, 1: np.int32, 2: np.uint32, 3: np.uint32, 4: np.uint64}[which]
        a = np.empty(nb.value // np.dtype(dt).itemsize, dtype=dt)
        _check(lib().pa_run_array(self.h, which, _ptr(a), ctypes.byref(nb)))
        return a


class Batch:
    """Independent runs advanced together by the same launches (optionally a hipGraph)."""

    def __init__(self, params_list, use_graph=True):
        arr = (Params * len(params_list))(*params_list)
        h = ctypes.c_void_p()
        _check(lib().pa_batch_create(arr, len(params_list), int(use_graph), ctypes.byref(h)))
        self.h = h
        self.n = len(params_list)

    def run(self):
        _check(lib().pa_batch_run(self.h))

    def time_ms(self):
        t = ctypes.c_double(0)
        _check(lib().pa_batch_time(self.h, ctypes.byref(t)))
        return t.value

    def results(self, i, max_rows=10000):
        buf = (Row * max_rows)()
        n = ctypes.c_int(0)
        _check(lib().pa_batch_results(self.h, i, buf, max_rows, ctypes.byref(n)))
        return _rows(buf, n.value)

    def close(self):
        if self.h:
            lib().pa_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching server: submit runs, tick, poll, fetch results."""

    def __init__(self, hbm_budget_bytes, max_runs=256, steps_per_tick=32, use_graph=True,
                 policy=None):
        h = ctypes.c_void_p()
        _check(lib().pa_server_create(hbm_budget_bytes, max_runs, ctypes.byref(h)))
        self.h = h
        _check(lib().pa_server_set_chunk(self.h, steps_per_tick, int(use_graph)))
        if policy is not None:
            _check(lib().pa_server_set_policy(self.h, policy))

    def submit(self, params):
        t = ctypes.c_int64(0)
        _check(lib().pa_server_submit(self.h, ctypes.byref(params), ctypes.byref(t)))
        return t.value

    def tick(self):
        c = ctypes.c_int(0)
        _check(lib().pa_server_tick(self.h, ctypes.byref(c)))
        return c.value

    def done(self, ticket):
        d = ctypes.c_int(0)
        _check(lib().pa_server_poll(self.h, ticket, ctypes.byref(d)))
        return bool(d.value)

    def results(self, ticket, max_rows=10000):
        buf = (Row * max_rows)()
        n = ctypes.c_int(0)
        _check(lib().pa_server_results(self.h, ticket, buf, max_rows, ctypes.byref(n)))
        return _rows(buf, n.value)

    def latency(self, ticket):
        a = ctypes.c_double(0)
        b = ctypes.c_double(0)
        _check(lib().pa_server_latency(self.h, ticket, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def reserve(self, nbytes):
        """Grow the device pool of run arenas to `nbytes` (at most the budget) up front."""
        _check(lib().pa_server_reserve(self.h, nbytes))

    def memory(self):
        """(device bytes held by the server's pools, device allocations made so far)."""
        r = ctypes.c_size_t(0)
        n = ctypes.c_int(0)
        _check(lib().pa_server_memory(self.h, ctypes.byref(r), ctypes.byref(n)))
        return r.value, n.value

    def stats(self):
        q = ctypes.c_int(0)
        r = ctypes.c_int(0)
        u = ctypes.c_size_t(0)
        _check(lib().pa_server_stats(self.h, ctypes.byref(q), ctypes.byref(r), ctypes.byref(u)))
        return q.value, r.value, u.value

    def close(self):
        if self.h:
            lib().pa_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def row_fields(row):
    return [row.B, row.E, row.C, row.M, row.M2, row.M4, row.BF, row.S, row.R, row.lnQ]


def format_rows(rows, precision=11):
    """Results-file text as the reference writes it (ostream precision P, SSC:679-690)."""
    out = []
    for r in rows:
        f = row_fields(r)
        out.append(" ".join(["%.*g" % (precision, v) for v in f[:8]] + ["%d" % f[8],
                                                                        "%.*g" % (precision, f[9])]) + "\n")
    return "".join(out)
