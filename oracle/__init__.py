"""ctypes binding of the CPU oracle (test infrastructure only).

The oracle is a host C++/OpenMP restatement of PAisingSSC.cu / PAisingMSC.cu.  Only the
tests, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use it, as
the checker; the product package ``paising_amd`` never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

SSC, MSC = 0, 1
(U_ROCRAND_UNIFORM, U_XOR21_P1, U_XOR21_HALF, U_XOR21_SWAP, U_U64_XHI, U_U64_YHI,
 U32_OPEN_CLOSED, U32_MIDPOINT, U32_CLOSED_OPEN) = range(9)
# The reference fixtures select one 32-bit word per uniform double: (v+1)/2^32 reproduces
# the CUDA >= 10.1 SSC fixture, (v+1/2)/2^32 the CUDA <= 10.0 one; MSC matches either.
U_DEFAULT = U32_OPEN_CLOSED
U_LEGACY = U32_MIDPOINT


class Params(ctypes.Structure):
    _fields_ = [
        ("L", ctypes.c_int32), ("coding", ctypes.c_int32), ("msbits", ctypes.c_int32),
        ("Rinit", ctypes.c_int32), ("eqsweeps", ctypes.c_uint32), ("dB", ctypes.c_double),
        ("Binit", ctypes.c_double), ("Bfin", ctypes.c_double), ("runs", ctypes.c_int32),
        ("seed", ctypes.c_uint64), ("initial_sequence", ctypes.c_uint64),
        ("umap", ctypes.c_int32), ("max_steps", ctypes.c_int32), ("threads", ctypes.c_int32),
        ("adaptive", ctypes.c_int32), ("pad0", ctypes.c_int32), ("min_overlap", ctypes.c_double),
        ("max_overlap", ctypes.c_double), ("mhr", ctypes.c_int32), ("mhr_iter", ctypes.c_int32),
    ]


class Row(ctypes.Structure):
    _fields_ = [
        ("B", ctypes.c_double), ("E", ctypes.c_double), ("C", ctypes.c_double),
        ("M", ctypes.c_double), ("M2", ctypes.c_double), ("M4", ctypes.c_double),
        ("BF", ctypes.c_double), ("S", ctypes.c_double), ("R", ctypes.c_int32),
        ("Rprev", ctypes.c_int32), ("lnQ", ctypes.c_double), ("q", ctypes.c_double),
        ("sums", ctypes.c_double * 5), ("sumE_int", ctypes.c_int64),
        ("sumM_int", ctypes.c_int64), ("roff_hash", ctypes.c_uint64),
        ("spin_hash", ctypes.c_uint64), ("IS_after", ctypes.c_uint64),
    ]


_lib = None


def build(force=False):
    """Compile liboracle.so and the CLI with the committed Makefile."""
    if force or not os.path.exists(LIB):
        subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        P, D, I, U64, U32 = (ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_uint64,
                             ctypes.c_uint32)
        sig = {
            "ora_run": (I, [ctypes.POINTER(Params), P, I, P, P, P]),
            "ora_run_ex": (I, [ctypes.POINTER(Params), P, I, P, P, P, P, P]),
            "ora_mhr_analysis": (None, [I, I, P, P, P, P, I, P]),
            "ora_philox": (None, [P, P, P]),
            "ora_stream_word": (U32, [U64, U64, U64]),
            "ora_uniform": (D, [U32, U32, I]),
            "ora_exp": (D, [D]),
            "ora_log": (D, [D]),
            "ora_block_reduce": (D, [P, I]),
            "ora_small_block_reduce": (D, [P, I]),
            "ora_ssc_init": (None, [I, I, U64, U64, I, P]),
            "ora_ssc_sweep": (None, [I, I, P, U32, U64, U64, D, I]),
            "ora_ssc_energy": (None, [I, I, P, P, P]),
            "ora_msc_init": (None, [I, I, I, I, U64, U64, P, P]),
            "ora_msc_sweep": (None, [I, I, I, P, U32, U64, U64, D]),
            "ora_msc_energy": (None, [I, I, I, P, P, P]),
            "ora_ssc_tau": (ctypes.c_int64, [I, P, I, D, D, U64, U64, I, P]),
            "ora_msc_tau": (ctypes.c_int64, [I, I, P, P, I, I, D, D, U64, U64, I, P]),
            "ora_ssc_q": (D, [I, P, D, D]),
            "ora_msc_q": (D, [I, I, P, P, D, D]),
            "ora_lnq": (D, [D, D, D, I]),
            "ora_ssc_overlap": (D, [I, P, I, D, D]),
            "ora_msc_overlap": (D, [I, I, P, P, I, I, D, D]),
            "ora_ssc_averages": (None, [I, P, P, P]),
            "ora_msc_averages": (None, [I, I, P, P, P, P]),
            "ora_ssc_resample": (None, [I, I, P, P, P]),
            "ora_msc_resample": (None, [I, I, I, P, P, P, P, I, P, P, P]),
            "ora_ssc_table": (None, [D, P]),
            "ora_msc_table": (None, [D, P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def run(L=64, coding=SSC, msbits=32, Rinit=20000, eqsweeps=100, dB=0.005, Binit=0.0,
        Bfin=1.0, runs=1, seed=100, initial_sequence=0, umap=U_DEFAULT, max_steps=0,
        threads=0, max_rows=10000, adaptive=0, min_overlap=0.0, max_overlap=0.0):
    """Run the oracle; returns (rc, list of per-run row lists, elapsed_s, total_pop)."""
    rc, out, el, tot, _, _ = run_mhr(L, coding, msbits, Rinit, eqsweeps, dB, Binit, Bfin, runs,
                                     seed, initial_sequence, umap, max_steps, threads, max_rows,
                                     adaptive, min_overlap, max_overlap, mhr=0)
    return rc, out, el, tot


def run_mhr(L=64, coding=SSC, msbits=32, Rinit=20000, eqsweeps=100, dB=0.005, Binit=0.0,
            Bfin=1.0, runs=1, seed=100, initial_sequence=0, umap=U_DEFAULT, max_steps=0,
            threads=0, max_rows=10000, adaptive=0, min_overlap=0.0, max_overlap=0.0, mhr=1,
            mhr_iter=1):
    """run() of the MHR build; also returns per-run energy histograms (N+1 ints) and
    PA_MHR_results rows ([B, E/N, C/N, BF/N per iteration])."""
    p = Params(L, coding, msbits, Rinit, eqsweeps, dB, Binit, Bfin, runs,
               seed & (2**64 - 1), initial_sequence, umap, max_steps, threads, adaptive, 0,
               min_overlap, max_overlap, mhr, mhr_iter)
    N = L * L
    cols = 1 + 3 * max(1, mhr_iter)
    hist = np.zeros(runs * (N + 1), dtype=np.int32) if mhr else None
    mrows = np.zeros(runs * max_rows * cols, dtype=np.float64) if mhr else None
    rows = (Row * (max_rows * runs))()
    nrows = np.zeros(runs, dtype=np.int32)
    el = ctypes.c_double(0)
    tot = ctypes.c_double(0)
    rc = lib().ora_run_ex(ctypes.byref(p), rows, max_rows, _ptr(nrows), ctypes.byref(el),
                          ctypes.byref(tot), _ptr(hist) if mhr else None,
                          _ptr(mrows) if mhr else None)
    out = [[rows[r * max_rows + i] for i in range(int(nrows[r]))] for r in range(runs)]
    hists = [hist[r * (N + 1):(r + 1) * (N + 1)].tolist() for r in range(runs)] if mhr else None
    mhr_rows = ([[mrows[(r * max_rows + i) * cols:(r * max_rows + i + 1) * cols].tolist()
                  for i in range(int(nrows[r]))] for r in range(runs)] if mhr else None)
    return rc, out, el.value, tot.value, hists, mhr_rows


def row_fields(row):
    return [row.B, row.E, row.C, row.M, row.M2, row.M4, row.BF, row.S, row.R, row.lnQ]


def format_mhr_rows(rows, precision=11):
    """PA_MHR_results_run_r.dat lines (SSC:776-786): every value followed by a space."""
    return "".join("".join("%.*g " % (precision, v) for v in r) + "\n" for r in rows)


def format_rows(rows, precision=11):
    """Rows as the reference writes them: ostream precision P (= %.Pg), SSC:679-690."""
    lines = []
    for r in rows:
        f = row_fields(r)
        parts = ["%.*g" % (precision, v) for v in f[:8]] + ["%d" % f[8], "%.*g" % (precision, f[9])]
        lines.append(" ".join(parts) + "\n")
    return "".join(lines)
