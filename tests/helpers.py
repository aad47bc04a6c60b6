"""Shared helpers: run the same configuration through the oracle and the engine and compare
per-step records exactly (integers, hashes and the bit patterns of the FP64 values)."""
import oracle
import paising_amd as pa

# engine uniform convention -> oracle mapping id
UMAP = {pa.UNIFORM_CUDA101: oracle.U32_OPEN_CLOSED, pa.UNIFORM_CUDA100: oracle.U32_MIDPOINT}


def oracle_rows(L, coding, msbits, Rinit, eqsweeps, dB, Bfin, seed, steps, uniform=0,
                initial_sequence=0, threads=0, **kw):
    rc, runs, _, _ = oracle.run(L=L, coding=coding, msbits=msbits, Rinit=Rinit,
                                eqsweeps=eqsweeps, dB=dB, Bfin=Bfin, seed=seed,
                                initial_sequence=initial_sequence, umap=UMAP[uniform],
                                max_steps=steps, threads=threads, **kw)
    assert rc == 0
    return runs[0]


def engine_params(L, coding, msbits, Rinit, eqsweeps, dB, Bfin, seed, steps, uniform=0,
                  initial_sequence=0, trace=1, **kw):
    return pa.make_params(L=L, coding=coding, msbits=msbits, Rinit=Rinit, eqsweeps=eqsweeps,
                          dB=dB, Bfin=Bfin, seed=seed, initial_sequence=initial_sequence,
                          uniform=uniform, max_steps=steps, trace=trace, **kw)


def compare_rows(orows, erows, check_hash=True):
    """Assert bit-exact agreement of every step record; returns the number of steps."""
    assert len(orows) == len(erows), (len(orows), len(erows))
    for i, (o, e) in enumerate(zip(orows, erows)):
        ctx = f"step {i}"
        assert o.R == e.R, (ctx, "R", o.R, e.R)
        assert o.B == e.B, (ctx, "B")
        if i == 0:
            continue
        assert o.Rprev == e.Rprev, (ctx, "Rprev")
        assert o.q.hex() == e.q.hex(), (ctx, "q", o.q, e.q)
        assert o.lnQ.hex() == e.lnQ.hex(), (ctx, "lnQ", o.lnQ, e.lnQ)
        if check_hash:
            assert o.roff_hash == e.roff_hash, (ctx, "roff_hash")
        assert o.sumE_int == e.sumE, (ctx, "sumE", o.sumE_int, e.sumE)
        assert o.sumM_int == e.sumM, (ctx, "sumM", o.sumM_int, e.sumM)
        for k in range(5):
            assert o.sums[k].hex() == e.sums[k].hex(), (ctx, "sums", k, o.sums[k], e.sums[k])
        if check_hash:
            assert o.spin_hash == e.spin_hash, (ctx, "spin_hash")
        assert o.IS_after == e.IS_after, (ctx, "IS", o.IS_after, e.IS_after)
        for f in ("E", "C", "M", "M2", "M4", "BF", "S"):
            assert getattr(o, f).hex() == getattr(e, f).hex(), (ctx, f, getattr(o, f), getattr(e, f))
    return len(orows)


def load_golden(name):
    """A committed oracle trace file tests/golden/oracle_<name>.json."""
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", f"oracle_{name}.json")
    with open(path) as f:
        return json.load(f)


def params_of(d, **over):
    """Engine parameters of a trace's parameter record."""
    kw = dict(d)
    kw.update(over)
    return pa.make_params(**kw)


def compare_trace(trace, erows, check_hash=True):
    """Assert that the engine's rows agree bit for bit with the first len(trace) records of
    an oracle trace (R, q, lnQ and FP64 sums as hex, integer sums, hashes, Philox counter)."""
    assert len(erows) >= len(trace), (len(erows), len(trace))
    for t, e in zip(trace, erows):
        ctx = f"step {t['i']}"
        assert t["R"] == e.R, (ctx, "R", t["R"], e.R)
        if t["i"] == 0:
            continue
        assert t["Rprev"] == e.Rprev, (ctx, "Rprev")
        assert t["q"] == e.q.hex(), (ctx, "q", t["q"], e.q.hex())
        assert t["lnQ"] == e.lnQ.hex(), (ctx, "lnQ")
        assert t["sums"] == [x.hex() for x in e.sums], (ctx, "sums")
        assert (t["sumE"], t["sumM"]) == (e.sumE, e.sumM), (ctx, "integer sums")
        if check_hash:
            assert t["roff_hash"] == e.roff_hash, (ctx, "roff_hash")
            assert t["spin_hash"] == e.spin_hash, (ctx, "spin_hash")
        assert t["IS"] == e.IS_after, (ctx, "IS")
    return len(trace)


def same_rows(a, b):
    """Bit-exact equality of two engine row lists (batched or served against plain)."""
    assert len(a) == len(b), (len(a), len(b))
    for i, (x, y) in enumerate(zip(a, b)):
        for f in ("R", "Rprev", "sumE", "sumM", "roff_hash", "spin_hash", "IS_after"):
            assert getattr(x, f) == getattr(y, f), (i, f)
        for f in ("q", "lnQ", "E", "C", "M", "M2", "M4", "BF", "S"):
            assert getattr(x, f).hex() == getattr(y, f).hex(), (i, f)
