"""Population dumps of the drop-in executables (EnergiesPopStore / LatticePopStore builds,
PAisingSSC.cu:24-25,563-605; PAisingMSC.cu:743-757): the dumped lattices reproduce the
oracle's per-step spin hashes and sums, and the dumped energies are the lattices' energies."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
import paising_amd as pa

pytestmark = pytest.mark.gpu
M64 = (1 << 64) - 1


def mix64(z):
    z = (z + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def spin_hash(reps, N):
    """sum over replicas and 8-byte chunks of mix64(mix64(index) ^ chunk), int8 gA then gB."""
    chunks = (N + 7) // 8
    h = 0
    for j, b in enumerate(reps):
        raw = b.astype(np.int8).tobytes()
        for c in range(chunks):
            v = int.from_bytes(raw[8 * c:8 * c + 8].ljust(8, b"\0"), "little")
            h = (h + mix64(mix64(j * chunks + c) ^ v)) & M64
    return h


def parse_lattices(text, L):
    """PA_lattice_i.dat -> (i, beta, R, [int8 gA||gB per replica], [2-D lattice per replica])."""
    head, *body = [l for l in text.splitlines() if l]
    m = re.match(r"i=(\d+); beta=([^;]+); R=(\d+);", head)
    Ld2 = L // 2
    reps, grids = [], []
    for line in body:
        rows = re.findall(r"\{([-0-9,]+)\}", line[len("P={"):])
        grid = np.array([[int(v) for v in r.split(",")] for r in rows])
        assert grid.shape == (L, L)
        gA = np.zeros(L * Ld2, dtype=np.int8)
        gB = np.zeros(L * Ld2, dtype=np.int8)
        for ty in range(L):   # A at x = 2tx + (ty & 1), B at 2tx + 1 - (ty & 1)
            for tx in range(Ld2):
                gA[ty * Ld2 + tx] = grid[ty, 2 * tx + (ty & 1)]
                gB[ty * Ld2 + tx] = grid[ty, 2 * tx + 1 - (ty & 1)]
        reps.append(np.concatenate([gA, gB]))
        grids.append(grid)
    return int(m.group(1)), float(m.group(2)), int(m.group(3)), reps, grids


def energy(grid):
    return int(-(grid * np.roll(grid, 1, 0)).sum() - (grid * np.roll(grid, 1, 1)).sum())


def test_ssc_energy_and_lattice_dumps(tmp_path):
    L, R = 8, 60
    r = subprocess.run([os.path.join(pa.BIN_DIR, "PAisingSSC"), "-e", "-l", "-R", str(R), "-t", "3",
                        "-d", "0.25", "-f", "1", "-s", "5", "-M", "2", "-L", str(L)],
                       capture_output=True, text=True, cwd=tmp_path, timeout=300)
    assert r.returncode == 0, r.stderr
    d = tmp_path / f"dataSSC_L{L}_R{R}_EqSw3_dB0.250000"
    rc, runs, _, _ = oracle.run(L=L, coding=oracle.SSC, Rinit=R, eqsweeps=3, dB=0.25, Bfin=1.0,
                                seed=5, runs=2)
    assert rc == 0
    for i in range(1, len(runs[0])):
        o = runs[0][i]
        k, beta, Ri, reps, grids = parse_lattices((d / f"PA_lattice_{i}.dat").read_text(), L)
        assert (k, Ri) == (i, o.R) and beta == pytest.approx(o.B, rel=1e-10)
        assert spin_hash(reps, L * L) == o.spin_hash
        assert sum(energy(g) for g in grids) == o.sumE_int
        assert sum(int(g.sum()) for g in grids) == o.sumM_int
        # energies files hold the last run's populations (run 1 overwrites run 0)
        e1 = [int(x) for x in (d / f"PA_energies_{i}.dat").read_text().split()]
        assert len(e1) == runs[1][i].R and sum(e1) == runs[1][i].sumE_int
    assert not (d / f"PA_lattice_{len(runs[0])}.dat").exists()


def test_msc_energy_dumps(tmp_path):
    r = subprocess.run([os.path.join(pa.BIN_DIR, "PAisingMSC"), "-e", "-R", "300", "-t", "2", "-d",
                        "0.25", "-f", "1", "-s", "9", "-L", "8"], capture_output=True, text=True,
                       cwd=tmp_path, timeout=300)
    assert r.returncode == 0, r.stderr
    d = tmp_path / "dataMSC_L8_R300_EqSw2_dB0.250000"
    rc, runs, _, _ = oracle.run(L=8, coding=oracle.MSC, Rinit=300, eqsweeps=2, dB=0.25, Bfin=1.0,
                                seed=9)
    assert rc == 0
    for i in range(1, len(runs[0])):
        e = [int(x) for x in (d / f"PA_energies_{i}.dat").read_text().split()]
        assert len(e) == runs[0][i].R and sum(e) == runs[0][i].sumE_int
        assert all((x + 128) % 4 == 0 and -128 <= x <= 128 for x in e)
