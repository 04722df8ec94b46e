"""End-to-end runs of the drop-in driver on the device.

The `gpu_DSM` executable reads input.dat like the reference CLI (CLI/main.cu:20-56,
common/main_cuda.cu:34-185) and writes tau/G[_ID].dat, the -s binary and the -distr
histograms.  Its printed numbers are checked against the oracle run on the same chains
(mirror-FP32 samples, FP64 ensemble sums, printed with 6 significant digits), and the
-s file against the oracle's chain state bit for bit.
"""
import struct
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _write_input(path, beta, nk, n, kappa, cd, eq_type, dt, T, pd=0):
    lines = [repr(beta), str(nk), str(n), " ".join(repr(float(k)) for k in kappa), str(cd), str(pd), str(eq_type),
             repr(dt), repr(T)]
    path.write_text("\n".join(lines) + "\n")


def _run(cwd, *args):
    import gpu_dsm_amd as g
    r = subprocess.run([g.EXE_PATH, *args], cwd=str(cwd), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
    return r.stdout


def _read_save(path):
    """-s file (common/ensemble.cu:342-354): int32 N_cha, then per chain a 16-byte head
    (Z, time, unused, stall), Z float4 records (Qx, Qy, Qz, N), Z floats (link rates)."""
    b = path.read_bytes()
    n = struct.unpack_from("<i", b, 0)[0]
    off, out = 4, []
    for _ in range(n):
        Z, t, _, stall = struct.unpack_from("<iffi", b, off)
        off += 16
        qn = np.frombuffer(b, np.float32, 4 * Z, off).reshape(Z, 4).copy()
        off += 16 * Z
        w = np.frombuffer(b, np.float32, Z, off).copy()
        off += 4 * Z
        out.append((Z, t, stall, qn, w))
    assert off == len(b)
    return out


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_cli_flow_run_matches_oracle(tmp_path, oracle):
    beta, nk, n, T, dt = 1.0, 20, 101, 10.0, 1.0
    kappa = [0, 0.05, 0, 0, 0, 0, 0, 0, 0]
    _write_input(tmp_path / "input.dat", beta, nk, n, kappa, 1, 0, dt, T)
    out = _run(tmp_path, "0", "-s", "save.bin", "-distr", "--quiet")
    assert "Calculation time" in out
    rows = np.loadtxt(tmp_path / "tau.dat")
    assert rows.shape == (10, 9)
    np.testing.assert_allclose(rows[:, 0], np.arange(1, 11) * dt)
    P = oracle.Params(beta, nk, 1, kappa=kappa)
    O = oracle.Ensemble(P, seed=0, n=n, salt=0)
    _, s = O.advance(t_end=T, delta=dt, k0=1, k_end=11, samples=True, flow_samples=True)
    ref = s.astype(np.float64).sum(axis=1) / n
    np.testing.assert_allclose(rows[:, 1:], ref, rtol=2e-5, atol=2e-6)
    st = O.state()
    saved = _read_save(tmp_path / "save.bin")
    assert len(saved) == n
    for i, (Z, _, stall, qn, w) in enumerate(saved):
        assert Z == st["Z"][i] and stall == 0
        assert np.array_equal(qn[:, 3].astype(np.int32), st["N"][i, :Z])
        assert np.array_equal(_bits(qn[:, :3]), _bits(st["Q"][i, :Z]))
        assert np.array_equal(_bits(w[:Z - 1]), _bits(st["w"][i, :Z - 1]))
    for name in ("distr_Z.dat", "distr_N.dat", "distr_Q.dat"):
        d = np.loadtxt(tmp_path / name)
        assert d.ndim == 2 and d.shape[0] > 1 and np.all(np.isfinite(d))


@pytest.mark.parametrize("eq_type,name", [(1, "G"), (2, "MSD")])
def test_cli_gt_run_matches_oracle(tmp_path, oracle, eq_type, name):
    """equilibrium_type 1 writes G_<job>.dat, 2 writes MSD_<job>.dat (common/main_cuda.cu:142-155,
    common/ensemble.cu:324-335)."""
    beta, nk, n, T, dt, job = 1.0, 20, 128, 100.0, 1.0, 7
    p, m = 16, 2
    _write_input(tmp_path / "input.dat", beta, nk, n, [0] * 9, 1, eq_type, dt, T)
    out = _run(tmp_path, str(job), "--corr-p", str(p), "--corr-m", str(m), "--quiet")
    assert ("%s(t) calc..." % name) in out
    g = np.loadtxt(tmp_path / ("%s_%d.dat" % (name, job)))
    s = int(np.ceil(np.log(T / dt / p) / np.log(m))) + 1
    P = oracle.Params(beta, nk, 1)
    O = oracle.Ensemble(P, seed=job, n=n, salt=0)
    _, samples = O.advance(t_end=T, delta=dt, k0=0, k_end=int(T), samples=True, msd=eq_type == 2)
    corr = oracle.Correlator(p, m, s, n, type=eq_type - 1)
    for row in samples:
        corr.add_row(row)
    lag_o, g_o = corr.result()
    k = g.shape[0]
    assert 0 < k <= len(lag_o)
    np.testing.assert_allclose(g[:, 0], lag_o[:k])
    np.testing.assert_allclose(g[:, 1], g_o[:k], rtol=2e-5, atol=1e-6 * abs(g_o[0 if eq_type == 1 else 1]))


def test_cli_pd_flag_gex_background_matches_oracle(tmp_path, oracle):
    """PD_flag = 1 reads a, b, mp, Mk from polydisp.dat (common/main_cuda.cu:95-103) and
    draws CD partners from the GEX background; G(t) and the -s state match the oracle."""
    beta, nk, n, T, dt, job = 1.0, 30, 96, 60.0, 1.0, 3
    gex = (1.5, 1.0, 40000.0, 120.0)
    p, m = 16, 2
    _write_input(tmp_path / "input.dat", beta, nk, n, [0] * 9, 1, 1, dt, T, pd=1)
    (tmp_path / "polydisp.dat").write_text(" ".join(repr(x) for x in gex) + "\n")
    _run(tmp_path, str(job), "--corr-p", str(p), "--corr-m", str(m), "-s", "save.bin", "--quiet")
    g = np.loadtxt(tmp_path / ("G_%d.dat" % job))
    s = int(np.ceil(np.log(T / dt / p) / np.log(m))) + 1
    P = oracle.Params(beta, nk, 1, partner="gex", gex=gex)
    O = oracle.Ensemble(P, seed=job, n=n, salt=0)
    _, samples = O.advance(t_end=T, delta=dt, k0=0, k_end=int(T), samples=True)
    corr = oracle.Correlator(p, m, s, n)
    for row in samples:
        corr.add_row(row)
    lag_o, g_o = corr.result()
    k = g.shape[0]
    np.testing.assert_allclose(g[:, 1], g_o[:k], rtol=2e-5, atol=1e-6 * abs(g_o[0]))
    st = O.state()
    for i, (Z, _, stall, qn, w) in enumerate(_read_save(tmp_path / "save.bin")):
        assert Z == st["Z"][i]
        assert np.array_equal(_bits(w[:Z - 1]), _bits(st["w"][i, :Z - 1]))


def test_cli_pd_flag_without_polydisp_file_fails(tmp_path):
    _write_input(tmp_path / "input.dat", 1.0, 20, 16, [0] * 9, 1, 1, 1.0, 20.0, pd=1)
    import gpu_dsm_amd as g
    r = subprocess.run([g.EXE_PATH, "0", "--quiet"], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "polydisp.dat" in r.stdout


def test_cli_load_rejects_size_mismatch(tmp_path):
    _write_input(tmp_path / "input.dat", 1.0, 20, 16, [0] * 9, 1, 1, 1.0, 20.0)
    _run(tmp_path, "0", "-s", "s.bin", "--quiet")
    _write_input(tmp_path / "input.dat", 1.0, 20, 17, [0] * 9, 1, 1, 1.0, 20.0)
    import gpu_dsm_amd as g
    r = subprocess.run([g.EXE_PATH, "0", "-l", "s.bin", "--quiet"], cwd=str(tmp_path), capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "ensemble size mismatch" in r.stdout
