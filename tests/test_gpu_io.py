"""Files and state access of the drop-in surface, on the device.

  * -s / -l binary files (common/ensemble.cu:342-354,486-504, common/chain.cu:236-250):
    engine save/load round trip; a CLI run saved with -s and continued with -l matches
    the oracle continued from the same file; a file written by hand in the reference's
    layout (unused head float and last rate slot holding garbage, nonzero clocks) loads
    and runs like the oracle from that state; a truncated file is an error, not a crash.
  * -distr histograms (common/ensemble.cu:356-484) against the saved state.
  * dsm_stress (common/ensemble.cu:210-221 over common/ensemble_kernel.cu:504-538)
    against an FP64 evaluation of the state on the host.
"""
import struct
import subprocess

import numpy as np
import pytest

from test_gpu_end_to_end import _bits, _read_save, _run, _write_input

pytestmark = pytest.mark.gpu


def _write_reference_file(path, Z, N, Q, w, t, junk=7.5, last_rate=1.0e30):
    """The reference's -s layout: int32 N_cha; per chain {int32 Z, float time, float
    unused, int32 stall}, Z x float4 {Qx, Qy, Qz, N}, Z x float rate (the last unused)."""
    out = [struct.pack("<i", len(Z))]
    for i in range(len(Z)):
        z = int(Z[i])
        out.append(struct.pack("<iffi", z, float(t[i]), junk, 0))
        rec = np.zeros((z, 4), np.float32)
        rec[:, :3] = Q[i, :z]
        rec[:, 3] = N[i, :z]
        out.append(rec.tobytes())
        r = np.array(w[i, :z], np.float32)
        r[z - 1] = last_rate
        out.append(r.tobytes())
    path.write_bytes(b"".join(out))


def _oracle_from(oracle, P, seed, n, Z, N, Q, w, t, next_sample):
    O = oracle.Ensemble(P, seed=seed, n=n, salt=0)
    O.set_state(Z, N, Q, w, t=t, ev=np.zeros(n, np.uint64), dt_prev=np.zeros(n, np.float32),
                next_sample=np.full(n, next_sample, np.int64))
    return O


def test_engine_save_load_roundtrip(engine, tmp_path):
    n, nk = 300, 60
    a = engine.add_ensemble(n, nk, 1.0, 1, seed=4)
    engine.init_equilibrium(a)
    engine.advance(a, 7.0)
    engine.save(a, tmp_path / "s.bin", t_ref=5.0)
    b = engine.add_ensemble(n, nk, 1.0, 1, seed=4)
    engine.load(b, tmp_path / "s.bin")
    sa, sb = engine.state(a), engine.state(b)
    assert np.array_equal(sa["Z"], sb["Z"])
    on = np.arange(sa["N"].shape[1])[None, :] < sa["Z"][:, None]
    assert np.array_equal(np.where(on, sa["N"], 0), sb["N"])
    assert np.array_equal(_bits(np.where(on[..., None], sa["Q"], 0)), _bits(sb["Q"]))
    assert np.array_equal(_bits(sa["w"]), _bits(sb["w"]))
    assert np.array_equal(sb["t"], (sa["t"] - 5.0).astype(np.float32).astype(np.float64))
    assert np.all(sb["ev"] == 0)
    c = engine.add_ensemble(n + 1, nk, 1.0, 1, seed=4)
    with pytest.raises(Exception, match="mismatch"):
        engine.load(c, tmp_path / "s.bin")
    for e in (a, b, c):
        engine.remove_ensemble(e)


def test_cli_save_then_load_and_continue(tmp_path, oracle):
    """-s after a flow run, then a second run with -l continues from that file: the second
    run's tau rows and -s file equal the oracle continued from the same loaded state
    (clocks relative to the first run's end, jump counters restarting at 0)."""
    beta, nk, n, dt, job = 1.0, 24, 97, 1.0, 5
    kappa = [0, 0.04, 0, 0, 0, 0, 0, 0, 0]
    _write_input(tmp_path / "input.dat", beta, nk, n, kappa, 1, 0, dt, 6.0)
    _run(tmp_path, str(job), "-s", "first.bin", "--quiet")
    T2 = 8.0
    _write_input(tmp_path / "input.dat", beta, nk, n, kappa, 1, 0, dt, T2)
    out = _run(tmp_path, str(job), "-l", "first.bin", "-s", "second.bin", "--quiet")
    assert "loading chain conformations from first.bin" in out
    first = _read_save(tmp_path / "first.bin")
    cap = nk
    Z = np.array([c[0] for c in first], np.int32)
    N = np.zeros((n, cap), np.int32); Q = np.zeros((n, cap, 3), np.float32); w = np.zeros((n, cap), np.float32)
    t = np.array([c[1] for c in first], np.float64)
    for i, (z, _, _, qn, ww) in enumerate(first):
        N[i, :z] = qn[:, 3].astype(np.int32); Q[i, :z] = qn[:, :3]; w[i, :z - 1] = ww[:z - 1]
    P = oracle.Params(beta, nk, 1, kappa=kappa)
    O = _oracle_from(oracle, P, job, n, Z, N, Q, w, t, 1)
    k_end = int(T2 / dt) + 1
    _, s = O.advance(t_end=T2, delta=dt, k0=1, k_end=k_end, samples=True, flow_samples=True)
    rows = np.loadtxt(tmp_path / ("tau_%d.dat" % job))
    assert rows.shape == (k_end - 1, 9)
    np.testing.assert_allclose(rows[:, 1:], s.astype(np.float64).sum(axis=1) / n, rtol=2e-5, atol=2e-6)
    st = O.state()
    for i, (z, _, stall, qn, ww) in enumerate(_read_save(tmp_path / "second.bin")):
        assert z == st["Z"][i] and stall == 0
        assert np.array_equal(qn[:, 3].astype(np.int32), st["N"][i, :z])
        assert np.array_equal(_bits(qn[:, :3]), _bits(st["Q"][i, :z]))


def test_cli_loads_reference_layout(tmp_path, oracle):
    """A file in the reference's layout with garbage in the fields it leaves unset, and
    chain clocks that are not zero, loads and runs G(t) like the oracle from that state."""
    beta, nk, n, T, dt, job = 1.0, 30, 64, 40.0, 1.0, 9
    p, m = 16, 2
    P = oracle.Params(beta, nk, 1)
    init = oracle.Ensemble(P, seed=123, n=n).state()
    rng = np.random.default_rng(2)
    t0 = rng.uniform(0.0, 0.7, n).astype(np.float32).astype(np.float64)
    _write_reference_file(tmp_path / "ref.bin", init["Z"], init["N"], init["Q"], init["w"], t0)
    _write_input(tmp_path / "input.dat", beta, nk, n, [0] * 9, 1, 1, dt, T)
    _run(tmp_path, str(job), "-l", "ref.bin", "--corr-p", str(p), "--corr-m", str(m), "--quiet")
    g = np.loadtxt(tmp_path / ("G_%d.dat" % job))
    w = init["w"].copy()
    w[np.arange(n), init["Z"] - 1] = 0.0
    O = _oracle_from(oracle, P, job, n, init["Z"], init["N"], init["Q"], w, t0, 0)
    _, samples = O.advance(t_end=T, delta=dt, k0=0, k_end=int(T), samples=True)
    s = int(np.ceil(np.log(T / dt / p) / np.log(m))) + 1
    corr = oracle.Correlator(p, m, s, n)
    for row in samples:
        corr.add_row(row)
    lag_o, g_o = corr.result()
    k = g.shape[0]
    np.testing.assert_allclose(g[:, 0], lag_o[:k])
    np.testing.assert_allclose(g[:, 1], g_o[:k], rtol=2e-5, atol=1e-6 * abs(g_o[0]))


def test_truncated_load_file_is_an_error(engine, tmp_path):
    n, nk = 50, 20
    a = engine.add_ensemble(n, nk, 1.0, 1, seed=4)
    engine.init_equilibrium(a)
    engine.save(a, tmp_path / "s.bin")
    b = tmp_path / "s.bin"
    full = b.read_bytes()
    for cut in (2, 4 + 10, len(full) // 2, len(full) - 1):
        (tmp_path / "cut.bin").write_bytes(full[:cut])
        with pytest.raises(Exception, match="-5"):
            engine.load(a, tmp_path / "cut.bin")
    with pytest.raises(Exception, match="-5"):
        engine.load(a, tmp_path / "missing.bin")
    # a record whose N do not sum to Nk is rejected as an invalid file
    bad = bytearray(full)
    struct.pack_into("<f", bad, 4 + 16 + 12, 1000.0)  # N of the first strand of chain 0
    (tmp_path / "bad.bin").write_bytes(bytes(bad))
    with pytest.raises(Exception, match="-5"):
        engine.load(a, tmp_path / "bad.bin")
    engine.remove_ensemble(a)
    _write_input(tmp_path / "input.dat", 1.0, nk, n, [0] * 9, 1, 1, 1.0, 10.0)
    import gpu_dsm_amd as g
    (tmp_path / "cut.bin").write_bytes(full[:len(full) // 2])
    r = subprocess.run([g.EXE_PATH, "0", "-l", "cut.bin", "--quiet"], cwd=str(tmp_path), capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0 and "file error" in r.stdout


def test_distr_files_match_state(tmp_path):
    """distr_Z (pdf of Z), distr_N (pdf of N over all strands) and distr_Q (CDF of |Q|
    over internal strands, every 10th point, ties sharing the lower probability) computed
    from the saved state by the reference's rules (common/ensemble.cu:356-484)."""
    beta, nk, n = 1.0, 40, 400
    _write_input(tmp_path / "input.dat", beta, nk, n, [0, 0.02, 0, 0, 0, 0, 0, 0, 0], 1, 0, 1.0, 5.0)
    _run(tmp_path, "0", "-s", "s.bin", "-distr", "--quiet")
    chains = _read_save(tmp_path / "s.bin")
    Z = np.array([c[0] for c in chains])
    dz = np.loadtxt(tmp_path / "distr_Z.dat")
    zs = np.arange(Z.min(), Z.max() + 1)
    np.testing.assert_array_equal(dz[:, 0], zs)
    np.testing.assert_allclose(dz[:, 1], [np.float32(np.sum(Z == z)) / np.float32(n) for z in zs], rtol=1e-5)
    Ns = np.concatenate([c[3][:, 3].astype(np.int64) for c in chains])
    dn = np.loadtxt(tmp_path / "distr_N.dat")
    vs = np.arange(Ns.min(), Ns.max() + 1)
    np.testing.assert_array_equal(dn[:, 0], vs)
    np.testing.assert_allclose(dn[:, 1], [np.float32(np.sum(Ns == v)) / np.float32(Ns.size) for v in vs], rtol=1e-5,
                               atol=1e-12)
    q = []
    for z, _, _, qn, _ in chains:
        x, y, zz = qn[1:z - 1, 0], qn[1:z - 1, 1], qn[1:z - 1, 2]
        q.append(np.sqrt((x * x + y * y) + zz * zz).astype(np.float32))
    q = np.sort(np.concatenate(q))
    P = np.arange(q.size, dtype=np.float64) / q.size
    ties = np.nonzero(q[1:] == q[:-1])[0]
    P[ties] = (ties + 1) / q.size
    dq = np.loadtxt(tmp_path / "distr_Q.dat")
    sel = np.arange(q.size // 10) * 10
    assert dq.shape[0] == sel.size
    np.testing.assert_allclose(dq[:, 0], q[sel], rtol=1e-5)
    np.testing.assert_allclose(dq[:, 1], P[sel].astype(np.float32), rtol=1e-5, atol=1e-7)


def test_stress_of_current_state(engine):
    """dsm_stress against an FP64 evaluation of the stored state on the host, with and
    without a deformation pending from the last jump (dt_prev)."""
    n, nk = 500, 50
    kappa = np.array([0.01, 0.05, 0, -0.02, 0, 0.03, 0, 0, -0.01])
    a = engine.add_ensemble(n, nk, 1.0, 1, kappa=kappa, seed=3)
    engine.init_equilibrium(a)
    engine.advance(a, 3.0)
    st = engine.state(a)
    for dtp in (None, 0.125):
        if dtp is not None:
            st["dt_prev"][:] = dtp
            engine.set_state(a, st["Z"], st["N"], st["Q"], st["w"], t=st["t"], ev=st["ev"], stall=st["stall"],
                             dt_prev=st["dt_prev"], next_sample=st["next_sample"])
        got, healthy = engine.stress(a)
        Q = st["Q"].astype(np.float64)
        if dtp is not None:
            K = kappa.astype(np.float32).astype(np.float64).reshape(3, 3)
            Q = Q + dtp * np.einsum("ij,nsj->nsi", K, Q)
        on = np.arange(st["N"].shape[1])[None, :] < st["Z"][:, None]
        Nn = np.where(on, st["N"], 1).astype(np.float64)
        t3 = np.where(on, 3.0 / Nn, 0.0)
        x, y, z = Q[..., 0], Q[..., 1], Q[..., 2]
        want = [-(t3 * x * x).sum(1), -(t3 * y * y).sum(1), -(t3 * z * z).sum(1), -(t3 * x * y).sum(1),
                -(t3 * y * z).sum(1), -(t3 * x * z).sum(1), np.where(on, np.sqrt(x * x + y * y + z * z), 0).sum(1),
                st["Z"].astype(np.float64)]
        want = np.array([v.mean() for v in want])
        assert healthy == n
        np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-11)
    engine.remove_ensemble(a)
