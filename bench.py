"""Throughput of the KPM Chebyshev path on one MI355X.

Workload (the precision form of the KPM DOS benchmark): honeycomb lattice of
1024 x 1024 two-atom cells (N = 2,097,152), periodic, t = -2.7 eV, onsite
disorder U[-0.5, 0.5] eV (seeded), ELL width K = 4, spectral bound
B = (3|t| + W/2) * 1.01; M = 2048 Chebyshev moments for R = 16 random-phase
vectors (seeds 1000..1015), complex64 recursion with FP64 moment dots, and
the Jackson-damped DOS on 4096 energies.

One step = the whole moment recursion (M - 1 fused SpMV steps over the R
vectors) plus the DOS reconstruction, inputs resident in HBM.  The metric is
N*M*R units per second.  The JSON line also carries
  roofline      the fused step kernel's algorithmic bytes per launch divided
                by its mean launch duration (HIP events on the launch
                stream around runs of 8 consecutive step launches, one run
                every `--profile-every` steps of the timed region)
                against the 8 TB/s HBM peak, plus HBM traffic per launch from
                the committed rocprofv3 counter summary;
  cpu_baseline  the complex128 CPU oracle (the reference algorithm, one
                vector at a time) on a bounded sample of the same workload;
  dos_err_vs_c128  max |DOS_c64 - DOS_c128| / max |DOS_c128| on the same
                inputs (the complex128 GPU path is pinned to the oracle by the
                tests);
  doubling      the same moments from M/2 steps of the doubling form
                (Hermitian H), with its own roofline and DOS error;
  plain_c128    the reference's structure on the GPU (complex128, one vector
                per recursion, one moment per step) and the headline's and
                doubling form's ratios over it;
  mu_err_vs_c128  max_n,r |mu_c64 - mu_c128| / mu_0;
  config1 / batch / grad / kubo  the other four workloads (plain complex128
                DOS with the oracle timed on the full workload; 64 disorder
                realisations in one batch; the adjoint gradient; the Kubo
                double moments), each with its own metric, the event-timed
                roofline of its dominant kernel and its accuracy field
                (parity with the oracle, error against complex128 / FP64, or
                the finite-difference check), over a few timed steps each
                (--sub-steps; --no-subs leaves them out).
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = ("Chebyshev SpMV-steps/s (N·M·R/s) + % HBM roofline at DOS err "
          "≤1e-5 vs FP64")
HBM_PEAK_GBS = 8000.0
PROFILES = ROOT / "profiles"
TRAFFIC_FILE = PROFILES / "cheb_step_c64_r16_honeycomb1024_pmc.md"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_workload(L, M, R, seed):
    import minijclsquant_amd as jcl

    H, Dx, Dy = jcl.honeycomb_lattice(L, t=-2.7, W=1.0, seed=seed)
    A = jcl.ell_matrix(H)
    B = (3 * 2.7 + 0.5) * 1.01
    A.bounds = np.array([-B, B])
    N = int(A.shape[0])
    r = np.stack([jcl.random_vector_generator(N, 1000 + s) for s in range(R)])
    return A, r, B


def profile_traffic(path, key, **shape):
    """HBM bytes per launch (FETCH_SIZE and WRITE_SIZE, calibrated) from a
    committed counter summary, if its `workload:` line names exactly this
    shape (every given key=value matches), else None."""
    if not path.exists():
        return None
    txt = path.read_text()
    m = re.search(rf"^{key}:\s*([0-9.eE+]+)", txt, re.M)
    w = re.search(r"^workload:(.*)$", txt, re.M)
    if not m or not w:
        return None
    have = {k: int(v) for k, v in re.findall(r"(\w+)=(\d+)", w.group(1))}
    if any(have.get(k) != int(v) for k, v in shape.items()):
        return None
    return float(m.group(1))


def traffic_from_profile(N, K, R, key="traffic_bytes_per_launch", g=None):
    """The headline workload's traffic figure (see profile_traffic); for
    passes of g < R vectors, the figure of one such pass's launch."""
    if g is not None and g < R:
        key = f"{key}_pass{g}"
    return profile_traffic(TRAFFIC_FILE, key, N=N, K=K, R=R)


def pass_note(N, K, R, g, mv, s_v, s_l, launch_ms):
    """Roofline fields for a recursion run in passes of g < R vectors: the
    bytes one step of all R vectors moves (the matrix once per pass) and
    their rate, and the bytes of the same step at R vectors per pass (the
    matrix once), over the same time."""
    per_vec = 3 * s_v + (s_l if s_l else 0)
    passes = -(-R // g)
    step_s = launch_ms / 1e3 * R / g
    all_bytes = passes * N * K * (mv + 4) + R * N * per_vec
    one_bytes = N * K * (mv + 4) + R * N * per_vec
    return {
        "passes_per_step": passes,
        "step_bytes_all_passes": all_bytes,
        "step_GBs_all_passes": round(all_bytes / step_s / 1e9, 1),
        "single_pass_bytes": one_bytes,
        "single_pass_equiv_GBs": round(one_bytes / step_s / 1e9, 1),
        "single_pass_equiv_frac": round(one_bytes / step_s / 1e9 / HBM_PEAK_GBS, 4),
        "pass_note": (f"the {R} vectors run as {passes} recursions of {g}: one pass's "
                      f"vectors and the matrix ({(N * K * (mv + 4) + 2 * g * N * s_v + g * N * (s_l or 0)) / 1e6:.0f} MB) "
                      "stay in the 256 MB Infinity Cache, so a launch reads much of its "
                      "algorithmic bytes from there; single_pass_equiv_* counts the bytes of "
                      "one pass of all R vectors over the same time"),
    }


def copy_ceiling(torch, nat, nbytes=2 << 30, reps=5):
    """HBM rate of the library's plain copy kernel (bytes read + written per
    second), timed with HIP events on the current stream."""
    src = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    nat.copy(dst, src, nbytes)
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(reps):
        nat.copy(dst, src, nbytes)
    t1.record()
    t1.synchronize()
    ms = t0.elapsed_time(t1) / reps
    del src, dst
    return 2 * nbytes / (ms / 1e3) / 1e9


def host_cpu():
    """(threads to use, logical CPUs visible, CPU model).  The threads are
    OMP_NUM_THREADS when set (the GPU box sets it to its per-GPU CPU share),
    else min(nproc, 16)."""
    nproc = os.cpu_count() or 1
    thr = int(os.environ.get("OMP_NUM_THREADS", min(nproc, 16)))
    model = None
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return thr, nproc, model


def cpu_baseline(A, r, target_s=12.0):
    """The reference's CPU path on host cores: its rec_A_tab loop over its own
    Cython kernels compiled into oracle/_ref (kind "reference"), else the C
    restatement of the oracle (kind "port"); complex128, one vector, a
    bounded number of moments sized to ~target_s seconds.  When both exist
    the port is timed on the same sample beside it."""
    from oracle import oracle as O
    from oracle import ref

    thr, nproc, model = host_cpu()
    Ht = A.data * (1 / A.bounds[1])
    N = int(A.shape[0])

    def port(m):
        return O.rec_A_tab(m, Ht, A.indices, A.len_row, r[0], n_threads=thr)

    def reference(m):
        return ref.rec_A_tab(m, Ht, A.indices, A.len_row, r[0], n_threads=thr)

    kind, run = ("reference", reference) if ref.available() else ("port", port)
    m_probe = 16
    t0 = time.perf_counter()
    run(m_probe)
    dt = time.perf_counter() - t0
    m = int(max(m_probe, min(4096, m_probe * target_s / max(dt, 1e-3))))
    t0 = time.perf_counter()
    mu_run = run(m)
    dt = time.perf_counter() - t0
    extra = {}
    if kind == "reference":
        t0 = time.perf_counter()
        mu_port = port(m)
        dt_p = time.perf_counter() - t0
        extra = {"port_value": N * m / dt_p,
                 "port_max_abs_dmu_over_mu0": float(np.max(np.abs(mu_port - mu_run)) / N)}
    return {
        "value": N * m / dt,
        "unit": "N*M*R/s",
        "cores": thr,
        "kind": kind,
        "sample": f"same H, R=1 vector, M={m} moments, complex128, {dt:.1f} s",
        **extra,
        "nproc": nproc,
        "cpu_model": model,
        "threads_note": "OpenMP threads = OMP_NUM_THREADS (the host's CPU share per GPU); "
                        "nproc is every logical CPU the process sees",
    }


def tiled_leg(torch, nat, h, vr, N, K, M, R, s_v, args, units, es):
    """The same moments through the two-steps-per-launch path (rows
    renumbered into LDS patches), plain and doubling forms, with the DOS; the
    kernel's algorithmic bytes are those of two fused steps."""
    plan = nat.TilePlan(h, R)
    info = plan.info()
    ws = torch.empty(plan.workspace_bytes(R, M, True), dtype=torch.uint8, device="cuda")
    mu = torch.empty((M, R), dtype=torch.complex128, device="cuda")

    def step(doubling):
        plan.moments(vr, M, doubling=doubling, workspace=ws, out=mu)
        return nat.dos(mu, es, nat.JKPM_KERNEL_JACKSON)

    out = {"patches": info["npatch"], "halo_rows": info["halo_rows"], "max_ext": info["max_ext"]}
    for doubling in (False, True):
        step(doubling)
        torch.cuda.synchronize()
        nat.profile_read(reset=True)
        nat.profile_enable(max(1, args.profile_every // 2))
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step(doubling)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        nat.profile_enable(0)
        ms_sum, n_prof = nat.profile_read(reset=True)
        launch_ms = ms_sum / max(n_prof, 1)
        nv = 4 if doubling else 5
        bytes_launch = N * K * (s_v + 4) + R * N * nv * s_v
        ach = bytes_launch / (launch_ms / 1e3) / 1e9
        key = "doubling_" if doubling else ""
        out[key + "value"] = units / (ms / 1e3)
        out[key + "ms_per_step"] = ms
        out[key + "roofline"] = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                                 "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                                 "algorithmic_bytes_per_launch": bytes_launch,
                                 "mean_launch_us": round(launch_ms * 1e3, 2),
                                 "launches_timed": n_prof}
    del plan
    return out


def wave_leg(torch, nat, h, vr, N, K, M, R, s_v, args, units, es):
    """The same moments through the wavefront path (two steps per launch, the
    work-groups of one XCD handing T_{n+1} over through L2 behind progress
    flags), plain and doubling forms, on the handle's tuned chunking; the
    kernel's algorithmic bytes are those of two fused steps.  The per-launch
    time covers the main launch and its fix-up and fold launches."""
    out = {}
    for doubling in (False, True):
        key = "doubling_" if doubling else ""
        try:
            plan = nat.WavePlan(h, R, doubling)
        except ValueError as e:  # chunking or shape the path does not take
            out[key + "skipped"] = str(e)
            continue
        info = plan.info()
        out.update({"lag": info["lag"], "tail_rows": info["tail_rows"],
                    "fixup_rows": info["fixup_rows"], "nchunks": info["nchunks"]})
        ws = torch.empty(plan.workspace_bytes(M), dtype=torch.uint8, device="cuda")
        mu = torch.empty((M, R), dtype=torch.complex128, device="cuda")

        def step():
            plan.moments(vr, M, workspace=ws, out=mu)
            return nat.dos(mu, es, nat.JKPM_KERNEL_JACKSON)

        step()
        torch.cuda.synchronize()
        reps = max(1, min(args.steps, 2))
        t0 = time.perf_counter()
        for _ in range(reps):
            step()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / reps
        launch_ms, n_prof = profiled(nat, step, max(1, args.profile_every // 2), nat.JKPM_PROF_WAVE)
        nv = 4 if doubling else 5
        bytes_launch = N * K * (s_v + 4) + R * N * nv * s_v
        out[key + "value"] = units / (ms / 1e3)
        out[key + "ms_per_step"] = ms
        out[key + "roofline"] = hbm_roofline(bytes_launch, launch_ms, n_prof, "wave2_kernel")
        del plan
    out["note"] = ("bound by the latency of its dependent loads like the one-step kernel: "
                   "fewer bytes per step do not shorten the chain, and the flag reads add "
                   "to it; the one-step kernel stays the headline path")
    return out


def plain_c128(torch, nat, _kpm, A, r, B, M, value, value_dbl, tune, reps=3):
    """The reference's structure on this GPU: complex128 recursion on
    complex128 matrix values (as the reference stores them), one vector at a
    time (R = 1), one moment per step, with its own tuned launch shape.
    Timed with HIP events on the current stream over `reps` recursions; the
    same path on the packed real copy of the (real) values is timed beside
    it."""
    N = int(A.shape[0])
    h = A.device("c128", 1.0 / B)
    v1 = _kpm.as_batch(r[0], N, "c128")
    ws = torch.empty(h.workspace_bytes(1, 1), dtype=torch.uint8, device="cuda")
    mu = torch.empty((1, M, 1), dtype=torch.complex128, device="cuda")

    def timed(real):
        h.use_real_values(real)
        launch = None
        if tune:
            from minijclsquant_amd import autotune

            launch = autotune.tuned(h, v1)
        h.moments(v1, M, workspace=ws, out=mu)
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0.record()
        for _ in range(reps):
            h.moments(v1, M, workspace=ws, out=mu)
        t1.record()
        t1.synchronize()
        return t0.elapsed_time(t1) / reps, launch

    ms_r, launch_r = timed(True) if h.is_real() else (None, None)
    ms, launch = timed(False)
    v = N * M / (ms / 1e3)
    v_r = None if ms_r is None else N * M / (ms_r / 1e3)
    # the reference-compatible host-pointer entry (rec_A_tab): upload of the
    # matrix and vector over PCIe, the recursion, download of the moments
    Ht = np.ascontiguousarray(A.data * (1.0 / B))
    nat.rec_A_tab_host(Ht, A.indices, int(A.len_row), r[0], M)
    t_h = time.perf_counter()
    nat.rec_A_tab_host(Ht, A.indices, int(A.len_row), r[0], M)
    host_s = time.perf_counter() - t_h
    return {
        "value": v,
        "host_entry_value": N * M / host_s,
        "host_entry_ms": host_s * 1e3,
        "unit": "N*M*R/s",
        "ms_per_recursion": ms,
        "sample": f"same H, R=1 vector, M={M} moments, complex128 vectors and matrix values, "
                  f"{reps} recursions",
        "launch": launch,
        "headline_over_plain": value / v,
        "doubling_over_plain": None if value_dbl is None else value_dbl / v,
        "real_values": None if v_r is None else {
            "value": v_r, "ms_per_recursion": ms_r, "launch": launch_r,
            "note": "the same path with the step kernels on the packed real copy of the values",
            "headline_over_it": value / v_r,
            "doubling_over_it": None if value_dbl is None else value_dbl / v_r},
    }


def event_ms(torch, fn, reps):
    """Mean milliseconds of fn() over reps calls, HIP events on the current
    stream."""
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(reps):
        fn()
    t1.record()
    t1.synchronize()
    return t0.elapsed_time(t1) / reps


def profiled(nat, fn, every=4, kind=0):
    """(mean ms per launch, launches) of one kind of kernel, from the
    library's HIP-event runs around its launches during one untimed call of
    fn() (kind 0: Chebyshev step launches; see jkpm_profile_read_kind)."""
    nat.profile_read(reset=True, kind=kind)
    nat.profile_enable(every)
    try:
        fn()
    finally:
        nat.profile_enable(0)
    ms, n = nat.profile_read(reset=True, kind=kind)
    return ms / max(n, 1), n


def hbm_roofline(bytes_launch, launch_ms, n, kernel, **extra):
    ach = bytes_launch / (launch_ms / 1e3) / 1e9
    out = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "kernel": kernel,
           "algorithmic_bytes_per_launch": int(bytes_launch),
           "mean_launch_us": round(launch_ms * 1e3, 2), "launches_timed": n}
    out.update(extra)
    return out


def base_line(metric, value, unit, args, ms, dtype, config, higher=True):
    return {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": max(1, args.warmup), "ms_per_step": ms, "higher_is_better": higher,
            "scaling": "weak", "vs_baseline": None, "dtype": dtype,
            "data": "synthetic (seeded lattice disorder and random-phase vectors)",
            "config": config}


def ingest_leg(blocks, torch, nat):
    """One-time cost of getting the realisations onto the device, two ways:
    host CSR -> row-major ELL per realisation (what ell_matrix does,
    jclsquant/ell_matrix.py:187-200), block stacking and upload; or stacking
    the CSR arrays and converting on the device (jkpm_ell_create_csr).  Both
    handles must hold identical columns and values."""
    import scipy.sparse as sp

    from minijclsquant_amd.batch import block_diagonal, stack_csr
    from minijclsquant_amd.ell_matrix import data_indices_cython_0

    mats = [sp.csr_matrix((d, ix, np.arange(int(n) + 1, dtype=np.int64) * int(k)),
                          shape=(int(n), int(n))) for d, ix, n, k, _ in blocks]
    scales = [b[4] for b in blocks]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ell_blocks = []
    for m, s in zip(mats, scales):
        d, ix, _ = data_indices_cython_0(5, m.shape[0], m.indptr, m.data, m.indices)
        ell_blocks.append((d, ix, m.shape[0], 5, s))
    data, idx, offsets, K = block_diagonal(ell_blocks)
    h_host = nat.DeviceELL(data, idx, int(offsets[-1]), K, 1.0, "c64")
    torch.cuda.synchronize()
    t_host = time.perf_counter() - t0
    del ell_blocks, data, idx
    t0 = time.perf_counter()
    indptr, indices, cdata, offs = stack_csr(mats, scales)
    h_dev = nat.DeviceELL.from_csr(indptr, indices, cdata, int(offs[-1]), 1.0, "c64")
    torch.cuda.synchronize()
    t_dev = time.perf_counter() - t0
    del indptr, indices, cdata
    same = bool(torch.equal(h_host.indices(), h_dev.indices()) and
                torch.equal(h_host.values(), h_dev.values()))
    if not same:
        raise RuntimeError("device CSR ingest differs from the host ELL build")
    del h_host, h_dev
    torch.cuda.synchronize()
    log(f"ingest: host CSR->ELL + upload {t_host:.2f} s, device CSR ingest {t_dev:.2f} s")
    return {"host_csr_to_ell_upload_s": round(t_host, 3),
            "device_csr_ingest_s": round(t_dev, 3), "identical": same,
            "note": "one-time, outside the timed region: row-major ELL per realisation on "
                    "the host + stacking + upload, against stacked CSR converted on the "
                    "device"}


def bench_batch(args, torch, jcl, nat, _kpm):
    """64 square-lattice realisations, L ~ U{128..1024}, W ~ U[0.5, 3],
    t = -1, K = 5, R = 8, M = 1024, complex64, one block-diagonal batch:
    the whole recursion of all realisations replayed as one HIP graph."""
    from minijclsquant_amd.batch import BatchKPM, square_lattice_ell

    G, R, M, K = 64, 8, 1024, 5
    rng = np.random.default_rng(2024)
    Ls = rng.integers(128, 1025, size=G)
    Ws = rng.uniform(0.5, 3.0, size=G)
    log(f"building {G} realisations, sum N = {int((Ls ** 2).sum())}")
    blocks, vecs = [], []
    for g in range(G):
        L = int(Ls[g])
        d, ix = square_lattice_ell(L, -1.0, float(Ws[g]), seed=5000 + g)
        blocks.append((d, ix, L * L, K, 1.0 / ((4 + Ws[g] / 2) * 1.01)))
        vecs.append(np.stack([jcl.random_vector_generator(L * L, 100 * g + s)
                              for s in range(R)]))
    ingest = ingest_leg(blocks, torch, nat)
    bk = BatchKPM(blocks, vecs, M, "c64", tune=not args.no_tune)
    N = int(bk.offsets[-1])
    log(f"launch shapes (variant, work-groups per CU): plain {bk.H.get_launch()}, "
        f"doubling {bk.H.get_launch(doubling=True)}")
    bk.capture()
    for _ in range(max(1, args.warmup)):  # tuning and allocation stay untimed
        bk.replay()
    torch.cuda.synchronize()
    ms = event_ms(torch, bk.replay, args.steps)
    log(f"batch graph replay: {ms:.1f} ms")
    ws_d = torch.empty(nat.moments_ex_workspace_bytes(None, bk.batch, R, M, True),
                       dtype=torch.uint8, device="cuda")
    mu_d = torch.empty((G, M, R), dtype=torch.complex128, device="cuda")

    def dbl():
        nat.moments_ex(None, bk.batch, bk.vr, M, True, workspace=ws_d, out=mu_d)

    dbl()
    ms_d = event_ms(torch, dbl, args.steps)
    diff = float((mu_d - bk.mu).abs().max() / bk.mu[:, 0].abs().max())
    units = N * M * R
    mv = 4 if bk.H.is_real() else 8  # matrix value bytes per slot (packed real copy)
    # the random-phase left vectors (v_l = v_r) are read as 32-bit phases
    s_l = 4 if nat.phase_pack(bk.vr)[1] else 8
    # vectors per recursion pass of each form (all R unless the tuner split them)
    g_pass = min(bk.H.get_pass() or R, R)
    gd_pass = min(bk.H.get_pass(doubling=True) or R, R)
    bytes_launch = N * K * (mv + 4) + g_pass * N * (3 * 8 + s_l)
    bytes_d = N * K * (mv + 4) + gd_pass * N * 3 * 8
    ach = bytes_launch * (M - 1) * (-(-R // g_pass)) / (ms / 1e3) / 1e9
    ach_d = bytes_d * max(1, M // 2) * (-(-R // gd_pass)) / (ms_d / 1e3) / 1e9
    # the step kernel alone: event-timed runs of launches of one eager pass
    # (a graph replay records no events)
    launch_ms, n_prof = profiled(nat, bk.run, every=16)
    torch.cuda.synchronize()
    # every realisation against its own complex128 recursion (the per-matrix
    # path the tests pin to the oracle): moments and the Jackson DOS on 1024
    # rescaled energies
    mu64 = bk.replay().cpu().numpy()  # [G, M, R]
    es = torch.as_tensor(np.linspace(-0.99, 0.99, 1024), device="cuda")
    mu_err = dos_err = 0.0
    for g, (d, ix, n_g, k_g, sc) in enumerate(blocks):
        hg = nat.DeviceELL(d, ix, n_g, k_g, sc, "c128")
        vg = _kpm.as_batch(vecs[g], n_g, "c128")
        m128 = hg.moments(vg, M)[0]
        mu_err = max(mu_err, float(np.max(np.abs(mu64[g] - m128.cpu().numpy())) / n_g))
        d128 = nat.dos(m128, es, nat.JKPM_KERNEL_JACKSON).cpu().numpy()
        d64 = nat.dos(torch.as_tensor(mu64[g], device="cuda"), es,
                      nat.JKPM_KERNEL_JACKSON).cpu().numpy()
        dos_err = max(dos_err, float(np.max(np.abs(d64 - d128)) / np.max(np.abs(d128))))
        del hg, vg, m128
    del blocks, vecs
    log(f"batch vs complex128: max |dmu|/mu0 {mu_err:.2e}, DOS err {dos_err:.2e}")
    out = base_line("batched realisations: Chebyshev SpMV-steps/s (sum N * M * R / s)",
                    units / (ms / 1e3), "N*M*R/s", args, ms, "c64",
                    {"workload": f"{G} square-lattice realisations, L ~ U{{128..1024}} "
                                 "(seed 2024), W ~ U[0.5, 3], t = -1, K = 5, R = 8, "
                                 "M = 1024, complex64 recursion + FP64 moments, one "
                                 "block-diagonal batch replayed as one HIP graph",
                     "G": G, "sum_N": N, "K": K, "M": M, "R": R})
    out["roofline"] = hbm_roofline(bytes_launch, launch_ms, n_prof, "cheb_step_kernel",
                                   whole_replay_GBs=round(ach, 1),
                                   whole_replay_frac=round(ach / HBM_PEAK_GBS, 4),
                                   left_vector_bytes_per_element=s_l,
                                   vectors_per_pass=g_pass)
    # (the committed counters are those of launches carrying all R vectors)
    out["roofline"]["traffic"] = profile_traffic(
        PROFILES / "cheb_step_c64_r8_batch64_square_trace.md", "traffic_bytes_per_launch",
        N=N, K=K, R=R) if g_pass == R else None
    out["mu_err_vs_c128"] = mu_err
    out["dos_err_vs_c128"] = dos_err
    out["accuracy_note"] = ("max over the 64 realisations and all n of |mu_c64 - mu_c128| / mu_0, "
                            "and of the Jackson DOS error on 1024 rescaled energies")
    out["doubling"] = {"value": units / (ms_d / 1e3), "ms_per_step": ms_d,
                       "frac": round(ach_d / HBM_PEAK_GBS, 4),
                       "max_rel_moment_diff_vs_plain": diff}
    out["ingest"] = ingest
    return out


def bench_config1(args, torch, jcl, nat, _kpm):
    """The reference's own CPU-runnable case: plain KPM DOS in complex128 on a
    square 256 x 256 periodic lattice, t = -1, onsite U[-0.5, 0.5] (W = 1),
    K = 5, B = (4 + W/2) * 1.01, M = 512, R = 4 phase vectors (seeds 0..3),
    Jackson kernel, DOS on 1024 energies.  GPU: moments of the R vectors in
    one pass per step plus the DOS on the device.  CPU: the oracle (the
    reference algorithm, one vector at a time) on the same H and vectors, in
    full.  Parity of the two in the same line."""
    from oracle import oracle as O

    L, W, M, R, n_e = 256, 1.0, 512, 4, 1024
    B = (4 + W / 2) * 1.01
    H, _ = jcl.square_lattice(L, t=-1.0, W=W, seed=11)
    A = jcl.ell_matrix(H, bounds=[-B, B])
    N, K = int(A.shape[0]), int(A.len_row)
    r = np.stack([jcl.random_vector_generator(N, s) for s in range(R)])
    E = np.linspace(-B + 1e-3, B - 1e-3, n_e)
    es = torch.as_tensor(E / B, device="cuda")
    h = A.device("c128", 1.0 / B)
    vr = _kpm.as_batch(r, N, "c128")
    if not args.no_tune:
        from minijclsquant_amd import autotune

        autotune.tuned(h, vr)
    ws = torch.empty(h.workspace_bytes(R, 1), dtype=torch.uint8, device="cuda")
    mu = torch.empty((1, M, R), dtype=torch.complex128, device="cuda")

    dos_out = torch.empty(n_e, dtype=torch.float64, device="cuda")
    dws = torch.empty(M, dtype=torch.float64, device="cuda")
    assert float(es.abs().max()) < 1.0

    def step(stream=None):
        h.moments(vr, M, workspace=ws, out=mu, stream=stream)
        nat.dos(mu[0], es, nat.JKPM_KERNEL_JACKSON, stream=stream, out=dos_out, workspace=dws,
                check_range=False)

    for _ in range(max(1, args.warmup)):  # tuning and allocation stay untimed
        step()
    torch.cuda.synchronize()
    ms = event_ms(torch, step, args.steps)
    launch_ms, n_prof = profiled(nat, step, every=4)
    mu_gpu = mu[0].cpu().numpy()
    # Side figure: the same moments from M/2 steps (Hermitian H, v_l = v_r).
    if not args.no_tune:
        autotune.tuned(h, vr, doubling=True)
    ws_d = torch.empty(nat.moments_ex_workspace_bytes(h, None, R, M, True), dtype=torch.uint8,
                       device="cuda")
    mu_d = torch.empty((M, R), dtype=torch.complex128, device="cuda")
    dos_d = torch.empty(n_e, dtype=torch.float64, device="cuda")

    def step_d(stream=None):
        nat.moments_ex(h, None, vr, M, True, workspace=ws_d, out=mu_d, stream=stream)
        nat.dos(mu_d, es, nat.JKPM_KERNEL_JACKSON, stream=stream, out=dos_d, workspace=dws,
                check_range=False)

    for _ in range(max(1, args.warmup)):  # tuning and allocation stay untimed
        step_d()
    torch.cuda.synchronize()
    ms_d = event_ms(torch, step_d, args.steps)
    mu_dbl = mu_d.cpu().numpy()
    # the package's DOS table (GPU) for the parity check against the oracle's
    dos_gpu = jcl.kpm_dos(A, M, r, energies=E)[:, 1]
    # the oracle on the same inputs, every vector, timed on the host cores
    thr, nproc, model = host_cpu()
    Ht = A.data * (1 / B)
    t0 = time.perf_counter()
    mu_cpu = np.stack([O.rec_A_tab(M, Ht, A.indices, K, r[k], n_threads=thr) for k in range(R)], 1)
    dos_cpu = O.dos(mu_cpu.mean(1), E, A.bounds[0], A.bounds[1], True)
    cpu_s = time.perf_counter() - t0
    err_mu = float(np.max(np.abs(mu_gpu - mu_cpu)) / N)
    err_dos = float(np.max(np.abs(dos_gpu - dos_cpu)) / np.max(np.abs(dos_cpu)))
    # the reference's own CPU kernels (oracle/_ref) on the same full workload
    from oracle import ref

    ref_s = err_ref = None
    if ref.available():
        t0 = time.perf_counter()
        mu_ref = np.stack([ref.rec_A_tab(M, Ht, A.indices, K, r[k], n_threads=thr)
                           for k in range(R)], 1)
        O.dos(mu_ref.mean(1), E, A.bounds[0], A.bounds[1], True)
        ref_s = time.perf_counter() - t0
        err_ref = float(np.max(np.abs(mu_gpu - mu_ref)) / N)
    units = N * M * R
    out = base_line("plain KPM DOS, complex128: N * M * R / s", units / (ms / 1e3), "N*M*R/s",
                    args, ms, "c128",
                    {"workload": "square 256x256 periodic, t = -1, onsite U[-0.5, 0.5] (W = 1, "
                                 "seed 11), K = 5, B = (4 + W/2) * 1.01, M = 512, R = 4 phase "
                                 "vectors (seeds 0..3), Jackson kernel, DOS on 1024 energies",
                     "N": N, "K": K, "M": M, "R": R})
    mv = 8 if h.is_real() else 16  # matrix value bytes per slot (packed real copy)
    bytes_launch = N * K * (mv + 4) + R * N * 4 * 16
    ach = bytes_launch * (M - 1) / (ms / 1e3) / 1e9
    out["roofline"] = hbm_roofline(
        bytes_launch, launch_ms, n_prof, "cheb_step_kernel",
        whole_step_GBs=round(ach, 1),
        note="the 23 MB working set is cache-resident (256 MB Infinity Cache) and each launch "
             "is latency-bound, so this is not an HBM-bound figure; traffic is what leaves "
             "the L2, served by the Infinity Cache")
    out["roofline"]["traffic"] = profile_traffic(
        PROFILES / "cheb_step_c128_r4_square256_pmc.md", "traffic_bytes_per_launch",
        N=N, K=K, R=R)
    if ref_s is not None:
        out["cpu_baseline"] = {"value": units / ref_s, "unit": "N*M*R/s", "cores": thr,
                               "kind": "reference", "nproc": nproc, "cpu_model": model,
                               "sample": f"the full workload (R = {R} vectors, M = {M}, DOS): "
                                         f"the reference's rec_A_tab loop over its compiled "
                                         f"aAxby/vdot, {ref_s:.2f} s",
                               "port_value": units / cpu_s,
                               "max_abs_dmu_over_mu0_vs_gpu": err_ref}
    else:
        out["cpu_baseline"] = {"value": units / cpu_s, "unit": "N*M*R/s", "cores": thr,
                               "kind": "port", "nproc": nproc, "cpu_model": model,
                               "sample": f"the full workload (R = {R} vectors, M = {M}, DOS), "
                                         f"complex128 oracle, {cpu_s:.2f} s"}
    out["gpu_over_cpu"] = (units / (ms / 1e3)) / out["cpu_baseline"]["value"]
    out["note"] = ("511 step launches of a 23 MB working set take ~5 us each; capturing the "
                   "step into a HIP graph measured the same (2.78 vs 2.76 ms), so the launches "
                   "are kernel-bound, not host-bound")
    out["parity"] = {"max_abs_dmu_over_mu0": err_mu, "dos_rel_err": err_dos, "bar": 1e-10}
    out["doubling"] = {"value": units / (ms_d / 1e3), "ms_per_step": ms_d,
                       "steps_per_recursion": max(1, M // 2),
                       "max_abs_dmu_over_mu0": float(np.max(np.abs(mu_dbl - mu_cpu)) / N),
                       "note": "the same M moments from M/2 Chebyshev steps (mu_2k = 2<T_k|T_k> - "
                               "mu_0, mu_2k+1 = 2<T_k|T_k+1> - mu_1; Hermitian H), checked "
                               "against the oracle's moments"}
    return out


def bench_grad(args, torch, jcl, nat, _kpm):
    """Square 512 x 512, onsite U[-1, 1], t_ij = -1 + N(0, 0.05^2), M = 1024,
    R = 8, complex128: loss = sum_E w(E) DOS(E) (256 energies, Gaussian w),
    forward (moments + DOS) and backward (adjoint recursion) timed."""
    from minijclsquant_amd.autograd import (chebyshev_moments, kpm_dos_torch, lattice_values,
                                            param_slots)

    L, M, R, n_e, sig = 512, 1024, 8, 256, 0.05
    H, _ = jcl.square_lattice(L, t=-1.0, seed=2024, W=2.0, bond_noise=sig)
    A = jcl.ell_matrix(H)
    B = (4 * (1 + 6 * sig) + 1.0) * 1.01
    N, K = int(A.shape[0]), int(A.len_row)
    slot_param, pairs = param_slots(A)
    rows = np.arange(A.data.size) // K
    eps0 = np.real(A.data[(rows == A.indices) & (A.valid_row_values == 1)])
    Hc = A.A.tocsr()
    t0 = np.real(np.asarray(Hc[pairs[:, 0], pairs[:, 1]]).ravel())
    r = np.stack([jcl.random_vector_generator(N, 300 + s) for s in range(R)])
    rd = torch.from_numpy(r.T.copy()).cuda()
    idx = torch.from_numpy(A.indices).cuda()
    es = torch.as_tensor(np.linspace(-B + 0.05, B - 0.05, n_e) / B, device="cuda")
    w = torch.as_tensor(np.random.default_rng(2025).normal(size=n_e), device="cuda")
    eps = torch.tensor(eps0, device="cuda", requires_grad=True)
    hop = torch.tensor(t0, device="cuda", requires_grad=True)

    seg = [None]  # None: every T_n kept when it fits (it does: 34 GB), else sqrt(M) checkpoints

    def forward():
        vals = lattice_values(eps, hop, slot_param)
        mu = chebyshev_moments(vals, idx, rd, M, scale=1.0 / B, seg=seg[0])
        return (w * kpm_dos_torch(mu, es)).sum()

    def both():
        eps.grad = None
        hop.grad = None
        forward().backward()

    for _ in range(max(1, args.warmup)):  # tuning and allocation stay untimed
        both()
    torch.cuda.synchronize()
    with torch.no_grad():
        ms_f = event_ms(torch, forward, args.steps)
        launch_ms, n_prof = profiled(nat, forward, every=16)
    ms_fb = event_ms(torch, both, args.steps)
    # adjoint gradient against central finite differences (h = 1e-6,
    # complex128 forward passes) on n_fd seeded coordinates
    both()
    g_all = torch.cat([eps.grad, hop.grad]).cpu().numpy()
    coords = np.random.default_rng(2026).choice(g_all.size, args.fd_coords, replace=False)
    fd = []
    with torch.no_grad():
        base = torch.cat([eps, hop]).detach().clone()
        for c in coords:
            vals = []
            for sgn in (1.0, -1.0):
                p = base.clone()
                p[c] += sgn * 1e-6
                eps.data.copy_(p[:N])
                hop.data.copy_(p[N:])
                vals.append(float(forward()))
            fd.append((vals[0] - vals[1]) / 2e-6)
        eps.data.copy_(base[:N])
        hop.data.copy_(base[N:])
    fd = np.array(fd)
    fd_err = float(np.linalg.norm(g_all[coords] - fd) / np.linalg.norm(fd))
    log(f"gradient vs finite differences ({len(coords)} coordinates): {fd_err:.2e}")
    units = N * M * R
    out = base_line("differentiation: forward + adjoint backward, N * M * R / s",
                    units / (ms_fb / 1e3), "N*M*R/s", args, ms_fb, "c128",
                    {"workload": "square 512x512 periodic, onsite U[-1, 1], "
                                 "t_ij = -1 + N(0, 0.05^2), K = 5, M = 1024, R = 8, "
                                 "complex128, loss = sum_E w(E) DOS(E) on 256 energies; "
                                 "gradients for all onsite energies and bonds",
                     "N": N, "K": K, "M": M, "R": R, "params": int(N + pairs.shape[0])})
    with torch.no_grad():  # real onsite energies and hoppings: the packed real copy
        real = bool((lattice_values(eps, hop, slot_param).imag == 0).all())
    mv = 8 if real else 16
    bytes_launch = N * K * (mv + 4) + R * N * 4 * 16
    out["forward_ms"] = ms_f
    out["backward_ms"] = ms_fb - ms_f
    out["backward_over_forward"] = (ms_fb - ms_f) / ms_f
    out["roofline"] = hbm_roofline(
        bytes_launch, launch_ms, n_prof, "cheb_step_kernel (forward recursion)",
        forward_whole_GBs=round(bytes_launch * (M - 1) / (ms_f / 1e3) / 1e9, 1),
        note="the 160 MB step working set stays partly in the 256 MB Infinity Cache")
    out["roofline"]["traffic"] = profile_traffic(
        PROFILES / "grad_adjoint_square512_pmc.md", "traffic_bytes_per_launch", N=N, K=K, R=R)
    out["grad_vs_fd"] = {"rel_err": fd_err, "coordinates": int(len(coords)), "h": 1e-6,
                         "bar": 1e-4}
    # the same with sqrt(M) checkpoints and recomputation (small memory)
    seg[0] = 0
    both()
    torch.cuda.synchronize()
    ms_fb_ck = event_ms(torch, both, args.steps)
    out["backward_mode"] = ("the forward keeps every T_n (M vector batches, 34 GB); the backward "
                            "runs the adjoint steps and folds only")
    out["sqrtM_checkpoints"] = {"ms_per_step": ms_fb_ck, "backward_ms": ms_fb_ck - ms_f,
                                "backward_over_forward": (ms_fb_ck - ms_f) / ms_f}
    return out


def kubo_mu_nm_fp64(torch, nat, _kpm, A, B, v_data, r, M, chunk=1 << 17):
    """mu_mn = (1/R) sum_r <r| v T_m(H~) v T_n(H~) |r> in complex128: both
    Chebyshev stacks from the complex128 step path (jkpm_moments_keep, the
    recursion the tests pin to the oracle), Rk_n = v T_n r by jkpm_aAxby, and
    the contraction sum_k conj(L[m, k]) Rk[n, k] as a complex128 GEMM in
    chunks of k.  Returns a complex128 [M, M] device tensor."""
    N, K = int(A.shape[0]), int(A.len_row)
    R = r.shape[0]
    H = nat.DeviceELL(A.data, A.indices, N, K, 1.0 / B, "c128")
    V = nat.DeviceELL.like(H, v_data, 1.0, "c128")
    rr = _kpm.as_batch(r, N, "c128")
    _, T = H.moments_keep(rr, M)  # [M, N, R]: T_n(H~) r
    Rk = torch.zeros((M, N, R), dtype=torch.complex128, device="cuda")
    for n in range(M):
        V.aAxby(1.0, 0.0, T[n], Rk[n])
    del T
    vr = Rk[0].clone()
    _, L = H.moments_keep(vr, M)  # L_m = T_m(H~) (v r)
    L2 = L[:M].reshape(M, N * R)
    R2 = Rk.reshape(M, N * R)
    mu = torch.zeros((M, M), dtype=torch.complex128, device="cuda")
    for k0 in range(0, N * R, chunk):
        k1 = min(N * R, k0 + chunk)
        a = L2[:, k0:k1].conj().contiguous()
        b = R2[:, k0:k1].transpose(0, 1).contiguous()
        mu += torch.mm(a, b)
        del a, b
    del L, Rk, L2, R2
    torch.cuda.synchronize()
    return mu / R


def bench_kubo(args, torch, jcl, nat, _kpm):
    """Honeycomb 512 x 512 cells (N = 524,288), t = -2.7, W = 1 eV, v_x on
    H's slots, M = 512, R = 4: mu_nm (complex64 stacks, FP32 MFMA GEMM,
    FP64 combine) and sigma on 512 energies; the GEMM and the stack steps
    timed alone by the library's event runs; mu_nm and sigma against the
    complex128 stacks contracted in FP64."""
    from minijclsquant_amd.kubo import velocity_handle

    Lc, M, R, n_e = 512, 512, 4, 512
    H, Dx, Dy = jcl.honeycomb_lattice(Lc, t=-2.7, W=1.0, seed=2024)
    A = jcl.ell_matrix(H, Dx, Dy)
    B = (3 * 2.7 + 0.5) * 1.01
    A.bounds = np.array([-B, B])
    N = int(A.shape[0])
    Hd, Vd = velocity_handle(A, "x")
    r = np.stack([jcl.random_vector_generator(N, 700 + s) for s in range(R)])
    rr = _kpm.as_batch(r, N, "c64")
    ws = torch.empty(nat.kubo_workspace_bytes(Hd, R, M), dtype=torch.uint8, device="cuda")
    mu = torch.empty((M, M), dtype=torch.complex128, device="cuda")
    es = torch.as_tensor(np.linspace(-0.99, 0.99, n_e), device="cuda")

    def step():
        nat.kubo_mu_nm(Hd, Vd, rr, M, workspace=ws, out=mu)
        return nat.kubo_sigma(mu, es, 0, B)

    for _ in range(max(1, args.warmup)):  # tuning and allocation stay untimed
        step()
    torch.cuda.synchronize()
    ms = event_ms(torch, step, args.steps)
    gemm_ms, n_gemm = profiled(nat, step, every=16, kind=nat.JKPM_PROF_KUBO_GEMM)
    kstep_ms, n_kstep = profiled(nat, step, every=16, kind=nat.JKPM_PROF_KUBO_STEP)
    sigma = step().cpu().numpy()
    mu64 = mu.cpu().numpy()
    ws = None
    torch.cuda.empty_cache()
    log("Kubo: complex128 stacks contracted in FP64 for the error")
    ref = kubo_mu_nm_fp64(torch, nat, _kpm, A, B, jcl.modifier_velocity(A, "x"), r, M)
    sig_ref = nat.kubo_sigma(ref, es, 0, B).cpu().numpy()
    ref = ref.cpu().numpy()
    torch.cuda.empty_cache()
    mu_err = float(np.max(np.abs(mu64 - ref)) / np.max(np.abs(ref)))
    sig_err = float(np.max(np.abs(sigma.real - sig_ref.real)) / np.max(np.abs(sig_ref.real)))
    log(f"Kubo error vs FP64: mu_nm {mu_err:.2e}, sigma {sig_err:.2e}")
    flops = 8.0 * N * M * M * R
    tf = flops / (ms / 1e3) / 1e12
    tf_gemm = flops / (gemm_ms / 1e3) / 1e12
    out = base_line("Kubo double moments: mu_nm GEMM TFLOP/s (8 N M^2 R per step)",
                    tf, "TFLOP/s", args, ms, "c64 stacks, fp32 MFMA",
                    {"workload": "honeycomb 512x512 cells periodic, t = -2.7, W = 1 eV, "
                                 "v_x on H's slots, M = 512, R = 4, complex64 stacks, FP32 "
                                 "MFMA accumulate, FP64 combine, sigma on 512 energies",
                     "N": N, "M": M, "R": R})
    # per slot: H's value (4 bytes packed when real), v's value (4 when purely
    # imaginary), the column; per row: 7 complex64 vector streams
    h_b = 4 if Hd.is_real() else 8
    v_b = 4 if Vd.is_imag() else 8
    kstep_bytes = N * 4 * (h_b + v_b + 4) + R * N * 8 * 7
    kubo_pmc = PROFILES / "kubo_step_and_gemm_honeycomb512_pmc.md"
    out["roofline"] = {"bound": "mfma", "achieved": round(tf_gemm, 2), "peak": 157.3,
                       "unit": "TFLOP/s", "frac": round(tf_gemm / 157.3, 4),
                       "traffic": profile_traffic(kubo_pmc, "gemm_traffic_bytes_per_launch",
                                                  N=N, M=M, R=R),
                       "kernel": "kubo_gemm3_kernel", "mean_launch_us": round(gemm_ms * 1e3, 1),
                       "launches_timed": n_gemm,
                       "note": "complex flops 8 N M^2 R over the GEMM launch (three real "
                               "products per complex one execute 3/4 of them)",
                       "whole_step_TFLOPs": round(tf, 2),
                       "whole_step_frac": round(tf / 157.3, 4)}
    out["stack_step"] = hbm_roofline(kstep_bytes, kstep_ms, n_kstep, "kubo_step_kernel",
                                     steps_per_call=M - 1, h_value_bytes=h_b,
                                     v_value_bytes=v_b)
    out["stack_step"]["traffic"] = profile_traffic(kubo_pmc, "step_traffic_bytes_per_launch",
                                                   N=N, M=M, R=R)
    out["mu_nm_err_vs_fp64"] = mu_err
    out["sigma_err_vs_fp64"] = sig_err
    out["accuracy_note"] = ("max |mu_c64 - mu_fp64| / max |mu_fp64| and the same for the "
                            "Kubo-Greenwood sigma on 512 rescaled energies; the FP64 reference "
                            "contracts complex128 stacks of the same H, v_x and vectors")
    return out


def probe_lattices(torch, jcl, nat, _kpm, steps=9):
    """Counter probe (see --probe-lattices): the step kernel on four shapes,
    each with row chunks of exactly one lattice stride (2048 rows), in this
    order: P16, D16, H8, S8."""
    H, _, _ = jcl.honeycomb_lattice(1024, t=-2.7, W=1.0, seed=2024)
    A = jcl.ell_matrix(H)
    B = (3 * 2.7 + 0.5) * 1.01
    N = int(A.shape[0])
    h = A.device("c64", 1 / B)
    for R, dbl, tag in ((16, False, "P16"), (16, True, "D16"), (8, False, "H8")):
        r = np.stack([jcl.random_vector_generator(N, 1000 + s) for s in range(R)])
        vr = _kpm.as_batch(r, N, "c64")
        h.set_launch(38, 4, dbl)
        if dbl:
            nat.moments_ex(h, None, vr, 2 * steps, True)
        else:
            h.moments(vr, steps + 1)
        torch.cuda.synchronize()
        log(f"{tag}: N={N} R={R} {'doubling' if dbl else 'plain'} done")
    H, _ = jcl.square_lattice(2048, t=-1.0, W=1.0, seed=7)
    A = jcl.ell_matrix(H)
    N = int(A.shape[0])
    h = A.device("c64", 1 / (4.5 * 1.01))
    r = np.stack([jcl.random_vector_generator(N, 2000 + s) for s in range(8)])
    vr = _kpm.as_batch(r, N, "c64")
    h.set_launch(38, 8, False)
    h.moments(vr, steps + 1)
    torch.cuda.synchronize()
    log(f"S8: N={N} R=8 plain done")
    print(json.dumps({"probe": "lattices", "shapes": ["P16", "D16", "H8", "S8"],
                      "step_launches_each": steps}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="dos",
                    choices=["dos", "config1", "batch", "grad", "kubo"],
                    help="dos: the headline precision-form DOS line; config1: the plain "
                         "complex128 DOS with the CPU oracle on the same inputs; batch / grad "
                         "/ kubo: the realisation-batch, differentiation and Kubo workloads")
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cells", type=int, default=1024)
    ap.add_argument("--moments", type=int, default=2048)
    ap.add_argument("--vectors", type=int, default=16)
    ap.add_argument("--energies", type=int, default=4096)
    ap.add_argument("--dtype", default="c64")
    ap.add_argument("--profile-every", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-dos-check", action="store_true")
    ap.add_argument("--no-doubling", action="store_true")
    ap.add_argument("--no-complex-left", action="store_true",
                    help="skip the plain recursion with v_l read as complex numbers")
    ap.add_argument("--no-tiles", action="store_true",
                    help="skip the two-steps-per-launch (LDS patch) path")
    ap.add_argument("--no-wave", action="store_true",
                    help="skip the two-steps-per-launch wavefront path")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the plain complex128 one-vector GPU recursion")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default launch shape instead of the tuned one")
    ap.add_argument("--no-subs", action="store_true",
                    help="default workload only: skip the config1 / batch / grad / kubo "
                         "sub-objects")
    ap.add_argument("--sub-steps", type=int, default=3,
                    help="timed steps of each sub-workload (at most --steps)")
    ap.add_argument("--fd-coords", type=int, default=16,
                    help="grad workload: coordinates of the finite-difference check")
    ap.add_argument("--probe-lattices", action="store_true",
                    help="short run for hardware counters: 8 step launches (after a first "
                         "step) of four shapes whose row chunks are one lattice stride: "
                         "honeycomb 1024x1024 cells R=16 plain and doubling and R=8 "
                         "(variant 38, 4 work-groups per CU), square 2048x2048 R=8 "
                         "(variant 38, 8 work-groups per CU)")
    ap.add_argument("--probe", action="store_true",
                    help="short run for hardware counters: 3 copies of 1 GiB with the "
                         "library's copy kernel, then 16 moments of the workload (15 "
                         "steps) and 32 moments in the doubling form (16 steps)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the KPM path "
              "shards over vectors and disorder realisations with one final all-reduce "
              "of the moments, which this package does not run", file=sys.stderr)
        sys.exit(2)

    import torch

    import minijclsquant_amd as jcl
    from minijclsquant_amd import _kpm, _native as nat

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    torch.cuda.set_device(0)
    if args.probe_lattices:
        probe_lattices(torch, jcl, nat, _kpm)
        return
    subs = {"config1": bench_config1, "batch": bench_batch, "grad": bench_grad,
            "kubo": bench_kubo}
    if args.workload != "dos":
        print(json.dumps(subs[args.workload](args, torch, jcl, nat, _kpm)), flush=True)
        return
    L, M, R = args.cells, args.moments, args.vectors
    log(f"building workload: honeycomb {L}x{L} cells, M={M}, R={R}, {args.dtype}")
    A, r, B = build_workload(L, M, R, seed=2024)
    N, K = int(A.shape[0]), int(A.len_row)
    h = A.device(args.dtype, 1.0 / B)
    vr = _kpm.as_batch(r, N, args.dtype)
    ws = torch.empty(h.workspace_bytes(R, 1), dtype=torch.uint8, device="cuda")
    mu = torch.empty((1, M, R), dtype=torch.complex128, device="cuda")
    E = np.linspace(-B + 1e-3, B - 1e-3, args.energies)
    es = torch.as_tensor(E / B, device="cuda")

    def one_step(doubling=False):
        if doubling:
            nat.moments_ex(h, None, vr, M, True, workspace=ws_dbl, out=mu[0])
        else:
            h.moments(vr, M, workspace=ws, out=mu)
        return nat.dos(mu[0], es, nat.JKPM_KERNEL_JACKSON)

    if args.probe:
        # launch shapes from the tuning table (tuned here only when absent)
        launch = {}
        if not args.no_tune:
            from minijclsquant_amd import autotune

            launch["plain"] = autotune.tuned(h, vr)
            launch["doubling"] = autotune.tuned(h, vr, doubling=True)
        nbytes = 1 << 30
        src = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            nat.copy(dst, src, nbytes)
        h.moments(vr, 16, workspace=ws)
        nat.moments_ex(h, None, vr, 32, True)
        torch.cuda.synchronize()
        print(json.dumps({"probe": "done", "copy_bytes": nbytes, "N": N, "K": K, "R": R,
                          "step_launches": 15, "doubling_step_launches": 16,
                          "launch": launch}), flush=True)
        return

    if not args.no_tune:
        from minijclsquant_amd import autotune

        v, m = autotune.tuned(h, vr)
        log(f"launch shape: variant {v}, {m} work-groups per CU")
        if not args.no_doubling:
            v, m = autotune.tuned(h, vr, doubling=True)
            log(f"doubling-form launch shape: variant {v}, {m} work-groups per CU")
        ws = torch.empty(h.workspace_bytes(R, 1), dtype=torch.uint8, device="cuda")
    ws_dbl = torch.empty(nat.moments_ex_workspace_bytes(h, None, R, M, True), dtype=torch.uint8,
                         device="cuda")
    for i in range(args.warmup):
        one_step()
        log(f"warmup step {i + 1}/{args.warmup} done")
    torch.cuda.synchronize()
    copy_gbs = copy_ceiling(torch, nat)
    log(f"copy kernel: {copy_gbs:.0f} GB/s (read + write)")

    def timed(doubling):
        nat.profile_read(reset=True)
        nat.profile_enable(args.profile_every)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            d = one_step(doubling)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        nat.profile_enable(0)
        ms_sum, n_prof = nat.profile_read(reset=True)
        return (t1 - t0) * 1e3 / args.steps, ms_sum / max(n_prof, 1), n_prof, d

    ms_per_step, launch_ms, n_prof, dos = timed(False)
    log(f"timed {args.steps} steps: {ms_per_step:.2f} ms per step")

    units = N * M * R
    value = units / (ms_per_step / 1e3)
    s_v = 8 if args.dtype == "c64" else 16
    mv = s_v // 2 if h.is_real() else s_v  # matrix value bytes per slot (packed real copy)
    # complex64: the random-phase left vectors (v_l = v_r) are read as 32-bit phases
    phases = args.dtype == "c64" and nat.phase_pack(vr)[1]
    s_l = 4 if phases else s_v
    # vectors per recursion pass (the tuner may run the R vectors in passes)
    g = h.get_pass() or R
    g = min(g, R)
    bytes_launch = N * K * (mv + 4) + g * N * (3 * s_v + s_l)
    achieved = bytes_launch / (launch_ms / 1e3) / 1e9
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic_from_profile(N, K, R, g=g),
        "kernel": "cheb_step_kernel",
        "algorithmic_bytes_per_launch": bytes_launch,
        "matrix_value_bytes_per_slot": mv,
        "left_vector_bytes_per_element": s_l,
        "vectors_per_pass": g,
        "mean_launch_us": round(launch_ms * 1e3, 2),
        "launches_timed": n_prof,
        "copy_kernel_GBs": round(copy_gbs, 1),
        "frac_of_copy": round(achieved / copy_gbs, 4),
    }
    if g < R:
        roofline.update(pass_note(N, K, R, g, mv, s_v, s_l, launch_ms))

    d128 = None
    dos_err = None
    mu_err = None
    if not args.no_dos_check and args.dtype == "c64":
        log("complex128 reference run for the moment and DOS errors")
        mu128 = jcl.kpm_moments(A, M, r, dtype="c128")  # [R, M]
        d128 = nat.dos(mu128.transpose(0, 1).contiguous(), es,
                       nat.JKPM_KERNEL_JACKSON).cpu().numpy()
        dos_err = float(np.max(np.abs(dos.cpu().numpy() - d128)) / np.max(np.abs(d128)))
        mu_err = float((mu[0].transpose(0, 1) - mu128).abs().max() / N)
        del mu128

    dbl = None
    if not args.no_doubling:
        one_step(True)
        torch.cuda.synchronize()
        ms_d, launch_d, n_d, dos_d = timed(True)
        log(f"doubling form: {ms_d:.2f} ms per step")
        gd = min(h.get_pass(doubling=True) or R, R)
        bytes_d = N * K * (mv + 4) + gd * N * 3 * s_v
        ach_d = bytes_d / (launch_d / 1e3) / 1e9
        dbl = {
            "value": units / (ms_d / 1e3),
            "ms_per_step": ms_d,
            "steps_per_recursion": max(1, M // 2),
            "note": ("same M moments from M/2 Chebyshev steps: mu_2k = 2<T_k|T_k> - mu_0, "
                     "mu_2k+1 = 2<T_k|T_k+1> - mu_1 (Hermitian H); no v_l stream"),
            "roofline": {"bound": "hbm", "achieved": round(ach_d, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(ach_d / HBM_PEAK_GBS, 4),
                         "traffic": traffic_from_profile(N, K, R,
                                                         "doubling_traffic_bytes_per_launch",
                                                         g=gd),
                         "algorithmic_bytes_per_launch": bytes_d,
                         "vectors_per_pass": gd,
                         "mean_launch_us": round(launch_d * 1e3, 2), "launches_timed": n_d},
            "dos_err_vs_c128": (None if d128 is None else
                                float(np.max(np.abs(dos_d.cpu().numpy() - d128))
                                      / np.max(np.abs(d128)))),
        }
        if gd < R:
            dbl["roofline"].update(pass_note(N, K, R, gd, mv, s_v, 0, launch_d))

    complex_left = None
    if phases and not args.no_complex_left:
        # the same plain recursion reading v_l as complex64 (its own tuned launch)
        h.use_left_phases(False)
        if not args.no_tune:
            autotune.tuned(h, vr)
        h.moments(vr, M, workspace=ws, out=mu)
        torch.cuda.synchronize()
        ms_c, launch_c, n_c, _ = timed(False)
        shape_c = list(h.get_launch())
        g_c = min(h.get_pass() or R, R)
        h.use_left_phases(True)
        if not args.no_tune:
            autotune.tuned(h, vr)
        bytes_c = N * K * (mv + 4) + g_c * N * 4 * s_v
        ach_c = bytes_c / (launch_c / 1e3) / 1e9
        complex_left = {
            "value": units / (ms_c / 1e3), "ms_per_step": ms_c,
            "launch": shape_c,
            "roofline": {"bound": "hbm", "achieved": round(ach_c, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(ach_c / HBM_PEAK_GBS, 4),
                         "algorithmic_bytes_per_launch": bytes_c, "vectors_per_pass": g_c,
                         "mean_launch_us": round(launch_c * 1e3, 2), "launches_timed": n_c},
            "headline_over_it": value / (units / (ms_c / 1e3)),
            "note": "the plain recursion with v_l read as complex64 pairs (8 bytes per "
                    "element) instead of 32-bit phases",
        }
        log(f"complex left vectors: {ms_c:.2f} ms per step")

    tiles = None
    if not args.no_tiles:
        tiles = tiled_leg(torch, nat, h, vr, N, K, M, R, s_v, args, units, es)
        log(f"two steps per launch: {tiles['ms_per_step']:.2f} ms per step, doubling "
            f"{tiles['doubling_ms_per_step']:.2f} ms")

    wave = None
    if not args.no_wave:
        wave = wave_leg(torch, nat, h, vr, N, K, M, R, s_v, args, units, es)
        if "ms_per_step" in wave:
            log(f"wavefront two steps per launch: {wave['ms_per_step']:.2f} ms per step")

    plain = None
    if not args.no_plain:
        plain = plain_c128(torch, nat, _kpm, A, r, B, M, units / (ms_per_step / 1e3),
                           None if dbl is None else dbl["value"], not args.no_tune)
        log(f"plain complex128 R=1 recursion: {plain['ms_per_recursion']:.1f} ms per "
            f"{M}-moment recursion")

    cpu = None
    if not args.no_cpu:
        log("CPU baseline (oracle, complex128, bounded sample)")
        cpu = cpu_baseline(A, r)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "SpMV-steps/s (N*M*R per second)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic (seeded lattice disorder and random-phase vectors)",
        "config": {
            "workload": (f"KPM DOS precision form: honeycomb {L}x{L} cells periodic, "
                         f"t=-2.7, W=1 eV, K={K}, M={M}, R={R}, "
                         f"{args.dtype} recursion + FP64 moments, DOS on "
                         f"{args.energies} energies"),
            "N": N, "K": K, "M": M, "R": R, "bound": B,
        },
        "dos_err_vs_c128": dos_err,
        "mu_err_vs_c128": mu_err,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "doubling": dbl,
        "complex_left": complex_left,
        "plain_c128": plain,
        "two_step_tiles": tiles,
        "two_step_wave": wave,
    }
    if not args.no_subs:
        # the other workloads, each with its own metric, dominant-kernel
        # roofline and accuracy field; a few timed steps each
        sub_args = argparse.Namespace(**vars(args))
        sub_args.steps = max(1, min(args.steps, args.sub_steps))
        sub_args.warmup = 1
        del h, vr, ws, ws_dbl, mu
        for name, fn in subs.items():
            torch.cuda.empty_cache()
            log(f"sub-workload {name}")
            t0 = time.perf_counter()
            sub = fn(sub_args, torch, jcl, nat, _kpm)
            for k in ("n_gpus", "higher_is_better", "scaling", "vs_baseline", "data"):
                sub.pop(k, None)
            sub["wall_s"] = round(time.perf_counter() - t0, 1)
            out[name] = sub
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
