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
                stream, every `--profile-every`-th launch of the timed steps)
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
                doubling form's ratios over it.
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
TRAFFIC_FILE = ROOT / "profiles" / "cheb_step_c64_r16_honeycomb1024_pmc.md"


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


def traffic_from_profile(N, K, R):
    """HBM bytes per launch from the committed counter summary, if it was
    measured on this exact workload shape."""
    if not TRAFFIC_FILE.exists():
        return None
    txt = TRAFFIC_FILE.read_text()
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", txt)
    w = re.search(r"workload:\s*N=(\d+)\s+K=(\d+)\s+R=(\d+)", txt)
    if not m or not w or tuple(map(int, w.groups())) != (N, K, R):
        return None
    return float(m.group(1))


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


def cpu_baseline(A, r, target_s=15.0):
    """complex128 oracle (reference algorithm) on host cores: one vector, a
    bounded number of moments sized to ~target_s seconds."""
    from oracle import oracle as O

    thr = int(os.environ.get("OMP_NUM_THREADS", min(os.cpu_count() or 1, 16)))
    Ht = A.data * (1 / A.bounds[1])
    N = int(A.shape[0])
    m_probe = 16
    t0 = time.perf_counter()
    O.rec_A_tab(m_probe, Ht, A.indices, A.len_row, r[0], n_threads=thr)
    dt = time.perf_counter() - t0
    m = int(max(m_probe, min(4096, m_probe * target_s / max(dt, 1e-3))))
    t0 = time.perf_counter()
    O.rec_A_tab(m, Ht, A.indices, A.len_row, r[0], n_threads=thr)
    dt = time.perf_counter() - t0
    return {
        "value": N * m / dt,
        "unit": "N*M*R/s",
        "cores": thr,
        "kind": "port",
        "sample": f"same H, R=1 vector, M={m} moments, complex128, {dt:.1f} s",
    }


def plain_c128(torch, nat, _kpm, A, r, B, M, value, value_dbl, reps=3):
    """The reference's structure on this GPU: complex128 recursion, one
    vector at a time (R = 1), one moment per step, default launch shape.
    Timed with HIP events on the current stream over `reps` recursions."""
    N = int(A.shape[0])
    h = A.device("c128", 1.0 / B)
    v1 = _kpm.as_batch(r[0], N, "c128")
    ws = torch.empty(h.workspace_bytes(1, 1), dtype=torch.uint8, device="cuda")
    mu = torch.empty((1, M, 1), dtype=torch.complex128, device="cuda")
    h.moments(v1, M, workspace=ws, out=mu)
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(reps):
        h.moments(v1, M, workspace=ws, out=mu)
    t1.record()
    t1.synchronize()
    ms = t0.elapsed_time(t1) / reps
    v = N * M / (ms / 1e3)
    return {
        "value": v,
        "unit": "N*M*R/s",
        "ms_per_recursion": ms,
        "sample": f"same H, R=1 vector, M={M} moments, complex128, {reps} recursions",
        "headline_over_plain": value / v,
        "doubling_over_plain": None if value_dbl is None else value_dbl / v,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
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
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the plain complex128 one-vector GPU recursion")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default launch shape instead of the tuned one")
    ap.add_argument("--probe", action="store_true",
                    help="short run for hardware counters: 3 copies of 1 GiB with the "
                         "library's copy kernel, then 16 moments of the workload")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the KPM path "
              "shards over vectors and disorder realisations with one final all-reduce "
              "of the moments, which this build does not run", file=sys.stderr)
        sys.exit(2)

    import torch

    import minijclsquant_amd as jcl
    from minijclsquant_amd import _kpm, _native as nat

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    torch.cuda.set_device(0)
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
        nbytes = 1 << 30
        src = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            nat.copy(dst, src, nbytes)
        h.moments(vr, 16, workspace=ws)
        torch.cuda.synchronize()
        print(json.dumps({"probe": "done", "copy_bytes": nbytes, "N": N, "K": K, "R": R,
                          "step_launches": 15}), flush=True)
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
    bytes_launch = N * K * (s_v + 4) + R * N * 4 * s_v
    achieved = bytes_launch / (launch_ms / 1e3) / 1e9
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic_from_profile(N, K, R),
        "kernel": "cheb_step_kernel",
        "algorithmic_bytes_per_launch": bytes_launch,
        "mean_launch_us": round(launch_ms * 1e3, 2),
        "launches_timed": n_prof,
        "copy_kernel_GBs": round(copy_gbs, 1),
        "frac_of_copy": round(achieved / copy_gbs, 4),
    }

    d128 = None
    dos_err = None
    if not args.no_dos_check and args.dtype == "c64":
        log("complex128 reference run for the DOS error")
        d128 = jcl.kpm_dos(A, M, r, energies=E, dtype="c128")[:, 1]
        dos_err = float(np.max(np.abs(dos.cpu().numpy() - d128)) / np.max(np.abs(d128)))

    dbl = None
    if not args.no_doubling:
        one_step(True)
        torch.cuda.synchronize()
        ms_d, launch_d, n_d, dos_d = timed(True)
        log(f"doubling form: {ms_d:.2f} ms per step")
        bytes_d = N * K * (s_v + 4) + R * N * 3 * s_v
        ach_d = bytes_d / (launch_d / 1e3) / 1e9
        dbl = {
            "value": units / (ms_d / 1e3),
            "ms_per_step": ms_d,
            "steps_per_recursion": max(1, M // 2),
            "note": ("same M moments from M/2 Chebyshev steps: mu_2k = 2<T_k|T_k> - mu_0, "
                     "mu_2k+1 = 2<T_k|T_k+1> - mu_1 (Hermitian H); no v_l stream"),
            "roofline": {"bound": "hbm", "achieved": round(ach_d, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(ach_d / HBM_PEAK_GBS, 4),
                         "algorithmic_bytes_per_launch": bytes_d,
                         "mean_launch_us": round(launch_d * 1e3, 2), "launches_timed": n_d},
            "dos_err_vs_c128": (None if d128 is None else
                                float(np.max(np.abs(dos_d.cpu().numpy() - d128))
                                      / np.max(np.abs(d128)))),
        }

    plain = None
    if not args.no_plain:
        plain = plain_c128(torch, nat, _kpm, A, r, B, M, units / (ms_per_step / 1e3),
                           None if dbl is None else dbl["value"])
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
        "roofline": roofline,
        "cpu_baseline": cpu,
        "doubling": dbl,
        "plain_c128": plain,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
