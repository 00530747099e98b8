#!/usr/bin/env python3
"""EVD wall time of the gfx950 two-stage symmetric eigensolver on one MI355X.

Default workload: n = 16384, fp64, A = Q diag(lam) Q^T (Q: Householder-QR factor of
a seeded Gaussian matrix, lam sorted iid U(-1,1)), bandwidth b = 64, block k = 512,
all eigenvalues and eigenvectors. One step = one full cuev_dsymm_eig_solve call
(DBBR -> bulge chasing -> device D&C beside the GPU build of Q_b^T Q_s^T -> final
GEMM), input already resident in HBM; the copy of A that the solver consumes is
made before each step's timer starts.

--only chase|bcback|copy runs one kernel of interest and nothing else (for hardware
counter runs): one bulge chase of a random packed band of the configured n and b, the
bc_back application of that chase's Q_b to the identity, or a plain 1 GiB device copy
(the calibration of the FETCH_SIZE / WRITE_SIZE counters).

Prints ONE JSON line: value = mean EVD wall seconds per solve (lower is better),
per-stage times, correctness against the exact eigenvalues, the roofline of the
dominant GPU kernel measured with HIP events inside the library, and host LAPACK
dsyevd as the CPU baseline (default: on the very n = 16384 matrix the GPU solved,
copied to the host). With --mixed the FP32 + FP64 refinement form is timed on the
same matrix as well.

The default q16k run adds the metric's other two numbers as extra keys (skip them
with --no-extras), and "values_only", the eigenvalues-only path (band reduction, chase
and Sturm multisection) timed on the same matrix: "g49k", one timed n = 49152 solve
after one untimed one, with its
accuracy checked over column blocks of V; and "mixed", the mixed16k matrix (n = 16384,
geometric clustered spectrum) solved by the FP64 path and by the mixed FP32 +
refinement path, with both times, the speedup and the accuracy of each.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_MFMA = 78.6e12   # MI355X dense FP64 matrix peak, FLOP/s
PEAK_FP32_MFMA = 157.3e12
PEAK_HBM = 8.0e12          # bytes/s
# single-launch sites whose HIP-event durations the timed steps record (the roofline's
# candidates for the dominant kernel)
ROOFLINE_SITES = ("bc.chase", "bt.bc_back", "bt.gemm")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _blas_threads():
    try:
        from threadpoolctl import threadpool_info
        return max([d.get("num_threads", 0) for d in threadpool_info()
                    if d.get("internal_api") == "openblas"] or [os.cpu_count()])
    except Exception:
        return os.cpu_count()


def cpu_baseline_same(A_host, lam_exact=None):
    """Host LAPACK dsyevd (jobz=V, uplo=L; SciPy's OpenBLAS, all its threads) on the very
    matrix the GPU solved, copied to the host; generation and the copy are excluded."""
    import numpy as np
    import scipy.linalg as sl
    n = A_host.shape[0]
    A = np.asfortranarray(A_host)
    t0 = time.perf_counter()
    w, V = sl.eigh(A, lower=True, driver="evd", overwrite_a=True, check_finite=False)
    t = time.perf_counter() - t0
    out = {"value": round(t, 3), "unit": "s", "cores": int(_blas_threads()), "kind": "port",
           "sample": f"LAPACK dsyevd (jobz=V, uplo=L, SciPy OpenBLAS) on the same n={n} fp64 "
                     f"matrix the GPU solved (copied to the host), measured once, not scaled"}
    if lam_exact is not None:
        le = np.sort(np.asarray(lam_exact))
        out["eig_err_vs_exact_over_norm2"] = float(np.abs(w - le).max() / np.abs(le).max())
    return out


def cpu_baseline(n_sample, n_full, seed=1):
    """Host LAPACK dsyevd (SciPy's OpenBLAS) on an n_sample matrix of the same construction,
    scaled by (n_full / n_sample)^3 (used where the full size would take too long)."""
    import numpy as np
    import scipy.linalg as sl
    threads = _blas_threads()
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n_sample, n_sample)))
    lam = np.sort(rng.uniform(-1.0, 1.0, n_sample))
    A = (Q * lam) @ Q.T
    A = (A + A.T) / 2
    del Q
    t0 = time.perf_counter()
    w, V = sl.eigh(A, driver="evd", overwrite_a=True, check_finite=False)
    t = time.perf_counter() - t0
    scaled = t * (n_full / n_sample) ** 3
    return {"value": round(scaled, 3), "unit": "s", "cores": int(threads), "kind": "port",
            "sample": f"LAPACK dsyevd (jobz=V, SciPy OpenBLAS) on one n={n_sample} fp64 "
                      f"Q diag(lam) Q^T matrix: {t:.2f} s measured, scaled by (n/{n_sample})^3 "
                      f"to n={n_full}"}


def measured_traffic(tag, n, b):
    """HBM bytes per launch of `tag` from the committed counter summary (FETCH_SIZE and
    WRITE_SIZE, copy-calibrated), when it was measured for this n and b; else None."""
    if tag != "bc.chase":
        return None
    path = os.path.join(ROOT, "profiles", "bc_chase_q16k_pmc.txt")
    try:
        text = open(path).read()
    except OSError:
        return None
    if f"n = {n}, b = {b}" not in text:
        return None
    for line in text.splitlines():
        if line.startswith("# traffic_bytes_per_launch"):
            return float(line.split()[-1])
    return None


def residual_orth(torch, A0, V, lam, cols=2048):
    from cuev_amd import workloads
    return workloads.residual_orth(A0, V, lam, cols)


def run_only(which, n, b, k):
    """One dispatch of interest (after one untimed repeat), for counter collection."""
    import torch
    import cuev_amd as cv
    if which == "copy":
        src = torch.randn(1 << 27, dtype=torch.float64, device="cuda")  # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(2):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"only": "copy", "bytes_read": src.numel() * 8, "bytes_written": dst.numel() * 8}))
        return
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    band = torch.randn((n, 2 * b), generator=g, dtype=torch.float64, device="cuda")  # packed, column-major
    with cv.Handle(torch.float64, n, b, k) as h:
        d = torch.empty(n, dtype=torch.float64, device="cuda")
        e = torch.empty(n, dtype=torch.float64, device="cuda")
        import ctypes
        fn = cv.lib().cuev_d_bc_chase
        for _ in range(2):
            B = band.clone()
            torch.cuda.synchronize()
            t = time.perf_counter()
            st = fn(h.h, ctypes.c_void_p(B.data_ptr()), ctypes.c_void_p(d.data_ptr()),
                    ctypes.c_void_p(e.data_ptr()))
            torch.cuda.synchronize()
            dt = time.perf_counter() - t
            assert st == 0, cv.lib().cuev_last_error_message()
        if which == "bcback":
            # the chase has left its reflectors in the handle's U: apply Q_b to I
            del band, B
            M = torch.eye(n, dtype=torch.float64, device="cuda")
            res = {"only": "bcback", "n": n, "b": b}
            for name, f in (("forward", h.bc_back), ("transposed", h.bc_back_t)):
                for _ in range(2):
                    torch.cuda.synchronize()
                    t = time.perf_counter()
                    f(M, n)
                    torch.cuda.synchronize()
                    dt = time.perf_counter() - t
                res[name] = {"seconds": dt, "tflops": 2.0 * n ** 3 / dt / 1e12}
            print(json.dumps(res))
            return
    hops = n * n / (2.0 * b)
    print(json.dumps({"only": "chase", "n": n, "b": b, "seconds": dt,
                      "algorithmic_bytes": hops * 4 * b * b * 8}))


def run_diff(n, steps, warmup):
    """Forward and backward eigh on the differentiation workload (A = Q diag(lam) Q^T, lam
    equispaced in [-1, 1]) in fp64 and fp32. Cotangents of L = sum_i w_i lam_i +
    ||V_k^T B||_F^2 (k = 128, B n x 64): lambar = w, Vbar[:, :k] = 2 B B^T V_k. Backward
    work = three n^3 products (V^T Vbar, V K, (V K) V^T) = 6 n^3 flops."""
    import torch
    import cuev_amd as cv
    from cuev_amd import workloads
    A0, lam_exact = workloads.make("diff4k", n=n, device="cuda")
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    w = torch.randn(n, generator=g, dtype=torch.float64, device="cuda")
    B = torch.randn((n, 64), generator=g, dtype=torch.float64, device="cuda")
    k = 128
    out = {"metric": f"eigh forward + backward wall ms at n={n}", "n": n, "k": k}
    for name, dt in (("f64", torch.float64), ("f32", torch.float32)):
        A = A0.to(dt)
        fw, bw = [], []
        for i in range(warmup + steps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            lam, V = cv.symm_eig_solve(A)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            Vk = V[:, :k]
            Vbar = torch.zeros_like(V)
            Vbar[:, :k] = 2.0 * (B.to(dt) @ (B.to(dt).T @ Vk))
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            Abar, masked = cv.eig_backward(V, lam, Vbar, w.to(dt), gap_tol=1e-6 if dt == torch.float64 else 1e-5)
            torch.cuda.synchronize()
            t3 = time.perf_counter()
            if i >= warmup:
                fw.append(t1 - t0)
                bw.append(t3 - t2)
        fwd = sum(fw) / len(fw)
        bwd = sum(bw) / len(bw)
        asym = (Abar - Abar.T).abs().max().item() / Abar.abs().max().item()
        err = (lam.double() - lam_exact).abs().max().item()
        out[name] = {"forward_ms": round(fwd * 1e3, 3), "backward_ms": round(bwd * 1e3, 3),
                     "backward_tflops": round(6.0 * n ** 3 / bwd / 1e12, 2), "masked_pairs": masked,
                     "abar_asymmetry": asym, "eig_err_vs_exact": err}
        log(f"[bench] diff {name}: {out[name]}")
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="q16k", choices=["q16k", "n1024", "g49k", "mixed16k", "diff4k"],
                    help="diff4k: forward + backward eigh at n=4096 in fp64 and fp32 (own JSON line)")
    ap.add_argument("--n", type=int, default=0)
    ap.add_argument("--b", type=int, default=64)
    ap.add_argument("--k", type=int, default=512)
    ap.add_argument("--bt", default="auto", choices=["auto", "direct", "reordered"])
    ap.add_argument("--dc", default="auto", choices=["auto", "host", "device"],
                    help="tridiagonal D&C: host LAPACK dstedc or the device-resident form")
    ap.add_argument("--mixed", action="store_true", help="also time the mixed FP32+refinement form")
    ap.add_argument("--cpu", default="same", choices=["same", "sample", "none"],
                    help="CPU baseline: dsyevd on the same matrix (default), on an n=--cpu-n sample "
                         "scaled by n^3, or none")
    ap.add_argument("--cpu-n", type=int, default=6144, help="n of the dsyevd CPU sample for --cpu sample")
    ap.add_argument("--no-extras", action="store_true",
                    help="q16k: skip the n=49152 solve and the mixed16k FP64/mixed pair")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--only", choices=["chase", "bcback", "copy"], default=None,
                    help="run one kernel of interest and exit (hardware counter runs); bcback "
                         "applies the Q_b of one chase (and its transpose) to the identity, alone on the GPU")
    a = ap.parse_args()

    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}): the single-GPU "
              "solver has no multi-GPU path; see DESIGN.md for how it would shard.", file=sys.stderr)
        sys.exit(2)

    import torch
    import cuev_amd as cv
    from cuev_amd import workloads

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    n_default = {"q16k": 16384, "n1024": 1024, "g49k": 49152, "mixed16k": 16384, "diff4k": 4096}[a.workload]
    n = a.n or n_default
    b, k = a.b, a.k
    if a.only:
        run_only(a.only, n, b, k)
        return
    if a.workload == "diff4k":
        run_diff(a.n or 4096, a.steps, a.warmup)
        return
    if a.workload == "n1024" and not a.n:
        b, k = 32, 256
    log(f"[bench] generating workload {a.workload} n={n} on the GPU")
    t0 = time.perf_counter()
    A0, lam_exact = workloads.make(a.workload, n=n, device="cuda")
    torch.cuda.synchronize()
    log(f"[bench] generated in {time.perf_counter() - t0:.1f} s; lapack={cv.lapack_name()}")

    # column-major buffer (strides (1, n)) consumed in place by the solver; A0 is exactly
    # symmetric, so filling it through its transposed view copies A0 itself
    work = torch.empty_like(A0)

    def one(precision="native", timer=True):
        work.mT.copy_(A0)
        torch.cuda.synchronize()
        t = time.perf_counter()
        lam, V, tm = cv.symm_eig_solve(work.mT, overwrite=True, bandwidth=b, block=k, bt_mode=a.bt,
                                       dc_mode=a.dc, precision=precision, timer=True)
        torch.cuda.synchronize()
        return time.perf_counter() - t, lam, V, tm

    for i in range(a.warmup):
        dt, _, _, _ = one()
        log(f"[bench] warmup {i}: {dt:.3f} s")

    # the timed steps record events only around the single-launch sites the roofline
    # reads: every recorded site puts two event records on the stream, which hold
    # back-to-back kernels apart by about 10 us each (about 1% of a solve if every
    # site were recorded)
    cv.profile_enable(True, only=ROOFLINE_SITES)
    cv.profile_read(reset=True)
    times, timers = [], []
    for i in range(a.steps):
        dt, lam, V, tm = one()
        times.append(dt)
        timers.append(tm)
        log(f"[bench] step {i}: {dt:.3f} s  dbbr={tm['dbbr_ms']:.1f} bc={tm['bc_ms']:.1f} "
            f"dc={tm['dc_ms']:.1f} bt_after_dc={tm['bt_ms']:.1f} (overlapped {tm['overlap_ms']:.1f}) ms")
    prof = cv.profile_read(reset=True)
    K = len(times)
    mean_s = sum(times) / K
    # every launch site, over one more (untimed) solve, for the per-kernel table
    cv.profile_enable(True)
    cv.profile_read(reset=True)
    one()
    kprof = cv.profile_read(reset=True)
    cv.profile_enable(False)

    # ---- correctness of the last solve
    check = {}
    if not a.no_check:
        lam64 = lam.double()
        res, orth = residual_orth(torch, A0, V, lam64)
        a2 = lam64.abs().max().item()
        check = {"residual": res, "orthogonality": orth,
                 "residual_bench_solver_form": res * n}
        if lam_exact is not None:
            check["eig_err_vs_exact_over_norm2"] = (lam64 - lam_exact).abs().max().item() / a2
        check["pass_1e-13"] = bool(res <= 1e-13 and orth <= 1e-13 and res * n <= 1e-13)
        log(f"[bench] residual {res:.2e} orth {orth:.2e} {check}")

    # ---- per-kernel rates (library HIP events over the timed steps)
    kernels = {}
    for tag, r in kprof.items():
        per_launch_ms = r["ms"] / max(1, r["count"])
        ent = {"ms_per_step": r["ms"], "launches_per_step": r["count"],
               "avg_launch_ms": per_launch_ms}
        if r["flops"] > 0 and r["ms"] > 0:
            ent["tflops"] = r["flops"] / (r["ms"] * 1e-3) / 1e12
        if r["bytes"] > 0 and r["ms"] > 0:
            ent["gbps"] = r["bytes"] / (r["ms"] * 1e-3) / 1e9
        kernels[tag] = ent

    roofline = dominant_roofline(prof, n, b)

    # ---- stages (means over the timed steps)
    stages = {key: sum(t[key] for t in timers) / K for key in
              ("dbbr_ms", "bc_ms", "dc_ms", "bt_ms", "total_ms", "bt_copyin_ms", "bt_qb_ms",
               "bt_qs_ms", "bt_copyout_ms", "overlap_ms")}
    flops16_3 = 16.0 / 3.0 * n ** 3
    dbbr_tf = (4.0 / 3.0 * n ** 3) / (stages["dbbr_ms"] * 1e-3) / 1e12 if stages["dbbr_ms"] else None

    # eigenvalues only (band reduction, chase, Sturm multisection; no back-transform, no
    # D&C) on the same matrix: one untimed and one timed call
    values = None
    if not a.no_extras:
        lam_v = None
        for i in range(2):
            work.mT.copy_(A0)
            torch.cuda.synchronize()
            t = time.perf_counter()
            lam_v, tmv = cv.symm_eig_values(work.mT, overwrite=True, bandwidth=b, block=k, timer=True)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t
        a2 = lam_v.double().abs().max().item()
        values = {"seconds": round(dt, 4), "speedup_vs_eigenpairs": round(mean_s / dt, 3),
                  "multisection_ms": round(tmv["dc_ms"], 2)}
        if lam_exact is not None:
            values["eig_err_vs_exact_over_norm2"] = (lam_v.double() - lam_exact).abs().max().item() / a2
        log(f"[bench] eigenvalues only: {dt:.3f} s {values}")
        del lam_v

    mixed = None
    if a.mixed:
        try:
            mixed = time_mixed(torch, cv, one, A0, lam_exact, a.warmup, a.steps, mean_s)
        except cv.CuevError as e:
            mixed = {"error": str(e)}

    del lam, V
    A_host = None
    if a.cpu == "same":
        A_host = A0.cpu().numpy()
    del work, A0
    cv.cache_release()
    torch.cuda.empty_cache()

    extras_on = a.workload == "q16k" and not a.no_extras and n == 16384
    g49k = mixed16k = None
    if extras_on:
        g49k = run_g49k(torch, cv, workloads)
        mixed16k = run_mixed16k(torch, cv, workloads)
        if mixed is None:
            mixed = mixed16k

    cpu = None
    if a.cpu == "same":
        log(f"[bench] CPU baseline: dsyevd on the same n={n} matrix")
        cpu = cpu_baseline_same(A_host, lam_exact.cpu().numpy() if lam_exact is not None else None)
        del A_host
        log(f"[bench] CPU baseline {cpu}")
    elif a.cpu == "sample" and a.cpu_n:
        log(f"[bench] CPU baseline: dsyevd n={a.cpu_n}")
        cpu = cpu_baseline(min(a.cpu_n, n), n)
        log(f"[bench] CPU baseline {cpu}")

    out = {
        "metric": f"EVD wall s at n={n} fp64 (two-stage, all eigenpairs)",
        "value": round(mean_s, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": K,
        "warmup": a.warmup,
        "ms_per_step": round(mean_s * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{a.workload}: n={n}, b={b}, k={k}, A=Q diag(lam) Q^T" if a.workload == "q16k"
                   else f"{a.workload}: n={n}, b={b}, k={k}", "n": n, "b": b, "k": k,
                   "bt_mode": a.bt, "dc_mode": a.dc, "parallelism": "single GPU"},
        "stages_ms": {kk: round(v, 2) for kk, v in stages.items()},
        "effective_tflops_16_3": round(flops16_3 / mean_s / 1e12, 3),
        "effective_tflops_4_3": round(flops16_3 / 4.0 / mean_s / 1e12, 3),
        "dbbr_tflops": round(dbbr_tf, 3) if dbbr_tf else None,
        "dbbr_frac_fp64_peak": round(dbbr_tf / 78.6, 4) if dbbr_tf else None,
        "check": check,
        "kernels": {t: {kk: (round(v, 4) if isinstance(v, float) else v) for kk, v in e.items()}
                    for t, e in kernels.items()},
        "kernels_source": "library HIP events around every launch site over one untimed solve "
                          "after the timed steps (the timed steps record only " +
                          ", ".join(ROOFLINE_SITES) + ")",
        "roofline": roofline,
        "cpu_baseline": cpu,
        "values_only": values,
        "g49k": g49k,
        "mixed": mixed,
        "lapack": cv.lapack_name(),
    }
    print(json.dumps(out), flush=True)


def dominant_roofline(prof, n, b):
    """Roofline of the single-launch kernel site with the most GPU time per solve.

    bc_chase (HBM/L2-bound): achieved = the algorithmic bytes of the model this build's
    library records for it (per hop b columns loaded, b stored, each b x 2b: 4 b^2 S;
    n^2 / (2b) hops), over its HIP-event duration; "frac_model_5p5bn2S" gives the same
    duration against the 5.5 b n^2 S L2-traffic model of a kernel that re-reads its
    window every hop. MFMA-bound kernels: the flops the library records per launch."""
    single = {"bc.chase": "hbm", "bt.bc_back": "mfma", "bt.gemm": "mfma"}
    dom = max((t for t in single if t in prof), key=lambda t: prof[t]["ms"], default=None)
    if not dom:
        return None
    r = prof[dom]
    dur = r["ms"] / r["count"] * 1e-3
    if single[dom] == "hbm":
        per_launch = r["bytes"] / r["count"]
        ach = per_launch / dur / 1e9
        alt = 5.5 * b * n * n * 8 / dur / 1e9
        return {"kernel": dom, "bound": "hbm", "achieved": round(ach, 2), "peak": 8000.0,
                "unit": "GB/s", "frac": round(ach / 8000.0, 4),
                "traffic": measured_traffic(dom, n, b),
                "traffic_source": "committed counter run profiles/bc_chase_q16k_pmc.txt "
                                  "(FETCH_SIZE + WRITE_SIZE, copy-calibrated), not measured by this run",
                "frac_model_5p5bn2S": round(alt / 8000.0, 4),
                "avg_launch_ms": round(dur * 1e3, 3)}
    per_launch = r["flops"] / r["count"]
    ach = per_launch / dur / 1e12
    return {"kernel": dom, "bound": "mfma", "achieved": round(ach, 3), "peak": 78.6,
            "unit": "TFLOP/s", "frac": round(ach / 78.6, 4), "traffic": None,
            "avg_launch_ms": round(dur * 1e3, 3)}


def time_mixed(torch, cv, one, A0, lam_exact, warmup, steps, fp64_s):
    """The mixed FP32 + refinement form on the matrix `one` solves."""
    for i in range(warmup):
        one("mixed")
    mt = []
    for i in range(steps):
        dt, lam_m, V_m, tm_m = one("mixed")
        mt.append(dt)
    # per-site event times over one more (untimed) mixed solve
    cv.profile_enable(True)
    cv.profile_read(reset=True)
    one("mixed")
    mprof = cv.profile_read(reset=True)
    cv.profile_enable(False)
    res_m, orth_m = residual_orth(torch, A0, V_m, lam_m.double())
    n = A0.shape[0]
    ms = sum(mt) / len(mt)
    mixed = {"seconds": round(ms, 4), "fp64_seconds": round(fp64_s, 4),
             "speedup_vs_fp64": round(fp64_s / ms, 4),
             "residual": res_m, "residual_bench_solver_form": res_m * n, "orthogonality": orth_m,
             "refine_sweeps": tm_m["refine_sweeps"], "n_cluster_cols": tm_m["n_cluster_cols"],
             "stages_ms": {kk: round(float(v), 2) for kk, v in tm_m.items() if kk.endswith("_ms")},
             # library events per mixed solve (the Rayleigh-Ritz step's inner FP64
             # solve shows under the FP64 stage names)
             "kernels_ms": {t: round(r["ms"], 2) for t, r in mprof.items()}}
    if lam_exact is not None:
        mixed["eig_err_vs_exact_over_norm2"] = \
            (lam_m.double() - lam_exact).abs().max().item() / lam_exact.abs().max().item()
    log(f"[bench] mixed {mixed}")
    return mixed


def _solve_timed(torch, cv, work, A0, **kw):
    work.mT.copy_(A0)
    torch.cuda.synchronize()
    t = time.perf_counter()
    lam, V, tm = cv.symm_eig_solve(work.mT, overwrite=True, timer=True, **kw)
    torch.cuda.synchronize()
    return time.perf_counter() - t, lam, V, tm


def run_g49k(torch, cv, workloads):
    """One untimed and one timed fp64 solve of g49k (n = 49152, A = (G + G^T)/2, G iid
    N(0, 1/n), seed 2, b = 64, k = 512), accuracy over column blocks of V."""
    n = 49152
    log(f"[bench] g49k: generating n={n}")
    A0, _ = workloads.make("g49k", n=n, device="cuda")
    work = torch.empty_like(A0)
    dt0, lam, V, _ = _solve_timed(torch, cv, work, A0, bandwidth=64, block=512)
    log(f"[bench] g49k untimed solve: {dt0:.2f} s")
    del lam, V
    dt, lam, V, tm = _solve_timed(torch, cv, work, A0, bandwidth=64, block=512)
    log(f"[bench] g49k timed solve: {dt:.2f} s {tm}")
    del work
    res, orth = residual_orth(torch, A0, V, lam.double())
    out = {"workload": "g49k: n=49152, b=64, k=512, A=(G+G^T)/2, G iid N(0,1/n), seed 2",
           "seconds": round(dt, 3), "untimed_first_solve_s": round(dt0, 3),
           "effective_tflops_16_3": round(16.0 / 3.0 * n ** 3 / dt / 1e12, 3),
           "dbbr_frac_fp64_peak": round(4.0 / 3.0 * n ** 3 / (tm["dbbr_ms"] * 1e-3) / 78.6e12, 4),
           "stages_ms": {kk: round(float(v), 2) for kk, v in tm.items() if kk.endswith("_ms")},
           "residual": res, "residual_bench_solver_form": res * n, "orthogonality": orth,
           "pass_1e-13": bool(res * n <= 1e-13 and orth <= 1e-13)}
    log(f"[bench] g49k {out}")
    del A0, V, lam
    cv.cache_release()
    torch.cuda.empty_cache()
    return out


def run_mixed16k(torch, cv, workloads, steps=2):
    """mixed16k (n = 16384, lam geometric 1..1e-6 with random signs, 10% in a 1e-9
    cluster): the FP64 path and the mixed path, each one untimed + `steps` timed solves."""
    n = 16384
    log(f"[bench] mixed16k: generating n={n}")
    A0, lam_exact = workloads.make("mixed16k", n=n, device="cuda")
    work = torch.empty_like(A0)

    def one(precision="native"):
        return _solve_timed(torch, cv, work, A0, bandwidth=64, block=512, precision=precision)

    one()
    ft = []
    for i in range(steps):
        dt, lam, V, _ = one()
        ft.append(dt)
    fp64_s = sum(ft) / len(ft)
    res, orth = residual_orth(torch, A0, V, lam.double())
    fp64_err = (lam.double() - lam_exact).abs().max().item() / lam_exact.abs().max().item()
    del lam, V
    log(f"[bench] mixed16k fp64: {fp64_s:.3f} s residual {res:.2e} orth {orth:.2e} eig err {fp64_err:.2e}")
    try:
        out = time_mixed(torch, cv, one, A0, lam_exact, 1, steps, fp64_s)
    except cv.CuevError as e:
        out = {"error": str(e), "fp64_seconds": round(fp64_s, 4)}
    out["workload"] = "mixed16k: n=16384, b=64, k=512, lam geometric 1..1e-6, random signs, 10% in a 1e-9 cluster"
    out["fp64"] = {"seconds": round(fp64_s, 4), "residual": res, "residual_bench_solver_form": res * n,
                   "orthogonality": orth, "eig_err_vs_exact_over_norm2": fp64_err}
    del work, A0
    cv.cache_release()
    torch.cuda.empty_cache()
    return out


if __name__ == "__main__":
    main()
