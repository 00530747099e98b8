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
dsyevd on a bounded sample as the CPU baseline. With --mixed the FP32 + FP64
refinement form is timed on the same matrix as well.
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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(n_sample, n_full, seed=1):
    """Host LAPACK dsyevd (SciPy's OpenBLAS) on an n_sample matrix of the same construction."""
    import numpy as np
    import scipy.linalg as sl
    try:
        from threadpoolctl import threadpool_info
        threads = max([d.get("num_threads", 0) for d in threadpool_info()
                       if d.get("internal_api") == "openblas"] or [os.cpu_count()])
    except Exception:
        threads = os.cpu_count()
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
    """||A V - V diag(lam)||_F / (n ||A||_F) and ||V^T V - I||_F / n, accumulated over
    column blocks of V so that the check needs O(n * cols) extra memory (at n = 49152
    the solver's cached workspace and three n x n matrices already fill most of HBM)."""
    n = A0.shape[0]
    an = torch.linalg.matrix_norm(A0).item()
    r2 = o2 = 0.0
    for c0 in range(0, n, cols):
        c1 = min(n, c0 + cols)
        Vc = V[:, c0:c1]
        R = A0 @ Vc
        R.sub_(Vc * lam[c0:c1])
        r2 += torch.linalg.matrix_norm(R).item() ** 2
        G = Vc.T @ V
        G[:, c0:c1].diagonal().sub_(1.0)
        o2 += torch.linalg.matrix_norm(G).item() ** 2
        del R, G
    return r2 ** 0.5 / (n * an), o2 ** 0.5 / n


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
    ap.add_argument("--cpu-n", type=int, default=6144, help="n of the dsyevd CPU sample (0: skip)")
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

    work = torch.empty_like(A0)

    def one(precision="native", timer=True):
        work.copy_(A0)
        torch.cuda.synchronize()
        t = time.perf_counter()
        lam, V, tm = cv.symm_eig_solve(work, overwrite=True, bandwidth=b, block=k, bt_mode=a.bt, dc_mode=a.dc,
                                       precision=precision, timer=True)
        torch.cuda.synchronize()
        return time.perf_counter() - t, lam, V, tm

    for i in range(a.warmup):
        dt, _, _, _ = one()
        log(f"[bench] warmup {i}: {dt:.3f} s")

    cv.profile_enable(True)
    cv.profile_read(reset=True)
    times, timers = [], []
    for i in range(a.steps):
        dt, lam, V, tm = one()
        times.append(dt)
        timers.append(tm)
        log(f"[bench] step {i}: {dt:.3f} s  dbbr={tm['dbbr_ms']:.1f} bc={tm['bc_ms']:.1f} "
            f"dc={tm['dc_ms']:.1f} bt_after_dc={tm['bt_ms']:.1f} (overlapped {tm['overlap_ms']:.1f}) ms")
    prof = cv.profile_read(reset=True)
    cv.profile_enable(False)
    K = len(times)
    mean_s = sum(times) / K

    # ---- correctness of the last solve
    check = {}
    if not a.no_check:
        lam64 = lam.double()
        Vd = V.double()
        res, orth = residual_orth(torch, A0, Vd, lam64)
        a2 = lam64.abs().max().item()
        check = {"residual": res, "orthogonality": orth,
                 "residual_bench_solver_form": res * n}
        if lam_exact is not None:
            check["eig_err_vs_exact_over_norm2"] = (lam64 - lam_exact).abs().max().item() / a2
        check["pass_1e-13"] = bool(res <= 1e-13 and orth <= 1e-13)
        log(f"[bench] residual {res:.2e} orth {orth:.2e} {check}")

    # ---- per-kernel rates (library HIP events over the timed steps)
    kernels = {}
    for tag, r in prof.items():
        per_launch_ms = r["ms"] / max(1, r["count"])
        ent = {"ms_per_step": r["ms"] / K, "launches_per_step": r["count"] / K,
               "avg_launch_ms": per_launch_ms}
        if r["flops"] > 0 and r["ms"] > 0:
            ent["tflops"] = r["flops"] / (r["ms"] * 1e-3) / 1e12
        if r["bytes"] > 0 and r["ms"] > 0:
            ent["gbps"] = r["bytes"] / (r["ms"] * 1e-3) / 1e9
        kernels[tag] = ent

    # dominant single-kernel launch site by GPU time per step
    single = {"bc.chase": "hbm", "bt.bc_back": "mfma", "bt.gemm": "mfma", "dbbr.syr2k": "mfma",
              "dbbr.symm": "mfma"}
    dom = max((t for t in single if t in prof), key=lambda t: prof[t]["ms"], default=None)
    roofline = None
    if dom:
        r = prof[dom]
        dur = r["ms"] / r["count"] * 1e-3
        if single[dom] == "hbm":
            per_launch = r["bytes"] / r["count"]
            ach = per_launch / dur / 1e9
            roofline = {"kernel": dom, "bound": "hbm", "achieved": round(ach, 2), "peak": 8000.0,
                        "unit": "GB/s", "frac": round(ach / 8000.0, 4),
                        "traffic": measured_traffic(dom, n, b),
                        "avg_launch_ms": round(dur * 1e3, 3)}
        else:
            per_launch = r["flops"] / r["count"]
            ach = per_launch / dur / 1e12
            roofline = {"kernel": dom, "bound": "mfma", "achieved": round(ach, 3), "peak": 78.6,
                        "unit": "TFLOP/s", "frac": round(ach / 78.6, 4), "traffic": None,
                        "avg_launch_ms": round(dur * 1e3, 3)}

    # ---- stages (means over the timed steps)
    stages = {key: sum(t[key] for t in timers) / K for key in
              ("dbbr_ms", "bc_ms", "dc_ms", "bt_ms", "total_ms", "bt_copyin_ms", "bt_qb_ms",
               "bt_qs_ms", "bt_copyout_ms", "overlap_ms")}
    flops16_3 = 16.0 / 3.0 * n ** 3
    dbbr_tf = (4.0 / 3.0 * n ** 3) / (stages["dbbr_ms"] * 1e-3) / 1e12 if stages["dbbr_ms"] else None

    mixed = None
    if a.mixed:
        try:
            for i in range(a.warmup):
                one("mixed")
            mt = []
            cv.profile_enable(True)
            cv.profile_read(reset=True)
            for i in range(a.steps):
                dt, lam_m, V_m, tm_m = one("mixed")
                mt.append(dt)
            mprof = cv.profile_read(reset=True)
            cv.profile_enable(False)
            res_m, orth_m = residual_orth(torch, A0, V_m.double(), lam_m.double())
            mixed = {"seconds": sum(mt) / len(mt), "speedup_vs_fp64": mean_s / (sum(mt) / len(mt)),
                     "residual": res_m, "orthogonality": orth_m,
                     "refine_sweeps": tm_m["refine_sweeps"],
                     "stages_ms": {kk: round(float(v), 2) for kk, v in tm_m.items()
                                   if kk.endswith("_ms")},
                     # library events per mixed solve (the Rayleigh-Ritz step's inner FP64
                     # solve shows under the FP64 stage names)
                     "kernels_ms": {t: round(r["ms"] / len(mt), 2) for t, r in mprof.items()}}
            if lam_exact is not None:
                mixed["eig_err_vs_exact_over_norm2"] = \
                    (lam_m.double() - lam_exact).abs().max().item() / lam_exact.abs().max().item()
            log(f"[bench] mixed {mixed}")
        except cv.CuevError as e:
            mixed = {"error": str(e)}

    cpu = None
    if a.cpu_n:
        log(f"[bench] CPU baseline: dsyevd n={a.cpu_n}")
        del work
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
        "roofline": roofline,
        "cpu_baseline": cpu,
        "mixed": mixed,
        "lapack": cv.lapack_name(),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
