"""Benchmark: SH_to_spat + spat_to_SH transform pairs per second on one MI355X.

Workload (one step): synthesis then analysis of a batch of 16 independent FP64 fields at
lmax = mmax = 1023 on the Gauss grid nlat = 1536, nphi = 3072 (theta-contiguous), with
a_lm ~ N(0, (l+1)^-2) per component (SplitMix64/Box-Muller, seed 1 + field), inputs
resident in HBM.  value = 16 * steps / elapsed  (pairs/s, whole job).

roofline: the dominant kernel (largest share of the timed region) timed live with HIP
events recorded by the library on its compute stream; achieved = algorithmic FLOPs per
launch (2*nlat*NLM per field for a Legendre kernel, 2.5*nlat*nphi*log2(nphi) per field for
an FFT kernel) / average launch duration, against the FP64 peak 78.6 TFLOP/s.
cpu_baseline: the CPU oracle (oracle/, an OpenMP FP64 restatement of SHTns's transform)
timed on this host's cores on one-field pairs of the same workload.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

LMAX, NLAT, NPHI, BATCH, SEED = 1023, 1536, 3072, 16, 1
# CPU pairs/s at lmax=1023, 1536x3072, one field, measured on the same 8-vCPU host type:
# SHTns 3.5.2 (OpenMP + MKL FFT, time_SHT protocol, 26.56 ms per pair) and this repository's
# oracle port (median of 10 pairs, 0.964 s per pair).
SHTNS_PAIRS_S_8CPU, PORT_PAIRS_S_8CPU = 37.7, 1.038
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def nlm_of(lmax):
    return (lmax + 1) * (lmax + 2) // 2


def leg_flops(lmax, nlat):
    return 2.0 * nlat * nlm_of(lmax)


def fft_flops(nlat, nphi):
    return 2.5 * nlat * nphi * math.log2(nphi)


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "sht_pair_l1023_b16_hbm_traffic.txt")
# rows of the FP64 kernels of this workload in that summary
TRAFFIC_ROWS = {"legendre_synth": ("premix_kernel<64>", "legendre_synth_kernel<4, 0, 0>"),
                "legendre_anal": ("legendre_anal_kernel<12, 1, 2, 0, 0, 1, 1>",), "fft_synth": ("fft_synth",),
                "fft_anal": ("fft_anal",)}


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, the copy-kernel
    calibration) from the committed rocprofv3 --pmc summary of this workload, or None."""
    try:
        lines = open(TRAFFIC_PROFILE).read().splitlines()
    except OSError:
        return None
    total, found = 0.0, 0
    for prefix in TRAFFIC_ROWS.get(kernel, ()):
        for ln in lines:
            if ln.startswith(prefix):
                parts = ln.split()
                total += (float(parts[-2]) + float(parts[-1])) * 1e6
                found += 1
                break
    return total if found == len(TRAFFIC_ROWS.get(kernel, ())) and found else None


def cpu_baseline(budget_s=10.0):
    """Time the CPU oracle on one-field SH_to_spat + spat_to_SH pairs of the workload."""
    from oracle.oracle import Oracle, set_threads
    from shtns_amd.synthetic import spectral_field
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except Exception:
        pass
    set_threads(cores)
    o = Oracle(LMAX, LMAX, 1, NLAT, NPHI)
    q = spectral_field(LMAX, LMAX, seed=SEED)
    v = o.synth(q)          # warm-up pair (not timed)
    o.analys(v)
    n, t0 = 0, time.perf_counter()
    while True:
        v = o.synth(q)
        o.analys(v)
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s or n >= 50:
            break
    rate = n / el
    return {"value": rate, "unit": "pairs/s", "cores": cores, "kind": "port",
            "sample": f"{n} one-field pairs (lmax=1023, 1536x3072, FP64) of the same workload, "
                      f"{el:.1f} s, OpenMP over {cores} threads",
            # The port is not SHTns: on one 8-vCPU host the port ran PORT_PAIRS_S_8CPU and SHTns
            # (OpenMP, MKL FFT, AVX-512 kernels) SHTNS_PAIRS_S_8CPU pairs/s at this size; scaled
            # by that ratio, SHTns on this host would run about `shtns_equiv_value` pairs/s.
            "port_over_shtns_8cpu": round(PORT_PAIRS_S_8CPU / SHTNS_PAIRS_S_8CPU, 4),
            "shtns_equiv_value": round(rate * SHTNS_PAIRS_S_8CPU / PORT_PAIRS_S_8CPU, 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-prec", action="store_true", help="skip the FP32 / FP32-split contraction legs")
    ap.add_argument("--cpu-budget", type=float, default=10.0)
    ap.add_argument("--layout", choices=["theta", "phi"], default="theta",
                    help="spatial layout (the metric is quoted on theta-contiguous)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="after the timed loop, copy a 512 MiB buffer once (reference kernel of known "
                         "traffic for hardware-counter runs)")
    ap.add_argument("--workload", choices=["pair", "uneven", "vector", "cplx", "host"], default="pair",
                    help="pair: the headline scalar pair at lmax=1023 (default); uneven: 256 transforms "
                         "of uneven lmax from one graph plus one lmax=4095 field; vector: spheroidal/"
                         "toroidal pairs at lmax=511; cplx: complex-valued scalar pairs at lmax=1023; "
                         "host: one field through the host-pointer API (pinned shtns_malloc buffers, "
                         "transfers included)")
    ap.add_argument("--tune", action="store_true", help="autotune the kernel shapes (cushtns_tune) before timing")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "sharding across GPUs is not built.", file=sys.stderr)
        sys.exit(2)
    if args.workload == "uneven":
        return bench_uneven(args)
    if args.workload == "vector":
        return bench_vector(args)
    if args.workload == "cplx":
        return bench_cplx(args)
    if args.workload == "host":
        return bench_host(args)

    import numpy as np
    import torch

    import shtns_amd
    from shtns_amd import _lib
    from shtns_amd.synthetic import spectral_batch

    lib = _lib.load()
    dev = torch.device("cuda:0")
    s = shtns_amd.sht(LMAX)
    lay = shtns_amd.SHT_THETA_CONTIGUOUS if args.layout == "theta" else shtns_amd.SHT_PHI_CONTIGUOUS
    s.set_grid(NLAT, NPHI, flags=lay | shtns_amd.sht_gauss, polar_opt=0.0)
    nlm, nspat = s.nlm, s.nspat
    q = torch.from_numpy(spectral_batch(LMAX, LMAX, BATCH, seed=SEED)).to(dev)
    v = torch.empty((BATCH,) + tuple(s.spat_shape), dtype=torch.float64, device=dev)
    q2 = torch.empty_like(q)
    stream = torch.cuda.current_stream(dev)
    lib.cushtns_set_streams(s._cfg, ctypes.c_void_p(stream.cuda_stream), ctypes.c_void_p(stream.cuda_stream))
    qp, vp, q2p = (ctypes.c_void_p(t.data_ptr()) for t in (q, v, q2))
    tuned = None
    if args.tune:
        from shtns_amd.batch import tune
        tuned = tune(s, BATCH)

    def step():
        lib.cu_SH_to_spat_batch(s._cfg, qp, vp, LMAX, BATCH, nlm, nspat)
        lib.cu_spat_to_SH_batch(s._cfg, vp, q2p, LMAX, BATCH, nspat, nlm)

    names = ["legendre_synth", "legendre_anal", "fft_synth", "fft_anal"]
    flops = [leg_flops(LMAX, NLAT) * BATCH, leg_flops(LMAX, NLAT) * BATCH,
             fft_flops(NLAT, NPHI) * BATCH, fft_flops(NLAT, NPHI) * BATCH]

    def timed_loop(prec):
        """W untimed steps, then K steps between two device synchronizations; the library's
        HIP events on its compute stream time each kernel inside the timed region."""
        lib.cushtns_set_precision(s._cfg, prec)
        for _ in range(args.warmup):
            step()
        _lib.check_error("warmup")
        torch.cuda.synchronize()
        lib.shtns_profile(s._cfg, 1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        _lib.check_error("timed loop")
        ms = (ctypes.c_double * 8)()
        cnt = (ctypes.c_int * 8)()
        lib.shtns_profile_get(s._cfg, ms, cnt)
        lib.shtns_profile(s._cfg, 0)
        elapsed = t1 - t0
        # One transform call may launch a kernel several times (premix + Legendre, or chunks);
        # times are summed per call: per_call_ms = total / steps.
        kern = {}
        for k, nm in enumerate(names):
            if cnt[k]:
                per_call = ms[k] / args.steps * 1e-3
                kern[nm] = {"per_call_ms": per_call * 1e3, "launches_per_call": cnt[k] / args.steps,
                            "share": ms[k] * 1e-3 / elapsed, "tflops": flops[k] / per_call / 1e12}
        return elapsed, kern

    pairs = BATCH * args.steps
    elapsed, kern = timed_loop(-1)
    value = pairs / elapsed

    if args.calib_copy:
        a = torch.empty(64 * 1024 * 1024, dtype=torch.float64, device=dev).fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)     # reads 512 MiB, writes 512 MiB
        torch.cuda.synchronize()
        del a, b

    # accuracy of the timed output (round trip), outside the timed region
    err = float(torch.linalg.norm((q2 - q).flatten()) / torch.linalg.norm(q.flatten()))

    # reduced-precision contractions on the same inputs and timing protocol (recurrence,
    # seeds, FFT and final sums stay FP64), errors against the FP64 outputs above
    prec_legs = {}
    if not args.no_prec:
        v64 = v.clone()
        q64 = q2.clone()
        for name, code in (("int8_sliced", 3), ("fp32_split", 2), ("fp32", 1)):
            el, kp = timed_loop(code)
            ev = float(torch.linalg.norm((v - v64).flatten()) / torch.linalg.norm(v64.flatten()))
            eq = float(torch.linalg.norm((q2 - q64).flatten()) / torch.linalg.norm(q64.flatten()))
            pv = pairs / el
            prec_legs[name] = {"pairs_per_s": round(pv, 2), "speedup_vs_fp64": round(pv / value, 3),
                               "synth_rel_l2_vs_fp64": ev, "pair_rel_l2_vs_fp64": eq,
                               "legendre_ms": {k: round(kp[k]["per_call_ms"], 4) for k in ("legendre_synth", "legendre_anal")
                                               if k in kp}}
        lib.cushtns_set_precision(s._cfg, -1)

    dom = max(kern, key=lambda n: kern[n]["share"])
    dk = kern[dom]
    traffic = profiled_traffic(dom) if args.layout == "theta" else None
    roof = {"bound": "mfma", "achieved": round(dk["tflops"], 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(dk["tflops"] / FP64_PEAK_TFLOPS, 4), "traffic": traffic, "kernel": dom,
            "traffic_source": os.path.relpath(TRAFFIC_PROFILE, ROOT) if traffic else None}
    alg_total = (leg_flops(LMAX, NLAT) + fft_flops(NLAT, NPHI)) * 2 * pairs
    out = {
        "metric": "SH_to_spat+spat_to_SH pairs/s at lmax=1023 (FP32-split vs FP64); % FP64 peak",
        "value": round(value, 2),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (SplitMix64/Box-Muller a_lm ~ N(0,(l+1)^-2), seed 1+field)",
        "config": {"workload": "SH_to_spat+spat_to_SH, lmax=mmax=1023, Gauss nlat=1536 nphi=3072, "
                               f"{args.layout}-contiguous, batch 16 FP64 fields",
                   "lmax": LMAX, "nlat": NLAT, "nphi": NPHI, "batch": BATCH, "parallelism": "single GPU",
                   "tuned_shapes": tuned},
        "fp64_frac_of_peak_whole_pair": round(alg_total / elapsed / 1e12 / FP64_PEAK_TFLOPS, 4),
        "round_trip_rel_l2": err,
        "kernels": {k: {kk: round(vv, 4) if isinstance(vv, float) else vv for kk, vv in d.items()}
                    for k, d in kern.items()},
        "roofline": roof,
        "precision_legs": prec_legs,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_budget)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


from shtns_amd.synthetic import UNEVEN_LMAX, UNEVEN_N, UNEVEN_SEED_DATA, uneven_lmaxes  # noqa: E402


def bench_uneven(args):
    """256 scalar transform pairs of uneven lmax, Gauss grids nlat = 3(lmax+1)/2, nphi = 2 nlat,
    one HIP graph per direction (cushtns_batch_*), FP64 and FP32-split; then one lmax=4095
    field on 6144 x 12288 (FP64)."""
    import torch
    import shtns_amd
    from shtns_amd import _lib
    from shtns_amd.batch import UnevenBatch, tune
    from shtns_amd.synthetic import spectral_field
    lmaxes = uneven_lmaxes()
    cfg = {}
    for L in UNEVEN_LMAX:
        nlat = 3 * (L + 1) // 2
        sh = shtns_amd.sht(L)
        sh.set_grid(nlat, 2 * nlat, flags=shtns_amd.SHT_THETA_CONTIGUOUS | shtns_amd.sht_gauss, polar_opt=0.0)
        cfg[L] = sh
    shts = [cfg[L] for L in lmaxes]
    alg = sum(2 * (leg_flops(L, 3 * (L + 1) // 2) + fft_flops(3 * (L + 1) // 2, 3 * (L + 1))) for L in lmaxes)
    out = {}
    results = {}
    for name, prec in (("fp64", shtns_amd.SHT_PREC_F64), ("fp32_split", shtns_amd.SHT_PREC_SPLIT)):
        b = UnevenBatch(shts, prec=prec, tune=args.tune)
        qs = []
        for i, L in enumerate(lmaxes):
            q = torch.from_numpy(spectral_field(L, L, seed=UNEVEN_SEED_DATA + i))
            b.arrays(i)[0].copy_(q)
            qs.append(q)
        q0 = [b.arrays(i)[0].clone() for i in range(len(b))]
        for _ in range(args.warmup):
            b.synth()
            b.analys()
        torch.cuda.synchronize()
        # restore inputs (the round trips of the warm-up rewrote them with their own output)
        for i in range(len(b)):
            b.arrays(i)[0].copy_(q0[i])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            b.synth()
            b.analys()
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        _lib.check_error("uneven batch")
        results[name] = [b.arrays(i)[0].clone() for i in range(len(b))]
        num = sum(float(torch.linalg.norm(results[name][i] - q0[i]) ** 2) for i in range(len(b)))
        den = sum(float(torch.linalg.norm(q0[i]) ** 2) for i in range(len(b)))
        out[name] = {"pairs_per_s": round(UNEVEN_N * args.steps / el, 2), "ms_per_step": round(el / args.steps * 1e3, 4),
                     "round_trip_rel_l2": (num / den) ** 0.5,
                     "fp64_frac_of_peak": round(alg * args.steps / el / 1e12 / FP64_PEAK_TFLOPS, 4)}
        b.close()
    num = sum(float(torch.linalg.norm(results["fp32_split"][i] - results["fp64"][i]) ** 2) for i in range(UNEVEN_N))
    den = sum(float(torch.linalg.norm(results["fp64"][i]) ** 2) for i in range(UNEVEN_N))
    out["fp32_split"]["rel_l2_vs_fp64"] = (num / den) ** 0.5
    out["fp32_split"]["speedup_vs_fp64"] = round(out["fp32_split"]["pairs_per_s"] / out["fp64"]["pairs_per_s"], 3)
    # one lmax = 4095 field
    L4, n4 = 4095, 6144
    s4 = shtns_amd.sht(L4)
    s4.set_grid(n4, 2 * n4, flags=shtns_amd.SHT_THETA_CONTIGUOUS | shtns_amd.sht_gauss, polar_opt=0.0)
    q4 = torch.from_numpy(spectral_field(L4, L4, seed=UNEVEN_SEED_DATA + UNEVEN_N)).cuda()
    v4 = s4.synth(q4)
    s4.analys(v4)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    v4 = s4.synth(q4)
    r4 = s4.analys(v4)
    torch.cuda.synchronize()
    el4 = time.perf_counter() - t0
    f4 = 2 * (leg_flops(L4, n4) + fft_flops(n4, 2 * n4))
    out["lmax4095"] = {"pair_ms": round(el4 * 1e3, 3), "round_trip_rel_l2": float(torch.linalg.norm(r4 - q4) / torch.linalg.norm(q4)),
                       "fp64_frac_of_peak": round(f4 / el4 / 1e12 / FP64_PEAK_TFLOPS, 4)}
    counts = {L: lmaxes.count(L) for L in UNEVEN_LMAX}
    print(json.dumps({
        "metric": "uneven batch: SH_to_spat+spat_to_SH pairs/s, 256 transforms with lmax in {63..1023}",
        "value": out["fp64"]["pairs_per_s"], "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": out["fp64"]["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (SplitMix64/Box-Muller a_lm ~ N(0,(l+1)^-2), lmax draw seed 4, data seed 5+i)",
        "config": {"workload": "256 scalar transforms, lmax uniform in {63,127,255,511,1023}, Gauss nlat=3(lmax+1)/2, "
                               "nphi=2 nlat, theta-contiguous, one HIP graph per direction; plus one lmax=4095 "
                               "6144x12288 FP64 field", "counts": counts, "tuned": bool(args.tune),
                   "parallelism": "single GPU"},
        "legs": out}))


def bench_vector(args):
    """32 spheroidal/toroidal pairs at lmax=511 on 768 x 1536 (SHsphtor_to_spat then
    spat_to_SHsphtor), S,T ~ N(0,(l+1)^-3) per component, seed 2; FP64 and FP32 contraction."""
    import torch
    import shtns_amd
    from shtns_amd import _lib
    from shtns_amd.synthetic import spectral_batch
    L, nlat, nphi, nb = 511, 768, 1536, 32
    sh = shtns_amd.sht(L)
    sh.set_grid(nlat, nphi, flags=shtns_amd.SHT_THETA_CONTIGUOUS | shtns_amd.sht_gauss, polar_opt=0.0)
    S = torch.from_numpy(spectral_batch(L, L, nb, seed=2, power=3.0)).cuda()
    T = torch.from_numpy(spectral_batch(L, L, nb, seed=2 + nb, power=3.0)).cuda()
    S[:, 0] = 0
    T[:, 0] = 0
    nlm1 = (L + 2) * (L + 3) // 2
    alg = 2 * (2 * 2 * nlat * nlm1 + 2 * fft_flops(nlat, nphi)) * nb
    out = {}
    ref = None
    for name, prec in (("fp64", shtns_amd.SHT_PREC_F64), ("fp32_split", shtns_amd.SHT_PREC_SPLIT),
                       ("fp32", shtns_amd.SHT_PREC_F32)):
        sh.set_precision(prec)
        for _ in range(args.warmup):
            vt, vp = sh.synth(S, T)
            s2, t2 = sh.analys(vt, vp)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            vt, vp = sh.synth(S, T)
            s2, t2 = sh.analys(vt, vp)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        _lib.check_error("vector")
        err = float(torch.sqrt((torch.linalg.norm(s2 - S) ** 2 + torch.linalg.norm(t2 - T) ** 2) /
                               (torch.linalg.norm(S) ** 2 + torch.linalg.norm(T) ** 2)))
        leg = {"pairs_per_s": round(nb * args.steps / el, 2), "ms_per_step": round(el / args.steps * 1e3, 4),
               "round_trip_rel_l2": err, "fp64_frac_of_peak": round(alg * args.steps / el / 1e12 / FP64_PEAK_TFLOPS, 4)}
        if ref is None:
            ref = (vt.clone(), vp.clone())
        else:
            leg["synth_rel_l2_vs_fp64"] = float(torch.sqrt((torch.linalg.norm(vt - ref[0]) ** 2 + torch.linalg.norm(vp - ref[1]) ** 2) /
                                                           (torch.linalg.norm(ref[0]) ** 2 + torch.linalg.norm(ref[1]) ** 2)))
        out[name] = leg
    sh.set_precision(shtns_amd.SHT_PREC_AUTO)
    print(json.dumps({
        "metric": "vector SHsphtor_to_spat+spat_to_SHsphtor pairs/s at lmax=511",
        "value": out["fp64"]["pairs_per_s"], "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": out["fp64"]["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (SplitMix64/Box-Muller S,T ~ N(0,(l+1)^-3), seeds 2+f, 34+f; S_00=T_00=0)",
        "config": {"workload": "32 (S,T) pairs, lmax=mmax=511, Gauss 768x1536 theta-contiguous", "lmax": L,
                   "nlat": nlat, "nphi": nphi, "batch": nb, "parallelism": "single GPU"},
        "legs": out}))


def bench_cplx(args):
    """16 complex-valued fields at lmax=mmax=1023 on 1536 x 3072 (SH_to_spat_cplx then
    spat_cplx_to_SH, batched on the device): coefficients a_lm for -l <= m <= l, real and
    imaginary parts ~ N(0, (l+1)^-2), seeded.  One complex pair is the work of two real pairs."""
    import numpy as np
    import torch
    import shtns_amd
    from shtns_amd import _lib
    L, nlat, nphi, nb = 1023, 1536, 3072, 16
    sh = shtns_amd.sht(L)
    sh.set_grid(nlat, nphi, flags=shtns_amd.SHT_THETA_CONTIGUOUS | shtns_amd.sht_gauss, polar_opt=0.0)
    rng = np.random.default_rng(6)
    lfac = np.array([1.0 / (l + 1) for l in range(L + 1) for _ in range(2 * l + 1)])
    A = torch.from_numpy((rng.standard_normal((nb, sh.nlm_cplx)) + 1j * rng.standard_normal((nb, sh.nlm_cplx)))
                         * lfac).cuda()
    for _ in range(args.warmup):
        z = sh.synth_cplx(A)
        a2 = sh.analys_cplx(z)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        z = sh.synth_cplx(A)
        a2 = sh.analys_cplx(z)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    _lib.check_error("cplx")
    err = float(torch.linalg.norm(a2 - A) / torch.linalg.norm(A))
    nlm = (L + 1) * (L + 2) // 2
    alg = 2 * 2 * (2 * nlat * nlm + fft_flops(nlat, nphi)) * nb     # two real fields per complex field
    print(json.dumps({
        "metric": "complex-valued SH_to_spat_cplx+spat_cplx_to_SH pairs/s at lmax=1023",
        "value": round(nb * args.steps / el, 2), "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(el / args.steps * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (numpy default_rng(6) a_lm, re/im ~ N(0,(l+1)^-2))",
        "config": {"workload": "16 complex fields, lmax=mmax=1023, Gauss 1536x3072 theta-contiguous", "lmax": L,
                   "nlat": nlat, "nphi": nphi, "batch": nb, "parallelism": "single GPU"},
        "round_trip_rel_l2": err,
        "fp64_frac_of_peak": round(alg * args.steps / el / 1e12 / FP64_PEAK_TFLOPS, 4)}))


def bench_host(args):
    """One field at lmax=1023 on 1536 x 3072 through the host-pointer API SH_to_spat /
    spat_to_SH (shtns.h:262-267) on pinned shtns_malloc arrays: each call copies the input to
    the GPU, transforms and copies the result back (the published GPU figure for SHTns,
    doc/main.dox:684-696, times this mode: P100 + NVLink, 3.46 ms / 4.24 ms per transform)."""
    import numpy as np
    import shtns_amd
    from shtns_amd import _lib
    from shtns_amd.synthetic import spectral_field
    lib = _lib.load()
    s = shtns_amd.sht(LMAX)
    s.set_grid(NLAT, NPHI, flags=shtns_amd.SHT_THETA_CONTIGUOUS | shtns_amd.sht_gauss, polar_opt=0.0)
    nlm, nspat = s.nlm, s.nspat
    pq, pv = lib.shtns_malloc(16 * nlm), lib.shtns_malloc(8 * nspat)
    try:
        q = np.ctypeslib.as_array((ctypes.c_double * (2 * nlm)).from_address(pq)).view(np.complex128)
        v = np.ctypeslib.as_array((ctypes.c_double * nspat).from_address(pv)).reshape(s.spat_shape)
        q0 = spectral_field(LMAX, LMAX, seed=SEED)
        q[:] = q0
        for _ in range(max(args.warmup, 1)):
            s.SH_to_spat(q, v)
            s.spat_to_SH(v, q)
        ts, ta = [], []
        for _ in range(args.steps):
            t0 = time.perf_counter()
            s.SH_to_spat(q, v)
            t1 = time.perf_counter()
            s.spat_to_SH(v, q)
            ts.append(t1 - t0)
            ta.append(time.perf_counter() - t1)
        err = float(np.linalg.norm(q - q0) / np.linalg.norm(q0))
    finally:
        lib.shtns_free(ctypes.c_void_p(pq))
        lib.shtns_free(ctypes.c_void_p(pv))
    ms_s, ms_a = 1e3 * float(np.median(ts)), 1e3 * float(np.median(ta))
    print(json.dumps({
        "metric": "host-pointer SH_to_spat + spat_to_SH at lmax=1023, transfers included",
        "value": round(1e3 / (ms_s + ms_a), 2), "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_s + ms_a, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (SplitMix64/Box-Muller a_lm ~ N(0,(l+1)^-2), seed 1)",
        "config": {"workload": "one field, lmax=mmax=1023, Gauss 1536x3072 theta-contiguous, host arrays from "
                               "shtns_malloc (pinned), copies to and from the GPU inside each call",
                   "parallelism": "single GPU"},
        "SH_to_spat_ms": round(ms_s, 4), "spat_to_SH_ms": round(ms_a, 4),
        "bytes_moved_per_transform": 16 * nlm + 8 * nspat,
        "round_trip_rel_l2": err}))


if __name__ == "__main__":
    main()
