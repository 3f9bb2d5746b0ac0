"""Benchmark: [X,Y] scalar IMSRG(2) commutators per second at emax=10 on one MI355X.

Workload: Z = [X, Y] with X anti-Hermitian and Y Hermitian, HO basis emax=10
(132 orbits), O16 reference, two-body blocks (R +/- R^T)/sqrt(2n) with seeded
N(0,1) entries (synthetic data; the library's counter-based generator).
Operators are resident in HBM before the timed region; each step is one full
commutator (all eight IMSRG(2) terms).

Prints one JSON line.  ``value`` is the mixed-precision rate when the mixed
result meets max|dZ|/max|Z| < 1e-7 and |dE0| < 1e-6 against the FP64 result,
else the FP64 rate; both rates are reported.  ``roofline`` is for the dominant
kernel, from HIP events recorded on the library stream during the timed steps;
``cpu_baseline`` times the oracle (C/OpenMP restatement, BLAS GEMMs) on the
same workload on this host (median of five commutators).  ``--workload`` selects the other measured
workloads (batched pairs, emax=12 and emax=14 precision, Magnus flow), one JSON line each.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_MFMA = 78.6e12   # FLOP/s, MI355X dense FP64 matrix
PEAK_FP32_MFMA = 157.3e12  # FLOP/s, MI355X dense FP32 matrix
PEAK_BF16_MFMA = 2.5e15    # FLOP/s, MI355X dense bf16 matrix
B3_PRODUCTS = 6            # bf16 MFMA products per FP64 product in the bf16x3 kernels
PEAK_HBM = 8.0e12          # B/s

PHASES = ["zerobody", "onebody", "comm122", "pphh_gemm", "pandya_gather", "ph_gemm", "inverse_pandya", "fold_mirror"]


def algorithmic_counts(ms):
    """Per-commutator algorithmic work of the reference algorithm over the
    basis tables (src/ModelSpace.cc:1120-1201, src/TwoBodyChannel.cc:142-188):
      F_A = sum_ch 2 n^2 (n_pp + 2 n_hh + n_ph)   pp/hh GEMMs (src/Commutator.cc:1176-1183)
      F_C = sum_cc 8 c^2 n_ph,cc                   ph GEMM (src/Commutator.cc:1789)
      B_I = z sum_cc 2 c^2 + 8 sum_ch n(n+1)/2     inverse Pandya: one read of Zbar over the
                                                   CC channels with ph kets (Zbar is zero in
                                                   the others), one store per upper Z element;
                                                   z = 8 for FP64 Zbar, 4 for the FP32 Zbar of
                                                   the FP32-product ph GEMM
    Returns (F_A, F_A executed, F_C, B_I with FP64 Zbar, B_I with FP32 Zbar)."""
    orb = ms.orbits
    occ = [o["occ"] for o in orb]
    FA = 0
    executed_A = 0
    upper = 0
    for ch, c in enumerate(ms.channels):
        n = c["n"]
        if n == 0:
            continue
        npp = nhh = nph = 0
        for p, q in ms.kets(ch):
            hp, hq = occ[p] > 1e-6, occ[q] > 1e-6
            if hp and hq:
                nhh += 1
            elif hp or hq:
                nph += 1
            else:
                npp += 1
        FA += 2 * n * n * (npp + 2 * nhh + nph)
        executed_A += 2 * n * n * (npp + nhh)
        upper += n * (n + 1) // 2
    # cross-coupled channels (J', parity, |tz_p - tz_q|/2): kets p <= q, no Pauli rule
    cc = {}
    for p in range(len(orb)):
        for q in range(p, len(orb)):
            a, b = orb[p], orb[q]
            key0 = ((a["l"] + b["l"]) & 1, abs(a["tz2"] - b["tz2"]) // 2)
            ph = (occ[p] > 1e-6) or (occ[q] > 1e-6)  # the reference's CC ph list holds hh and ph kets
            for J in range(abs(a["j2"] - b["j2"]) // 2, (a["j2"] + b["j2"]) // 2 + 1):
                e = cc.setdefault((J,) + key0, [0, 0])
                e[0] += 1
                e[1] += ph
    FC = sum(8 * c * c * nph for c, nph in cc.values() if nph > 0)
    zbar = sum(2 * c * c for c, nph in cc.values() if nph > 0)
    return FA, executed_A, FC, 8 * zbar + 8 * upper, 4 * zbar + 8 * upper


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE and WRITE_SIZE scaled by a plain-copy calibration), or
    None when no summary for it exists."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        return t["kernels"][kernel]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def _sum_or_none(*v):
    return None if any(x is None for x in v) else sum(v)


def _errors(Z, R, np):
    scale = max(np.max(np.abs(R.TwoBody)), np.max(np.abs(R.OneBody)))
    rel = float(max(np.max(np.abs(Z.TwoBody - R.TwoBody)), np.max(np.abs(Z.OneBody - R.OneBody))) / scale)
    return rel, abs(Z.ZeroBody - R.ZeroBody)


def workload_batched(args, ia, np):
    """48 independent pairs, emax drawn uniformly from {4,5,6,7} (seed 3000),
    pair i seeded 3000+2i+1 / 3000+2i+2: one grouped launch sequence for all
    pairs against a loop of single-pair commutators."""
    rng = np.random.default_rng(3000)
    emaxes = [int(e) for e in rng.integers(4, 8, size=48)]
    spaces = {e: ia.ModelSpace(e, "O16", device=0) for e in sorted(set(emaxes))}
    pairs = []
    for i, e in enumerate(emaxes):
        ms = spaces[e]
        pairs.append((ia.DeviceOperator(ms, -1).fill_random(-1, 3000 + 2 * i + 1),
                      ia.DeviceOperator(ms, 1).fill_random(1, 3000 + 2 * i + 2), ia.DeviceOperator(ms, 1)))

    def sync():
        for ms in spaces.values():
            ms.synchronize()

    reps = max(args.steps, 5)
    out = {}
    for name, prec in (("fp64", ia.PREC_FP64), ("mixed", ia.PREC_MIXED)):
        for _ in range(2):
            ia.Commutator.CommutatorScalarScalarBatched(pairs, precision=prec)
            for X, Y, Z in pairs:
                ia.Commutator.CommutatorScalarScalar(X, Y, precision=prec, out=Z)
        sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            ia.Commutator.CommutatorScalarScalarBatched(pairs, precision=prec)
        sync()
        tb = time.perf_counter() - t0
        ref = [Z.to_host() for _, _, Z in pairs]
        t0 = time.perf_counter()
        for _ in range(reps):
            for X, Y, Z in pairs:
                ia.Commutator.CommutatorScalarScalar(X, Y, precision=prec, out=Z)
        sync()
        tl = time.perf_counter() - t0
        same = all(np.array_equal(r.TwoBody, Z.to_host().TwoBody) for r, (_, _, Z) in zip(ref, pairs))
        out[name] = {"batched_pairs_per_s": 48 * reps / tb, "looped_pairs_per_s": 48 * reps / tl,
                     "speedup": tl / tb, "batched_equals_looped": bool(same)}
    return {"metric": "independent [X,Y] pairs/sec, 48 pairs with emax in {4,5,6,7}: one grouped launch vs a loop",
            "value": out["fp64"]["batched_pairs_per_s"], "unit": "pairs/s", "n_gpus": 1, "steps": reps,
            "higher_is_better": True, "dtype": "f64", "data": "synthetic (seeded, library generator)",
            "config": {"workload": "48 scalar IMSRG(2) commutator pairs, O16 reference, emax drawn from {4,5,6,7}",
                       "emax_counts": {e: emaxes.count(e) for e in sorted(set(emaxes))}},
            "fp64": out["fp64"], "mixed": out["mixed"]}


def workload_emax12(args, ia, np, emax=12):
    """One commutator at emax=12 (182 orbits; or emax=14, 240 orbits): FP64,
    mixed and split-product rates and the errors of both reduced-precision
    modes against the FP64 GPU result.  Operators seeded 1000*emax+1 / +2."""
    ms = ia.ModelSpace(emax, "O16", device=0)
    X = ia.DeviceOperator(ms, -1).fill_random(-1, 1000 * emax + 1)
    Y = ia.DeviceOperator(ms, 1).fill_random(1, 1000 * emax + 2)
    Z = ia.DeviceOperator(ms, 1)
    res = {}
    hosts = {}
    for name, prec in (("fp64", ia.PREC_FP64), ("mixed", ia.PREC_MIXED), ("split", ia.PREC_SPLIT)):
        ia.Commutator.CommutatorScalarScalar(X, Y, precision=prec, out=Z)
        ms.synchronize()
        steps = max(3, min(args.steps, 5))
        ms.set_profiling(True)
        ms.phase_times()
        t0 = time.perf_counter()
        for _ in range(steps):
            ia.Commutator.CommutatorScalarScalar(X, Y, precision=prec, out=Z)
        ms.synchronize()
        t = time.perf_counter() - t0
        phase = np.array(ms.phase_times())
        ms.set_profiling(False)
        hosts[name] = Z.to_host()
        res[name] = {"commutators_per_s": steps / t, "phase_us": dict(zip(PHASES, [round(x, 1) for x in phase]))}
    rel, de0 = _errors(hosts["mixed"], hosts["fp64"], np)
    rel_s, de0_s = _errors(hosts["split"], hosts["fp64"], np)
    res["split"].update({"rel_dZ": rel_s, "abs_dE0": de0_s})
    d = np.abs(hosts["mixed"].TwoBody - hosts["fp64"].TwoBody)
    return {"metric": f"mixed-precision error at emax={emax} vs the FP64 GPU result", "value": rel,
            "unit": "max-abs dZ / max |Z|", "n_gpus": 1, "higher_is_better": False,
            "dtype": "mixed: bf16x3 MFMA pp/hh + f32 MFMA ph + f64; split: 3 f32 MFMA products + f64",
            "data": f"synthetic (seeds {1000 * emax + 1}/{1000 * emax + 2})",
            "config": {"workload": f"scalar IMSRG(2) commutator, HO emax={emax}, O16 reference", "norb": ms.norb,
                       "two_body_doubles": ms.tbsize},
            "abs_dE0": de0, "max_abs_dZ": float(np.max(d)),
            "max_rel_dZ_elementwise_over_1e-3_max": float(np.max(d / np.maximum(np.abs(hosts["fp64"].TwoBody),
                                                                                1e-3 * np.max(np.abs(hosts["fp64"].TwoBody))))),
            "fp64": res["fp64"], "mixed": res["mixed"], "split": res["split"]}


def workload_flow(args, ia, np):
    """Magnus-Euler IMSRG(2) flow at emax=10: one-body diagonal 10 (2n+l+1.5)
    plus N(0, 0.5) off-diagonals inside each (l, j, tz) block, two-body
    (R+R^T)/sqrt(2n) (seed 5001), White generator, ds = 0.5, BCH stopped at
    nested norm < 1e-8; E0(s) of the mixed-precision flow against the FP64 one."""
    ms = ia.ModelSpace(10, "O16", device=0)
    H = ia.DeviceOperator(ms, 1).fill_random(1, 5001)
    Hh = H.to_host()
    rng = np.random.default_rng(5001)
    Hh.OneBody[:] = 0.0
    Hh.ZeroBody = 0.0
    blocks = {}
    for i, o in enumerate(ms.orbits):
        blocks.setdefault((o["l"], o["j2"], o["tz2"]), []).append(i)
    for blk in blocks.values():
        for x, a in enumerate(blk):
            o = ms.orbits[a]
            Hh.OneBody[a, a] = 10.0 * (2 * o["n"] + o["l"] + 1.5)
            for b in blk[x + 1:]:
                v = rng.normal(0.0, 0.5)
                Hh.OneBody[a, b] = Hh.OneBody[b, a] = v
    H.upload(Hh)
    steps = 40
    res = {}
    for name, prec in (("fp64", ia.PREC_FP64), ("mixed", ia.PREC_MIXED)):
        t0 = time.perf_counter()
        Hs, e0, en, nc = ia.IMSRGSolver.Solve_magnus_euler(H, ds=0.5, smax=0.5 * steps, eta_criterion=1e-6,
                                                           max_steps=steps, mode=1, transform_thr=1e-8,
                                                           precision=prec)
        t = time.perf_counter() - t0
        res[name] = {"steps": len(e0), "seconds": t, "commutators": int(np.sum(nc)),
                     "commutators_per_s": float(np.sum(nc)) / t, "E0_final": float(e0[-1]) if len(e0) else None,
                     "eta_norm_final": float(en[len(e0)]) if len(en) > len(e0) else None,
                     "E0": [float(x) for x in e0]}
    # the same FP64 flow with every nested commutator replayed from a captured hipGraph
    ms.set_graphs(True)
    t0 = time.perf_counter()
    Hs, e0g, en, nc = ia.IMSRGSolver.Solve_magnus_euler(H, ds=0.5, smax=0.5 * steps, eta_criterion=1e-6,
                                                        max_steps=steps, mode=1, transform_thr=1e-8,
                                                        precision=ia.PREC_FP64)
    tg = time.perf_counter() - t0
    ms.set_graphs(False)
    res["fp64_graphs"] = {"seconds": tg, "commutators_per_s": float(np.sum(nc)) / tg,
                          "E0_equal_to_eager": bool(np.array_equal(e0g, np.array(res["fp64"]["E0"])))}
    # launch overhead matters at small sizes: the same flow recipe at emax=4
    ms4 = ia.ModelSpace(4, "O16", device=0)
    H4 = ia.DeviceOperator(ms4, 1).fill_random(1, 5001)
    H4h = H4.to_host()
    H4h.OneBody[:] = 0.0
    for i, o in enumerate(ms4.orbits):
        H4h.OneBody[i, i] = 10.0 * (2 * o["n"] + o["l"] + 1.5)
    H4.upload(H4h)
    small = {}
    for name, on in (("eager", False), ("graphs", True)):
        ms4.set_graphs(on)
        ia.IMSRGSolver.Solve_magnus_euler(H4, ds=0.5, smax=2.0, max_steps=4, mode=1, transform_thr=1e-8)
        t0 = time.perf_counter()
        _, e04, _, nc4 = ia.IMSRGSolver.Solve_magnus_euler(H4, ds=0.5, smax=0.5 * steps, eta_criterion=1e-6,
                                                           max_steps=steps, mode=1, transform_thr=1e-8)
        t4 = time.perf_counter() - t0
        small[name] = {"seconds": t4, "steps": len(e04), "commutators": int(np.sum(nc4)),
                       "commutators_per_s": float(np.sum(nc4)) / t4}
    ms4.set_graphs(False)
    res["emax4_flow"] = small
    n = min(len(res["fp64"]["E0"]), len(res["mixed"]["E0"]))
    drift = max(abs(a - b) for a, b in zip(res["fp64"]["E0"][:n], res["mixed"]["E0"][:n])) if n else None
    return {"metric": "Magnus IMSRG(2) flow at emax=10: E0(s) drift of the mixed-precision flow vs FP64",
            "value": drift, "unit": "max_s |E0_mixed(s) - E0_fp64(s)|", "n_gpus": 1, "higher_is_better": False,
            "dtype": "mixed: bf16x3 MFMA pp/hh + f32 MFMA ph + f64", "data": "synthetic Hamiltonian (seed 5001)",
            "config": {"workload": "Magnus-Euler flow, White generator, ds=0.5, BCH nested-norm cut 1e-8, "
                                   "HO emax=10, O16 reference", "max_steps": steps},
            "fp64": res["fp64"], "mixed": res["mixed"], "fp64_graphs": res["fp64_graphs"],
            "emax4_flow_eager_vs_graphs": res["emax4_flow"]}


def workload_normal_order(args, ia, np):
    """Operator::DoNormalOrdering of a scalar Hermitian operator at emax=10 (seed 6001,
    E0 = 0.5) into another operator: device time per call (the copy of the operator,
    its two-body norm, which the reference tests against 1e-7, and the two kernels),
    the oracle's time for the same call on one host core, and the difference of the
    two results."""
    emax = 10
    ms = ia.ModelSpace(emax, "O16", device=0)
    H = ia.DeviceOperator(ms, 1).fill_random(1, 6001)
    Hh = H.to_host()
    Hh.ZeroBody = 0.5
    H.upload(Hh)
    out = ia.DeviceOperator(ms, 1)
    for _ in range(3):
        H.DoNormalOrdering(out=out)
    ms.synchronize()
    reps = max(args.steps, 20)
    t0 = time.perf_counter()
    for _ in range(reps):
        H.DoNormalOrdering(out=out)
    ms.synchronize()
    us = (time.perf_counter() - t0) / reps * 1e6
    got = out.to_host()
    line = {"metric": f"normal ordering (Operator::DoNormalOrdering) of a scalar operator at emax={emax}",
            "value": us, "unit": "us/call", "n_gpus": 1, "steps": reps, "higher_is_better": False, "dtype": "f64",
            "data": "synthetic (seeded, library generator)",
            "config": {"workload": f"DoNormalOrdering, HO emax={emax}, O16 reference, {len(ms.orbits)} orbits"},
            "E0_after": float(got.ZeroBody)}
    if not args.no_cpu_baseline:
        from oracle.pyoracle import OracleModelSpace
        orc = OracleModelSpace(emax, 16, 8)
        op = {"zero": float(Hh.ZeroBody), "one": np.asarray(Hh.OneBody), "two": np.asarray(Hh.TwoBody), "herm": 1}
        samples = []
        for _ in range(5):
            t0 = time.perf_counter()
            ref = orc.normal_order(op, 1)
            samples.append(time.perf_counter() - t0)
        scale = max(np.max(np.abs(ref["one"])), 1.0)
        line["cpu_baseline"] = {"value": float(np.median(samples)) * 1e6, "unit": "us/call", "cores": 1,
                                "kind": "port", "sample": "median of 5 calls of the oracle's C restatement "
                                                          "of DoNormalOrdering2 on the same operator"}
        line["gpu_vs_cpu_rel_maxabs"] = float(np.max(np.abs(np.asarray(got.OneBody) - ref["one"])) / scale)
        line["gpu_vs_cpu_dE0"] = float(abs(got.ZeroBody - ref["zero"]))
    return line


def workload_transform(args, ia, np):
    """HartreeFock::TransformToHFBasis of a scalar Hermitian operator at emax=10 (seed
    6101) with a basis change that is orthogonal and block diagonal in (l, j, tz)
    (seed 6102, as a Hartree-Fock C): device time per call (C's upload, the channel
    matrices D, and C^T O C plus D^T V D per channel as rocBLAS DGEMMs), its FP64
    rate against the MFMA peak, and the oracle's time for the same call (its channel
    products through the host BLAS, one thread)."""
    emax = 10
    ms = ia.ModelSpace(emax, "O16", device=0)
    H = ia.DeviceOperator(ms, 1).fill_random(1, 6101)
    rng = np.random.default_rng(6102)
    Cm = np.eye(ms.norb)
    blocks = {}
    for i, o in enumerate(ms.orbits):
        blocks.setdefault((o["l"], o["j2"], o["tz2"]), []).append(i)
    for blk in blocks.values():
        Q, R = np.linalg.qr(rng.normal(size=(len(blk), len(blk))))
        Cm[np.ix_(blk, blk)] = Q * np.sign(np.diag(R))
    Cm = np.asfortranarray(Cm)
    out = ia.DeviceOperator(ms, 1)
    # the same C with one entry of 1e-300 between two blocks takes the dense DGEMM path
    Cdense = Cm.copy()
    Cdense[0, 1] = 1e-300

    def timed(Cx, reps):
        for _ in range(3):
            H.TransformToHFBasis(Cx, out=out)
        t0 = time.perf_counter()
        for _ in range(reps):
            H.TransformToHFBasis(Cx, out=out)
        return (time.perf_counter() - t0) / reps

    reps = max(args.steps, 10)
    sec_dense = timed(Cdense, reps)
    dense = np.asarray(out.to_host().TwoBody).copy()
    sec = timed(Cm, reps)
    flops = 4.0 * sum(c["n"] ** 3 for c in ms.channels) + 4.0 * ms.norb ** 3
    got = out.to_host()
    tb = 8.0 * ms.tbsize
    line = {"metric": f"basis transformation (HartreeFock::TransformToHFBasis) of a scalar operator at emax={emax}",
            "value": sec * 1e6, "unit": "us/call", "n_gpus": 1, "steps": reps, "higher_is_better": False,
            "dtype": "f64", "data": "synthetic (seeded operator and orthogonal block-diagonal C)",
            "config": {"workload": f"TransformToHFBasis, HO emax={emax}, O16 reference, {len(ms.orbits)} orbits"},
            "roofline": {"bound": "hbm", "achieved": 4 * tb / sec / 1e9, "peak": 8000.0, "unit": "GB/s",
                         "frac": 4 * tb / sec / 1e9 / 8000.0, "traffic": None,
                         "note": "block-sparse path, whole call (C upload, D blocks, T = V D, D^T T, one-body): "
                                 "V read, T written and read, V' written = 4 x 8 B per two-body element"},
            "dense_dgemm_path_us": sec_dense * 1e6,
            "dense_dgemm_tflops": flops / sec_dense / 1e12,
            "dense_vs_sparse_rel_maxabs": float(np.max(np.abs(dense - np.asarray(got.TwoBody))) /
                                                np.max(np.abs(dense)))}
    if not args.no_cpu_baseline:
        from oracle import pyoracle
        blas = pyoracle.use_blas(True)
        orc = pyoracle.OracleModelSpace(emax, 16, 8)
        Hh = H.to_host()
        op = {"zero": float(Hh.ZeroBody), "one": np.asarray(Hh.OneBody), "two": np.asarray(Hh.TwoBody), "herm": 1}
        t0 = time.perf_counter()
        ref = orc.transform_basis(op, Cm)
        cpu = time.perf_counter() - t0
        pyoracle.use_blas(False)
        line["cpu_baseline"] = {"value": cpu * 1e6, "unit": "us/call", "cores": 1, "kind": "port",
                                "sample": f"one call of the oracle's restatement, channel products through {blas} "
                                          "on one thread"}
        two = np.asarray(got.TwoBody)
        line["gpu_vs_cpu_rel_maxabs"] = float(max(
            np.max(np.abs(two - ref["two"])) / np.max(np.abs(ref["two"])),
            np.max(np.abs(np.asarray(got.OneBody) - ref["one"])) / np.max(np.abs(ref["one"]))))
    return line


def workload_hf(args, ia, np):
    """HartreeFock::Solve + GetNormalOrderedH at emax=10 for the synthetic flow
    Hamiltonian (one-body 10(2n+l+3/2) plus N(0, 0.5) inside each (l, j, tz) block,
    two-body (R+R^T)/sqrt(2n), seed 7300): device monopoles, host iteration with the
    reference's density mixing, then the normal-ordered Hamiltonian in the HF basis.
    Beside it the oracle's restatement of Solve (numpy, its monopoles through the
    oracle's C GetTBMEmonopole) on one host core."""
    emax = 10
    ms = ia.ModelSpace(emax, "O16", device=0)
    rng = np.random.default_rng(7300)
    norb = ms.norb
    one = np.zeros((norb, norb), order="F")
    blocks = {}
    for i, o in enumerate(ms.orbits):
        blocks.setdefault((o["l"], o["j2"], o["tz2"]), []).append(i)
    for blk in blocks.values():
        for a in blk:
            o = ms.orbits[a]
            one[a, a] = 10.0 * (2 * o["n"] + o["l"] + 1.5)
        for x in range(len(blk)):
            for y in range(x + 1, len(blk)):
                one[blk[x], blk[y]] = one[blk[y], blk[x]] = rng.normal(0.0, 0.5)
    two = np.zeros(ms.tbsize)
    for c in ms.channels:
        n = c["n"]
        if n:
            R = rng.normal(size=(n, n)) / np.sqrt(n)
            two[c["off"]:c["off"] + n * n] = ((R + R.T) / np.sqrt(2.0)).reshape(-1, order="F")
    H = ia.Operator(ms, 1, 0.0, one, two).to_device()
    out = ia.DeviceOperator(ms, 1)
    hf = ia.HartreeFock(H).Solve()
    hf.GetNormalOrderedH(out=out)
    reps = max(min(args.steps, 10), 3)
    t0 = time.perf_counter()
    for _ in range(reps):
        hf = ia.HartreeFock(H).Solve()
    t_solve = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        hf.GetNormalOrderedH(out=out)
    t_no = (time.perf_counter() - t0) / reps
    line = {"metric": f"Hartree-Fock (HartreeFock::Solve + GetNormalOrderedH) at emax={emax}",
            "value": (t_solve + t_no) * 1e3, "unit": "ms/call", "n_gpus": 1, "steps": reps,
            "higher_is_better": False, "dtype": "f64", "data": "synthetic Hamiltonian (seed 7300)",
            "config": {"workload": f"HF with frozen occupations, HO emax={emax}, O16 reference, {norb} orbits"},
            "solve_ms": t_solve * 1e3, "normal_ordered_h_ms": t_no * 1e3, "iterations": hf.iterations,
            "converged": hf.converged, "EHF": hf.EHF}
    if not args.no_cpu_baseline:
        from oracle.pyoracle import OracleModelSpace
        orc = OracleModelSpace(emax, 16, 8)
        op = {"zero": 0.0, "one": one, "two": two, "herm": 1}
        t0 = time.perf_counter()
        ref = orc.hf_solve(op)
        cpu = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": cpu * 1e3, "unit": "ms/call (Solve only)", "cores": 1, "kind": "port",
                                "sample": "one call of the oracle's Solve (numpy and Python loops, monopoles "
                                          "through its C GetTBMEmonopole); not a tuned CPU code"}
        line["gpu_vs_cpu_dEHF"] = float(abs(hf.EHF - ref["EHF"]))
        line["gpu_vs_cpu_C_maxabs"] = float(np.max(np.abs(hf.C - ref["C"])))
        line["same_iterations"] = hf.iterations == ref["iterations"]
    return line


def me2j_record_count(emax, np):
    """Records (a b c d J) of a Darmstadt me2j file to emax with E2max = 2 emax: pairs
    (nlj1 >= nlj2) in file order, every pair (nlj3, nlj4) up to it in that order with
    the same parity, one record per common J (src/ReadWrite.cc:877-935)."""
    orb = []
    for e in range(emax + 1):
        for l in range(e % 2, e + 1, 2):
            for tj in range(abs(2 * l - 1), 2 * l + 2, 2):
                orb.append((l, tj))
    pairs = [(i1, i2) for i1 in range(len(orb)) for i2 in range(i1 + 1)]
    par = np.array([(orb[a][0] + orb[b][0]) % 2 for a, b in pairs])
    jmin = np.array([abs(orb[a][1] - orb[b][1]) // 2 for a, b in pairs])
    jmax = np.array([(orb[a][1] + orb[b][1]) // 2 for a, b in pairs])
    nJ = np.minimum(jmax[:, None], jmax[None, :]) - np.maximum(jmin[:, None], jmin[None, :]) + 1
    ok = (par[:, None] == par[None, :]) & (nJ > 0)
    return int(np.sum(np.tril(np.where(ok, nJ, 0))))


def workload_pipeline(args, ia, np):
    """From an interaction file to the IMSRG(2) ground-state energy at emax=10: a
    synthetic charge-independent me2j file (binary, N(0, 0.1^2) elements, seed 7400)
    with the one-body part 10(2n+l+3/2) on the diagonal, read
    (ReadBareTBME_Darmstadt), Hartree-Fock (Solve), GetNormalOrderedH, then the
    Magnus-Euler flow (White generator, ds = 0.5, BCH nested-norm cut 1e-8, mixed
    precision) to ||eta|| < 1e-6.  Wall time of each stage; the file is written
    before the timing."""
    import tempfile
    emax = 10
    ms = ia.ModelSpace(emax, "O16", device=0)
    nrec = me2j_record_count(emax, np)
    rng = np.random.default_rng(7400)
    vals = (0.1 * rng.normal(size=(nrec, 4))).astype(np.float32)
    vals[:, 2] = vals[:, 1]
    vals[:, 3] = vals[:, 1]
    one = np.asfortranarray(np.diag([10.0 * (2 * o["n"] + o["l"] + 1.5) for o in ms.orbits]))
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "H.me2j.bin")
        with open(path, "wb") as f:
            f.write(b"synthetic me2j".ljust(255, b" "))
            f.write(vals.astype("<f4").tobytes())
        H = ia.DeviceOperator(ms, 1)
        H.upload(ia.Operator(ms, 1, 0.0, one, np.zeros(ms.tbsize)))
        ms.synchronize()
        t0 = time.perf_counter()
        n = ia.ReadWrite().ReadBareTBME_Darmstadt(path, H, emax)
        t1 = time.perf_counter()
    hf = ia.HartreeFock(H).Solve()
    t2 = time.perf_counter()
    HNO = hf.GetNormalOrderedH()
    t3 = time.perf_counter()
    Hs, e0, en, nc = ia.IMSRGSolver.Solve_magnus_euler(HNO, ds=0.5, smax=50.0, eta_criterion=1e-6, max_steps=60,
                                                       mode=1, transform_thr=1e-8, precision=ia.PREC_MIXED)
    ms.synchronize()
    t4 = time.perf_counter()
    return {"metric": f"interaction file to IMSRG(2) ground-state energy at emax={emax}", "value": t4 - t0,
            "unit": "s", "n_gpus": 1, "higher_is_better": False,
            "dtype": "mixed: bf16x3 MFMA pp/hh + f32 MFMA ph + f64 (flow); f64 (reader, HF, normal ordering)",
            "data": f"synthetic charge-independent me2j file, {n} records (seed 7400)",
            "config": {"workload": f"ReadBareTBME_Darmstadt -> HartreeFock -> GetNormalOrderedH -> Magnus-Euler "
                                   f"flow, HO emax={emax}, O16 reference"},
            "read_s": t1 - t0, "hf_solve_s": t2 - t1, "hf_iterations": hf.iterations, "EHF": hf.EHF,
            "normal_ordered_h_s": t3 - t2, "flow_s": t4 - t3, "flow_steps": int(len(e0)),
            "flow_commutators": int(np.sum(nc)), "eta_norm_final": float(en[len(e0)]), "E0_final": float(e0[-1])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--emax", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--autotune-cache", default=None,
                    help="text cache of the autotuner: settings found there are applied without timing trials")
    ap.add_argument("--workload", default="commutator",
                    choices=["commutator", "batched", "emax12", "emax14", "flow", "normal_order", "transform", "hf",
                             "pipeline"],
                    help="commutator: the headline metric line; batched: 48 independent pairs at emax 4-7; "
                         "emax12 / emax14: precision report at emax=12 / 14; flow: Magnus flow at emax=10; "
                         "normal_order: DoNormalOrdering at emax=10; transform: TransformToHFBasis at emax=10; "
                         "hf: HartreeFock Solve + GetNormalOrderedH at emax=10; pipeline: me2j file -> HF -> "
                         "normal ordering -> Magnus flow at emax=10")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one MI355X; runs on more than one GPU are not part of this build "
              "(independent operators shard one process per GPU, see the sharding notes)", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import imsrg_amd as ia

    if args.workload != "commutator":
        line = {"batched": workload_batched, "emax12": workload_emax12,
                "emax14": lambda a, i, n: workload_emax12(a, i, n, emax=14), "flow": workload_flow,
                "normal_order": workload_normal_order, "transform": workload_transform,
                "hf": workload_hf, "pipeline": workload_pipeline}[args.workload](
            args, ia, np)
        print(json.dumps(line), flush=True)
        return

    emax = args.emax
    ms = ia.ModelSpace(emax, "O16", device=0)
    X = ia.DeviceOperator(ms, -1).fill_random(-1, 10001)
    Y = ia.DeviceOperator(ms, 1).fill_random(1, 10002)
    Z = ia.DeviceOperator(ms, 1)

    # plain device-to-device copy of one operator buffer (known size), used to
    # calibrate the FETCH_SIZE / WRITE_SIZE counters of profiling runs
    Zc = ia.DeviceOperator(ms, 1)
    Zc.copy_from(Y)
    ms.synchronize()

    def timed(precision):
        """warm-up, then the timed steps; the library records HIP events around
        every phase of the timed steps on the stream that phase runs on, with no
        host synchronisation per step, and averages them afterwards; returns the
        rate, the per-phase averages (us) and the result on the host"""
        for _ in range(args.warmup):
            ia.Commutator.CommutatorScalarScalar(X, Y, precision=precision, out=Z)
        ms.synchronize()
        ms.set_profiling(True)
        ms.phase_times()  # nothing pending from earlier calls
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ia.Commutator.CommutatorScalarScalar(X, Y, precision=precision, out=Z)
        ms.synchronize()
        t = time.perf_counter() - t0
        phases = np.array(ms.phase_times())
        ms.set_profiling(False)
        return args.steps / t, t, phases, Z.to_host()

    # launch settings chosen by the autotuner on this basis, per precision (set-up, untimed)
    tuning = {"fp64": ms.autotune(ia.PREC_FP64, args.autotune_cache)}
    fp64_rate, t_fp64, phase64, Zref = timed(ia.PREC_FP64)
    tuning["mixed"] = ms.autotune(ia.PREC_MIXED, args.autotune_cache)
    mixed_rate, t_mixed, phasemx, Zmx = timed(ia.PREC_MIXED)

    # host-buffer drop-in (operators in host memory, PCIe transfers included);
    # reported beside the device-resident rate, never as the value
    hX, hY = X.to_host(), Y.to_host()
    ia.Commutator.CommutatorScalarScalar(hX, hY)
    nh = 3
    t0 = time.perf_counter()
    for _ in range(nh):
        ia.Commutator.CommutatorScalarScalar(hX, hY)
    host_rate = nh / (time.perf_counter() - t0)
    scale = max(np.max(np.abs(Zref.TwoBody)), np.max(np.abs(Zref.OneBody)))
    rel_dz = float(max(np.max(np.abs(Zmx.TwoBody - Zref.TwoBody)), np.max(np.abs(Zmx.OneBody - Zref.OneBody)))
                   / scale)
    de0 = abs(Zmx.ZeroBody - Zref.ZeroBody)
    mixed_ok = rel_dz < 1e-7 and de0 < 1e-6
    # the headline value is the mixed-precision rate when it meets the accuracy
    # bar, else the FP64 rate
    use_mixed = mixed_ok
    rate = mixed_rate if use_mixed else fp64_rate
    t_run = t_mixed if use_mixed else t_fp64
    phase_avg = phasemx if use_mixed else phase64

    # the same phases run one at a time (pp/hh GEMM not overlapped): each
    # kernel's own duration without a concurrent kernel on the GPU
    def isolated(precision, tun):
        ms.set_tuning(**{**{k: v for k, v in tun.items() if k != "us_per_commutator"}, "overlap_pphh": 0})
        ia.Commutator.CommutatorScalarScalar(X, Y, precision=precision, out=Z)
        ms.synchronize()
        ms.set_profiling(True)
        ms.phase_times()
        for _ in range(max(3, min(args.steps, 5))):
            ia.Commutator.CommutatorScalarScalar(X, Y, precision=precision, out=Z)
        ph = np.array(ms.phase_times())
        ms.set_profiling(False)
        ms.set_tuning(**{k: v for k, v in tun.items() if k != "us_per_commutator"})
        return ph

    tun = tuning["mixed" if use_mixed else "fp64"]
    phase_iso = isolated(ia.PREC_MIXED if use_mixed else ia.PREC_FP64, tun)

    # ---- roofline of the dominant kernel of the reported run ----
    # (phase durations are HIP events recorded around each launch during the
    # timed steps, on the stream the kernel runs on; with the pp/hh GEMM
    # overlapping other phases, durations include the sharing of the GPU)
    FA, execA, FC, BI64, BI32 = algorithmic_counts(ms)
    # the FP32-product ph GEMM (mixed_ph_fp32 = 1) leaves Zbar in FP32
    zbar32 = use_mixed and tun["mixed_ph_fp32"] == 1
    BI = BI32 if zbar32 else BI64
    dom = int(np.argmax(phase_avg))
    t_dom = phase_avg[dom] * 1e-6
    if PHASES[dom] == "pphh_gemm":
        # the kernel that ran: under mixed precision the tuning picks bf16x3 (2),
        # FP32 (1) or FP64 (0) MFMA.  The bf16x3 kernel's peak for the FP64
        # algorithm's flops is the bf16 matrix peak over its six products.
        fam = tun["mixed_pphh_fp32"] if use_mixed else 0
        peak = {0: PEAK_FP64_MFMA, 1: PEAK_FP32_MFMA, 2: PEAK_BF16_MFMA / B3_PRODUCTS}[fam]
        kname = {0: "k_pphh_big_f64", 1: "k_pphh_big_mixed", 2: "k_pphh_big_b3p"}[fam]
        ach = FA / t_dom
        roof = {"kernel": ("k_pack_pphh + " if fam == 2 else "") + kname +
                          (" + k_syr2k<0> (channels n <= %d)" % tun["pphh_small_tile_max"]
                           if tun["pphh_small_tile_max"] > 0 else ""),
                "bound": "mfma", "achieved": ach / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                "frac": ach / peak,
                "traffic": (profiled_traffic(kname) if fam != 2 else
                            _sum_or_none(profiled_traffic("k_pack_pphh"), profiled_traffic(kname))),
                "algorithmic_flops": FA,
                "executed_flops": execA}
    elif PHASES[dom] == "inverse_pandya":
        ach = BI / t_dom
        kname = "k_invpandya<float>" if zbar32 else "k_invpandya<double>"
        roof = {"kernel": kname, "bound": "hbm", "achieved": ach / 1e9, "peak": PEAK_HBM / 1e9,
                "unit": "GB/s", "frac": ach / PEAK_HBM, "traffic": profiled_traffic(kname),
                "algorithmic_bytes": BI}
    else:
        roof = {"kernel": PHASES[dom], "bound": "hbm", "achieved": None, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                "frac": None, "traffic": None}
    roof["launch_us"] = phase_avg[dom]
    if roof["achieved"] is not None:  # the same kernel run alone (no overlapping kernel)
        t_iso = phase_iso[dom] * 1e-6
        work = FA if PHASES[dom] == "pphh_gemm" else BI
        ach_iso = work / t_iso / (1e12 if roof["unit"] == "TFLOP/s" else 1e9)
        roof["isolated"] = {"launch_us": phase_iso[dom], "achieved": ach_iso, "frac": ach_iso / roof["peak"]}
        if roof.get("traffic"):
            # the profiled HBM bytes per launch over the same launch run alone:
            # how close the kernel's actual traffic comes to the HBM peak
            roof["isolated"]["hbm_traffic_GBps"] = roof["traffic"] / t_iso / 1e9
            roof["isolated"]["hbm_traffic_frac"] = roof["traffic"] / t_iso / PEAK_HBM

    line = {
        "metric": "[X,Y] commutators/sec at emax=10: mixed vs FP64 GPU, at |dE0|<1e-6",
        "value": rate,
        "unit": "commutators/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_run / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ({2: "bf16x3 MFMA products + f32/f64 accumulation", 1: "f32 MFMA products + f64 accumulation",
                   0: "f64"}[tuning["mixed"]["mixed_pphh_fp32"]] if use_mixed else "f64"),
        "data": "synthetic (seeded N(0,1)/sqrt(dim) operators, X anti-Hermitian, Y Hermitian)",
        "config": {"workload": f"scalar IMSRG(2) commutator [X,Y], HO emax={emax}, O16 reference, "
                               f"{ms.norb} orbits", "emax": emax, "norb": ms.norb,
                   "two_body_doubles": ms.tbsize, "precision": "mixed" if use_mixed else "fp64"},
        "fp64_commutators_per_s": fp64_rate,
        "mixed_commutators_per_s": mixed_rate,
        "mixed_vs_fp64": {"speedup": mixed_rate / fp64_rate, "max_abs_dZ_rel": rel_dz, "abs_dE0": de0,
                          "meets_bar": mixed_ok},
        "host_buffer_fp64_commutators_per_s": host_rate,
        "tuning": tuning,
        "phase_us": dict(zip(PHASES, [round(x, 2) for x in phase_avg])),
        "phase_us_fp64": dict(zip(PHASES, [round(x, 2) for x in phase64])),
        "phase_us_isolated": dict(zip(PHASES, [round(x, 2) for x in phase_iso])),
        "roofline": roof,
    }

    # ---- CPU baseline: the oracle on the same workload, on this host ----
    # The oracle restates the reference's OpenMP path; its channel GEMMs go
    # through BLAS dgemm (scipy's OpenBLAS, one thread per OpenMP thread) as the
    # reference's go through Armadillo.  One warm-up, then the median of five
    # commutators, with per-phase times under the reference's timer names.
    if not args.no_cpu_baseline:
        try:
            from oracle import pyoracle
            blas = pyoracle.use_blas(True)
            orc = pyoracle.OracleModelSpace(emax, 16, 8)
            Xh, Yh = X.to_host(), Y.to_host()
            xo = {"zero": Xh.ZeroBody, "one": Xh.OneBody, "two": Xh.TwoBody, "herm": -1}
            yo = {"zero": Yh.ZeroBody, "one": Yh.OneBody, "two": Yh.TwoBody, "herm": 1}
            R = orc.commutator(xo, yo)
            pyoracle.timers(reset=True)
            samples = []
            for _ in range(5):
                t0 = time.perf_counter()
                R = orc.commutator(xo, yo)
                samples.append(time.perf_counter() - t0)
            ncalls, phase_s = pyoracle.timers(reset=True)
            tc = float(np.median(samples))
            scale = max(np.max(np.abs(R["two"])), np.max(np.abs(R["one"])))
            err = float(max(np.max(np.abs(Zref.TwoBody - R["two"])), np.max(np.abs(Zref.OneBody - R["one"]))) / scale)
            line["cpu_baseline"] = {
                "value": 1.0 / tc, "unit": "commutators/s", "cores": pyoracle.lib().orc_num_threads(),
                "kind": "port",
                "sample": f"median of 5 FP64 commutators of the same emax={emax} workload (after one warm-up) by "
                          "the OpenMP C restatement of the reference path (oracle/), channel GEMMs through BLAS "
                          f"dgemm ({blas or 'plain loops: scipy BLAS unavailable'})",
                "samples_s": [round(x, 4) for x in samples],
                "gemm": f"BLAS dgemm, {blas}" if blas else "plain loops",
                "phase_us": {k: round(1e6 * v / max(ncalls, 1), 1) for k, v in phase_s.items()},
                "gpu_vs_cpu_rel_maxabs": err,
                "gpu_vs_cpu_dE0": abs(Zref.ZeroBody - R["zero"])}
        except Exception as e:  # the baseline is reported, never the measured value
            line["cpu_baseline"] = {"value": None, "unit": "commutators/s", "cores": None, "kind": "port",
                                    "sample": f"failed: {e}"}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
