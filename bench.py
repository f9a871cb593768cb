"""Benchmark of the fused fold pipeline on one MI355X.

Workload: causal attention written as a fold pipeline (scores = SUM of
q*k grouped by h, tq, tk with tq >= tk, softmax-weighted sum of v), 8 heads,
T = 4096, d = 64, q and k ~ N(0,1)/8, v and the upstream gradient ~ N(0,1),
synthetic data from a seed, inputs resident in HBM.  A step is one forward
plus backward pass (dq, dk, dv of SUM(out * g)) in the mixed form: BF16x3
MFMA products with FP32 accumulation.  The same step in plain FP32 (FP32
MFMA) is timed beside it, and both are checked against the FP64 GPU form.

value = algorithmic GFLOP/s of the mixed step: 14 * H * d * P FLOP for P
allowed (query, key) pairs per head (forward 2 products, backward 5).
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

H, T, D = 8, 4096, 64
SCALE = 1.0 / 8.0
PEAK_BF16 = 2.5e15       # MI355X dense BF16 MFMA, FLOP/s
PEAK_FP32 = 157.3e12
# HBM bytes one launch of the dK/dV kernel (BF16x3, d = 64) moves at this
# workload: rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 KiB, the factors a
# 256 MiB device copy reads back (profiles/attention_causal_h8_t4096_d64_hbm_traffic.md).
DKDV_TRAFFIC_BYTES = (29260.7 * 2.0 + 32768.0) * 1024


def pairs_per_head(t: int) -> int:
    return t * (t + 1) // 2


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-secondary", action="store_true",
                    help="skip the secondary lines (dense fold, ranking, batched and large folds, smoother, "
                         "regridding, chain, sparse product, attention over Arrow, gradient memory)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run one plain 256 MiB device copy first (a known-size kernel for counter scaling)")
    ap.add_argument("--line", choices=["dense_fold", "ranking", "batched_folds", "large_fold", "smoother",
                                       "regridding", "chain", "sparse_product", "attention_arrow",
                                       "gradient_memory"],
                    help="print one secondary line alone (for profiling one workload) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the library has no multi-GPU path (asked for {args.gpus})",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from einfold_amd import _lib, data, device
    from einfold_amd._lib import lib, check

    dev = torch.device("cuda:0")
    check(lib().ef_init(0))
    if args.copy_probe:
        src = torch.empty(256 << 20, dtype=torch.uint8, device=dev).fill_(1)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.line:
        one = {"dense_fold": lambda: dense_fold_line(dev, not args.no_cpu), "ranking": lambda: ranking_line(dev, not args.no_cpu),
               "batched_folds": lambda: batched_folds_line(dev), "large_fold": lambda: large_fold_line(dev),
               "smoother": smoother_line,
               "regridding": regrid_line, "chain": lambda: chain_line(not args.no_cpu),
               "sparse_product": lambda: sparse_product_line(not args.no_cpu),
               "attention_arrow": attention_arrow_line,
               "gradient_memory": lambda: gradient_memory_line(dev)}[args.line]
        print(json.dumps({args.line: one()}), flush=True)
        return
    q64, k64, v64 = data.attention_tables(H, T, D, seed=3)
    g64 = data.normal(3, 4, H * T * D).reshape(H, T, D)
    to = lambda x, dt: torch.from_numpy(np.ascontiguousarray(x)).to(dt).to(dev)
    q32, k32, v32, g32 = (to(x, torch.float32) for x in (q64, k64, v64, g64))
    qd, kd, vd, gd = (to(x, torch.float64) for x in (q64, k64, v64, g64))
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def make_desc(precision):
        desc = _lib.CAttentionDesc()
        desc.batches, desc.queries, desc.keys, desc.features, desc.values = H, T, T, D, D
        desc.scale = SCALE
        desc.allowed = None
        desc.causal, desc.causal_offset, desc.window = 1, 0, 0
        desc.precision = _lib.PRECISION[precision]
        return desc

    def buffers(qx, dtype):
        return dict(out=torch.empty_like(qx), m=torch.empty((H, T), dtype=dtype, device=dev),
                    l=torch.empty((H, T), dtype=dtype, device=dev),
                    dq=torch.empty_like(qx), dk=torch.empty_like(qx), dv=torch.empty_like(qx))

    P = lambda t: ctypes.c_void_p(t.data_ptr())

    def step(desc, q, k, v, g, bf):
        """One training step of the pipeline: the forward (out and the row
        statistics), then the backward (dq, dk, dv) as one entry point."""
        L = lib()
        if "scratch" not in bf:
            bf["scratch"] = device.attention_scratch(desc, dev)
        sc = P(bf["scratch"])
        check(L.ef_attention_fwd(ctypes.byref(desc), P(q), P(k), P(v), P(bf["out"]), P(bf["m"]),
                                 P(bf["l"]), None, sc, stream))
        check(L.ef_attention_bwd(ctypes.byref(desc), P(q), P(k), P(v), P(bf["out"]), P(g), P(bf["m"]),
                                 P(bf["l"]), P(bf["dq"]), P(bf["dk"]), P(bf["dv"]), sc, stream))

    def kv_stage_ms(desc, q, k, v, g, bf, reps):
        """The dK/dV stage alone (its counter reset, the kernel and the sum of
        its parts), after a full step filled the scratch: the roofline's
        kernel, timed by events around ef_attention_bwd_kv."""
        L = lib()
        sc = P(bf["scratch"])
        check(L.ef_attention_bwd_prepare(ctypes.byref(desc), P(q), P(k), P(v), P(bf["out"]), P(g), sc, stream))
        run = lambda: check(L.ef_attention_bwd_kv(ctypes.byref(desc), P(k), P(v), P(bf["m"]), P(bf["l"]), sc,
                                                  P(bf["dk"]), P(bf["dv"]), stream))
        return _events(run, reps)

    def timed(desc, q, k, v, g, bf, steps, warmup, kv_time=False):
        for _ in range(warmup):
            step(desc, q, k, v, g, bf)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record()
        for i in range(steps):
            step(desc, q, k, v, g, bf)
        stop.record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ms = start.elapsed_time(stop) / steps
        kv_ms = kv_stage_ms(desc, q, k, v, g, bf, steps) if kv_time else None
        return ms, kv_ms, wall

    # FP64 reference (the GPU's FP64 form), once.
    ref = buffers(qd, torch.float64)
    step(make_desc("fp64"), qd, kd, vd, gd, ref)
    torch.cuda.synchronize()

    def rel(x, r):
        return float(torch.linalg.vector_norm(x.double() - r) / torch.linalg.vector_norm(r))

    def errors(bf):
        return {name: rel(bf[name], ref[name]) for name in ("out", "dq", "dk", "dv")}

    mixed = buffers(q32, torch.float32)
    mixed_desc = make_desc("bf16x3")
    ms, kv_ms, wall = timed(mixed_desc, q32, k32, v32, g32, mixed, args.steps, args.warmup, kv_time=True)
    mixed_err = errors(mixed)
    plain = buffers(q32, torch.float32)
    plain_desc = make_desc("fp32")
    fp32_ms, fp32_kv_ms, _ = timed(plain_desc, q32, k32, v32, g32, plain, args.steps, args.warmup, kv_time=True)
    plain_err = errors(plain)

    pairs = H * pairs_per_head(T)
    flop_step = 14 * D * pairs
    flop_kv = 8 * D * pairs            # S, dP, dV, dK products of the dK/dV kernel
    value = flop_step / (ms * 1e-3) / 1e9
    achieved = flop_kv / (kv_ms * 1e-3)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline()
    secondary = None
    if not args.no_secondary:
        secondary = {"dense_fold": dense_fold_line(dev, not args.no_cpu), "ranking": ranking_line(dev, not args.no_cpu),
                     "batched_folds": batched_folds_line(dev), "large_fold": large_fold_line(dev),
                     "smoother": smoother_line(),
                     "regridding": regrid_line(), "chain": chain_line(not args.no_cpu),
                     "sparse_product": sparse_product_line(not args.no_cpu),
                     "attention_arrow": attention_arrow_line(),
                     "gradient_memory": gradient_memory_line(dev)}

    line = {
        "metric": "fold GFLOP/s + attention fwd+bwd ms: mixed vs plain FP32, 1 MI355X",
        "value": round(value, 1),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x3 (fp32 accumulate)",
        "data": "synthetic (SplitMix64 seed 3), inputs resident in HBM",
        "config": {"workload": "causal attention fold pipeline fwd+bwd, 8 heads x T=4096 x d=64",
                   "heads": H, "seq_len": T, "head_dim": D, "mask": "causal", "parallelism": "none"},
        "attention_fwd_bwd_ms": {"mixed_bf16x3": round(ms, 4), "plain_fp32": round(fp32_ms, 4),
                                 "speedup": round(fp32_ms / ms, 3)},
        "rel_error_vs_fp64": {"mixed_bf16x3": mixed_err, "plain_fp32": plain_err},
        "roofline": {"kernel": "attn_bwd_dkdv_kernel<bf16x3, 64>", "bound": "mfma",
                     "achieved": round(achieved / 1e12, 2), "peak": PEAK_BF16 / 1e12,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_BF16, 4),
                     "issued_frac": round(3 * achieved / PEAK_BF16, 4),
                     "avg_ms": round(kv_ms, 4), "traffic": round(DKDV_TRAFFIC_BYTES),
                     "traffic_note": "HBM bytes per launch from rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-calibrated",
                     "fp32_kernel": {"avg_ms": round(fp32_kv_ms, 4),
                                     "frac_of_fp32_peak": round(flop_kv / (fp32_kv_ms * 1e-3) / PEAK_FP32, 4)}},
        "cpu_baseline": cpu,
        "secondary": secondary,
    }
    print(json.dumps(line), flush=True)


def _events(fn, reps, stream_fn=None):
    """Average ms of fn() over reps, by HIP events on torch's stream."""
    import torch
    fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def chain_line(with_cpu: bool):
    """A chain of three dense products over Arrow (kernel API fold_chain):
    A[i, j] B[j, k] C[k, l] with i = k = 256, j = l = 1024, every cell, FP64
    values U[-1, 1]; (A B) C is cheaper, 2 * 2 * 256^2 * 1024 FLOP in the
    two dense contractions (exact reference order, FP64 VALU).  Arrow to
    Arrow on the device, and the oracle's chain (the same two folds, its C
    contraction) on one core, whose result the device's must equal bit for
    bit."""
    import pyarrow as pa
    from einfold_amd import data, ir, kernels

    def grid(n0, n1, salt, names):
        v = data.unit(17, salt, n0 * n1) * 2 - 1
        return pa.table({names[0]: np.repeat(np.arange(n0), n1), names[1]: np.tile(np.arange(n1), n0),
                         names[2]: v})
    ni, nj, nk, nl = 256, 1024, 256, 1024
    a, b, c = grid(ni, nj, 1, "ijv"), grid(nj, nk, 2, "jkv"), grid(nk, nl, 3, "klv")
    EQ = ir.KeyEquality.EQUAL
    fold = ir.Fold([EQ] * 4, [ir.Operand((ir.Key(0, 0), ir.Key(1, 1)), 2),
                              ir.Operand((ir.Key(1, 0), ir.Key(2, 1)), 2),
                              ir.Operand((ir.Key(2, 0), ir.Key(3, 1)), 2)], [ir.Output(0), ir.Output(3)])
    out = kernels.fold_chain(fold, a, b, c)
    walls = []
    for _ in range(5):
        t0 = time.perf_counter()
        out = kernels.fold_chain(fold, a, b, c)
        walls.append(time.perf_counter() - t0)
    assert out is not None and sum(x.num_rows for x in out) == ni * nl
    flop = 2.0 * ni * nj * nk + 2.0 * ni * nk * nl
    ms = float(np.median(walls)) * 1e3
    line = {"workload": "chain A[256x1024] B[1024x256] C[256x1024] over Arrow (fold_chain), FP64, "
                        "(A B) C in exact dense order", "gflop": round(flop / 1e9, 3),
            "arrow_to_arrow_ms": round(ms, 2), "gflop_per_s_arrow_to_arrow": round(flop / (ms * 1e-3) / 1e9, 1)}
    if with_cpu:
        from oracle import reference as R
        sch = R.schema(fold, [a.schema, b.schema, c.schema])
        t0 = time.perf_counter()
        want = R.fold_chain(fold, [a.to_batches(), b.to_batches(), c.to_batches()],
                            [a.schema, b.schema, c.schema], sch)
        line["cpu_port_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
        line["cpu_port_cores"] = 1
        got_v = pa.Table.from_batches(out).column(2).to_numpy()
        want_v = pa.Table.from_batches(want).column(2).to_numpy()
        if not np.array_equal(got_v.view(np.uint64), want_v.view(np.uint64)):
            raise SystemExit("chain: device sums differ from the oracle's")
        line["bit_exact_vs_oracle"] = True
    return line


def dense_fold_line(dev, with_cpu: bool):
    """The plain dense matrix-product fold: a(i, j, v), b(j, k, v) over every
    cell of a 256 x 256 grid, Int64 keys, v ~ U[-1, 1]; 2 * 256^3 FLOP.  The
    contraction kernel alone (device-resident), the operator from Arrow
    batches to Arrow batches, and the oracle's C restatement of dense.rs's
    contraction on one core."""
    import torch
    from einfold_amd import data, device, ir, kernels
    n = 256
    a, b = data.grid_matmul(n)
    A = torch.from_numpy(a.column("v").to_numpy().reshape(n, n).copy()).to(dev)
    B = torch.from_numpy(b.column("v").to_numpy().reshape(1, n, n).copy()).to(dev)
    held = torch.full((n, n), 2, dtype=torch.int32, device=dev)
    flop = 2.0 * n ** 3
    kms = _events(lambda: device.dense_accumulate_f64(A, held, B, n), 50)
    EQ = ir.KeyEquality.EQUAL
    fold = ir.Fold([EQ] * 3, [ir.Operand((ir.Key(0, 0), ir.Key(1, 1)), 2),
                              ir.Operand((ir.Key(1, 0), ir.Key(2, 1)), 2)], [ir.Output(0), ir.Output(2)])
    kernels.fold_dense(fold, a, b)
    walls = []
    for _ in range(10):
        t0 = time.perf_counter()
        out = kernels.fold_dense(fold, a, b)
        walls.append(time.perf_counter() - t0)
    assert out is not None and sum(x.num_rows for x in out) == n * n
    line = {"workload": "dense fold 256^3 grid (a.i, b.k, SUM(a.v*b.v)), FP64, bit-exact reference order",
            "kernel_ms": round(kms, 4), "kernel_gflops": round(flop / kms / 1e6, 1),
            "arrow_to_arrow_ms": round(float(np.median(walls)) * 1e3, 3),
            "bound": "fp64 valu (exact order; no FMA contraction)",
            "frac_of_fp64_peak": round(flop / (kms * 1e-3) / 78.6e12, 4)}
    if with_cpu:
        import oracle
        Ah = a.column("v").to_numpy().reshape(n, n)
        Bh = b.column("v").to_numpy().reshape(n, n)
        hh = np.full((n, n), 2, np.uint8)
        t0 = time.perf_counter()
        reps = 5
        for _ in range(reps):
            oracle.dense_accumulate(Ah, hh, Bh, n, n, n, n)
        dt = (time.perf_counter() - t0) / reps
        line["cpu_port_ms"] = round(dt * 1e3, 3)
        line["cpu_port_cores"] = 1
        # The host plan the operator replaces (a hash join on j, then a hash
        # aggregate by (i, k)), stood in for by pyarrow's Acero engine on this
        # host's cores; it is not DataFusion.
        import pyarrow as pa
        import pyarrow.compute as pc
        plans = []
        for _ in range(3):
            t0 = time.perf_counter()
            j = a.join(b, keys="j", join_type="inner", left_suffix="_a", right_suffix="_b")
            j = j.append_column("p", pc.multiply(j.column("v_a"), j.column("v_b")))
            g = j.group_by(["i", "k"]).aggregate([("p", "sum")])
            plans.append(time.perf_counter() - t0)
        assert g.num_rows == n * n
        line["hash_plan_ms"] = round(float(np.median(plans)) * 1e3, 1)
        line["hash_plan"] = (f"pyarrow Acero hash join + hash aggregate ({j.num_rows} joined rows), "
                             f"{pa.cpu_count()} threads: a stand-in for DataFusion's plan, not DataFusion")
    return line


def batched_folds_line(dev):
    """64 independent dense Float32 folds, M, N and K each log-uniform in
    [32, 2048] (seed 9), in one persistent graph-replayed launch; the
    autotuner picks the tile shape.  FLOP = sum of 2 M N K."""
    import torch
    from einfold_amd import data, device
    shapes = data.log_uniform_shapes(64, 32, 2048, seed=9)
    g = torch.Generator().manual_seed(9)
    probs = [(torch.randn(m, k, generator=g).to(dev), torch.randn(k, n, generator=g).to(dev),
              torch.empty(m, n, device=dev)) for m, n, k in shapes]
    flop = float(sum(2.0 * m * n * k for m, n, k in shapes))
    line = {"workload": "64 uneven dense FP32 folds, M,N,K log-uniform in [32, 2048]", "gflop": round(flop / 1e9, 2)}
    for precision in ("bf16x3", "fp32"):
        plan = device.GroupedPlan(probs, precision)
        ms = _events(plan.run, 10)
        a, b, c = probs[0]
        err = float(torch.linalg.vector_norm(c.double() - a.double() @ b.double()) /
                    torch.linalg.vector_norm(a.double() @ b.double()))
        tf = flop / (ms * 1e-3) / 1e12
        peak = PEAK_BF16 if precision == "bf16x3" else PEAK_FP32
        line[precision] = {"ms": round(ms, 4), "tflops": round(tf, 1), "tile": plan.tile,
                           "frac_of_peak": round(tf / (peak / 1e12), 4),
                           "issued_frac": round((3 if precision == "bf16x3" else 1) * tf / (peak / 1e12), 4),
                           "rel_err_fold0": err}
        plan.close()
    return line


def large_fold_line(dev):
    """One large dense Float32 fold, 8192 x 8192 x 8192 (2 M N K = 1.1e12
    FLOP), through the grouped product with the tile shape its autotuner
    picks, in each precision: FP32 (FP32 MFMA), BF16x3 (three BF16 products)
    and BF16.  A run is the operand prep plus the GEMM.  Fractions are of
    the MFMA peak of the arithmetic issued."""
    import torch
    from einfold_amd import device
    n = 8192
    g = torch.Generator().manual_seed(4)
    a = torch.randn(n, n, generator=g).to(dev)
    b = torch.randn(n, n, generator=g).to(dev)
    c = torch.empty(n, n, device=dev)
    flop = 2.0 * n * n * n
    line = {"workload": "dense FP32 fold 8192 x 8192 x 8192", "tflop": round(flop / 1e12, 3)}
    rows = torch.arange(0, n, 1024, device=dev)
    want = a[rows].double() @ b.double()
    for precision, issued, peak in (("fp32", 1, PEAK_FP32), ("bf16x3", 3, PEAK_BF16), ("bf16", 1, PEAK_BF16)):
        plan = device.GroupedPlan([(a, b, c)], precision)
        ms = _events(plan.run, 5)
        err = float(torch.linalg.vector_norm(c[rows].double() - want) / torch.linalg.vector_norm(want))
        tf = flop / (ms * 1e-3) / 1e12
        line[precision] = {"ms": round(ms, 3), "tflops": round(tf, 1), "tile": plan.tile,
                           "frac_of_issued_peak": round(issued * tf / (peak / 1e12), 4),
                           "rel_err_8_rows": err}
        plan.close()
    # The vendor library's GEMM (torch.matmul, hipBLASLt) on the same
    # operands: what this GPU sustains on a plain product of this size, for
    # scale next to the published peaks.
    for name, dt in (("library_bf16", torch.bfloat16), ("library_fp32", torch.float32)):
        x, y = a.to(dt), b.to(dt)
        lib_ms = _events(lambda: torch.matmul(x, y), 5)
        line[name] = {"ms": round(lib_ms, 3), "tflops": round(flop / (lib_ms * 1e-3) / 1e12, 1)}
        del x, y
    return line


def smoother_line():
    """Gaussian kernel smoothing (config 5) as the fused pipeline's general
    FP64 path: scores -power(x - y, 2)/h over the cross join of x (20k rows)
    and y (20k rows), values U[0, 100] (seed 5), h = 4, weights exp with the
    kept maximum shift, normalised per kept group, applied to w(j) ~ N(0, 1).
    Arrow tables in host memory in, Arrow out (transfers included)."""
    import time
    import pyarrow as pa
    from einfold_amd import data, ir, kernels
    n = m = 20000
    xs = pa.table({"i": np.arange(n, dtype=np.int64), "x": data.unit(5, 1, n) * 100.0})
    ys = pa.table({"j": np.arange(m, dtype=np.int64), "y": data.unit(5, 2, m) * 100.0})
    ws = pa.table({"j": np.arange(m, dtype=np.int64), "w": data.normal(5, 3, m)})
    x = ir.Expression.col(0, 1)
    y = ir.Expression.col(1, 1)
    score = -((x - y) ** ir.Expression.lit(2.0)) / ir.Expression.lit(4.0)
    producer = ir.Fold([ir.KeyEquality.EQUAL] * 2, [ir.Operand((ir.Key(0, 0),), None),
                                                    ir.Operand((ir.Key(1, 0),), None)],
                       [ir.Output(0), ir.Output(1)], row=score)
    p = ir.Pipeline(producer, ir.Exp(ir.Scale(), ir.Shift.KEPT_MAXIMUM), [], ir.Operand((ir.Key(1, 0),), 1), [0],
                    ir.Normalizer.KEPT_GROUPS)
    out = kernels.pipeline(p, xs, ys, ws)
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        out = kernels.pipeline(p, xs, ys, ws)
        times.append((time.perf_counter() - t0) * 1e3)
    ms = float(np.median(times))
    rows = sum(b.num_rows for b in out.batches)
    return {"workload": "Gaussian smoother 20k x 20k, h=4 (general FP64 pipeline path)", "pairs": n * m,
            "rows_out": rows, "ms_arrow_to_arrow": round(ms, 2),
            "gpairs_per_s": round(n * m / (ms * 1e-3) / 1e9, 2)}


def regrid_line():
    """Regridding (config 5) as a sorting-path fold: out(dst, t) = SUM(w.w *
    x.v) FROM w JOIN x ON w.src = x.src GROUP BY dst, t, at the stated scale:
    1M dst cells with 16 random src each over 1M src cells, 365 time steps
    (16M weight rows, 365M value rows, 5.84e9 multiply-adds, 365M output
    groups).  Int32 keys, Float32 weights and values (accept_float32: the
    FP64 precision widens them exactly, so the sums are the reference's
    Float64 sums).  Arrow in host memory in, Arrow out (transfers included);
    the median of three calls.  Checked against the closed form
    sum over a cell's 16 (src, w) of w * x[src, t] in float64 on 4096
    sampled dst cells (all 365 steps each), within 16 eps times the sum of
    magnitudes."""
    import time
    import pyarrow as pa
    from einfold_amd import data, ir, kernels
    n_dst, per, n_src, steps = 1_000_000, 16, 1_000_000, 365
    dst = np.repeat(np.arange(n_dst, dtype=np.int32), per)
    src = (data.next_u64(7, 1, n_dst * per) % np.uint64(n_src)).astype(np.int32)
    wv = data.unit(7, 2, n_dst * per).astype(np.float32)
    w = pa.table({"dst": dst, "src": src, "w": wv})
    xv = data.normal(7, 3, n_src * steps).astype(np.float32)
    xs = pa.table({"src": np.repeat(np.arange(n_src, dtype=np.int32), steps),
                   "t": np.tile(np.arange(steps, dtype=np.int32), n_src), "v": xv})
    eq = ir.KeyEquality.EQUAL
    f = ir.Fold([eq] * 3, [ir.Operand((ir.Key(0, 0), ir.Key(1, 1)), 2), ir.Operand((ir.Key(1, 0), ir.Key(2, 1)), 2)],
                [ir.Output(0), ir.Output(2)])
    opt = kernels.Options(accept_float32=True)
    times = []
    out = None
    for _ in range(3):
        out = None
        t0 = time.perf_counter()
        out = kernels.fold(f, w, xs, options=opt)
        times.append((time.perf_counter() - t0) * 1e3)
    ms = float(np.median(times))
    t = pa.Table.from_batches(out.batches)
    rows = t.num_rows
    # Every (dst, t) group exists, ordered by dst then t: row dst * 365 + t.
    sample = np.sort(data.next_u64(7, 4, 4096) % np.uint64(n_dst)).astype(np.int64)
    at = (sample[:, None] * steps + np.arange(steps)[None, :]).reshape(-1)
    got_dst = t.column(0).take(pa.array(at)).to_numpy()
    got_t = t.column(1).take(pa.array(at)).to_numpy()
    got_v = t.column(2).take(pa.array(at)).to_numpy()
    X = xv.astype(np.float64).reshape(n_src, steps)
    cols = (sample[:, None] * per + np.arange(per)[None, :])
    S, Wt = src[cols], wv[cols].astype(np.float64)                    # [cells, 16]
    want = np.einsum("cs,cst->ct", Wt, X[S]).reshape(-1)
    mag = np.einsum("cs,cst->ct", np.abs(Wt), np.abs(X[S])).reshape(-1)
    keys_ok = bool(np.array_equal(got_dst, np.repeat(sample, steps)) and
                   np.array_equal(got_t, np.tile(np.arange(steps), len(sample))))
    err = float(np.max(np.abs(got_v - want) / np.maximum(mag, 1e-300)))
    fma = n_dst * per * steps
    return {"workload": "regridding 1M dst x 16 src over 1M src, 365 steps, Float32 factors (config 5)",
            "algorithm": out.algorithm, "fma": fma, "rows_out": rows, "ms_arrow_to_arrow": round(ms, 1),
            "ms_calls": [round(x, 1) for x in times],
            "gflop_per_s": round(2 * fma / (ms * 1e-3) / 1e9, 2),
            "sample_check": {"cells": int(len(sample)) * steps, "keys_equal": keys_ok,
                             "max_err_over_magnitude": err, "bound": 16 * float(np.finfo(float).eps),
                             "ok": keys_ok and err <= 16 * float(np.finfo(float).eps)}}


def sparse_product_line(with_cpu: bool):
    """Config 2's sorting-path product (radix ranking, counting sorts,
    Gustavson's product) through ef_fold, Arrow in and out: a(i, j, v) and
    b(j, k, v) with 100k rows each, j = UInt64 hashes of Zipf(1.1) ids over
    1M ids (as config 2), i and k = UInt64 hashes of 4096 ids each (config 2
    has 64k: at this row count its groups would be about all 1.5e8 joined
    pairs, 3.6 GB of Arrow), v ~ N(0, 1) as Float32 (accept_float32; the
    FP64 precision widens them exactly).  FLOP = 2 per joined pair.  With
    the CPU leg, the oracle's restatement on one core, whose groups and sums
    the device's must equal bit for bit."""
    import pyarrow as pa
    import pyarrow.compute as pc
    from einfold_amd import data, ir, kernels
    rows = 100_000
    a = pa.table({"i": pa.array(data.hash64(data.zipf_ids(8, 1, rows, 4096, 0.0), 0x11), pa.uint64()),
                  "j": pa.array(data.hash64(data.zipf_ids(8, 2, rows, 1 << 20, 1.1), 0x6A), pa.uint64()),
                  "v": pa.array(data.normal(8, 3, rows).astype(np.float32))})
    b = pa.table({"j": pa.array(data.hash64(data.zipf_ids(8, 4, rows, 1 << 20, 1.1), 0x6A), pa.uint64()),
                  "k": pa.array(data.hash64(data.zipf_ids(8, 5, rows, 4096, 0.0), 0x22), pa.uint64()),
                  "v": pa.array(data.normal(8, 6, rows).astype(np.float32))})
    EQ = ir.KeyEquality.EQUAL
    f = ir.Fold([EQ] * 3, [ir.Operand((ir.Key(0, 0), ir.Key(1, 1)), 2), ir.Operand((ir.Key(1, 0), ir.Key(2, 1)), 2)],
                [ir.Output(0), ir.Output(2)])
    opts = kernels.Options(accept_float32=True)
    out = kernels.fold(f, a.to_batches(max_chunksize=65536), b.to_batches(max_chunksize=65536), options=opts)
    walls = []
    for _ in range(3):
        t0 = time.perf_counter()
        out = kernels.fold(f, a.to_batches(max_chunksize=65536), b.to_batches(max_chunksize=65536), options=opts)
        walls.append(time.perf_counter() - t0)
    ja, jb = a.column(1).to_numpy(), b.column(0).to_numpy()
    ua, ca = np.unique(ja, return_counts=True)
    ub, cb = np.unique(jb, return_counts=True)
    common, ia, ib = np.intersect1d(ua, ub, return_indices=True)
    pairs = int((ca[ia].astype(np.int64) * cb[ib]).sum())
    groups = sum(x.num_rows for x in out.batches)
    ms = float(np.median(walls)) * 1e3
    line = {"workload": "sparse product, 100k x 100k rows, UInt64 hashed keys, j Zipf(1.1) over 1M ids, i and k over "
                        "4096 ids, Float32 factors (sorting path, Arrow to Arrow)",
            "algorithm": out.algorithm, "joined_pairs": pairs, "hottest_j_rows": [int(ca.max()), int(cb.max())],
            "groups": groups, "arrow_to_arrow_ms": round(ms, 2),
            "gflop_per_s": round(2 * pairs / (ms * 1e-3) / 1e9, 2)}
    if with_cpu:
        from oracle import reference as R
        wa = pa.table([pc.cast(c, pa.float64()) if c.type == pa.float32() else c for c in a.columns],
                      names=a.column_names)
        wb = pa.table([pc.cast(c, pa.float64()) if c.type == pa.float32() else c for c in b.columns],
                      names=b.column_names)
        t0 = time.perf_counter()
        _, want = R.fold(f, wa.to_batches(), wb.to_batches(), None, R.schema(f, [wa.schema, wb.schema]))
        line["cpu_port_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
        line["cpu_port_cores"] = 1
        tg, tw = pa.Table.from_batches(out.batches), pa.Table.from_batches(want)
        same = tg.num_rows == tw.num_rows and all(
            np.array_equal(tg.column(c).to_numpy().view(np.uint64), tw.column(c).to_numpy().view(np.uint64))
            for c in range(3))
        if not same:
            raise SystemExit("sparse product: device groups or sums differ from the oracle's")
        line["bit_exact_vs_oracle"] = True
        # The host plan this path replaces (a hash join on j, then a hash
        # aggregate by (i, k)), stood in for by pyarrow's Acero engine on this
        # host's cores -- not DataFusion -- at 20k rows per side (its join
        # materializes every joined pair), beside the device at that size.
        small = 20_000
        sa = pa.table({"i": a.column(0).slice(0, small), "j": a.column(1).slice(0, small),
                       "va": pc.cast(a.column(2).slice(0, small), pa.float64())})
        sb = pa.table({"j": b.column(0).slice(0, small), "k": b.column(1).slice(0, small),
                       "vb": pc.cast(b.column(2).slice(0, small), pa.float64())})
        plans = []
        for _ in range(3):
            t0 = time.perf_counter()
            jn = sa.join(sb, keys="j", join_type="inner")
            jn = jn.append_column("p", pc.multiply(jn.column("va"), jn.column("vb")))
            g = jn.group_by(["i", "k"]).aggregate([("p", "sum")])
            plans.append(time.perf_counter() - t0)
            if plans[-1] > 30.0:   # time cap: one run is enough
                break
        dev_small = []
        sa32, sb32 = a.slice(0, small), b.slice(0, small)
        kernels.fold(f, sa32, sb32, options=opts)
        for _ in range(3):
            t0 = time.perf_counter()
            got = kernels.fold(f, sa32, sb32, options=opts)
            dev_small.append(time.perf_counter() - t0)
        assert sum(x.num_rows for x in got.batches) == g.num_rows
        line["hash_plan_20k"] = {"joined_rows": jn.num_rows, "groups": g.num_rows,
                                 "hash_plan_ms": round(float(np.median(plans)) * 1e3, 1),
                                 "device_arrow_to_arrow_ms": round(float(np.median(dev_small)) * 1e3, 2),
                                 "hash_plan": f"pyarrow Acero hash join + hash aggregate, {pa.cpu_count()} threads "
                                              "(30 s cap per run): a stand-in for DataFusion's plan, not DataFusion"}
    return line


def attention_pipeline_ir():
    """Causal attention as the reference's fold pipeline over q(h, t, d, v),
    k(h, t, d, v), v(h, t, d, v) (pipeline_vs_sql.rs:622-690): producer
    dimensions 0 = tq, 1 = d, 2 = tk, 3 = h, the pipeline's own 4 = the
    value's d; SUM(q.v * k.v) grouped by (h, tq, tk) where tq >= tk,
    weights exp(s - MAX(s) OVER (h, tq)) over their sum, applied to v."""
    from einfold_amd import ir
    EQ = ir.KeyEquality.EQUAL
    q = ir.Operand((ir.Key(3, 0), ir.Key(0, 1), ir.Key(1, 2)), 3)
    k = ir.Operand((ir.Key(3, 0), ir.Key(2, 1), ir.Key(1, 2)), 3)
    producer = ir.Fold([EQ] * 4, [q, k], [ir.Output(3), ir.Output(0), ir.Output(2)],
                       conditions=[ir.Condition(0, ir.Comparison.GREATER_OR_EQUAL, 2, 0)])
    values = ir.Operand((ir.Key(3, 0), ir.Key(2, 1), ir.Key(4, 2)), 3)
    return ir.Pipeline(producer, ir.Exp(), [EQ], values, [3, 0, 4], ir.Normalizer.KEPT_GROUPS)


def attention_arrow_line():
    """Config 3 through the drop-in boundary: ef_pipeline on Arrow tables in
    host memory (q, k and v 2M rows each: h, t, d INT32 and v FLOAT32, 8
    heads x T = 4096 x d = 64), causal, the SQL's exp(s - max) weights,
    Arrow out (2M rows); host-to-device and device-to-host transfers and
    the Arrow import and export included.  BF16x3 and FP32, each against
    the same call on the values widened to Float64 (the FP64 path)."""
    import pyarrow as pa
    from einfold_amd import data, kernels
    q, k, v = data.attention_tables(H, T, D, seed=3)
    q, k, v = (x.astype(np.float32) for x in (q, k, v))
    t32 = data.attention_arrow(q, k, v, pa.float32())
    t64 = data.attention_arrow(q.astype(np.float64), k.astype(np.float64), v.astype(np.float64), pa.float64())
    p = attention_pipeline_ir()

    def values(piped):
        return np.concatenate([b.column(b.num_columns - 1).to_numpy() for b in piped.batches])

    want = values(kernels.pipeline(p, *t64))
    line = {"workload": "causal attention pipeline over Arrow, 8 heads x T=4096 x d=64 (2M rows per operand)",
            "rows_out": int(want.size)}
    for prec in ("bf16x3", "fp32"):
        opts = kernels.Options(accept_float32=True, precision=prec)
        got = kernels.pipeline(p, *t32, opts)
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            got = kernels.pipeline(p, *t32, opts)
            times.append((time.perf_counter() - t0) * 1e3)
        g = values(got)
        line[f"ms_arrow_to_arrow_{prec}"] = round(sorted(times)[1], 2)
        line[f"rel_error_vs_fp64_{prec}"] = float(np.linalg.norm(g - want) / np.linalg.norm(want))
        line["tiles_visited"], line["tiles"] = got.tiles_visited, got.tiles
    return line


def gradient_memory_line(dev):
    """Config 4's memory report: device memory of the gradient step (forward,
    then dq, dk, dv of SUM(out * g)) at 4 heads x T = 2048 x d = 64, causal,
    beyond the inputs: the recomputing backward (saved row max and row sum,
    its scratch included) against a stored-scores variant (torch's FP32
    autograd of the same causal attention, which keeps the T x T weights of
    every head).  Peaks from torch's allocator, which also serves the
    library's scratch."""
    import torch
    from einfold_amd import data, device
    hh, tt, dd = 4, 2048, 64
    q64, k64, v64 = data.attention_tables(hh, tt, dd, seed=4)
    g64 = data.normal(4, 4, hh * tt * dd).reshape(hh, tt, dd)
    q, k, v, g = (torch.from_numpy(np.ascontiguousarray(x)).float().to(dev) for x in (q64, k64, v64, g64))
    torch.cuda.synchronize()

    def peak(fn):
        torch.cuda.synchronize()
        torch.cuda.reset_peak_memory_stats(dev)
        base = torch.cuda.memory_allocated(dev)
        keep = fn()
        torch.cuda.synchronize()
        used = torch.cuda.max_memory_allocated(dev) - base
        del keep
        torch.cuda.empty_cache()
        return used

    def recompute():
        out, m, l = device.attention_fwd(q, k, v, SCALE, causal=True, precision="bf16x3")
        return out, device.attention_bwd(q, k, v, out, g, m, l, SCALE, causal=True, precision="bf16x3")

    def stored():
        qq, kk, vv = (x.clone().requires_grad_() for x in (q, k, v))
        s = torch.matmul(qq, kk.transpose(1, 2).contiguous()) * SCALE
        mask = torch.ones(tt, tt, dtype=torch.bool, device=dev).tril()
        p = torch.softmax(s.masked_fill(~mask, float("-inf")), dim=-1)
        out = torch.matmul(p, vv)
        (out * g).sum().backward()
        return out.detach(), qq.grad, kk.grad, vv.grad

    mib = lambda b: round(b / 2 ** 20, 2)
    r, s = peak(recompute), peak(stored)
    return {"workload": "attention gradient step, 4 heads x T=2048 x d=64, causal (FP32 inputs)",
            "recompute_peak_mib": mib(r), "stored_scores_peak_mib": mib(s),
            "saved_state_mib": {"recompute_row_stats": mib(2 * 4 * hh * tt),
                                "stored_scores": mib(4 * hh * tt * tt)},
            "ratio": round(s / max(r, 1), 1)}


def ranking_line(dev, with_cpu: bool = False):
    """Ranking of the sorting path's join key (hash dedup, then the distinct
    words ordered by bucket): j over both operands,
    8M UInt64 hashes of Zipf(1.1) ids over 1M ids.  Algorithmic traffic 12 B
    per key (read the 8-byte word, write the 4-byte rank).  With the CPU
    leg, the oracle's C restatement of discriminate.rs's ranking on one
    core over the same keys, whose ranks and first rows must equal the
    device's bit for bit."""
    import ctypes
    import torch
    from einfold_amd import data
    from einfold_amd._lib import lib, check
    rows = 4 << 20
    a_j = data.hash64(data.zipf_ids(5, 1, rows, 1 << 20, 1.1), 0x6A)
    b_j = data.hash64(data.zipf_ids(5, 2, rows, 1 << 20, 1.1), 0x6A)
    words = torch.from_numpy(np.concatenate([a_j, b_j]).view(np.int64)).to(dev)
    n = words.numel()
    rank = torch.empty(n, dtype=torch.int32, device=dev)
    first = torch.empty(n, dtype=torch.int32, device=dev)
    scratch = torch.empty(lib().ef_rank_scratch_bytes(n), dtype=torch.uint8, device=dev)
    k = ctypes.c_size_t(0)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    P = lambda t: ctypes.c_void_p(t.data_ptr())
    run = lambda: check(lib().ef_rank_u64(P(words), n, P(rank), P(first), ctypes.byref(k), P(scratch), stream))
    ms = _events(run, 10)
    gbs = 12.0 * n / (ms * 1e-3) / 1e9
    line = {"workload": "ranking of 8M UInt64 join keys (Zipf 1.1 over 1M ids): hash dedup, distinct words ordered by bucket", "keys": n,
            "distinct": int(k.value), "ms": round(ms, 4), "gb_per_s": round(gbs, 1),
            "frac_of_hbm_peak": round(gbs / 8000.0, 4)}
    # Config 2's group keys: i (and k alike), 4M UInt64 hashes of uniform
    # ids over 64k.
    gi = torch.from_numpy(data.hash64(data.zipf_ids(5, 3, rows, 1 << 16, 0.0), 0x11).view(np.int64)).to(dev)
    ki = ctypes.c_size_t(0)
    rank_i = torch.empty(rows, dtype=torch.int32, device=dev)
    first_i = torch.empty(rows, dtype=torch.int32, device=dev)
    run_i = lambda: check(lib().ef_rank_u64(P(gi), rows, P(rank_i), P(first_i), ctypes.byref(ki), P(scratch), stream))
    ms_i = _events(run_i, 10)
    gbs_i = 12.0 * rows / (ms_i * 1e-3) / 1e9
    line["group_keys"] = {"workload": "ranking of 4M UInt64 group keys (uniform over 64k ids), config 2's i and k",
                          "keys": rows, "distinct": int(ki.value), "ms": round(ms_i, 4), "gb_per_s": round(gbs_i, 1),
                          "frac_of_hbm_peak": round(gbs_i / 8000.0, 4)}
    if with_cpu:
        import oracle
        host = np.concatenate([a_j, b_j])
        t0 = time.perf_counter()
        want_rank, want_first = oracle.ranks_of_words(host)
        line["cpu_port_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
        line["cpu_port_cores"] = 1
        torch.cuda.synchronize()
        same = (np.array_equal(rank.cpu().numpy().view(np.uint32), want_rank) and
                np.array_equal(first[:k.value].cpu().numpy().view(np.uint32), want_first))
        if not same:
            raise SystemExit("ranking: device ranks differ from the oracle's")
        line["bit_exact_vs_oracle"] = True
    return line


def _cpu_info():
    """The host's CPU model, its core count, and the threads this process
    may use (its affinity, capped by OMP_NUM_THREADS where set)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0))
    cap = os.environ.get("OMP_NUM_THREADS")
    if cap and cap.isdigit() and int(cap) > 0:
        usable = min(usable, int(cap))
    return model, os.cpu_count(), usable


def cpu_baseline():
    """einfold's fused CPU kernel for this workload's forward pass (the
    reference has no backward), restated for speed as the reference runs
    it: X tiles of 64 queries as tasks over threads (fused.rs:412-428), each
    64 x 64 score block and weighted sum a register-blocked FP64 product
    (gemm.rs:25-46), libm exp; built -O3 with AVX2 + FMA where the host has
    them (oracle/ef_cpu_baseline.c).  All 8 heads at T=4096, d=64, causal,
    on every thread this process may use; a single-thread figure over one
    head beside it.  Checked against numpy float64 on the first head."""
    import oracle
    from einfold_amd import data
    q, k, v = data.attention_tables(H, T, D, seed=3)
    lo = np.zeros(T, np.uint64)
    hi = np.arange(1, T + 1, dtype=np.uint64)
    model, ncpu, threads = _cpu_info()
    _, isa = oracle.cpu_lib()
    kts = [np.ascontiguousarray(k[h].T) for h in range(H)]
    oracle.cpu_attention_fwd(q[0, :256], kts[0], v[0], SCALE, lo[:256], hi[:256], threads)   # warm
    t0 = time.perf_counter()
    outs = [oracle.cpu_attention_fwd(q[h], kts[h], v[h], SCALE, lo, hi, threads) for h in range(H)]
    dt = time.perf_counter() - t0
    t1 = time.perf_counter()
    oracle.cpu_attention_fwd(q[0], kts[0], v[0], SCALE, lo, hi, 1)
    dt1 = time.perf_counter() - t1
    flop_head = 4 * D * pairs_per_head(T)
    s = (q[0] @ k[0].T) * SCALE
    s = np.where(np.tril(np.ones((T, T), bool)), s, -np.inf)
    w = np.exp(s - s.max(axis=1, keepdims=True))
    want = (w @ v[0]) / w.sum(axis=1, keepdims=True)
    err = float(np.max(np.abs(outs[0] - want)) / np.max(np.abs(v[0])))
    if err > 1e-9:
        raise SystemExit(f"cpu baseline: differs from numpy float64 ({err})")
    return {"value": round(H * flop_head / dt / 1e9, 2), "unit": "GFLOP/s", "cores": threads, "kind": "port",
            "seconds": round(dt, 3), "single_thread_gflops": round(flop_head / dt1 / 1e9, 2),
            "cpu_model": model, "nproc": ncpu, "isa": isa, "max_err_vs_numpy": err,
            "sample": "the full forward pass of the workload (8 heads x T=4096 x d=64, causal, 17.2 GFLOP), "
                      "FP64, einfold's fused tile loop restated in C (X tiles over threads, register-blocked "
                      "tile products, libm exp); the single-thread figure is one head on one thread"}


if __name__ == "__main__":
    main()
