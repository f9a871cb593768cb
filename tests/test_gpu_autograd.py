"""The backward kernels through torch.autograd, checked three ways:

* FP64 central differences (eps = 1e-3) of SUM(out * g) on 256 random
  entries of each of q, k and v, at 4 heads x T = 2048 x d = 64 (the
  differences' own O(eps^2) truncation sets the 1e-5 bound);
* torch float64 autograd of the same attention on the CPU;
* the left-joined gradient folds that ddx writes for dQ, dK and dV
  (docs/spikes/s31-ddx-gradient-steps), each left-joined from the
  differentiated table with 0.0 for positions nothing reaches, run by the
  dense fold operator (ef_fold_dense).
"""
import numpy as np
import pyarrow as pa
import pytest
import torch

from einfold_amd import autograd, data, device, ir, kernels

pytestmark = pytest.mark.gpu


def _tables(h, t, d, seed):
    q, k, v = data.attention_tables(h, t, d, seed=seed)
    g = data.normal(seed, 4, h * t * d).reshape(h, t, d)
    return q, k, v, g


def _loss(q, k, v, g, scale, precision):
    return (autograd.fused_attention(q, k, v, scale, True, precision) * g).sum()


def test_central_differences_fp64(cuda):
    """4 heads, T = 2048, d = 64 (the gradient workload): the FP64 backward
    against (L(x + eps) - L(x - eps)) / 2 eps at 256 entries per tensor."""
    h, t, d, eps, scale = 4, 2048, 64, 1e-3, 0.125
    q, k, v, g = (torch.from_numpy(np.ascontiguousarray(x)).to(cuda) for x in _tables(h, t, d, 5))
    qg, kg, vg = (x.clone().requires_grad_(True) for x in (q, k, v))
    _loss(qg, kg, vg, g, scale, "fp64").backward()
    rng = np.random.default_rng(0)
    for name, base, grad in (("q", q, qg.grad), ("k", k, kg.grad), ("v", v, vg.grad)):
        idx = rng.choice(base.numel(), 256, replace=False)
        fd = np.empty(256)
        with torch.no_grad():
            for n, i in enumerate(idx):
                args = {"q": q, "k": k, "v": v}
                plus, minus = base.clone(), base.clone()
                plus.view(-1)[i] += eps
                minus.view(-1)[i] -= eps
                lp = _loss(*(plus if nm == name else args[nm] for nm in ("q", "k", "v")), g, scale, "fp64")
                lm = _loss(*(minus if nm == name else args[nm] for nm in ("q", "k", "v")), g, scale, "fp64")
                fd[n] = (lp - lm).item() / (2 * eps)
        got = grad.view(-1)[torch.from_numpy(idx).to(cuda)].cpu().numpy()
        err = np.linalg.norm(got - fd) / np.linalg.norm(fd)
        assert err < 1e-5, (name, err)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-5), ("bf16x3", 3e-5)])
def test_autograd_against_torch_float64(cuda, precision, tol):
    h, t, d, scale = 2, 384, 64, 0.125
    q, k, v, g = _tables(h, t, d, 6)
    dt = torch.float64 if precision == "fp64" else torch.float32
    xs = [torch.from_numpy(np.ascontiguousarray(x)).to(dt).to(cuda).requires_grad_(True) for x in (q, k, v)]
    gt = torch.from_numpy(g).to(dt).to(cuda)
    _loss(*xs, gt, scale, precision).backward()
    r = [torch.from_numpy(np.ascontiguousarray(x)).double().requires_grad_(True) for x in (q, k, v)]
    s = scale * r[0] @ r[1].transpose(1, 2)
    mask = torch.ones(t, t, dtype=torch.bool).tril()
    p = torch.softmax(s.masked_fill(~mask, float("-inf")), dim=-1)
    ((p @ r[2]) * torch.from_numpy(g)).sum().backward()
    for x, want in zip(xs, r):
        got = x.grad.double().cpu()
        err = float(torch.linalg.vector_norm(got - want.grad) / torch.linalg.vector_norm(want.grad))
        assert err < tol, err


@pytest.mark.parametrize("wrt", ["q", "k", "v"])
def test_gradients_equal_the_left_joined_gradient_folds(cuda, wrt):
    """dQ, dK and dV from the backward kernels (FP64) against ddx's gradient
    SQL (docs/spikes/s31-ddx-gradient-steps/README.md:57-67), each run as a
    dense fold left-joined from the differentiated table with missing 0.0
    (crates/einfold-datafusion/tests/detect_vs_host.rs:1306-1367):

      dQ(h, tq, d) = SUM over tk of dS(h, tq, tk) * scale * K(h, tk, d)
      dK(h, tk, d) = SUM over tq of dS(h, tq, tk) * scale * Q(h, tq, d)
      dV(h, tk, d) = SUM over tq of P(h, tq, tk) * dO(h, tq, d)

    with P the causal softmax and dS = P * (dP - rowsum(dO * O)), dP = dO V^T,
    given as tables of the allowed (tq, tk) pairs only: the masked triangle
    is holes, which the domain's default turns into 0.0."""
    h, t, d, scale = 2, 256, 32, 0.125
    q, k, v, g = _tables(h, t, d, 7)
    xs = [torch.from_numpy(np.ascontiguousarray(x)).to(cuda).requires_grad_(True) for x in (q, k, v)]
    _loss(*xs, torch.from_numpy(g).to(cuda), scale, "fp64").backward()
    grad = xs["qkv".index(wrt)].grad.cpu().numpy()
    s = scale * np.einsum("htd,hsd->hts", q, k)
    allowed = np.tril(np.ones((t, t), bool))
    s = np.where(allowed, s, -np.inf)
    p = np.exp(s - s.max(-1, keepdims=True))
    p /= p.sum(-1, keepdims=True)
    o = np.einsum("hts,hsd->htd", p, v)
    dp = np.einsum("htd,hsd->hts", g, v)
    ds = p * (dp - (g * o).sum(-1, keepdims=True))
    hh, tq, tk = np.nonzero(np.broadcast_to(allowed, (h, t, t)))
    w = (p if wrt == "v" else ds * scale)[hh, tq, tk]
    ptab = pa.table({"h": hh.astype(np.int64), "tq": tq.astype(np.int64), "tk": tk.astype(np.int64), "w": w})
    gh, gt, gd = np.meshgrid(np.arange(h), np.arange(t), np.arange(d), indexing="ij")
    cols = lambda name, x: {"h": gh.reshape(-1), name: gt.reshape(-1), "d": gd.reshape(-1), "v": x.reshape(-1)}
    other = {"q": pa.table(cols("tk", k)), "k": pa.table(cols("tq", q)), "v": pa.table(cols("tq", g))}[wrt]
    dom = pa.table({"h": gh.reshape(-1), "t": gt.reshape(-1), "d": gd.reshape(-1)})
    EQ = ir.KeyEquality.EQUAL
    # dims 0 = h, 1 = tq, 2 = tk, 3 = d
    a_op = ir.Operand((ir.Key(0, 0), ir.Key(1, 1), ir.Key(2, 2)), 3)
    if wrt == "q":   # K's rows are keys: joined on tk, grouped by (h, tq, d)
        b_op = ir.Operand((ir.Key(0, 0), ir.Key(2, 1), ir.Key(3, 2)), 3)
        outs = [0, 1, 3]
    else:            # Q's or dO's rows are queries: joined on tq, grouped by (h, tk, d)
        b_op = ir.Operand((ir.Key(0, 0), ir.Key(1, 1), ir.Key(3, 2)), 3)
        outs = [0, 2, 3]
    fold = ir.Fold([EQ] * 4, [a_op, b_op], [ir.Output(x, ir.DomainKey(i)) for i, x in enumerate(outs)],
                   domain=ir.Domain([0, 1, 2], 0.0))
    out = kernels.fold_dense(fold, ptab, other, dom)
    assert out is not None, "the dense algorithm takes the masked operand under the 0.0 default"
    tab = pa.Table.from_batches(out)
    assert np.array_equal(tab.column(1).to_numpy().reshape(h, t, d)[0, :, 0], np.arange(t))
    got = tab.column(3).to_numpy().reshape(h, t, d)
    err = np.max(np.abs(got - grad)) / np.max(np.abs(grad))
    assert err < 1e-11, (wrt, err)


@pytest.mark.parametrize("precision,tol", [("bf16x3", 1e-5), ("fp32", 2e-6)])
def test_mixed_backward_at_the_gradient_workload_against_fp64(cuda, precision, tol):
    """The gradient workload at full size (4 heads, T = 2048, d = 64,
    causal): out, dq, dk and dv of the mixed (BF16x3) and plain FP32 forms
    against the FP64 form on the same device, each within the 1e-5 relative
    bar (norm of the difference over the norm of the FP64 result)."""
    print(f"full-size backward {precision}", flush=True)
    h, t, d, scale = 4, 2048, 64, 0.125
    q, k, v, g = _tables(h, t, d, 11)
    t64 = [torch.from_numpy(np.ascontiguousarray(x)).to(cuda) for x in (q, k, v, g)]
    t32 = [x.float() for x in t64]
    o64, m64, l64 = device.attention_fwd(*t64[:3], scale, causal=True, precision="fp64")
    g64 = device.attention_bwd(*t64[:3], o64, t64[3], m64, l64, scale, causal=True, precision="fp64")
    o32, m32, l32 = device.attention_fwd(*t32[:3], scale, causal=True, precision=precision)
    g32 = device.attention_bwd(*t32[:3], o32, t32[3], m32, l32, scale, causal=True, precision=precision)
    for name, got, want in zip(("out", "dq", "dk", "dv"), (o32,) + tuple(g32), (o64,) + tuple(g64)):
        err = float(torch.linalg.vector_norm(got.double() - want) / torch.linalg.vector_norm(want))
        assert err < tol, (name, err)
