"""Many uneven dense products in one persistent, graph-replayed launch
(the batched dense folds with Float32 factors): accuracy against float64
products per precision, bits independent of the tile shape the autotuner
picks and of replays, and edge shapes."""
import numpy as np
import pytest
import torch

from einfold_amd import data, device

pytestmark = pytest.mark.gpu

TOL = {"fp32": 2e-6, "bf16x3": 2e-5, "bf16": 2e-2}


def _problems(shapes, dev, seed=0):
    g = torch.Generator().manual_seed(seed)
    out = []
    for m, n, k in shapes:
        a = torch.randn(m, k, generator=g).to(dev)
        b = torch.randn(k, n, generator=g).to(dev)
        c = torch.full((m, n), float("nan"), device=dev)
        out.append((a, b, c))
    return out


def _check(problems, tol):
    for a, b, c in problems:
        want = a.double() @ b.double()
        err = float(torch.linalg.vector_norm(c.double() - want) / torch.linalg.vector_norm(want))
        assert err < tol, (tuple(a.shape), tuple(b.shape), err)


@pytest.mark.parametrize("precision", ["fp32", "bf16x3", "bf16"])
def test_uneven_products(cuda, precision):
    shapes = data.log_uniform_shapes(24, 32, 512, seed=3)
    problems = _problems(shapes, cuda)
    plan = device.GroupedPlan(problems, precision)
    plan.run()
    torch.cuda.synchronize()
    _check(problems, TOL[precision])
    plan.close()


def test_bits_independent_of_tile_and_replay(cuda):
    shapes = [(100, 200, 300), (33, 17, 1000), (256, 256, 64), (1, 5, 7), (130, 129, 31)]
    p64 = _problems(shapes, cuda, seed=1)
    p128 = [(a, b, torch.empty_like(c)) for a, b, c in p64]
    p256 = [(a, b, torch.empty_like(c)) for a, b, c in p64]
    plan64 = device.GroupedPlan(p64, "bf16x3", tile=64)
    plan128 = device.GroupedPlan(p128, "bf16x3", tile=128)
    plan256 = device.GroupedPlan(p256, "bf16x3", tile=256)
    assert (plan64.tile, plan128.tile, plan256.tile) == (64, 128, 256)
    plan64.run()
    plan128.run()
    plan256.run()
    torch.cuda.synchronize()
    first = [c.clone() for _, _, c in p64]
    for (_, _, c64), (_, _, c128), (_, _, c256) in zip(p64, p128, p256):
        assert torch.equal(c64, c128)
        assert torch.equal(c64, c256)
    for _ in range(3):
        plan64.run()
    torch.cuda.synchronize()
    for c0, (_, _, c) in zip(first, p64):
        assert torch.equal(c0, c)
    _check(p64, TOL["bf16x3"])


@pytest.mark.parametrize("precision", ["bf16", "bf16x3", "fp32"])
def test_large_tiles(cuda, precision):
    """The 256-row tiles (256 x 256 for BF16, 256 x 128 with 32-wide k
    steps for BF16x3 and FP32) on shapes with partial tiles at every
    edge."""
    shapes = [(300, 520, 200), (257, 255, 65), (512, 512, 128), (1, 300, 33), (600, 1, 64)]
    problems = _problems(shapes, cuda, seed=5)
    plan = device.GroupedPlan(problems, precision, tile=256)
    assert plan.tile == 256
    plan.run()
    torch.cuda.synchronize()
    _check(problems, TOL[precision])
    plan.close()


def test_autotuner_picks_a_tile(cuda):
    problems = _problems(data.log_uniform_shapes(8, 64, 256, seed=4), cuda)
    plan = device.GroupedPlan(problems, "fp32")
    assert plan.tile in (64, 128, 256)
    plan.run()
    torch.cuda.synchronize()
    _check(problems, TOL["fp32"])


def test_empty_dimensions(cuda):
    """k = 0 gives a zero product (nothing is read); m = 0 or n = 0 writes
    nothing; the other products in the same launch are unaffected."""
    shapes = [(5, 7, 0), (0, 3, 4), (3, 0, 4), (64, 64, 64), (70, 9, 0)]
    problems = _problems(shapes, cuda, seed=2)
    plan = device.GroupedPlan(problems, "bf16x3")
    plan.run()
    torch.cuda.synchronize()
    for (m, n, k), (a, b, c) in zip(shapes, problems):
        if k == 0:
            assert torch.equal(c, torch.zeros_like(c))
        elif m and n:
            _check([(a, b, c)], TOL["bf16x3"])
    plan.close()
