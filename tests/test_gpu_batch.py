"""Batched instances (lgh_prior_create_mat_batch, lgh_glr_compute_batch) on
an MI355X: every instance of a batch equals the single call bit for bit
(the workers only change which stream runs it: the same kernels in the
same order, and every factorization's result depends only on its input),
instances that share a prior or a B are handled in order by one worker,
and a failing instance is reported by index with every output released."""
import numpy as np
import pytest

import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev(lgh):
    lgh.init()
    return lgh


def problems():
    out = []
    for k, side in enumerate((16, 20, 24, 28, 22, 18)):
        Bs, Zs, M = W.psf2d_problem(side)
        out.append((Bs, Zs, M, 12 + 4 * k))
    # a block with a whole panel and a narrow last one (ell = 74: 64 + 10),
    # which a batch worker solves on its one stream
    Bs, Zs, M = W.psf2d_problem(30)
    out.append((Bs, Zs, M, 74))
    return out


@pytest.mark.parametrize("mixed", [True, False])
def test_batch_equals_single_calls(dev, mixed):
    probs = problems()
    x = dev.xopts(mixed=mixed)
    priors = dev.Prior.batch([(Z, M) for _, Z, M, _ in probs], workers=4,
                             blocked_mode=3)
    single_priors = [dev.Prior.from_mat(Z, M, blocked_mode=3) for _, Z, M, _ in probs]
    for a, b in zip(priors, single_priors):
        assert a.info() == b.info()
    items = [(B, pr, dev.glr_opts(ell=ell, seed=0))
             for (B, _, _, ell), pr in zip(probs, priors)]
    gs = dev.GLR.batch(items, x=x, workers=4)
    for (B, _, _, ell), pr, g in zip(probs, single_priors, gs):
        g1 = dev.GLR(B, pr, dev.glr_opts(ell=ell, seed=0), x=x)
        assert g.kept == g1.kept
        assert np.array_equal(g.lam, g1.lam)
        assert np.array_equal(g.U, g1.U)
        rhs = np.random.default_rng(ell).standard_normal(B.shape[0])
        assert np.array_equal(g.solve(1e-2, rhs), g1.solve(1e-2, rhs))


def test_batch_shared_objects_and_failure(dev):
    probs = problems()[:3]
    B0, Z0, M0, _ = probs[0]
    shared = dev.Prior.from_mat(Z0, M0, blocked_mode=3)
    Bm0 = dev.Mat.aij(B0)
    # three instances on one prior and one B, plus two others
    items = [(Bm0, shared, dev.glr_opts(ell=10, seed=0)),
             (Bm0, shared, dev.glr_opts(ell=14, seed=1)),
             (Bm0, shared, dev.glr_opts(ell=18, seed=2))]
    for B, Z, M, ell in probs[1:]:
        items.append((B, dev.Prior.from_mat(Z, M, blocked_mode=3),
                      dev.glr_opts(ell=ell, seed=0)))
    x = dev.xopts(mixed=True)
    gs = dev.GLR.batch(items, x=x, workers=8)
    for (B, pr, o), g in zip(items, gs):
        g1 = dev.GLR(B, pr, o, x=x)
        assert np.array_equal(g.lam, g1.lam)
    # instance 3 asks for ell > n: the batch fails and names it
    bad = list(items)
    B3, pr3, _ = bad[3]
    bad[3] = (B3, pr3, dev.glr_opts(ell=10 ** 6, seed=0))
    with pytest.raises(dev.LGHError, match="instance 3"):
        dev.GLR.batch(bad, x=x, workers=4)
