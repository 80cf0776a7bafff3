"""CPU pins of the FP64 reference used by the adjoint tests (tests/adjoint_ref.py):
its forward substeps against the oracle's f64 mode (same algorithm, with the
fixed-point p2g sums the reference replaces by real sums), and its autograd
gradient against central finite differences with the discrete choices frozen."""
import numpy as np
import torch

from splashtank_amd import capi, scenes
from oracle.oracle import Oracle
from tests.adjoint_ref import RefSim
from tests.helpers import opts_for


def _scene():
    return scenes.dam_break(d=1.0 / 32, iters=20)  # 16^3 cells, 7,904 particles


def test_reference_forward_matches_oracle_f64():
    sc = _scene()
    o = Oracle(opts_for(sc), fp64=True)
    o.upload(sc.pos, sc.vel)
    prm = capi.make_params(**sc.params)
    K = 3
    for _ in range(K):
        o.step(prm)
    ref = RefSim(o.dims, sc.params)
    p, v, _ = ref.rollout(torch.tensor(sc.pos[:, :3], dtype=torch.float64), torch.tensor(sc.vel[:, :3], dtype=torch.float64), K)
    dp = np.abs(p.numpy() - o.dump("pos")[:, :3]).max()
    dv = np.abs(v.numpy() - o.dump("vel")[:, :3]).max()
    vs = np.abs(o.dump("vel")[:, :3]).max()
    # the oracle rounds every p2g term to 1/65536: differences at that level only
    assert dp < 1e-6, dp
    assert dv < 2e-4 * vs, (dv, vs)


def test_reference_gradient_matches_central_differences():
    sc = _scene()
    o = Oracle(opts_for(sc), fp64=True)
    ref = RefSim(o.dims, sc.params)
    K = 3
    rng = np.random.default_rng(0)
    n = sc.n
    A = torch.tensor(rng.standard_normal((n, 3)))
    B = torch.tensor(rng.standard_normal((n, 3)))
    pos0 = torch.tensor(sc.pos[:, :3], dtype=torch.float64, requires_grad=True)
    vel0 = torch.tensor(sc.vel[:, :3] + 0.05 * rng.standard_normal((n, 3)), dtype=torch.float64, requires_grad=True)
    pK, vK, rec = ref.rollout(pos0, vel0, K)
    ((A * pK).sum() + (B * vK).sum()).backward()
    eps = 1e-6
    for _ in range(4):
        dp = torch.tensor(rng.standard_normal((n, 3)))
        dv = torch.tensor(rng.standard_normal((n, 3)))

        def f(s):
            with torch.no_grad():
                a, b, _ = ref.rollout(pos0 + s * eps * dp, vel0 + s * eps * dv, K, frozen=rec)
                return float((A * a).sum() + (B * b).sum())

        fd = (f(1) - f(-1)) / (2 * eps)
        an = float((pos0.grad * dp).sum() + (vel0.grad * dv).sum())
        assert abs(fd - an) <= 1e-7 * abs(an), (fd, an)
