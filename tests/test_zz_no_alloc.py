"""sld_model_eval_noalloc, the model evaluation for callers that cannot
allocate device memory: through the C ABI, into caller device buffers, it
must give bitwise what sld_model_eval gives, on two calls in a row (the
second on a smaller context, so the shared static scratch holds the first
call's values), and refuse a context larger than its fixed scratch."""
import numpy as np
import pytest

from spilady_amd import configs, fit
from spilady_amd.engine import SldError

pytestmark = pytest.mark.gpu


def _confs(n, nu, seed0):
    return [configs.make_bcc_fe(nu, 300.0, seed0 + k, tilt_max_deg=30.0, displace_sigma=0.08) for k in range(n)]


def _both(confs, theta):
    import torch
    m = fit.PotentialModel(confs)
    try:
        m.set_params(theta)
        E, F, H = m.eval()
        dev = torch.device("cuda")
        Ed = torch.full((m.nconf,), np.nan, dtype=torch.float64, device=dev)
        Fd = torch.full((m.ntot, 3), np.nan, dtype=torch.float64, device=dev)
        Hd = torch.full((m.ntot, 3), np.nan, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        print(f"noalloc: {m.nconf} configurations, {m.ntot} atoms", flush=True)
        m.engine._rc(m.L.sld_model_eval_noalloc(m.h, Ed.data_ptr(), Fd.data_ptr(), Hd.data_ptr()))
        torch.cuda.synchronize()
        return (E, F, H), (Ed.cpu().numpy(), Fd.cpu().numpy(), Hd.cpu().numpy())
    finally:
        m.close()


def test_noalloc_eval_equals_eval_twice_then_smaller():
    th = fit.perturbed_params(fit.default_params(), 0.02, 999)
    for confs in (_confs(6, 10, 300), _confs(3, 7, 400)):
        ref, got = _both(confs, th)
        for name, a, b in zip("EFH", ref, got):
            assert np.array_equal(a, b), (name, np.max(np.abs(a - b)))


def test_noalloc_refuses_contexts_beyond_its_scratch():
    m = fit.PotentialModel(_confs(9, 26, 500))   # 9 x 35,152 atoms > 262,144 atom slots
    try:
        m.set_params(fit.default_params())
        with pytest.raises(SldError):
            m.engine._rc(m.L.sld_model_eval_noalloc(m.h, None, None, None))
    finally:
        m.close()
