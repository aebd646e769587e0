"""fpc_stage_diagnostics_noalloc, the allocation-free form of the diagnostics stage
(its block partial sums in the library's static device array instead of an
allocated buffer), against fpc_stage_diagnostics on the same fields: the same
six sums bit for bit, on two calls in a row, the second on a smaller channel,
and equal to the diagnostics the simulation computed itself after its step."""
import ctypes

import numpy as np
import pytest

import fpc_amd as F
from helpers import CFG1, TINY, bits_equal

pytestmark = pytest.mark.gpu


def _diag_both(kw, seed):
    sim = F.Sim(F.default_params(**dict(kw, init_profile=1, noise_amp=1e-3, seed=seed)))
    sim.step(2)
    L = F.lib()
    F.check(L.fpc_DeviceSynchronize(), "sync")
    ops = sim.flow_ops()
    ptr = {f: ctypes.c_void_p(sim.device_ptr(f)) for f in ("u", "v", "w", "p", "qout")}
    n_p = F.Setup(sim.params).count("Np")
    ramp = sim.scalar("ramp")
    import torch
    outs = []
    for fn in (L.fpc_stage_diagnostics, L.fpc_stage_diagnostics_noalloc):
        out = torch.full((8,), -1.0, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        F.check(fn(ctypes.byref(ops), ptr["u"], ptr["v"], ptr["w"], ptr["p"], ptr["qout"], ramp, n_p,
                   ctypes.c_void_p(out.data_ptr())), fn.__name__)
        F.check(L.fpc_DeviceSynchronize(), "sync")
        outs.append(out.cpu().numpy()[:6])
    own = np.array([sim.scalar(k) for k in ("div_l2", "umax", "vmax", "wmax", "fx", "fy")])
    return outs[0], outs[1], own


def test_noalloc_diagnostics_equal_the_allocating_stage(torch_cuda):
    L = F.lib()
    L.fpc_ClearError(-1)  # error bits are sticky per thread (as HYPRE's): drop those of earlier calls
    F.check(L.fpc_Initialize(), "Initialize")
    big = _diag_both(CFG1, seed=11)
    small = _diag_both(TINY, seed=12)  # second call: a smaller channel, the array keeps the first one's sums
    for ref, noalloc, own in (big, small):
        assert bits_equal(noalloc, ref), (noalloc, ref)
        assert bits_equal(ref, own), (ref, own)
        assert np.all(np.isfinite(ref)) and ref[1] > 0
    assert not np.array_equal(big[0], small[0])


def test_noalloc_diagnostics_rejects_bad_arguments(torch_cuda):
    L = F.lib()
    ops = F.FlowOps()
    z = ctypes.c_void_p(0)
    assert L.fpc_stage_diagnostics_noalloc(None, z, z, z, z, z, 0.0, 1, z) == F.ERROR_ARG
    L.fpc_ClearError(F.ERROR_ARG)
    assert L.fpc_stage_diagnostics_noalloc(ctypes.byref(ops), z, z, z, z, z, 0.0, 1, z) == F.ERROR_ARG
    L.fpc_ClearError(F.ERROR_ARG)
