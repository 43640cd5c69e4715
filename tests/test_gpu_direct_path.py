"""computeControl's direct path against the graph-replayed path.

Without use_graph the built-in controllers launch a call directly: the iteration arguments are kernel arguments, the
rollout reads x0 from mapped host memory, the last combine workgroup smooths and constrains the new mean (the host
functions' arithmetic) and leaves it on the device as the next call's nominal, uploaded again only when the host
changes its sequence. With use_graph the call is the captured copy + iteration graph and the host post-processes.
Both must give the same bits: the control sequence, the raw mean, the weight statistics and the state trajectory
over calls with slides, an explicit set of the sequence, several iterations per call, colored noise, and the
device-resident loop and sampled-trajectory export that follow a direct call.
"""
import ctypes as ct

import numpy as np
import pytest

from gpu_util import banner
from mppi_generic_vendor_amd import workloads
from mppi_generic_vendor_amd.controller import VanillaMPPIController

pytestmark = pytest.mark.gpu


def _problems():
    m = workloads.gaussian_blob_map(256, 256, blobs=16, seed=7)
    yield workloads.cartpole(num_rollouts=512, num_timesteps=50)
    yield workloads.double_integrator_map(num_rollouts=1024, num_timesteps=64, colored=True, cost_map=m)
    yield workloads.bicycle_nn(num_rollouts=256, num_timesteps=40, residual=False, cost_map=m)


def _device_mean(ctrl):
    from mppi_generic_vendor_amd import hip as hipmod

    hipmod.synchronize()
    _, _, mean_d = ctrl.device_buffers()
    m = np.empty((ctrl.T, ctrl.C), np.float32)
    hipmod.check(hipmod._load().hipMemcpy(m.ctypes.data, ct.c_void_p(mean_d), m.nbytes, hipmod.D2H), "hipMemcpy")
    return m


@pytest.mark.parametrize("num_iters", [1, 2])
def test_direct_calls_equal_graph_calls(hip, engine, num_iters):
    for prob in _problems():
        banner(f"direct vs graph: {prob.name} K={prob.num_rollouts} T={prob.num_timesteps} iters={num_iters}")
        a = VanillaMPPIController(prob, num_iters=num_iters, use_graph=False)
        b = VanillaMPPIController(prob, num_iters=num_iters, use_graph=True)
        rng = np.random.default_rng(5)
        x = prob.x0.copy()
        for call in range(6):
            if call == 2:
                a.slideControlSequence(1)
                b.slideControlSequence(1)
            if call == 4:
                u = (0.1 * rng.standard_normal((prob.num_timesteps, prob.C))).astype(np.float32)
                a.updateImportanceSampler(u)
                b.updateImportanceSampler(u)
            a.computeControl(x, 1)
            b.computeControl(x, 1)
            np.testing.assert_array_equal(a.getRawMean(), b.getRawMean())
            np.testing.assert_array_equal(a.getControlSeq(), b.getControlSeq())
            np.testing.assert_array_equal(a.getTargetStateSeq(), b.getTargetStateSeq())
            sa, sb = a.getWeightStats(), b.getWeightStats()
            assert (sa.baseline, sa.argmin, sa.normalizer) == (sb.baseline, sb.argmin, sb.normalizer)
            x = (x + 0.01 * (call + 1)).astype(np.float32)
        # the device loop starts from the last call's x0 and mean on both paths
        a.stepDevice(3)
        b.stepDevice(3)
        np.testing.assert_array_equal(_device_mean(a), _device_mean(b))
        a.close()
        b.close()


def test_direct_path_sampled_trajectories(hip, engine):
    """The sampled-trajectory export after direct calls (nominal kept on the device, x0 read from host memory) equals
    the export after graph calls."""
    from mppi_generic_vendor_amd import capi

    prob = workloads.double_integrator_map(num_rollouts=512, num_timesteps=48, colored=False,
                                           cost_map=workloads.gaussian_blob_map(256, 256, blobs=16, seed=7))
    banner("direct vs graph: sampled-trajectory export")
    T, C, O, n = prob.num_timesteps, prob.C, prob.dyn.output_dim, 1 + 3 + 2
    outs = []
    for graph in (False, True):
        c = VanillaMPPIController(prob, num_iters=1, use_graph=graph)
        capi.call("mppi_controller_set_sampled_trajectories", c._h, 3, 2)
        got = []
        for _ in range(3):
            c.computeControl(prob.x0, 1)
            cnt = ct.c_int32()
            arrs = [np.zeros(n, np.int32), np.zeros((n, T, C), np.float32), np.zeros((n, T, O), np.float32),
                    np.zeros((n, T), np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32)]
            capi.call("mppi_controller_get_sampled_trajectories", c._h, ct.byref(cnt), *[a.ctypes.data for a in arrs])
            assert cnt.value == n
            got.append(arrs)
        outs.append(got)
        c.close()
    for ca, cb in zip(outs[0], outs[1]):
        for a, b in zip(ca, cb):
            np.testing.assert_array_equal(a, b)
