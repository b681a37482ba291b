"""cf_driven_force_noalloc (no device allocation; scratch in a file-scope
device array) against cf_driven_force, through the C ABI: the same bits for
a device array and for the resident iterate, on two calls in a row whose
second has the smaller grid, and the size limit refused."""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not gpu_available():
        pytest.fail("no GPU visible: the gpu tests must run on an MI355X")


def _case(grid, seed):
    import torch
    from channel_flow_amd.config import ChannelConfig, FlowParams, poiseuille_state
    from channel_flow_amd.solver import Solver
    from oracle import oracle as O
    cfg = ChannelConfig(*grid, gamma_metric=1.5)
    Qn = poiseuille_state(cfg, FlowParams(), seed=seed, stream=0)
    Qo = poiseuille_state(cfg, FlowParams(), seed=seed, stream=1)
    s = Solver(cfg)
    met = s.Jacobian_all(1.5)
    s.load_state(Qn, Qn)
    U2 = np.ascontiguousarray(O.state_from_Q(cfg, Qo, met["J"])[1])
    d = torch.from_numpy(U2).to("cuda:0")
    torch.cuda.synchronize()
    return s, met, U2, d


@pytest.mark.parametrize("order", [((24, 12, 16), (9, 10, 13)), ((40, 20, 24), (16, 12, 16))])
def test_gpu_driven_force_noalloc_matches(order):
    from channel_flow_amd.solver import Solver  # noqa: F401
    print(f"noalloc: grids {order}", flush=True)
    for n, grid in enumerate(order):
        s, met, U2, d = _case(grid, 30 + n)
        try:
            # a device array in the reference layout
            want = s.Driven_force(U2_=U2, J=met["J"])
            got = s.Driven_force_noalloc(U2_dev=d.data_ptr())
            assert got == want, (grid, got, want)
            # the resident iterate
            want = s.Driven_force()
            got = s.Driven_force_noalloc()
            assert got == want, (grid, got, want)
            # again, after the other path has used the scratch
            assert s.Driven_force_noalloc(U2_dev=d.data_ptr()) == s.Driven_force(U2_=U2, J=met["J"])
        finally:
            s.close()


def test_gpu_driven_force_noalloc_refuses_large_grid():
    from channel_flow_amd.config import ChannelConfig, FlowParams, poiseuille_state
    from channel_flow_amd.solver import CfError, Solver
    cfg = ChannelConfig(520, 8, 8, gamma_metric=1.5)
    with Solver(cfg) as s:
        s.Jacobian_all(1.5, want_arrays=False)
        Qn = poiseuille_state(cfg, FlowParams(), seed=3, stream=0)
        s.load_state(Qn, Qn)
        with pytest.raises(CfError, match="CF_ERR_ARG"):
            s.Driven_force_noalloc()
        s.Driven_force()  # the allocating entry point has no such limit
