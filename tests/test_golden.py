"""Golden fixtures of the cube workload: the oracle reproduces them exactly,
and (on a GPU) so does the HIP path."""
import numpy as np
import pytest

from luz_amd import api
from tests.golden.make_golden import HERE, LIGHT_VARIANTS, cube_case, gbuffer_digest


@pytest.mark.parametrize("variant", list(LIGHT_VARIANTS))
def test_oracle_reproduces_golden(variant):
    gold = np.load(HERE / f"cube_{variant}.npz")
    _, g, _, sh, ao, st = cube_case(variant)
    assert gbuffer_digest(g) == str(gold["gbuffer_sha256"])
    np.testing.assert_array_equal(sh, gold["shadow"])
    np.testing.assert_array_equal(ao, gold["ao"])
    assert (st.rays, st.node_visits, st.tri_tests, st.inst_visits) == (
        int(gold["rays"]), int(gold["node_visits"]), int(gold["tri_tests"]), int(gold["inst_visits"]))


@pytest.mark.gpu
@pytest.mark.parametrize("variant", list(LIGHT_VARIANTS))
def test_gpu_reproduces_golden(ctx, variant):
    gold = np.load(HERE / f"cube_{variant}.npz")
    s, g, sb, _, _, _ = cube_case(variant)
    gs = api.GPUScene(ctx, s)
    W, H, cam = s.camera()
    _, _, mats = s.nodes()
    gd = api.GBufferDevice(ctx, W, H)
    ctx.primary_gbuffer(gs.tlas, cam, mats, gd.desc())
    assert gbuffer_digest(gd.download()) == str(gold["gbuffer_sha256"])
    dsh, dao = ctx.buffer(W * H), ctx.buffer(W * H)
    ctx.trace_visibility(gs.tlas, gd.desc(), sb, dsh, dao)
    np.testing.assert_array_equal(dsh.download(np.uint8, (1, W * H)), gold["shadow"])
    np.testing.assert_array_equal(dao.download(np.uint8, W * H), gold["ao"])
    gs.close()
