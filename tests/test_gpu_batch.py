"""pf2_batch_step: one launch per kernel of the step for several simulations
at once gives every simulation the bits of its own step (pf2_step_with_field),
for FP32-mixed and FP64 handles, in 2-D and 3-D, with uneven sizes, lattices
and parameters, and with the batch's membership changing between steps."""
import numpy as np
import pytest

from parity_util import bitwise_equal
from pfem2_amd import FP32_MIXED, FP64, Batch, Pf2Error, ParticleStage
from pfem2_amd import mesh as M

pytestmark = pytest.mark.gpu


def _cases(dim):
    out = []
    if dim == 2:
        specs = [(12, 4, 3, 0.02, 2), (60, 12, 4, 0.01, 2), (30, 9, 5, 0.015, 3), (8, 8, 4, 0.03, 1),
                 (90, 20, 3, 0.01, 2)]
        for k, (nx, ny, q, dt, mx) in enumerate(specs):
            m = M.channel_2d(nx, ny, 5.0, 1.0)
            V = M.poiseuille_2d(m, 0.5 + 0.3 * k, 1.0, 0.02, seed=80 + k)
            out.append((m, V, [q, q], dt, mx))
    else:
        for k, (n, q) in enumerate([(6, 2), (8, 2), (5, 3)]):
            m = M.pipe_3d(n, n, n + 2, 0.5, 1.0)
            V = M.pipe_flow(m, 1.0, 0.5, eps=0.05, seed=90 + k)
            out.append((m, V, [q, q, q], 0.02, 2))
    return out


def _stage(case, precision, seed):
    m, V, q, dt, mx = case
    st = ParticleStage(m, dt, 8, mx, q, precision)
    st.upload_fields(V, V)
    cell, xi = M.jittered_lattice(m, q, 0.05, seed)
    st.insert_particles(cell, xi)
    return st


@pytest.mark.parametrize("dim,precision", [(2, FP32_MIXED), (2, FP64), (3, FP32_MIXED), (3, FP64)])
def test_batch_step_equals_own_steps(dim, precision):
    import torch
    cases = _cases(dim)
    solo = [_stage(c, precision, 100 + k) for k, c in enumerate(cases)]
    bat = [_stage(c, precision, 100 + k) for k, c in enumerate(cases)]
    fields = [torch.from_numpy(np.ascontiguousarray(c[1])).to("cuda") for c in cases]
    torch.cuda.synchronize()
    b = Batch(len(cases))
    for step in range(4):
        # the membership changes: simulation 1 sits out step 2, simulation 3 steps 0-1 only
        members = [k for k in range(len(cases)) if not (k == 1 and step == 2) and not (k == 3 and step >= 2)]
        for k in members:
            solo[k].step_with_field(fields[k].data_ptr())
        b.step([bat[k] for k in members], [fields[k].data_ptr() for k in members])
        for k in members:
            pa, pb = solo[k].particles(), bat[k].particles()
            for key in ("cell", "id", "x", "xi", "v"):
                assert np.array_equal(pa[key], pb[key]), f"step {step} sim {k} {key}"
            assert bitwise_equal(solo[k].download_field(), bat[k].download_field()), f"step {step} sim {k} V"
            assert solo[k].particle_counter() == bat[k].particle_counter()
            assert solo[k].stats() == bat[k].stats()
    print(f"batch dim {dim} precision {precision}: {len(cases)} simulations bitwise equal", flush=True)
    b.close()


def test_batch_refuses_mixed_sets():
    import torch
    c2 = _cases(2)[0]
    a = _stage(c2, FP32_MIXED, 1)
    d = _stage(c2, FP64, 2)
    f = torch.from_numpy(np.ascontiguousarray(c2[1])).to("cuda")
    torch.cuda.synchronize()
    b = Batch(4)
    with pytest.raises(Pf2Error, match="one dimension and precision"):
        b.step([a, d], [f.data_ptr(), f.data_ptr()])
    with pytest.raises(Pf2Error, match="twice"):
        b.step([a, a], [f.data_ptr(), f.data_ptr()])
    with pytest.raises(Pf2Error, match="more simulations"):
        b.step([a] * 5, [f.data_ptr()] * 5)
    b.close()
