"""pf2_distribute_particle_velocities_to_grid_static (projection sums in the
library's static device array, no allocation) against
pf2_distribute_particle_velocities_to_grid on twin handles holding the same
particles: the projected field, the resident field and the particles are
bitwise equal.  Two calls in a row, the second on a smaller mesh (the array
keeps the first call's sums; nothing in it needs to start at zero)."""
import numpy as np
import pytest

from parity_util import bitwise_equal
from pfem2_amd import FP32_MIXED, FP64, ParticleStage
from pfem2_amd import mesh as M

pytestmark = pytest.mark.gpu


def _twins(m, V, precision, seed):
    stages = []
    cell, xi = M.jittered_lattice(m, [3, 3], 0.05, seed)
    for _ in range(2):
        g = ParticleStage(m, 0.004, 5, 3, [3, 3], precision)
        g.upload_fields(V, V)
        g.insert_particles(cell, xi)
        g.correct_particles_velocities()
        g.move_particles()
        stages.append(g)
    return stages


@pytest.mark.parametrize("precision", [FP64, FP32_MIXED])
def test_static_projection_equals_allocating_entry(precision):
    import torch
    for scale, seed in ((0.15, 71), (0.08, 72)):
        m = M.cylinder_channel(scale, seed=seed)
        V = M.cylinder_flow(m, 1.0, eps=0.02, seed=seed + 10)
        a, b = _twins(m, V, precision, seed + 20)
        print(f"static projection: precision {precision}, {m.n_cells} cells", flush=True)
        Va = a.distribute_particle_velocities_to_grid()
        out = torch.full((m.dim, m.n_dofs), float("nan"), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        b.distribute_particle_velocities_to_grid_static(out.data_ptr())
        b.synchronize()
        Vb = out.cpu().numpy()
        assert bitwise_equal(Va, Vb), f"{m.n_cells} cells: static projection differs"
        assert bitwise_equal(b.download_field(), Va)
        pa, pb = a.particles(), b.particles()
        for k in ("cell", "id", "x", "xi", "v"):
            assert np.array_equal(pa[k], pb[k]), k
        # the step after the projection sees the same field on both handles
        a.step()
        b.step()
        assert bitwise_equal(a.download_field(), b.download_field())
        a.close()
        b.close()
