"""The C++ facade (include/pfem2_amd.hpp, the reference's pfem2Solver /
pfem2ParticleHandler names over the C ABI) compiles and links; on the GPU, a
driver written against it (tests/cpp/driver_replay.cpp, the reference
driver's call order) gives the same bits as the same calls through the
Python binding."""
import os
import subprocess

import numpy as np
import pytest

from pfem2_amd import FP32_MIXED, FP64, ParticleStage
from pfem2_amd import mesh as M

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "build", "driver_replay")


def _build():
    subprocess.check_call(["bash", os.path.join(HERE, "cpp", "build.sh")])


def test_facade_compiles_and_links():
    _build()
    assert os.access(BIN, os.X_OK)


def _write_input(path, m, V, steps, dt, S, MAX, q, precision):
    q3 = list(q) + [1] * (3 - len(q))
    hdr = np.array([m.dim, m.n_cells, m.n_vertices, m.n_dofs, steps, S, MAX, *q3, precision], dtype=np.int32)
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(np.array([dt], dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(m.vertices, dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(m.cells, dtype=np.int32).tobytes())
        f.write(np.ascontiguousarray(V, dtype=np.float64).tobytes())


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [FP64, FP32_MIXED])
def test_driver_replay_matches_python(tmp_path, precision):
    if not os.access(BIN, os.X_OK):
        _build()
    m = M.cylinder_channel(0.05, seed=71)
    V = M.cylinder_flow(m, 1.0, eps=0.02, seed=72)
    steps, dt, S, MAX, q = 5, 0.003, 6, 2, [3, 3]
    inp, out = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    _write_input(inp, m, V, steps, dt, S, MAX, q, precision)
    print(f"driver replay precision={precision}", flush=True)
    subprocess.run([BIN, inp, out], check=True, timeout=120)
    raw = open(out, "rb").read()
    n = int(np.frombuffer(raw[:8], dtype=np.int64)[0])
    cnt = int(np.frombuffer(raw[8:12], dtype=np.uint32)[0])
    cell_sum = int(np.frombuffer(raw[12:20], dtype=np.int64)[0])
    Vc = np.frombuffer(raw[20:], dtype=np.float64).reshape(m.dim, m.n_dofs)

    st = ParticleStage(m, dt, S, MAX, q, precision)
    st.upload_fields(V, V)
    st.seed_particles()
    for _ in range(steps):
        st.upload_fields(V, None)
        st.correct_particles_velocities()
        st.move_particles()
        Vp = st.distribute_particle_velocities_to_grid()
    p = st.particles()
    assert n == st.n_global_particles()
    assert cnt == st.particle_counter()
    assert cell_sum == int(p["cell"].astype(np.int64).sum())
    assert np.array_equal(Vc, Vp, equal_nan=True)
    st.close()
