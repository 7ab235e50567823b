import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def c0_model():
    """C0 cantilever box (6-tet split, -1 kN z on the tip face), RCM ordered."""
    from wnfea_amd.mesh.rcm import apply_rcm_to_model
    from wnfea_amd.mesh.solid_generator import generate_c3d10_structured_block
    m, _ = generate_c3d10_structured_block(1.0, 0.1, 0.1, 40, 4, 4, tip_load_total=-1000.0, tets_per_hex=6,
                                           load_direction="z")
    mr, n2o, o2n = apply_rcm_to_model(m)
    return m, mr, n2o


@pytest.fixture(scope="session")
def c0_oracle(c0_model):
    """Oracle FP64 CSR system and exact solution of the RCM-ordered C0."""
    import numpy as np
    from scipy.sparse.linalg import spsolve
    from oracle import oracle as O
    _, mr, _ = c0_model
    E, nu = mr.element_moduli()
    K, F = O.assemble_csr(mr.mesh_nodes, mr.solid_elements, E, nu, mr.fixed_dof_mask(3), mr.load_vector(3))
    u = spsolve(K.tocsc(), F)
    return K, F, np.asarray(u)
