"""Shared fixtures.  GPU tests carry @pytest.mark.gpu and call the HIP path
through the C-ABI; results are compared against the CPU oracle (oracle/)."""
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)
FIXTURES = os.path.join(REPO, "tests", "golden", "fixtures")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    pyoracle.load()
    return pyoracle


@pytest.fixture(scope="session")
def lib():
    from siconos_amd import capi
    return capi.load()


def fixture_path(name):
    return os.path.join(FIXTURES, name)


_UNZIPPED = {}


def fixture_file(name):
    """Path of a fixture; a `.gz` copy is decompressed once into a temporary
    directory."""
    plain = os.path.join(FIXTURES, name)
    if os.path.exists(plain):
        return plain
    if name not in _UNZIPPED:
        import gzip
        import tempfile
        out = os.path.join(tempfile.mkdtemp(prefix="fixtures_"), name)
        with gzip.open(plain + ".gz", "rb") as f, open(out, "wb") as g:
            g.write(f.read())
        _UNZIPPED[name] = out
    return _UNZIPPED[name]


FC2D_COLLECTION = [
    # data_collection_fc2d_1.c, in its order
    "FrictionContactProblem00374.dat", "FrictionContactProblem00394.dat",
    "FC2D_SliderCrankLagrangian00000.dat", "FC2D_SliderCrankLagrangian00001.dat",
    "FrictionContactProblem00727.dat", "FrictionContactProblem00031.dat",
    "FrictionContactProblem00071.dat", "FrictionContactProblem00237.dat",
    "fc2d_granularflowonwall_00667_00000.dat", "fc2d_granularflowonwall_02020_00213.dat",
    "fc2d_granularflowonwall_07002_00801.dat",
]


def read_ac_inputs():
    """ACinputs.dat (reference test data): n, then 3n reactions, 3n
    velocities, n mu, 3n rho (fc3d_AlartCurnierFunctions_test.c:34-70)."""
    v = np.loadtxt(fixture_path("ACinputs.dat"))
    n = int(v[0])
    R = v[1:1 + 3 * n].reshape(n, 3)
    U = v[1 + 3 * n:1 + 6 * n].reshape(n, 3)
    MU = v[1 + 6 * n:1 + 7 * n]
    RHO = v[1 + 7 * n:1 + 10 * n].reshape(n, 3)
    return R, U, MU, RHO


def load_fixture(name):
    """(W as SparseBlock3, q, mu) read by the product's host reader."""
    from siconos_amd import numerics as N
    p = N.FrictionContactProblem.from_file(fixture_path(name))
    return p.sparse_block(), p.q, p.mu


def natural_residual(W, q, mu, r):
    """Independent numpy evaluation of the fc3d natural-map residual
    (the quantity of fc3d_compute_error.c:70-122) for cross-checks."""
    D = W.to_dense() if W.nc <= 3000 else None
    u = (D @ r + q) if D is not None else None
    if u is None:
        u = q.copy()
        for i in range(W.nc):
            for k in range(W.row_ptr[i], W.row_ptr[i + 1]):
                j = W.col_idx[k]
                u[3 * i:3 * i + 3] += W.block(k) @ r[3 * j:3 * j + 3]
    e2 = 0.0
    for i in range(W.nc):
        rr = r[3 * i:3 * i + 3]
        uu = u[3 * i:3 * i + 3]
        m = mu[i]
        w = np.array([rr[0] - uu[0] - m * np.hypot(uu[1], uu[2]), rr[1] - uu[1], rr[2] - uu[2]])
        p = cone_projection(w, m)
        e2 += np.sum((rr - p) ** 2)
    scale = max(np.linalg.norm(q), np.linalg.norm(r), np.linalg.norm(u))
    e = np.sqrt(e2)
    return (e / scale if scale > np.finfo(float).eps else e), u


def cone_projection(x, mu):
    """Euclidean projection on {|x_t| <= mu x_n} from its definition."""
    nt = np.hypot(x[1], x[2])
    if nt <= mu * x[0]:
        return x.copy()
    if mu * nt <= -x[0]:
        return np.zeros(3)
    rn = (x[0] + mu * nt) / (1.0 + mu * mu)
    return np.array([rn, mu * rn * x[1] / nt, mu * rn * x[2] / nt])
